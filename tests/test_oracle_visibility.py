"""Pins the oracle's visibility definition without a GPU.

Closed forms (convex shapes, L-shape, U-channel), brute force vs the CPU
BVH, endpoint symmetry, exact rotation equivariance (axis permutations and
sign flips are exact in float32, and the parity ray direction must not
matter on closed meshes), the cage/parity subset relation and the mask.
"""
import numpy as np
import pytest

import meshes
import oracle
from visacd_amd import workloads as wl


def test_convex_has_no_edges_and_blocks_its_mesh_edges():
    v, f = wl.icosphere(2)
    V = wl.normalize_like_reference(v).astype(np.float32)
    T = f.astype(np.uint32)
    e, st = oracle.visibility(V, V, T, brute=True)
    assert len(e) == 0
    assert st["blocked"] == 3 * T.shape[0] // 2  # every mesh edge lies on the surface


def test_l_shape_closed_form():
    V, T = meshes.l_shape()
    e, _ = oracle.visibility(V, V, T, brute=True)
    # the notch is x>=1, y>=1 (cells of 0.25): edges join vertices x in {2,3}
    # on the wall y=1 with vertices y in {2,3} on the wall x=1, any z
    P = np.round(V / 0.25).astype(int)
    expect = set()
    for i in range(len(P)):
        for j in range(i + 1, len(P)):
            a, b = P[i], P[j]
            for u, w in ((a, b), (b, a)):
                if u[1] == 1 and u[0] in (2, 3) and w[0] == 1 and w[1] in (2, 3):
                    expect.add((i, j))
    assert len(expect) == 16
    assert set(map(tuple, e.tolist())) == expect


def test_u_channel_closed_form():
    V, T = meshes.u_channel()
    e, _ = oracle.visibility(V, V, T, brute=True)
    P = np.round(V / 0.25).astype(int)
    expect = set()
    for i in range(len(P)):
        for j in range(i + 1, len(P)):
            a, b = P[i], P[j]
            # across the notch x in [1, 2], y in [1, 2]: walls x=1 and x=2,
            # except pairs lying on the notch floor y=1
            walls = {a[0], b[0]} == {1, 2} and a[1] in (1, 2) and b[1] in (1, 2)
            if walls and not (a[1] == 1 and b[1] == 1):
                expect.add((i, j))
    assert len(expect) == 12
    assert set(map(tuple, e.tolist())) == expect


@pytest.mark.parametrize("name", ["torus-small", "poly0", "poly1", "ico3"])
def test_bvh_equals_brute_force(name):
    if name == "torus-small":
        V, T = wl.torus(nu=16, nv=8, seed=3)
    elif name == "ico3":
        V, T = wl.perturbed_icosphere(2, 0.4, 9)
    else:
        rng = np.random.default_rng(int(name[-1]))
        cells = {(x, y, z) for x in range(3) for y in range(3) for z in range(2)
                 if rng.uniform() < 0.7} | {(0, 0, 0)}
        V, T = meshes.polycube(sorted(cells))
    a, sa = oracle.visibility(V, V, T, brute=True)
    b, sb = oracle.visibility(V, V, T, brute=False)
    np.testing.assert_array_equal(a, b)
    assert sa == sb


def _remap(edges, perm_inv):
    out = set()
    for i, j in edges.tolist():
        a, b = perm_inv[i], perm_inv[j]
        out.add((min(a, b), max(a, b)))
    return out


def test_endpoint_symmetry():
    V, T = wl.perturbed_icosphere(2, 0.4, 21)
    n = V.shape[0]
    e, _ = oracle.visibility(V, V, T)
    rev = np.arange(n)[::-1].copy()  # point i of the reversed list is vertex n-1-i
    e2, _ = oracle.visibility(V[rev], V, T)
    assert _remap(e2, rev) == set(map(tuple, e.tolist()))


@pytest.mark.parametrize("perm,sign", [((1, 2, 0), (1, 1, 1)), ((2, 0, 1), (-1, 1, -1)),
                                       ((0, 1, 2), (-1, -1, 1))])
def test_rotation_equivariance(perm, sign):
    for V, T in (wl.perturbed_icosphere(2, 0.4, 5), meshes.l_shape(), wl.torus(16, 8, seed=1)[::1]):
        V2 = (V[:, list(perm)] * np.array(sign, np.float32)).astype(np.float32)
        T2 = T.copy()
        if np.prod(sign) * (1 if perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1) < 0:
            T2 = T2[:, [0, 2, 1]]  # keep outward orientation under a reflection
        a, _ = oracle.visibility(V, V, T)
        b, _ = oracle.visibility(V2, V2, np.ascontiguousarray(T2))
        np.testing.assert_array_equal(a, b)


def test_cage_edges_are_parity_edges():
    V, T = wl.perturbed_icosphere(2, 0.3, 4)
    cage = (meshes.cage_of(V, 1.05), T)
    ec, _ = oracle.visibility(V, V, T, cage=cage)
    ep, _ = oracle.visibility(V, V, T)
    assert len(ec) > 0
    assert set(map(tuple, ec.tolist())) <= set(map(tuple, ep.tolist()))


def test_mask_skips_old_pairs():
    V, T = wl.perturbed_icosphere(2, 0.4, 8)
    n = V.shape[0]
    rng = np.random.default_rng(0)
    mask = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    full, _ = oracle.visibility(V, V, T)
    part, st = oracle.visibility(V, V, T, mask=mask)
    keep = [(i, j) for i, j in full.tolist() if mask[i] or mask[j]]
    assert set(map(tuple, part.tolist())) == set(keep)
    nm = int((mask == 0).sum())
    assert st["masked"] == nm * (nm - 1) // 2


def test_mask_bits_above_zero_do_not_skip_without_groups():
    """Only bit 0 marks a new point: a pair is skipped when neither point has
    it; bits 1-7 (points on a group plane) change nothing without groups."""
    V, T = wl.perturbed_icosphere(2, 0.4, 8)
    n = V.shape[0]
    rng = np.random.default_rng(1)
    mask = rng.choice(np.array([0, 1, 2, 3, 6], np.uint8), size=n, p=[0.4, 0.15, 0.15, 0.15, 0.15])
    full, _ = oracle.visibility(V, V, T)
    part, st = oracle.visibility(V, V, T, mask=mask)
    keep = [(i, j) for i, j in full.tolist() if (mask[i] | mask[j]) & 1]
    assert set(map(tuple, part.tolist())) == set(keep)
    old = int(((mask & 1) == 0).sum())
    assert st["masked"] == old * (old - 1) // 2


def _clipped(V, T, plane):
    """The positive-side part of the host clip, float32, with its new
    (cap) vertices and the cap triangles as group 0."""
    from visacd_amd import _visacd as H

    m = H.Mesh()
    m.vertices = H.VecArray3d(np.asarray(V, np.float64).tolist())
    m.triangles = H.make_vecarray3i(np.asarray(T, np.int32))
    parts, _mp, _mn, flags = H.clip(m, list(plane))
    part = parts[0]
    Pv = np.array([list(x) for x in part.vertices], np.float64)
    Pt = np.array([list(x) for x in part.triangles], np.int64)
    new = np.asarray(flags[0], bool)
    on = np.abs(Pv @ np.asarray(plane[:3]) + plane[3]) <= 1e-9 * np.linalg.norm(plane[:3])
    grp = np.where(on[Pt].all(1), 0, -1).astype(np.int8)
    return Pv.astype(np.float32), Pt.astype(np.uint32), new, on, grp


def test_cap_rule_keeps_edges_across_the_gap_of_a_u_channel():
    """The U-channel cut through both arms (y = 0.375, exact in float32)
    leaves two rectangular caps in one plane.  Pairs of cap points across
    the gap between the arms cross air and are edges; pairs within one cap
    run into it.  With exactly planar caps the cap rule (mask bit 1 on the
    cap points, the cap triangles as group 0) must give exactly the edges of
    the plain mask (new points as 1), and the gap-crossing pairs must be
    among them."""
    V, T = meshes.u_channel()
    plane = (0.0, -1.0, 0.0, 0.375)  # positive side: y < 0.375 (the lower part)
    Pv, Pt, new, on, grp = _clipped(V, T, plane)
    assert new.sum() == 16 and (on & new).sum() == 16 and (grp == 0).sum() >= 4
    m1 = new.astype(np.uint8)
    e1, _ = oracle.visibility(Pv, Pv, Pt, mask=m1)
    m_rule = (new.astype(np.uint8) | (2 * on).astype(np.uint8))
    e2, st2 = oracle.visibility(Pv, Pv, Pt, mask=m_rule, groups=(grp, [plane]))
    np.testing.assert_array_equal(e2, e1)
    cap = np.where(new)[0]
    left = [i for i in cap if Pv[i, 0] < 0.3]
    right = [i for i in cap if Pv[i, 0] > 0.4]
    assert len(left) == 8 and len(right) == 8
    edges = set(map(tuple, e2.tolist()))
    across = {(min(a, b), max(a, b)) for a in left for b in right}
    assert len(across & edges) >= 9
    within = {(min(a, b), max(a, b)) for a in left for b in left if a != b}
    # diagonals of one rectangular cap run through it: never edges
    assert not (within & edges) - set(map(tuple, e1.tolist()))


def test_cap_rule_only_removes_in_plane_pairs_that_enter_the_cap():
    """On a rounded cut (a general plane: the cap vertices are rounded to
    float32) the cap rule may only turn edges into non-edges, and only pairs
    of two points on the cap plane; it removes every in-plane pair whose
    direction enters a cap triangle at either end (checked on the chords of
    the cap that pass through the part's section)."""
    V, T = wl.perturbed_icosphere(3, 0.3, 21)
    nrm = np.array([0.3, 0.5, -0.2])
    nrm /= np.linalg.norm(nrm)
    plane = (*nrm, 0.05)
    Pv, Pt, new, on, grp = _clipped(V, T, plane)
    assert (grp == 0).sum() > 16
    m1 = new.astype(np.uint8)
    m_rule = (m1 | (2 * on).astype(np.uint8))
    e1, _ = oracle.visibility(Pv, Pv, Pt, mask=m1)
    e2, _ = oracle.visibility(Pv, Pv, Pt, mask=m_rule, groups=(grp, [plane]))
    s1, s2 = set(map(tuple, e1.tolist())), set(map(tuple, e2.tolist()))
    assert s2 <= s1
    for i, j in s1 - s2:
        assert on[i] and on[j]
    # no in-plane pair left an edge enters a cap triangle at either end
    # (wedge test restated with the Fraction-pinned orient2d)
    k = int(np.argmax(np.abs(nrm)))
    caps = [t for t in range(Pt.shape[0]) if grp[t] == 0]

    def fan_size(a):
        return int(np.all(Pv[Pt] == Pv[a], axis=2).any(axis=1).sum())

    def enters(a, b):  # an end with more than 16 incident triangles is not tested
        p, q = Pv[a], Pv[b]
        if fan_size(a) > 16:
            return False
        for t in caps:
            tri = [Pv[v] for v in Pt[t]]
            at = [x for x in range(3) if np.array_equal(tri[x], p)]
            if not at:
                continue
            v1, v2 = tri[(at[0] + 1) % 3], tri[(at[0] + 2) % 3]
            o = oracle.orient2d(p, v1, v2, k)
            o1, o2 = oracle.orient2d(p, v1, q, k), oracle.orient2d(p, q, v2, k)
            if o != 0 and (o1 or o2) and o * o1 >= 0 and o * o2 >= 0:
                return True
        return False

    in_plane = [(i, j) for i, j in s2 if on[i] and on[j]]
    assert not any(enters(i, j) or enters(j, i) for i, j in in_plane)
    removed = [(i, j) for i, j in s1 - s2]
    assert all(enters(i, j) or enters(j, i) for i, j in removed)


def test_k_order_and_pair_decode():
    n = 1000
    for k in [0, 1, 998, 999, 12345, n * (n - 1) // 2 - 1]:
        i, j = oracle.k_to_ij(k, n)
        assert 0 <= i < j < n
        assert i * (2 * n - i - 1) // 2 + j - i - 1 == k
