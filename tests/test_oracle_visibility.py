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


def test_mask_value_two_skips_pairs_among_twos():
    """Mask 2: a new point whose pairs with the other 2-points are skipped;
    its pairs with 0- and 1-points are tested as for 1."""
    V, T = wl.perturbed_icosphere(2, 0.4, 8)
    n = V.shape[0]
    rng = np.random.default_rng(1)
    mask = rng.choice(np.array([0, 1, 2], np.uint8), size=n, p=[0.6, 0.2, 0.2])
    full, _ = oracle.visibility(V, V, T)
    part, st = oracle.visibility(V, V, T, mask=mask)
    keep = [(i, j) for i, j in full.tolist()
            if (mask[i] or mask[j]) and not (mask[i] == 2 and mask[j] == 2)]
    assert set(map(tuple, part.tolist())) == set(keep)
    n0, n2 = int((mask == 0).sum()), int((mask == 2).sum())
    assert st["masked"] == n0 * (n0 - 1) // 2 + n2 * (n2 - 1) // 2


def test_mask_cap_bits_skip_pairs_on_one_cut_plane():
    """Bits 2-7 mark points on older cut planes: 4 = old point on plane 1,
    6 = new point on the last cap and on plane 1.  Skipped: old-old pairs
    (0 and 4 are old), pairs of two 6s (shared last cap) and 4-6 pairs
    (shared plane 1); 1-x, 0-6 and the rest are tested as usual."""
    V, T = wl.perturbed_icosphere(2, 0.4, 8)
    n = V.shape[0]
    rng = np.random.default_rng(3)
    mask = rng.choice(np.array([0, 1, 4, 6], np.uint8), size=n, p=[0.4, 0.2, 0.2, 0.2])
    full, _ = oracle.visibility(V, V, T)
    part, st = oracle.visibility(V, V, T, mask=mask)
    n0, n4, n6 = (int((mask == v).sum()) for v in (0, 4, 6))
    old = n0 + n4
    assert st["masked"] == old * (old - 1) // 2 + n6 * (n6 - 1) // 2 + n4 * n6
    tested = {(0, 1), (1, 1), (1, 4), (1, 6), (0, 6)}
    keep = [(i, j) for i, j in full.tolist()
            if tuple(sorted((int(mask[i]), int(mask[j])))) in tested]
    assert set(map(tuple, part.tolist())) == set(keep)


def test_k_order_and_pair_decode():
    n = 1000
    for k in [0, 1, 998, 999, 12345, n * (n - 1) // 2 - 1]:
        i, j = oracle.k_to_ij(k, n)
        assert 0 <= i < j < n
        assert i * (2 * n - i - 1) // 2 + j - i - 1 == k
