"""Pins the oracle's exact predicates against exact rational arithmetic.

Every orientation determinant of float32 inputs is recomputed with Python
Fractions (exact) on random, near-degenerate (1-ulp perturbations of exactly
coplanar / collinear configurations) and exactly degenerate inputs.
"""
from fractions import Fraction as F

import numpy as np
import pytest

import oracle


def _fr(v):
    return [F(float(x)) for x in np.asarray(v, np.float32)]


def det3(r0, r1, r2):
    return (r0[0] * (r1[1] * r2[2] - r1[2] * r2[1]) - r0[1] * (r1[0] * r2[2] - r1[2] * r2[0]) +
            r0[2] * (r1[0] * r2[1] - r1[1] * r2[0]))


def o3_exact(a, b, c, d):
    a, b, c, d = map(_fr, (a, b, c, d))
    rows = [[a[k] - d[k] for k in range(3)], [b[k] - d[k] for k in range(3)],
            [c[k] - d[k] for k in range(3)]]
    v = det3(*rows)
    return (v > 0) - (v < 0)


def o3_mid_exact(a, b, c, p, q):
    a, b, c, p, q = map(_fr, (a, b, c, p, q))
    m = [(p[k] + q[k]) / 2 for k in range(3)]
    rows = [[a[k] - m[k] for k in range(3)], [b[k] - m[k] for k in range(3)],
            [c[k] - m[k] for k in range(3)]]
    v = det3(*rows)
    return (v > 0) - (v < 0)


AX = {0: (1, 2), 1: (2, 0), 2: (0, 1)}


def o2_exact(a, b, c, k, mid=None):
    u, v = AX[k]
    a, b = map(_fr, (a, b))
    if mid is not None:
        p, q = map(_fr, mid)
        c = [(p[t] + q[t]) / 2 for t in range(3)]
    else:
        c = _fr(c)
    val = (a[u] - c[u]) * (b[v] - c[v]) - (a[v] - c[v]) * (b[u] - c[u])
    return (val > 0) - (val < 0)


def _nudge(x, steps):
    x = np.float32(x)
    for _ in range(abs(steps)):
        x = np.nextafter(x, np.float32(np.inf if steps > 0 else -np.inf), dtype=np.float32)
    return x


def _cases(rng, n):
    out = []
    for t in range(n):
        kind = t % 4
        if kind == 0:  # generic
            pts = rng.uniform(-1, 1, size=(4, 3)).astype(np.float32)
        elif kind == 1:  # exactly coplanar on a small integer lattice
            base = rng.integers(-4, 5, size=(3, 3)).astype(np.float64)
            w = rng.integers(-3, 4, size=2)
            d = base[0] + w[0] * (base[1] - base[0]) + w[1] * (base[2] - base[0])
            pts = (np.vstack([base, d]) * 0.125).astype(np.float32)
        elif kind == 2:  # coplanar lattice with one coordinate nudged by 1-2 ulps
            base = rng.integers(-4, 5, size=(3, 3)).astype(np.float64)
            w = rng.integers(-3, 4, size=2)
            d = base[0] + w[0] * (base[1] - base[0]) + w[1] * (base[2] - base[0])
            pts = (np.vstack([base, d]) * 0.375).astype(np.float32)
            i, k = rng.integers(0, 4), rng.integers(0, 3)
            pts[i, k] = _nudge(pts[i, k], int(rng.choice([-2, -1, 1, 2])))
        else:  # nearly coplanar, general float values
            a, b, c = rng.uniform(-1, 1, size=(3, 3))
            s, r = rng.uniform(-1, 1, 2)
            d = a + s * (b - a) + r * (c - a)
            pts = np.vstack([a, b, c, d]).astype(np.float32)
        out.append(pts)
    return out


def test_orient3d_matches_rationals():
    rng = np.random.default_rng(0)
    zeros = 0
    for pts in _cases(rng, 800):
        a, b, c, d = pts
        e = o3_exact(a, b, c, d)
        zeros += e == 0
        assert oracle.orient3d(a, b, c, d) == e
    assert zeros > 100  # the exact-degenerate cases were exercised


def test_orient3d_mid_matches_rationals():
    rng = np.random.default_rng(1)
    for pts in _cases(rng, 400):
        a, b, c, p = pts
        # q chosen so that the midpoint is exactly d when possible
        q = (2.0 * p.astype(np.float64) - rng.uniform(-1, 1, 3) * 0.0).astype(np.float32)
        if rng.uniform() < 0.5:
            q = rng.uniform(-1, 1, 3).astype(np.float32)
        assert oracle.orient3d_mid(a, b, c, p, q) == o3_mid_exact(a, b, c, p, q)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_orient2d_matches_rationals(k):
    rng = np.random.default_rng(2 + k)
    zeros = 0
    for pts in _cases(rng, 400):
        a, b, c, d = pts
        e = o2_exact(a, b, d, k)
        zeros += e == 0
        assert oracle.orient2d(a, b, d, k) == e
        e2 = o2_exact(a, b, None, k, mid=(c, d))
        assert oracle.orient2d_mid(a, b, c, d, k) == e2
    assert zeros > 10


def test_orient3d_sign_conventions():
    a, b, c = [1, 0, 0], [0, 1, 0], [0, 0, 1]
    assert oracle.orient3d(a, b, c, [0, 0, 0]) == 1  # det[a-d, b-d, c-d] = det(I)
    assert oracle.orient3d(b, a, c, [0, 0, 0]) == -1
    assert oracle.orient2d([0, 0, 0], [1, 0, 0], [0, 1, 0], 2) == 1  # ccw in (x, y)


# ------------------------------------------------------------- contact rule
T0 = (np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32),
      np.array([0, 1, 0], np.float32))


@pytest.mark.parametrize("p,q,expect", [
    ([0.2, 0.2, 1], [0.2, 0.2, -1], True),      # through the interior
    ([2, 2, 1], [2, 2, -1], False),             # misses
    ([0.5, 0, 1], [0.5, 0, -1], True),          # through an edge
    ([0, 0, 1], [0, 0, -1], True),              # through a vertex
    ([0.2, 0.2, 1], [0.2, 0.2, 0.5], False),    # stops above
    ([0.2, 0.2, 1], [0.2, 0.2, 0], True),       # ends on the face (not a vertex)
    ([-1, 0.25, 0], [2, 0.25, 0], True),        # coplanar, crossing
    ([-1, 2, 0], [2, 2, 0], False),             # coplanar, disjoint
    ([0, 0, 0], [0, 0, 1], False),              # from a vertex, leaves the plane
    ([0, 0, 0], [1, 1, 0], True),               # from a vertex, into the triangle
    ([0, 0, 0], [-1, -1, 0], False),            # from a vertex, away in-plane
    ([0, 0, 0], [2, 0, 0], True),               # from a vertex along an edge
    ([0, 0, 0], [1, 0, 0], True),               # the edge itself
    ([1, 0, 0], [5, 5, 5], False),              # from a vertex, off-plane
])
def test_contact_cases(p, q, expect):
    assert oracle.contact(p, q, *T0) == expect
    assert oracle.contact(q, p, *T0) == expect  # symmetric in the endpoints


def test_degenerate_triangle_never_touched():
    a, b, c = [0, 0, 0], [1, 1, 1], [2, 2, 2]
    assert not oracle.contact([0.5, 0.5, 0], [0.5, 0.5, 2], a, b, c)
    assert not oracle.contact([1, 1, 1], [3, 0, 0], a, b, c)


# ------------------------------------------------------------- ray parity
def _icosphere(level):
    from visacd_amd import workloads as wl

    v, f = wl.icosphere(level)
    return v.astype(np.float32), f.astype(np.uint32)


def _parity(p, q, V, T):
    cnt = 0
    for t in T:
        cnt += oracle.ray_cross(p, q, V[t[0]], V[t[1]], V[t[2]])
    return cnt & 1


def test_ray_parity_watertight_through_vertices_and_edges():
    """Rays aimed exactly at mesh vertices / edge midpoints still give the
    right inside/outside answer (the symbolic perturbation is consistent)."""
    V, T = _icosphere(1)
    rng = np.random.default_rng(4)
    checked = 0
    for vid in rng.choice(V.shape[0], 12, replace=False):
        y, z = V[vid, 1], V[vid, 2]
        for x, inside in ((np.float32(0.0), True), (np.float32(-3.0), False)):
            # the +x ray from (x, y, z) passes exactly through vertex vid
            m = np.array([x, y, z], np.float32)
            r2 = float(np.dot(m.astype(np.float64), m.astype(np.float64)))
            if x == 0.0 and r2 > 0.6:
                continue
            assert _parity(m, m, V, T) == int(inside)
            checked += 1
    assert checked >= 12


def test_ray_cross_matches_rationals_generic():
    rng = np.random.default_rng(5)
    for _ in range(300):
        a, b, c = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        p, q = rng.uniform(-1, 1, size=(2, 3)).astype(np.float32)
        s1 = o2_exact(a, b, None, 0, mid=(p, q))
        s2 = o2_exact(b, c, None, 0, mid=(p, q))
        s3 = o2_exact(c, a, None, 0, mid=(p, q))
        if 0 in (s1, s2, s3):
            continue
        expect = s1 == s2 == s3 and o3_mid_exact(a, b, c, p, q) == s1
        assert oracle.ray_cross(p, q, a, b, c) == expect
