"""Host driver services of the visacd module (CPU only, no GPU calls)."""
import numpy as np
import pytest

import meshes
from visacd_amd import workloads as wl

vis = pytest.importorskip("visacd")
hs = pytest.importorskip("visacd_amd._visacd")


def _mesh(V, T):
    m = vis.Mesh()
    m.vertices = vis.VecArray3d([list(map(float, v)) for v in V])
    m.triangles = vis.make_vecarray3i(np.asarray(T, np.int32))
    return m


def _arrays(m):
    return np.array([list(v) for v in m.vertices]), np.array([list(t) for t in m.triangles])


def _closed(T):
    d = set()
    for a, b, c in T:
        for e in ((a, b), (b, c), (c, a)):
            assert e not in d, "edge used twice in one direction"
            d.add(e)
    return all((b, a) in d for a, b in d)


def test_surface_types():
    m = vis.Mesh()
    assert len(m.vertices) == 0
    m.vertices.append([0.0, 1.0, 2.0])
    assert list(m.vertices[0]) == [0.0, 1.0, 2.0]
    t = vis.make_vecarray3i(np.array([[0, 1, 2], [2, 3, 4]], np.int32))
    assert len(t) == 2 and list(t[1]) == [2, 3, 4]
    c = vis.Config()
    assert (c.return_parts, c.score_mode, c.flat_surface_min_area, c.use_flat_surfaces,
            c.flat_surface_k, c.use_merging) == (False, "concavity", 0.1, True, 2.0, False)
    vis.config.flat_surface_k = 3.0
    assert vis.config.flat_surface_k == 3.0  # module attribute is the global object
    vis.config.flat_surface_k = 2.0
    vis.set_seed(7)
    r = vis.ProcessResult
    assert {"parts", "concavity", "num_parts"} <= set(dir(r))


@pytest.mark.parametrize("plane", [(1, 0, 0, 0.1), (0.3, 0.5, -0.2, 0.05), (0, 0, 1, 0.0)])
def test_clip_is_watertight_and_conserves_volume(plane):
    V, T = wl.perturbed_icosphere(3, 0.3, 3)
    m = _mesh(V, T)
    vol = hs.mesh_volume(m)
    parts, mp, mn, flags = hs.clip(m, np.asarray(plane, float).tolist())
    assert len(parts) == 2
    total = 0.0
    for p, f in zip(parts, flags):
        pv, pt = _arrays(p)
        assert _closed(pt.tolist())
        total += hs.mesh_volume(p)
        n_old = sum(1 for x in f if not x)
        assert all(not x for x in f[:n_old]) and all(f[n_old:])  # cap vertices last
    assert abs(total - vol) < 1e-9 * max(1.0, abs(vol)) + 1e-12
    # 1-based maps: positive side vertices in increasing id order
    n = np.asarray(plane[:3], float)
    side = np.asarray(V, float) @ n + plane[3]
    pos = np.flatnonzero(side > 1e-6)
    assert [mp[i] for i in pos] == list(range(1, len(pos) + 1))


def test_clip_nonconvex_cap_with_holes():
    # torus cut by its equatorial plane z=0 gives an annulus-shaped cap pair
    V, T = wl.torus(nu=32, nv=16, jitter=0.0)
    m = _mesh(V, T)
    parts, _, _, _ = hs.clip(m, [0.0, 0.0, 1.0, 1e-3])
    assert len(parts) == 2
    for p in parts:
        assert _closed(_arrays(p)[1].tolist())
    # cut through the hole axis: x=0 makes two C-shaped pieces, each closed
    parts, _, _, _ = hs.clip(m, [1.0, 0.0, 0.0, 1e-3])
    for p in parts:
        assert _closed(_arrays(p)[1].tolist())


def test_hull_of_sphere_and_cube():
    V, T = wl.icosphere(3)
    h = hs.convex_hull(_mesh(V, T))
    assert abs(hs.mesh_volume(h) - hs.mesh_volume(_mesh(V, T))) < 1e-9
    assert _closed(_arrays(h)[1].tolist())
    Vb, Tb = meshes.box((0, 0, 0), (2, 2, 2))
    hb = hs.convex_hull(_mesh(Vb, Tb))
    assert abs(hs.mesh_volume(hb) - 0.125) < 1e-12


def test_concavity_convex_vs_concave():
    V, T = wl.icosphere(3)
    m = _mesh(V.astype(np.float64), T)
    h = hs.convex_hull(m)
    assert hs.concavity(m, h, 3000, 42) < 0.02
    Vl, Tl = meshes.l_shape()
    ml = _mesh(Vl, Tl)
    hl = hs.convex_hull(ml)
    c = hs.concavity(ml, hl, 3000, 42)
    assert 0.05 < c < 0.6


def test_separate_disjoint():
    A, TA = meshes.box((0, 0, 0), (1, 1, 1))
    B, TB = meshes.box((3, 0, 0), (4, 1, 1))
    V = np.concatenate([A, B])
    T = np.concatenate([TA, TB + len(A)])
    parts = hs.separate_disjoint(_mesh(V, T))
    assert len(parts) == 2


def test_flat_surfaces_of_a_box():
    V, T = meshes.box((0, 0, 0), (4, 4, 4))  # 1 x 1 x 1 after the 0.25 scale
    planes = hs.flat_surface_planes(_mesh(V, T), 0.1)
    assert len(planes) == 6
    for a, b, c, d in planes:
        assert abs(np.linalg.norm([a, b, c]) - 1) < 1e-12


def _cone(n=64, h=2e-6):
    """Closed cone: apex just above z = 0 (outside the 1e-6 dead zone), base
    ring at z = -1, base centre below.  The plane z = 0 crosses the n slanted
    edges within about 2e-6 of the apex, so the crossings lie closer together
    than float32 resolution at the mesh's scale."""
    ang = 2 * np.pi * np.arange(n) / n
    V = [[0.0, 0.0, h]] + [[np.cos(a), np.sin(a), -1.0] for a in ang] + [[0.0, 0.0, -1.0]]
    T = []
    for i in range(n):
        a, b = 1 + i, 1 + (i + 1) % n
        T.append([0, a, b])           # side, outward
        T.append([n + 1, b, a])       # base, outward (down)
    return np.array(V), np.array(T)


def test_clip_welds_crossings_closer_than_float32_resolution():
    V, T = _cone()
    m = _mesh(V, T)
    vol = hs.mesh_volume(m)
    assert vol > 0
    parts, mp, mn, flags = hs.clip(m, [0.0, 0.0, 1.0, 0.0])
    assert len(parts) == 2
    total = 0.0
    for p, f in zip(parts, flags):
        pv, pt = _arrays(p)
        assert _closed(pt.tolist())
        assert len(f) == len(pv)
        n_old = sum(1 for x in f if not x)
        assert all(not x for x in f[:n_old]) and all(f[n_old:])
        new = np.flatnonzero(f)
        # no two new vertices left within the weld distance, nor at one float32 position
        q = pv[new]
        d = np.linalg.norm(q[:, None] - q[None], axis=2) + np.eye(len(q))
        assert d.min() >= 4e-7 * np.abs(pv).max()
        assert len(np.unique(q.astype(np.float32), axis=0)) == len(q)
        total += hs.mesh_volume(p)
    # welding moves cap points by less than 1e-6 next to the tiny cross-section
    assert abs(total - vol) < 1e-6 * vol
    # the maps of old vertices are unchanged by the weld
    assert mp[0] == 1 and all(mn[i] == i for i in range(1, len(V)))
