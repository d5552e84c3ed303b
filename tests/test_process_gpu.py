"""visacd.process end to end on the GPU (drop-in surface of the reference)."""
import numpy as np
import pytest

import meshes
from visacd_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _mesh(vis, V, T):
    m = vis.Mesh()
    m.vertices = vis.VecArray3d([list(map(float, v)) for v in V])
    m.triangles = vis.make_vecarray3i(np.asarray(T, np.int32))
    return m


def _volume(m):
    V = np.array([list(v) for v in m.vertices])
    T = np.array([list(t) for t in m.triangles])
    return np.einsum("ij,ij->i", V[T[:, 0]], np.cross(V[T[:, 1]], V[T[:, 2]])).sum() / 6


def _run(vis, V, T, conc, n, **cfg):
    for k, v in cfg.items():
        setattr(vis.config, k, v)
    vis.set_seed(42)
    return vis.process(_mesh(vis, V, T), conc, n)


def test_process_parts_partition_the_volume():
    import visacd as vis

    vis.engine.verbose = True
    V, T = wl.perturbed_icosphere(3, 0.4, 1)
    print("process ico3", flush=True)
    r = _run(vis, V, T, 0.02, 6, return_parts=True)
    assert 2 <= r.num_parts == len(r.parts)
    vol = _volume(_mesh(vis, V, T))
    total = sum(_volume(p) for p in r.parts)
    assert abs(total - vol) < 1e-6 * abs(vol) + 1e-9
    assert np.isfinite(r.concavity) and r.concavity >= 0
    vis.config.return_parts = False


def test_process_hulls_deterministic_and_concavity_drops():
    import visacd as vis

    V, T = meshes.l_shape()
    print("process L", flush=True)
    r1 = _run(vis, V, T, 0.01, 4)
    r2 = _run(vis, V, T, 0.01, 4)
    assert r1.num_parts == r2.num_parts >= 2
    assert r1.concavity == r2.concavity
    for a, b in zip(r1.parts, r2.parts):
        assert [list(v) for v in a.vertices] == [list(v) for v in b.vertices]
    r0 = _run(vis, V, T, 0.01, 1)  # no split allowed
    assert r0.num_parts == 1
    assert r1.concavity < r0.concavity


def test_process_edge_mode_and_sphere_planes():
    import visacd as vis

    V, T = wl.torus(nu=24, nv=12, seed=3)
    print("process torus edge mode", flush=True)
    vis.engine.planes = "sphere"
    vis.engine.num_planes = 256
    try:
        r = _run(vis, V, T, 0.04, 4, score_mode="edge")
        assert r.num_parts >= 2
    finally:
        vis.engine.planes = "bisector"
        vis.engine.num_planes = 2500
        vis.config.score_mode = "concavity"


def test_process_with_point_limit():
    import visacd as vis

    V, T = wl.perturbed_icosphere(4, 0.3, 2)
    print("process ico4 with max_points", flush=True)
    vis.engine.max_points = 512
    try:
        r = _run(vis, V, T, 0.04, 4, return_parts=True)
        assert r.num_parts >= 2
    finally:
        vis.engine.max_points = 0
        vis.config.return_parts = False


def _parts(r):
    return [[list(v) for v in p.vertices] for p in r.parts]


@pytest.mark.parametrize("conc,n", [(0.001, 9), (0.06, 40)])
def test_batched_frontier_gives_the_plain_decomposition(conc, n):
    """The batched frontier replays the plain loop on the tree of parts it has
    split (speculatively, several per round): the parts, their order and the
    final concavity must be exactly the plain loop's, whether the part budget
    (conc 0.001) or the threshold (conc 0.06) ends the run, for any round
    width and with or without eager visibility."""
    import visacd as vis
    from visacd_amd import _visacd

    V, T = wl.perturbed_icosphere(3, 0.4, 1)
    print("batched vs plain ico3", conc, n, flush=True)
    try:
        vis.engine.mode = "plain"
        r0 = _run(vis, V, T, conc, n, return_parts=True)
        for width, eager in ((64, True), (3, False), (1, True)):
            vis.engine.mode = "batched"
            vis.engine.batch_width = width
            vis.engine.eager_visibility = eager
            r1 = _run(vis, V, T, conc, n, return_parts=True)
            prof = _visacd.last_profile()
            print("  width", width, "eager", eager, "rounds", prof.get("batched.rounds"),
                  "split", prof.get("batched.parts_split"), flush=True)
            assert r1.num_parts == r0.num_parts and r1.concavity == r0.concavity, (width, eager)
            assert _parts(r1) == _parts(r0), (width, eager)
    finally:
        vis.engine.mode = "plain"
        vis.engine.batch_width = 64
        vis.engine.eager_visibility = False
        vis.config.return_parts = False
    assert r0.num_parts >= 3
    vol = _volume(_mesh(vis, V, T))
    assert abs(sum(_volume(p) for p in r0.parts) - vol) < 1e-6 * abs(vol) + 1e-9


def test_learned_proposal_with_full_sweep_matches_exact():
    """With an audit ratio no ranking can meet, every ranked part falls back to
    the full sweep, so the learned mode must cut exactly like the exact one."""
    import visacd as vis
    from visacd_amd import _visacd

    V, T = wl.perturbed_icosphere(4, 0.4, 5)
    print("process learned vs exact ico4", flush=True)
    vis.engine.planes = "sphere"
    vis.engine.num_planes = 1024
    try:
        for mode in ("plain", "batched"):
            vis.engine.mode = mode
            vis.engine.proposal = "exact"
            r0 = _run(vis, V, T, 0.03, 6, return_parts=True)
            vis.engine.proposal = "learned"
            vis.engine.proposal_audit_ratio = 1e30
            r1 = _run(vis, V, T, 0.03, 6, return_parts=True)
            prof = _visacd.last_profile()
            assert prof["proposal.parts"] == prof["proposal.fallbacks"] >= 1, mode
            assert r1.num_parts == r0.num_parts and r1.concavity == r0.concavity, mode
            assert _parts(r1) == _parts(r0), mode
    finally:
        vis.engine.proposal = "exact"
        vis.engine.proposal_audit_ratio = 0.98
        vis.engine.mode = "plain"
        vis.engine.planes = "bisector"
        vis.engine.num_planes = 2500
        vis.config.return_parts = False


def test_learned_proposal_with_trained_model_is_deterministic():
    """The shipped model ranks the planes; 512 of 4096 are scored exactly per
    ranked cut, the result repeats, and the parts still partition the volume."""
    import visacd as vis
    from visacd_amd import _visacd

    V, T = wl.perturbed_icosphere(4, 0.4, 6)
    print("process learned ico4", flush=True)
    vis.engine.planes = "sphere"
    vis.engine.num_planes = 4096
    vis.engine.proposal = "learned"
    try:
        r1 = _run(vis, V, T, 0.03, 6, return_parts=True)
        prof = _visacd.last_profile()
        r2 = _run(vis, V, T, 0.03, 6, return_parts=True)
    finally:
        vis.engine.proposal = "exact"
        vis.engine.planes = "bisector"
        vis.engine.num_planes = 2500
        vis.config.return_parts = False
    assert prof["proposal.parts"] >= 1
    ranked = prof["proposal.parts"]
    fb = prof.get("proposal.fallbacks", 0)
    # ranked cuts score top 256 + audit 256 (+ flat planes), fallbacks all 4096
    if fb < ranked:
        assert prof["scoring.planes"] < prof["proposal.ranked_planes"] + fb * 4096
    assert r1.num_parts == r2.num_parts >= 2 and r1.concavity == r2.concavity
    assert _parts(r1) == _parts(r2)
    vol = _volume(_mesh(vis, V, T))
    assert abs(sum(_volume(p) for p in r1.parts) - vol) < 1e-6 * abs(vol) + 1e-9


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_process_c1_matches_cpu_oracle_loop(mode):
    """The torus workload (1,152 vertices, every pair tested, 256 sphere
    planes per cut, concavity 0.04, 8 parts, seed 42) decomposed by
    visacd.process must give exactly the parts and final concavity of the
    CPU restatement of the loop whose visibility and scores come from the
    FP64 oracle (oracle/decompose.py); host distances on both sides."""
    import visacd as vis
    from oracle import decompose as D

    W = wl.torus_c1()
    print("c1 process vs cpu loop", mode, flush=True)
    ref_parts, ref_h = D.decompose(W.vertices, W.triangles, 0.04, 8, seed=42, num_planes=256,
                                   planes="sphere")
    vis.engine.mode = mode
    vis.engine.planes = "sphere"
    vis.engine.num_planes = 256
    vis.engine.gpu_distance = False
    try:
        r = _run(vis, W.vertices.astype(np.float64), W.triangles, 0.04, 8, return_parts=True)
    finally:
        vis.engine.mode = "plain"
        vis.engine.planes = "bisector"
        vis.engine.num_planes = 2500
        vis.engine.gpu_distance = True
        vis.config.return_parts = False
    print("  parts", r.num_parts, len(ref_parts), "concavity", r.concavity, ref_h, flush=True)
    assert r.num_parts == len(ref_parts) >= 4
    for p, (rv, rt) in zip(r.parts, ref_parts):
        assert np.array_equal(np.array([list(v) for v in p.vertices]), rv)
        assert np.array_equal(np.array([list(t) for t in p.triangles]), rt)
    assert r.concavity == ref_h


def _dumbbell():
    """Two 3x3x3 polycube boxes joined by a 1x1x2 bar along x."""
    cells = [(x, y, z) for x in range(3) for y in range(3) for z in range(3)]
    cells += [(x, 1, 1) for x in (3, 4)]
    cells += [(x, y, z) for x in range(5, 8) for y in range(3) for z in range(3)]
    return meshes.polycube(cells)


@pytest.mark.parametrize("shape,n", [("lshape", 2), ("dumbbell", 3)])
def test_process_closed_form_shapes(shape, n):
    """Shapes whose decomposition is known in form (the reference's default
    bisector planes, seed 42, host distances): the L-shape splits into two
    parts at its notch, one of them exactly convex; the dumbbell splits into
    its two boxes and the bar between them, the bar exactly convex and the
    two boxes alike.  The cuts are bisector planes of visibility edges, so
    they lie near, not on, the notch and the bar's ends (points on a cut
    plane are in the 1e-6 dead zone and their edges would not count).  Both
    equal the CPU restatement of the loop on the FP64 oracle, part for part."""
    import visacd as vis
    from oracle import decompose as D

    V, T = meshes.l_shape() if shape == "lshape" else _dumbbell()
    V = V.astype(np.float64)
    print("closed form", shape, n, flush=True)
    ref_parts, ref_h = D.decompose(V, T, 0.001, n, seed=42)
    vis.engine.gpu_distance = False
    try:
        r = _run(vis, V, T, 0.001, n, return_parts=True)
    finally:
        vis.engine.gpu_distance = True
        vis.config.return_parts = False
    assert r.num_parts == len(ref_parts) == n
    for p, (rv, rt) in zip(r.parts, ref_parts):
        assert np.array_equal(np.array([list(v) for v in p.vertices]), rv)
        assert np.array_equal(np.array([list(t) for t in p.triangles]), rt)
    assert r.concavity == ref_h
    from visacd_amd import _visacd as H

    vols, hulls, xr = [], [], []
    for p in r.parts:
        vols.append(_volume(p))
        hulls.append(H.mesh_volume(H.convex_hull(p)))
        x = np.array([list(v) for v in p.vertices])[:, 0]
        xr.append((x.min(), x.max()))
    total = _volume(_mesh(vis, V, T))
    assert abs(sum(vols) - total) < 1e-9 * abs(total) + 1e-12
    convex = [abs(h - v) < 1e-9 * abs(v) + 1e-12 for v, h in zip(vols, hulls)]
    print("  x ranges", xr, "convex", convex, "concavity", r.concavity, flush=True)
    if shape == "lshape":
        # the L spans x in [0, 0.75]; its notch is at x = 0.25 (or y = 0.25)
        assert sum(convex) >= 1
        cut = [b for a, b in xr if b < 0.7] + [a for a, b in xr if a > 0.05]
        assert cut and all(abs(c - 0.25) < 0.02 for c in cut)
    else:
        # boxes at x in [0, 0.75] and [1.25, 2.0], the bar between
        order = np.argsort([a for a, b in xr])
        (a0, b0), (a1, b1), (a2, b2) = (xr[i] for i in order)
        assert convex[order[1]]  # the bar
        assert abs(b0 - 0.75) < 0.05 and abs(a2 - 1.25) < 0.05 and a1 == b0 and b1 == a2
        assert abs(vols[order[0]] - vols[order[2]]) < 1e-3 * abs(vols[order[0]])


def test_process_cow_reference_sample():
    """The reference's own sample input (data/cow.obj, via tests/golden/cow.npz)
    with its CLI defaults (concavity 0.04, 40 parts): plain and batched give
    the same decomposition, the parts partition the volume and the
    concavity falls below the undivided mesh's."""
    import os

    import visacd as vis

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "cow.npz"))
    V, T = d["vertices"], d["triangles"]
    print("process cow", V.shape, T.shape, flush=True)
    try:
        vis.engine.mode = "plain"
        r0 = _run(vis, V, T, 0.04, 40, return_parts=True)
        vis.engine.mode = "batched"
        r1 = _run(vis, V, T, 0.04, 40, return_parts=True)
        vis.engine.mode = "plain"
        r_one = _run(vis, V, T, 0.04, 1)
    finally:
        vis.engine.mode = "plain"
        vis.config.return_parts = False
    print("  parts", r0.num_parts, "concavity", r0.concavity, r_one.concavity, flush=True)
    assert r0.num_parts >= 2 and r1.num_parts == r0.num_parts and r1.concavity == r0.concavity
    assert _parts(r1) == _parts(r0)
    vol = _volume(_mesh(vis, V, T))
    assert abs(sum(_volume(p) for p in r0.parts) - vol) < 1e-6 * abs(vol)
    assert r0.concavity < r_one.concavity


def test_process_cage_mode_plain_and_batched_agree():
    """engine.visibility = "cage": the reference's rule on the engine's cages
    (the input's at distance 0.03, each new part's at 0.02, on a 1/40 grid).
    Plain and batched give the same decomposition, whose parts partition the
    volume."""
    import visacd as vis
    from visacd_amd import _visacd

    V, T = wl.perturbed_icosphere(3, 0.4, 1)
    print("process cage mode", flush=True)
    try:
        vis.engine.visibility = "cage"
        vis.engine.mode = "plain"
        r0 = _run(vis, V, T, 0.02, 6, return_parts=True)
        prof = _visacd.last_profile()
        vis.engine.mode = "batched"
        r1 = _run(vis, V, T, 0.02, 6, return_parts=True)
    finally:
        vis.engine.visibility = "parity"
        vis.engine.mode = "plain"
        vis.config.return_parts = False
    print("  parts", r0.num_parts, r0.concavity, "cage builds", prof.get("visibility.cage"), flush=True)
    assert prof["visibility.cage"][1] >= 2
    assert r0.num_parts >= 2 and r1.num_parts == r0.num_parts and r1.concavity == r0.concavity
    assert _parts(r1) == _parts(r0)
    vol = _volume(_mesh(vis, V, T))
    assert abs(sum(_volume(p) for p in r0.parts) - vol) < 1e-6 * abs(vol)


def test_process_remeshes_an_open_input():
    """An input with a hole (not closed) is remeshed first (engine.remesh
    "auto", the reference's scale 30 and level 0.55/30) and decomposed; with
    "always" a closed input is remeshed too; the parts partition the
    remeshed volume, slightly larger than the input's."""
    import visacd as vis
    from visacd_amd import _visacd

    V, T = wl.perturbed_icosphere(3, 0.3, 2)
    open_T = np.ascontiguousarray(T[4:])
    print("process open input", flush=True)
    try:
        r = _run(vis, V, open_T, 0.03, 5, return_parts=True)
        prof = _visacd.last_profile()
        vis.engine.remesh = "always"
        r2 = _run(vis, V, T, 0.03, 5, return_parts=True)
        prof2 = _visacd.last_profile()
    finally:
        vis.engine.remesh = "auto"
        vis.config.return_parts = False
    assert "remesh" in prof and "remesh" in prof2
    vol = _volume(_mesh(vis, V, T))
    for res in (r, r2):
        total = sum(_volume(p) for p in res.parts)
        print("  parts", res.num_parts, "volume", total, vol, flush=True)
        assert res.num_parts >= 2
        assert vol < total < 1.15 * vol
