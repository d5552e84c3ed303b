"""Allocation-free scoring entry point (vacd_score_planes_noalloc_d).

It must give exactly what vacd_score_planes_d gives on the same device
inputs, on consecutive calls that share the library's fixed scratch array
(the second call smaller than the first, so stale sums or records of the
first call would show), and reject sizes beyond its stated maximum.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(ctx, torch, n_pts, n_planes, seed):
    from visacd_amd import workloads as wl

    V, T = wl.perturbed_icosphere(3, 0.3, seed)  # 642 vertices, 1,280 triangles
    bvh = ctx.bvh(V, T)
    P_pts = np.ascontiguousarray(V[:n_pts], dtype=np.float32)
    edges = ctx.visibility(P_pts, bvh)
    bvh.close()
    planes = wl.sphere_planes(P_pts, n_planes, seed + 1).astype(np.float32)
    dev = torch.device("cuda:0")
    d_planes = torch.from_numpy(np.ascontiguousarray(planes)).to(dev)
    d_pts = torch.from_numpy(P_pts).to(dev)
    d_edges = torch.from_numpy(np.ascontiguousarray(edges, dtype=np.uint32).view(np.int32)).to(dev)
    return planes, P_pts, edges, d_planes, d_pts, d_edges


def _run(ctx, torch, fn, d_planes, P, d_pts, n, d_edges, E):
    dev = torch.device("cuda:0")
    q = torch.full((max(P, 1),), -7, dtype=torch.int64, device=dev)
    f = torch.full((max(P, 1),), -7.0, dtype=torch.float32, device=dev)
    am = torch.full((1,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()  # the context enqueues on its own stream
    fn(d_planes, P, d_pts, n, d_edges, E, q, f, am)
    ctx.sync()
    return q[:P].cpu().numpy(), f[:P].cpu().numpy(), int(am.cpu().item())


def test_noalloc_scoring_matches_pooled_entry_point(ctx):
    import torch

    import oracle

    results = []
    # the second call is smaller (fewer planes, fewer edges) than the first
    for n_pts, n_planes, seed in ((642, 3000, 11), (300, 700, 12)):
        planes, V, edges, d_planes, d_pts, d_edges = _case(ctx, torch, n_pts, n_planes, seed)
        P, n, E = planes.shape[0], V.shape[0], edges.shape[0]
        assert E > 500 and P == n_planes
        print(f"noalloc case n={n} P={P} E={E}", flush=True)
        ref = _run(ctx, torch, ctx.score_planes_d, d_planes, P, d_pts, n, d_edges, E)
        got = _run(ctx, torch, ctx.score_planes_noalloc_d, d_planes, P, d_pts, n, d_edges, E)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]
        oq = oracle.score_q32(planes, V, edges)
        assert np.array_equal(got[0], oq) and got[2] == oracle.argmax_i64(oq)
        results.append(got)
    assert results[0][0].shape != results[1][0].shape


def test_noalloc_scoring_empty_inputs_and_limits(ctx):
    import torch

    from visacd_amd import native

    dev = torch.device("cuda:0")
    am = torch.full((1,), 5, dtype=torch.int32, device=dev)
    pts = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ctx.score_planes_noalloc_d(None, 0, pts, 4, None, 0, None, None, am)
    ctx.sync()
    assert int(am.cpu().item()) == -1
    # planes but no edges: every score 0, argmax the first plane
    planes = torch.zeros((3, 4), dtype=torch.float32, device=dev)
    q = torch.full((3,), 9, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    ctx.score_planes_noalloc_d(planes, 3, pts, 4, None, 0, q, None, am)
    ctx.sync()
    assert q.cpu().tolist() == [0, 0, 0] and int(am.cpu().item()) == 0
    # beyond the maximum sizes: refused on the host before any launch
    with pytest.raises(native.VacdError) as ei:
        ctx.score_planes_noalloc_d(planes, native.VACD_NOALLOC_MAX_PLANES + 1, pts, 4, None, 0,
                                   q, None, am)
    assert ei.value.code == native.VACD_ERR_LIMIT
    edges = torch.zeros((2,), dtype=torch.int32, device=dev)
    with pytest.raises(native.VacdError) as ei:
        ctx.score_planes_noalloc_d(planes, 3, pts, 4, edges, native.VACD_NOALLOC_MAX_EDGES + 1,
                                   q, None, am)
    assert ei.value.code == native.VACD_ERR_LIMIT
