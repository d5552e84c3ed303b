"""Batched frontier: many parts in one visibility launch equal per-part results."""
import numpy as np
import pytest

import meshes
import oracle
from visacd_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _point_tri_dist(p, a, b, c):
    """Brute force: distance to the plane projection if inside, else to the edges."""
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    bc = [np.dot(np.cross(b - a, q - a), n), np.dot(np.cross(c - b, q - b), n),
          np.dot(np.cross(a - c, q - c), n)]
    if min(bc) >= 0:
        return abs(np.dot(p - a, n))
    best = np.inf
    for u, v in ((a, b), (b, c), (c, a)):
        t = np.clip(np.dot(p - u, v - u) / np.dot(v - u, v - u), 0, 1)
        best = min(best, np.linalg.norm(p - (u + t * (v - u))))
    return best


def test_max_distance_matches_brute_force(ctx):
    V, T = wl.perturbed_icosphere(2, 0.3, 31)
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1.5, 1.5, size=(200, 3)).astype(np.float32)
    print("max distance", flush=True)
    b = ctx.bvh(V, T)
    got = ctx.max_distance(pts, b)
    V64 = V.astype(np.float64)
    ref = max(min(_point_tri_dist(p.astype(np.float64), V64[t[0]], V64[t[1]], V64[t[2]])
                  for t in T) for p in pts)
    assert abs(got - ref) <= 1e-5 * max(1.0, ref)
    assert ctx.max_distance(V, b) < 1e-6  # mesh vertices lie on the mesh


def test_max_distance_batch_matches_bvh_form_and_brute_force(ctx):
    """The batched brute-force distance equals the BVH walk's value for every
    query of a mixed batch: meshes of several sizes (more than one triangle
    chunk), points inside, outside and on the mesh, an empty query and one
    with no triangles."""
    rng = np.random.default_rng(5)
    queries, want = [], []
    for s, (level, m) in enumerate(((1, 40), (3, 700), (4, 300), (2, 1))):
        V, T = wl.perturbed_icosphere(level, 0.3, 40 + s)
        pts = rng.uniform(-1.4, 1.4, size=(m, 3)).astype(np.float32)
        pts[: m // 4] = V[rng.integers(0, V.shape[0], m // 4)]  # some on the mesh
        print("distance batch query", s, "level", level, "points", m, flush=True)
        b = ctx.bvh(V, T)
        want.append(ctx.max_distance(pts, b))
        b.close()
        queries.append((pts, V, T))
    V, T = wl.perturbed_icosphere(1, 0.2, 3)
    queries.append((np.zeros((0, 3), np.float32), V, T))  # no points: 0
    queries.append((V[:5], V, np.zeros((0, 3), np.uint32)))  # no triangles: inf
    got = ctx.max_distance_batch(queries)
    # same minimum over the same float32 triangle distances (up to the
    # compiler's contraction choices in the two kernels)
    assert np.allclose(got[:4], np.array(want, dtype=np.float32), rtol=1e-6, atol=1e-7), (got, want)
    assert got[4] == 0.0 and np.isinf(got[5])
    # brute force in float64 on the small query
    pts, V, T = queries[0]
    V64 = V.astype(np.float64)
    ref = max(min(_point_tri_dist(p.astype(np.float64), V64[t[0]], V64[t[1]], V64[t[2]])
                  for t in T) for p in pts)
    assert abs(got[0] - ref) <= 1e-5 * max(1.0, ref)
    with pytest.raises(Exception):
        ctx.max_distance_batch([(pts, V, T + 1000)])  # vertex id out of range


def _dist_to_mesh(P, V, T):
    """float64 distance of every point of P to the triangle mesh (V, T)."""
    V = V.astype(np.float64)
    A, B, C = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    out = np.empty(P.shape[0])
    for i0 in range(0, P.shape[0], 256):
        Q = P[i0:i0 + 256].astype(np.float64)[:, None, :]
        n = np.cross(B - A, C - A)
        n = n / np.linalg.norm(n, axis=1, keepdims=True)
        h = np.einsum("ptk,tk->pt", Q - A, n)
        X = Q - h[..., None] * n
        ins = np.ones(h.shape, bool)
        for U, W in ((A, B), (B, C), (C, A)):
            ins &= np.einsum("ptk,tk->pt", np.cross(np.broadcast_to(W - U, X.shape), X - U), n) >= 0
        d = np.where(ins, np.abs(h), np.inf)
        for U, W in ((A, B), (B, C), (C, A)):
            e = W - U
            t = np.clip(np.einsum("ptk,tk->pt", Q - U, e) / np.einsum("tk,tk->t", e, e), 0, 1)
            d = np.minimum(d, np.linalg.norm(Q - (U + t[..., None] * e), axis=-1))
        out[i0:i0 + 256] = d.min(1)
    return out


def test_max_distance_batch_part_against_hull(ctx):
    """The concavity queries: surface samples of a dented mesh against its
    convex hull and hull samples against the mesh, where most samples lie
    near the other surface and the seed's lower bound stops them early; the
    maximum equals the float64 brute force."""
    from scipy.spatial import ConvexHull

    rng = np.random.default_rng(12)
    queries, refs = [], []
    for seed in (3, 4):
        V, T = wl.perturbed_icosphere(3, 0.45, seed)
        H = ConvexHull(V.astype(np.float64))
        hv = V[H.vertices]
        remap = np.full(V.shape[0], -1)
        remap[H.vertices] = np.arange(H.vertices.size)
        hT = remap[H.simplices].astype(np.uint32)

        def samples(VV, TT, m):
            t = rng.integers(0, TT.shape[0], m)
            r1, r2 = np.sqrt(rng.random(m))[:, None], rng.random(m)[:, None]
            a, b, c = VV[TT[t, 0]], VV[TT[t, 1]], VV[TT[t, 2]]
            return (a * (1 - r1) + b * (r1 * (1 - r2)) + c * (r1 * r2)).astype(np.float32)

        sa, sb = samples(V, T, 3000), samples(hv, hT, 3000)
        queries += [(sa, hv, hT), (sb, V, T)]
        refs += [_dist_to_mesh(sa, hv, hT).max(), _dist_to_mesh(sb, V, T).max()]
    print("part/hull distance queries", len(queries), flush=True)
    got = ctx.max_distance_batch(queries)
    np.testing.assert_allclose(got, np.array(refs), rtol=1e-5, atol=1e-6)
    assert (np.array(refs) > 1e-3).all()


def test_plane_extents_batch_equal_host_float64(ctx):
    """Extents along unit normals equal numpy's float64 ((n0 p0 + n1 p1) + n2 p2)
    bit for bit, for queries of several sizes (more than one LDS tile of
    points, more than one block of normals), an empty point set included."""
    rng = np.random.default_rng(9)
    queries = []
    for P, n in ((5, 3), (300, 2500), (1000, 700), (7, 0)):
        N = rng.normal(size=(P, 3))
        N /= np.linalg.norm(N, axis=1, keepdims=True)
        queries.append((N, rng.uniform(-1, 1, size=(n, 3))))
    print("extent batch", [(q[0].shape[0], q[1].shape[0]) for q in queries], flush=True)
    got = ctx.plane_extents_batch(queries)
    for (N, pts), g in zip(queries, got):
        if pts.shape[0] == 0:
            assert np.all(g[:, 0] == 1e300) and np.all(g[:, 1] == -1e300)
            continue
        d = N[:, 0, None] * pts[None, :, 0] + N[:, 1, None] * pts[None, :, 1]
        d = d + N[:, 2, None] * pts[None, :, 2]
        assert np.array_equal(g[:, 0], d.min(axis=1)) and np.array_equal(g[:, 1], d.max(axis=1))


def test_batch_matches_oracle_per_part(ctx):
    import torch

    dev = torch.device("cuda:0")
    cases = []
    for s in range(5):
        V, T = wl.perturbed_icosphere(2 + s % 2, 0.2 + 0.05 * s, 100 + s)
        cases.append((V, T, None))
    Vl, Tl = meshes.l_shape()
    cases.append((Vl, Tl, None))
    # a part with a mask, one with a single point (no pairs), an open mesh
    Vm, Tm = wl.perturbed_icosphere(2, 0.3, 7)
    rng = np.random.default_rng(0)
    cases.append((Vm, Tm, (rng.uniform(size=Vm.shape[0]) < 0.2).astype(np.uint8)))
    cases.append((Vm[:1].copy(), Tm, None))
    Vo, To = wl.perturbed_icosphere(2, 0.3, 8)
    cases.append((Vo, np.ascontiguousarray(To[2:]), None))
    print("batch", len(cases), "parts", flush=True)
    keep = []
    parts = []
    for V, T, m in cases:
        b = ctx.bvh(V if V.shape[0] > 1 else Vm, T)
        d_pts = torch.from_numpy(np.ascontiguousarray(V)).to(dev)
        d_m = torch.from_numpy(m).to(dev) if m is not None else None
        keep.append((b, d_pts, d_m))
        parts.append((d_pts, V.shape[0], d_m, b, None))
    cap = sum(V.shape[0] * (V.shape[0] - 1) // 2 for V, _, _ in cases)
    d_pairs = torch.zeros((cap, 2), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    d_off = torch.zeros(len(parts) + 1, dtype=torch.int64, device=dev)
    ref = [oracle.visibility(V, V if V.shape[0] > 1 else Vm, T, mask=m)[0] for V, T, m in cases]
    from visacd_amd import native

    # default (these parts are too small for clipped walks), then every
    # part's walks clipped at both ends where the fans allow it
    for clip in (1 << 25, 1):
        ctx.set_option(native.VACD_OPT_CLIP_ENDS, clip)
        try:
            ctx.visibility_batch_d(parts, d_pairs, cap, d_count, d_off)
            torch.cuda.synchronize()
        finally:
            ctx.set_option(native.VACD_OPT_CLIP_ENDS, 1 << 25)
        if clip == 1:
            assert ctx.visibility_counters()["clipped_walks"] > 0
        pairs = d_pairs.cpu().numpy().astype(np.uint32)
        off = d_off.cpu().numpy()
        assert off[-1] == int(d_count.item())
        assert (np.diff(off) >= 0).all()
        for t in range(len(cases)):
            np.testing.assert_array_equal(pairs[off[t]:off[t + 1]], ref[t],
                                          err_msg=f"part {t}, clip {clip}")


def test_score_batch_matches_oracle_per_part(ctx):
    """Parts of very different sizes (none, one plane, no edges, thousands)
    in one batched scoring call: each part's Q32.32 sums and first-max argmax
    equal the oracle's."""
    rng = np.random.default_rng(11)
    sizes = [(0, 50, 40), (1, 30, 25), (700, 200, 0), (3000, 900, 5000), (5000, 400, 2500),
             (2049, 300, 777)]
    parts = []
    for P, n, E in sizes:
        nrm = rng.normal(size=(P, 3))
        nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-12)
        planes = np.concatenate([nrm, rng.uniform(-0.5, 0.5, (P, 1))], axis=1).astype(np.float32)
        pts = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        e = np.sort(rng.integers(0, n, size=(E, 2)), axis=1).astype(np.uint32)
        e = e[e[:, 0] != e[:, 1]]
        parts.append((planes, pts, e[np.lexsort((e[:, 1], e[:, 0]))]))
    print("score batch", [(p[0].shape[0], p[2].shape[0]) for p in parts], flush=True)
    qs, am = ctx.score_planes_batch(parts)
    for k, (pl, pt, ed) in enumerate(parts):
        oq = oracle.score_q32(pl, pt, ed)
        np.testing.assert_array_equal(qs[k], oq, err_msg=f"part {k}")
        assert am[k] == (oracle.argmax_i64(oq) if pl.shape[0] else -1), k



def test_batched_c3_step_matches_oracle_and_graph_replays(ctx):
    """The 64-part C3 step (perturbed icospheres, levels 2-5, at most 2,048
    points and 1,024 planes each): one batched visibility call and one
    batched scoring call on its edges.  Per part, 1,500 random pairs and 500
    edges are decided as the oracle decides them, and every plane's Q32.32
    score and the first maximum equal the oracle's.  The step captured in a
    HIP graph and replayed three times (eager work between replays) gives the
    same edges, scores and argmax each time."""
    import torch

    Ws = wl.batched_c3()
    dev = torch.device("cuda:0")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    bvhs, pts, planes = [], [], []
    for W in Ws:
        bvhs.append(ctx.bvh(torch.from_numpy(W.vertices).to(dev),
                            torch.from_numpy(W.triangles.astype(np.int32)).to(dev)))
        pts.append(torch.from_numpy(W.points).to(dev))
        planes.append(torch.from_numpy(W.planes).to(dev))
    ns = [W.n_points for W in Ws]
    S = sum(n * (n - 1) // 2 for n in ns)
    print("c3 parts", len(Ws), "segments", S, flush=True)
    d_pairs = torch.empty((S, 2), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    d_off = torch.zeros(len(Ws) + 1, dtype=torch.int64, device=dev)
    vis_parts = [(pts[i], ns[i], None, bvhs[i], None) for i in range(len(Ws))]
    ctx.visibility_batch_d(vis_parts, d_pairs, S, d_count, d_off)
    torch.cuda.synchronize()
    off = d_off.cpu().numpy()
    pairs = d_pairs.cpu().numpy().astype(np.uint32)
    score_parts = [(planes[i], Ws[i].planes.shape[0], pts[i], ns[i],
                    d_pairs.data_ptr() + 8 * int(off[i]), int(off[i + 1] - off[i]))
                   for i in range(len(Ws))]
    P = sum(W.planes.shape[0] for W in Ws)
    d_q = torch.zeros(P, dtype=torch.int64, device=dev)
    d_am = torch.zeros(len(Ws), dtype=torch.int32, device=dev)
    ctx.score_planes_batch_d(score_parts, d_q, d_am)
    torch.cuda.synchronize()
    q = d_q.cpu().numpy()
    am = d_am.cpu().numpy()
    rng = np.random.default_rng(9)
    po = 0
    for t, W in enumerate(Ws):
        e = pairs[off[t]:off[t + 1]]
        n = ns[t]
        k = e[:, 0].astype(np.int64) * (2 * n - e[:, 0].astype(np.int64) - 1) // 2 + \
            e[:, 1].astype(np.int64) - e[:, 0].astype(np.int64) - 1
        assert (np.diff(k) > 0).all(), t
        Sp = n * (n - 1) // 2
        ks = np.unique(np.concatenate([rng.integers(0, Sp, 1500),
                                       k[rng.integers(0, len(k), 500)] if len(k) else []]).astype(np.int64))
        det = oracle.visibility_sample(W.points, W.vertices, W.triangles, ks)
        np.testing.assert_array_equal(np.isin(ks, k), det == 1, err_msg=f"part {t}")
        Pp = W.planes.shape[0]
        oq = oracle.score_q32(W.planes, W.points, e)
        np.testing.assert_array_equal(q[po:po + Pp], oq, err_msg=f"part {t}")
        assert am[t] == oracle.argmax_i64(oq), t
        po += Pp

    def step():
        ctx.visibility_batch_d(vis_parts, d_pairs, S, d_count, d_off)
        ctx.score_planes_batch_d(score_parts, d_q, d_am)

    step()  # warm-up: buffers and staging regions sized before the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=s):
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        step()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    total = int(off[-1])
    for r in range(3):
        d_count.zero_()
        d_q.zero_()
        d_am.fill_(-7)
        d_pairs[:total].fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        assert int(d_count.item()) == total, r
        np.testing.assert_array_equal(d_off.cpu().numpy(), off)
        np.testing.assert_array_equal(d_pairs[:total].cpu().numpy().astype(np.uint32), pairs[:total])
        np.testing.assert_array_equal(d_q.cpu().numpy(), q)
        np.testing.assert_array_equal(d_am.cpu().numpy(), am)
        _ = torch.ones(4096, device=dev).sum().item()  # eager work between replays
    for bb in bvhs:
        bb.close()
