"""Benchmark of the visibility-edge engine on one MI355X.

Headline workload (default, ``--workload c5``): perturbed icosphere,
subdivision level 7 with four Gaussian dents (163,842 v / 327,680 t,
amplitude 0.5, seed 42), seeded 16,384-vertex subsample, so one step tests
all 134,209,536 vertex-pair segments against the 327,680-triangle BVH in
midpoint-parity mode.  Inputs are resident in HBM before timing.

A "step" is one visibility pass (traversal + exact fallback + k-ordered
compaction).  ``value`` = segments per second of that step, in
Gsegments/s.  Beside it: the LBVH build of the same mesh, plane scoring of
its 16,384 candidate planes on the pass's edges, the same pass on the C2
workload (icosphere L6, 4,096 points) with its walk-variant sweeps, the
proposal model, the 64-part batched step (C3), the decompositions of C2 and
C5 in plain, batched+exact and batched+learned form (``acd_speedup``), and
the CPU baselines: the FP64 oracle's visibility and scoring rates on
samples of the same workload and the C1 decomposition on the CPU oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c2|c1]
                       [--no-cpu] [--cpu-seconds S]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from visacd_amd import workloads as wl

    if name == "c2":
        return wl.icosphere_c2()
    if name == "c5":
        return wl.icosphere_c5()
    if name == "c1":
        return wl.torus_c1()
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(W, seconds):
    import numpy as np

    import oracle

    rng = np.random.default_rng(123)
    # rate estimate from two calls (the first also pays the CPU BVH build)
    rate = 0.0
    for m in (2000, 100000):
        ks = rng.integers(0, W.n_segments, m)
        t0 = time.perf_counter()
        oracle.visibility_sample(W.points, W.vertices, W.triangles, ks)
        rate = m / max(time.perf_counter() - t0, 1e-6)
    m = int(min(W.n_segments, max(2000, rate * seconds)))
    ks = rng.integers(0, W.n_segments, m)
    t0 = time.perf_counter()
    oracle.visibility_sample(W.points, W.vertices, W.triangles, ks)
    dt = time.perf_counter() - t0
    return {"value": m / dt / 1e9, "unit": "Gsegments/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"{m} uniformly drawn pairs of the same workload, FP64 oracle "
                      f"(exact predicates, CPU BVH), {dt:.1f} s"}


def window_rule(W, m=200000):
    """The reference's ray window (float32 ray from p over t in [0.001, len],
    lib/ptx/ray_segments.cu:77-111) beside exact contact with the self mesh
    on a sample of the workload's pairs: counts of agreement and of each
    cause of disagreement (the oracle's restatement, a checker)."""
    import numpy as np

    import oracle

    rng = np.random.default_rng(7)
    ks = np.unique(rng.integers(0, W.n_segments, m))
    tau = 1e-6 * float((W.vertices.max(0) - W.vertices.min(0)).max())
    t0 = time.perf_counter()
    bits = oracle.window_hits(W.points, W.vertices, W.triangles, tau, ks=ks)
    r = oracle.window_classes(bits)
    r["tau"] = tau
    r["sample"] = f"{ks.size} uniformly drawn pairs, {time.perf_counter() - t0:.1f} s"
    return r


def acd_wall(W, mode, proposal="exact"):
    """Full decomposition of the workload's mesh through visacd.process:
    mode "plain" (the reference's loop) or "batched" (frontier rounds);
    proposal "exact" (every candidate plane scored) or "learned" (the
    proposal model ranks them, the top 256 and 256 audit planes are scored)."""
    import numpy as np

    import visacd as vis
    from visacd_amd import _visacd

    m = vis.Mesh()
    m.vertices = vis.VecArray3d(W.vertices.astype(np.float64).tolist())
    m.triangles = vis.make_vecarray3i(W.triangles.astype(np.int32))
    vis.engine.verbose = False
    vis.engine.mode = mode
    vis.engine.max_points = W.n_points
    vis.engine.planes = "sphere"
    vis.engine.num_planes = int(W.planes.shape[0])
    vis.engine.proposal = proposal
    vis.config.return_parts = False
    vis.set_seed(int(W.meta.get("seed", 42)))
    conc = float(W.meta.get("concavity", 0.04))
    nparts = int(W.meta.get("num_parts", 32))
    t0 = time.perf_counter()
    r = vis.process(m, conc, nparts)
    dt = time.perf_counter() - t0
    raw = _visacd.last_profile()
    prof = {k: round(v[0], 4) for k, v in raw.items() if isinstance(v, tuple)}
    counters = {k: v for k, v in raw.items() if not isinstance(v, tuple)}
    vis.engine.mode = "plain"
    vis.engine.proposal = "exact"
    return {"workload": W.name, "mode": mode, "proposal": proposal, "wall_s": dt,
            "parts": int(r.num_parts), "counters": counters,
            "concavity": float(r.concavity), "concavity_threshold": conc, "max_parts": nparts,
            "points_per_part": W.n_points, "planes_per_cut": int(W.planes.shape[0]),
            "phase_seconds": prof}


def traffic_kernel(name):
    """Kernels of one default visibility pass whose HBM traffic is reported:
    the fan pass (k_visibility) and the walks of its queue (k_walk_front /
    k_walk_split instances other than the exact redo, whose last template
    argument is true)."""
    import re

    m = re.search(r"(k_visibility|k_walk_front|k_walk_split)<([^>]*)>", name)
    if not m:
        return None
    if m.group(1) != "k_visibility" and m.group(2).replace(" ", "").endswith("true"):
        return None
    return m.group(0)


def pmc_traffic(workload):
    """HBM bytes of one default visibility pass (fan pass + queue walk), from
    two rocprofv3 --pmc runs of this script in probe mode (FETCH_SIZE, then
    WRITE_SIZE), each scaled by what the counter reads for a 256 MiB copy.
    Returns (bytes, detail) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    vals = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(td, counter)
            cmd = [prof, "--pmc", counter, "-f", "csv", "-d", d, "-o", "pmc", "--",
                   sys.executable, "-u", os.path.abspath(__file__), "--pmc-probe",
                   "--workload", workload]
            env = dict(os.environ, TMPDIR="/tmp")
            proc = subprocess.Popen(cmd, env=env, stdout=subprocess.DEVNULL,
                                    stderr=subprocess.DEVNULL, start_new_session=True)
            try:
                rc = proc.wait(timeout=90)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, 9)  # the profiler and the probe it started
                proc.wait()
                return None, f"rocprofv3 --pmc {counter} timed out"
            # the profiler may exit non-zero after the program ended with its
            # results written; a missing counter file is caught below
            if rc != 0:
                log(f"rocprofv3 --pmc {counter} exited with {rc}")
            per = {"cal": 0.0, "pass": 0.0}
            for root, _, files in os.walk(d):
                for fn in files:
                    if not fn.endswith("counter_collection.csv"):
                        continue
                    rows = [r for r in csv.DictReader(open(os.path.join(root, fn)))
                            if r["Counter_Name"] == counter]
                    # the last default pass of the probe: its last dispatch of each kernel
                    last = {}
                    for r in rows:
                        name = r["Kernel_Name"]
                        v = float(r["Counter_Value"])
                        if "copyBuffer" in name:
                            per["cal"] = max(per["cal"], v)
                        k = traffic_kernel(name)
                        if k:
                            last[k] = v
                    per["pass"] = sum(last.values())
            if per["cal"] <= 0 or per["pass"] <= 0:
                return None, f"no counter data in the {counter} run"
            vals[counter] = per["pass"] * (262144.0 / per["cal"]) * 1024.0  # KiB -> bytes
    return vals["FETCH_SIZE"] + vals["WRITE_SIZE"], vals


def pmc_probe(workload):
    """Two default visibility passes and a 256 MiB calibration copy."""
    import numpy as np
    import torch

    from visacd_amd import native

    W = make_workload(workload)
    dev = torch.device("cuda:0")
    d_V = torch.from_numpy(W.vertices).to(dev)
    d_T = torch.from_numpy(W.triangles.astype(np.int32)).to(dev)
    d_pts = torch.from_numpy(W.points).to(dev)
    ctx = native.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    bvh = ctx.bvh(d_V, d_T)
    d_pairs = torch.empty((W.n_segments, 2), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    for _ in range(2):
        ctx.visibility_d(d_pts, W.n_points, bvh, d_pairs=d_pairs, cap=W.n_segments,
                         d_count=d_count)
    cal = torch.empty(64 << 20, dtype=torch.float32, device=dev)
    cal2 = torch.empty_like(cal)
    cal2.copy_(cal)
    torch.cuda.synchronize()
    bvh.close()
    ctx.close()


def mlp_flops_per_part(E, P, D=256):
    """Multiply-adds x 2 of one part's forward pass: the shared edge layers
    (input padded to 8), the pooled projection once, and per plane the plane
    term, the 256 x 256 layer and the output dot product."""
    enc = 2 * E * (8 * D + 2 * D * D)
    pooled = 2 * D * D
    head = P * (2 * 4 * D + 2 * D * D + 2 * D)
    return enc + pooled + head


def proposal_bench(steps):
    """Forward pass of the shipped proposal model for B parts x 512 edges x
    16,384 planes (B = 1, 16, 64), timed with events on the launch stream."""
    import numpy as np
    import torch

    import visacd
    from visacd_amd import native

    dev = torch.device("cuda:0")
    ctx = native.Context(0)
    stream = torch.cuda.Stream()  # the kernels and the events on one stream
    ctx.set_stream(stream.cuda_stream)
    m = native.Mlp(ctx, seed=0)
    m.set(np.fromfile(visacd.engine.proposal_weights, dtype="<f4"))
    E, P = 512, 16384
    res = {}
    torch.cuda.synchronize()
    for B in (1, 16, 64):
        g = torch.Generator().manual_seed(B)
        feat = (torch.rand((B, E, 7), generator=g) * 2 - 1).to(dev)
        nrm = torch.randn((B, P, 3), generator=g)
        nrm = nrm / nrm.norm(dim=-1, keepdim=True)
        planes = torch.cat([nrm, torch.rand((B, P, 1), generator=g)], -1).to(dev).contiguous()
        out = torch.empty(B * P, device=dev)
        torch.cuda.synchronize()
        for _ in range(2):
            m.forward_d(B, E, feat, P, planes, out)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            m.forward_d(B, E, feat, P, planes, out)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        fl = B * mlp_flops_per_part(E, P)
        res[str(B)] = {"ms": ms, "parts_per_s": B / (ms * 1e-3),
                       "gflop": fl / 1e9, "tflops": fl / (ms * 1e-3) / 1e12}
    m.close()
    ctx.close()
    best = res["64"]
    return {"edges_per_part": E, "planes_per_part": P, "by_parts": res,
            "roofline": {"bound": "mfma", "achieved": best["tflops"],
                         "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": best["tflops"] / MFMA_BF16_PEAK_TFLOPS,
                         "flop_per_part": mlp_flops_per_part(E, P), "parts": 64}}


def c3_batched(steps):
    """Batched frontier step on 64 independent perturbed icospheres (levels
    2-5, at most 2,048 points and 1,024 planes each): one batched visibility
    launch set for all parts, then one batched scoring launch set on each
    part's edges.  The knobs are autotuned, then the step is captured into a
    HIP graph and replayed; every replay is checked against the eager step."""
    import numpy as np
    import torch

    from visacd_amd import autotune, native
    from visacd_amd import workloads as wl

    Ws = wl.batched_c3()
    dev = torch.device("cuda:0")
    ctx = native.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    bvhs, pts, planes = [], [], []
    for W in Ws:
        bvhs.append(ctx.bvh(torch.from_numpy(W.vertices).to(dev),
                            torch.from_numpy(W.triangles.astype(np.int32)).to(dev)))
        pts.append(torch.from_numpy(W.points).to(dev))
        planes.append(torch.from_numpy(W.planes).to(dev))
    ns = [W.n_points for W in Ws]
    S = sum(n * (n - 1) // 2 for n in ns)
    cap = S
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    d_off = torch.zeros(len(Ws) + 1, dtype=torch.int64, device=dev)
    vis_parts = [(pts[i], ns[i], None, bvhs[i], None) for i in range(len(Ws))]

    def vis():
        ctx.visibility_batch_d(vis_parts, d_pairs, cap, d_count, d_off)

    vis()
    torch.cuda.synchronize()
    off = d_off.cpu().numpy()
    base = d_pairs.data_ptr()
    score_parts = [(planes[i], Ws[i].planes.shape[0], pts[i], ns[i], base + 8 * int(off[i]),
                    int(off[i + 1] - off[i])) for i in range(len(Ws))]
    P = sum(W.planes.shape[0] for W in Ws)
    d_q = torch.zeros(P, dtype=torch.int64, device=dev)
    d_am = torch.zeros(len(Ws), dtype=torch.int32, device=dev)

    def score():
        ctx.score_planes_batch_d(score_parts, d_q, d_am)

    def step():
        vis()
        score()

    best, table = autotune.tune(ctx, step, torch.cuda.synchronize)
    log("c3 autotune best", best)
    step()
    torch.cuda.synchronize()
    ref = (int(d_count.item()), d_off.cpu().clone(), d_q.cpu().clone(), d_am.cpu().clone())
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / steps
    # capture the step into a graph (buffers and staging regions are sized by
    # the warm-up calls above) and replay it
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=s):
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        step()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ok = True
    for _ in range(3):  # replays interleaved with eager work, each checked
        d_count.zero_()
        d_am.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        ok &= int(d_count.item()) == ref[0] and torch.equal(d_off.cpu(), ref[1])
        ok &= torch.equal(d_q.cpu(), ref[2]) and torch.equal(d_am.cpu(), ref[3])
        _ = torch.ones(1024, device=dev).sum().item()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / steps
    ok &= int(d_count.item()) == ref[0] and torch.equal(d_am.cpu(), ref[3])
    for b in bvhs:
        b.close()
    ctx.close()
    return {"workload": "64 perturbed icospheres L2-L5, <=2048 points, 1024 planes each",
            "parts": len(Ws), "segments": S, "edges": ref[0], "planes": P,
            "eager_ms_per_step": eager_ms, "graph_ms_per_step": graph_ms,
            "gsegments_per_s_graph": S / (graph_ms * 1e-3) / 1e9,
            "graph_replays_match_eager": bool(ok), "autotune_best": best,
            "autotune_table": [{**c, "ms": round(t, 4)} for c, t in table]}


def vis_pass(ctx, W, steps, warmup, sweeps=False):
    """One workload's visibility pass: wall time per pass over `steps` passes
    (inputs resident in HBM), HIP-event time of the fan pass + walk kernels
    on the launch stream, the instrumented node/triangle work per segment
    and the algorithmic bytes that gives; optionally the walk-variant sweeps."""
    import numpy as np
    import torch

    from visacd_amd import native

    dev = torch.device("cuda:0")
    d_V = torch.from_numpy(W.vertices).to(dev)
    d_T = torch.from_numpy(W.triangles.astype(np.int32)).to(dev)
    d_pts = torch.from_numpy(W.points).to(dev)
    n, S = W.n_points, W.n_segments
    ctx.timing(reset=True)
    bvh = ctx.bvh(d_V, d_T)
    binfo = bvh.info()
    bt = ctx.timing(reset=True)
    cap = S
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)

    def step():
        ctx.visibility_d(d_pts, n, bvh, d_pairs=d_pairs, cap=cap, d_count=d_count)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    E = int(d_count.item())
    st = ctx.visibility_stats()
    ctx.timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    tv = ctx.timing(reset=True)
    if int(d_count.item()) != E:
        raise SystemExit("visibility result changed between steps")
    kms = tv["vis_ms"] / max(1, tv["vis_launches"])

    def timed(k):
        step()
        ctx.timing(reset=True)
        for _ in range(k):
            step()
        t = ctx.timing(reset=True)
        return t["vis_ms"] / max(1, t["vis_launches"])

    out = {"edges": E, "blocked": st["blocked"], "exact_fallback_pairs": st["exact_calls"],
           "walked_pairs": st["walked"], "ms_per_pass": ms, "kernel_ms": kms,
           "gsegments_per_s": S / (ms * 1e-3) / 1e9,
           "kernel_gsegments_per_s": S / (kms * 1e-3) / 1e9,
           "bvh_build_ms": bt["bvh_ms"], "bvh_depth": binfo["depth"]}
    if sweeps:
        ctx.set_option(native.VACD_OPT_FRONT_LANES, 0)  # depth-first walks for this sweep
        lpp = {}
        for k in (1, 2, 4, 8, 16):
            ctx.set_option(native.VACD_OPT_LANES_PER_PAIR, k)
            lpp[k] = timed(max(2, steps // 4))
        ctx.set_option(native.VACD_OPT_LANES_PER_PAIR, 1)
        front = {}
        for fl in (0, 4, 8, 16, 32, 64):
            ctx.set_option(native.VACD_OPT_FRONT_LANES, fl)
            front[fl] = timed(max(2, steps // 4))
        ctx.set_option(native.VACD_OPT_FRONT_LANES, 8)
        ctx.set_option(native.VACD_OPT_FAN_PAIRS, 0)
        walk_all = timed(max(2, steps // 4))
        ctx.set_option(native.VACD_OPT_FAN_PAIRS, 1)
        torch.cuda.synchronize()
        if int(d_count.item()) != E:
            raise SystemExit("visibility result differs between walk variants")
        out.update({"kernel_ms_by_lanes_per_pair": lpp, "kernel_ms_by_front_lanes": front,
                    "kernel_ms_walk_every_pair": walk_all})
    # walks clipped to the segment outside the endpoints' clearance balls
    out["clipped_walks"] = ctx.visibility_counters()["clipped_walks"]
    # algorithmic work per segment (instrumented kernel variant, untimed)
    ctx.set_option(native.VACD_OPT_COUNT_WORK, 1)
    step()
    stw = ctx.visibility_stats()
    ctx.set_option(native.VACD_OPT_COUNT_WORK, 0)
    ctx.timing(reset=True)
    nps, tps = stw["node_visits"] / S, stw["tri_tests"] / S
    # a binary node visit loads one 64-byte node (both child boxes and ids);
    # a triangle test its 9 coordinates (36 B); a segment its endpoints
    # (24 B) and one result bit (counted as 1 B)
    bps = 24 + 1 + 64 * nps + 36 * tps
    out.update({"node_visits_per_segment": nps, "tri_tests_per_segment": tps,
                "alg_bytes_per_segment": bps,
                "achieved_gbs": bps * S / (kms * 1e-3) / 1e9})
    return out, (bvh, d_pts, d_pairs, E, d_V, d_T)


# LBVH algorithmic bytes per triangle: its 9 coordinates read (36 B), the
# 8-byte key/value pair written by each of the 4 radix passes and one
# 64-byte node written
LBVH_BYTES_PER_TRI = 36 + 8 * 4 + 64


def bvh_roofline(ctx, d_V, d_T, n_tris, reps):
    """Builds of the workload's LBVH timed with HIP events around the whole
    build on the context stream (every build kernel, no host waits)."""
    ctx.timing(reset=True)
    for _ in range(reps):
        ctx.bvh(d_V, d_T).close()
    t = ctx.timing(reset=True)
    ms = t["bvh_ms"] / max(1, t["bvh_launches"])
    ach = LBVH_BYTES_PER_TRI * n_tris / (ms * 1e-3) / 1e9
    return {"triangles": n_tris, "build_ms": ms, "builds": int(t["bvh_launches"]),
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "bytes_per_triangle": LBVH_BYTES_PER_TRI}}


def cpu_scoring_baseline(W, edges, seconds):
    """Oracle Q32.32 scoring (all host threads) on a plane sample of the workload."""
    import numpy as np

    import oracle

    E = edges.shape[0]
    P = 8
    t0 = time.perf_counter()
    oracle.score_q32(W.planes[:P], W.points, edges)
    dt = max(time.perf_counter() - t0, 1e-6)
    P = int(min(W.planes.shape[0], max(8, P * seconds / dt)))
    t0 = time.perf_counter()
    oracle.score_q32(W.planes[:P], W.points, edges)
    dt = time.perf_counter() - t0
    return {"value": P * E / dt / 1e9, "unit": "Gtests/s", "cores": oracle.num_threads(),
            "kind": "port", "sample": f"{P} planes x {E} visibility edges of the same workload, "
                                      f"oracle Q32.32 scoring, {dt:.1f} s"}


def c1_decomposition(seconds_cap=120.0):
    """C1 (torus, 1,152 vertices, all pairs, 256 sphere planes per cut,
    concavity 0.04, 8 parts, seed 42): the plain loop on the CPU oracle
    (oracle/decompose.py, all host threads) and visacd.process in plain
    mode on the GPU, both with host distances in the concavity; the two
    decompositions must be identical."""
    import numpy as np

    import oracle
    import visacd as vis
    from oracle import decompose as D
    from visacd_amd import workloads as wl

    W = wl.torus_c1()
    t0 = time.perf_counter()
    ref, ref_h = D.decompose(W.vertices, W.triangles, 0.04, 8, seed=42, num_planes=256,
                             planes="sphere")
    cpu_s = time.perf_counter() - t0
    m = vis.Mesh()
    m.vertices = vis.VecArray3d(W.vertices.astype(np.float64).tolist())
    m.triangles = vis.make_vecarray3i(W.triangles.astype(np.int32))
    vis.engine.verbose = False
    vis.engine.mode = "plain"
    vis.engine.planes = "sphere"
    vis.engine.num_planes = 256
    vis.engine.max_points = 0
    vis.engine.gpu_distance = False
    vis.config.return_parts = True
    vis.set_seed(42)
    t0 = time.perf_counter()
    r = vis.process(m, 0.04, 8)
    gpu_s = time.perf_counter() - t0
    vis.engine.gpu_distance = True
    vis.config.return_parts = False
    same = r.num_parts == len(ref) and r.concavity == ref_h and all(
        np.array_equal(np.array([list(v) for v in p.vertices]), rv) for p, (rv, _rt) in zip(r.parts, ref))
    return {"value": cpu_s, "unit": "s (wall, C1 decomposition)", "cores": oracle.num_threads(),
            "kind": "port", "sample": "the whole C1 decomposition: plain loop with the FP64 oracle "
                                      "as visibility and scoring engine",
            "gpu_plain_wall_s": gpu_s, "parts": int(r.num_parts), "concavity": float(r.concavity),
            "identical_to_gpu": bool(same)}


def acd_speedup(runs):
    """The C5 decomposition target: batched+learned at least 5x faster than plain, parts
    +-1, final concavity within 2 %."""
    by = {(r["mode"], r["proposal"]): r for r in runs if r["workload"].startswith("icosphere-L7")}
    p, b, be = by.get(("plain", "exact")), by.get(("batched", "learned")), by.get(("batched", "exact"))
    if not p or not b:
        return None
    sp = p["wall_s"] / b["wall_s"]
    dc = (b["concavity"] - p["concavity"]) / p["concavity"]
    out = {"workload": p["workload"], "plain_exact_s": p["wall_s"],
           "batched_learned_s": b["wall_s"], "speedup": sp, "target_speedup": 5.0,
           "parts_plain": p["parts"], "parts_batched_learned": b["parts"],
           "delta_parts": b["parts"] - p["parts"], "concavity_plain": p["concavity"],
           "concavity_batched_learned": b["concavity"], "delta_concavity_rel": dc,
           "met": {"speedup": sp >= 5.0, "parts": abs(b["parts"] - p["parts"]) <= 1,
                   "concavity": abs(dc) <= 0.02}}
    if be:
        out["batched_exact_s"] = be["wall_s"]
        out["batched_exact_speedup"] = p["wall_s"] / be["wall_s"]
        out["batched_exact_identical"] = (be["parts"] == p["parts"]
                                          and be["concavity"] == p["concavity"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--score-steps", type=int, default=10)
    ap.add_argument("--no-acd", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the HBM traffic runs")
    ap.add_argument("--no-c3", action="store_true", help="skip the batched 64-part step")
    ap.add_argument("--no-mlp", action="store_true", help="skip the proposal model timing")
    ap.add_argument("--pmc-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--acd-run", metavar="WORKLOAD:MODE:PROPOSAL",
                    help="only one decomposition, e.g. c5:batched:exact (for profiling)")
    ap.add_argument("--vis-only", action="store_true",
                    help="only the headline visibility passes (for profiling)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only the proposal model timing (for profiling)")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe(args.workload)
        return
    if args.mlp_only:
        print(json.dumps(proposal_bench(max(3, args.steps // 2))), flush=True)
        return
    if args.acd_run:
        wname, mode, prop = args.acd_run.split(":")
        print(json.dumps(acd_wall(make_workload(wname), mode, prop)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "this engine runs on one GPU.", flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    from visacd_amd import native

    W = make_workload(args.workload)
    log("workload", W.describe())
    dev = torch.device("cuda:0")
    ctx = native.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.set_option(native.VACD_OPT_TIMING, 1)
    vis, (bvh, d_pts, d_pairs, E, d_V, d_T) = vis_pass(ctx, W, args.steps, args.warmup)
    log("visibility", W.name, vis)
    if args.vis_only:
        print(json.dumps(vis), flush=True)
        return
    n, S, P = W.n_points, W.n_segments, W.planes.shape[0]
    lbvh = bvh_roofline(ctx, d_V, d_T, int(W.triangles.shape[0]), max(3, args.steps // 2))
    log("lbvh", lbvh)

    # plane scoring on the visibility edges of this workload
    d_planes = torch.from_numpy(W.planes).to(dev)
    d_q = torch.zeros(P, dtype=torch.int64, device=dev)
    d_am = torch.zeros(1, dtype=torch.int32, device=dev)
    for _ in range(2):
        ctx.score_planes_d(d_planes, P, d_pts, n, d_pairs, E, d_q32=d_q, d_argmax=d_am)
    torch.cuda.synchronize()
    ctx.timing(reset=True)
    ts0 = time.perf_counter()
    for _ in range(args.score_steps):
        ctx.score_planes_d(d_planes, P, d_pts, n, d_pairs, E, d_q32=d_q, d_argmax=d_am)
    torch.cuda.synchronize()
    ts1 = time.perf_counter()
    tsc = ctx.timing(reset=True)
    score_ms = (ts1 - ts0) * 1e3 / args.score_steps
    score_kernel_ms = tsc["score_ms"] / max(1, tsc["score_launches"])
    tests = P * E
    score = {
        "planes": P, "edges": E, "ms_per_call": score_ms, "kernel_ms": score_kernel_ms,
        "gtests_per_s": tests / (score_ms * 1e-3) / 1e9,
        "kernel_gtests_per_s": tests / (score_kernel_ms * 1e-3) / 1e9,
        "roofline": {"bound": "fp32-valu", "achieved": 12 * tests / (score_kernel_ms * 1e-3) / 1e12,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": 12 * tests / (score_kernel_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                     "flop_per_test": 12},
        "argmax": int(d_am.item()),
    }
    log("scoring", score)
    edges_host = d_pairs[:E].cpu().numpy().astype(np.uint32) if not args.no_cpu else None
    bvh.close()
    # the C2 pass (icosphere L6, 4,096 points) with the walk-variant sweeps
    vis_c2 = None
    if args.workload != "c2":
        W2 = make_workload("c2")
        vis_c2, (b2, *_rest) = vis_pass(ctx, W2, args.steps, args.warmup, sweeps=True)
        b2.close()
        vis_c2["workload"] = W2.name
        log("visibility c2", vis_c2)
    ctx.close()

    line = {
        "metric": "ACD wall-time @concavity 0.04 (plain vs batched+learned); Gsegments/s vis",
        "value": vis["gsegments_per_s"],
        "unit": "Gsegments/s",
        "value_is": "visibility segments per second of one full visibility pass (all pairs of "
                    "the workload's points against its whole mesh), wall time over the timed passes",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": vis["ms_per_pass"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 geometry, f64 filters + exact predicates",
        "data": "synthetic (seeded procedural mesh)",
        "config": {"workload": W.name, "vertices": int(W.vertices.shape[0]),
                   "triangles": int(W.triangles.shape[0]), "points": n, "segments": S,
                   "planes": P, "mode": "midpoint parity", "seed": W.meta.get("seed")},
        "visibility": vis,
        "roofline": {"bound": "hbm", "achieved": vis["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": vis["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fan pass + queue walk (k_visibility, k_walk_*)",
                     "alg_bytes_per_pass": vis["alg_bytes_per_segment"] * S},
        "lbvh": lbvh,
        "scoring": score,
    }
    if vis_c2:
        line["visibility_c2"] = vis_c2
    if not args.no_pmc:
        traffic, detail = pmc_traffic(args.workload)
        log("pmc traffic", traffic, detail)
        line["roofline"]["traffic"] = traffic
        if traffic is None:
            line["roofline"]["traffic_note"] = detail
        else:
            line["roofline"]["traffic_unit"] = "bytes per visibility pass (fan pass + queue walk)"
    if not args.no_mlp:
        line["proposal_mlp"] = proposal_bench(max(3, args.steps // 2))
        log("mlp", line["proposal_mlp"])
    if not args.no_c3:
        line["batched_frontier"] = c3_batched(max(3, args.steps // 2))
        log("c3", {k: v for k, v in line["batched_frontier"].items() if k != "autotune_table"})
    if not args.no_acd:
        from visacd_amd import workloads as wl

        runs = []
        for WW in (wl.icosphere_c2(), wl.icosphere_c5()):
            for mode, prop in (("plain", "exact"), ("batched", "exact"), ("batched", "learned")):
                runs.append(acd_wall(WW, mode, prop))
                log("acd", {k: v for k, v in runs[-1].items() if k != "phase_seconds"})
        line["acd"] = runs
        line["acd_speedup"] = acd_speedup(runs)
    if not args.no_cpu:
        cb = cpu_baseline(W, args.cpu_seconds)
        cb["scoring"] = cpu_scoring_baseline(W, edges_host, args.cpu_seconds / 2)
        cb["c1_decomposition"] = c1_decomposition()
        cb["window_rule"] = window_rule(W)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
