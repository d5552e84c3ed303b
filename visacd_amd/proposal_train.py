"""Training of the learned plane proposal (the vacd_mlp model of include/vacd.h).

The model reads 512 visibility edges of a part and a candidate plane and
predicts the plane's cut length (the sum of the lengths of the edges whose
endpoints lie strictly on opposite sides, the score of lib/cuda/
plane_intersections.cu:43-66) divided by the largest cut length among the
sample's planes: only the order of a part's planes is used, so the target is
normalised per part.  Its training targets are the engine's own exact scores
(vacd_score_planes, Q32.32):

* sample i (seed 10000 + i) is a perturbed icosphere of level U{3, 4, 5} and
  amplitude U(0.1, 0.5), normalised (longest bbox side 2), then cut down to
  the side of 0, 1 or 2 random planes (each through a point at a U(0.25,
  0.75) quantile of the vertices along its normal).  A segment between two
  points on one side of a plane stays on that side, so the visibility edges
  among the kept vertices are those of the clipped part, minus the cap;
* at most 2048 kept vertices take part in the visibility pairs (seeded);
* 256 planes per sample: 192 with normals uniform on the sphere and offsets
  uniform over the kept vertices' extent, 64 perpendicular bisectors of
  random visibility edges (the reference's candidates,
  lib/src/process.cpp:28-72);
* the inputs are built by the driver's own code (_visacd.proposal_inputs), so
  training and inference see identical features.

Adam (lr 1e-3, betas 0.9/0.999) on the mean squared error, batches of 32
samples x 256 planes; the last 5 % of the samples are held out and the
ranking quality on them is reported: for each held-out sample, the best exact
score among the k top-ranked planes over the best exact score of all 256.

Run:  python -m visacd_amd.proposal_train --samples 10000 --out weights.f32
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from visacd_amd import native, workloads

N_PLANES = 256
N_SPHERE = 192
MAX_POINTS = 2048


def make_sample(ctx, seed: int):
    """(features (512, 7), plane inputs (256, 4), targets (256,), info) or None
    when the sample has no visibility edge (convex)."""
    from visacd_amd import _visacd

    rng = np.random.default_rng(seed)
    level = int(rng.integers(3, 6))
    amp = float(rng.uniform(0.1, 0.5))
    n_cuts = int(rng.integers(0, 3))
    v, t = workloads.perturbed_icosphere(level, amp, seed)
    v = workloads.normalize_like_reference(v)
    keep = np.ones(v.shape[0], dtype=bool)
    for _ in range(n_cuts):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        proj = v @ n
        o = np.quantile(proj[keep], rng.uniform(0.25, 0.75))
        keep &= proj >= o
    crop = np.flatnonzero(keep)
    pos = workloads.subsample_vertices(crop.size, MAX_POINTS, seed + 1)  # positions in crop
    sub = crop[pos]
    bvh = ctx.bvh(v.astype(np.float32), t.astype(np.uint32))
    pts = v[sub].astype(np.float32)
    pairs = ctx.visibility(pts, bvh)
    bvh.close()
    if pairs.shape[0] == 0:
        return None
    edges = pos[pairs].astype(np.uint32)
    planes = np.empty((N_PLANES, 4), dtype=np.float64)
    planes[:N_SPHERE] = workloads.sphere_planes(v[crop], N_SPHERE, seed + 2)
    pick = rng.integers(0, pairs.shape[0], N_PLANES - N_SPHERE)
    a = pts[pairs[pick, 0]].astype(np.float64)
    b = pts[pairs[pick, 1]].astype(np.float64)
    nrm = b - a
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-12)
    planes[N_SPHERE:, :3] = nrm
    planes[N_SPHERE:, 3] = -np.einsum("ij,ij->i", nrm, 0.5 * (a + b))
    q, _, _ = ctx.score_planes(planes.astype(np.float32), pts, pairs)
    target = (q.astype(np.float64) / max(1, int(q.max()))).astype(np.float32)
    feat, pin, _ = _visacd.proposal_inputs(v[crop], edges, planes.astype(np.float32).astype(np.float64),
                                           int(native_seed(seed)))
    info = {"level": level, "amplitude": amp, "cuts": n_cuts, "points": int(sub.size),
            "edges": int(pairs.shape[0])}
    return feat, pin, target, info


def native_seed(seed: int) -> int:
    """Seed of a sample's edge draw (any fixed function of the sample seed)."""
    return (seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF


def make_dataset(ctx, n: int, seed0: int = 10000, log=print):
    feats, planes, targets, infos = [], [], [], []
    t0 = time.perf_counter()
    for i in range(n):
        s = make_sample(ctx, seed0 + i)
        if s is not None:
            feats.append(s[0])
            planes.append(s[1])
            targets.append(s[2])
            infos.append(s[3])
        if (i + 1) % 500 == 0:
            log(f"samples {i + 1}/{n} ({len(feats)} with edges) {time.perf_counter() - t0:.1f} s")
    return (np.stack(feats), np.stack(planes), np.stack(targets), infos)


def ranking_quality(pred: np.ndarray, target: np.ndarray, ks=(1, 4, 16)) -> dict:
    """Mean over samples of max(target[top-k by pred]) / max(target), and the
    share of samples where that ratio is at least 0.98."""
    out = {}
    best = target.max(1)
    ok = best > 0
    order = np.argsort(-pred, axis=1, kind="stable")
    for k in ks:
        top = np.take_along_axis(target, order[:, :k], 1).max(1)
        r = top[ok] / best[ok]
        out[f"top{k}_ratio"] = float(r.mean())
        out[f"top{k}_ge_0.98"] = float((r >= 0.98).mean())
    rng = np.random.default_rng(0)
    rnd = np.stack([rng.permutation(target.shape[1]) for _ in range(target.shape[0])])
    for k in ks:
        top = np.take_along_axis(target, rnd[:, :k], 1).max(1)
        out[f"random{k}_ratio"] = float((top[ok] / best[ok]).mean())
    return out


def train(ctx, data, epochs: int, batch: int = 32, lr: float = 1e-3, seed: int = 1,
          holdout: float = 0.05, log=print):
    import torch

    feat, planes, target, _ = data
    N = feat.shape[0]
    n_val = max(1, int(round(holdout * N)))
    n_tr = N - n_val
    dev = torch.device("cuda")
    F = torch.from_numpy(feat).to(dev)
    Q = torch.from_numpy(planes).to(dev)
    Y = torch.from_numpy(target).to(dev)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    m = native.Mlp(ctx, seed=seed)
    g = torch.Generator().manual_seed(seed)
    E, P = feat.shape[1], planes.shape[1]
    curve = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.randperm(n_tr, generator=g)
        losses = []
        for b0 in range(0, n_tr - batch + 1, batch):
            idx = perm[b0:b0 + batch].to(dev)
            f = F.index_select(0, idx).contiguous()
            q = Q.index_select(0, idx).contiguous()
            y = Y.index_select(0, idx).contiguous()
            losses.append(m.train_step_d(batch, E, f, P, q, y, lr))
        val = predict(m, feat[n_tr:], planes[n_tr:])
        mse = float(((val - target[n_tr:]) ** 2).mean())
        rq = ranking_quality(val, target[n_tr:])
        curve.append({"epoch": ep + 1, "train_mse": float(np.mean(losses)), "val_mse": mse, **rq,
                      "seconds": time.perf_counter() - t0})
        log(json.dumps(curve[-1]))
    return m, curve


def predict(m, feat: np.ndarray, planes: np.ndarray, chunk: int = 64) -> np.ndarray:
    out = []
    for i in range(0, feat.shape[0], chunk):
        out.append(m.forward(feat[i:i + chunk], planes[i:i + chunk]))
    return np.concatenate(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", required=True, help="float32 parameter file to write")
    ap.add_argument("--log", default=None, help="JSON summary to write")
    a = ap.parse_args(argv)

    import torch  # noqa: F401  (shares the HIP runtime before libvacd loads)

    def log(msg):
        print(msg, flush=True)

    ctx = native.Context(0)
    t0 = time.perf_counter()
    data = make_dataset(ctx, a.samples, log=log)
    t_data = time.perf_counter() - t0
    log(f"dataset: {data[0].shape[0]} samples in {t_data:.1f} s")
    m, curve = train(ctx, data, a.epochs, a.batch, a.lr, log=log)
    w = m.get()
    w.astype("<f4").tofile(a.out)
    infos = data[3]
    summary = {
        "samples_drawn": a.samples, "samples_with_edges": int(data[0].shape[0]),
        "levels": {str(k): sum(1 for x in infos if x["level"] == k) for k in (3, 4, 5)},
        "cuts": {str(k): sum(1 for x in infos if x["cuts"] == k) for k in (0, 1, 2)},
        "edges_mean": float(np.mean([x["edges"] for x in infos])),
        "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "dataset_seconds": t_data,
        "final": curve[-1], "curve": curve, "params": int(w.size),
    }
    if a.log:
        with open(a.log, "w") as f:
            json.dump(summary, f, indent=1)
    m.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
