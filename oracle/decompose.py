"""CPU restatement of the plain decomposition loop, with the FP64 oracle as
its visibility and scoring engine.

TEST INFRASTRUCTURE ONLY (like the rest of ``oracle/``): the checker of
``visacd.process`` in plain mode and the CPU baseline of the C1 decomposition
wall time.  It follows lib/src/process.cpp:238-347 step for step as the
engine's driver restates it -- normalise, flat-surface planes, disjoint
parts, initial visibility, then per step the part of highest concavity
(first maximum), its candidate planes from its own seeded generator, exact
Q32.32 scores with flat planes times flat_surface_k (first maximum), the
clip with inherited edges, the new parts' masked visibility (with the cap
rule), hulls and the final concavity -- with every GPU call replaced by the
oracle: ``oracle.visibility`` for the visibility edges, ``oracle.score_q32``
for the scores, ``oracle.bisector_planes`` for the candidate sampler and
numpy for the sphere planes' extents.  The host geometry services (clip and
its bookkeeping, connected pieces, hulls, surface sampling and distances,
flat surfaces) are the engine's own host code, shared: what this pins is the
loop and every result the GPU feeds into it.  Concavity uses host distances,
so compare against ``visacd.engine.gpu_distance = False``.
"""
from __future__ import annotations

import numpy as np

import oracle

_M64 = (1 << 64) - 1


def mix64(x: int) -> int:
    """splitmix64 finaliser (the driver's part-seed mixer)."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def kid_seed(parent: int, k: int) -> int:
    return mix64((parent * 31 + k) & _M64)


def plane_seed(s: int, attempt: int) -> int:
    return mix64(s ^ ((0x5EED + attempt) & _M64)) & 0xFFFFFFFF


def root_seed(seed: int) -> int:
    """Two draws of std::mt19937(seed) (set_seed), as process() takes them."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    a, b = int(bg.random_raw()), int(bg.random_raw())
    return ((a << 32) | b) & _M64


def _eval(pl, v):
    """a x + b y + c z + d evaluated as the host's Plane::eval (left to right)."""
    return v[:, 0] * pl[0] + v[:, 1] * pl[1] + v[:, 2] * pl[2] + pl[3]


def cap_groups(V, T, cuts):
    """Planar groups of a part's BVH: the caps of its last 7 cuts with at
    least 16 triangles, all three vertices within 1e-9 |n| of the plane."""
    grp = np.full(T.shape[0], -1, np.int8)
    planes = []
    for c in range(len(cuts) - 1, -1, -1):
        if len(planes) >= 7:
            break
        q = cuts[c]
        l = np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2])
        if not l > 0:
            continue
        on = np.abs(_eval(q, V)) <= 1e-9 * l
        mine = (grp < 0) & on[T].all(1)
        if mine.sum() < 16:
            continue
        grp[mine] = len(planes)
        planes.append(np.asarray(q, np.float64))
    return grp, planes


def _bvh_groups(V32, T, grp, planes):
    """The groups the planar BVH build keeps: a group whose float32 vertices
    lie farther than 1e-4 of the mesh extent from its (float) plane is dropped."""
    if not planes:
        return None
    ext = float((V32.max(0).astype(np.float64) - V32.min(0).astype(np.float64)).max())
    grp = grp.copy()
    for g, p in enumerate(planes):
        l = np.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
        nf = (np.asarray(p, np.float64) / l).astype(np.float32).astype(np.float64)
        idx = np.where(grp == g)[0]
        x = V32[T[idx].ravel()].astype(np.float64)
        tau = np.abs(nf[0] * x[:, 0] + nf[1] * x[:, 1] + nf[2] * x[:, 2] + nf[3]).max() if idx.size else 0.0
        if tau > 1e-4 * ext:
            grp[idx] = -1
    return grp, np.array(planes)


def point_mask(V, is_new, points, planes):
    m = np.where(is_new[points], 1, 0).astype(np.uint8)
    for g, q in enumerate(planes):
        l = np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2])
        on = np.abs(_eval(q, V[points])) <= 1e-9 * l
        m |= (on * (2 << g)).astype(np.uint8)
    return m


def visibility(part, with_mask):
    """Visibility edges of a part as vertex-id pairs (its points' pairs)."""
    V, T, pts = part["V"], part["T"], part["points"]
    if len(pts) < 2:
        return np.zeros((0, 2), np.uint32)
    V32 = V.astype(np.float32)
    grp, planes = cap_groups(V, T, part["cuts"])
    mask = point_mask(V, part["is_new"], pts, planes) if with_mask else None
    if with_mask and not (mask & 1).any():
        return np.zeros((0, 2), np.uint32)
    groups = _bvh_groups(V32, T, grp, planes) if with_mask else None
    e, _ = oracle.visibility(V32[pts], V32, T.astype(np.uint32), mask=mask, groups=groups)
    return np.asarray(pts, np.uint32)[e]


def candidates(part, seed, attempt, num_planes, kind, flat):
    from visacd_amd import _visacd as H

    ps = plane_seed(seed, attempt)
    if kind == "sphere":
        nrm, frac = H.sphere_draws(num_planes, ps)
        p = part["V"][part["points"]]
        proj = (nrm[:, 0:1] * p[:, 0] + nrm[:, 1:2] * p[:, 1]) + nrm[:, 2:3] * p[:, 2]
        lo, hi = proj.min(1), proj.max(1)
        planes = np.concatenate([nrm, -(lo + frac * (hi - lo))[:, None]], 1)
    else:
        planes = oracle.bisector_planes(part["V"], part["E"], num_planes, ps)
    ss = planes.shape[0]
    if len(flat):
        planes = np.concatenate([planes.reshape(-1, 4), np.asarray(flat, np.float64)], 0)
    return planes.reshape(-1, 4), ss


def _mesh(H, V, T):
    m = H.Mesh()
    m.vertices = H.VecArray3d(V.tolist())
    m.triangles = H.make_vecarray3i(np.asarray(T, np.int32))
    return m


def decompose(V, T, concavity, num_parts, seed=42, num_planes=2500, planes="bisector",
              resolution=3000, sample_seed=42, flat_min_area=0.1, use_flat=True):
    """The plain loop on the CPU oracle.  Returns (parts as (V, T) arrays in
    the input frame, final concavity)."""
    from visacd_amd import _visacd as H

    V = np.asarray(V, np.float64)
    lo, hi = V.min(0), V.max(0)
    ln = float((hi - lo).max())
    mid = (hi + lo) / 2
    Vn = 2.0 * (V - mid) / ln
    mesh = _mesh(H, Vn, T)
    flat = [list(p) for p in H.flat_surface_planes(mesh, flat_min_area)] if use_flat else []
    root = root_seed(seed)
    parts = []
    for i, r in enumerate(H.separate_disjoint(mesh)):
        pv = np.array([list(v) for v in r.vertices], np.float64)
        pt = np.array([list(t) for t in r.triangles], np.int64)
        part = {"V": pv, "T": pt, "is_new": np.zeros(pv.shape[0], bool),
                "points": np.arange(pv.shape[0]), "cuts": np.zeros((0, 4)),
                "seed": mix64((root + i) & _M64), "attempt": 0, "conc": None}
        part["E"] = visibility(part, False)
        parts.append(part)

    def conc_of(p):
        if p["conc"] is None:
            m = _mesh(H, p["V"], p["T"])
            p["conc"] = H.concavity(m, H.convex_hull(m), resolution, sample_seed)
        return p["conc"]

    for _step in range(num_parts - 1):
        key = [conc_of(p) for p in parts]
        idx, best = -1, -1.0
        for i, k in enumerate(key):
            if k > best:
                idx, best = i, k
        if idx < 0 or best < concavity:
            break
        part = parts.pop(idx)
        pl, ss = candidates(part, part["seed"], part["attempt"], num_planes, planes, flat)
        if pl.shape[0] == 0 or part["E"].shape[0] == 0:
            parts.append(part)
            break
        q = oracle.score_q32(pl.astype(np.float32), part["V"].astype(np.float32), part["E"])
        b = H.select_plane(q.tolist(), ss)
        kids = H.split_part(part["V"], part["T"], part["E"], part["points"], part["cuts"],
                            part["seed"], list(pl[b]))
        if kids is None:
            part["attempt"] += 1
            parts.append(part)
            continue
        for kv, kt, kn, ke, kp, kc, ks in kids:
            kid = {"V": kv, "T": kt.astype(np.int64), "is_new": np.asarray(kn, bool),
                   "points": np.asarray(kp, np.int64), "cuts": kc, "seed": int(ks),
                   "attempt": 0, "conc": None}
            new = visibility(kid, True)
            kid["E"] = np.concatenate([np.asarray(ke, np.uint32).reshape(-1, 2), new], 0)
            parts.append(kid)
    h = 0.0
    out = []
    for p in parts:
        m = _mesh(H, p["V"], p["T"])
        hull = H.convex_hull(m)
        h = max(h, H.hausdorff_boundary(m, hull, 10000, sample_seed))
        out.append((p["V"] / 2 * ln + mid, p["T"]))
    return out, h
