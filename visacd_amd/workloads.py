"""Seeded synthetic meshes, vertex subsamples and candidate-plane sets.

These define the inputs of the visibility and scoring paths.  Geometry is
generated in float64 with numpy's PCG64 (``numpy.random.default_rng(seed)``),
normalised the way the reference normalises a mesh before decomposition
(bounding-box centre to the origin, longest side to 2:
``lib/src/core.cpp:262-291``), then rounded to float32 once.  Every consumer
(HIP kernels, the FP64 oracle, tests, bench) sees the same float32 arrays.

Draw order inside each generator is part of its definition and is documented
per function, so a seed fixes the workload bit for bit.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "normalize_like_reference",
    "torus",
    "icosphere",
    "perturbed_icosphere",
    "subsample_vertices",
    "sphere_planes",
    "Workload",
    "torus_c1",
    "icosphere_c2",
    "icosphere_c5",
    "batched_c3",
]


def normalize_like_reference(v: np.ndarray) -> np.ndarray:
    """Centre the bbox at the origin and scale the longest side to 2.

    Restates ``Mesh::normalize`` (``lib/src/core.cpp:262-291``) in float64.
    """
    v = np.asarray(v, dtype=np.float64)
    lo = v.min(axis=0)
    hi = v.max(axis=0)
    m_len = float((hi - lo).max())
    mid = (hi + lo) / 2.0
    return 2.0 * (v - mid) / m_len


def torus(R: float = 1.0, r: float = 0.35, nu: int = 48, nv: int = 24,
          jitter: float = 0.002, seed: int = 42):
    """Closed torus, ``nu`` steps around the ring, ``nv`` around the tube.

    Vertex (u, v) = ((R + r cos v) cos u, (R + r cos v) sin u, r sin v) with
    u = 2*pi*iu/nu, v = 2*pi*iv/nv, index iu*nv + iv.  Each quad gives two
    outward-facing triangles.  Jitter: one ``rng.uniform(-jitter, jitter,
    (V, 3))`` draw added before normalisation.
    """
    rng = np.random.default_rng(seed)
    iu, iv = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    u = 2.0 * np.pi * iu.ravel() / nu
    w = 2.0 * np.pi * iv.ravel() / nv
    x = (R + r * np.cos(w)) * np.cos(u)
    y = (R + r * np.cos(w)) * np.sin(u)
    z = r * np.sin(w)
    v = np.stack([x, y, z], axis=1)
    v = v + rng.uniform(-jitter, jitter, size=v.shape)
    a = iu.ravel() * nv + iv.ravel()
    b = ((iu.ravel() + 1) % nu) * nv + iv.ravel()
    c = ((iu.ravel() + 1) % nu) * nv + (iv.ravel() + 1) % nv
    d = iu.ravel() * nv + (iv.ravel() + 1) % nv
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 0)
    v = normalize_like_reference(v).astype(np.float32)
    return v, tris.astype(np.uint32)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
        dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level: int):
    """Unit icosphere: 10*4^L + 2 vertices, 20*4^L outward triangles (float64).

    Each level splits every triangle into four at its edge midpoints
    (midpoints pushed to the unit sphere); new vertices are numbered after
    the old ones in order of the sorted unique edge key lo*V + hi.
    """
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], 0)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * nv + hi
        uniq, inv = np.unique(key, return_inverse=True)
        ul, uh = uniq // nv, uniq % nv
        mid = v[ul] + v[uh]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], 0)
        m = (inv + nv).reshape(3, -1).T  # m01, m12, m20 per face
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([
            np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)], 0)
    return v, f


def perturbed_icosphere(level: int, amplitude: float, seed: int,
                        n_terms: int = 8, dents: int = 0,
                        dent_depth: float = 0.3, dent_sigma: float = 0.15):
    """Icosphere with radius 1 + A * sum_i a_i sin(k_i . n + phi_i) (+ dents).

    Draw order from ``default_rng(seed)``: a = uniform(0.2, 1, n_terms)
    (then normalised to sum 1); kdir = standard_normal((n_terms, 3));
    kmag = uniform(2, 6, n_terms); phi = uniform(0, 2*pi, n_terms); then, if
    ``dents``: centres = standard_normal((dents, 3)) normalised to S^2.  A
    dent subtracts depth*exp(-theta^2 / (2 sigma^2)) from the radius, theta
    being the angle between the vertex direction and the dent centre.
    """
    rng = np.random.default_rng(seed)
    n, f = icosphere(level)
    a = rng.uniform(0.2, 1.0, n_terms)
    a = a / a.sum()
    kdir = rng.standard_normal((n_terms, 3))
    kmag = rng.uniform(2.0, 6.0, n_terms)
    k = kdir * kmag[:, None]
    phi = rng.uniform(0.0, 2.0 * np.pi, n_terms)
    r = 1.0 + amplitude * (np.sin(n @ k.T + phi[None, :]) * a[None, :]).sum(1)
    if dents:
        cen = rng.standard_normal((dents, 3))
        cen /= np.linalg.norm(cen, axis=1, keepdims=True)
        theta = np.arccos(np.clip(n @ cen.T, -1.0, 1.0))
        r = r - (dent_depth * np.exp(-theta ** 2 / (2.0 * dent_sigma ** 2))).sum(1)
    v = normalize_like_reference(n * r[:, None]).astype(np.float32)
    return v, f.astype(np.uint32)


def subsample_vertices(n_vertices: int, k: int, seed: int) -> np.ndarray:
    """Seeded k-vertex subsample (sorted ascending; all vertices if k >= V)."""
    if k >= n_vertices:
        return np.arange(n_vertices, dtype=np.int64)
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_vertices, size=k, replace=False)).astype(np.int64)


def sphere_planes(points: np.ndarray, n_planes: int, seed: int) -> np.ndarray:
    """Candidate planes (a, b, c, d) with a*x + b*y + c*z + d = 0, float32.

    Normals uniform on S^2 (normalised standard_normal((P, 3))); offsets
    o ~ uniform over [min, max] of n.p over ``points`` (one
    uniform(0, 1, P) draw, o = lo + u*(hi - lo)); d = -o.
    """
    rng = np.random.default_rng(seed)
    nrm = rng.standard_normal((n_planes, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    u = rng.uniform(0.0, 1.0, n_planes)
    proj = np.asarray(points, np.float64) @ nrm.T
    lo, hi = proj.min(0), proj.max(0)
    off = lo + u * (hi - lo)
    return np.concatenate([nrm, -off[:, None]], 1).astype(np.float32)


class Workload:
    """A mesh, the points whose pairs are tested, and candidate planes."""

    def __init__(self, name, vertices, triangles, point_ids, planes, meta):
        self.name = name
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.uint32)
        self.point_ids = np.asarray(point_ids, dtype=np.int64)
        self.points = np.ascontiguousarray(self.vertices[self.point_ids])
        self.planes = np.ascontiguousarray(planes, dtype=np.float32)
        self.meta = dict(meta)

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_segments(self) -> int:
        n = self.n_points
        return n * (n - 1) // 2

    def describe(self) -> dict:
        return {"workload": self.name, "vertices": int(self.vertices.shape[0]),
                "triangles": int(self.triangles.shape[0]),
                "points": self.n_points, "segments": self.n_segments,
                "planes": int(self.planes.shape[0]), **self.meta}


def torus_c1(seed: int = 42) -> Workload:
    """Torus R=1, r=0.35, 48x24 (1,152 v / 2,304 t), all 662,976 pairs, 256 planes."""
    v, t = torus(seed=seed)
    ids = np.arange(v.shape[0])
    planes = sphere_planes(v, 256, seed + 1)
    return Workload("torus-48x24", v, t, ids, planes,
                    {"seed": seed, "concavity": 0.04, "num_parts": 8})


def icosphere_c2(seed: int = 42, n_sub: int = 4096, n_planes: int = 4096) -> Workload:
    """Icosphere L6 (40,962 v / 81,920 t), amplitude 0.3, 4,096-vertex subsample."""
    v, t = perturbed_icosphere(6, 0.3, seed)
    ids = subsample_vertices(v.shape[0], n_sub, seed + 1)
    planes = sphere_planes(v[ids], n_planes, seed + 2)
    return Workload("icosphere-L6-sub4096", v, t, ids, planes,
                    {"seed": seed, "amplitude": 0.3, "concavity": 0.04,
                     "num_parts": 32})


def icosphere_c5(seed: int = 42, n_sub: int = 16384, n_planes: int = 16384) -> Workload:
    """Icosphere L7 (163,842 v / 327,680 t), amplitude 0.5 + 4 dents, 16,384 subsample."""
    v, t = perturbed_icosphere(7, 0.5, seed, dents=4)
    ids = subsample_vertices(v.shape[0], n_sub, seed + 1)
    planes = sphere_planes(v[ids], n_planes, seed + 2)
    return Workload("icosphere-L7-dents-sub16384", v, t, ids, planes,
                    {"seed": seed, "amplitude": 0.5, "dents": 4,
                     "concavity": 0.04, "num_parts": 64})


def batched_c3(n_meshes: int = 64, base_seed: int = 1000, n_sub: int = 2048,
               n_planes: int = 1024):
    """64 perturbed icospheres, level U{2..5}, amplitude U(0.1, 0.5), seed 1000+i.

    Mesh i draws (level, amplitude) from ``default_rng(1000+i)`` in that
    order (integers(2, 6), uniform(0.1, 0.5)); its geometry uses the same
    seed, its subsample seed 1000+i+1 and its planes seed 1000+i+2.
    """
    out = []
    for i in range(n_meshes):
        s = base_seed + i
        rng = np.random.default_rng(s)
        level = int(rng.integers(2, 6))
        amp = float(rng.uniform(0.1, 0.5))
        v, t = perturbed_icosphere(level, amp, s)
        ids = subsample_vertices(v.shape[0], n_sub, s + 1)
        planes = sphere_planes(v[ids], n_planes, s + 2)
        out.append(Workload(f"icosphere-L{level}-a{amp:.3f}", v, t, ids, planes,
                            {"seed": s, "level": level, "amplitude": amp}))
    return out
