"""ctypes front end of the FP64 CPU oracle (``oracle/oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product
(``visacd_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
SRC = os.path.join(_HERE, "oracle.cpp")

# g++ only; -ffp-contract=off keeps the float32 emulation free of implicit
# fused multiply-adds (the explicit fmaf calls define the arithmetic).
BUILD_CMD = ["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
             "-ffp-contract=off", "-fno-fast-math", SRC, "-o", LIB_PATH]

_lib = None


def build() -> str:
    subprocess.run(BUILD_CMD, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        i64 = ctypes.c_int64
        vp = ctypes.c_void_p
        L.orc_orient3d.argtypes = [vp] * 4
        L.orc_orient3d_mid.argtypes = [vp] * 5
        L.orc_orient2d.argtypes = [vp] * 3 + [ctypes.c_int]
        L.orc_orient2d_mid.argtypes = [vp] * 4 + [ctypes.c_int]
        L.orc_contact.argtypes = [vp] * 5
        L.orc_ray_cross.argtypes = [vp] * 5
        L.orc_k_to_ij.argtypes = [i64, i64, vp, vp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_visibility.argtypes = [vp, i64, vp, vp, i64, vp, i64, vp, i64, vp, i64,
                                     ctypes.c_int, vp, i64, vp, vp, vp, ctypes.c_int]
        L.orc_visibility.restype = i64
        L.orc_visibility_sample.argtypes = [vp, i64, vp, vp, i64, vp, i64, vp, i64, vp,
                                            i64, ctypes.c_int, vp, i64, vp, vp, vp,
                                            ctypes.c_int]
        L.orc_score_fp64.argtypes = [vp, i64, vp, i64, vp, i64, vp]
        L.orc_score_q32.argtypes = [vp, i64, vp, i64, vp, i64, vp]
        L.orc_argmax_i64.argtypes = [vp, i64]
        L.orc_argmax_i64.restype = i64
        L.orc_bisector_planes.argtypes = [vp, vp, i64, ctypes.c_int, ctypes.c_uint32, vp]
        L.orc_bisector_planes.restype = i64
        L.orc_mt19937.argtypes = [ctypes.c_uint32, i64, vp]
        L.orc_predicates_batch.argtypes = [ctypes.c_int, i64, vp, ctypes.c_int, vp]
        L.orc_window_hits.argtypes = [vp, i64, vp, vp, i64, vp, i64, ctypes.c_double, vp]
        L.orc_uniform_int.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, i64, vp]
        del fp
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data


def orient3d(a, b, c, d) -> int:
    a, b, c, d = (_f32(x) for x in (a, b, c, d))
    return lib().orc_orient3d(_ptr(a), _ptr(b), _ptr(c), _ptr(d))


def orient3d_mid(a, b, c, p, q) -> int:
    a, b, c, p, q = (_f32(x) for x in (a, b, c, p, q))
    return lib().orc_orient3d_mid(_ptr(a), _ptr(b), _ptr(c), _ptr(p), _ptr(q))


def orient2d(a, b, c, k: int) -> int:
    a, b, c = (_f32(x) for x in (a, b, c))
    return lib().orc_orient2d(_ptr(a), _ptr(b), _ptr(c), k)


def orient2d_mid(a, b, p, q, k: int) -> int:
    a, b, p, q = (_f32(x) for x in (a, b, p, q))
    return lib().orc_orient2d_mid(_ptr(a), _ptr(b), _ptr(p), _ptr(q), k)


def contact(p, q, a, b, c) -> bool:
    p, q, a, b, c = (_f32(x) for x in (p, q, a, b, c))
    return bool(lib().orc_contact(_ptr(p), _ptr(q), _ptr(a), _ptr(b), _ptr(c)))


def ray_cross(p, q, a, b, c) -> bool:
    p, q, a, b, c = (_f32(x) for x in (p, q, a, b, c))
    return bool(lib().orc_ray_cross(_ptr(p), _ptr(q), _ptr(a), _ptr(b), _ptr(c)))


def k_to_ij(k: int, n: int):
    i = ctypes.c_int64()
    j = ctypes.c_int64()
    lib().orc_k_to_ij(k, n, ctypes.byref(i), ctypes.byref(j))
    return i.value, j.value


def _mesh_args(V, T):
    if V is None:
        return None, 0, None, 0, None, None
    V = _f32(V)
    T = np.ascontiguousarray(T, dtype=np.uint32)
    return _ptr(V), V.shape[0], _ptr(T), T.shape[0], V, T


def _groups(groups):
    """(tri_group int8 [T] (-1 none), planes float64 [G, 4]) -> ctypes args."""
    if groups is None:
        return None, None, 0, ()
    g = np.ascontiguousarray(groups[0], dtype=np.int8)
    pl = np.ascontiguousarray(groups[1], dtype=np.float64).reshape(-1, 4)
    return _ptr(g), _ptr(pl), int(pl.shape[0]), (g, pl)


def visibility(points, V, T, mask=None, cage=None, brute=False, cap=None, groups=None):
    """Visibility edges (E, 2) uint32 in k order and detail counts.

    groups = (tri_group, planes): the triangle groups of mask bits 1-7 (the
    cap rule), as given to the engine's planar BVH build.
    Returns (pairs, stats) with stats = {none, edge, masked, zero, blocked}.
    """
    pts = _f32(points)
    n = pts.shape[0]
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    sv, snv, st, snt, _k1, _k2 = _mesh_args(V, T)
    if cage is not None:
        cv, cnv, ct, cnt, _k3, _k4 = _mesh_args(*cage)
    else:
        cv, cnv, ct, cnt = None, 0, None, 0
    if cap is None:
        cap = max(1, n * (n - 1) // 2)
    out = np.zeros((cap, 2), dtype=np.uint32)
    stats = np.zeros(5, dtype=np.int64)
    gp, pp, ng, _keep = _groups(groups)
    e = lib().orc_visibility(_ptr(pts), n, _ptr(m), sv, snv, st, snt, cv, cnv, ct, cnt,
                             1 if brute else 0, _ptr(out), cap, _ptr(stats), gp, pp, ng)
    if e < 0:
        raise RuntimeError("oracle: edge buffer too small")
    names = ["none", "edge", "masked", "zero", "blocked"]
    return out[:e].copy(), dict(zip(names, stats.tolist()))


def visibility_sample(points, V, T, ks, mask=None, cage=None, brute=False, groups=None):
    """Per-pair detail codes (0 none, 1 edge, 2 masked, 3 zero, 4 blocked) for pair ids ks."""
    pts = _f32(points)
    n = pts.shape[0]
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    ks = np.ascontiguousarray(ks, dtype=np.int64)
    sv, snv, st, snt, _k1, _k2 = _mesh_args(V, T)
    if cage is not None:
        cv, cnv, ct, cnt, _k3, _k4 = _mesh_args(*cage)
    else:
        cv, cnv, ct, cnt = None, 0, None, 0
    out = np.zeros(ks.shape[0], dtype=np.uint8)
    gp, pp, ng, _keep = _groups(groups)
    lib().orc_visibility_sample(_ptr(pts), n, _ptr(m), sv, snv, st, snt, cv, cnv, ct, cnt,
                                1 if brute else 0, _ptr(ks), ks.shape[0], _ptr(out), gp, pp, ng)
    return out


WINDOW_EXACT, WINDOW_HIT, WINDOW_NO_TMIN, WINDOW_SHORT, WINDOW_GRAZING = 1, 2, 4, 8, 16


def window_hits(points, V, T, tau, ks=None):
    """The reference's ray-window rule (lib/ptx/ray_segments.cu:77-111,
    float32 watertight test, t in [0.001, len]) beside the exact contact,
    per pair k (all pairs in k order when ks is None): bit WINDOW_EXACT
    exact contact, WINDOW_HIT window hit, WINDOW_NO_TMIN a hit with t in
    [0, len], WINDOW_SHORT a hit with t in [0.001, len - tau],
    WINDOW_GRAZING every triangle on which the two disagree lies within tau
    of grazing."""
    pts = _f32(points)
    n = pts.shape[0]
    sv, _snv, st, snt, _k1, _k2 = _mesh_args(V, T)
    if ks is not None:
        ks = np.ascontiguousarray(ks, dtype=np.int64)
        m = ks.shape[0]
    else:
        m = n * (n - 1) // 2
    out = np.zeros(m, dtype=np.uint8)
    lib().orc_window_hits(_ptr(pts), n, sv, st, snt, _ptr(ks), m, float(tau), _ptr(out))
    return out


def window_classes(bits):
    """Counts of the pairs on which the window rule and the exact contact
    agree or disagree, by cause: 'tmin' (only contacts within 0.001 of p,
    which the window skips), 'q_end' (the window's only hits lie within tau
    of q, where rounding decides t <= len), 'grazing' (every disagreeing
    triangle within tau of grazing) and 'other'."""
    b = np.asarray(bits)
    ex = (b & WINDOW_EXACT) != 0
    win = (b & WINDOW_HIT) != 0
    miss = ex & ~win  # contact the window rule misses
    extra = ~ex & win  # a window hit without contact
    tmin = miss & ((b & WINDOW_NO_TMIN) != 0)
    q_end = extra & ((b & WINDOW_SHORT) == 0)
    graze = ((miss & ~tmin) | (extra & ~q_end)) & ((b & WINDOW_GRAZING) != 0)
    other = ((miss & ~tmin) | (extra & ~q_end)) & ~graze
    return {"pairs": int(b.size), "agree": int((ex == win).sum()), "missed": int(miss.sum()),
            "extra": int(extra.sum()), "tmin": int(tmin.sum()), "q_end": int(q_end.sum()),
            "grazing": int(graze.sum()), "other": int(other.sum())}


def score_fp64(planes, points, edges):
    pl = _f32(planes)
    pts = _f32(points)
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(pl.shape[0], dtype=np.float64)
    lib().orc_score_fp64(_ptr(pl), pl.shape[0], _ptr(pts), pts.shape[0], _ptr(e), e.shape[0],
                         _ptr(out))
    return out


def score_q32(planes, points, edges):
    pl = _f32(planes)
    pts = _f32(points)
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(pl.shape[0], dtype=np.int64)
    lib().orc_score_q32(_ptr(pl), pl.shape[0], _ptr(pts), pts.shape[0], _ptr(e), e.shape[0],
                        _ptr(out))
    return out


def argmax_i64(scores) -> int:
    s = np.ascontiguousarray(scores, dtype=np.int64)
    return int(lib().orc_argmax_i64(_ptr(s), s.shape[0]))


def bisector_planes(vertices64, edges, num_planes: int, seed: int):
    V = np.ascontiguousarray(vertices64, dtype=np.float64)
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros((max(num_planes, 1), 4), dtype=np.float64)
    n = lib().orc_bisector_planes(_ptr(V), _ptr(e), e.shape[0], num_planes, seed, _ptr(out))
    return out[:n].copy()


def num_threads() -> int:
    return int(lib().orc_num_threads())


def mt19937(seed: int, n: int) -> np.ndarray:
    """The first n raw outputs of std::mt19937(seed)."""
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt19937(seed, n, _ptr(out))
    return out


def uniform_int(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    """n draws of std::uniform_int_distribution<int>(lo, hi) on std::mt19937(seed)."""
    out = np.zeros(n, dtype=np.int32)
    lib().orc_uniform_int(seed, lo, hi, n, _ptr(out))
    return out


def predicates_batch(kind: int, inputs, k: int = 0) -> np.ndarray:
    """Exact signs of n predicates: kind 0 orient3d (12 floats per case), 1
    orient3d with a midpoint (15), 2 orient2d in projection k (9), 3 orient2d
    with a midpoint (12)."""
    x = _f32(inputs)
    stride = {0: 12, 1: 15, 2: 9, 3: 12}[kind]
    n = x.size // stride
    out = np.zeros(n, dtype=np.int32)
    lib().orc_predicates_batch(kind, n, _ptr(x), k, _ptr(out))
    return out
