"""ctypes binding of libvacd (the C ABI in ``include/vacd.h``).

This module is the Python-side caller of the HIP engine: it loads the
in-tree ``libvacd.so`` and fails loudly when it is missing -- there is no
CPU fallback.  Host (numpy) and device (torch tensor) entry points mirror
the C ABI one to one.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvacd.so")

VACD_OK = 0
VACD_ERR_INVALID = 1
VACD_ERR_HIP = 2
VACD_ERR_CAPACITY = 3
VACD_ERR_LIMIT = 4
VACD_NOALLOC_MAX_PLANES = 65536
VACD_NOALLOC_MAX_EDGES = 1048576

# Every symbol include/vacd.h declares (checked by tests/test_abi.py).
EXPORTED = [
    "vacd_ctx_create", "vacd_ctx_destroy", "vacd_last_error", "vacd_ctx_stream",
    "vacd_ctx_set_stream", "vacd_sync", "vacd_ctx_set_option", "vacd_ctx_timing", "vacd_bvh_build", "vacd_bvh_build_planar",
    "vacd_bvh_build_d",
    "vacd_bvh_destroy", "vacd_bvh_info", "vacd_bvh_contact_info", "vacd_bvh_export",
    "vacd_visibility",
    "vacd_visibility_d", "vacd_visibility_stats", "vacd_visibility_counters", "vacd_visibility_deferred",
    "vacd_visibility_batch_d", "vacd_visibility_batch", "vacd_score_planes",
    "vacd_score_planes_d", "vacd_score_planes_noalloc_d", "vacd_score_planes_batch", "vacd_score_planes_batch_d",
    "vacd_max_distance", "vacd_max_distance_d", "vacd_max_distance_batch", "vacd_plane_extents_batch",
    "vacd_sdf_remesh", "vacd_sdf_grid",
    "vacd_gemm_bf16_d", "vacd_mlp_param_count",
    "vacd_mlp_create", "vacd_mlp_destroy", "vacd_mlp_set_params", "vacd_mlp_get_params", "vacd_mlp_get_grads",
    "vacd_mlp_forward_d", "vacd_mlp_forward", "vacd_mlp_train_step_d", "classify_and_rate_planes",
]


class VacdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vacd error {code}: {msg}")
        self.code = code


class VisStats(ctypes.Structure):
    _fields_ = [("segments", ctypes.c_int64), ("masked", ctypes.c_int64),
                ("blocked", ctypes.c_int64), ("edges", ctypes.c_int64),
                ("exact_calls", ctypes.c_int64), ("node_visits", ctypes.c_int64),
                ("tri_tests", ctypes.c_int64), ("walked", ctypes.c_int64)]


class Part(ctypes.Structure):
    _fields_ = [("pts", ctypes.c_void_p), ("n", ctypes.c_int64), ("new_mask", ctypes.c_void_p),
                ("self", ctypes.c_void_p), ("cage", ctypes.c_void_p)]


class DistanceQuery(ctypes.Structure):
    _fields_ = [("pts", ctypes.c_void_p), ("m", ctypes.c_int64), ("V", ctypes.c_void_p),
                ("nV", ctypes.c_int64), ("T", ctypes.c_void_p), ("nT", ctypes.c_int64)]


class ExtentQuery(ctypes.Structure):
    _fields_ = [("normals", ctypes.c_void_p), ("P", ctypes.c_int64), ("pts", ctypes.c_void_p),
                ("n", ctypes.c_int64)]


class ScorePart(ctypes.Structure):
    _fields_ = [("planes", ctypes.c_void_p), ("P", ctypes.c_int64), ("pts", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("edges", ctypes.c_void_p), ("E", ctypes.c_int64)]


class Timing(ctypes.Structure):
    _fields_ = [("vis_ms", ctypes.c_double), ("vis_launches", ctypes.c_int64),
                ("score_ms", ctypes.c_double), ("score_launches", ctypes.c_int64),
                ("bvh_ms", ctypes.c_double), ("bvh_launches", ctypes.c_int64)]


VACD_OPT_TIMING = 1
VACD_OPT_COUNT_WORK = 2
VACD_OPT_FAN_PAIRS = 4
VACD_OPT_SCORE_CHUNK = 5
VACD_OPT_LANES_PER_PAIR = 6
VACD_OPT_FRONT_LANES = 8
VACD_OPT_FRONT_MAX = 9
VACD_OPT_CLIP_ENDS = 10
VACD_OPT_ROTATED_RAYS = 11


_lib = None


def lib():
    """Load libvacd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m visacd_amd.build`")
        from visacd_amd import share_hip_runtime

        share_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        L.vacd_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.vacd_ctx_destroy.argtypes = [vp]
        L.vacd_last_error.argtypes = [vp]
        L.vacd_last_error.restype = ctypes.c_char_p
        L.vacd_ctx_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.vacd_ctx_set_stream.argtypes = [vp, vp]
        L.vacd_sync.argtypes = [vp]
        L.vacd_ctx_set_option.argtypes = [vp, ctypes.c_int, i64]
        L.vacd_ctx_timing.argtypes = [vp, ctypes.POINTER(Timing), ctypes.c_int]
        L.vacd_bvh_build.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(vp)]
        L.vacd_bvh_build_d.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(vp)]
        L.vacd_bvh_build_planar.argtypes = [vp, vp, i64, vp, i64, vp, vp, ctypes.c_int,
                                            ctypes.POINTER(vp)]
        L.vacd_bvh_destroy.argtypes = [vp]
        L.vacd_bvh_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i32), vp]
        L.vacd_bvh_export.argtypes = [vp, vp, vp]
        L.vacd_bvh_contact_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i32)]
        L.vacd_visibility.argtypes = [vp, vp, i64, vp, vp, vp, vp, i64, ctypes.POINTER(i64)]
        L.vacd_visibility_d.argtypes = [vp, vp, i64, vp, vp, vp, vp, i64, vp]
        L.vacd_visibility_stats.argtypes = [vp, ctypes.POINTER(VisStats)]
        L.vacd_visibility_counters.argtypes = [vp, vp, ctypes.c_int32]
        L.vacd_visibility_deferred.argtypes = [vp, vp, i64, ctypes.POINTER(i64)]
        L.vacd_visibility_batch_d.argtypes = [vp, ctypes.POINTER(Part), ctypes.c_int32, vp, i64,
                                              vp, vp]
        L.vacd_visibility_batch.argtypes = [vp, ctypes.POINTER(Part), ctypes.c_int32, vp, i64,
                                            ctypes.POINTER(i64), vp]
        L.vacd_score_planes.argtypes = [vp, vp, i64, vp, i64, vp, i64, vp, vp,
                                        ctypes.POINTER(i32)]
        L.vacd_score_planes_d.argtypes = [vp, vp, i64, vp, i64, vp, i64, vp, vp, vp]
        L.vacd_score_planes_noalloc_d.argtypes = [vp, vp, i64, vp, i64, vp, i64, vp, vp, vp]
        L.vacd_score_planes_batch.argtypes = [vp, ctypes.POINTER(ScorePart), ctypes.c_int32, vp,
                                              vp]
        L.vacd_score_planes_batch_d.argtypes = [vp, ctypes.POINTER(ScorePart), ctypes.c_int32, vp,
                                                vp]
        L.vacd_max_distance.argtypes = [vp, vp, i64, vp, ctypes.POINTER(ctypes.c_float)]
        L.vacd_max_distance_d.argtypes = [vp, vp, i64, vp, vp]
        L.vacd_max_distance_batch.argtypes = [vp, ctypes.POINTER(DistanceQuery), ctypes.c_int32,
                                              vp]
        L.vacd_plane_extents_batch.argtypes = [vp, ctypes.POINTER(ExtentQuery), ctypes.c_int32, vp]
        L.vacd_sdf_remesh.argtypes = [vp, vp, i64, vp, i64, ctypes.c_double, ctypes.c_double, vp,
                                      i64, vp, i64, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.vacd_sdf_grid.argtypes = [vp, vp, i64, vp, i64, vp, ctypes.c_double, vp, vp, vp]
        ci = ctypes.c_int
        L.vacd_gemm_bf16_d.argtypes = [vp, ci, ci, ci, ci, ci, vp, ci, vp, ci, vp, ci, ci, vp, ci]
        L.vacd_mlp_param_count.restype = i64
        L.vacd_mlp_create.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.vacd_mlp_destroy.argtypes = [vp]
        L.vacd_mlp_set_params.argtypes = [vp, vp, vp]
        L.vacd_mlp_get_params.argtypes = [vp, vp, vp]
        L.vacd_mlp_get_grads.argtypes = [vp, vp, vp]
        L.vacd_mlp_forward_d.argtypes = [vp, vp, ci, ci, vp, ci, vp, vp]
        L.vacd_mlp_forward.argtypes = [vp, vp, ci, ci, vp, ci, vp, vp]
        L.vacd_mlp_train_step_d.argtypes = [vp, vp, ci, ci, vp, ci, vp, vp, ctypes.c_float,
                                            ctypes.POINTER(ctypes.c_float)]
        L.classify_and_rate_planes.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int]
        L.classify_and_rate_planes.restype = None
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):  # a raw device pointer
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


class Context:
    """One vacd context (device + stream + pooled buffers)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        self._check(lib().vacd_ctx_create(device, ctypes.byref(self._h)), None)
        self.device = device

    def _check(self, rc, h="self"):
        if rc != VACD_OK:
            handle = self._h if h == "self" else None
            msg = lib().vacd_last_error(handle).decode(errors="replace")
            raise VacdError(rc, msg)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().vacd_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(lib().vacd_ctx_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def set_stream(self, stream_ptr):
        self._check(lib().vacd_ctx_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))

    def sync(self):
        self._check(lib().vacd_sync(self._h))

    def set_option(self, option: int, value: int):
        self._check(lib().vacd_ctx_set_option(self._h, option, value))

    def timing(self, reset: bool = True) -> dict:
        t = Timing()
        self._check(lib().vacd_ctx_timing(self._h, ctypes.byref(t), 1 if reset else 0))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    # -- BVH -----------------------------------------------------------------
    def bvh(self, vertices, tris, tri_group=None, planes=None) -> "Bvh":
        """BVH from numpy (host) or torch (device) arrays.  tri_group (int8
        per triangle, -1 = none) and planes ((G, 4) float64) give planar
        groups (host arrays only; see vacd_bvh_build_planar)."""
        h = ctypes.c_void_p()
        if tri_group is not None:
            V = np.ascontiguousarray(vertices, dtype=np.float32)
            T = np.ascontiguousarray(tris, dtype=np.uint32)
            g = np.ascontiguousarray(tri_group, dtype=np.int8)
            pl = np.ascontiguousarray(planes, dtype=np.float64).reshape(-1, 4)
            assert g.shape == (T.shape[0],)
            rc = lib().vacd_bvh_build_planar(self._h, _p(V), V.shape[0], _p(T), T.shape[0],
                                             _p(g), _p(pl), pl.shape[0], ctypes.byref(h))
            self._check(rc)
            return Bvh(h)
        if hasattr(vertices, "data_ptr"):
            rc = lib().vacd_bvh_build_d(self._h, _p(vertices), vertices.shape[0], _p(tris),
                                        tris.shape[0], ctypes.byref(h))
        else:
            V = np.ascontiguousarray(vertices, dtype=np.float32)
            T = np.ascontiguousarray(tris, dtype=np.uint32)
            rc = lib().vacd_bvh_build(self._h, _p(V), V.shape[0], _p(T), T.shape[0],
                                      ctypes.byref(h))
        self._check(rc)
        return Bvh(h)

    # -- visibility ----------------------------------------------------------
    def visibility(self, points, self_bvh: "Bvh", cage: "Bvh" = None, new_mask=None,
                   cap: int | None = None):
        """Visibility edges (E, 2) uint32 in k order, from host points."""
        pts = np.ascontiguousarray(points, dtype=np.float32)
        n = pts.shape[0]
        m = None if new_mask is None else np.ascontiguousarray(new_mask, dtype=np.uint8)
        if cap is None:
            cap = max(1, min(n * (n - 1) // 2, 1 << 22))
        while True:
            out = np.empty((cap, 2), dtype=np.uint32)
            cnt = ctypes.c_int64()
            rc = lib().vacd_visibility(self._h, _p(pts), n, _p(m), self_bvh.h,
                                       cage.h if cage is not None else None, _p(out), cap,
                                       ctypes.byref(cnt))
            if rc == VACD_ERR_CAPACITY:
                cap = int(cnt.value)
                continue
            self._check(rc)
            return out[: cnt.value].copy()

    def visibility_d(self, d_pts, n, self_bvh, cage=None, d_mask=None, d_pairs=None, cap=0,
                     d_count=None):
        self._check(lib().vacd_visibility_d(self._h, _p(d_pts), n, _p(d_mask), self_bvh.h,
                                            cage.h if cage is not None else None,
                                            _p(d_pairs), cap, _p(d_count)))

    def visibility_batch_d(self, parts, d_pairs, cap, d_count, d_part_off):
        """parts: list of (d_pts, n, d_mask or None, Bvh, cage Bvh or None)."""
        arr = (Part * len(parts))()
        for t, (p, n, m, b, c) in enumerate(parts):
            arr[t] = Part(_p(p), n, _p(m), b.h, c.h if c is not None else None)
        self._check(lib().vacd_visibility_batch_d(self._h, arr, len(parts), _p(d_pairs), cap,
                                                  _p(d_count), _p(d_part_off)))

    def max_distance(self, points, mesh_bvh: "Bvh") -> float:
        """Largest distance from the host points to the BVH's mesh."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = ctypes.c_float()
        self._check(lib().vacd_max_distance(self._h, _p(pts), pts.shape[0], mesh_bvh.h,
                                            ctypes.byref(out)))
        return float(out.value)

    def max_distance_batch(self, queries) -> np.ndarray:
        """Largest distance of each query's points to its mesh, in one call.
        queries: list of (points [m,3], V [nV,3], T [nT,3]) host arrays."""
        keep, arr = [], (DistanceQuery * max(1, len(queries)))()
        for k, (pts, V, T) in enumerate(queries):
            pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
            V = np.ascontiguousarray(V, dtype=np.float32).reshape(-1, 3)
            T = np.ascontiguousarray(T, dtype=np.uint32).reshape(-1, 3)
            keep += [pts, V, T]
            arr[k] = DistanceQuery(_p(pts), pts.shape[0], _p(V), V.shape[0], _p(T), T.shape[0])
        out = np.zeros(len(queries), dtype=np.float32)
        self._check(lib().vacd_max_distance_batch(self._h, arr, len(queries), _p(out)))
        return out

    def plane_extents_batch(self, queries) -> list:
        """[min, max] of n.p over each query's points for each of its normals.
        queries: list of (normals [P,3], points [n,3]) float64 host arrays."""
        keep, arr, sizes = [], (ExtentQuery * max(1, len(queries)))(), []
        for k, (N, pts) in enumerate(queries):
            N = np.ascontiguousarray(N, dtype=np.float64).reshape(-1, 3)
            pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
            keep += [N, pts]
            arr[k] = ExtentQuery(_p(N), N.shape[0], _p(pts), pts.shape[0])
            sizes.append(N.shape[0])
        out = np.zeros((max(1, sum(sizes)), 2), dtype=np.float64)
        self._check(lib().vacd_plane_extents_batch(self._h, arr, len(queries), _p(out)))
        return np.split(out[: sum(sizes)], np.cumsum(sizes)[:-1])

    def sdf_remesh(self, vertices, tris, scale: float, level_set: float):
        """The surface at distance level_set outside the mesh on a 1 / scale
        grid: (vertices float32 (n, 3), triangles uint32 (m, 3))."""
        V = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        vc, tc = 4 * V.shape[0] + 1024, 4 * T.shape[0] + 2048
        while True:
            vo = np.empty((vc, 3), np.float32)
            to = np.empty((tc, 3), np.uint32)
            nv, nt = ctypes.c_int64(), ctypes.c_int64()
            rc = lib().vacd_sdf_remesh(self._h, _p(V), V.shape[0], _p(T), T.shape[0], scale,
                                       level_set, _p(vo), vc, _p(to), tc, ctypes.byref(nv),
                                       ctypes.byref(nt))
            if rc == VACD_ERR_CAPACITY:
                vc, tc = max(vc, nv.value), max(tc, nt.value)
                continue
            self._check(rc)
            return vo[:nv.value].copy(), to[:nt.value].copy()

    def sdf_grid(self, vertices, tris, lo, h: float, n):
        """Signed distances (negative inside) at lo + (i, j, k) h, as an array
        of shape (n[2], n[1], n[0]), and the sign method (0 parity, 1 winding)."""
        V = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        out = np.empty(int(np.prod(nn)), np.float32)
        mode = ctypes.c_int32()
        self._check(lib().vacd_sdf_grid(self._h, _p(V), V.shape[0], _p(T), T.shape[0], _p(lo), h,
                                        _p(nn), _p(out), ctypes.byref(mode)))
        return out.reshape(nn[2], nn[1], nn[0]), int(mode.value)

    COUNTER_NAMES = ("masked", "blocked", "rotated_rays", "edges", "malformed", "node_visits",
                     "tri_tests", "walked", "exact_calls", "exact_tri_tests", "undecided_no_fan_p",
                     "undecided_fan_p", "undecided_no_fan_q", "undecided_fan_q",
                     "undecided_both_outwards", "clipped_walks")

    def visibility_counters(self) -> dict:
        """Raw work counters of the last visibility call (fan-pass reasons
        only with VACD_OPT_COUNT_WORK)."""
        out = np.zeros(16, dtype=np.int64)
        self._check(lib().vacd_visibility_counters(self._h, _p(out), 16))
        return {k: int(v) for k, v in zip(self.COUNTER_NAMES, out)}

    def visibility_stats(self) -> dict:
        s = VisStats()
        self._check(lib().vacd_visibility_stats(self._h, ctypes.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in VisStats._fields_}

    def visibility_deferred(self) -> np.ndarray:
        """Pair ids k the last visibility call redid with exact predicates."""
        cnt = ctypes.c_int64()
        self._check(lib().vacd_visibility_deferred(self._h, None, 0, ctypes.byref(cnt)))
        out = np.empty(max(1, cnt.value), dtype=np.int64)
        self._check(lib().vacd_visibility_deferred(self._h, _p(out), out.size, ctypes.byref(cnt)))
        return np.sort(out[:cnt.value])

    # -- scoring ---------------------------------------------------------------
    def score_planes(self, planes, points, edges):
        """(score_q32 int64[P], score_f float32[P], argmax) from host arrays."""
        pl = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
        P = pl.shape[0]
        q = np.zeros(P, dtype=np.int64)
        f = np.zeros(P, dtype=np.float32)
        am = ctypes.c_int32(-1)
        self._check(lib().vacd_score_planes(self._h, _p(pl), P, _p(pts), pts.shape[0], _p(e),
                                            e.shape[0], _p(q), _p(f), ctypes.byref(am)))
        return q, f, int(am.value)

    def score_planes_d(self, d_planes, P, d_pts, n, d_edges, E, d_q32=None, d_f=None,
                       d_argmax=None):
        self._check(lib().vacd_score_planes_d(self._h, _p(d_planes), P, _p(d_pts), n,
                                              _p(d_edges), E, _p(d_q32), _p(d_f),
                                              _p(d_argmax)))

    def score_planes_noalloc_d(self, d_planes, P, d_pts, n, d_edges, E, d_q32=None, d_f=None,
                               d_argmax=None):
        """vacd_score_planes_noalloc_d: as score_planes_d, with the library's
        fixed scratch array instead of the context's pool."""
        self._check(lib().vacd_score_planes_noalloc_d(self._h, _p(d_planes), P, _p(d_pts), n,
                                                      _p(d_edges), E, _p(d_q32), _p(d_f),
                                                      _p(d_argmax)))

    def score_planes_batch(self, parts):
        """parts: list of (planes (P, 4), points (n, 3), edges (E, 2)) host arrays.
        Returns (list of int64 score arrays, int32 argmax per part)."""
        keep = []
        arr = (ScorePart * max(1, len(parts)))()
        for k, (pl, pt, ed) in enumerate(parts):
            pl = np.ascontiguousarray(pl, dtype=np.float32).reshape(-1, 4)
            pt = np.ascontiguousarray(pt, dtype=np.float32).reshape(-1, 3)
            ed = np.ascontiguousarray(ed, dtype=np.uint32).reshape(-1, 2)
            keep += [pl, pt, ed]
            arr[k] = ScorePart(_p(pl), pl.shape[0], _p(pt), pt.shape[0], _p(ed), ed.shape[0])
        P = [int(keep[3 * k].shape[0]) for k in range(len(parts))]
        q = np.zeros(max(1, sum(P)), dtype=np.int64)
        am = np.zeros(max(1, len(parts)), dtype=np.int32)
        self._check(lib().vacd_score_planes_batch(self._h, arr, len(parts), _p(q), _p(am)))
        offs = np.cumsum([0] + P)
        return [q[offs[k]:offs[k + 1]].copy() for k in range(len(parts))], am[:len(parts)].copy()

    def score_planes_batch_d(self, parts, d_q32, d_argmax):
        """parts: list of (d_planes, P, d_pts, n, d_edges, E) device pointers/tensors."""
        arr = (ScorePart * max(1, len(parts)))()
        for k, (pl, P, pt, n, ed, E) in enumerate(parts):
            arr[k] = ScorePart(_p(pl), P, _p(pt), n, _p(ed), E)
        self._check(lib().vacd_score_planes_batch_d(self._h, arr, len(parts), _p(d_q32),
                                                    _p(d_argmax)))


    def gemm_bf16_d(self, ta, tb, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, d_bias=None,
                    relu=False, acc=False):
        self._check(lib().vacd_gemm_bf16_d(self._h, int(ta), int(tb), M, N, K, _p(d_A), lda,
                                           _p(d_B), ldb, _p(d_bias), int(relu), int(acc),
                                           _p(d_C), ldc))


class Mlp:
    """The learned plane proposal model of a context (see include/vacd.h)."""

    def __init__(self, ctx: "Context", seed: int = 0):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        ctx._check(lib().vacd_mlp_create(ctx._h, seed, ctypes.byref(self.h)))

    @staticmethod
    def param_count() -> int:
        return int(lib().vacd_mlp_param_count())

    def get(self) -> np.ndarray:
        out = np.empty(self.param_count(), dtype=np.float32)
        self.ctx._check(lib().vacd_mlp_get_params(self.ctx._h, self.h, _p(out)))
        return out

    def grads(self) -> np.ndarray:
        """Gradient of the last training step's loss (parameter layout)."""
        out = np.empty(self.param_count(), dtype=np.float32)
        self.ctx._check(lib().vacd_mlp_get_grads(self.ctx._h, self.h, _p(out)))
        return out

    def set(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.param_count()
        self.ctx._check(lib().vacd_mlp_set_params(self.ctx._h, self.h, _p(p)))

    def forward_d(self, B, E, d_feat, P, d_planes, d_out) -> None:
        self.ctx._check(lib().vacd_mlp_forward_d(self.ctx._h, self.h, B, E, _p(d_feat), P,
                                                 _p(d_planes), _p(d_out)))

    def forward(self, feat: np.ndarray, planes: np.ndarray) -> np.ndarray:
        """Host arrays feat [B][E][7], planes [B][P][4] -> predictions [B][P]."""
        f = np.ascontiguousarray(feat, dtype=np.float32)
        q = np.ascontiguousarray(planes, dtype=np.float32)
        B, E = f.shape[0], f.shape[1]
        P = q.shape[1]
        assert f.shape == (B, E, 7) and q.shape == (B, P, 4)
        out = np.empty((B, P), dtype=np.float32)
        self.ctx._check(lib().vacd_mlp_forward(self.ctx._h, self.h, B, E, _p(f), P, _p(q),
                                               _p(out)))
        return out

    def train_step_d(self, B, E, d_feat, P, d_planes, d_target, lr) -> float:
        loss = ctypes.c_float()
        self.ctx._check(lib().vacd_mlp_train_step_d(self.ctx._h, self.h, B, E, _p(d_feat), P,
                                                    _p(d_planes), _p(d_target), lr,
                                                    ctypes.byref(loss)))
        return float(loss.value)

    def close(self):
        if self.h:
            lib().vacd_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlp_unpack(p: np.ndarray) -> dict:
    """Named views of a parameter vector (layout of the C side)."""
    D = 256
    shapes = [("W1", (D, 8)), ("b1", (D,)), ("W2", (D, D)), ("b2", (D,)), ("W3", (D, D)),
              ("b3", (D,)), ("Wz", (D, D)), ("Wp", (D, 4)), ("c1", (D,)), ("V2", (D, D)),
              ("c2", (D,)), ("v3", (D,)), ("c3", (1,))]
    out, o = {}, 0
    for name, shp in shapes:
        k = int(np.prod(shp))
        out[name] = p[o:o + k].reshape(shp)
        o += k
    assert o == p.size
    return out


class Bvh:
    def __init__(self, h):
        self.h = h

    def info(self):
        n = ctypes.c_int64()
        d = ctypes.c_int32()
        b = np.zeros(6, dtype=np.float32)
        rc = lib().vacd_bvh_info(self.h, ctypes.byref(n), ctypes.byref(d), _p(b))
        if rc != VACD_OK:
            raise VacdError(rc, lib().vacd_last_error(None).decode())
        return {"n_tris": n.value, "depth": d.value, "bounds": b}

    def contact_info(self):
        """(primitives, depth) of the contact tree; (0, 0) when not built."""
        n, d = ctypes.c_int64(), ctypes.c_int32()
        rc = lib().vacd_bvh_contact_info(self.h, ctypes.byref(n), ctypes.byref(d))
        if rc != VACD_OK:
            raise VacdError(rc, lib().vacd_last_error(None).decode())
        return int(n.value), int(d.value)

    def export(self):
        """(nodes (N-1, 16) float32 view, tris (N, 12) float32 view)."""
        nt = self.info()["n_tris"]
        nodes = np.zeros((max(nt - 1, 1), 16), dtype=np.float32)
        tris = np.zeros((nt, 12), dtype=np.float32)
        rc = lib().vacd_bvh_export(self.h, _p(nodes), _p(tris))
        if rc != VACD_OK:
            raise VacdError(rc, lib().vacd_last_error(None).decode())
        return nodes, tris

    def close(self):
        if self.h:
            lib().vacd_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def classify_and_rate_planes(planes, points, edges):
    """Reference-signature shim: float32 scores, synchronous."""
    pl = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(pl.shape[0], dtype=np.float32)
    lib().classify_and_rate_planes(_p(pl), _p(pts), _p(e), _p(out), pl.shape[0], pts.shape[0],
                                   e.shape[0])
    return out
