"""Error paths of the C ABI on the GPU.

* A result larger than the caller's buffer: VACD_ERR_CAPACITY, with the
  total still reported, the first `cap` edges (k order) written, and the
  context usable afterwards -- single-part and batched visibility, remesh.
* Inputs beyond a documented limit: VACD_ERR_LIMIT, raised by the host-side
  argument checks before any device work (points per part, grid nodes per
  axis), and the context usable afterwards.
"""
import ctypes

import numpy as np
import pytest

import oracle
from visacd_amd import native
from visacd_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _mesh():
    V, T = wl.perturbed_icosphere(2, 0.3, 5)
    return np.ascontiguousarray(V, np.float32), np.ascontiguousarray(T, np.uint32)


def test_visibility_capacity_error_reports_total_and_prefix(ctx):
    V, T = _mesh()
    b = ctx.bvh(V, T)
    o, _ = oracle.visibility(V, V, T)
    cap = o.shape[0] // 3
    out = np.zeros((cap, 2), np.uint32)
    cnt = ctypes.c_int64(-1)
    print("capacity", V.shape[0], "edges", o.shape[0], "cap", cap, flush=True)
    rc = native.lib().vacd_visibility(ctx.handle, native._p(V), V.shape[0], None, b.h, None,
                                      native._p(out), cap, ctypes.byref(cnt))
    assert rc == native.VACD_ERR_CAPACITY
    assert b"more edges" in native.lib().vacd_last_error(ctx.handle)
    assert cnt.value == o.shape[0]
    np.testing.assert_array_equal(out, o[:cap])
    # zero capacity: only the count
    rc = native.lib().vacd_visibility(ctx.handle, native._p(V), V.shape[0], None, b.h, None,
                                      None, 0, ctypes.byref(cnt))
    assert rc == native.VACD_ERR_CAPACITY and cnt.value == o.shape[0]
    np.testing.assert_array_equal(ctx.visibility(V, b), o)  # the context still works


def test_visibility_batch_capacity_error_reports_offsets(ctx):
    V, T = _mesh()
    V2, T2 = wl.torus(nu=16, nv=8, seed=3)
    V2 = np.ascontiguousarray(V2, np.float32)
    T2 = np.ascontiguousarray(T2, np.uint32)
    b1, b2 = ctx.bvh(V, T), ctx.bvh(V2, T2)
    o1, _ = oracle.visibility(V, V, T)
    o2, _ = oracle.visibility(V2, V2, T2)
    parts = (native.Part * 2)()
    parts[0] = native.Part(native._p(V), V.shape[0], None, b1.h, None)
    parts[1] = native.Part(native._p(V2), V2.shape[0], None, b2.h, None)
    total = o1.shape[0] + o2.shape[0]
    cap = o1.shape[0] + o2.shape[0] // 2
    out = np.zeros((cap, 2), np.uint32)
    off = np.zeros(3, np.int64)
    cnt = ctypes.c_int64(-1)
    print("batch capacity", total, cap, flush=True)
    rc = native.lib().vacd_visibility_batch(ctx.handle, parts, 2, native._p(out), cap,
                                            ctypes.byref(cnt), native._p(off))
    assert rc == native.VACD_ERR_CAPACITY
    assert cnt.value == total
    assert off.tolist() == [0, o1.shape[0], total]
    np.testing.assert_array_equal(out[: o1.shape[0]], o1)
    np.testing.assert_array_equal(out[o1.shape[0]:], o2[: cap - o1.shape[0]])
    big = np.zeros((total, 2), np.uint32)
    rc = native.lib().vacd_visibility_batch(ctx.handle, parts, 2, native._p(big), total,
                                            ctypes.byref(cnt), native._p(off))
    assert rc == native.VACD_OK
    np.testing.assert_array_equal(big, np.concatenate([o1, o2]))


def test_remesh_capacity_error_reports_sizes(ctx):
    V, T = _mesh()
    V = wl.normalize_like_reference(V).astype(np.float32)
    Vo, To = ctx.sdf_remesh(V, T, 30.0, 0.55 / 30)
    nv, nt = ctypes.c_int64(-1), ctypes.c_int64(-1)
    vbuf = np.zeros((8, 3), np.float32)
    tbuf = np.zeros((8, 3), np.uint32)
    print("remesh capacity", Vo.shape, To.shape, flush=True)
    rc = native.lib().vacd_sdf_remesh(ctx.handle, native._p(V), V.shape[0], native._p(T),
                                      T.shape[0], 30.0, 0.55 / 30, native._p(vbuf), 8,
                                      native._p(tbuf), 8, ctypes.byref(nv), ctypes.byref(nt))
    assert rc == native.VACD_ERR_CAPACITY
    assert (nv.value, nt.value) == (Vo.shape[0], To.shape[0])
    Vb, Tb = ctx.sdf_remesh(V, T, 30.0, 0.55 / 30)  # deterministic, and the context works
    np.testing.assert_array_equal(Vb, Vo)
    np.testing.assert_array_equal(Tb, To)


def test_limits_are_checked_before_device_work(ctx):
    V, T = _mesh()
    b = ctx.bvh(V, T)
    cnt = ctypes.c_int64(-1)
    out = np.zeros((4, 2), np.uint32)
    print("limits", flush=True)
    # more points than a pair id can address (the points are never read)
    rc = native.lib().vacd_visibility(ctx.handle, native._p(V), (1 << 26) + 1, None, b.h, None,
                                      native._p(out), 4, ctypes.byref(cnt))
    assert rc == native.VACD_ERR_LIMIT
    assert b"2^26" in native.lib().vacd_last_error(ctx.handle)
    parts = (native.Part * 1)()
    parts[0] = native.Part(native._p(V), (1 << 26) + 1, None, b.h, None)
    off = np.zeros(2, np.int64)
    rc = native.lib().vacd_visibility_batch(ctx.handle, parts, 1, native._p(out), 4,
                                            ctypes.byref(cnt), native._p(off))
    assert rc == native.VACD_ERR_LIMIT
    # a signed-distance grid beyond 2048 nodes per axis
    with pytest.raises(native.VacdError) as ei:
        ctx.sdf_grid(V, T, np.zeros(3), 0.01, np.array([4096, 4, 4], np.int32))
    assert ei.value.code == native.VACD_ERR_LIMIT
    o, _ = oracle.visibility(V, V, T)
    np.testing.assert_array_equal(ctx.visibility(V, b), o)  # the context still works
