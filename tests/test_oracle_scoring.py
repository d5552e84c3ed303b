"""Pins the oracle's plane-score definitions (FP64 and float32/Q32.32)."""
import numpy as np

import oracle
from visacd_amd import workloads as wl


def test_axis_plane_closed_form():
    pts = np.array([[-1, 0, 0], [1, 0, 0], [-0.5, 1, 0], [0.25, 1, 0], [0, 2, 0],
                    [3e-7, 5, 0], [-1, 5, 0]], np.float32)
    # edges: crosses (len 2), crosses (len 0.75), touches plane (x=0),
    # inside the dead zone (|x| < 1e-6): none count except the first two
    edges = np.array([[0, 1], [2, 3], [4, 1], [5, 6]], np.uint32)
    planes = np.array([[1, 0, 0, 0]], np.float32)
    assert oracle.score_fp64(planes, pts, edges)[0] == 2.75
    assert oracle.score_q32(planes, pts, edges)[0] == int(2.75 * 2 ** 32)


def test_q32_agrees_with_fp64():
    W = wl.torus_c1()
    e, _ = oracle.visibility(W.points, W.vertices, W.triangles)
    q = oracle.score_q32(W.planes, W.points, e)
    d = oracle.score_fp64(W.planes, W.points, e)
    qf = q.astype(np.float64) * 2.0 ** -32
    big = d > 1e-3 * d.max()
    assert big.sum() > 100
    np.testing.assert_allclose(qf[big], d[big], rtol=1e-5)


def test_q32_order_independent():
    W = wl.torus_c1()
    e, _ = oracle.visibility(W.points, W.vertices, W.triangles)
    rng = np.random.default_rng(0)
    q1 = oracle.score_q32(W.planes[:32], W.points, e)
    q2 = oracle.score_q32(W.planes[:32], W.points, e[rng.permutation(len(e))])
    q3 = oracle.score_q32(W.planes[:32], W.points, e[:, ::-1])  # endpoint order
    np.testing.assert_array_equal(q1, q2)
    np.testing.assert_array_equal(q1, q3)


def test_argmax_first_maximum():
    assert oracle.argmax_i64(np.array([3, 7, 7, 1], np.int64)) == 1
    assert oracle.argmax_i64(np.array([5], np.int64)) == 0


def test_bisector_planes_properties():
    W = wl.torus_c1()
    e, _ = oracle.visibility(W.points, W.vertices, W.triangles)
    V64 = W.points.astype(np.float64)
    pl = oracle.bisector_planes(V64, e, 200, 7)
    assert 0 < len(pl) <= 200
    np.testing.assert_allclose(np.linalg.norm(pl[:, :3], axis=1), 1.0, rtol=1e-12)
    cos5 = np.cos(np.deg2rad(5.0))
    for a in range(len(pl)):
        dots = np.abs(pl[a + 1:, :3] @ pl[a, :3])
        close = (dots > cos5) & (np.abs(pl[a + 1:, 3] - pl[a, 3]) < 1e-3)
        assert not close.any()
    # deterministic for a seed, different for another
    np.testing.assert_array_equal(pl, oracle.bisector_planes(V64, e, 200, 7))
    assert not np.array_equal(pl, oracle.bisector_planes(V64, e, 200, 8))
    # every plane bisects one visibility edge
    mids = (V64[e[:, 0]] + V64[e[:, 1]]) / 2
    for p in pl[:20]:
        assert np.min(np.abs(mids @ p[:3] + p[3])) < 1e-12
