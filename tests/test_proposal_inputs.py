"""Inputs of the learned plane proposal (host code shared by the driver and the
trainer): the part frame, the edge sample and the plane transform."""
import numpy as np
import pytest

import visacd  # noqa: F401  (loads the host driver)
from visacd_amd import _visacd


def _inputs(v, e, pl, seed=5):
    return _visacd.proposal_inputs(np.asarray(v, float), np.asarray(e, np.uint32),
                                   np.asarray(pl, float), seed)


def test_frame_and_features_on_a_hand_example():
    v = [[0, 0, 0], [2, 1, 1], [1, 0.5, 0], [0, 1, 1]]
    feat, q, fr = _inputs(v, [[1, 0], [2, 3]], [[2, 0, 0, -2], [0, 0, 1, 0]])
    # bbox [0,2]x[0,1]x[0,1]: centre (1, .5, .5), half the longest side 1
    assert fr == (1.0, 0.5, 0.5, 1.0)
    # edge (1, 0): endpoints ordered lexicographically in the frame
    np.testing.assert_allclose(feat[0], [-1, -0.5, -0.5, 1, 0.5, 0.5, np.sqrt(6)], rtol=1e-6)
    np.testing.assert_allclose(feat[1], [-1, 0.5, 0.5, 0, 0, -0.5, 1.5], rtol=1e-6)
    # fewer than 512 edges: the list repeats cyclically
    assert feat.shape == (512, 7)
    np.testing.assert_array_equal(feat[2::2], np.repeat(feat[:1], 255, 0))
    np.testing.assert_array_equal(feat[3::2], np.repeat(feat[1:2], 255, 0))
    # x = 1 is the frame's x' = 0 (scaled input normalised); z = 0 is z' = -0.5
    np.testing.assert_allclose(q, [[1, 0, 0, 0], [0, 0, 1, 0.5]], atol=1e-7)
    # the same planes with the opposite orientation give the same inputs
    _, q2, _ = _inputs(v, [[1, 0], [2, 3]], [[-1, 0, 0, 1], [0, 0, -3, 0]])
    np.testing.assert_allclose(q2, q, atol=1e-7)


def test_stratified_sample_takes_one_edge_per_stratum():
    rng = np.random.default_rng(0)
    n = 200
    v = rng.uniform(-1, 1, (n, 3))
    E = 512 * 3 + 100
    e = np.stack([rng.integers(0, n, E), rng.integers(0, n, E)], 1)
    e[:, 1] = np.where(e[:, 1] == e[:, 0], (e[:, 0] + 1) % n, e[:, 1])
    feat, _, fr = _inputs(v, e, np.zeros((0, 4)), seed=11)
    c, s = np.array(fr[:3]), fr[3]
    u = (v[e[:, 0]] - c) / s
    w = (v[e[:, 1]] - c) / s
    for k in range(512):
        a, b = k * E // 512, (k + 1) * E // 512
        # the drawn edge is one of the edges of stratum k (either orientation)
        fwd = np.abs(np.concatenate([u[a:b], w[a:b]], 1) - feat[k, :6]).max(1)
        bwd = np.abs(np.concatenate([w[a:b], u[a:b]], 1) - feat[k, :6]).max(1)
        assert min(fwd.min(), bwd.min()) < 1e-6, k
        np.testing.assert_allclose(feat[k, 6], np.linalg.norm(feat[k, 3:6] - feat[k, :3]),
                                   rtol=1e-6)
        assert tuple(feat[k, :3]) <= tuple(feat[k, 3:6])
    # the draw depends on the seed
    feat2, _, _ = _inputs(v, e, np.zeros((0, 4)), seed=12)
    assert not np.array_equal(feat, feat2)


def test_inputs_are_invariant_to_moving_and_scaling_the_part():
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, (50, 3))
    e = np.stack([np.arange(0, 49), np.arange(1, 50)], 1)
    n = rng.standard_normal((20, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pl = np.concatenate([n, rng.uniform(-0.5, 0.5, (20, 1))], 1)
    f0, q0, _ = _inputs(v, e, pl)
    shift, scale = np.array([3.0, -2.0, 0.5]), 7.0
    v1 = v * scale + shift
    # the same planes in the moved frame: n.x1 + d1 = 0 with d1 = scale d - n.shift
    pl1 = np.concatenate([n, (scale * pl[:, 3] - n @ shift)[:, None]], 1)
    f1, q1, _ = _inputs(v1, e, pl1)
    np.testing.assert_allclose(f1, f0, atol=1e-5)
    np.testing.assert_allclose(q1, q0, atol=1e-5)


def test_rejects_out_of_range_edges():
    with pytest.raises(IndexError):
        _inputs([[0, 0, 0], [1, 1, 1]], [[0, 2]], [[1, 0, 0, 0]])
