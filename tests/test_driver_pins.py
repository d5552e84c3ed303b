"""Pins of the driver's candidate sampling and plane choice (CPU).

* std::mt19937, the reference's random_engine (lib/include/core.hpp:13): the
  C++ standard's known answer (the 10,000th output of a default-constructed
  engine is 4123659995, [rand.predef]) and agreement with numpy's
  independent MT19937 under the same (legacy) seeding.
* std::uniform_int_distribution<int> over a power-of-two range 2^k draws
  the top k bits of one engine output (every libstdc++ algorithm, with no
  rejection for such ranges): a closed form for the index draws of
  lib/src/process.cpp:38-43.
* The oracle's bisector sampler against that closed form (first plane =
  bisector of the edge the first draw names), and the driver's sampler
  against the oracle's, bit for bit.
* The driver's plane choice: first maximum, flat-surface planes times
  config.flat_surface_k (lib/src/process.cpp:128-134).
"""
import numpy as np
import pytest

import oracle
from visacd_amd import workloads as wl


def test_mt19937_known_answer():
    out = oracle.mt19937(5489, 10000)  # 5489: the default seed
    assert int(out[-1]) == 4123659995


def test_mt19937_matches_numpy_for_seed_42():
    bg = np.random.MT19937(0)
    bg._legacy_seeding(42)
    ref = np.array([bg.random_raw() for _ in range(2000)], dtype=np.uint64)
    np.testing.assert_array_equal(oracle.mt19937(42, 2000).astype(np.uint64), ref)


@pytest.mark.parametrize("k", [1, 5, 12])
def test_uniform_int_power_of_two_is_top_bits(k):
    raw = oracle.mt19937(7, 500).astype(np.uint64)
    draws = oracle.uniform_int(7, 0, (1 << k) - 1, 500)
    np.testing.assert_array_equal(draws.astype(np.uint64), raw >> np.uint64(32 - k))


def _edges_mesh(n_edges_pow2):
    V, T = wl.perturbed_icosphere(3, 0.3, 3)
    e, _ = oracle.visibility(V, V, T)
    assert e.shape[0] >= n_edges_pow2
    return V.astype(np.float64), e[:n_edges_pow2]


def test_oracle_bisector_first_plane_closed_form():
    """With E = 2^k edges the first draw names edge (first output >> (32-k));
    the first plane is that edge's perpendicular bisector: unit normal
    (p2 - p1)/|p2 - p1| (component-wise division), d = -n . (p1 + p2)/2."""
    V, e = _edges_mesh(256)
    for seed in (1, 42, 12345):
        pl = oracle.bisector_planes(V, e, 64, seed)
        idx = int(oracle.mt19937(seed, 1)[0]) >> (32 - 8)
        p1, p2 = V[e[idx, 0]], V[e[idx, 1]]
        d = p2 - p1
        l = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        n = np.array([d[0] / l, d[1] / l, d[2] / l])
        m = (p1 + p2) * 0.5
        assert np.array_equal(pl[0], np.array([n[0], n[1], n[2], -(n[0] * m[0] + n[1] * m[1] + n[2] * m[2])]))


@pytest.mark.parametrize("E,P,seed", [(256, 64, 1), (1000, 2500, 42), (7, 50, 9), (300, 40, 77)])
def test_driver_bisector_planes_equal_oracle(E, P, seed):
    from visacd_amd import _visacd

    V, e = _edges_mesh(E)
    a = np.asarray(_visacd.bisector_planes(V, e, P, seed), dtype=np.float64).reshape(-1, 4)
    b = oracle.bisector_planes(V, e, P, seed)
    assert a.shape == b.shape and a.shape[0] >= 1
    np.testing.assert_array_equal(a, b)
    if E <= 7:  # few edges: the 5 x P draw bound and the duplicate test end the loop
        assert a.shape[0] <= E


def test_select_plane_first_max_and_flat_factor():
    from visacd_amd import _visacd

    q = lambda *v: [int(x * 2**32) for x in v]  # noqa: E731
    k0 = _visacd.config.flat_surface_k
    try:
        _visacd.config.flat_surface_k = 2.0
        assert _visacd.select_plane(q(1.0, 3.0, 2.0), 2) == 2  # flat 2.0 x 2 beats 3.0
        assert _visacd.select_plane(q(1.0, 3.0, 1.5), 2) == 1  # 1.5 x 2 ties 3.0: first wins
        assert _visacd.select_plane(q(5.0, 5.0), 2) == 0  # ties: first maximum
        assert _visacd.select_plane(q(1.0, 3.0, 2.0), 3) == 1  # no flat planes
        _visacd.config.flat_surface_k = 1.0
        assert _visacd.select_plane(q(1.0, 3.0, 2.0), 2) == 1
        assert _visacd.select_plane([], 0) == -1
    finally:
        _visacd.config.flat_surface_k = k0
