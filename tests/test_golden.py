"""Golden fixtures (tests/golden/*.npz) against the oracle (CPU) and the HIP engine (GPU)."""
import os

import numpy as np
import pytest

import meshes
import oracle
from visacd_amd import workloads as wl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _edges_from_bits(bits, n):
    S = n * (n - 1) // 2
    ks = np.flatnonzero(np.unpackbits(bits)[:S])
    out = np.zeros((len(ks), 2), np.uint32)
    for t, k in enumerate(ks):
        out[t] = oracle.k_to_ij(int(k), n)
    return out


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_torus_fixture_oracle():
    g = _load("torus_c1.npz")
    W = wl.torus_c1()
    assert np.array_equal(np.frombuffer(W.vertices.tobytes()[:64], np.uint8), g["vertices_sha"])
    e, st = oracle.visibility(W.points, W.vertices, W.triangles)
    np.testing.assert_array_equal(e, _edges_from_bits(g["edge_bits"], int(g["n"])))
    assert st["blocked"] == int(g["blocked"])
    q = oracle.score_q32(W.planes, W.points, e)
    np.testing.assert_array_equal(q, g["scores_q32"])
    assert oracle.argmax_i64(q) == int(g["argmax"])


def test_cage_fixture_oracle():
    g = _load("ico3_cage.npz")
    V, T = wl.perturbed_icosphere(3, 0.3, 11)
    ec, _ = oracle.visibility(V, V, T, cage=(meshes.cage_of(V, 1.06), T))
    np.testing.assert_array_equal(ec, _edges_from_bits(g["cage_bits"], int(g["n"])))


@pytest.mark.gpu
def test_torus_fixture_gpu(ctx):
    g = _load("torus_c1.npz")
    W = wl.torus_c1()
    print("golden torus on GPU", flush=True)
    b = ctx.bvh(W.vertices, W.triangles)
    e = ctx.visibility(W.points, b)
    np.testing.assert_array_equal(e, _edges_from_bits(g["edge_bits"], int(g["n"])))
    q, f, am = ctx.score_planes(W.planes, W.points, e)
    np.testing.assert_array_equal(q, g["scores_q32"])
    assert am == int(g["argmax"])


@pytest.mark.gpu
def test_cage_fixture_gpu(ctx):
    g = _load("ico3_cage.npz")
    V, T = wl.perturbed_icosphere(3, 0.3, 11)
    print("golden cage on GPU", flush=True)
    b = ctx.bvh(V, T)
    cb = ctx.bvh(meshes.cage_of(V, 1.06), T)
    np.testing.assert_array_equal(ctx.visibility(V, b, cage=cb),
                                  _edges_from_bits(g["cage_bits"], int(g["n"])))
    np.testing.assert_array_equal(ctx.visibility(V, b),
                                  _edges_from_bits(g["parity_bits"], int(g["n"])))
