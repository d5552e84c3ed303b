"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

The reference ships no golden vectors for this path (no tests, no expected
outputs), so the fixtures are produced by the oracle after it has been
pinned by tests/test_oracle_*.py, and freeze its answers for regression
checks of both the oracle and the HIP engine.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import meshes  # noqa: E402
import oracle  # noqa: E402
from visacd_amd import workloads as wl  # noqa: E402


def k_of(edges, n):
    i = edges[:, 0].astype(np.int64)
    j = edges[:, 1].astype(np.int64)
    return i * (2 * n - i - 1) // 2 + j - i - 1


def bitmask(edges, n):
    S = n * (n - 1) // 2
    bits = np.zeros(S, dtype=bool)
    bits[k_of(edges, n)] = True
    return np.packbits(bits)


def main():
    W = wl.torus_c1()
    e, st = oracle.visibility(W.points, W.vertices, W.triangles)
    q = oracle.score_q32(W.planes, W.points, e)
    np.savez_compressed(
        os.path.join(HERE, "torus_c1.npz"),
        edge_bits=bitmask(e, W.n_points), n=W.n_points, n_edges=len(e),
        blocked=st["blocked"], scores_q32=q, argmax=oracle.argmax_i64(q),
        vertices_sha=np.frombuffer(W.vertices.tobytes()[:64], np.uint8))
    # cage mode on a small perturbed icosphere with a radially scaled cage
    V, T = wl.perturbed_icosphere(3, 0.3, 11)
    cage = meshes.cage_of(V, 1.06)
    ec, _ = oracle.visibility(V, V, T, cage=(cage, T))
    ep, _ = oracle.visibility(V, V, T)
    np.savez_compressed(os.path.join(HERE, "ico3_cage.npz"), n=V.shape[0],
                        cage_bits=bitmask(ec, V.shape[0]), parity_bits=bitmask(ep, V.shape[0]))
    print("torus edges", len(e), "argmax", oracle.argmax_i64(q), "cage", len(ec), "parity", len(ep))


if __name__ == "__main__":
    main()
