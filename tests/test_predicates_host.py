"""The engine's device predicates (visacd_amd/csrc/predicates.hpp), built for
the host, against the oracle's exact predicates (themselves pinned against
rational arithmetic in tests/test_oracle_predicates.py): the complete form,
the integer-grid path and the expansion path of orient3d, orient3d with a
midpoint, orient2d and orient2d with a midpoint, on generic inputs, lattice
inputs with exact zeros and inputs whose exponents spread beyond the integer
path's 15-bit window.  A second build of the same check with the oracle runs
under AddressSanitizer and UndefinedBehaviorSanitizer."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "predicates_host.cpp")
FLAGS = ["-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


@pytest.fixture(scope="module")
def ph(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ph") / "libph.so")
    subprocess.run(["g++", *FLAGS, "-shared", "-fPIC", SRC, "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.ph_eval.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                          ctypes.c_void_p]
    return L


def _cases(kind, stride, n, seed):
    rng = np.random.default_rng(seed)
    generic = rng.uniform(-1, 1, (n, stride))
    lattice = rng.integers(-8, 9, (n, stride)) / 8.0
    wide = rng.uniform(-1, 1, (n, stride)) * np.exp2(rng.integers(-40, 13, (n, stride)))
    # nearly degenerate: last point (or points) on the plane/line of the first, rounded
    near = rng.uniform(-1, 1, (n, stride))
    base = near[:, :3]
    e1, e2 = near[:, 3:6] - base, near[:, 6:9] - base
    s, t = rng.uniform(-1, 1, (n, 1)), rng.uniform(-1, 1, (n, 1))
    near[:, 9:12] = base + s * e1 + t * e2 if stride >= 12 else near[:, 9:12]
    # exactly degenerate: the last points on the line through the first two
    # (lattice values, so every coordinate is exact): the sign is 0
    degen = rng.integers(-8, 9, (n, stride)) / 8.0
    last = 3 if kind in (0, 2) else 6  # the last point, or both midpoint ends
    for j in range(stride // 3 - last // 3, stride // 3):
        m = rng.integers(-3, 4, (n, 1))
        degen[:, 3 * j:3 * j + 3] = degen[:, 0:3] + m * (degen[:, 3:6] - degen[:, 0:3])
    return np.concatenate([generic, lattice, wide, near, degen]).astype(np.float32)


@pytest.mark.parametrize("kind,stride", [(0, 12), (1, 15), (2, 9), (3, 12)])
def test_device_predicates_every_path_equals_oracle(ph, kind, stride):
    x = _cases(kind, stride, 20000, 100 + kind)
    for k in (range(3) if kind >= 2 else [0]):
        n = x.shape[0]
        got = np.zeros(3 * n, np.int32)
        ph.ph_eval(kind, n, x.ctypes.data, k, got.ctypes.data)
        got = got.reshape(n, 3)
        want = oracle.predicates_batch(kind, x, k)
        np.testing.assert_array_equal(got[:, 0], want)
        np.testing.assert_array_equal(got[:, 2], want)
        ok = got[:, 1] != 2
        np.testing.assert_array_equal(got[ok, 1], want[ok])
        assert ok.sum() > n // 2 and (~ok).sum() > n // 8  # both paths exercised
        assert (want == 0).sum() > 10000  # exact zeros among the cases


def test_predicates_and_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "ph_check")
    cmd = ["g++", *FLAGS, "-g", "-O1", "-fopenmp", "-DPH_MAIN", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           SRC, oracle.SRC, "-o", exe]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, env=env)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    assert "mismatches 0" in r.stdout
