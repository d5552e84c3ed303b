"""The C ABI library loads and exports every function include/vacd.h declares."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "vacd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(vacd_\w+|classify_and_rate_planes)\s*\(", txt))
    return sorted(names)


def test_header_symbols_exported():
    from visacd_amd import native

    lib = ctypes.CDLL(native.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []
    assert sorted(native.EXPORTED) == declared


def test_error_convention_without_device():
    """Context creation on a machine without a GPU fails with a message, not a crash."""
    from visacd_amd import native

    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(native.VacdError) as ei:
        native.Context(0)
    assert ei.value.code in (native.VACD_ERR_HIP, native.VACD_ERR_INVALID)


def test_workloads_are_deterministic():
    from visacd_amd import workloads as wl

    a = wl.icosphere_c2()
    b = wl.icosphere_c2()
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.planes, b.planes)
    assert a.vertices.shape == (40962, 3) and a.triangles.shape == (81920, 3)
    assert a.n_points == 4096 and a.n_segments == 8386560
    t = wl.torus_c1()
    assert t.vertices.shape == (1152, 3) and t.n_segments == 662976
    lo, hi = t.vertices.min(0), t.vertices.max(0)
    assert np.isclose((hi - lo).max(), 2.0, atol=1e-6)  # reference normalisation
