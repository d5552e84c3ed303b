"""Build recipe for the native parts of visacd_amd.

* ``libvacd.so``  -- the HIP kernels and the C ABI of ``include/vacd.h``,
  compiled for gfx950 only.
* ``_visacd*.so`` -- the pybind11 module with the reference's Python surface
  (``visacd.process``, ``Mesh``, ``Config``, ...), linked against libvacd.

Run ``python -m visacd_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBVACD = os.path.join(HERE, "libvacd.so")

HIP_SOURCES = ["vacd_api.hip", "lbvh.hip", "visibility.hip", "scoring.hip", "distance.hip", "mlp.hip"]
HOST_SOURCES = []  # filled in by the host driver sources when present

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE sqrtf / division so float32 edge lengths match the host bit for bit
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_libvacd(jobs: int = 4) -> str:
    objs = []
    procs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", LIBVACD])
    return LIBVACD


def build_module() -> str | None:
    """pybind11 module ``visacd_amd/_visacd`` (host C++ linked to libvacd)."""
    src = os.path.join(CSRC, "host", "bindings.cpp")
    if not os.path.exists(src):
        return None
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_visacd" + suffix)
    host_srcs = sorted(
        os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
        if f.endswith(".cpp"))
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host"),
           *host_srcs, "-L" + HERE, "-lvacd", "-Wl,-rpath,$ORIGIN", "-o", out]
    _run(cmd)
    return out


def build_all() -> None:
    build_libvacd()
    build_module()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
