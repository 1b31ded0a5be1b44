"""Build recipe for the HIP library (gfx950) -- run as ``python -m tissuefolding_paperversion_amd.build``.

Produces ``tissuefolding_paperversion_amd/lib/libtfnr.so`` in-tree with hipcc, the C++ driver
``lib/tf_driver`` (g++) and the copy-kernel calibration program ``lib/tf_copy_calib`` (hipcc).  Packed FP32
instructions are disabled for all device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libtfnr.so")
SOURCES = ["tf_api.hip", "tf_assemble.hip", "tf_krylov.hip", "tf_adjoint.hip", "tf_packing.hip", "tf_growth.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TF_OFFLOAD_ARCH", "gfx950")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
            "-Wno-unused-variable", "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + flags() + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtfnr.so")
    os.replace(tmp, LIB)
    # C++ host driver over the facade header (plain g++, links the library above)
    drv = os.path.join(LIBDIR, "tf_driver")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "tf_driver.cpp"), "-L" + LIBDIR, "-ltfnr",
           "-Wl,-rpath,$ORIGIN", "-o", drv]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building tf_driver")
    # copy kernel of known size, the calibration for the profiler's HBM byte counters
    cal = os.path.join(LIBDIR, "tf_copy_calib")
    cmd = [HIPCC] + [f for f in flags() if f not in ("-shared", "-fPIC")] + [os.path.join(CSRC, "tf_copy_calib.hip"),
                                                                       "-o", cal]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building tf_copy_calib")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
