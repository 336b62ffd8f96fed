"""Build the native libraries of ldsc_amd in-tree.

* libldsc_l2.so   -- the HIP engine for gfx950 behind include/ldsc_l2.h
* libldsc_synth.so -- the seeded synthetic PLINK generator (host C, OpenMP)
* libldsc_calib.so -- measured MFMA and HBM copy peaks for bench.py's roofline
* libldsc_io.so    -- the host BIM reader and .l2.ldscore.gz writer (include/ldsc_io.h)

Run ``python -m ldsc_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("LDSC_OFFLOAD_ARCH", "gfx950")

ENGINE_SOURCES = ["host_windows.cpp", "engine.hip", "sketch.hip", "server.cpp", "noalloc.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-Wall", "-Wno-unused-result",
    # keep the compiler from forming packed FP32 ops in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libldsc_l2.so")
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs, "-lpthread"])
    os.replace(tmp, out)
    return out


def build_synth() -> str:
    out = os.path.join(HERE, "libldsc_synth.so")
    _run(["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-o", out,
          os.path.join(CSRC, "synth.c"), "-lm"])
    return out


def build_calib() -> str:
    out = os.path.join(HERE, "libldsc_calib.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "calib.hip")])
    os.replace(tmp, out)
    return out


def build_io() -> str:
    out = os.path.join(HERE, "libldsc_io.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall", "-o", tmp,
          os.path.join(CSRC, "io.cpp"), "-lz"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_synth()
    build_io()
    build_engine()
    build_calib()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
