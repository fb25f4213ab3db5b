"""Build recipe of the native libraries (no CMake): g++ for the host-only libraries and
hipcc for the gfx950 HIP library. Outputs are written in-tree so they travel with the
sources; they are git-ignored."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "allen_amd")
CSRC = os.path.join(PKG, "csrc")
HIP_LIB = os.path.join(PKG, "libvelo_amd.so")
SYNTH_LIB = os.path.join(PKG, "synth", "libvelo_synth.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["velo_api.hip", "velo_clustering.hip", "velo_retina.hip", "velo_search.hip", "velo_consolidate.hip",
               "velo_scorer.hip", "velo_pv.hip", "velo_microbench.hip", "velo_mdf.cpp"]
# -ffp-contract=off: every float expression is evaluated as written (bit parity with the
# CPU target). Packed-FP32 instructions are disabled for the device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_synth():
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", SYNTH_LIB,
          os.path.join(PKG, "synth", "velo_synth.cpp"), "-lpthread"])


def build_hip():
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", tmp] + srcs)
    os.replace(tmp, HIP_LIB)


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_synth()
    build_oracle()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "oracle": build_oracle}[w]()
