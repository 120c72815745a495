"""Builds libtiger_chem.so in-tree for gfx950 with hipcc.

Flags: -O3, C++17, --offload-arch=gfx950, and packed-FP32 instruction
selection disabled for device code (-Xclang -target-feature -Xclang
-packed-fp32-ops).  Usage: ``python -m tiger_nga_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtiger_chem.so")
SOURCES = ["chem_bdf.hip", "chem_capi.cpp", "tables_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TGR_OFFLOAD_ARCH", "gfx950")


def command(out=OUT, extra=()):
    return [HIPCC, "-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(HERE, "..", "include"), *extra,
            *[os.path.join(CSRC, s) for s in SOURCES], "-o", out]


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
