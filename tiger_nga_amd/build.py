"""Builds libtiger_chem.so in-tree for gfx950 with hipcc, then the Fortran
ISO_C_BINDING module and the Fortran test caller with flang.

Flags: -O3, C++17, --offload-arch=gfx950, and packed-FP32 instruction
selection disabled for device code (-Xclang -target-feature -Xclang
-packed-fp32-ops); a use of an uninitialised variable is an error.  Usage: ``python -m tiger_nga_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtiger_chem.so")
SOURCES = ["chem_bdf.hip", "chem_lane.hip", "props.hip", "surrogate.hip", "mlp.hip", "noalloc.hip", "diffusion.hip",
           "chem_capi.cpp", "tables_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TGR_OFFLOAD_ARCH", "gfx950")


def command(out=OUT, extra=()):
    return [HIPCC, "-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Werror=uninitialized",
            "-I" + os.path.join(HERE, "..", "include"), *extra,
            *[os.path.join(CSRC, s) for s in SOURCES], "-o", out]


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_SRC = os.path.join(HERE, "fortran", "tiger_chem_mod.f90")
FORTRAN_OUT = os.path.join(HERE, "fortran", "_build")
DRIVER_SRC = os.path.join(HERE, "..", "tests", "fortran", "drive_step.f90")
DRIVER_OUT = os.path.join(HERE, "..", "tests", "fortran", "_build", "drive_step")


def build_fortran(verbose=False):
    """Compiles the ISO_C_BINDING module (tiger_chem.mod + object) and the
    Fortran test caller linked against libtiger_chem.so."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(FORTRAN_OUT, exist_ok=True)
    obj = os.path.join(FORTRAN_OUT, "tiger_chem_mod.o")
    cmds = [[FLANG, "-O2", "-fPIC", "-c", FORTRAN_SRC, "-module-dir", FORTRAN_OUT, "-o", obj]]
    if os.path.exists(DRIVER_SRC):
        os.makedirs(os.path.dirname(DRIVER_OUT), exist_ok=True)
        rel = os.path.relpath(HERE, os.path.dirname(DRIVER_OUT))
        cmds.append([FLANG, "-O2", DRIVER_SRC, obj, "-I", FORTRAN_OUT, "-L" + HERE, "-ltiger_chem",
                     "-Wl,-rpath,$ORIGIN/" + rel, "-o", DRIVER_OUT])
    for c in cmds:
        if verbose:
            print(" ".join(c))
        subprocess.check_call(c)
    return DRIVER_OUT


OBJ_DIR = os.path.join(HERE, "_build")


def build(verbose=False, jobs=None):
    """Compiles each source to an object in parallel (the translation units
    share no device code), then links the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    base = command(extra=())[:-len(SOURCES) - 2]  # compiler and flags, without sources and output
    base = [a for a in base if a not in ("-shared",)]
    objs = [os.path.join(OBJ_DIR, s + ".o") for s in SOURCES]

    def compile_one(k):
        cmd = base + ["-c", os.path.join(CSRC, SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    n = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(compile_one, range(len(SOURCES))))
    tmp = OUT + ".tmp"
    link = [HIPCC, "--offload-arch=" + ARCH, "-fPIC", "-shared", *objs, "-o", tmp]
    if verbose:
        print(" ".join(link))
    subprocess.check_call(link)
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
