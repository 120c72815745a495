"""Builds the CPU wave-emulator library of the solver (test infrastructure):
tests/emu/_build/libwave_emu.so from the device sources compiled with g++
-DTGR_WAVE_EMU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
CSRC = os.path.join(ROOT, "tiger_nga_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libwave_emu.so")


def build(extra=()):
    """Compiles to a file of this process and renames it into place, so that
    test workers building at the same time never load a partly written
    library."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = "%s.%d.tmp" % (OUT, os.getpid())
    cmd = ["g++", "-std=c++20", "-O2", "-g", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-DTGR_WAVE_EMU",
           "-I" + HERE, "-I" + CSRC, *extra, os.path.join(HERE, "emu_driver.cpp"),
           os.path.join(CSRC, "tables_host.cpp"), "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
