"""The C-ABI library loads without a GPU and exports every entry point that
include/tiger_chem.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(tgr_\w+)\s*\(", txt))
    return sorted(names)


def test_header_declares_the_entry_points():
    names = declared()
    for must in ("tgr_chem_create", "tgr_chem_destroy", "tgr_chem_source_chemistry",
                 "tgr_chem_source_chemistry_device", "tgr_chem_last_error", "tgr_mlp_train_step",
                 "tgr_mlp_forward"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from tiger_nga_amd import finitechem
    lib = ctypes.CDLL(finitechem.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_fails_loudly_without_library(tmp_path):
    from tiger_nga_amd import finitechem
    import pytest
    saved = finitechem._lib
    finitechem._lib = None
    try:
        with pytest.raises(Exception):
            finitechem.load_library(str(tmp_path / "missing.so"))
    finally:
        finitechem._lib = saved
