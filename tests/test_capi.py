"""The C-ABI library loads and exports every symbol include/impes_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "impes_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(impes_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "impes_fgmres_st7" in syms and "impes_amg_setup_csr" in syms
    assert len(syms) >= 20


def test_library_exports_all_declared_symbols():
    import impes_hip

    lib = ctypes.CDLL(impes_hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    import impes_hip

    bound = set(impes_hip.exported_symbols())
    assert set(declared_symbols()) <= bound, set(declared_symbols()) - bound


def test_abi_version_and_error_channel():
    import impes_hip

    lib = impes_hip.load()
    assert lib.impes_abi_version() == 2
    # a call with invalid arguments fails cleanly (argument checks run before any HIP call)
    rc = lib.impes_amg_create(None, None)
    assert rc < 0
    assert b"null out" in lib.impes_last_error()
