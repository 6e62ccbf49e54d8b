"""The C ABI library loads and exports every entry point include/simkit_amd.h declares (no GPU needed)."""

import ctypes
import os
import re

import pytest

from simkit_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "simkit_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(sk_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    syms = declared_symbols()
    assert len(syms) >= 20
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m simkit_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_only_calls():
    lib = _lib.load()
    assert lib.sk_version() == 1
    assert lib.sk_status_string(_lib.SK_ERR_LS_FAIL) == b"line search failed"
    # argument validation happens before any device work
    assert lib.sk_mesh_create(None, None, 0, 0, ctypes.byref(ctypes.c_void_p())) == _lib.SK_ERR_INVALID
    assert b"empty mesh" in lib.sk_last_error()
    assert lib.sk_psd_project(None, 0, 9, 7, None, None) == _lib.SK_ERR_INVALID
