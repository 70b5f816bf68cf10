"""The C ABI library loads on a machine without a GPU and exports every entry point that
include/ibh.h declares; the Python binding covers the same list.  No compute calls."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ibh.h")
LIB = os.path.join(ROOT, "immersedboundary.jl_amd", "libibhip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ibh_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("ibh_create", "ibh_setup", "ibh_advance", "ibh_get_field", "ibh_op_lapinv",
                 "ibh_op_B"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libibhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.ibh_version() == 1


def test_python_binding_lists_every_symbol():
    from ibhip import _lib

    assert sorted(_lib.EXPORTED) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libibhip.so not built")
def test_create_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from ibhip import _lib

    L = _lib.load()
    d = _lib.CaseDesc()
    d.nx, d.ny, d.nlev, d.h = 32, 32, 2, 0.1
    h = C.c_void_p()
    assert L.ibh_create(C.byref(d), C.byref(h)) == _lib.EDEVICE
    d.nx = 30  # not a multiple of 4
    assert L.ibh_create(C.byref(d), C.byref(h)) == _lib.EINVAL
