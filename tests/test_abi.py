"""The C-ABI library loads and exports every symbol include/mpas_sw_amd.h
declares (no GPU call is made here)."""

import ctypes
import os
import re

from mpas_sw_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mpas_sw_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(msw_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_abi_version_and_error_channel():
    lib = _lib.load()
    assert lib.msw_abi_version() == 1
    assert isinstance(lib.msw_last_error(), bytes)
    names = lib.msw_diag_names().decode().split(",")
    kinds = lib.msw_diag_kinds().decode().split(",")
    assert len(names) == len(kinds) == 13


def test_create_rejects_bad_arguments_without_gpu():
    # argument validation happens before any device call
    lib = _lib.load()
    out = ctypes.c_void_p()
    rc = lib.msw_create(None, None, 0, 1, None, ctypes.byref(out))
    assert rc == -1 and not out
    assert b"NULL" in lib.msw_last_error()
    cfg = _lib.Config(300.0, 0.5, 0.0, 0.0, 2, 2, 0, 0, 0)
    d = _lib.MeshDesc()
    rc = lib.msw_create(ctypes.byref(d), ctypes.byref(cfg), 0, 3, None, ctypes.byref(out))
    assert rc == -1 and b"power of two" in lib.msw_last_error()


def test_python_constants_match_the_header():
    src = open(os.path.join(ROOT, "include", "mpas_sw_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    consts = {k: int(v) for k, v in re.findall(r"\b(MSW_[A-Z0-9]+)\s*=\s*(-?\d+)", src)}
    for name in ("MSW_OK", "MSW_EINVAL", "MSW_ENOMEM", "MSW_EHIP", "MSW_EUNSUPPORTED",
                 "MSW_FP64", "MSW_FP32"):
        assert getattr(_lib, name) == consts[name], name
