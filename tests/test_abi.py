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


def test_create_rejects_corrupt_meshes_without_gpu():
    """msw_create checks the mesh on the host before any device call: an
    out-of-range connectivity index, a non-positive metric or a permutation
    that is not a bijection gives MSW_EINVAL naming the entity (1-based,
    reference order), never a fault."""
    import copy
    import numpy as np
    from mpas_sw_amd.core import mesh_desc
    from mpas_sw_amd.mesh import get_mesh, sfc_permutation
    from mpas_sw_amd.testcases import initial_conditions
    mesh = get_mesh(2)
    phys = initial_conditions(mesh, 5)
    perm = {k: np.ascontiguousarray(p, np.int32) for k, p in zip("cev", sfc_permutation(mesh))}
    cfg = _lib.Config(900.0, 0.5, 0.0, 0.0, 5, 2, 0, 0, 0)
    lib = _lib.load()
    E, C, V = mesh.nEdges, mesh.nCells, mesh.nVertices

    def attempt(m, ph, pm):
        d, keep = mesh_desc(m, ph, pm)
        out = ctypes.c_void_p()
        rc = lib.msw_create(ctypes.byref(d), ctypes.byref(cfg), _lib.MSW_FP64, 1, None, ctypes.byref(out))
        assert not out.value
        return rc, lib.msw_last_error().decode()

    cases = [("edgesOnEdge", (0, 7), E + 5, "edgesOnEdge out of range at edge 8"),
             ("edgesOnEdge", (2, 11), 0, "edgesOnEdge out of range at edge 12"),
             ("cellsOnEdge", (1, 3), C + 1, "cellsOnEdge out of range at edge 4"),
             ("verticesOnEdge", (0, 9), -2, "verticesOnEdge out of range at edge 10"),
             ("edgesOnCell", (4, 20), E + 1, "edgesOnCell out of range at cell 21"),
             ("cellsOnVertex", (2, 30), 0, "cellsOnVertex out of range at vertex 31")]
    for field, idx, bad, msg in cases:
        m = copy.copy(mesh)
        a = np.array(getattr(mesh, field), copy=True)
        a[idx] = bad
        setattr(m, field, a)
        rc, err = attempt(m, phys, perm)
        assert rc == _lib.MSW_EINVAL and msg in err, (field, rc, err)
    ph = dict(phys)
    ph["dcEdge"] = phys["dcEdge"].copy()
    ph["dcEdge"][5] = 0.0
    rc, err = attempt(mesh, ph, perm)
    assert rc == _lib.MSW_EINVAL and "dcEdge/dvEdge at edge 6" in err
    pm = dict(perm)
    pm["e"] = perm["e"].copy()
    pm["e"][3] = pm["e"][4]
    rc, err = attempt(mesh, phys, pm)
    assert rc == _lib.MSW_EINVAL and "bijection" in err
