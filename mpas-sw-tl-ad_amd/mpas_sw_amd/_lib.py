"""ctypes binding of libmpas_sw_amd.so (the C ABI in include/mpas_sw_amd.h).

The shared library is built in-tree by :func:`build` (hipcc, gfx950).  Loading
fails loudly when it is missing: there is no CPU fallback for the stepped
path.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(HERE, "libmpas_sw_amd.so")

SOURCES = ("msw_kernels.hip", "msw_runtime.cpp")
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # FP contraction off: gathers reproduce the reference's rounding exactly
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]

# roctx ranges around the entry points (rocprofv3 --marker-trace)
LINK_FLAGS = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"]

MSW_FP64 = 0
MSW_FP32 = 1

# status codes (enum msw_status)
MSW_OK = 0
MSW_EINVAL = -1
MSW_ENOMEM = -2
MSW_EHIP = -3
MSW_EUNSUPPORTED = -4


HOSTMATH_PATH = os.path.join(HERE, "libmsw_hostmath.so")


def build_hostmath() -> str:
    """Compile the host libm loops of the initial conditions (gcc -O2, no
    fast-math) in-tree; returns the path."""
    cc = os.environ.get("CC", "gcc")
    tmp = HOSTMATH_PATH + ".tmp"
    # -fno-builtin: the libm calls as written (GCC would otherwise split
    # sincos into sin and cos, whose results can differ in the last bit)
    cmd = [cc, "-O2", "-fno-builtin", "-fPIC", "-shared", "-o", tmp, os.path.join(CSRC, "msw_hostmath.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOSTMATH_PATH)
    return HOSTMATH_PATH


def build(verbose: bool = False) -> str:
    """Compile the HIP library (and the host libm loops) in-tree; returns the
    HIP library's path."""
    build_hostmath()
    hipcc = os.environ.get("HIPCC", "hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, *srcs, *LINK_FLAGS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(" ".join(cmd))
    return LIB_PATH


class MeshDesc(C.Structure):
    _fields_ = [
        ("nCells", C.c_int32), ("nEdges", C.c_int32), ("nVertices", C.c_int32),
        ("maxEdges", C.c_int32), ("maxEdges2", C.c_int32), ("vertexDegree", C.c_int32),
        ("cellsOnEdge", C.c_void_p), ("verticesOnEdge", C.c_void_p),
        ("nEdgesOnCell", C.c_void_p), ("edgesOnCell", C.c_void_p),
        ("nEdgesOnEdge", C.c_void_p), ("edgesOnEdge", C.c_void_p),
        ("weightsOnEdge", C.c_void_p), ("cellsOnVertex", C.c_void_p),
        ("kiteAreasOnVertex", C.c_void_p), ("dcEdge", C.c_void_p), ("dvEdge", C.c_void_p),
        ("areaCell", C.c_void_p), ("areaTriangle", C.c_void_p), ("fVertex", C.c_void_p),
        ("h_s", C.c_void_p), ("latCell", C.c_void_p), ("lonCell", C.c_void_p),
        ("coeffs_reconstruct", C.c_void_p),
        ("cellPerm", C.c_void_p), ("edgePerm", C.c_void_p), ("vertexPerm", C.c_void_p),
    ]


class Config(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("apvm_upwinding", C.c_double), ("visc2", C.c_double),
        ("visc4", C.c_double), ("test_case", C.c_int32), ("thickness_adv_order", C.c_int32),
        ("monotonic", C.c_int32), ("wind_stress", C.c_int32), ("bottom_drag", C.c_int32),
    ]


def mesh_desc(mesh, phys, perm):
    """msw_mesh_desc of a mesh and its scaled metrics (``phys``, as from
    :func:`mpas_sw_amd.testcases.initial_conditions`), with the internal
    order ``perm`` = {"c", "e", "v"} (new -> old, int32).  Returns the
    descriptor and the host arrays it points to, which must outlive the
    msw_create call that reads them."""
    keep = []

    def f8(x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data

    def i4(x):
        a = np.ascontiguousarray(x, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data

    d = MeshDesc()
    d.nCells, d.nEdges, d.nVertices = mesh.nCells, mesh.nEdges, mesh.nVertices
    d.maxEdges, d.maxEdges2, d.vertexDegree = mesh.maxEdges, mesh.maxEdges2, mesh.vertexDegree
    d.cellsOnEdge = i4(mesh.cellsOnEdge.T)
    d.verticesOnEdge = i4(mesh.verticesOnEdge.T)
    d.nEdgesOnCell = i4(mesh.nEdgesOnCell)
    d.edgesOnCell = i4(mesh.edgesOnCell.T)
    d.nEdgesOnEdge = i4(mesh.nEdgesOnEdge)
    d.edgesOnEdge = i4(mesh.edgesOnEdge.T)
    d.weightsOnEdge = f8(mesh.weightsOnEdge.T)
    d.cellsOnVertex = i4(mesh.cellsOnVertex.T)
    kite = np.asarray(phys["kiteAreasOnVertex"])
    d.kiteAreasOnVertex = f8(kite.T if kite.ndim == 2 else kite)
    for name in ("dcEdge", "dvEdge", "areaCell", "areaTriangle", "fVertex", "h_s"):
        setattr(d, name, f8(phys[name]))
    d.latCell = f8(mesh.latCell)
    d.lonCell = f8(phys.get("lonCell", mesh.lonCell))
    cr = phys.get("coeffs_reconstruct")
    if cr is not None:
        cr = np.asarray(cr, np.float64)
        # (3, maxEdges, C) reference layout -> entity-major flat
        d.coeffs_reconstruct = f8(np.transpose(cr, (2, 1, 0)) if cr.ndim == 3 else cr)
    d.cellPerm = i4(perm["c"])
    d.edgePerm = i4(perm["e"])
    d.vertexPerm = i4(perm["v"])
    return d, keep


# every symbol include/mpas_sw_amd.h declares
EXPORTS = (
    "msw_create", "msw_destroy", "msw_set_config", "msw_set_forcing", "msw_set_mesh_scaling",
    "msw_set_option",
    "msw_set_graphs", "msw_fwd_steps",
    "msw_base_steps", "msw_tlm_steps", "msw_adj_step", "msw_adj_steps", "msw_diagnostics",
    "msw_diagnostics_adj",
    "msw_diag_names", "msw_diag_kinds", "msw_reconstruct", "msw_reconstruct_adj", "msw_invariants", "msw_dot", "msw_nonfinite",
    "msw_permute", "msw_device_bytes", "msw_abi_version", "msw_last_error", "msw_dot_noalloc",
    "msw_set_instances", "msw_set_tend_output",
)

OPT_TL_FUSED_FLUX = 1   # msw_option
OPT_XCD_STRIPE = 2
OPT_BLOCK = 3
OPT_PERSISTENT = 4
OPT_BARRIER_POLLS = 5

_lib = None


class MSWError(RuntimeError):
    """A negative status from the HIP library, with its message."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


def load(path=None):
    """Load the library (no GPU call happens here).  ``path`` binds the same
    signatures to another build of the C ABI (not cached)."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is None and not os.path.isfile(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the GPU path has no CPU fallback)")
    lib = C.CDLL(path or LIB_PATH)
    vp, i32, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    sig = {
        "msw_create": ([C.POINTER(MeshDesc), C.POINTER(Config), i32, i32, vp, C.POINTER(vp)], i32),
        "msw_destroy": ([vp], i32),
        "msw_set_config": ([vp, C.POINTER(Config)], i32),
        "msw_set_forcing": ([vp, vp], i32),
        "msw_set_mesh_scaling": ([vp, vp, vp], i32),
        "msw_set_option": ([vp, i32, i32], i32),
        "msw_set_graphs": ([vp, i32], i32),
        "msw_fwd_steps": ([vp, vp, vp, i32], i32),
        "msw_base_steps": ([vp, vp, vp, i32, vp, vp], i32),
        "msw_tlm_steps": ([vp, vp, vp, vp, vp, i32, vp, vp], i32),
        "msw_adj_step": ([vp, vp, vp, vp, vp, vp, vp], i32),
        "msw_adj_steps": ([vp, vp, vp, vp, vp, i32], i32),
        "msw_diagnostics": ([vp, vp, vp, vp], i32),
        "msw_diagnostics_adj": ([vp, vp, vp, vp, vp, vp], i32),
        "msw_diag_names": ([], C.c_char_p),
        "msw_diag_kinds": ([], C.c_char_p),
        "msw_reconstruct": ([vp, vp, vp], i32),
        "msw_reconstruct_adj": ([vp, vp, vp], i32),
        "msw_invariants": ([vp, vp, vp, vp], i32),
        "msw_dot": ([vp, vp, vp, vp, vp, vp, vp, vp], i32),
        "msw_nonfinite": ([vp, vp, vp, vp], i32),
        "msw_dot_noalloc": ([i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp], i32),
        "msw_set_instances": ([vp, i32, vp, vp, vp], i32),
        "msw_set_tend_output": ([vp, vp, vp, vp, vp], i32),
        "msw_permute": ([vp, i32, i32, i32, i32, vp, vp], i32),
        "msw_device_bytes": ([vp], i64),
        "msw_abi_version": ([], i32),
        "msw_last_error": ([], C.c_char_p),
    }
    for name, (argt, rest) in sig.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = rest
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().msw_last_error().decode(errors="replace")
        raise MSWError(rc, msg)
    return rc
