"""ctypes binding of the C oracle (oracle/libmsw_oracle.so).

Test infrastructure only (it checks results): imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product package.  ``Oracle`` mirrors the reference module: named state,
diagnostic and tendency arrays plus the L1 kernels and the RK4 drivers
(``src/mpas_sw_module.f90:495-1002``).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsw_oracle.so")

_i = C.POINTER(C.c_int)
_d = C.POINTER(C.c_double)


class _Mesh(C.Structure):
    _fields_ = [
        ("nCells", C.c_int), ("nEdges", C.c_int), ("nVertices", C.c_int),
        ("maxEdges", C.c_int), ("maxEdges2", C.c_int), ("vertexDegree", C.c_int),
        ("cellsOnEdge", _i), ("verticesOnEdge", _i),
        ("nEdgesOnCell", _i), ("edgesOnCell", _i), ("cellsOnCell", _i),
        ("nEdgesOnEdge", _i), ("edgesOnEdge", _i), ("weightsOnEdge", _d),
        ("edgesOnVertex", _i), ("cellsOnVertex", _i), ("kiteAreasOnVertex", _d),
        ("dcEdge", _d), ("dvEdge", _d), ("areaCell", _d), ("areaTriangle", _d),
        ("fVertex", _d), ("h_s", _d), ("u_src", _d),
        ("meshScalingDel2", _d), ("meshScalingDel4", _d),
        ("latCell", _d), ("lonCell", _d), ("coeffs_reconstruct", _d),
    ]


class _Cfg(C.Structure):
    _fields_ = [("dt", C.c_double), ("apvm", C.c_double), ("visc2", C.c_double),
                ("visc4", C.c_double), ("test_case", C.c_int), ("adv_order", C.c_int),
                ("monotonic", C.c_int), ("wind_stress", C.c_int), ("bottom_drag", C.c_int)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        lib = C.CDLL(LIB_PATH)
        lib.mswo_create.restype = C.c_void_p
        lib.mswo_create.argtypes = [C.POINTER(_Mesh), C.POINTER(_Cfg)]
        lib.mswo_destroy.argtypes = [C.c_void_p]
        lib.mswo_set_config.argtypes = [C.c_void_p, C.POINTER(_Cfg)]
        lib.mswo_array.restype = _d
        lib.mswo_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long)]
        for n in ("mswo_diag", "mswo_tend"):
            getattr(lib, n).argtypes = [C.c_void_p, _d, _d]
        for n in ("mswo_diag_tlm", "mswo_tend_tlm", "mswo_diag_adj", "mswo_tend_adj"):
            getattr(lib, n).argtypes = [C.c_void_p, _d, _d, _d, _d]
        lib.mswo_reconstruct.argtypes = [C.c_void_p, _d]
        lib.mswo_reconstruct_tlm.argtypes = [C.c_void_p, _d, _d]
        lib.mswo_reconstruct_adj.argtypes = [C.c_void_p, _d, _d]
        for n in ("mswo_init_block", "mswo_rk4", "mswo_rk4_tlm"):
            getattr(lib, n).argtypes = [C.c_void_p]
        lib.mswo_rk4_adj.argtypes = [C.c_void_p]
        lib.mswo_rk4_adj.restype = C.c_int
        _lib = lib
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Oracle:
    """One model instance.  ``mesh`` is a mesh_like object with the
    reference-layout arrays (``mpas_sw_amd.mesh.Mesh``); ``phys`` holds the
    fields a test case produces: scaled dcEdge, dvEdge, areaCell,
    areaTriangle, kiteAreasOnVertex, and fVertex, h_s, coeffs_reconstruct
    (3, maxEdges, nCells)."""

    def __init__(self, mesh, phys, dt, test_case=2, apvm=0.5, visc2=0.0, visc4=0.0,
                 wind_stress=False, bottom_drag=False, u_src=None):
        lib = load()
        E = mesh.nEdges
        keep = {}

        def f8(x):
            a = np.ascontiguousarray(x, dtype=np.float64)
            keep[id(a)] = a
            return a

        def i4(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            keep[id(a)] = a
            return a

        m = _Mesh()
        m.nCells, m.nEdges, m.nVertices = mesh.nCells, mesh.nEdges, mesh.nVertices
        m.maxEdges, m.maxEdges2, m.vertexDegree = mesh.maxEdges, mesh.maxEdges2, mesh.vertexDegree
        m.cellsOnEdge = _ptr(i4(mesh.cellsOnEdge.T), _i)
        m.verticesOnEdge = _ptr(i4(mesh.verticesOnEdge.T), _i)
        m.nEdgesOnCell = _ptr(i4(mesh.nEdgesOnCell), _i)
        m.edgesOnCell = _ptr(i4(mesh.edgesOnCell.T), _i)
        m.cellsOnCell = _ptr(i4(mesh.cellsOnCell.T), _i)
        m.nEdgesOnEdge = _ptr(i4(mesh.nEdgesOnEdge), _i)
        m.edgesOnEdge = _ptr(i4(mesh.edgesOnEdge.T), _i)
        m.weightsOnEdge = _ptr(f8(mesh.weightsOnEdge.T), _d)
        m.edgesOnVertex = _ptr(i4(mesh.edgesOnVertex.T), _i)
        m.cellsOnVertex = _ptr(i4(mesh.cellsOnVertex.T), _i)
        m.kiteAreasOnVertex = _ptr(f8(np.asarray(phys["kiteAreasOnVertex"]).reshape(3, -1).T
                                      if np.ndim(phys["kiteAreasOnVertex"]) == 2
                                      else phys["kiteAreasOnVertex"]), _d)
        for name in ("dcEdge", "dvEdge", "areaCell", "areaTriangle", "fVertex", "h_s"):
            setattr(m, name, _ptr(f8(phys[name]), _d))
        m.u_src = _ptr(f8(np.zeros(E) if u_src is None else u_src), _d)
        m.meshScalingDel2 = None
        m.meshScalingDel4 = None
        m.latCell = _ptr(f8(mesh.latCell), _d)
        m.lonCell = _ptr(f8(phys.get("lonCell", mesh.lonCell)), _d)
        cr = np.asarray(phys["coeffs_reconstruct"], np.float64)
        # accept (3, maxEdges, C) Fortran order flattened or a (C, maxEdges, 3) array
        cr = cr.reshape(-1) if cr.ndim == 1 else np.ascontiguousarray(cr).reshape(-1)
        m.coeffs_reconstruct = _ptr(f8(cr), _d)
        self.cfg = _Cfg(dt, apvm, visc2, visc4, test_case, 2, 0, int(wind_stress), int(bottom_drag))
        self._h = lib.mswo_create(C.byref(m), C.byref(self.cfg))
        if not self._h:
            raise MemoryError("mswo_create failed")
        self._lib = lib
        self.mesh = mesh

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.mswo_destroy(h)
            self._h = None

    def array(self, name):
        n = C.c_long(0)
        p = self._lib.mswo_array(self._h, name.encode(), C.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def __getitem__(self, name):
        return self.array(name)

    # L2 drivers
    def init_block(self):
        self._lib.mswo_init_block(self._h)

    def rk4(self, n=1):
        for _ in range(n):
            self._lib.mswo_rk4(self._h)

    def rk4_tlm(self, n=1):
        for _ in range(n):
            self._lib.mswo_rk4_tlm(self._h)

    def rk4_adj(self):
        if self._lib.mswo_rk4_adj(self._h) != 0:
            raise MemoryError("mswo_rk4_adj")

    # L1 kernels
    def diag(self, h, u):
        h = np.ascontiguousarray(h, np.float64); u = np.ascontiguousarray(u, np.float64)
        self._lib.mswo_diag(self._h, _ptr(h, _d), _ptr(u, _d))

    def tend(self, h, u):
        h = np.ascontiguousarray(h, np.float64); u = np.ascontiguousarray(u, np.float64)
        self._lib.mswo_tend(self._h, _ptr(h, _d), _ptr(u, _d))

    def diag_tlm(self, h, h_tl, u, u_tl):
        a = [np.ascontiguousarray(x, np.float64) for x in (h, h_tl, u, u_tl)]
        self._lib.mswo_diag_tlm(self._h, *[_ptr(x, _d) for x in a])

    def tend_tlm(self, h, h_tl, u, u_tl):
        a = [np.ascontiguousarray(x, np.float64) for x in (h, h_tl, u, u_tl)]
        self._lib.mswo_tend_tlm(self._h, *[_ptr(x, _d) for x in a])

    def reconstruct(self, u):
        """mpas_reconstruct_2d of u into self["uReconstruct*"]."""
        u = np.ascontiguousarray(u, np.float64)
        self._lib.mswo_reconstruct(self._h, _ptr(u, _d))

    def reconstruct_tlm(self, u, u_tl):
        """mpas_reconstruct_2d_tlm: self["uReconstruct*"] and self["uReconstruct*_tl"]."""
        u = np.ascontiguousarray(u, np.float64)
        u_tl = np.ascontiguousarray(u_tl, np.float64)
        self._lib.mswo_reconstruct_tlm(self._h, _ptr(u, _d), _ptr(u_tl, _d))

    def diag_adj(self, h, u):
        """sw_compute_solve_diagnostics_adj at (h,u) applied to the seeds in
        the diagnostic adjoint arrays (self["<name>_ad"]), which it zeroes:
        returns (h_ad, u_ad)."""
        h = np.ascontiguousarray(h, np.float64); u = np.ascontiguousarray(u, np.float64)
        h_ad = np.zeros_like(h); u_ad = np.zeros_like(u)
        self._lib.mswo_diag_adj(self._h, _ptr(h, _d), _ptr(h_ad, _d), _ptr(u, _d), _ptr(u_ad, _d))
        return h_ad, u_ad

    def tend_full_adj(self, h, u, g_h, g_u):
        """Adjoint of the full tendency operator at (h,u): returns
        (h_ad, u_ad) = F'(h,u)^T (g_h, g_u), via diag, tend_adj, diag_adj."""
        h = np.ascontiguousarray(h, np.float64); u = np.ascontiguousarray(u, np.float64)
        self.diag(h, u)
        self["tend_h_ad"][:] = g_h
        self["tend_u_ad"][:] = g_u
        h_ad = np.zeros_like(h); u_ad = np.zeros_like(u)
        self._lib.mswo_tend_adj(self._h, _ptr(h, _d), _ptr(h_ad, _d), _ptr(u, _d), _ptr(u_ad, _d))
        self._lib.mswo_diag_adj(self._h, _ptr(h, _d), _ptr(h_ad, _d), _ptr(u, _d), _ptr(u_ad, _d))
        return h_ad, u_ad
