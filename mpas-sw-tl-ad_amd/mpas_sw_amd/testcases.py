"""Williamson et al. (1992) test cases, host side.

Restates ``sw_test_case_1/2/5/6`` of ``src/mpas_sw_testcases.f90`` in the
reference's expression order.  Transcendentals go through Python's ``math``
module (the C library's libm, the same the reference Fortran links) element by
element, so the initial conditions are the reference's to the bit rather than
depending on NumPy's vectorised kernels.  Squares (``x**2.0`` and ``x**2``
in the Fortran) are written as products, since the reference's compiler emits
``x*x`` for them and libm's ``pow(x, 2.0)`` can differ from that product in the
last bit; every other real exponent stays a ``pow`` call, as it does there.

Constants are the reference's (``src/mpas_sw_module.f90:8-11``); with the
reference's ``-fdefault-real-8`` build every literal is a double.
"""

from __future__ import annotations

import ctypes
import ctypes.util
import math
import os

import numpy as np

PII = 3.141592653589793
A = 6371229.0
OMEGA = 7.29212e-5
GRAVITY = 9.80616


_LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
_LIBM.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double)]
_LIBM.sincos.restype = None

# The same libm calls in C loops (csrc/msw_hostmath.c, built next to this
# file): one call per element, so the same bits as the per-element Python
# path below, which is used when the helper is not built.
_VEC = None
_HOSTMATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmsw_hostmath.so")
if os.path.isfile(_HOSTMATH):
    _VEC = ctypes.CDLL(_HOSTMATH)
    _dp = ctypes.POINTER(ctypes.c_double)
    _VEC.msw_vsincos.argtypes = [ctypes.c_long, _dp, _dp, _dp]
    _VEC.msw_vsin.argtypes = [ctypes.c_long, _dp, _dp]
    _VEC.msw_vcos.argtypes = [ctypes.c_long, _dp, _dp]
    _VEC.msw_vpow.argtypes = [ctypes.c_long, _dp, ctypes.c_double, _dp]
    for _fn in ("msw_vsincos", "msw_vsin", "msw_vcos", "msw_vpow"):
        getattr(_VEC, _fn).restype = None


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _sincos(x):
    """(sin x, cos x) through libm's sincos: where the reference evaluates both
    of one argument its compiler emits a single sincos call, whose results can
    differ from separate sin and cos calls in the last bit."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    s = np.empty(x.size)
    c = np.empty(x.size)
    if _VEC is not None:
        _VEC.msw_vsincos(x.size, _dptr(x), _dptr(s), _dptr(c))
        return s, c
    ps, pc = ctypes.c_double(), ctypes.c_double()
    f = _LIBM.sincos
    for i, t in enumerate(x.tolist()):
        f(t, ctypes.byref(ps), ctypes.byref(pc))
        s[i] = ps.value
        c[i] = pc.value
    return s, c


def _f(fn, x):
    return np.fromiter(map(fn, np.asarray(x, np.float64).tolist()), np.float64, count=np.size(x))


def _vec1(name, fn, x):
    if _VEC is None:
        return _f(fn, x)
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.empty(x.size)
    getattr(_VEC, name)(x.size, _dptr(x), _dptr(out))
    return out


def _sin(x):
    return _vec1("msw_vsin", math.sin, x)


def _cos(x):
    return _vec1("msw_vcos", math.cos, x)


def _pow(x, y):
    if _VEC is None:
        return np.fromiter((math.pow(t, y) for t in np.asarray(x, np.float64).tolist()), np.float64,
                           count=np.size(x))
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.empty(x.size)
    _VEC.msw_vpow(x.size, _dptr(x), float(y), _dptr(out))
    return out


def scale_mesh(mesh):
    """Scaled metric fields of a unit-sphere mesh (``src/mpas_sw_testcases.f90:91-104``).

    Returns a dict; the mesh object itself is not modified.
    """
    a2 = A * A
    return {
        "xCell": mesh.xCell * A, "yCell": mesh.yCell * A, "zCell": mesh.zCell * A,
        "xVertex": mesh.xVertex * A, "yVertex": mesh.yVertex * A, "zVertex": mesh.zVertex * A,
        "xEdge": mesh.xEdge * A, "yEdge": mesh.yEdge * A, "zEdge": mesh.zEdge * A,
        "dvEdge": mesh.dvEdge * A, "dcEdge": mesh.dcEdge * A,
        "areaCell": mesh.areaCell * a2, "areaTriangle": mesh.areaTriangle * a2,
        "kiteAreasOnVertex": mesh.kiteAreasOnVertex * a2,
    }


def _u_from_psi(psi, mesh, dvEdge):
    """u = -(psi(v2) - psi(v1))/dvEdge (``src/mpas_sw_testcases.f90:115-119``)."""
    v1 = mesh.verticesOnEdge[0] - 1
    v2 = mesh.verticesOnEdge[1] - 1
    return -1.0 * (psi[v2] - psi[v1]) / dvEdge


def _coriolis(lat, lon, alpha):
    slat, clat = _sincos(lat)
    return 2.0 * OMEGA * (-_cos(lon) * clat * math.sin(alpha) + slat * math.cos(alpha))


def _zonal_psi(mesh, u0, alpha):
    slat, clat = _sincos(mesh.latVertex)
    return -A * u0 * (slat * math.cos(alpha) - _cos(mesh.lonVertex) * clat * math.sin(alpha))


def _geostrophic_h(lat, lon, u0, gh0, alpha):
    slat, clat = _sincos(lat)
    s = (-_cos(lon) * clat * math.sin(alpha) + slat * math.cos(alpha))
    return (gh0 - (A * OMEGA * u0 + 0.5 * (u0 * u0)) * (s * s)) / GRAVITY


def sphere_distance(lat1, lon1, lat2, lon2, radius):
    """``src/mpas_sw_testcases.f90:321-339``."""
    s1 = math.sin(0.5 * (lat2 - lat1))
    s2 = math.sin(0.5 * (lon2 - lon1))
    arg1 = math.sqrt(s1 * s1 + math.cos(lat1) * math.cos(lat2) * (s2 * s2))
    return 2.0 * radius * math.asin(arg1)


def aa(theta):
    """``src/mpas_sw_testcases.f90:341-360``."""
    w = K = 7.848e-6
    R = 4.0
    c = math.cos(theta)
    return (0.5 * w * (2.0 * OMEGA + w) * (c * c) +
            0.25 * (K * K) * math.pow(c, 2.0 * R) *
            ((R + 1.0) * (c * c) + 2.0 * (R * R) - R - 2.0 -
             2.0 * (R * R) * math.pow(c, -2.0)))


def bb(theta):
    """``src/mpas_sw_testcases.f90:362-379``."""
    w = K = 7.848e-6
    R = 4.0
    c = math.cos(theta)
    return ((2.0 * (OMEGA + w) * K / ((R + 1.0) * (R + 2.0))) * math.pow(c, R) *
            (((R * R) + 2.0 * R + 2.0) - ((R + 1.0) * c) * ((R + 1.0) * c)))


def cc(theta):
    """``src/mpas_sw_testcases.f90:381-398``."""
    K = 7.848e-6
    R = 4.0
    c = math.cos(theta)
    return 0.25 * (K * K) * math.pow(c, 2.0 * R) * ((R + 1.0) * (c * c) - R - 2.0)


def initial_conditions(mesh, case):
    """Initial state of Williamson case ``case`` (1, 2, 5 or 6).

    Returns a dict with the scaled metrics of :func:`scale_mesh` plus
    ``h`` [C], ``u`` [E], ``h_s`` [C], ``fVertex`` [V], ``fEdge`` [E],
    ``lonCell`` [C] (TC5 wraps it into [0, 2pi), ``src/mpas_sw_testcases.f90:222``)
    and ``tracer`` [C].
    """
    out = scale_mesh(mesh)
    C, E, V = mesh.nCells, mesh.nEdges, mesh.nVertices
    h_s = np.zeros(C)
    fV = np.zeros(V)
    fE = np.zeros(E)
    tracer = np.zeros(C)
    lonCell = mesh.lonCell.copy()
    dv = out["dvEdge"]
    if case == 1:                                      # :1-66
        u0 = 2.0 * PII * A / (12.0 * 86400.0)
        h0 = 1000.0
        theta_c, lambda_c, alpha = 0.0, 3.0 * PII / 2.0, PII / 4.0
        slat, clat = _sincos(mesh.latVertex)
        psi = -A * u0 * (slat * math.cos(alpha) - _cos(mesh.lonVertex) * clat * math.sin(alpha))
        u = _u_from_psi(psi, mesh, dv)
        h = np.empty(C)
        for i in range(C):
            r = sphere_distance(theta_c, lambda_c, float(mesh.latCell[i]), float(mesh.lonCell[i]), A)
            h[i] = (h0 / 2.0) * (1.0 + math.cos(PII * r * 3.0 / A)) if r < A / 3.0 else h0 / 2.0
    elif case == 2:                                    # :68-145
        u0 = 2.0 * PII * A / (12.0 * 86400.0)
        gh0, alpha = 29400.0, 0.0
        u = _u_from_psi(_zonal_psi(mesh, u0, alpha), mesh, dv)
        fE = _coriolis(mesh.latEdge, mesh.lonEdge, alpha)
        fV = _coriolis(mesh.latVertex, mesh.lonVertex, alpha)
        h = _geostrophic_h(mesh.latCell, mesh.lonCell, u0, gh0, alpha)
    elif case == 5:                                    # :147-254
        u0, gh0, hs0 = 20.0, 5960.0 * GRAVITY, 2000.0
        theta_c, lambda_c, rr, alpha = PII / 6.0, 3.0 * PII / 2.0, PII / 9.0, 0.0
        u = _u_from_psi(_zonal_psi(mesh, u0, alpha), mesh, dv)
        fE = _coriolis(mesh.latEdge, mesh.lonEdge, alpha)
        fV = _coriolis(mesh.latVertex, mesh.lonVertex, alpha)
        lonCell = np.where(lonCell < 0.0, lonCell + 2.0 * PII, lonCell)
        rr2 = rr * rr
        for i in range(C):
            dl = float(lonCell[i]) - lambda_c
            dt = float(mesh.latCell[i]) - theta_c
            d2 = dl * dl + dt * dt
            r = math.sqrt(min(rr2, d2))
            h_s[i] = hs0 * (1.0 - r / rr)
            tracer[i] = 1.0 - r / rr
        h = _geostrophic_h(mesh.latCell, lonCell, u0, gh0, alpha) - h_s
    elif case == 6:                                    # :256-319 (f stays 0)
        h0, w, K, R = 8000.0, 7.848e-6, 7.848e-6, 4.0
        lat, lon = mesh.latVertex, mesh.lonVertex
        slat, clat = _sincos(lat)
        psi = (-A * A * w * slat + A * A * K * _pow(clat, R) * slat * _cos(R * lon))
        u = _u_from_psi(psi, mesh, dv)
        h = np.empty(C)
        for i in range(C):
            th = float(mesh.latCell[i])
            ln = float(mesh.lonCell[i])
            h[i] = (GRAVITY * h0 + A * A * aa(th) + A * A * bb(th) * math.cos(R * ln) +
                    A * A * cc(th) * math.cos(2.0 * R * ln)) / GRAVITY
    else:
        raise ValueError("test case must be 1, 2, 5 or 6")
    out.update(h=np.asarray(h, np.float64), u=np.asarray(u, np.float64), h_s=h_s, fVertex=fV,
               fEdge=fE, lonCell=lonCell, tracer=tracer)
    return out
