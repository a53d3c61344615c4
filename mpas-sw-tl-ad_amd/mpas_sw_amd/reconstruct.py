"""Host set-up of the cell-centre velocity reconstruction.

Restates ``mpas_initialize_vectors`` (``src/mpas_sw_operators.f90:2-116``)
and ``mpas_init_reconstruct`` with its RBF solve
(``src/mpas_sw_operators.f90:162-279, 413-625``): per cell, an
inverse-multiquadric vector RBF interpolation in the cell's tangent plane,
solved by scaled-partial-pivot Gaussian elimination (``elgs``/``mpas_legs``).
The elimination is vectorised over all cells with the same edge count and
follows the reference's operation order.  This runs once per mesh on the
host; the per-step reconstruction itself is a GPU kernel.
"""

from __future__ import annotations

import numpy as np


def _unit(v):
    mag = np.sqrt(v[..., 0] ** 2 + v[..., 1] ** 2 + v[..., 2] ** 2)
    return v / mag[..., None]


def _sum3(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def initialize_vectors(xc, xe, cellsOnEdge, edgesOnCell):
    """Returns (localVerticalUnitVectors [C,3], edgeNormalVectors [E,3],
    cellTangentPlane [C,2,3]) for a sphere (``src/mpas_sw_operators.f90:2-116``)."""
    lvu = _unit(xc)
    c1 = cellsOnEdge[0] - 1
    c2 = cellsOnEdge[1] - 1
    env = _unit(xc[c2] - xc[c1])
    e1 = edgesOnCell[0] - 1
    n = env[e1]
    ndr = _sum3(n, lvu)
    xhat = _unit(n - ndr[:, None] * lvu)
    r = lvu
    yhat = np.stack([r[:, 1] * xhat[:, 2] - r[:, 2] * xhat[:, 1],
                     r[:, 2] * xhat[:, 0] - r[:, 0] * xhat[:, 2],
                     r[:, 0] * xhat[:, 1] - r[:, 1] * xhat[:, 0]], 1)
    yhat = _unit(yhat)
    return lvu, env, np.stack([xhat, yhat], 1)


def _elgs(A):
    """Scaled partial-pivot elimination in place on a batch (nb,N,N);
    returns the pivot order (``src/mpas_sw_operators.f90:560-625``)."""
    nb, N, _ = A.shape
    ar = np.arange(nb)
    indx = np.tile(np.arange(N), (nb, 1))
    Cs = np.max(np.abs(A), axis=2)
    for j in range(N - 1):
        rows = indx[:, j:]
        pi = np.abs(A[ar[:, None], rows, j]) / Cs[ar[:, None], rows]
        k = j + np.argmax(pi, axis=1)
        tmp = indx[ar, j].copy()
        indx[ar, j] = indx[ar, k]
        indx[ar, k] = tmp
        pj_row = indx[:, j]
        for i in range(j + 1, N):
            ri = indx[:, i]
            pj = A[ar, ri, j] / A[ar, pj_row, j]
            A[ar, ri, j] = pj
            for kk in range(j + 1, N):
                A[ar, ri, kk] = A[ar, ri, kk] - pj * A[ar, pj_row, kk]
    return indx


def _legs(A, b):
    """Solve with the elimination of ``_elgs`` (``src/mpas_sw_operators.f90:529-558``)."""
    nb, N, _ = A.shape
    ar = np.arange(nb)
    indx = _elgs(A)
    b = b.copy()
    for i in range(N - 1):
        for j in range(i + 1, N):
            b[ar, indx[:, j]] = b[ar, indx[:, j]] - A[ar, indx[:, j], i] * b[ar, indx[:, i]]
    x = np.zeros((nb, N))
    x[:, N - 1] = b[ar, indx[:, N - 1]] / A[ar, indx[:, N - 1], N - 1]
    for i in range(N - 2, -1, -1):
        xi = b[ar, indx[:, i]]
        for j in range(i + 1, N):
            xi = xi - A[ar, indx[:, i], j] * x[:, j]
        x[:, i] = xi / A[ar, indx[:, i], i]
    return x


def init_reconstruct(xc, xe, env, ctp, nEdgesOnCell, edgesOnCell):
    """Coefficients (3, maxEdges, C) of ``mpas_init_reconstruct``
    (``src/mpas_sw_operators.f90:162-279``); zero in unused slots."""
    C = xc.shape[0]
    maxE = edgesOnCell.shape[0]
    out = np.zeros((3, maxE, C))
    for n in np.unique(nEdgesOnCell):
        cells = np.nonzero(nEdgesOnCell == n)[0]
        nb = cells.size
        eids = (edgesOnCell[:n, cells] - 1).T                   # (nb, n)
        loc = xe[eids]                                          # (nb, n, 3)
        nrm = env[eids]
        cc = xc[cells]                                          # (nb, 3)
        d = cc[:, None, :] - loc
        r = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
        alpha = np.zeros(nb)
        for i in range(n):
            alpha = alpha + r[:, i]
        alpha = alpha / n
        p1 = ctp[cells, 0, :]
        p2 = ctp[cells, 1, :]
        ps = np.stack([_sum3(loc, p1[:, None, :]), _sum3(loc, p2[:, None, :])], -1)   # (nb,n,2)
        pu = np.stack([_sum3(nrm, p1[:, None, :]), _sum3(nrm, p2[:, None, :])], -1)
        pd = np.stack([_sum3(cc, p1), _sum3(cc, p2)], -1)                              # (nb,2)
        N = n + 2
        M = np.zeros((nb, N, N))
        a2 = alpha ** 2
        for j in range(n):
            for i in range(j, n):
                dd = ps[:, i, :] - ps[:, j, :]
                rsq = (dd[:, 0] ** 2 + dd[:, 1] ** 2) / a2
                rbf = 1 / np.sqrt(1 + rsq)
                dot = pu[:, i, 0] * pu[:, j, 0] + pu[:, i, 1] * pu[:, j, 1]
                M[:, i, j] = rbf * dot
                M[:, j, i] = M[:, i, j]
        rhs = np.zeros((nb, N, 2))
        for j in range(n):
            dd = pd - ps[:, j, :]
            rsq = (dd[:, 0] ** 2 + dd[:, 1] ** 2) / a2
            rhs[:, j, :] = (1 / np.sqrt(1 + rsq))[:, None] * pu[:, j, :]
        for i in range(n):
            M[:, i, n:n + 2] = pu[:, i, :]
            M[:, n:n + 2, i] = M[:, i, n:n + 2]
        rhs[:, n, 0] = 1.0
        rhs[:, n + 1, 1] = 1.0
        c1 = _legs(M.copy(), rhs[:, :, 0])
        c2 = _legs(M.copy(), rhs[:, :, 1])
        for k in range(3):
            out[k, :n, cells] = (p1[:, k][:, None] * c1[:, :n] + p2[:, k][:, None] * c2[:, :n])
    return out


def reconstruct_coefficients(scaled, mesh):
    """Coefficients for a test-case-scaled mesh (dict from ``testcases``)."""
    xc = np.stack([scaled["xCell"], scaled["yCell"], scaled["zCell"]], 1)
    xe = np.stack([scaled["xEdge"], scaled["yEdge"], scaled["zEdge"]], 1)
    lvu, env, ctp = initialize_vectors(xc, xe, mesh.cellsOnEdge, mesh.edgesOnCell)
    return init_reconstruct(xc, xe, env, ctp, mesh.nEdgesOnCell, mesh.edgesOnCell)
