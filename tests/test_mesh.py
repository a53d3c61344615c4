"""Mesh generator invariants (the reference's mesh files are not available,
so the generated meshes are pinned by geometry and by the TRiSK identities)."""

import numpy as np
import pytest

from mpas_sw_amd.mesh import get_mesh, sfc_permutation


@pytest.mark.parametrize("level", [0, 1, 2, 3, 5])
def test_counts_and_areas(level):
    m = get_mesh(level)
    assert m.nCells == 10 * 4 ** level + 2
    assert m.nEdges == 30 * 4 ** level
    assert m.nVertices == 20 * 4 ** level
    assert m.nCells - m.nEdges + m.nVertices == 2
    assert abs(m.areaCell.sum() - 4 * np.pi) < 1e-12
    assert abs(m.areaTriangle.sum() - 4 * np.pi) < 1e-12
    assert np.abs(m.kiteAreasOnVertex.sum(0) - m.areaTriangle).max() < 1e-15
    assert np.all(m.nEdgesOnCell[:12] == 5) and np.all(m.nEdgesOnCell[12:] == 6)
    # every edge's edgesOnEdge = other edges of its two cells
    assert np.all(m.nEdgesOnEdge == (m.nEdgesOnCell[m.cellsOnEdge[0] - 1] +
                                     m.nEdgesOnCell[m.cellsOnEdge[1] - 1] - 2))


@pytest.mark.parametrize("level", [2, 4])
def test_trisk_weights(level):
    m = get_mesh(level)
    E = m.nEdges
    # antisymmetry w(e,e') dc(e)/dv(e') = -w(e',e) dc(e')/dv(e)
    scale = np.abs(m.weightsOnEdge).max()
    for e in range(0, E, 7):
        for j in range(m.nEdgesOnEdge[e]):
            ep = m.edgesOnEdge[j, e] - 1
            lst = list(m.edgesOnEdge[:m.nEdgesOnEdge[ep], ep] - 1)
            jj = lst.index(e)
            a = m.weightsOnEdge[j, e] * m.dcEdge[e] / m.dvEdge[ep]
            b = -m.weightsOnEdge[jj, ep] * m.dcEdge[ep] / m.dvEdge[e]
            assert abs(a - b) < 1e-13 * scale
    # v = sum w u(eoe) approximates the tangential component (along k x n)
    xe = np.stack([m.xEdge, m.yEdge, m.zEdge], 1)
    xc = np.stack([m.xCell, m.yCell, m.zCell], 1)
    c1, c2 = m.cellsOnEdge - 1
    n = xc[c2] - xc[c1]
    n -= (n * xe).sum(1)[:, None] * xe
    n /= np.linalg.norm(n, axis=1)[:, None]
    t = np.cross(xe, n)
    vel = np.cross(np.array([0.2, 0.3, 0.9])[None, :], xe)
    u = (vel * n).sum(1)
    vex = (vel * t).sum(1)
    v = np.zeros(E)
    for j in range(12):
        eo = m.edgesOnEdge[j] - 1
        ok = eo >= 0
        v[ok] += m.weightsOnEdge[j][ok] * u[eo[ok]]
    assert np.corrcoef(v, vex)[0, 1] > 0.99
    # vertex order: circulation is counter-clockwise (tangent from v1 to v2)
    xv = np.stack([m.xVertex, m.yVertex, m.zVertex], 1)
    d = xv[m.verticesOnEdge[1] - 1] - xv[m.verticesOnEdge[0] - 1]
    assert np.all((d * t).sum(1) > 0)


def test_sfc_permutation():
    m = get_mesh(4)
    pc, pe, pv = sfc_permutation(m)
    for p, n in ((pc, m.nCells), (pe, m.nEdges), (pv, m.nVertices)):
        assert np.array_equal(np.sort(p), np.arange(n))
    # locality: consecutive cells in the new order are close on the sphere
    xc = np.stack([m.xCell, m.yCell, m.zCell], 1)[pc]
    step = np.linalg.norm(np.diff(xc, axis=0), axis=1)
    assert np.median(step) < 2.5 * np.median(m.dcEdge)
    pm = m.permuted(pc, pe, pv)
    assert abs(pm.areaCell.sum() - m.areaCell.sum()) < 1e-12
    deg = np.bincount(pm.cellsOnEdge.ravel() - 1, minlength=pm.nCells)
    assert np.array_equal(deg, pm.nEdgesOnCell)
