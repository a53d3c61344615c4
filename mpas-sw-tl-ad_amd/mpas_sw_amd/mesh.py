"""Spherical Voronoi mesh generator in MPAS conventions.

The reference reads its mesh from an MPAS ``x1.<nCells>.grid.nc`` file
(``src/mpas_sw_driver.py:87-119``).  Those files are not available, so this
module builds the level-``n`` mesh itself: the dual of an icosahedron whose
faces are bisected ``n`` times, every new point projected to the unit sphere.

Counts: nCells = 10*4**n + 2, nEdges = 30*4**n, nVertices = 20*4**n.

Conventions (the ones the reference's stencil loops rely on):

* Indices are 1-based in the returned "reference layout" arrays, as in the
  MPAS file (``src/mpas_sw_module.f90:253-272``).
* Edge ``e`` joins ``cellsOnEdge[:, e]`` = (c1, c2); its unit normal points
  from c1 to c2; ``verticesOnEdge`` = (v1, v2) is ordered so that the tangent
  ``t = k x n`` points from v1 to v2 (``k`` = local vertical).  With this,
  circulation around a vertex is counter-clockwise
  (``src/mpas_sw_dyncore.f90:209-214``) and gradPVt is along ``t``
  (``src/mpas_sw_dyncore.f90:296-301``).
* ``edgesOnCell`` / ``verticesOnCell`` run counter-clockwise seen from
  outside the sphere; ``verticesOnCell[i]`` lies between ``edgesOnCell[i]``
  and ``edgesOnCell[i+1]``; ``cellsOnCell[i]`` is across ``edgesOnCell[i]``.
* ``cellsOnVertex`` runs counter-clockwise; ``edgesOnVertex[i]`` joins
  ``cellsOnVertex[i]`` and ``cellsOnVertex[i+1]``; ``kiteAreasOnVertex[i]`` is
  the part of the vertex's triangle inside cell ``cellsOnVertex[i]``.
* ``edgesOnEdge`` / ``weightsOnEdge`` are the TRiSK reconstruction of the
  tangential velocity (Thuburn et al. 2009, Ringler et al. 2010): the other
  edges of c1 counter-clockwise starting after ``e``, then those of c2.  With
  R(i,v) = kite(i,v)/areaCell(i) and the partial sum S_k over the vertices
  passed going counter-clockwise from ``e`` to the k-th edge,
  ``w = (1/2 - S_k) * s(e,i) * s(e_k,i) * dvEdge(e_k)/dcEdge(e)`` where
  ``s(e,i) = +1`` when the normal of ``e`` points out of cell ``i``.  The sign
  makes ``v = sum w*u(eoe)`` the component along ``t``; the mesh tests check it
  against an analytic flow and check the antisymmetry
  ``w(e,e') dc(e) dv(e) = -w(e',e) dc(e') dv(e')``.
* Lengths and areas are on the unit sphere; the test-case routines scale them
  by the Earth radius (``src/mpas_sw_testcases.f90:91-104``).
"""

from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

MAX_EDGES = 6
MAX_EDGES2 = 12
VERTEX_DEGREE = 3


def _normalize(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def _arc(a, b):
    """Great-circle angle between unit vectors (rows), robust for small arcs."""
    cr = np.linalg.norm(np.cross(a, b), axis=-1)
    dt = np.einsum("ij,ij->i", a, b)
    return np.arctan2(cr, dt)


def _tri_area(a, b, c):
    """Signed spherical-triangle area (positive for counter-clockwise a,b,c).

    tan(E/2) = a.(b x c) / (1 + a.b + b.c + c.a)  (Van Oosterom & Strackee).
    """
    num = np.einsum("ij,ij->i", a, np.cross(b, c))
    den = 1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
    return 2.0 * np.arctan2(num, den)


def _unique_pairs(es, n):
    """np.unique(es, axis=0, return_inverse=True) for rows (a, b) of point
    indices with a < b < n: the same unique rows in the same (lexicographic)
    order and the same inverse, from a 1-D key a * n + b."""
    key = es[:, 0].astype(np.int64) * n + es[:, 1]
    uk, inv = np.unique(key, return_inverse=True)
    return np.stack([uk // n, uk % n], 1), inv


def _icosahedron():
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
         [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v = _normalize(v)
    # make every face counter-clockwise seen from outside
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def _bisect(v, f):
    """One bisection level: every triangle -> 4, midpoints projected."""
    nf = f.shape[0]
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
    es = np.sort(e, axis=1)
    uniq, inv = _unique_pairs(es, v.shape[0])
    inv = inv.reshape(-1)
    mid = _normalize(v[uniq[:, 0]] + v[uniq[:, 1]])
    nv = v.shape[0]
    vnew = np.concatenate([v, mid], axis=0)
    m01 = nv + inv[0:nf]
    m12 = nv + inv[nf:2 * nf]
    m20 = nv + inv[2 * nf:3 * nf]
    f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
    fnew = np.concatenate([
        np.stack([f0, m01, m20], 1),
        np.stack([m01, f1, m12], 1),
        np.stack([m20, m12, f2], 1),
        np.stack([m01, m12, m20], 1)], axis=0)
    return vnew, fnew


@dataclasses.dataclass
class Mesh:
    """A spherical Voronoi mesh.  Arrays follow the MPAS file layout
    transposed to what the reference's Fortran module holds: connectivity is
    ``(slot, entity)`` and 1-based; per-entity fields are 1-D."""

    level: int
    nCells: int
    nEdges: int
    nVertices: int
    xCell: np.ndarray
    yCell: np.ndarray
    zCell: np.ndarray
    latCell: np.ndarray
    lonCell: np.ndarray
    xEdge: np.ndarray
    yEdge: np.ndarray
    zEdge: np.ndarray
    latEdge: np.ndarray
    lonEdge: np.ndarray
    xVertex: np.ndarray
    yVertex: np.ndarray
    zVertex: np.ndarray
    latVertex: np.ndarray
    lonVertex: np.ndarray
    cellsOnEdge: np.ndarray        # (2, E)
    verticesOnEdge: np.ndarray     # (2, E)
    nEdgesOnCell: np.ndarray       # (C,)
    edgesOnCell: np.ndarray        # (6, C)
    verticesOnCell: np.ndarray     # (6, C)
    cellsOnCell: np.ndarray        # (6, C)
    cellsOnVertex: np.ndarray      # (3, V)
    edgesOnVertex: np.ndarray      # (3, V)
    kiteAreasOnVertex: np.ndarray  # (3, V)
    nEdgesOnEdge: np.ndarray       # (E,)
    edgesOnEdge: np.ndarray        # (12, E)
    weightsOnEdge: np.ndarray      # (12, E)
    dcEdge: np.ndarray
    dvEdge: np.ndarray
    angleEdge: np.ndarray
    areaCell: np.ndarray
    areaTriangle: np.ndarray
    meshDensity: np.ndarray

    maxEdges: int = MAX_EDGES
    maxEdges2: int = MAX_EDGES2
    vertexDegree: int = VERTEX_DEGREE

    # names of the arrays the reference driver reads (src/mpas_sw_driver.py:87-119)
    FIELDS = (
        "latCell", "lonCell", "xCell", "yCell", "zCell",
        "latEdge", "lonEdge", "xEdge", "yEdge", "zEdge",
        "latVertex", "lonVertex", "xVertex", "yVertex", "zVertex",
        "cellsOnEdge", "nEdgesOnCell", "nEdgesOnEdge", "edgesOnCell", "edgesOnEdge",
        "weightsOnEdge", "dvEdge", "dcEdge", "angleEdge", "areaCell", "areaTriangle",
        "cellsOnCell", "verticesOnCell", "verticesOnEdge", "edgesOnVertex",
        "cellsOnVertex", "kiteAreasOnVertex", "meshDensity")

    def digest(self) -> str:
        h = hashlib.sha256()
        for name in self.FIELDS:
            h.update(np.ascontiguousarray(getattr(self, name)).tobytes())
        return h.hexdigest()

    def permuted(self, pc, pe, pv) -> "Mesh":
        """Return the same mesh with entities renumbered: new entity ``k`` is
        old entity ``p[k]`` (0-based permutations of cells/edges/vertices)."""
        pc = np.asarray(pc); pe = np.asarray(pe); pv = np.asarray(pv)
        ic = np.empty_like(pc); ic[pc] = np.arange(pc.size)
        ie = np.empty_like(pe); ie[pe] = np.arange(pe.size)
        iv = np.empty_like(pv); iv[pv] = np.arange(pv.size)

        def remap(idx, inv):
            out = idx.copy()
            m = idx > 0
            out[m] = inv[idx[m] - 1] + 1
            return out

        kw = {}
        for f in dataclasses.fields(self):
            kw[f.name] = getattr(self, f.name)
        for name in ("xCell", "yCell", "zCell", "latCell", "lonCell", "areaCell", "meshDensity", "nEdgesOnCell"):
            kw[name] = kw[name][pc]
        for name in ("xEdge", "yEdge", "zEdge", "latEdge", "lonEdge", "dcEdge", "dvEdge", "angleEdge", "nEdgesOnEdge"):
            kw[name] = kw[name][pe]
        for name in ("xVertex", "yVertex", "zVertex", "latVertex", "lonVertex", "areaTriangle"):
            kw[name] = kw[name][pv]
        kw["cellsOnEdge"] = remap(self.cellsOnEdge[:, pe], ic)
        kw["verticesOnEdge"] = remap(self.verticesOnEdge[:, pe], iv)
        kw["edgesOnEdge"] = remap(self.edgesOnEdge[:, pe], ie)
        kw["weightsOnEdge"] = self.weightsOnEdge[:, pe]
        kw["edgesOnCell"] = remap(self.edgesOnCell[:, pc], ie)
        kw["verticesOnCell"] = remap(self.verticesOnCell[:, pc], iv)
        kw["cellsOnCell"] = remap(self.cellsOnCell[:, pc], ic)
        kw["cellsOnVertex"] = remap(self.cellsOnVertex[:, pv], ic)
        kw["edgesOnVertex"] = remap(self.edgesOnVertex[:, pv], ie)
        kw["kiteAreasOnVertex"] = self.kiteAreasOnVertex[:, pv]
        return Mesh(**kw)


def _latlon(x):
    lat = np.arcsin(np.clip(x[:, 2], -1.0, 1.0))
    lon = np.arctan2(x[:, 1], x[:, 0])
    lon = np.where(lon < 0.0, lon + 2.0 * np.pi, lon)
    return lat, lon


def icosahedral_mesh(level: int) -> Mesh:
    """Build the level-``level`` bisected-icosahedron Voronoi mesh."""
    if level < 0 or level > 11:
        raise ValueError("level must be in 0..11")
    pts, tri = _icosahedron()
    for _ in range(level):
        pts, tri = _bisect(pts, tri)
    C = pts.shape[0]
    V = tri.shape[0]

    # ---- Voronoi vertices: circumcentres of the Delaunay triangles
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    xv = _normalize(np.cross(b - a, c - a))

    # ---- Delaunay edges <-> Voronoi edges
    te = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], axis=0)  # directed, ccw in tri
    tri_of = np.concatenate([np.arange(V)] * 3)
    es = np.sort(te, axis=1)
    uniq, inv = _unique_pairs(es, C)
    inv = inv.reshape(-1)
    E = uniq.shape[0]
    assert E == 30 * 4 ** level and C == 10 * 4 ** level + 2 and V == 20 * 4 ** level
    c1 = uniq[:, 0].copy()
    c2 = uniq[:, 1].copy()
    # the two triangles sharing each Delaunay edge
    order = np.argsort(inv, kind="stable")
    tpair = tri_of[order].reshape(E, 2)
    xe = _normalize(pts[c1] + pts[c2])
    k = xe
    nvec = pts[c2] - pts[c1]
    nvec = _normalize(nvec - np.einsum("ij,ij->i", nvec, k)[:, None] * k)
    tvec = np.cross(k, nvec)
    va, vb = tpair[:, 0], tpair[:, 1]
    swap = np.einsum("ij,ij->i", xv[vb] - xv[va], tvec) < 0
    v1 = np.where(swap, vb, va)
    v2 = np.where(swap, va, vb)

    dc = _arc(pts[c1], pts[c2])
    dv = _arc(xv[v1], xv[v2])
    lat_e, lon_e = _latlon(xe)
    east = np.stack([-np.sin(lon_e), np.cos(lon_e), np.zeros_like(lon_e)], 1)
    north = np.stack([-np.sin(lat_e) * np.cos(lon_e), -np.sin(lat_e) * np.sin(lon_e), np.cos(lat_e)], 1)
    angle = np.arctan2(np.einsum("ij,ij->i", nvec, north), np.einsum("ij,ij->i", nvec, east))

    # ---- per-vertex (triangle) connectivity, counter-clockwise
    cov = tri.copy()  # already ccw
    # edge id of each directed triangle side: side i joins cov[i] and cov[i+1]
    side_edge = inv.reshape(3, V).T  # (V,3): sides (0,1),(1,2),(2,0)
    eov = side_edge
    area_tri = _tri_area(a, b, c)
    # kites: part of triangle v inside cell cov[i]: quad (c_i, m_{i,i+1}, v, m_{i-1,i})
    kites = np.empty((V, 3))
    for i in range(3):
        ci = pts[cov[:, i]]
        m_next = xe[eov[:, i]]
        m_prev = xe[eov[:, (i + 2) % 3]]
        kites[:, i] = _tri_area(ci, m_next, xv) + _tri_area(ci, xv, m_prev)

    # ---- per-cell connectivity, counter-clockwise around the cell
    # incidence cell -> edges
    inc_c = np.concatenate([c1, c2])
    inc_e = np.concatenate([np.arange(E), np.arange(E)])
    nbr = np.concatenate([c2, c1])
    # local frame at each cell
    P = pts[inc_c]
    ref = np.where(np.abs(P[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    ex = _normalize(np.cross(ref, P))
    ey = np.cross(P, ex)
    d = xe[inc_e] - P
    ang = np.arctan2(np.einsum("ij,ij->i", d, ey), np.einsum("ij,ij->i", d, ex))
    o = np.lexsort((ang, inc_c))
    inc_c, inc_e, nbr, ang = inc_c[o], inc_e[o], nbr[o], ang[o]
    nEoC = np.bincount(inc_c, minlength=C)
    start = np.concatenate([[0], np.cumsum(nEoC)[:-1]])
    slot = np.arange(inc_c.size) - start[inc_c]
    eoc = np.zeros((C, MAX_EDGES), dtype=np.int64) - 1
    coc = np.zeros((C, MAX_EDGES), dtype=np.int64) - 1
    eoc[inc_c, slot] = inc_e
    coc[inc_c, slot] = nbr
    # vertex between edgesOnCell[i] and edgesOnCell[i+1]: the vertex of edge i
    # that is also on edge i+1
    voc = np.zeros((C, MAX_EDGES), dtype=np.int64) - 1
    ev = np.stack([v1, v2], 1)
    for i in range(MAX_EDGES):
        has = i < nEoC
        cells = np.nonzero(has)[0]
        ei = eoc[cells, i]
        inext = (i + 1) % nEoC[cells]
        ej = eoc[cells, inext]
        a0, a1 = ev[ei, 0], ev[ei, 1]
        b0, b1 = ev[ej, 0], ev[ej, 1]
        common = np.where((a0 == b0) | (a0 == b1), a0, a1)
        voc[cells, i] = common
    # areaCell: sum of the cell's kites, so that the kite fractions R(i,v)
    # used by the TRiSK weights add up to one for every cell
    area_cell = np.zeros(C)
    for i in range(3):
        np.add.at(area_cell, cov[:, i], kites[:, i])

    # ---- TRiSK edgesOnEdge / weightsOnEdge
    # kite area of (cell, vertex): look up via the vertex's cellsOnVertex
    def kite_of(cells, verts):
        out = np.zeros(cells.size)
        for i in range(3):
            m = cov[verts, i] == cells
            out[m] = kites[verts[m], i]
        return out

    eoe = np.zeros((E, MAX_EDGES2), dtype=np.int64) - 1
    woe = np.zeros((E, MAX_EDGES2))
    nEoE = np.zeros(E, dtype=np.int64)
    # position of edge e in each of its cells' edge lists
    pos1 = np.zeros(E, dtype=np.int64)
    pos2 = np.zeros(E, dtype=np.int64)
    for i in range(MAX_EDGES):
        m = eoc[c1, i] == np.arange(E)
        pos1[m] = i
        m = eoc[c2, i] == np.arange(E)
        pos2[m] = i
    edges = np.arange(E)
    for which, cell, pos in ((0, c1, pos1), (1, c2, pos2)):
        n = nEoC[cell]
        s_e = 1.0 if which == 0 else -1.0  # normal of e points out of c1
        rsum = np.zeros(E)
        for kk in range(1, MAX_EDGES):
            act = kk < n
            idx = np.nonzero(act)[0]
            cc = cell[idx]
            # vertex passed between e_{kk-1} and e_kk = verticesOnCell[pos+kk-1]
            vpass = voc[cc, (pos[idx] + kk - 1) % n[idx]]
            rsum[idx] += kite_of(cc, vpass) / area_cell[cc]
            ek = eoc[cc, (pos[idx] + kk) % n[idx]]
            s_k = np.where(c1[ek] == cc, 1.0, -1.0)
            w = (0.5 - rsum[idx]) * s_e * s_k * dv[ek] / dc[idx]
            j = nEoE[idx]
            eoe[idx, j] = ek
            woe[idx, j] = w
            nEoE[idx] += 1

    lat_c, lon_c = _latlon(pts)
    lat_v, lon_v = _latlon(xv)

    def one(x):
        return (x + 1).astype(np.int32)

    def conn(x):
        y = np.where(x >= 0, x + 1, 0).astype(np.int32)
        return np.ascontiguousarray(y.T)

    return Mesh(
        level=level, nCells=C, nEdges=E, nVertices=V,
        xCell=pts[:, 0].copy(), yCell=pts[:, 1].copy(), zCell=pts[:, 2].copy(),
        latCell=lat_c, lonCell=lon_c,
        xEdge=xe[:, 0].copy(), yEdge=xe[:, 1].copy(), zEdge=xe[:, 2].copy(),
        latEdge=lat_e, lonEdge=lon_e,
        xVertex=xv[:, 0].copy(), yVertex=xv[:, 1].copy(), zVertex=xv[:, 2].copy(),
        latVertex=lat_v, lonVertex=lon_v,
        cellsOnEdge=np.ascontiguousarray(np.stack([one(c1), one(c2)], 0)),
        verticesOnEdge=np.ascontiguousarray(np.stack([one(v1), one(v2)], 0)),
        nEdgesOnCell=nEoC.astype(np.int32),
        edgesOnCell=conn(eoc), verticesOnCell=conn(voc), cellsOnCell=conn(coc),
        cellsOnVertex=conn(cov), edgesOnVertex=conn(eov),
        kiteAreasOnVertex=np.ascontiguousarray(kites.T),
        nEdgesOnEdge=nEoE.astype(np.int32),
        edgesOnEdge=conn(eoe), weightsOnEdge=np.ascontiguousarray(woe.T),
        dcEdge=dc, dvEdge=dv, angleEdge=angle,
        areaCell=area_cell, areaTriangle=area_tri,
        meshDensity=np.ones(C),
    )


def _hilbert3d_key(x, bits=16):
    """Skilling's transpose algorithm: 3-D Hilbert index of points in [-1,1]^3.

    The coordinates (bits <= 21 each) are transformed in uint32 with
    branch-free masks (m * P for "where m, xor P"); the 3*bits-bit index is
    interleaved in uint64."""
    assert bits <= 21
    q = np.clip(((x + 1.0) * 0.5 * (2 ** bits - 1)).round(), 0, 2 ** bits - 1).astype(np.uint32)
    X = [q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy()]
    n = 3
    M = np.uint32(1 << (bits - 1))
    Q = M
    while Q > 1:
        P = np.uint32(Q - 1)
        for i in range(n):
            m = ((X[i] & Q) != 0).view(np.uint8).astype(np.uint32)
            # invert where the bit is set
            X[0] ^= m * P
            # exchange where it is not
            t = ((X[0] ^ X[i]) & P) * (np.uint32(1) - m)
            X[0] ^= t
            X[i] ^= t
        Q = np.uint32(Q >> 1)
    for i in range(1, n):
        X[i] ^= X[i - 1]
    t = np.zeros_like(X[0])
    Q = M
    while Q > 1:
        m = ((X[n - 1] & Q) != 0).view(np.uint8).astype(np.uint32)
        t ^= m * np.uint32(Q - 1)
        Q = np.uint32(Q >> 1)
    for i in range(n):
        X[i] ^= t
    key = np.zeros(X[0].shape, dtype=np.uint64)
    for b in range(bits - 1, -1, -1):
        for i in range(n):
            key = (key << np.uint64(1)) | ((X[i] >> np.uint32(b)) & np.uint32(1)).astype(np.uint64)
    return key


def sfc_permutation(mesh: Mesh):
    """Space-filling-curve (3-D Hilbert) order of cells, edges and vertices.

    Returns 0-based permutations (pc, pe, pv): new entity k is old p[k].
    All three entity kinds are ordered by the same curve, so the edges and
    vertices of a run of consecutive cells are themselves nearly consecutive.
    """
    xc = np.stack([mesh.xCell, mesh.yCell, mesh.zCell], 1)
    xe = np.stack([mesh.xEdge, mesh.yEdge, mesh.zEdge], 1)
    xv = np.stack([mesh.xVertex, mesh.yVertex, mesh.zVertex], 1)
    pc = np.argsort(_hilbert3d_key(xc), kind="stable")
    pe = np.argsort(_hilbert3d_key(xe), kind="stable")
    pv = np.argsort(_hilbert3d_key(xv), kind="stable")
    return pc, pe, pv


_CACHE: dict = {}


def get_mesh(level: int) -> Mesh:
    """Cached ``icosahedral_mesh(level)`` (generator order, unit sphere)."""
    m = _CACHE.get(level)
    if m is None:
        m = icosahedral_mesh(level)
        _CACHE[level] = m
    return m


def union_mesh(meshes):
    """Pack several meshes into one index space (a disjoint union).

    Instance k's cells, edges and vertices follow those of instances < k;
    its connectivity is shifted by the same offsets, so one launch of any
    stage kernel advances every instance.  Returns (mesh, offsets) with
    offsets[k] = (cell0, edge0, vertex0) of instance k.
    """
    offs = []
    c0 = e0 = v0 = 0
    for m in meshes:
        offs.append((c0, e0, v0))
        c0 += m.nCells
        e0 += m.nEdges
        v0 += m.nVertices
    kw = {"level": -1, "nCells": c0, "nEdges": e0, "nVertices": v0}

    def cat(name):
        return np.concatenate([getattr(m, name) for m in meshes], axis=-1)

    def shifted(name, which):
        parts = []
        for m, (oc, oe, ov) in zip(meshes, offs):
            a = getattr(m, name).astype(np.int64)
            off = {"c": oc, "e": oe, "v": ov}[which]
            parts.append(np.where(a > 0, a + off, 0).astype(np.int32))
        return np.concatenate(parts, axis=-1)

    for f in dataclasses.fields(Mesh):
        n = f.name
        if n in kw or n in ("maxEdges", "maxEdges2", "vertexDegree"):
            continue
        if n in ("cellsOnEdge", "cellsOnCell", "cellsOnVertex"):
            kw[n] = shifted(n, "c")
        elif n in ("edgesOnCell", "edgesOnEdge", "edgesOnVertex"):
            kw[n] = shifted(n, "e")
        elif n in ("verticesOnEdge", "verticesOnCell"):
            kw[n] = shifted(n, "v")
        else:
            kw[n] = cat(n)
    return Mesh(**kw), offs


def union_sfc_permutation(meshes, offsets):
    """Per-instance space-filling-curve order, instances kept contiguous."""
    pcs, pes, pvs = [], [], []
    done = {}   # instances of one mesh share its curve order
    for m, (oc, oe, ov) in zip(meshes, offsets):
        if id(m) not in done:
            done[id(m)] = sfc_permutation(m)
        pc, pe, pv = done[id(m)]
        pcs.append(pc + oc)
        pes.append(pe + oe)
        pvs.append(pv + ov)
    return np.concatenate(pcs), np.concatenate(pes), np.concatenate(pvs)
