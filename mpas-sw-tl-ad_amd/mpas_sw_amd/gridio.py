"""MPAS grid files (``x1.<nCells>.grid.nc``) in NetCDF-3 through
``scipy.io.netcdf_file`` (netCDF4 is not available here).

The variables are the ones ``src/mpas_sw_driver.py:87-119`` reads, with the
MPAS on-disk layout: per-entity fields 1-D, connectivity ``(entity, slot)``
and 1-based, unit sphere.
"""

from __future__ import annotations

import numpy as np
from scipy.io import netcdf_file

from .mesh import Mesh

_SHAPES = {
    "latCell": ("nCells",), "lonCell": ("nCells",), "xCell": ("nCells",), "yCell": ("nCells",),
    "zCell": ("nCells",), "latEdge": ("nEdges",), "lonEdge": ("nEdges",), "xEdge": ("nEdges",),
    "yEdge": ("nEdges",), "zEdge": ("nEdges",), "latVertex": ("nVertices",),
    "lonVertex": ("nVertices",), "xVertex": ("nVertices",), "yVertex": ("nVertices",),
    "zVertex": ("nVertices",), "cellsOnEdge": ("nEdges", "TWO"), "nEdgesOnCell": ("nCells",),
    "nEdgesOnEdge": ("nEdges",), "edgesOnCell": ("nCells", "maxEdges"),
    "edgesOnEdge": ("nEdges", "maxEdges2"), "weightsOnEdge": ("nEdges", "maxEdges2"),
    "dvEdge": ("nEdges",), "dcEdge": ("nEdges",), "angleEdge": ("nEdges",),
    "areaCell": ("nCells",), "areaTriangle": ("nVertices",), "cellsOnCell": ("nCells", "maxEdges"),
    "verticesOnCell": ("nCells", "maxEdges"), "verticesOnEdge": ("nEdges", "TWO"),
    "edgesOnVertex": ("nVertices", "vertexDegree"), "cellsOnVertex": ("nVertices", "vertexDegree"),
    "kiteAreasOnVertex": ("nVertices", "vertexDegree"), "meshDensity": ("nCells",),
}
_INT = {"cellsOnEdge", "nEdgesOnCell", "nEdgesOnEdge", "edgesOnCell", "edgesOnEdge", "cellsOnCell",
        "verticesOnCell", "verticesOnEdge", "edgesOnVertex", "cellsOnVertex"}


def write_grid(mesh: Mesh, path: str) -> None:
    """Write ``mesh`` as an MPAS grid file (NetCDF-3 64-bit offset)."""
    with netcdf_file(path, "w", version=2) as f:
        dims = {"nCells": mesh.nCells, "nEdges": mesh.nEdges, "nVertices": mesh.nVertices,
                "maxEdges": mesh.maxEdges, "maxEdges2": mesh.maxEdges2,
                "vertexDegree": mesh.vertexDegree, "TWO": 2}
        for k, v in dims.items():
            f.createDimension(k, v)
        f.on_a_sphere = b"YES"
        f.sphere_radius = 1.0
        for name, shp in _SHAPES.items():
            arr = getattr(mesh, name)
            if len(shp) == 2:
                arr = arr.T               # (slot, entity) -> (entity, slot)
            typ = "i" if name in _INT else "d"
            var = f.createVariable(name, typ, shp)
            var[:] = np.ascontiguousarray(arr, dtype=np.int32 if typ == "i" else np.float64)


def read_grid(path: str) -> Mesh:
    """Read an MPAS grid file into a ``Mesh`` (reference layout)."""
    with netcdf_file(path, "r", mmap=False) as f:
        d = f.dimensions
        kw = {}
        for name, shp in _SHAPES.items():
            # NetCDF-3 data are big-endian: convert to native int32 / float64
            a = np.array(f.variables[name][:]).astype(np.int32 if name in _INT else np.float64)
            kw[name] = np.ascontiguousarray(a.T if len(shp) == 2 else a)
        C = int(d["nCells"])
        return Mesh(level=-1, nCells=C, nEdges=int(d["nEdges"]), nVertices=int(d["nVertices"]),
                    maxEdges=int(d["maxEdges"]), maxEdges2=int(d["maxEdges2"]),
                    vertexDegree=int(d["vertexDegree"]), **kw)


def dims(path: str) -> dict:
    with netcdf_file(path, "r", mmap=False) as f:
        return {k: int(v) for k, v in f.dimensions.items() if v is not None}


def variable(path: str, name: str) -> np.ndarray:
    with netcdf_file(path, "r", mmap=False) as f:
        a = np.array(f.variables[name][:])
        return a.astype(a.dtype.newbyteorder("="))


def attribute(path: str, name: str, default=None):
    with netcdf_file(path, "r", mmap=False) as f:
        return getattr(f, name, default)
