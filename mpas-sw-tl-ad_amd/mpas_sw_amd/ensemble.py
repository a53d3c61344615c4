"""Uneven batches: many mesh instances of different sizes in one packed
index space (``mesh.union_mesh``), stepped by one launch per stage kernel.

Each instance keeps its own space-filling-curve order inside its contiguous
block of the packed arrays; per-instance initial conditions (Williamson
cases) are concatenated the same way.
"""

from __future__ import annotations

import numpy as np

from .mesh import get_mesh, union_mesh, union_sfc_permutation
from .testcases import initial_conditions

_FIELDS = ("dcEdge", "dvEdge", "areaCell", "areaTriangle", "kiteAreasOnVertex", "fVertex",
           "fEdge", "h_s", "lonCell", "h", "u")


def packed_instances(specs):
    """specs: list of (level, test_case).  Returns (mesh, phys, offsets, perm)
    ready for ``SWModel(mesh, phys, dt, perm=perm)``."""
    meshes = [get_mesh(level) for level, _ in specs]
    mesh, offsets = union_mesh(meshes)
    ics = {}
    phys = {}
    for (level, tc), m in zip(specs, meshes):
        key = (level, tc)
        if key not in ics:
            ics[key] = initial_conditions(m, tc)
    for name in _FIELDS:
        phys[name] = np.concatenate([ics[k][name] for k in specs], axis=-1)
    perm = union_sfc_permutation(meshes, offsets)
    return mesh, phys, offsets, perm


def split(arr, offsets, sizes):
    """Views of a packed per-entity array, one per instance."""
    return [arr[o:o + n] for o, n in zip(offsets, sizes)]
