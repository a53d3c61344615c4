"""Shared set-up for the tests: meshes, initial conditions and the CPU oracle."""

from __future__ import annotations

import functools

import numpy as np

from mpas_sw_amd.mesh import get_mesh
from mpas_sw_amd.reconstruct import reconstruct_coefficients
from mpas_sw_amd.testcases import initial_conditions

SEED = 20261203


@functools.lru_cache(maxsize=None)
def case(level, test_case):
    """(mesh, phys) for a level and Williamson case; phys has the scaled
    metrics, f, h_s, initial h/u and the reconstruction coefficients."""
    mesh = get_mesh(level)
    phys = initial_conditions(mesh, test_case)
    phys["coeffs_reconstruct"] = reconstruct_coefficients(phys, mesh)
    return mesh, phys


def oracle_phys(phys):
    """The oracle wants coefficients flattened like the reference's
    (3, maxEdges, C) column-major array."""
    p = dict(phys)
    p["coeffs_reconstruct"] = np.transpose(phys["coeffs_reconstruct"], (2, 1, 0)).ravel()
    return p


def make_oracle(level, test_case, dt, **kw):
    from oracle.oracle import Oracle
    mesh, phys = case(level, test_case)
    o = Oracle(mesh, oracle_phys(phys), dt, test_case=test_case, **kw)
    o["h1"][:] = phys["h"]
    o["u1"][:] = phys["u"]
    o.init_block()
    return o


def perturbations(mesh, nmem, sigma_h=1.0, sigma_u=0.1, seed=SEED):
    """Seeded Gaussian perturbations per member, in reference order:
    member m draws dh then du from generator seed + m."""
    import torch
    hs, us = [], []
    for m in range(nmem):
        g = torch.Generator("cpu").manual_seed(seed + m)
        hs.append(torch.randn(mesh.nCells, generator=g, dtype=torch.float64).numpy() * sigma_h)
        us.append(torch.randn(mesh.nEdges, generator=g, dtype=torch.float64).numpy() * sigma_u)
    return np.stack(hs, 1), np.stack(us, 1)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden(name):
    """Load tests/golden/<name>.npz -> (arrays dict, meta dict); checks that
    the mesh generator still produces the mesh the vectors were made on."""
    import ast
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npz")
    z = np.load(path, allow_pickle=False)
    meta = ast.literal_eval(str(z["meta"]))
    arrays = {k: z[k] for k in z.files if k != "meta"}
    if "level" in meta:
        assert get_mesh(meta["level"]).digest() == meta["digest"], "mesh generator changed"
    return arrays, meta
