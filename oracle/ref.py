"""Runner for the reference Fortran model built into ``oracle/_ref/msw_ref``.

Test infrastructure (it checks results), not product code: only ``tests/``,
``tests/golden/make_golden.py`` and the ``cpu_baseline`` leg of ``bench.py``
use it.  It writes the flat input file that ``oracle/ref_main.f90`` reads
(mesh arrays in the order ``src/mpas_sw_driver.py:87-119`` sets them, plus
the configuration of ``read_configs``, ``src/mpas_sw_driver.py:12-30``), runs
the executable in a fresh process (the Fortran module can be allocated only
once per process, ``src/mpas_sw_module.f90:211-367``) and reads the named
float64 records it writes.
"""

from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "msw_ref")
MAGIC = 1297307462


def available() -> bool:
    return os.path.isfile(REF_EXE) and os.access(REF_EXE, os.X_OK)


# order in which ref_main reads the diagnostic adjoint seeds
DIAG_SEED_ORDER = ("v", "h_edge", "h_vertex", "circulation", "vorticity", "divergence", "ke",
                   "pv_edge", "pv_vertex", "pv_cell", "vorticity_cell", "gradPVn", "gradPVt")
DIAG_SEED_KINDS = "eevvvccevccee"   # entity of each, in that order


def write_input(path, mesh, test_case, dt, apvm=0.5, visc2=0.0, visc4=0.0, adv_order=2,
                monotonic=False, wind_stress=False, bottom_drag=False,
                state=None, tl=None, ad=None, u_src=None, recon_seeds=None, diag_seeds=None,
                scale_with_mesh=False):
    hdr = np.zeros(16, np.int32)
    hdr[0] = MAGIC
    hdr[1] = 1
    hdr[2:8] = [mesh.nCells, mesh.nEdges, mesh.nVertices, mesh.maxEdges, mesh.maxEdges2, mesh.vertexDegree]
    hdr[8] = test_case
    hdr[9] = adv_order
    hdr[10] = int(monotonic)
    hdr[11] = int(wind_stress)
    hdr[12] = int(bottom_drag)
    hdr[13] = state is not None
    hdr[14] = tl is not None
    hdr[15] = ad is not None
    cfg = np.zeros(8, np.float64)
    cfg[:4] = [dt, apvm, visc2, visc4]
    cfg[4] = 1.0 if u_src is not None else 0.0
    cfg[5] = 1.0 if recon_seeds is not None else 0.0
    cfg[6] = 1.0 if diag_seeds is not None else 0.0
    cfg[7] = 1.0 if scale_with_mesh else 0.0
    f8 = lambda x: np.asarray(x, np.float64).ravel(order="C")
    i4 = lambda x: np.asarray(x, np.int32).ravel(order="C")
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(cfg.tobytes())
        for n in ("latCell", "lonCell", "xCell", "yCell", "zCell",
                  "latEdge", "lonEdge", "xEdge", "yEdge", "zEdge",
                  "latVertex", "lonVertex", "xVertex", "yVertex", "zVertex"):
            f.write(f8(getattr(mesh, n)).tobytes())
        # (slot, entity) arrays: Fortran column-major == C-order of the (entity, slot) transpose
        f.write(i4(mesh.cellsOnEdge.T).tobytes())
        f.write(i4(mesh.nEdgesOnCell).tobytes())
        f.write(i4(mesh.nEdgesOnEdge).tobytes())
        f.write(i4(mesh.edgesOnCell.T).tobytes())
        f.write(i4(mesh.edgesOnEdge.T).tobytes())
        f.write(f8(mesh.weightsOnEdge.T).tobytes())
        for n in ("dvEdge", "dcEdge", "angleEdge", "areaCell", "areaTriangle"):
            f.write(f8(getattr(mesh, n)).tobytes())
        for n in ("cellsOnCell", "verticesOnCell", "verticesOnEdge", "edgesOnVertex", "cellsOnVertex"):
            f.write(i4(getattr(mesh, n).T).tobytes())
        f.write(f8(mesh.kiteAreasOnVertex.T).tobytes())
        f.write(f8(mesh.meshDensity).tobytes())
        for pair in (state, tl, ad):
            if pair is not None:
                f.write(f8(pair[0]).tobytes())
                f.write(f8(pair[1]).tobytes())
        if u_src is not None:
            f.write(f8(u_src).tobytes())
        if recon_seeds is not None:   # (5, C): X, Y, Z, zonal, meridional adjoint seeds
            f.write(f8(np.asarray(recon_seeds, np.float64).reshape(5, -1)).tobytes())
        if diag_seeds is not None:    # dict name -> array, the 13 diagnostics of DIAG_SEED_ORDER
            size = {"c": mesh.nCells, "e": mesh.nEdges, "v": mesh.nVertices}
            for name, kind in zip(DIAG_SEED_ORDER, DIAG_SEED_KINDS):
                a = f8(diag_seeds[name])
                if a.size != size[kind]:   # the reader takes raw streams: a wrong size would misalign
                    raise ValueError(f"diagnostic seed {name}: {a.size} values, expected {size[kind]}")
                f.write(a.tobytes())


def read_output(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        name = data[off:off + 32].decode().strip()
        off += 32
        n = int(np.frombuffer(data, np.int64, 1, off)[0])
        off += 8
        out[name] = np.frombuffer(data, np.float64, n, off).copy()
        off += 8 * n
    return out


def run(mesh, mode, nsteps=0, timeout=3600, **kw):
    """Run the reference in ``mode`` (see oracle/ref_main.f90) and return its records."""
    if not available():
        raise RuntimeError("reference build missing: run `make -C oracle ref`")
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        write_input(fin, mesh, **kw)
        env = dict(os.environ)
        # L7+ automatic arrays in the reference RK4 routines need a big stack
        cmd = ["bash", "-c", 'ulimit -s unlimited; exec "$0" "$@"', REF_EXE, fin, mode, str(nsteps), fout]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"msw_ref {mode} failed ({r.returncode}): {r.stdout}\n{r.stderr}")
        return read_output(fout)
