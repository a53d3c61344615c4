"""The reference-style driver (mpas_sw_driver / mpas_namelist) and the MPAS
grid file format."""

import os

import numpy as np
import pytest

from helpers import case, make_oracle

NML = """&sw_model
    config_test_case = {tc}
    config_time_integration = 'RK4'
    config_dt = {dt}
    config_calendar_type = 'gregorian_noleap'
    config_start_time = '2000-01-01_00:00:00'
    config_run_duration = '{dur}'
    config_h_ScaleWithMesh = false   ! lower-cased by the parser
    config_h_mom_eddy_visc2 = 0.0
    config_h_mom_eddy_visc4 = 0.0
    config_apvm_upwinding = 0.5
/
&io
    input = '{grid}'
    output = '{out}'
    output_interval = '01:00:00'
/
"""


def write_case(tmp_path, level, tc, dt, dur):
    from mpas_sw_amd.gridio import write_grid
    from mpas_sw_amd.mesh import get_mesh
    grid = os.path.join(tmp_path, f"x1.{get_mesh(level).nCells}.grid.nc")
    write_grid(get_mesh(level), grid)
    nml = os.path.join(tmp_path, "namelist.sw.test")
    with open(nml, "w") as f:
        f.write(NML.format(tc=tc, dt=dt, dur=dur, grid=grid, out=os.path.join(tmp_path, "out.nc")))
    return nml


def test_namelist_parser(tmp_path):
    import mpas_namelist
    nml_path = write_case(str(tmp_path), 2, 5, 900.0, "00_03:00:00")
    g = mpas_namelist.read_fortran_namelist(nml_path)
    assert g["sw_model"]["config_test_case"] == 5
    assert g["sw_model"]["config_dt"] == 900.0
    assert g["sw_model"]["config_h_scalewithmesh"] is False
    assert g["sw_model"]["config_run_duration"] == "00_03:00:00"
    nml = mpas_namelist.namelist(nml_path)
    assert nml.output_interval == 3600
    assert nml.nlatm["sw_model"]["config_dt"] == 900.0


def test_grid_roundtrip(tmp_path):
    from mpas_sw_amd.gridio import read_grid, write_grid
    from mpas_sw_amd.mesh import get_mesh
    m = get_mesh(3)
    p = os.path.join(str(tmp_path), "g.nc")
    write_grid(m, p)
    m2 = read_grid(p)
    for name in m.FIELDS:
        np.testing.assert_array_equal(getattr(m2, name), getattr(m, name), err_msg=name)
    assert m2.digest() == m.digest()


@pytest.mark.gpu
def test_run_sw_fwd_and_check_tlm(tmp_path):
    print("driver L3 tc5", flush=True)
    import importlib
    import module_sw_mpas._facade as fac
    importlib.reload(fac)
    import module_sw_mpas
    module_sw_mpas.mpas_sw_module = fac.mpas_sw_module
    import mpas_sw_driver
    importlib.reload(mpas_sw_driver)
    from scipy.io import netcdf_file
    nml = write_case(str(tmp_path), 3, 5, 900.0, "00_03:00:00")
    mpas_sw_driver.run_sw_fwd(nml)
    with netcdf_file(os.path.join(str(tmp_path), "out.nc"), "r", mmap=False) as f:
        h = np.array(f.variables["h"][:])
    # the driver loop includes the end time: 13 steps, outputs every hour
    o = make_oracle(3, 5, 900.0)
    o.rk4(12)
    np.testing.assert_array_equal(h[-1], o["h1"])
    assert h.shape[0] == 4
    # linearisation ratios tend to one
    importlib.reload(fac)
    module_sw_mpas.mpas_sw_module = fac.mpas_sw_module
    importlib.reload(mpas_sw_driver)
    res = mpas_sw_driver.check_tlm(nml, idx2=100, npts=9)
    r6 = res[6]
    assert abs(r6[1] - 1) < 1e-3 and abs(r6[2] - 1) < 1e-3


@pytest.mark.gpu
def test_run_sw_adj(tmp_path):
    """The flow of demo/run-adj.py (forward sweep keeping every step's start
    state, then sw_rk4_adj backwards from a unit seed at one cell and one
    edge) against the oracle's adjoint sweep over the same 13 steps."""
    print("driver run_sw_adj L3 tc5", flush=True)
    import importlib
    import module_sw_mpas._facade as fac
    importlib.reload(fac)
    import module_sw_mpas
    module_sw_mpas.mpas_sw_module = fac.mpas_sw_module
    import mpas_sw_driver
    importlib.reload(mpas_sw_driver)
    from helpers import rel_l2
    nml = write_case(str(tmp_path), 3, 5, 900.0, "00_03:00:00")
    u_ad, h_ad = mpas_sw_driver.run_sw_adj(nml, seed_index=7)
    o = make_oracle(3, 5, 900.0)
    traj = []
    for _ in range(13):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = 0.0
    o["u_ad1"][:] = 0.0
    o["h_ad1"][7] = 1.0
    o["u_ad1"][7] = 1.0
    for s in range(12, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    assert rel_l2(h_ad, o["h_ad1"]) < 1e-11
    assert rel_l2(u_ad, o["u_ad1"]) < 1e-11
