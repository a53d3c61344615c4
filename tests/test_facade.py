"""The drop-in module object: f2py attribute semantics (CPU) and the GPU
stepping routines against the oracle."""

import importlib

import numpy as np
import pytest

from helpers import case, make_oracle, perturbations, rel_l2


def fresh_module():
    import module_sw_mpas._facade as f
    importlib.reload(f)
    return f.mpas_sw_module


def load_mesh(mpsw, mesh):
    """What src/mpas_sw_driver.py:33-119 does, from a generated mesh."""
    mpsw.nvertlevels = 1
    mpsw.ncells, mpsw.nedges, mpsw.nvertices = mesh.nCells, mesh.nEdges, mesh.nVertices
    mpsw.maxedges, mpsw.maxedges2, mpsw.vertexdegree = 6, 12, 3
    mpsw.r3 = 3
    mpsw.ntracers = 1
    mpsw.nverticessolve = mesh.nVertices          # not a module variable: goes aside
    mpsw.var_allocation()
    mpsw.on_a_sphere = True
    mpsw.is_periodic = False
    mpsw.sphere_radius = 1.0
    for name in ("latCell", "lonCell", "xCell", "yCell", "zCell", "latEdge", "lonEdge", "xEdge",
                 "yEdge", "zEdge", "latVertex", "lonVertex", "xVertex", "yVertex", "zVertex",
                 "cellsOnEdge", "nEdgesOnCell", "nEdgesOnEdge", "edgesOnCell", "edgesOnEdge",
                 "weightsOnEdge", "dvEdge", "dcEdge", "angleEdge", "areaCell", "areaTriangle",
                 "cellsOnCell", "verticesOnCell", "verticesOnEdge", "edgesOnVertex",
                 "cellsOnVertex", "kiteAreasOnVertex", "meshDensity"):
        setattr(mpsw, name.lower(), getattr(mesh, name))


def test_attribute_semantics():
    mpsw = fresh_module()
    assert mpsw.a == 6371229.0 and mpsw.gravity == 9.80616 and mpsw.omega == 7.29212e-5
    assert mpsw.u is None                      # declared, not allocated
    mesh, _ = case(2, 2)
    load_mesh(mpsw, mesh)
    assert mpsw.u.shape == (2, 1, mesh.nEdges) and mpsw.u.flags.f_contiguous
    assert mpsw.cellsonedge.dtype == np.int32 and mpsw.cellsonedge.shape == (2, mesh.nEdges)
    mpsw.u[0, 0] = 3.0                         # aliasing view
    assert np.all(mpsw.u[0, 0] == 3.0)
    mpsw.config_dt = "600"                     # cast to the Fortran type
    assert mpsw.config_dt == 600.0
    mpsw.config_h_ScaleWithMesh = True         # mixed case never reaches the module
    assert mpsw.config_h_scalewithmesh is False
    assert mpsw.nverticessolve == mesh.nVertices
    with pytest.raises(RuntimeError):
        mpsw.var_allocation()                  # allocating twice is an error
    with pytest.raises(AttributeError):
        mpsw.a = 1.0


@pytest.mark.parametrize("tc", [2, 5, 6])
def test_test_cases_match_reference(tc):
    from helpers import golden
    g, _ = golden("l3_initial_conditions")
    mpsw = fresh_module()
    mesh, _ = case(3, tc)
    load_mesh(mpsw, mesh)
    getattr(mpsw, f"sw_test_case_{tc}")()
    np.testing.assert_array_equal(mpsw.h[0, 0], g[f"tc{tc}_h"])
    np.testing.assert_array_equal(mpsw.u[0, 0], g[f"tc{tc}_u"])
    np.testing.assert_array_equal(mpsw.fvertex, g[f"tc{tc}_fVertex"])


@pytest.mark.gpu
def test_facade_steps_on_gpu():
    print("facade L3 tc5", flush=True)
    mesh, phys = case(3, 5)
    mpsw = fresh_module()
    load_mesh(mpsw, mesh)
    mpsw.config_test_case = 5
    mpsw.config_dt = 900.0
    mpsw.sw_test_case_5()
    mpsw.var_allocation_tlm()
    mpsw.var_allocation_adj()
    mpsw.sw_mpas_init_block()
    o = make_oracle(3, 5, 900.0)
    np.testing.assert_array_equal(mpsw.ureconstructzonal[0], o["uReconstructZonal"])
    for _ in range(3):
        mpsw.sw_rk4()
    o.rk4(3)
    np.testing.assert_array_equal(mpsw.h[0, 0], o["h1"])
    np.testing.assert_array_equal(mpsw.u[1, 0], o["u1"])
    np.testing.assert_array_equal(mpsw.pv_edge[0], o["pv_edge"])
    np.testing.assert_array_equal(mpsw.ureconstructmeridional[0], o["uReconstructMeridional"])
    # TLM from the current state
    dh, du = perturbations(mesh, 1)
    mpsw.h_tl[0, 0] = dh[:, 0]
    mpsw.u_tl[0, 0] = du[:, 0]
    o["h_tl1"][:] = dh[:, 0]
    o["u_tl1"][:] = du[:, 0]
    mpsw.sw_rk4_tlm()
    o.rk4_tlm()
    np.testing.assert_array_equal(mpsw.h_tl[0, 0], o["h_tl1"])
    # adjoint of the next step
    o["h_ad1"][:] = dh[:, 0]
    o["u_ad1"][:] = du[:, 0]
    mpsw.h_ad[0, 0] = dh[:, 0]
    mpsw.u_ad[0, 0] = du[:, 0]
    mpsw.sw_rk4_adj()
    o.rk4_adj()
    assert rel_l2(mpsw.h_ad[0, 0], o["h_ad1"]) < 1e-11
    np.testing.assert_array_equal(mpsw.h[1, 0], o["h2"])
    lhs, rhs = mpsw.link_tlm_adj()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.gpu
@pytest.mark.parametrize("forcing", [False, True])
def test_facade_tendency_arrays_match_oracle(forcing):
    """mpsw.tend_h/tend_u (and _tl) after sw_rk4, sw_rk4_tlm and sw_rk4_adj
    hold the last RK stage's tendencies, as the reference's module arrays do
    (src/mpas_sw_module.f90:577-586, :700-712, :838-848; the tendency
    adjoint leaves them alone), and tend_*_ad are zero: equal to the oracle's
    module arrays bit for bit, level 3 TC5, also with del2, del4, bottom drag
    and wind stress on."""
    print(f"facade tendency arrays L3 tc5 forcing={forcing}", flush=True)
    mesh, phys = case(3, 5)
    mpsw = fresh_module()
    load_mesh(mpsw, mesh)
    mpsw.config_test_case = 5
    mpsw.config_dt = 900.0
    kw = {}
    if forcing:
        usrc = 0.1 * np.cos(mesh.latEdge) * np.cos(mesh.angleEdge) + 0.02 * np.sin(3.0 * mesh.lonEdge)
        kw = dict(visc2=2.0e5, visc4=1.0e15, wind_stress=True, bottom_drag=True, u_src=usrc)
        mpsw.config_h_mom_eddy_visc2 = kw["visc2"]
        mpsw.config_h_mom_eddy_visc4 = kw["visc4"]
        mpsw.config_wind_stress = True
        mpsw.config_bottom_drag = True
    mpsw.sw_test_case_5()
    mpsw.var_allocation_tlm()
    mpsw.var_allocation_adj()
    if forcing:
        mpsw.u_src[0] = usrc
    mpsw.sw_mpas_init_block()
    o = make_oracle(3, 5, 900.0, **kw)
    assert not np.any(mpsw.tend_u)                    # zero after var_allocation
    for _ in range(2):
        mpsw.sw_rk4()
    o.rk4(2)
    np.testing.assert_array_equal(mpsw.u[0, 0], o["u1"])
    assert np.any(o["tend_u"] != 0)
    np.testing.assert_array_equal(mpsw.tend_h[0], o["tend_h"])
    np.testing.assert_array_equal(mpsw.tend_u[0], o["tend_u"])
    dh, du = perturbations(mesh, 1)
    mpsw.h_tl[0, 0] = dh[:, 0]
    mpsw.u_tl[0, 0] = du[:, 0]
    o["h_tl1"][:] = dh[:, 0]
    o["u_tl1"][:] = du[:, 0]
    mpsw.sw_rk4_tlm()
    o.rk4_tlm()
    np.testing.assert_array_equal(mpsw.u_tl[0, 0], o["u_tl1"])
    for n in ("tend_h", "tend_u", "tend_h_tl", "tend_u_tl"):
        np.testing.assert_array_equal(getattr(mpsw, n)[0], o[n], err_msg=n)
    o["h_ad1"][:] = dh[:, 0]
    o["u_ad1"][:] = du[:, 0]
    mpsw.h_ad[0, 0] = dh[:, 0]
    mpsw.u_ad[0, 0] = du[:, 0]
    mpsw.sw_rk4_adj()
    o.rk4_adj()
    for n in ("tend_h", "tend_u"):
        np.testing.assert_array_equal(getattr(mpsw, n)[0], o[n], err_msg=n + " after sw_rk4_adj")
    for n in ("tend_h_ad", "tend_u_ad"):
        assert not np.any(getattr(mpsw, n)) and not np.any(o[n]), n


@pytest.mark.gpu
def test_facade_tl_reconstruction_against_reference_fortran():
    """mpsw.ureconstruct{zonal,meridional}_tl after sw_rk4_tlm, which the
    reference's TLM demo writes at every output time (demo/run-tlm.py:38-39),
    equal to the reference Fortran's (5 TC6 steps on level 3)."""
    from helpers import golden
    print("facade TL reconstruction L3 tc6", flush=True)
    g, meta = golden("tc6_l3_recon_tl")
    mesh, _ = case(3, 6)
    mpsw = fresh_module()
    load_mesh(mpsw, mesh)
    mpsw.config_test_case = 6
    mpsw.config_dt = meta["dt"]
    mpsw.sw_test_case_6()
    mpsw.var_allocation_tlm()
    mpsw.sw_mpas_init_block()
    np.testing.assert_array_equal(mpsw.h[0, 0], g["h0"])
    mpsw.h_tl[0, 0] = g["dh"]
    mpsw.u_tl[0, 0] = g["du"]
    for _ in range(meta["nsteps"]):
        mpsw.sw_rk4_tlm()
    np.testing.assert_array_equal(mpsw.u_tl[0, 0], g["u_tl"])
    np.testing.assert_array_equal(mpsw.ureconstructzonal_tl[0], g["zonal_tl"])
    np.testing.assert_array_equal(mpsw.ureconstructmeridional_tl[0], g["meridional_tl"])


@pytest.mark.gpu
def test_facade_diagnostic_adjoint_seeds_against_reference_fortran():
    """The reference's adjoint flow (ref run-adj: forward storing each step's
    start state, then sw_rk4_adj backwards) with seeds in the end-state
    adjoints and in all 13 diagnostic adjoint arrays: sw_rk4_adj consumes
    and zeroes the diagnostic seeds in its first call, and after three steps
    h_ad, u_ad are the reference Fortran's to 1e-11 (golden vectors written
    by the reference itself)."""
    print("facade diagnostic adjoint seeds L3 tc5 vs reference", flush=True)
    from helpers import golden
    g, meta = golden("tc5_l3_diag_seeds")
    n, dt = meta["nsteps"], meta["dt"]
    mesh, _ = case(3, 5)
    mpsw = fresh_module()
    load_mesh(mpsw, mesh)
    mpsw.config_test_case = 5
    mpsw.config_dt = dt
    mpsw.sw_test_case_5()
    mpsw.var_allocation_tlm()
    mpsw.var_allocation_adj()
    mpsw.sw_mpas_init_block()
    traj = []
    for _ in range(n):
        traj.append((mpsw.h[0, 0].copy(), mpsw.u[0, 0].copy()))
        mpsw.sw_rk4()
    mpsw.h_ad[0, 0] = g["dh"]
    mpsw.u_ad[0, 0] = g["du"]
    names = [k[5:] for k in g if k.startswith("seed_")]
    assert len(names) == 13
    for k in names:
        getattr(mpsw, k.lower() + "_ad")[0] = g[f"seed_{k}"]
    for s in range(n - 1, -1, -1):
        mpsw.h[0, 0] = traj[s][0]
        mpsw.u[0, 0] = traj[s][1]
        mpsw.sw_rk4_adj()
        if s == n - 1:
            for k in names:
                assert not getattr(mpsw, k.lower() + "_ad").any(), k
    assert rel_l2(mpsw.h_ad[0, 0], g["h_ad"]) < 1e-11
    assert rel_l2(mpsw.u_ad[0, 0], g["u_ad"]) < 1e-11


@pytest.mark.gpu
def test_facade_mesh_scaling_against_reference_fortran():
    """del2 and del4 viscosity with config_h_scalewithmesh on a non-uniform
    meshDensity, so every edge scales both terms by its own
    meshScalingDel2/4 factor (src/mpas_sw_dyncore.f90:362-392): after 6 steps
    the forward and the TLM equal the reference Fortran's bit for bit and the
    adjoint is within 1e-11 of it."""
    print("facade mesh-scaled viscosity L3 tc5 vs reference", flush=True)
    from helpers import golden
    g, meta = golden("tc5_l3_mesh_scaling")
    n, dt = meta["nsteps"], meta["dt"]
    mesh, _ = case(3, 5)

    def setup():
        mpsw = fresh_module()
        load_mesh(mpsw, mesh)
        mpsw.meshdensity = g["density"]
        mpsw.config_test_case = 5
        mpsw.config_dt = dt
        mpsw.config_h_mom_eddy_visc2 = meta["visc2"]
        mpsw.config_h_mom_eddy_visc4 = meta["visc4"]
        mpsw.config_h_scalewithmesh = True
        mpsw.sw_test_case_5()
        mpsw.var_allocation_tlm()
        mpsw.var_allocation_adj()
        mpsw.sw_mpas_init_block()
        assert mpsw.meshscalingdel2.min() < 0.8 and mpsw.meshscalingdel4.min() < 0.7
        return mpsw

    mpsw = setup()
    for _ in range(n):
        mpsw.sw_rk4()
    np.testing.assert_array_equal(mpsw.h[0, 0], g["h"])
    np.testing.assert_array_equal(mpsw.u[0, 0], g["u"])
    mpsw = setup()
    mpsw.h_tl[0, 0] = g["dh"]
    mpsw.u_tl[0, 0] = g["du"]
    for _ in range(n):
        mpsw.sw_rk4_tlm()
    np.testing.assert_array_equal(mpsw.h_tl[0, 0], g["h_tl"])
    np.testing.assert_array_equal(mpsw.u_tl[0, 0], g["u_tl"])
    np.testing.assert_array_equal(mpsw.h[0, 0], g["h_base"])
    mpsw = setup()
    traj = []
    for _ in range(n):
        traj.append((mpsw.h[0, 0].copy(), mpsw.u[0, 0].copy()))
        mpsw.sw_rk4()
    mpsw.h_ad[0, 0] = g["dh"]
    mpsw.u_ad[0, 0] = g["du"]
    for s in range(n - 1, -1, -1):
        mpsw.h[0, 0] = traj[s][0]
        mpsw.u[0, 0] = traj[s][1]
        mpsw.sw_rk4_adj()
    assert rel_l2(mpsw.h_ad[0, 0], g["h_ad"]) < 1e-11
    assert rel_l2(mpsw.u_ad[0, 0], g["u_ad"]) < 1e-11
