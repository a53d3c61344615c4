"""Pins the CPU oracle to the reference Fortran.

The golden vectors in tests/golden/ were written by the reference model
itself (oracle/_ref/msw_ref, compiled from the reference's src/*.f90 files,
see tests/golden/make_golden.py).  The oracle restates the same loops in the
same order, so every comparison here is bit-for-bit.
"""

import numpy as np
import pytest

from helpers import case, golden, make_oracle, oracle_phys


def _oracle_from(level, tc, dt, h0, u0):
    from oracle.oracle import Oracle
    mesh, phys = case(level, tc)
    o = Oracle(mesh, oracle_phys(phys), dt, test_case=tc)
    o["h1"][:] = h0
    o["u1"][:] = u0
    o.init_block()
    return o


def test_tc2_l5_one_day_bitwise():
    g, meta = golden("tc2_l5_fwd_1day")
    o = _oracle_from(5, 2, meta["dt"], g["h0"], g["u0"])
    o.rk4(meta["nsteps"])
    np.testing.assert_array_equal(o["h1"], g["h"])
    np.testing.assert_array_equal(o["u1"], g["u"])


def test_routines_bitwise():
    g, meta = golden("tc5_l2_routines")
    mesh, phys = case(2, 5)
    o = _oracle_from(2, 5, meta["dt"], g["ic_h"], g["ic_u"])
    o.diag(g["ic_h"], g["ic_u"])
    o.tend(g["ic_h"], g["ic_u"])
    for k in ("v", "h_edge", "h_vertex", "circulation", "vorticity", "divergence", "ke", "pv_edge",
              "pv_vertex", "pv_cell", "vorticity_cell", "gradPVn", "gradPVt", "tend_h", "tend_u"):
        np.testing.assert_array_equal(o[k], g[f"diag_{k}"], err_msg=k)
    o.diag_tlm(g["ic_h"], g["dh"], g["ic_u"], g["du"])
    o.tend_tlm(g["ic_h"], g["dh"], g["ic_u"], g["du"])
    for k in ("v_tl", "h_edge_tl", "h_vertex_tl", "vorticity_tl", "divergence_tl", "ke_tl",
              "pv_edge_tl", "pv_vertex_tl", "pv_cell_tl", "gradPVn_tl", "gradPVt_tl"):
        np.testing.assert_array_equal(o[k], g[f"diagtl_{k}"], err_msg=k)
    ha, ua = o.tend_full_adj(g["ic_h"], g["ic_u"], g["gh"], g["gu"])
    np.testing.assert_array_equal(ha, g["tad_h_ad"])
    np.testing.assert_array_equal(ua, g["tad_u_ad"])


@pytest.mark.parametrize("tc", [5, 6])
def test_tlm_and_adjoint_windows_bitwise(tc):
    g, meta = golden(f"tc{tc}_l4_tlm_adj")
    n, dt = meta["nsteps"], meta["dt"]
    o = _oracle_from(4, tc, dt, g["h0"], g["u0"])
    o["h_tl1"][:] = g["dh"]
    o["u_tl1"][:] = g["du"]
    o.rk4_tlm(n)
    np.testing.assert_array_equal(o["h_tl1"], g["h_tl"])
    np.testing.assert_array_equal(o["u_tl1"], g["u_tl"])
    np.testing.assert_array_equal(o["h1"], g["h"])
    o = _oracle_from(4, tc, dt, g["h0"], g["u0"])
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = g["dh"]
    o["u_ad1"][:] = g["du"]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    np.testing.assert_array_equal(o["h_ad1"], g["h_ad"])
    np.testing.assert_array_equal(o["u_ad1"], g["u_ad"])


def test_reconstruction_bitwise():
    g, meta = golden("l3_initial_conditions")
    o = make_oracle(3, 2, 600.0)
    np.testing.assert_array_equal(o["uReconstructZonal"], g["tc2_uReconstructZonal"])
    np.testing.assert_array_equal(o["uReconstructMeridional"], g["tc2_uReconstructMeridional"])


def test_tl_reconstruction_bitwise():
    """The reconstructed TL velocity sw_rk4_tlm leaves after its last
    mpas_reconstruct_2d_tlm (src/mpas_sw_dyncore_tlm.f90:772-870), 5 steps of
    TC6 on level 3; and the same operator applied directly to that TL state."""
    g, meta = golden("tc6_l3_recon_tl")
    o = _oracle_from(3, 6, meta["dt"], g["h0"], g["u0"])
    o["h_tl1"][:] = g["dh"]
    o["u_tl1"][:] = g["du"]
    o.rk4_tlm(meta["nsteps"])
    np.testing.assert_array_equal(o["h_tl1"], g["h_tl"])
    np.testing.assert_array_equal(o["u_tl1"], g["u_tl"])
    np.testing.assert_array_equal(o["uReconstructZonal_tl"], g["zonal_tl"])
    np.testing.assert_array_equal(o["uReconstructMeridional_tl"], g["meridional_tl"])
    # linear in u_tl and independent of u: the forward operator on u_tl
    o.reconstruct(g["u_tl"])
    np.testing.assert_array_equal(o["uReconstructZonal"], g["zonal_tl"])
    np.testing.assert_array_equal(o["uReconstructMeridional"], g["meridional_tl"])


def test_dissipation_and_forcing_bitwise():
    """del2 + del4 viscosity, bottom drag and wind stress (all on): forward,
    TLM and adjoint windows and the adjoint of one tendency evaluation."""
    from oracle.oracle import Oracle
    g, meta = golden("tc5_l3_dissipation")
    n, dt = meta["nsteps"], meta["dt"]
    mesh, phys = case(3, 5)
    kw = dict(test_case=5, visc2=meta["visc2"], visc4=meta["visc4"],
              wind_stress=meta["wind_stress"], bottom_drag=meta["bottom_drag"], u_src=g["u_src"])

    def fresh():
        o = Oracle(mesh, oracle_phys(phys), dt, **kw)
        o["h1"][:] = g["h0"]
        o["u1"][:] = g["u0"]
        o.init_block()
        return o

    o = fresh()
    o.rk4(n)
    np.testing.assert_array_equal(o["h1"], g["h_fwd"])
    np.testing.assert_array_equal(o["u1"], g["u_fwd"])
    o = fresh()
    o["h_tl1"][:] = g["dh"]
    o["u_tl1"][:] = g["du"]
    o.rk4_tlm(n)
    np.testing.assert_array_equal(o["h_tl1"], g["h_tl"])
    np.testing.assert_array_equal(o["u_tl1"], g["u_tl"])
    np.testing.assert_array_equal(o["h1"], g["h"])
    o = fresh()
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = g["dh"]
    o["u_ad1"][:] = g["du"]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    np.testing.assert_array_equal(o["h_ad1"], g["h_ad"])
    np.testing.assert_array_equal(o["u_ad1"], g["u_ad"])
    o = fresh()
    ha, ua = o.tend_full_adj(g["h0"], g["u0"], g["gh"], g["gu"])
    np.testing.assert_array_equal(ha, g["tad_h_ad"])
    np.testing.assert_array_equal(ua, g["tad_u_ad"])


def test_reconstruction_adjoint_seeds_bitwise():
    """Seeds in uReconstruct{X,Y,Z,Zonal,Meridional}_ad are applied to the
    end-state adjoint by the first adjoint step and consumed."""
    g, meta = golden("tc5_l3_recon_adjoint")
    n, dt = meta["nsteps"], meta["dt"]
    assert np.abs(g["u_ad"]).max() > 0
    o = _oracle_from(3, 5, dt, g["h0"], g["u0"])
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = 0.0
    o["u_ad1"][:] = 0.0
    for k, name in enumerate(("X", "Y", "Z", "Zonal", "Meridional")):
        o[f"uReconstruct{name}_ad"][:] = g["seeds"][k]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    np.testing.assert_array_equal(o["h_ad1"], g["h_ad"])
    np.testing.assert_array_equal(o["u_ad1"], g["u_ad"])
    assert not np.any(o["uReconstructZonal_ad"])


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_level6_one_day_forward_digest():
    """TC5 level 6, dt 150 s, 576 steps: the oracle's state after one day has
    the reference's SHA-256 (the initial state and the seeded TLM
    perturbation regenerated here match the reference's inputs too)."""
    g, meta = golden("tc5_l6_1day_digest")
    mesh, phys = case(6, 5)
    assert _sha(phys["h"]) == meta["sha_h0"] and _sha(phys["u"]) == meta["sha_u0"]
    rng = np.random.default_rng(meta["perturbation_seed"])
    dh, du = rng.standard_normal(mesh.nCells), 0.1 * rng.standard_normal(mesh.nEdges)
    assert _sha(dh) == meta["sha_dh"] and _sha(du) == meta["sha_du"]
    o = _oracle_from(6, 5, meta["dt"], phys["h"], phys["u"])
    o.rk4(meta["nsteps"])
    assert _sha(o["h1"]) == meta["sha_h"]
    assert _sha(o["u1"]) == meta["sha_u"]


def test_tc1_frozen_wind_bitwise():
    """TC1 (cosine-bell advection, src/mpas_sw_testcases.f90:1-66): the
    reference keeps u, and in the TLM u_tl, at their initial values
    (src/mpas_sw_module.f90:596-598,615,724-726,749-752) while h is advected;
    its adjoint sends every stage's u adjoint to the step-start state
    (:907-910,930-933)."""
    g, meta = golden("tc1_l3_frozen_wind")
    n, dt = meta["nsteps"], meta["dt"]
    mesh, phys = case(3, 1)
    o = make_oracle(3, 1, dt)
    o.rk4(n)
    np.testing.assert_array_equal(o["h1"], g["h"])
    np.testing.assert_array_equal(o["u1"], g["u"])
    np.testing.assert_array_equal(g["u"], phys["u"])
    o = make_oracle(3, 1, dt)
    o["h_tl1"][:] = g["dh"]
    o["u_tl1"][:] = g["du"]
    o.rk4_tlm(n)
    for k, ref in (("h1", "h_base"), ("u1", "u_base"), ("h_tl1", "h_tl"), ("u_tl1", "u_tl")):
        np.testing.assert_array_equal(o[k], g[ref], err_msg=k)
    np.testing.assert_array_equal(g["u_tl"], g["du"])
    o = make_oracle(3, 1, dt)
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = g["dh"]
    o["u_ad1"][:] = g["du"]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    np.testing.assert_array_equal(o["h_ad1"], g["h_ad"])
    np.testing.assert_array_equal(o["u_ad1"], g["u_ad"])


def test_thickness_advection_orders_are_order_two():
    """The reference's orders 3 and 4 (with and without config_monotonic)
    hash to the same TC5 forward, TLM and adjoint as order 2, because
    deriv_two stays zero (src/mpas_sw_module.f90:285-286); the oracle's
    order-2 run gives those digests."""
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    g, meta = golden("tc5_l3_adv_order_digest")
    n, dt = meta["nsteps"], meta["dt"]
    dig = meta["digests"]
    assert set(dig) == {"order2", "order3", "order3_monotonic", "order4"}
    assert all(v == dig["order2"] for v in dig.values())
    o = make_oracle(3, 5, dt)
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    fwd = (sha(o["h1"]), sha(o["u1"]))
    o = make_oracle(3, 5, dt)
    o["h_tl1"][:] = g["dh"]
    o["u_tl1"][:] = g["du"]
    o.rk4_tlm(n)
    tlm = (sha(o["h_tl1"]), sha(o["u_tl1"]))
    o["h_ad1"][:] = g["dh"]
    o["u_ad1"][:] = g["du"]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    adj = (sha(o["h_ad1"]), sha(o["u_ad1"]))
    assert list(fwd + tlm + adj) == dig["order2"]


def test_diagnostic_adjoint_seeds_bitwise():
    """Seeds put in the 13 diagnostic adjoint arrays are consumed by the
    first sw_rk4_adj at its end state (src/mpas_sw_module.f90:894-905) and
    zeroed; three adjoint steps with those seeds and end-state adjoints give
    the reference's h_ad, u_ad bit for bit."""
    g, meta = golden("tc5_l3_diag_seeds")
    n, dt = meta["nsteps"], meta["dt"]
    o = make_oracle(3, 5, dt)
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = g["dh"]
    o["u_ad1"][:] = g["du"]
    names = [k[5:] for k in g if k.startswith("seed_")]
    assert len(names) == 13
    for k in names:
        o[f"{k}_ad"][:] = g[f"seed_{k}"]
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
        if s == n - 1:
            for k in names:
                assert not o[f"{k}_ad"].any(), k
    np.testing.assert_array_equal(o["h_ad1"], g["h_ad"])
    np.testing.assert_array_equal(o["u_ad1"], g["u_ad"])
