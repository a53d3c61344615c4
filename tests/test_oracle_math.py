"""Pins of the oracle that do not rely on the reference run: conservation
laws, discrete identities and the tangent-linear / adjoint relations.  Each
is chosen so that a plausible slip (a dropped term, a flipped sign, a
transposed operand) fails it."""

import numpy as np

from helpers import case, make_oracle, oracle_phys, perturbations, rel_l2

GRAV = 9.80616


def test_mass_conservation():
    # sum_c A_c tend_h(c) = 0: every edge flux leaves one cell and enters the other
    mesh, phys = case(3, 5)
    o = make_oracle(3, 5, 900.0)
    o.tend(phys["h"], phys["u"])
    s = np.sum(phys["areaCell"] * o["tend_h"])
    scale = np.sum(np.abs(phys["areaCell"] * o["tend_h"]))
    assert abs(s) < 1e-13 * scale


def test_streamfunction_flow_is_divergence_free():
    # u = -(psi(v2)-psi(v1))/dv has zero discrete divergence in every cell
    mesh, phys = case(3, 2)
    o = make_oracle(3, 2, 600.0)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(mesh.nVertices)
    u = -(psi[mesh.verticesOnEdge[1] - 1] - psi[mesh.verticesOnEdge[0] - 1]) / phys["dvEdge"]
    o.diag(phys["h"], u)
    div = o["divergence"] * phys["areaCell"]
    ref = np.abs(u * phys["dvEdge"]).max()
    assert np.abs(div).max() < 1e-12 * ref


def test_nonlinear_coriolis_conserves_energy():
    # With apvm = 0 the TRiSK q.u_perp term does no work:
    # sum_e dc dv h_edge u q_e = 0 (Ringler et al. 2010)
    mesh, phys = case(3, 5)
    o = make_oracle(3, 5, 900.0, apvm=0.0)
    h, u = phys["h"], phys["u"]
    o.diag(h, u)
    o.tend(h, u)
    c1 = mesh.cellsOnEdge[0] - 1
    c2 = mesh.cellsOnEdge[1] - 1
    ke, hs = o["ke"], phys["h_s"]
    q = o["tend_u"] + (ke[c2] - ke[c1] + GRAV * (h[c2] + hs[c2] - h[c1] - hs[c1])) / phys["dcEdge"]
    w = phys["dcEdge"] * phys["dvEdge"] * o["h_edge"] * u * q
    assert abs(w.sum()) < 1e-12 * np.abs(w).sum()


def _run_fwd(level, tc, dt, h, u, n):
    from oracle.oracle import Oracle
    mesh, phys = case(level, tc)
    o = Oracle(mesh, oracle_phys(phys), dt, test_case=tc)
    o["h1"][:] = h
    o["u1"][:] = u
    o.init_block()
    o.rk4(n)
    return o["h1"].copy(), o["u1"].copy()


def test_tlm_linearisation_ratio():
    # ||M(x + e dx) - M(x)|| / ||e L dx|| -> 1 with an O(e) error
    level, tc, dt, n = 3, 5, 900.0, 12
    mesh, phys = case(level, tc)
    dh, du = perturbations(mesh, 1, sigma_h=10.0, sigma_u=1.0)
    dh, du = dh[:, 0], du[:, 0]
    h0, u0 = _run_fwd(level, tc, dt, phys["h"], phys["u"], n)
    o = make_oracle(level, tc, dt)
    o["h_tl1"][:] = dh
    o["u_tl1"][:] = du
    o.rk4_tlm(n)
    lh, lu = o["h_tl1"].copy(), o["u_tl1"].copy()
    errs = []
    for eps in (1e-2, 1e-3, 1e-4):
        h1, u1 = _run_fwd(level, tc, dt, phys["h"] + eps * dh, phys["u"] + eps * du, n)
        num = np.sqrt(np.sum((h1 - h0) ** 2) + np.sum((u1 - u0) ** 2))
        den = eps * np.sqrt(np.sum(lh ** 2) + np.sum(lu ** 2))
        errs.append(abs(num / den - 1.0))
    assert errs[2] < 1e-4
    assert errs[0] / errs[1] > 5 and errs[1] / errs[2] > 5   # first order in eps


def test_dot_product_identity_multistep():
    # <L x, y> = <x, L^T y> over 6 steps of sw_rk4_tlm / sw_rk4_adj
    level, tc, dt, n = 3, 6, 600.0, 6
    mesh, phys = case(level, tc)
    xh, xu = [a[:, 0] for a in perturbations(mesh, 1, seed=11)]
    yh, yu = [a[:, 0] for a in perturbations(mesh, 1, seed=12)]
    o = make_oracle(level, tc, dt)
    traj = []
    o["h_tl1"][:] = xh
    o["u_tl1"][:] = xu
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4_tlm()
    lhs = np.dot(o["h_tl1"], yh) + np.dot(o["u_tl1"], yu)
    o2 = make_oracle(level, tc, dt)
    o2["h_ad1"][:] = yh
    o2["u_ad1"][:] = yu
    for s in range(n - 1, -1, -1):
        o2["h1"][:] = traj[s][0]
        o2["u1"][:] = traj[s][1]
        o2.rk4_adj()
    rhs = np.dot(xh, o2["h_ad1"]) + np.dot(xu, o2["u_ad1"])
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


def test_adjoint_gradient_matches_finite_differences():
    # J = 1/2 sum A h(T)^2 ; dJ/dx0 from the adjoint vs central differences
    level, tc, dt, n = 3, 5, 900.0, 4
    mesh, phys = case(level, tc)
    A = phys["areaCell"]

    def J(h, u):
        hT, _ = _run_fwd(level, tc, dt, h, u, n)
        return 0.5 * np.sum(A * hT ** 2)

    o = make_oracle(level, tc, dt)
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    o["h_ad1"][:] = A * o["h1"]
    o["u_ad1"][:] = 0.0
    for s in range(n - 1, -1, -1):
        o["h1"][:] = traj[s][0]
        o["u1"][:] = traj[s][1]
        o.rk4_adj()
    gh, gu = o["h_ad1"].copy(), o["u_ad1"].copy()
    rng = np.random.default_rng(5)
    for _ in range(3):
        dh = rng.standard_normal(mesh.nCells)
        du = 0.1 * rng.standard_normal(mesh.nEdges)
        eps = 1e-3
        fd = (J(phys["h"] + eps * dh, phys["u"] + eps * du) -
              J(phys["h"] - eps * dh, phys["u"] - eps * du)) / (2 * eps)
        ad = np.dot(gh, dh) + np.dot(gu, du)
        assert abs(fd - ad) < 1e-6 * abs(ad)


def test_rk4_stage_weights_on_a_uniform_tendency():
    # A state at rest on a flat surface with h = const has zero tendency;
    # adding a constant to h is a fixed point of sw_rk4 (closed form)
    mesh, phys = case(2, 6)
    from oracle.oracle import Oracle
    p = dict(oracle_phys(phys))
    p["fVertex"] = np.zeros(mesh.nVertices)
    p["h_s"] = np.zeros(mesh.nCells)
    o = Oracle(mesh, p, 600.0, test_case=6)
    o["h1"][:] = 1000.0
    o["u1"][:] = 0.0
    o.init_block()
    o.rk4(3)
    assert np.all(o["h1"] == 1000.0) and np.all(o["u1"] == 0.0)
    assert rel_l2(o["h1"], np.full(mesh.nCells, 1000.0)) == 0.0


def test_dense_jacobians_on_the_12_cell_mesh():
    """Level 0 (12 cells, 30 edges): the 3-step TLM matrix, built column by
    column from unit perturbations, equals central differences of the
    forward, and the adjoint applied to unit seeds gives its rows (L^T e_i
    is row i of L, so the adjoint outputs stacked as rows rebuild the same
    matrix) -- the brute-force form of the tangent-linear and adjoint
    relations, on a mesh small enough to form them whole."""
    level, tc, dt, n = 0, 5, 1800.0, 3
    mesh, phys = case(level, tc)
    C, E = mesh.nCells, mesh.nEdges
    N = C + E

    def unit(k):
        x = np.zeros(N)
        x[k] = 1.0
        return x[:C], x[C:]

    M = np.empty((N, N))
    for k in range(N):
        o = make_oracle(level, tc, dt)
        o["h_tl1"][:], o["u_tl1"][:] = unit(k)
        o.rk4_tlm(n)
        M[:, k] = np.concatenate([o["h_tl1"], o["u_tl1"]])
    o = make_oracle(level, tc, dt)
    traj = []
    for _ in range(n):
        traj.append((o["h1"].copy(), o["u1"].copy()))
        o.rk4()
    A = np.empty((N, N))
    for i in range(N):
        o = make_oracle(level, tc, dt)
        o["h_ad1"][:], o["u_ad1"][:] = unit(i)
        for s in range(n - 1, -1, -1):
            o["h1"][:] = traj[s][0]
            o["u1"][:] = traj[s][1]
            o.rk4_adj()
        A[i, :] = np.concatenate([o["h_ad1"], o["u_ad1"]])
    assert np.abs(A - M).max() < 1e-12 * np.abs(M).max()
    assert np.abs(A - M.T).max() > 1e-3 * np.abs(M).max()   # L is far from symmetric
    # central differences, step 1e-3 m in h and 1e-4 m/s in u
    x0 = np.concatenate([phys["h"], phys["u"]])
    steps = np.concatenate([np.full(C, 1e-3), np.full(E, 1e-4)])
    FD = np.empty((N, N))
    for k in range(N):
        d = np.zeros(N)
        d[k] = steps[k]
        hp, up = _run_fwd(level, tc, dt, (x0 + d)[:C], (x0 + d)[C:], n)
        hm, um = _run_fwd(level, tc, dt, (x0 - d)[:C], (x0 - d)[C:], n)
        FD[:, k] = np.concatenate([hp - hm, up - um]) / (2 * steps[k])
    assert np.abs(FD - M).max() < 1e-6 * np.abs(M).max()
    # and the matrix is not trivial: every state variable reaches others
    assert np.count_nonzero(np.abs(M - np.eye(N)) > 1e-12 * np.abs(M).max()) > N * N // 4
