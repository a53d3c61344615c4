"""Binomial checkpoint schedule (CPU) and the differentiable RK4 window (GPU)."""

import numpy as np
import pytest
import torch

from helpers import case, rel_l2
from mpas_sw_amd.checkpoint import forward_cost, schedule


def simulate(n, s):
    """Play a schedule with integer 'states'; check every adjoint step sees
    its own step-start state, count forward steps and live snapshots."""
    snaps = {0}
    cur = None
    fwd = 0
    done = []
    peak = 1
    for act in schedule(n, s):
        if act[0] == "advance":
            _, a, b = act
            assert a in snaps
            fwd += b - a
            cur = b
        elif act[0] == "store":
            assert cur == act[1]
            snaps.add(act[1])
            peak = max(peak, len(snaps))
        elif act[0] == "drop":
            snaps.discard(act[1])
        else:
            k = act[1]
            assert k in snaps or cur == k
            done.append(k)
    assert done == list(range(n - 1, -1, -1))
    return fwd, peak


@pytest.mark.parametrize("n,s", [(1, 1), (7, 1), (10, 2), (50, 3), (100, 4), (2304, 32), (400, 6)])
def test_schedule_is_valid_and_optimal(n, s):
    fwd, peak = simulate(n, s)
    assert peak <= s
    assert fwd == forward_cost(n, s)


def test_cost_closed_form_small_cases():
    # one snapshot: quadratic; enough snapshots: every step once
    assert forward_cost(10, 1) == 45
    assert forward_cost(10, 9) == 9
    # beta(2,2) = 6 steps with 2 snapshots cost 2*6 - beta(3,1) = 8
    assert forward_cost(6, 2) == 8


@pytest.mark.gpu
def test_autograd_window_gradient():
    print("autograd window L3 tc5", flush=True)
    from mpas_sw_amd.checkpoint import rk4_window
    from mpas_sw_amd.core import SWModel
    tc, dt, n = 5, 900.0, 10
    mesh, phys = case(3, tc)
    m = SWModel(mesh, phys, dt, test_case=tc)
    A = m.to_internal(phys["areaCell"], "c", member=False)
    rng = np.random.default_rng(0)
    hobs = m.to_internal(phys["h"] + rng.standard_normal(mesh.nCells), "c", member=False)

    def loss(h0, u0, snaps):
        hT, uT = rk4_window(m, h0, u0, n, snaps)
        return 0.5 * (A * (hT - hobs) ** 2).sum()

    grads = {}
    for snaps in (n, 3, 1):
        h0 = m.to_internal(phys["h"], "c", member=False).requires_grad_(True)
        u0 = m.to_internal(phys["u"], "e", member=False).requires_grad_(True)
        J = loss(h0, u0, snaps)
        J.backward()
        grads[snaps] = (h0.grad.clone(), u0.grad.clone())
    # recomputation reproduces the stored trajectory exactly
    for s in (3, 1):
        assert torch.equal(grads[s][0], grads[n][0]) and torch.equal(grads[s][1], grads[n][1])
    # finite differences along random directions
    gh, gu = grads[n]
    h0 = m.to_internal(phys["h"], "c", member=False)
    u0 = m.to_internal(phys["u"], "e", member=False)
    for k in range(3):
        dh = m.to_internal(rng.standard_normal(mesh.nCells), "c", member=False)
        du = m.to_internal(0.1 * rng.standard_normal(mesh.nEdges), "e", member=False)
        eps = 1e-3
        with torch.no_grad():
            fd = (loss(h0 + eps * dh, u0 + eps * du, n) - loss(h0 - eps * dh, u0 - eps * du, n)) / (2 * eps)
        ad = (gh * dh).sum() + (gu * du).sum()
        assert abs(float(fd) - float(ad)) < 1e-6 * abs(float(ad))
