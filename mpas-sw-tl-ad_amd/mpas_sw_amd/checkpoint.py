"""Binomial checkpointing (Griewank's revolve) of the nonlinear trajectory and
a ``torch.autograd.Function`` over a window of RK4 steps.

The forward pass runs the FP64 nonlinear model on the GPU and keeps at most
``snapshots`` step-start states; the backward pass walks the window in
reverse, recomputing states from the nearest snapshot on the optimal
binomial schedule, and applies the adjoint RK4 step (``sw_rk4_adj``,
``src/mpas_sw_module.f90:783-1002``) on the GPU at each step.  With
``snapshots >= nsteps`` every state is stored and nothing is recomputed.

Schedule: reversing ``l`` steps with ``s`` snapshots costs
``T(l, s) = t*l - beta(s+1, t-1)`` forward steps, ``beta(s, t) = C(s+t, s)``,
``t`` the smallest integer with ``beta(s, t) >= l`` (Griewank & Walther,
ACM TOMS 26, 2000).  Each recursion advances to the split point ``m`` that
minimises ``m + T(l-m, s-1) + T(m, s)``.
"""

from __future__ import annotations

from functools import lru_cache
from math import comb

import numpy as np
import torch


def beta(s: int, t: int) -> int:
    return comb(s + t, s)


@lru_cache(maxsize=None)
def forward_cost(l: int, s: int) -> int:
    """Minimal number of forward steps to reverse ``l`` steps with ``s``
    snapshots (the start state counts as held)."""
    if l <= 1:
        return 0
    if s <= 0:
        raise ValueError("need at least one snapshot")
    if s == 1:
        return l * (l - 1) // 2
    t = 0
    while beta(s, t) < l:
        t += 1
    return t * l - beta(s + 1, t - 1)


@lru_cache(maxsize=None)
def best_split(l: int, s: int) -> int:
    """Steps to advance before storing the next snapshot."""
    if s <= 1 or l <= 2:
        return l - 1 if s <= 1 else 1
    m = np.arange(1, l)
    cost = m + np.array([forward_cost(l - k, s - 1) for k in m]) + np.array([forward_cost(k, s) for k in m])
    return int(m[np.argmin(cost)])


def schedule(nsteps: int, snapshots: int):
    """The action list of a binomial reversal of ``nsteps`` steps:
    ("advance", from, to), ("store", step), ("adjoint", step), ("drop", step).
    Used by :class:`RK4Window` and by the tests."""
    acts = []

    def rev(start, end, s):
        # state at `start` is held in a snapshot
        if end - start == 1:
            acts.append(("adjoint", start))
            return
        if s <= 1:
            # no free slot: recompute from `start` for every step
            for k in range(end - 1, start - 1, -1):
                if k > start:
                    acts.append(("advance", start, k))
                acts.append(("adjoint", k))
            return
        m = start + best_split(end - start, s)
        acts.append(("advance", start, m))
        acts.append(("store", m))
        rev(m, end, s - 1)
        acts.append(("drop", m))
        rev(start, m, s)

    if nsteps > 0:
        rev(0, nsteps, max(1, snapshots))
    return acts


class RK4Window(torch.autograd.Function):
    """(h0, u0) -> (hT, uT) after ``nsteps`` sw_rk4 steps, differentiable
    through the GPU adjoint.  Inputs are float64 internal-order tensors of a
    single state on ``model``'s device."""

    @staticmethod
    def forward(ctx, h0, u0, model, nsteps, snapshots):
        nsteps = int(nsteps)
        snapshots = max(1, int(snapshots))
        h = h0.detach().clone()
        u = u0.detach().clone()
        if snapshots >= nsteps:
            th = torch.empty((max(nsteps, 1), model.nC), dtype=torch.float64, device=h.device)
            tu = torch.empty((max(nsteps, 1), model.nE), dtype=torch.float64, device=h.device)
            model.base(h, u, nsteps, th, tu)
            ctx.traj = (th, tu)
        else:
            ctx.traj = None
        if ctx.traj is None:
            model.base(h, u, nsteps)
        ctx.model = model
        ctx.nsteps = nsteps
        ctx.snapshots = snapshots
        ctx.save_for_backward(h0.detach(), u0.detach())
        return h, u

    @staticmethod
    def backward(ctx, gh, gu):
        m = ctx.model
        h0, u0 = ctx.saved_tensors
        n = ctx.nsteps
        lam_h = (gh if gh is not None else torch.zeros_like(h0)).detach().to(m.dtype).clone().contiguous()
        lam_u = (gu if gu is not None else torch.zeros_like(u0)).detach().to(m.dtype).clone().contiguous()
        if m.batch != 1:
            raise ValueError("RK4Window needs a model with batch = 1")
        if ctx.traj is not None:
            th, tu = ctx.traj
            if n > 0:
                m.adj(th, tu, lam_h, lam_u, n)
        else:
            snaps = {0: (h0.clone(), u0.clone())}
            for act in schedule(n, ctx.snapshots):
                if act[0] == "advance":
                    _, a, b = act
                    hs, us = snaps[a]
                    h, u = hs.clone(), us.clone()
                    m.base(h, u, b - a)
                    cur = (b, h, u)
                elif act[0] == "store":
                    b, h, u = cur
                    snaps[b] = (h, u)
                elif act[0] == "drop":
                    snaps.pop(act[1], None)
                else:
                    k = act[1]
                    if k in snaps:
                        hs, us = snaps[k]
                    else:
                        b, hs, us = cur
                        assert b == k
                    m.adj_step(hs, us, lam_h, lam_u)
        return lam_h.to(torch.float64), lam_u.to(torch.float64), None, None, None


def rk4_window(model, h0, u0, nsteps, snapshots=32):
    """Differentiable nonlinear window: returns (hT, uT)."""
    return RK4Window.apply(h0, u0, model, nsteps, snapshots)
