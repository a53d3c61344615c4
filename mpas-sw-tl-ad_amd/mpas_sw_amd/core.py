"""Torch-facing host API over the HIP library.

``SWModel`` owns one library context (one mesh, one batch size, one member
precision) and steps caller-owned torch tensors on the GPU:

* states are in the library's internal (space-filling-curve) entity order,
  shape ``[n]`` for a single member or ``[n, B]`` with members innermost;
* nonlinear base states of the TLM/adjoint are float64 ``[n]``;
* member arrays are float64 or float32 (``precision``).

Host helpers convert reference-order numpy arrays to internal-order device
tensors and back.  Nothing here computes model physics on the CPU: every
step, diagnostic and reduction is a library call.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import mesh_desc  # noqa: F401  (re-exported)
from .mesh import Mesh, sfc_permutation

_DT = {"fp64": (torch.float64, _lib.MSW_FP64), "fp32": (torch.float32, _lib.MSW_FP32)}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class SWModel:
    """One mesh + configuration on the GPU.

    Parameters
    ----------
    mesh : Mesh in reference (generator/file) order, unit sphere.
    phys : dict with the test-case-scaled fields: dcEdge, dvEdge, areaCell,
        areaTriangle, kiteAreasOnVertex (3,V), fVertex, h_s, and optionally
        lonCell and coeffs_reconstruct (3, maxEdges, C).
    dt, apvm, test_case, visc2, visc4, ... : the reference's config_* values.
    batch : members per entity (power of two, 1..256).
    precision : "fp64" or "fp32" for member arrays.
    sfc : reorder entities along a 3-D Hilbert curve (default) or keep the
        reference order.
    perm : explicit (cell, edge, vertex) permutations, new -> old (overrides sfc).
    u_src : wind-stress source on edges (reference order), read when
        wind_stress is set; zero by default as after var_allocation.
    tl_fused_flux : form the TL tendency from the TL thickness flux (fewer
        gathers per member; rounding differs from the reference's term order).
    """

    def __init__(self, mesh: Mesh, phys: dict, dt: float, *, apvm: float = 0.5, test_case: int = 2,
                 visc2: float = 0.0, visc4: float = 0.0, thickness_adv_order: int = 2,
                 monotonic: bool = False, wind_stress: bool = False, bottom_drag: bool = False,
                 batch: int = 1, precision: str = "fp64", sfc: bool = True, device=None,
                 graphs: bool = False, perm=None, u_src=None, tl_fused_flux: bool = False):
        if precision not in _DT:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.lib = _lib.load()
        self.mesh = mesh
        self.batch = int(batch)
        self.precision = precision
        self.dtype, self._prec = _DT[precision]
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nC, self.nE, self.nV = mesh.nCells, mesh.nEdges, mesh.nVertices
        if perm is not None:
            pc, pe, pv = perm
        elif sfc:
            pc, pe, pv = sfc_permutation(mesh)
        else:
            pc, pe, pv = (np.arange(self.nC), np.arange(self.nE), np.arange(self.nV))
        self.perm = {"c": np.ascontiguousarray(pc, np.int32), "e": np.ascontiguousarray(pe, np.int32),
                     "v": np.ascontiguousarray(pv, np.int32)}
        self.cfg = _lib.Config(dt, apvm, visc2, visc4, test_case, thickness_adv_order,
                               int(monotonic), int(wind_stress), int(bottom_drag))
        d, keep = mesh_desc(mesh, phys, self.perm)   # keep: host arrays d points to
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            h = C.c_void_p()
            _lib.check(self.lib.msw_create(C.byref(d), C.byref(self.cfg), self._prec, self.batch,
                                           C.c_void_p(self.stream.cuda_stream), C.byref(h)))
        self._h = h
        if graphs:
            self.set_graphs(True)
        self._inv = {k: np.argsort(v).astype(np.int64) for k, v in self.perm.items()}
        if u_src is not None:
            self.set_forcing(u_src)
        if tl_fused_flux:
            self.set_option(_lib.OPT_TL_FUSED_FLUX, 1)
        self.stripe_log2 = 2
        self.block = 256

    # ------------------------------------------------------------------ utils
    def close(self):
        if getattr(self, "_h", None):
            self.lib.msw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, on: bool):
        _lib.check(self.lib.msw_set_graphs(self._h, int(bool(on))))

    def set_config(self, dt=None, apvm=None, visc2=None, visc4=None, wind_stress=None,
                   bottom_drag=None):
        if dt is not None:
            self.cfg.dt = dt
        if apvm is not None:
            self.cfg.apvm_upwinding = apvm
        if visc2 is not None:
            self.cfg.visc2 = visc2
        if visc4 is not None:
            self.cfg.visc4 = visc4
        if wind_stress is not None:
            self.cfg.wind_stress = int(bool(wind_stress))
        if bottom_drag is not None:
            self.cfg.bottom_drag = int(bool(bottom_drag))
        _lib.check(self.lib.msw_set_config(self._h, C.byref(self.cfg)))

    def set_option(self, option: int, value: int):
        _lib.check(self.lib.msw_set_option(self._h, int(option), int(value)))
        if option == _lib.OPT_XCD_STRIPE:
            self.stripe_log2 = int(value)
        elif option == _lib.OPT_BLOCK:
            self.block = int(value)

    def autotune(self, h, u, nsteps: int = 2, candidates=(0, 1, 2, 3, 4), blocks=(128, 256, 512),
                 seed: int = 0):
        """Choose the launch shape by timing a TLM sweep plus an adjoint sweep
        of `nsteps` from the FP64 base state (h, u) with random member
        perturbations, each candidate once after a warm-up: first the stripe
        (MSW_OPT_XCD_STRIPE: 2^k consecutive blocks per XCD turn) at the
        current block size, then the block size (MSW_OPT_BLOCK) at that
        stripe.  Sets the fastest pair and returns {(k, block): seconds}.
        (h, u) are not modified.  With 64 or more members a wavefront holds one
        entity's members, so the members per wavefront are fixed by the batch
        and the block size sets the entities per block."""
        g = torch.Generator(device="cpu").manual_seed(seed)
        shp_c, shp_e = self._shape("c"), self._shape("e")
        ht0 = torch.randn(shp_c, generator=g, dtype=torch.float64).to(self.device, self.dtype)
        ut0 = (0.1 * torch.randn(shp_e, generator=g, dtype=torch.float64)).to(self.device, self.dtype)
        th = torch.empty((nsteps, self.nC), dtype=torch.float64, device=self.device)
        tu = torch.empty((nsteps, self.nE), dtype=torch.float64, device=self.device)
        bufs = [torch.empty_like(h), torch.empty_like(u), torch.empty_like(ht0), torch.empty_like(ut0)]
        times = {}

        def run(k, blk):
            self.set_option(_lib.OPT_XCD_STRIPE, int(k))
            self.set_option(_lib.OPT_BLOCK, int(blk))
            for rep in range(2):   # the first run captures the step graphs
                for b, src in zip(bufs, (h, u, ht0, ut0)):
                    b.copy_(src)
                self.stream.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                self.tlm(bufs[0], bufs[1], bufs[2], bufs[3], nsteps, th, tu)
                self.adj(th, tu, bufs[2], bufs[3], nsteps)
                e1.record(self.stream)
                e1.synchronize()
            times[(int(k), int(blk))] = e0.elapsed_time(e1) / 1e3

        blk0 = self.block
        for k in candidates:
            run(k, blk0)
        k_best = min((kb for kb in times if kb[1] == blk0), key=times.get)[0]
        for blk in blocks:
            if (k_best, blk) not in times:
                run(k_best, blk)
        best = min(times, key=times.get)
        self.set_option(_lib.OPT_XCD_STRIPE, best[0])
        self.set_option(_lib.OPT_BLOCK, best[1])
        return times

    def set_instances(self, offsets):
        """The mesh is a disjoint union of instances (``mesh.union_mesh``):
        offsets[k] = (first cell, first edge, first vertex) of instance k,
        which the internal order keeps (``mesh.union_sfc_permutation``).  The
        persistent launch (``set_persistent``) cuts its work items there."""
        o = np.asarray(offsets, dtype=np.int32).reshape(-1, 3)
        cs, es, vs = (np.ascontiguousarray(o[:, i]) for i in range(3))
        _lib.check(self.lib.msw_set_instances(self._h, len(o), cs.ctypes.data, es.ctypes.data, vs.ctypes.data))

    def set_tend_output(self, tend_h=None, tend_u=None, tend_h_tl=None, tend_u_tl=None):
        """Keep the last RK stage's tendencies of every step (the module arrays
        tend_h, tend_u, tend_h_tl, tend_u_tl of the reference), internal order:
        fwd() writes the members' into tend_h/tend_u ([C, B], [E, B] of the
        member dtype); base(), tlm(), adj_step() and adj() write the FP64 base
        state's ([C], [E]) there; tlm() writes the members' TL tendencies into
        tend_h_tl/tend_u_tl.  None leaves an array unwritten; all None stops."""
        arrs = (tend_h, tend_u, tend_h_tl, tend_u_tl)
        for a in arrs:
            if a is not None and (not a.is_cuda or not a.is_contiguous()):
                raise ValueError("tendency outputs must be contiguous device tensors")
        self._tend_out = arrs   # the library keeps the pointers
        _lib.check(self.lib.msw_set_tend_output(self._h, *[None if a is None else a.data_ptr() for a in arrs]))

    def set_persistent(self, on: bool = True, barrier_polls: int | None = None):
        """Run fwd/tlm of a batch-1 model as one cooperative launch per call
        (MSW_OPT_PERSISTENT); results equal the per-stage launches bit for bit."""
        _lib.check(self.lib.msw_set_option(self._h, _lib.OPT_PERSISTENT, int(bool(on))))
        if barrier_polls is not None:
            _lib.check(self.lib.msw_set_option(self._h, _lib.OPT_BARRIER_POLLS, int(barrier_polls)))

    def set_mesh_scaling(self, del2=None, del4=None):
        """Per-edge factors of the del2 and del4 terms (meshScalingDel2/4 of
        compute_mesh_scaling, [E] reference order); None for both: all 1."""
        if (del2 is None) != (del4 is None):
            raise ValueError("del2 and del4 go together")
        if del2 is None:
            _lib.check(self.lib.msw_set_mesh_scaling(self._h, None, None))
            return
        d2 = self.to_internal(np.asarray(del2, np.float64).reshape(-1), "e", member=False)
        d4 = self.to_internal(np.asarray(del4, np.float64).reshape(-1), "e", member=False)
        _lib.check(self.lib.msw_set_mesh_scaling(self._h, _ptr(d2), _ptr(d4)))
        self.stream.synchronize()   # the copies read d2, d4 before they go

    def set_forcing(self, u_src):
        """Wind-stress source u_src [E] (reference order); None resets it to 0."""
        if u_src is None:
            _lib.check(self.lib.msw_set_forcing(self._h, None))
            return
        t = self.to_internal(np.asarray(u_src, np.float64).reshape(-1), "e", member=False)
        _lib.check(self.lib.msw_set_forcing(self._h, C.c_void_p(t.data_ptr())))
        self.stream.synchronize()

    @property
    def device_bytes(self) -> int:
        return int(self.lib.msw_device_bytes(self._h))

    def n(self, kind):
        return {"c": self.nC, "e": self.nE, "v": self.nV}[kind]

    def _shape(self, kind, batch=None):
        b = self.batch if batch is None else batch
        n = self.n(kind)
        return (n,) if b == 1 else (n, b)

    def empty(self, kind, member=True, batch=None):
        """Uninitialised internal-order device tensor for entity kind c/e/v."""
        dt = self.dtype if member else torch.float64
        b = (self.batch if batch is None else batch) if member else 1
        return torch.empty(self._shape(kind, b), dtype=dt, device=self.device)

    def zeros(self, kind, member=True, batch=None):
        t = self.empty(kind, member, batch)
        t.zero_()
        return t

    def to_internal(self, arr, kind, member=True):
        """Reference-order host array ([n] or [n,B]) -> internal-order device tensor."""
        a = np.asarray(arr)[self.perm[kind]]
        dt = self.dtype if member else torch.float64
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)

    def to_reference(self, t, kind):
        """Internal-order device tensor -> reference-order host float64 array."""
        a = t.detach().to("cpu", torch.float64).numpy()
        return a[self._inv[kind]]

    def _check(self, t, kind, member=True, name="array"):
        dt = self.dtype if member else torch.float64
        shp = self._shape(kind, self.batch if member else 1)
        if t.dtype != dt or tuple(t.shape) != shp or not t.is_contiguous() or t.device != self.device:
            raise ValueError(f"{name}: expected contiguous {dt} {shp} on {self.device}, got "
                             f"{t.dtype} {tuple(t.shape)} on {t.device}")

    # ----------------------------------------------------------------- steps
    def fwd(self, h, u, nsteps=1):
        """sw_rk4 x nsteps on the batch (h [C,B], u [E,B] of `precision`), in place."""
        self._check(h, "c", name="h"); self._check(u, "e", name="u")
        _lib.check(self.lib.msw_fwd_steps(self._h, _ptr(h), _ptr(u), int(nsteps)))

    def base(self, h, u, nsteps=1, traj_h=None, traj_u=None):
        """FP64 single-member forward, in place; optional trajectory storage
        (traj_h [nsteps, C], traj_u [nsteps, E] get the step-start states)."""
        self._check(h, "c", False, "h"); self._check(u, "e", False, "u")
        self._check_traj(traj_h, traj_u, nsteps)
        _lib.check(self.lib.msw_base_steps(self._h, _ptr(h), _ptr(u), int(nsteps), _ptr(traj_h), _ptr(traj_u)))

    def tlm(self, h, u, h_tl, u_tl, nsteps=1, traj_h=None, traj_u=None):
        """sw_rk4_tlm x nsteps: advances the FP64 base (h,u) and the batched
        TL state (h_tl,u_tl) in place."""
        self._check(h, "c", False, "h"); self._check(u, "e", False, "u")
        self._check(h_tl, "c", name="h_tl"); self._check(u_tl, "e", name="u_tl")
        self._check_traj(traj_h, traj_u, nsteps)
        _lib.check(self.lib.msw_tlm_steps(self._h, _ptr(h), _ptr(u), _ptr(h_tl), _ptr(u_tl), int(nsteps),
                                          _ptr(traj_h), _ptr(traj_u)))

    def adj_step(self, h, u, h_ad, u_ad, want_end=False):
        """sw_rk4_adj: (h,u) base at the step start; (h_ad,u_ad) updated in
        place from the step-end adjoint to the step-start adjoint.  Returns
        the nonlinear end state if ``want_end``."""
        self._check(h, "c", False, "h"); self._check(u, "e", False, "u")
        self._check(h_ad, "c", name="h_ad"); self._check(u_ad, "e", name="u_ad")
        he = ue = None
        if want_end:
            he = torch.empty_like(h)
            ue = torch.empty_like(u)
        _lib.check(self.lib.msw_adj_step(self._h, _ptr(h), _ptr(u), _ptr(h_ad), _ptr(u_ad), _ptr(he), _ptr(ue)))
        return (he, ue) if want_end else None

    def adj(self, traj_h, traj_u, h_ad, u_ad, nsteps):
        """nsteps adjoint steps backwards over a stored trajectory, in place."""
        self._check_traj(traj_h, traj_u, nsteps, required=True)
        self._check(h_ad, "c", name="h_ad"); self._check(u_ad, "e", name="u_ad")
        _lib.check(self.lib.msw_adj_steps(self._h, _ptr(traj_h), _ptr(traj_u), _ptr(h_ad), _ptr(u_ad), int(nsteps)))

    def _check_traj(self, th, tu, nsteps, required=False):
        if th is None and tu is None and not required:
            return
        if th is None or tu is None:
            raise ValueError("traj_h and traj_u go together")
        for t, n in ((th, self.nC), (tu, self.nE)):
            if t.dtype != torch.float64 or not t.is_contiguous() or t.device != self.device or t.numel() < nsteps * n:
                raise ValueError("trajectory buffers must be contiguous float64 [nsteps, n] on the device")

    # ------------------------------------------------------------ diagnostics
    def diagnostics(self, h, u):
        """All 13 diagnostic fields of one FP64 state, internal order."""
        self._check(h, "c", False, "h"); self._check(u, "e", False, "u")
        names = self.lib.msw_diag_names().decode().split(",")
        kinds = self.lib.msw_diag_kinds().decode().split(",")
        sizes = [self.n(k) for k in kinds]
        out = torch.empty(sum(sizes), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.msw_diagnostics(self._h, _ptr(h), _ptr(u), _ptr(out)))
        res, off = {}, 0
        for nm, k, s in zip(names, kinds, sizes):
            res[nm] = (out[off:off + s], k)
            off += s
        return res

    def diagnostics_adj(self, h, u, seeds, h_ad, u_ad):
        """h_ad += and u_ad += the adjoint of the diagnostics of the FP64 state
        (h, u) applied to `seeds` (msw_diagnostics_adj): a dict from the names
        of msw_diag_names to member tensors in internal order; missing names
        are zero.  h_ad, u_ad: member tensors, updated in place."""
        self._check(h, "c", False, "h"); self._check(u, "e", False, "u")
        self._check(h_ad, "c", name="h_ad"); self._check(u_ad, "e", name="u_ad")
        names = self.lib.msw_diag_names().decode().split(",")
        kinds = self.lib.msw_diag_kinds().decode().split(",")
        unknown = set(seeds) - set(names)
        if unknown:
            raise ValueError(f"unknown diagnostic seeds: {sorted(unknown)}")
        parts = []
        for nm, k in zip(names, kinds):
            t = seeds.get(nm)
            if t is None:
                parts.append(torch.zeros(self._shape(k), dtype=self.dtype, device=self.device).reshape(-1))
            else:
                self._check(t, k, name=nm)
                parts.append(t.reshape(-1))
        packed = torch.cat(parts)
        _lib.check(self.lib.msw_diagnostics_adj(self._h, _ptr(h), _ptr(u), _ptr(packed), _ptr(h_ad), _ptr(u_ad)))

    def reconstruct(self, u):
        """[5, C(, B)] = X, Y, Z, zonal, meridional of the batch `u`."""
        self._check(u, "e", name="u")
        out = torch.empty((5,) + self._shape("c"), dtype=self.dtype, device=self.device)
        _lib.check(self.lib.msw_reconstruct(self._h, _ptr(u), _ptr(out)))
        return out

    def invariants(self, h, u):
        """FP64 [mass, total energy] of one FP64 state (see msw_invariants)."""
        self._check(h, "c", member=False, name="h")
        self._check(u, "e", member=False, name="u")
        out = torch.empty(2, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.msw_invariants(self._h, _ptr(h), _ptr(u), _ptr(out)))
        return out

    def reconstruct_adj(self, seeds, u_ad):
        """u_ad += R^T seeds in place; seeds [5, C(, B)] = adjoints of X, Y, Z,
        zonal, meridional (the order `reconstruct` returns)."""
        self._check(u_ad, "e", name="u_ad")
        seeds = seeds.to(self.dtype).contiguous()
        if tuple(seeds.shape) != (5,) + self._shape("c"):
            raise ValueError(f"seeds must have shape {(5,) + self._shape('c')}")
        _lib.check(self.lib.msw_reconstruct_adj(self._h, _ptr(seeds), _ptr(u_ad)))
        return u_ad

    def dot(self, xh=None, xu=None, yh=None, yu=None, wh=None, wu=None):
        """Per-member FP64 inner products (see msw_dot); returns float64 [B]."""
        out = torch.empty(self.batch, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.msw_dot(self._h, _ptr(xh), _ptr(xu), _ptr(yh), _ptr(yu), _ptr(wh), _ptr(wu), _ptr(out)))
        return out

    def nonfinite(self, h, u):
        flags = torch.zeros(self.batch, dtype=torch.int32, device=self.device)
        _lib.check(self.lib.msw_nonfinite(self._h, _ptr(h), _ptr(u), _ptr(flags)))
        return flags

    def permute(self, src, kind, to_internal, batch=1):
        dst = torch.empty_like(src)
        prec = _lib.MSW_FP64 if src.dtype == torch.float64 else _lib.MSW_FP32
        _lib.check(self.lib.msw_permute(self._h, ord(kind), int(to_internal), prec, int(batch), _ptr(src), _ptr(dst)))
        return dst
