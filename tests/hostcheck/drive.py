"""Drive the host-check build of libmpas_sw_amd (the real msw_runtime.cpp
linked with mock_device.cpp, no GPU) through its C ABI, the way the GPU
tests and the bench do, and fail on anything the mock reports: a buffer
outside a live allocation, a race between the main and the side stream, a
capture that breaks HIP's rules, an error that does not reach the caller,
or graphs piling up.

    python drive.py <libmsw_hostcheck.so>

Run under AddressSanitizer (LD_PRELOAD of libasan); prints one line per
scenario and exits 1 on the first failure.
"""

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "mpas-sw-tl-ad_amd"))

from mpas_sw_amd import _lib  # noqa: E402
from mpas_sw_amd.mesh import get_mesh, sfc_permutation  # noqa: E402
from mpas_sw_amd.reconstruct import reconstruct_coefficients  # noqa: E402
from mpas_sw_amd.testcases import initial_conditions  # noqa: E402

lib = _lib.load(sys.argv[1])
vp = C.c_void_p
lib.mock_dev_alloc.restype = vp
lib.mock_dev_alloc.argtypes = [C.c_size_t]
lib.mock_dev_free.argtypes = [vp]
lib.mock_fail.argtypes = [C.c_int, C.c_long]
for f in ("mock_calls", "mock_violations", "mock_graphs_live", "mock_graphs_made", "mock_allocations"):
    getattr(lib, f).restype = C.c_long
lib.mock_calls.argtypes = [C.c_int]
lib.mock_message.restype = C.c_char_p
lib.mock_message.argtypes = [C.c_int]

COPY, LAUNCH, MALLOC = 0, 1, 2
FP64, FP32 = _lib.MSW_FP64, _lib.MSW_FP32

mesh = get_mesh(3)
phys = initial_conditions(mesh, 6)
phys["coeffs_reconstruct"] = reconstruct_coefficients(phys, mesh)
pc, pe, pv = sfc_permutation(mesh)
desc, keep = _lib.mesh_desc(mesh, phys, {"c": pc, "e": pe, "v": pv})
nC, nE, nV = mesh.nCells, mesh.nEdges, mesh.nVertices


def fail(msg):
    print("FAIL", msg, flush=True)
    for i in range(min(int(lib.mock_violations()), 40)):
        print("  ", lib.mock_message(i).decode(), flush=True)
    sys.exit(1)


def no_violations(where):
    if lib.mock_violations():
        fail(f"{where}: {lib.mock_violations()} violations")


def ok(rc, where):
    if rc != 0:
        fail(f"{where}: status {rc}: {lib.msw_last_error().decode()}")


class Dev:
    """Mock device buffers, freed together."""

    def __init__(self):
        self.ptrs = []

    def __call__(self, nbytes):
        p = lib.mock_dev_alloc(max(int(nbytes), 1))
        if not p:
            fail("mock allocation")
        self.ptrs.append(p)
        return vp(p)

    def free(self):
        for p in self.ptrs:
            lib.mock_dev_free(vp(p))
        self.ptrs = []


def config(**kw):
    c = _lib.Config(600.0, 0.5, 0.0, 0.0, 6, 2, 0, 0, 0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def create(prec, B, cfg=None):
    h = vp()
    ok(lib.msw_create(C.byref(desc), C.byref(cfg or config()), prec, B, None, C.byref(h)), "msw_create")
    return h


def state(dev, prec, B, n=3):
    es = 8 if prec == FP64 else 4
    return dict(h=dev(8 * nC), u=dev(8 * nE), ht=dev(es * nC * B), ut=dev(es * nE * B),
                th=dev(8 * n * nC), tu=dev(8 * n * nE))


def tlm_adj(ctx, s, n=3):
    ok(lib.msw_tlm_steps(ctx, s["h"], s["u"], s["ht"], s["ut"], n, s["th"], s["tu"]), "msw_tlm_steps")
    ok(lib.msw_adj_steps(ctx, s["th"], s["tu"], s["ht"], s["ut"], n), "msw_adj_steps")


def launch_shapes(prec, B):
    """The launch-mapping test's sequence and the autotuner's."""
    ctx = create(prec, B)
    ok(lib.msw_set_graphs(ctx, 1), "graphs")
    ok(lib.msw_set_option(ctx, _lib.OPT_TL_FUSED_FLUX, 1), "fused")
    for k, blk in ((0, 256), (2, 256), (5, 256), (2, 64), (1, 128), (3, 512), (7, 256), (7, 64)):
        dev = Dev()
        ok(lib.msw_set_option(ctx, _lib.OPT_XCD_STRIPE, k), "stripe")
        ok(lib.msw_set_option(ctx, _lib.OPT_BLOCK, blk), "block")
        s = state(dev, prec, B)
        tlm_adj(ctx, s)
        tlm_adj(ctx, s)   # replays the cached graphs
        dev.free()        # the caller frees its buffers between shapes
    dev = Dev()
    s = state(dev, prec, B, n=1)
    for k, blk in ((0, 256), (2, 256), (2, 128), (2, 512)):
        ok(lib.msw_set_option(ctx, _lib.OPT_XCD_STRIPE, k), "stripe")
        ok(lib.msw_set_option(ctx, _lib.OPT_BLOCK, blk), "block")
        for _ in range(2):
            tlm_adj(ctx, s, n=1)
    if lib.msw_set_option(ctx, _lib.OPT_BLOCK, 96) != _lib.MSW_EINVAL:   # the test's last call
        fail("a block of 96 threads was accepted")
    ok(lib.msw_destroy(ctx), "destroy")
    dev.free()
    no_violations(f"launch shapes prec={prec} B={B}")
    if lib.mock_graphs_live():
        fail(f"{lib.mock_graphs_live()} graphs left after msw_destroy")


def other_entry_points(prec, B, graphs):
    es = 8 if prec == FP64 else 4
    ctx = create(prec, B)
    ok(lib.msw_set_graphs(ctx, int(graphs)), "graphs")
    dev = Dev()
    s = state(dev, prec, B, n=4)
    ok(lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 2), "msw_fwd_steps")
    ok(lib.msw_base_steps(ctx, s["h"], s["u"], 2, s["th"], s["tu"]), "msw_base_steps traj")
    ok(lib.msw_base_steps(ctx, s["h"], s["u"], 2, None, None), "msw_base_steps")
    tlm_adj(ctx, s, n=4)
    he, ue = dev(8 * nC), dev(8 * nE)
    ok(lib.msw_adj_step(ctx, s["h"], s["u"], s["ht"], s["ut"], he, ue), "msw_adj_step end")
    ok(lib.msw_adj_step(ctx, s["h"], s["u"], s["ht"], s["ut"], None, None), "msw_adj_step")
    diag = dev(8 * (5 * nE + 4 * nV + 4 * nC))
    ok(lib.msw_diagnostics(ctx, s["h"], s["u"], diag), "msw_diagnostics")
    seeds = dev(es * B * (5 * nE + 4 * nV + 4 * nC))
    ok(lib.msw_diagnostics_adj(ctx, s["h"], s["u"], seeds, s["ht"], s["ut"]), "msw_diagnostics_adj")
    ok(lib.msw_diagnostics_adj(ctx, s["h"], s["u"], seeds, s["ht"], s["ut"]), "msw_diagnostics_adj again")
    inv = dev(16)
    for _ in range(2):
        ok(lib.msw_invariants(ctx, s["h"], s["u"], inv), "msw_invariants")
    r5 = dev(es * B * 5 * nC)
    ok(lib.msw_reconstruct(ctx, s["ut"], r5), "msw_reconstruct")
    ok(lib.msw_reconstruct_adj(ctx, r5, s["ut"]), "msw_reconstruct_adj")
    out = dev(8 * B)
    wh, wu = dev(8 * nC), dev(8 * nE)
    ok(lib.msw_dot(ctx, s["ht"], s["ut"], s["ht"], s["ut"], wh, wu, out), "msw_dot")
    ok(lib.msw_dot(ctx, s["ht"], None, None, None, None, None, out), "msw_dot h only")
    ok(lib.msw_dot_noalloc(nC, nE, B, prec, s["ht"], s["ut"], s["ht"], s["ut"], wh, wu, out, None),
       "msw_dot_noalloc")
    flags = dev(4 * B)
    ok(lib.msw_nonfinite(ctx, s["ht"], s["ut"], flags), "msw_nonfinite")
    tmp = dev(es * B * nE)
    ok(lib.msw_permute(ctx, ord("e"), 1, prec, B, s["ut"], tmp), "msw_permute in")
    ok(lib.msw_permute(ctx, ord("e"), 0, prec, B, tmp, s["ut"]), "msw_permute out")
    # dissipation, drag, wind, mesh scaling: scratch made on first use, graphs redone
    usrc, d2, d4 = dev(8 * nE), dev(8 * nE), dev(8 * nE)
    ok(lib.msw_set_config(ctx, C.byref(config(visc2=0.1, visc4=0.1, wind_stress=1, bottom_drag=1))), "set_config")
    ok(lib.msw_set_forcing(ctx, usrc), "set_forcing")
    ok(lib.msw_set_mesh_scaling(ctx, d2, d4), "set_mesh_scaling")
    ok(lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 2), "msw_fwd_steps dissipation")
    tlm_adj(ctx, s, n=4)
    ok(lib.msw_adj_step(ctx, s["h"], s["u"], s["ht"], s["ut"], he, ue), "msw_adj_step dissipation")
    ok(lib.msw_set_mesh_scaling(ctx, None, None), "mesh scaling off")
    ok(lib.msw_set_forcing(ctx, None), "forcing off")
    ok(lib.msw_set_config(ctx, C.byref(config(test_case=1))), "frozen wind")
    tlm_adj(ctx, s, n=4)
    ok(lib.msw_destroy(ctx), "destroy")
    dev.free()
    no_violations(f"entry points prec={prec} B={B} graphs={graphs}")


def tend_outputs(prec, B, graphs):
    """msw_set_tend_output with buffers of exactly the documented sizes: the
    members' tendencies for msw_fwd_steps, the FP64 base tendencies and the
    members' TL tendencies for the other steps (persistent option on: the
    steps must still run as stage launches that write them)."""
    es = 8 if prec == FP64 else 4
    ctx = create(prec, B)
    ok(lib.msw_set_graphs(ctx, int(graphs)), "graphs")
    if B == 1:
        ok(lib.msw_set_option(ctx, _lib.OPT_PERSISTENT, 1), "persistent")
    dev = Dev()
    s = state(dev, prec, B, n=3)
    mh, mu = dev(es * B * nC), dev(es * B * nE)
    ok(lib.msw_set_tend_output(ctx, mh, mu, None, None), "tend output, members")
    ok(lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 2), "msw_fwd_steps kept tendencies")
    bh, bu, th, tu = dev(8 * nC), dev(8 * nE), dev(es * B * nC), dev(es * B * nE)
    ok(lib.msw_set_tend_output(ctx, bh, bu, th, tu), "tend output, base and TL")
    for cfg in (config(), config(visc2=0.1, visc4=0.1, wind_stress=1, bottom_drag=1), config(test_case=1)):
        ok(lib.msw_set_config(ctx, C.byref(cfg)), "set_config")
        ok(lib.msw_base_steps(ctx, s["h"], s["u"], 2, None, None), "msw_base_steps kept tendencies")
        tlm_adj(ctx, s)
        tlm_adj(ctx, s)
        ok(lib.msw_adj_step(ctx, s["h"], s["u"], s["ht"], s["ut"], None, None), "msw_adj_step kept tendencies")
    ok(lib.msw_set_tend_output(ctx, None, None, None, None), "tend output off")
    tlm_adj(ctx, s)
    ok(lib.msw_destroy(ctx), "destroy")
    dev.free()
    no_violations(f"tendency outputs prec={prec} B={B} graphs={graphs}")


def persistent_path():
    """The persistent launch of a batch-1 context split into instances: its
    buffers, the error-flag read-back, and what it refuses."""
    dev = Dev()
    for prec in (FP64, FP32):
        ctx = create(prec, 1)
        s = state(dev, prec, 1, n=3)
        cs = (C.c_int32 * 3)(0, 200, 400)
        es = (C.c_int32 * 3)(0, 600, 1200)
        vs = (C.c_int32 * 3)(0, 400, 800)
        ok(lib.msw_set_instances(ctx, 3, cs, es, vs), "msw_set_instances")
        ok(lib.msw_set_option(ctx, _lib.OPT_PERSISTENT, 1), "persistent")
        ok(lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 3), "persistent forward")
        ok(lib.msw_tlm_steps(ctx, s["h"], s["u"], s["ht"], s["ut"], 3, s["th"], s["tu"]), "persistent TLM")
        ok(lib.msw_tlm_steps(ctx, s["h"], s["u"], s["ht"], s["ut"], 2, None, None), "persistent TLM, no trajectory")
        ok(lib.msw_adj_steps(ctx, s["th"], s["tu"], s["ht"], s["ut"], 3), "adjoint after a persistent TLM")
        bad = (C.c_int32 * 3)(0, 700, 600)
        if lib.msw_set_instances(ctx, 3, cs, bad, vs) != _lib.MSW_EINVAL:
            fail("msw_set_instances accepted decreasing starts")
        ok(lib.msw_set_config(ctx, C.byref(config(visc2=0.1))), "viscosity on")
        if lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 1) != _lib.MSW_EUNSUPPORTED:
            fail("the persistent path accepted a viscosity")
        ok(lib.msw_destroy(ctx), "destroy")
    ctx = create(FP64, 64)
    ok(lib.msw_set_option(ctx, _lib.OPT_PERSISTENT, 1), "persistent")
    s = state(dev, FP64, 64, n=1)
    if lib.msw_fwd_steps(ctx, s["ht"], s["ut"], 1) != _lib.MSW_EUNSUPPORTED:
        fail("the persistent path accepted 64 members")
    ok(lib.msw_destroy(ctx), "destroy")
    dev.free()
    no_violations("persistent path")


def errors_reach_the_caller():
    """An injected failure of every copy, of a launch and of every allocation
    in msw_create comes back as a status, and the context keeps working."""
    for graphs in (0, 1):
        ctx = create(FP64, 64)
        ok(lib.msw_set_graphs(ctx, graphs), "graphs")
        dev = Dev()
        s = state(dev, FP64, 64, n=3)
        calls = {}
        for name, fn in (
                ("msw_tlm_steps", lambda: lib.msw_tlm_steps(ctx, s["h"], s["u"], s["ht"], s["ut"], 3, s["th"], s["tu"])),
                ("msw_adj_steps", lambda: lib.msw_adj_steps(ctx, s["th"], s["tu"], s["ht"], s["ut"], 3)),
                ("msw_base_steps", lambda: lib.msw_base_steps(ctx, s["h"], s["u"], 3, s["th"], s["tu"])),
                ("msw_adj_step", lambda: lib.msw_adj_step(ctx, s["h"], s["u"], s["ht"], s["ut"], None, None))):
            ok(fn(), name)   # with graphs: captured here, replayed below
            lib.mock_fail(COPY, -1)
            ok(fn(), name)
            calls[name] = int(lib.mock_calls(COPY))
            if calls[name] == 0:
                fail(f"{name} made no copy to fail")
            for i in range(calls[name]):
                lib.mock_fail(COPY, i)
                rc = fn()
                if rc != _lib.MSW_EHIP:
                    fail(f"{name}: failed copy {i} of {calls[name]} returned {rc}, not MSW_EHIP")
                lib.mock_clear_fail()
                ok(fn(), f"{name} after a failed copy")
            if graphs:
                # copies made while a step graph is being captured
                ok(lib.msw_set_option(ctx, _lib.OPT_BLOCK, 256), "drop graphs")
                lib.mock_fail(COPY, -1)
                ok(fn(), name)
                n_cap = int(lib.mock_calls(COPY))
                for i in range(n_cap):
                    ok(lib.msw_set_option(ctx, _lib.OPT_BLOCK, 256), "drop graphs")
                    lib.mock_fail(COPY, i)
                    rc = fn()
                    lib.mock_clear_fail()
                    if rc != _lib.MSW_EHIP:
                        fail(f"{name}: failed copy {i} of {n_cap} while capturing returned {rc}")
                    ok(fn(), f"{name} after a failed copy while capturing")
            for drop in ((0, 1) if graphs else (0,)):
                if drop:   # a kernel launch that fails while its graph is captured
                    ok(lib.msw_set_option(ctx, _lib.OPT_BLOCK, 256), "drop graphs")
                lib.mock_fail(LAUNCH, 0)
                rc = fn()
                lib.mock_clear_fail()
                if rc != _lib.MSW_EHIP:
                    fail(f"{name}: failed launch (graphs {graphs}, recapture {drop}) returned {rc}")
                ok(fn(), f"{name} after a failed launch")
        ok(lib.msw_destroy(ctx), "destroy")
        dev.free()
        no_violations(f"error paths graphs={graphs}")
    base = int(lib.mock_allocations())
    lib.mock_fail(MALLOC, -1)
    ctx = create(FP32, 64)
    n_alloc = int(lib.mock_calls(MALLOC))
    ok(lib.msw_destroy(ctx), "destroy")
    for i in range(n_alloc):
        lib.mock_fail(MALLOC, i)
        h = vp()
        rc = lib.msw_create(C.byref(desc), C.byref(config()), FP32, 64, None, C.byref(h))
        lib.mock_clear_fail()
        if rc != _lib.MSW_ENOMEM or h.value:
            fail(f"msw_create with allocation {i} failing returned {rc}")
        if int(lib.mock_allocations()) != base:
            fail(f"msw_create with allocation {i} failing left {int(lib.mock_allocations()) - base} allocations")
    no_violations("allocation failures")


def graph_cache_is_bounded():
    ctx = create(FP64, 64)
    ok(lib.msw_set_graphs(ctx, 1), "graphs")
    dev = Dev()
    s = state(dev, FP64, 64, n=2)
    peak = 0
    for i in range(40):
        # a new member pair each time, as a caller allocating per call does
        d2 = Dev()
        ht, ut = d2(8 * 64 * nC), d2(8 * 64 * nE)
        ok(lib.msw_tlm_steps(ctx, s["h"], s["u"], ht, ut, 2, s["th"], s["tu"]), "tlm")
        ok(lib.msw_adj_steps(ctx, s["th"], s["tu"], ht, ut, 2), "adj")
        peak = max(peak, int(lib.mock_graphs_live()))
        d2.free()
    ok(lib.msw_destroy(ctx), "destroy")
    dev.free()
    no_violations("graph cache")
    if peak > 32:
        fail(f"graph cache grew to {peak} executable graphs over 40 buffer pairs")
    print(f"graph cache peak {peak} graphs over 40 buffer pairs", flush=True)


for prec, B in ((FP64, 4), (FP64, 64), (FP32, 64), (FP32, 1), (FP64, 256)):
    launch_shapes(prec, B)
    print(f"ok launch shapes prec={prec} B={B}", flush=True)
for prec, B in ((FP64, 1), (FP64, 64), (FP32, 64)):
    for graphs in (False, True):
        other_entry_points(prec, B, graphs)
        print(f"ok entry points prec={prec} B={B} graphs={graphs}", flush=True)
for prec, B in ((FP64, 1), (FP32, 1), (FP64, 64), (FP32, 64)):
    for graphs in (False, True):
        tend_outputs(prec, B, graphs)
        print(f"ok tendency outputs prec={prec} B={B} graphs={graphs}", flush=True)
persistent_path()
print("ok persistent path", flush=True)
errors_reach_the_caller()
print("ok error paths", flush=True)
graph_cache_is_bounded()
print(f"host check ok: {int(lib.mock_graphs_made())} graphs made, {int(lib.mock_calls(COPY))} copies", flush=True)
