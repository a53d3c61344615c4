"""Benchmark: batched TLM+adjoint member cell-steps/s on one MI355X.

Workload (the reference configuration the metric is quoted on): level-7
icosahedral Voronoi mesh (163842 cells, 491520 edges, 327680 vertices),
Williamson TC6 (Rossby-Haurwitz wave, f = 0 as in the reference), dt = 75 s,
64 seeded Gaussian perturbations (dh sigma 1 m, du sigma 0.1 m/s) stored
members-innermost, FP64.

One step = one tangent-linear RK4 step of all 64 members (which also
advances the shared FP64 nonlinear state and stores it) followed, after the
K-step TLM sweep, by the adjoint RK4 steps of all 64 members over the same K
steps in reverse (each recomputing its four nonlinear stage states from the
stored step-start state).  value = C * B * K / T with T the wall time of both
sweeps, inputs resident in HBM.

The TL tendency is formed from the fused TL thickness flux
(MSW_OPT_TL_FUSED_FLUX; same TLM to rounding).  Also reported: the same
batch in the reference's term order (bit-identical TLM), the plain unbatched
FP64 path (B = 1, same kernels, reference order) and its ratio, the
FP32-member variant, the same batch on TC6 with the physical Coriolis
parameter f = 2 Omega sin(lat) (the reference's TC6 leaves f = 0), the
algorithmic-bytes roofline, and the reference Fortran timed on this host's
CPU (one core) as cpu_baseline.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mpas-sw-tl-ad_amd"))
sys.path.insert(0, ROOT)

LEVEL, TC, DT, BATCH = 7, 6, 75.0, 64
SEED = 20261203
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = {"f64": 1024.0, "f32": 512.0}   # algorithmic, per member TLM+ADJ cell-step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def library_sha256():
    """SHA-256 of the HIP library this process loads (ties counter summaries
    to the code they measured)."""
    import hashlib
    from mpas_sw_amd import _lib
    with open(_lib.LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def git_head():
    try:
        r = subprocess.run(["git", "-C", ROOT, "rev-parse", "--short=12", "HEAD"], capture_output=True, text=True)
        return r.stdout.strip() or None
    except OSError:
        return None


def perturbations(mesh, nmem):
    import numpy as np
    import torch
    hs, us = [], []
    for m in range(nmem):
        g = torch.Generator("cpu").manual_seed(SEED + m)
        hs.append(torch.randn(mesh.nCells, generator=g, dtype=torch.float64).numpy() * 1.0)
        us.append(torch.randn(mesh.nEdges, generator=g, dtype=torch.float64).numpy() * 0.1)
    return np.stack(hs, 1), np.stack(us, 1)


def run_window(m, bufs, h0, u0, ht0, ut0, nsteps):
    """TLM sweep of nsteps (storing the trajectory) + adjoint sweep back.
    The same device buffers are reused on every call, so cached step graphs
    stay valid."""
    import torch
    h, u, ht, ut, th, tu = bufs
    h.copy_(h0)
    u.copy_(u0)
    ht.copy_(ht0)
    ut.copy_(ut0)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t0 = time.perf_counter()
    ev[0].record()
    m.tlm(h, u, ht, ut, nsteps, th, tu)
    ev[1].record()
    # adjoint seeded with the final TL state (the gradient of 1/2|dx(T)|^2)
    m.adj(th, tu, ht, ut, nsteps)
    ev[2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev[0].elapsed_time(ev[1]) / 1e3, ev[1].elapsed_time(ev[2]) / 1e3


def gpu_rate(mesh, phys, dh, du, batch, precision, steps, warmup, graphs=True, fused=True, tune=True,
             launch=None):
    import torch
    from mpas_sw_amd import _lib
    from mpas_sw_amd.core import SWModel
    m = SWModel(mesh, phys, DT, test_case=TC, batch=batch, precision=precision, graphs=graphs,
                tl_fused_flux=fused)
    h0 = m.to_internal(phys["h"], "c", member=False)
    u0 = m.to_internal(phys["u"], "e", member=False)
    sel = slice(0, batch)
    ht0 = m.to_internal(dh[:, sel] if batch > 1 else dh[:, 0], "c")
    ut0 = m.to_internal(du[:, sel] if batch > 1 else du[:, 0], "e")
    if launch is not None:        # a fixed (stripe, block) pair instead of the autotuner
        m.set_option(_lib.OPT_XCD_STRIPE, launch[0])
        m.set_option(_lib.OPT_BLOCK, launch[1])
        tune = {(m.stripe_log2, m.block): 0.0}
    elif tune:
        tune = m.autotune(h0, u0)     # launch stripe and block size, chosen by timing (not timed here)
        log(f"launch autotune B={batch} {precision} (stripe, block): {tune}")
    else:
        tune = {(m.stripe_log2, m.block): 0.0}
    n = max(steps, warmup)
    bufs = (torch.empty_like(h0), torch.empty_like(u0), torch.empty_like(ht0), torch.empty_like(ut0),
            torch.empty((n, m.nC), dtype=torch.float64, device=m.device),
            torch.empty((n, m.nE), dtype=torch.float64, device=m.device))
    if warmup > 0:
        run_window(m, bufs, h0, u0, ht0, ut0, warmup)
    wall, t_tlm, t_adj = run_window(m, bufs, h0, u0, ht0, ut0, steps)
    flags = m.nonfinite(bufs[2], bufs[3])
    ok = int(flags.sum().item()) == 0
    del m, bufs
    torch.cuda.empty_cache()
    T = t_tlm + t_adj
    return dict(seconds=T, wall=wall, t_tlm=t_tlm, t_adj=t_adj,
                rate=mesh.nCells * batch * steps / T, finite=ok, stripe_log2=min(tune, key=tune.get)[0],
                block=min(tune, key=tune.get)[1])


def calibration_copy(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of nbytes (read nbytes, write nbytes):
    the known-size kernel that scales the HBM byte counters."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"calibration copy: {reps} x {nbytes} B read + {nbytes} B written")
    del src, dst
    torch.cuda.empty_cache()


def traffic_from_counters(fetch_dir, write_dir, nsteps, out_json, out_md, command):
    """HBM bytes per TLM+ADJ step from two rocprofv3 --pmc runs of
    `bench.py --steps S --warmup W --no-cpu --no-plain --no-fp32 --no-physical-f --no-tune --calibrate`
    (FETCH_SIZE, then WRITE_SIZE; counters in KiB per dispatch).  Each counter
    is scaled by what it reads for the 1 GiB calibration copy; the library's
    kernels over the whole run are summed and divided by the S + W window
    steps of the FP64 batch."""
    import csv
    import glob

    def rows(d):
        f = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not f:
            raise FileNotFoundError(f"no counter_collection.csv under {d}")
        return list(csv.DictReader(open(f[0])))

    def scale(rs, counter, nbytes):
        vals = [float(r["Counter_Value"]) for r in rs
                if r["Counter_Name"] == counter and "copyBuffer" in r["Kernel_Name"]]
        big = max(vals)
        return nbytes / (big * 1024.0), big

    fr, wr = rows(fetch_dir), rows(write_dir)
    fs, fraw = scale(fr, "FETCH_SIZE", 1 << 30)
    ws, wraw = scale(wr, "WRITE_SIZE", 1 << 30)
    per = {}
    for rs, counter, sc, key in ((fr, "FETCH_SIZE", fs, "read"), (wr, "WRITE_SIZE", ws, "write")):
        for r in rs:
            if r["Counter_Name"] != counter or "msw::" not in r["Kernel_Name"]:
                continue
            n = r["Kernel_Name"].replace("void ", "").replace("msw::(anonymous namespace)::", "").split("(")[0]
            d = per.setdefault(n, {"read": 0.0, "write": 0.0, "calls": 0})
            d[key] += float(r["Counter_Value"]) * 1024.0 * sc
            if key == "read":
                d["calls"] += 1
    total = sum(d["read"] + d["write"] for d in per.values())
    res = {"hbm_bytes_per_step": total / nsteps, "window_steps": nsteps,
           "fetch_scale": fs, "write_scale": ws, "command": command,
           "library_sha256": library_sha256(), "commit": git_head(),
           "kernels": {k: {"read_bytes_per_call": v["read"] / max(v["calls"], 1),
                           "write_bytes_per_call": v["write"] / max(v["calls"], 1),
                           "calls": v["calls"]} for k, v in per.items()}}
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    with open(out_md, "w") as f:
        f.write("# HBM traffic of the batched TLM+ADJ step (level 7, TC6, 64 members, FP64)\n\n")
        f.write(f"Commands (from the repository root):\n\n```\n{command}\n```\n\n")
        f.write(f"Library: `libmpas_sw_amd.so` SHA-256 `{res['library_sha256']}` (commit {res['commit']}); "
                "bench.py reports this traffic only for a library with the same hash.\n\n")
        f.write(f"Calibration: the 1 GiB device copy reads FETCH_SIZE = {fraw:.0f} KiB and "
                f"WRITE_SIZE = {wraw:.0f} KiB, so FETCH_SIZE is scaled by {fs:.3f} and "
                f"WRITE_SIZE by {ws:.3f}.\n\n")
        f.write(f"HBM bytes per step (all kernels of the library, {nsteps} steps): "
                f"{total / nsteps / 1e9:.3f} GB\n\n")
        f.write("| kernel | calls | read MB/call | write MB/call |\n|---|---:|---:|---:|\n")
        for k, v in sorted(per.items(), key=lambda kv: -(kv[1]["read"] + kv[1]["write"])):
            c = max(v["calls"], 1)
            f.write(f"| `{k}` | {v['calls']} | {v['read'] / c / 1e6:.1f} | {v['write'] / c / 1e6:.1f} |\n")
    return res


def packed_workload(steps, warmup):
    """Uneven batch: 64 level-5 + 16 level-6 + 4 level-7 + 1 level-9 instances
    (4.59 M cells, alternating TC2 / TC6) in one packed index space; FP32
    forward and FP32 TLM members (FP64 base), dt 18.75 s.  Timed two ways:
    one launch per stage kernel for all instances (hipGraph per step), and
    the persistent path, one cooperative launch per call whose passes take
    (instance, patch) work items from a queue.  Returns cell-steps/s of each."""
    import torch
    from mpas_sw_amd import _lib
    from mpas_sw_amd.core import SWModel
    from mpas_sw_amd.ensemble import packed_instances
    from mpas_sw_amd.mesh import get_mesh
    t = time.time()
    for lv in (5, 6, 7, 9):
        get_mesh(lv)
        log(f"mesh level {lv}: {time.time() - t:.1f}s")
    levels = [5] * 64 + [6] * 16 + [7] * 4 + [9]
    specs = [(lv, 2 if i % 2 == 0 else 6) for i, lv in enumerate(levels)]
    mesh, phys, offs, perm = packed_instances(specs)
    log(f"packed set-up: {time.time() - t:.1f}s; C={mesh.nCells} E={mesh.nEdges} V={mesh.nVertices}")
    m = SWModel(mesh, phys, 18.75, test_case=2, perm=perm, precision="fp32", graphs=True)
    m.set_instances(offs)
    g = torch.Generator("cpu").manual_seed(SEED)
    dh = torch.randn(mesh.nCells, generator=g, dtype=torch.float64).numpy()
    du = 0.1 * torch.randn(mesh.nEdges, generator=g, dtype=torch.float64).numpy()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn(warmup)
        torch.cuda.synchronize()
        e0.record()
        fn(steps)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    out = {"workload": "uneven batch 64xL5 + 16xL6 + 4xL7 + 1xL9, TC2/TC6, dt 18.75 s, FP32 members",
           "cells": mesh.nCells, "edges": mesh.nEdges, "vertices": mesh.nVertices, "instances": len(specs),
           "steps": steps}
    ok = True
    for mode in ("stage_launches", "persistent"):
        m.set_persistent(mode == "persistent")
        h = m.to_internal(phys["h"], "c")
        u = m.to_internal(phys["u"], "e")
        t_fwd = timed(lambda n: m.fwd(h, u, n))
        hb = m.to_internal(phys["h"], "c", member=False)
        ub = m.to_internal(phys["u"], "e", member=False)
        ht = m.to_internal(dh, "c")
        ut = m.to_internal(du, "e")
        t_tlm = timed(lambda n: m.tlm(hb, ub, ht, ut, n))
        ok = ok and int(m.nonfinite(h, u).sum().item()) == 0 and int(m.nonfinite(ht, ut).sum().item()) == 0
        # the TLM again with the fused TL flux (the form the 64-member bench uses)
        m.set_option(_lib.OPT_TL_FUSED_FLUX, 1)
        ht, ut = m.to_internal(dh, "c"), m.to_internal(du, "e")
        t_tlf = timed(lambda n: m.tlm(hb, ub, ht, ut, n))
        m.set_option(_lib.OPT_TL_FUSED_FLUX, 0)
        ok = ok and int(m.nonfinite(ht, ut).sum().item()) == 0
        log(f"packed {mode}: forward {1e3 * t_fwd / steps:.3f} ms/step, TLM {1e3 * t_tlm / steps:.3f} ms/step, "
            f"TLM with the fused TL flux {1e3 * t_tlf / steps:.3f} ms/step")
        out[mode] = {"fwd_cell_steps_per_s": mesh.nCells * steps / t_fwd, "fwd_ms_per_step": 1e3 * t_fwd / steps,
                     "tlm_cell_steps_per_s": mesh.nCells * steps / t_tlm, "tlm_ms_per_step": 1e3 * t_tlm / steps,
                     "tlm_fused_flux_ms_per_step": 1e3 * t_tlf / steps}
    out["finite"] = ok
    return out


def gradient_workload(steps, snapshots):
    """Level 8 (655362 cells), TC5, dt 37.5 s: J = 1/2 sum A (h(T) - h_obs)^2,
    h_obs = the unperturbed forward h(T) + N(0, 10 m) (seed 20261203 + 1000);
    the gradient wrt (h0, u0) through the binomially checkpointed autograd
    window, with FP64 and with FP32 adjoint members.  Returns the wall time
    of each gradient and their relative difference."""
    import torch
    from mpas_sw_amd.checkpoint import forward_cost, rk4_window
    from mpas_sw_amd.core import SWModel
    from mpas_sw_amd.mesh import get_mesh
    from mpas_sw_amd.testcases import initial_conditions
    t = time.time()
    mesh = get_mesh(8)
    phys = initial_conditions(mesh, 5)
    log(f"level-8 set-up: {time.time() - t:.1f}s")
    res = {"workload": f"level-8 TC5 dt 37.5 s, {steps} steps, {snapshots} binomial snapshots, "
                       "J = 1/2 sum A (h(T) - h_obs)^2",
           "cells": mesh.nCells, "steps": steps, "snapshots": snapshots,
           "recomputed_forward_steps": forward_cost(steps, snapshots)}
    grads, hobs = {}, None
    for prec in ("fp64", "fp32"):
        m = SWModel(mesh, phys, 37.5, test_case=5, precision=prec, graphs=True)
        h0 = m.to_internal(phys["h"], "c", member=False)
        u0 = m.to_internal(phys["u"], "e", member=False)
        A = m.to_internal(phys["areaCell"], "c", member=False)
        if hobs is None:
            hT, uT = h0.clone(), u0.clone()
            m.base(hT, uT, steps)
            g = torch.Generator("cpu").manual_seed(SEED + 1000)
            noise = torch.randn(mesh.nCells, generator=g, dtype=torch.float64).numpy() * 10.0
            hobs = hT + m.to_internal(noise, "c", member=False)
            log("level-8 observation forward done")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hg, ug = h0.clone().requires_grad_(True), u0.clone().requires_grad_(True)
        hTg, _ = rk4_window(m, hg, ug, steps, snapshots)
        (0.5 * (A * (hTg - hobs) ** 2).sum()).backward()
        torch.cuda.synchronize()
        res[f"{prec}_gradient_s"] = time.perf_counter() - t0
        log(f"level-8 gradient {prec}: {res[f'{prec}_gradient_s']:.1f}s")
        grads[prec] = torch.cat([hg.grad, ug.grad])
        del m
    d = grads["fp32"] - grads["fp64"]
    res["fp32_vs_fp64_gradient_rel_l2"] = float(torch.linalg.norm(d) / torch.linalg.norm(grads["fp64"]))
    res["finite"] = bool(torch.isfinite(grads["fp64"]).all() and torch.isfinite(grads["fp32"]).all())
    return res


def cpu_baseline(mesh, phys, dh, du, steps=2):
    """Reference Fortran (oracle/_ref/msw_ref, one core) TLM and adjoint
    steps on the same mesh and case; member cell-steps/s of TLM+ADJ."""
    from oracle import ref
    if not ref.available():
        return None
    t_tlm = ref.run(mesh, "time_tlm", nsteps=steps, test_case=TC, dt=DT, tl=(dh[:, 0], du[:, 0]))["seconds"][0]
    t_adj = ref.run(mesh, "time_adj", nsteps=steps, test_case=TC, dt=DT, ad=(dh[:, 0], du[:, 0]))["seconds"][0]
    per_step = (t_tlm + t_adj) / steps
    return dict(value=mesh.nCells / per_step, unit="member cell-steps/s (TLM+ADJ)", cores=1,
                kind="reference",
                sample=f"reference Fortran sw_rk4_tlm x{steps} + sw_rk4_adj x{steps}, level 7 TC6 "
                       f"dt 75 s, 1 member, amdflang -O2, 1 thread; TLM {t_tlm / steps:.3f} s/step, "
                       f"ADJ {t_adj / steps:.3f} s/step")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=96)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--level", type=int, default=LEVEL)
    ap.add_argument("--no-graphs", action="store_true", help="launch stage kernels directly")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-physical-f", action="store_true",
                    help="skip the TC6 run with f = 2 Omega sin(lat)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch stripe (no autotuning runs before the timed window)")
    ap.add_argument("--launch", nargs=2, type=int, metavar=("STRIPE_LOG2", "BLOCK"),
                    help="fixed XCD stripe and block size for the batched runs (no autotuning)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 1 GiB device copy (reference for HBM counters)")
    ap.add_argument("--packed", action="store_true",
                    help="time the uneven-batch workload (64+16+4+1 instances, FP32) and exit")
    ap.add_argument("--gradient", action="store_true",
                    help="time the level-8 checkpointed gradient (FP64 and FP32 adjoint) and exit")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "`--no-cpu --no-plain --no-fp32 --no-physical-f --calibrate` into profiles/ and exit")
    args = ap.parse_args()
    if args.traffic_from:
        legs = "--no-cpu --no-plain --no-fp32 --no-physical-f --no-tune --calibrate"
        cmd = ("rocprofv3 --pmc FETCH_SIZE --output-format csv -d prof/fetch -o run -- python3 bench.py "
               f"--steps {args.steps} --warmup {args.warmup} {legs}\n"
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d prof/write -o run -- python3 bench.py "
               f"--steps {args.steps} --warmup {args.warmup} {legs}\n"
               f"python3 bench.py --steps {args.steps} --warmup {args.warmup} --traffic-from "
               "prof/fetch prof/write")
        r = traffic_from_counters(args.traffic_from[0], args.traffic_from[1], args.steps + args.warmup,
                                  os.path.join(ROOT, "profiles", "traffic_tlm_adj_l7_b64_fp64.json"),
                                  os.path.join(ROOT, "profiles", "traffic_tlm_adj_l7_b64_fp64.md"), cmd)
        print(json.dumps({"hbm_bytes_per_step": r["hbm_bytes_per_step"]}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; the batched members would shard "
              "across GPUs with no per-step exchange", file=sys.stderr)
        sys.exit(2)
    if args.packed:
        print(json.dumps(packed_workload(args.steps, args.warmup)), flush=True)
        return
    if args.gradient:
        print(json.dumps(gradient_workload(args.steps, 32)), flush=True)
        return
    import numpy as np
    import torch
    from mpas_sw_amd.mesh import get_mesh
    from mpas_sw_amd.testcases import initial_conditions

    t = time.time()
    mesh = get_mesh(args.level)
    phys = initial_conditions(mesh, TC)
    dh, du = perturbations(mesh, BATCH)
    log(f"setup L{args.level}: {time.time() - t:.1f}s; C={mesh.nCells} E={mesh.nEdges} V={mesh.nVertices}")

    if args.calibrate:
        calibration_copy()
    g = not args.no_graphs
    res = gpu_rate(mesh, phys, dh, du, BATCH, "fp64", args.steps, args.warmup, g, tune=not args.no_tune,
                   launch=args.launch)
    log(f"batched fp64: {res}")
    exact = None
    if not args.no_plain:
        exact = gpu_rate(mesh, phys, dh, du, BATCH, "fp64", args.steps, args.warmup, g, fused=False)
        log(f"batched fp64, reference term order: {exact}")
    plain = None
    if not args.no_plain:
        plain = gpu_rate(mesh, phys, dh, du, 1, "fp64", args.steps, args.warmup, g, fused=False)
        log(f"plain fp64: {plain}")
    f32 = None
    if not args.no_fp32:
        # at the default launch shape (stripe 2, 256 threads), without the
        # autotuner's option changes: the 64-member FP32 test of those changes
        # is kept off the GPU (a host fault whose cause is not found)
        f32 = gpu_rate(mesh, phys, dh, du, BATCH, "fp32", args.steps, args.warmup, g, tune=False)
        log(f"batched fp32: {f32}")
    physf = None
    if not args.no_physical_f:
        from mpas_sw_amd.testcases import OMEGA
        pf = dict(phys)
        pf["fVertex"] = 2.0 * OMEGA * np.sin(mesh.latVertex)
        pf["fEdge"] = 2.0 * OMEGA * np.sin(mesh.latEdge)
        physf = gpu_rate(mesh, pf, dh, du, BATCH, "fp64", args.steps, args.warmup, g, tune=False)
        log(f"batched fp64, TC6 with f = 2 Omega sin(lat): {physf}")

    K = args.steps
    bytes_step = BYTES_PER_CELL_STEP["f64"] * mesh.nCells * BATCH
    achieved = bytes_step * K / res["seconds"] / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "kernel": "one TLM+ADJ RK4 step of 64 members (all stage kernels, hipGraph)",
        "algorithmic_bytes_per_step": bytes_step,
        "avg_step_ms": round(1e3 * res["seconds"] / K, 4),
    }
    prof = os.path.join(ROOT, "profiles", "traffic_tlm_adj_l7_b64_fp64.json")
    if os.path.isfile(prof):
        # counter bytes measured on this very library only; otherwise left out
        with open(prof) as f:
            tr = json.load(f)
        if tr.get("library_sha256") == library_sha256():
            roofline["traffic"] = tr.get("hbm_bytes_per_step")
            roofline["traffic_source"] = f"profiles/traffic_tlm_adj_l7_b64_fp64.json (commit {tr.get('commit')})"
        else:
            roofline["traffic_source"] = "not measured for this library build (counter summary is of another build)"

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(mesh, phys, dh, du)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
    out = {
        "metric": "TLM+ADJ cell-steps/s at 163842 cells, batch 64 vs plain; % HBM peak",
        "value": res["rate"],
        "unit": "member cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["seconds"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated level-7 mesh, Williamson TC6, seeded perturbations)",
        "config": {"workload": "batched TLM+ADJ RK4, level-7 icosahedral Voronoi mesh, TC6, dt 75 s",
                   "cells": mesh.nCells, "edges": mesh.nEdges, "vertices": mesh.nVertices,
                   "batch": BATCH, "dt": DT, "window_steps": K,
                   "tl_tendency": "fused TL flux (MSW_OPT_TL_FUSED_FLUX)",
                   "xcd_stripe_log2": res["stripe_log2"], "block": res["block"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_reference_term_order": None if exact is None else {
            "value": exact["rate"], "ms_per_step": 1e3 * exact["seconds"] / K},
        "plain_fp64": None if plain is None else {"value": plain["rate"], "ms_per_step": 1e3 * plain["seconds"] / K},
        "batched_vs_plain": None if plain is None else res["rate"] / plain["rate"],
        "fp32": None if f32 is None else {
            "value": f32["rate"], "ms_per_step": 1e3 * f32["seconds"] / K,
            "xcd_stripe_log2": f32["stripe_log2"], "block": f32["block"],
            "roofline_frac": round(BYTES_PER_CELL_STEP["f32"] * mesh.nCells * BATCH * K /
                                   f32["seconds"] / 1e9 / HBM_PEAK_GBS, 4)},
        "tc6_physical_f": None if physf is None else {
            "value": physf["rate"], "ms_per_step": 1e3 * physf["seconds"] / K, "finite": physf["finite"]},
        "tlm_s": res["t_tlm"], "adj_s": res["t_adj"],
        "finite": res["finite"] and (f32 is None or f32["finite"]),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
