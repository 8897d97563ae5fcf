#!/usr/bin/env python3
"""Benchmark of the MI355X time step on the headline workload.

Workload (default, the largest single-GPU configuration): 3D hex box of
256^3 = 16,777,216 unit cells, inlet/outlet patches on all six sides, a
Gaussian acoustic pulse of 1e-3 p_inf (half-width 8 cells, seeded centre
jitter) in a uniform flow at M = 0.1, M1 central-upwind flux, inviscid, RK5
at CFL 0.5.  Synthetic mesh and fields.

A "step" is one full time step (5 RK stages) of every cell; the metric is
cell-updates/s = cells x steps / s with the state resident in HBM.  The
FP32 deviation-from-freestream path is the headline value; the FP64 path is
timed beside it, and the FP32 error against FP64 is measured after the
configuration's own step count (1000), independent of --steps.

Prints ONE JSON line (one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# host cores allotted to a one-GPU box: the CPU baseline uses at most these
BOX_CPU_SHARE = 16
OWN_STEPS = {"box3d": 1000, "cylinder2d": 2000, "vortex2d": 1000, "ogrid3d": 500}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(nc, nf, D, w_s, w_g):
    """Algorithmic bytes of one RK stage: q and dq read + written, 1/V and
    sigma per cell; S, d, w and the two cell indices per face."""
    return nc * (4 * (D + 2) * w_s + 2 * w_g) + nf * ((2 * D + 1) * w_g + 8)


def build_case(workload, solver=2, renumber="none"):
    """The workload and a one-line description of it."""
    from lfm_public_amd import cases
    if workload == "cylinder2d":
        n = 1024
        c = cases.cylinder_2d(n_theta=n, n_quad=n // 4, n_tri=3 * n // 8, solver=solver)
        desc = (f"2D mixed quad/tri annulus around a unit cylinder (10 % node jitter), M=0.3, 1e-4 p noise, "
                f"solver {solver} ({'M2 + minmod AUSM dissipation' if solver == 2 else 'M2, minmodExists'})"
                f", viscous, sponge Ls=5D, RK5")
    elif workload == "vortex2d":
        c = cases.isentropic_vortex(nx=256)
        desc = "2D 256x256 quad box, isentropic vortex (strength 5) at M=0.5, M1, inviscid, RK5, CFL 0.5"
    elif workload == "ogrid3d":
        c = cases.cylinder_3d(nr=256, nth=256, nz=64, solver=1, minmod=1)
        desc = ("3D cylinder O-grid 256x256x64 hexes, spanwise cyclic, M=0.2, mu from Re=3900, 1 % spanwise "
                "perturbation, M2 + minmodExists, sponge Ls=13, RK5")
    elif workload == "box3d":
        c = cases.acoustic_box(n=256)
        desc = ("3D 256^3 hex box, Gaussian pulse 1e-3 p_inf (half-width 8 cells) in M=0.1 flow, M1, inviscid, "
                "RK5, CFL 0.5")
    else:
        raise SystemExit(f"unknown workload {workload}")
    if renumber != "none":
        from lfm_public_amd import renumber as rn
        c = rn.renumber_case(c, renumber)
        desc += f"; cells renumbered ({renumber})"
    return c, desc


def cpu_sample_case(workload, case):
    """The CPU baseline's sample: the workload itself up to ~1.5M cells, else
    the same case family at a size one CPU solve finishes steps of in seconds."""
    from lfm_public_amd import cases
    if case.n_cells <= 1_500_000:
        return case, "the full mesh"
    if workload == "ogrid3d":
        c = cases.cylinder_3d(nr=64, nth=64, nz=32, solver=1, minmod=1)
    else:
        c = cases.acoustic_box(n=64)
    return c, f"the same case family at {c.n_cells} cells ({c.name})"


def launch_plan(prm, nstages):
    """Launches per time step by profiling group, as Solver::step issues them:
    {group: (launches, kernel name)}.  FP32 inviscid M1/M2 runs every stage
    (but a residual stage 0) as one stage_fused32 launch; other paths run
    grad_vis (when it has work), face_flux and cell_update per stage.  The
    last stage's launch also forms the CFL/dt block extrema, except on the
    FP64 3D path, where they keep a launch of their own.  Their final
    reduction is one single-block launch, or, inside a run that steps
    again, one extra block of the next step's first FP32 launch."""
    D, S = prm["dim"], prm["solver"]
    fp32 = prm["precision"] == 32
    visc = prm["mu"] != 0.0 or not prm.get("laminar", 1)
    sfx = "32" if fp32 else ""
    mode = 0 if prm.get("adjust_dt") else 1
    if fp32 and not visc and S < 2:
        first = 1 if prm.get("have_res") else 0
        plan = {"cell_update": (nstages - 1 - first, f"stage_fused32<{D}, {S}, -1>"),
                "cell_update_cfl": (1, f"stage_fused32<{D}, {S}, {mode}>")}
        if first:
            plan["face_flux"] = (first, f"face_flux32<{D}, {S}, false, false>")
        return plan
    rg = S == 2 and prm.get("minmod")
    plan = {"face_flux": (nstages, f"face_flux{sfx}<{D}, {S}, {str(visc).lower()}, false>")}
    if fp32 or D == 2:
        plan["cell_update"] = (nstages - 1, f"cell_update{sfx}<{D}, false, -1>")
        plan["cell_update_cfl"] = (1, f"cell_update{sfx}<{D}, false, {mode}>")
    else:
        plan["cell_update"] = (nstages, f"cell_update<{D}, false, -1>")
        plan["cfl_dt"] = (1, f"cfl_dt_cells<{D}, {mode}>")
    if visc or S == 2:
        # FP32 3D gradients read the face slots (the kernel's last parameter)
        slots = f", {str(D == 3).lower()}" if fp32 else ""
        plan["grad_vis"] = (nstages, f"grad_vis{sfx}<{D}, {2 if rg else 0}, {str(visc).lower()}, "
                                     f"{str(S == 2).lower()}{slots}>")
    return plan


class Restarts:
    """Keeps a solve within its configuration's own step count: before a run
    of n steps would take it past `limit` steps from the initial state, the
    initial state is loaded again.  Timing is independent of the state, but
    a small case warmed up for seconds would otherwise march tens of
    thousands of steps, past where its flow stays valid (the vortex leaves
    the box)."""

    def __init__(self, g, limit):
        from lfm_public_amd import capi
        self.g, self.limit, self.field = g, max(1, int(limit)), capi.FIELD_Q
        self.q0 = g.q()
        self.since = 0
        self.count = 0

    def before(self, n):
        if self.since + n > self.limit:
            self.g.set_field(self.field, self.q0)
            self.since = 0
            self.count += 1
        self.since += n


def warm_up(g, min_steps, min_s, already_s=0.0, restarts=None):
    """At least min_steps steps, and stepping until min_s seconds have passed
    including already_s of earlier stepping (GPU clocks settle); returns
    (steps, seconds of this call)."""
    t0 = time.perf_counter()
    n = 0
    chunk = max(1, min_steps)
    while True:
        if restarts:
            chunk = min(chunk, restarts.limit)
            restarts.before(chunk)
        g.run(chunk)
        g.synchronize()
        n += chunk
        el = time.perf_counter() - t0
        if n >= min_steps and el + already_s >= min_s:
            return n, el
        per_step = el / n
        chunk = int(min(2000, max(1, (min_s - already_s - el) / max(per_step, 1e-6))))


def time_path(case, precision, steps, warmup, warmup_s, nprof, err_steps=0, limit=None):
    """Create the solver; run err_steps steps from the initial state (the
    state then is kept for the FP32 error; the run also warms the GPU up);
    warm up; time `steps` steps; profile nprof more by kernel group.  With a
    step limit the warm-up, timed and profiled runs restart from the initial
    state rather than go past it (a timed run longer than the limit runs
    as it is)."""
    from lfm_public_amd import capi
    t0 = time.perf_counter()
    g = capi.GpuSolver.from_case(case, precision=precision)
    setup_s = time.perf_counter() - t0
    info = g.info()
    restarts = Restarts(g, limit) if limit else None
    q_err, err_s = None, 0.0
    if err_steps > 0:
        t0 = time.perf_counter()
        if restarts:
            restarts.before(err_steps)
        g.run(err_steps)
        g.synchronize()
        err_s = time.perf_counter() - t0
        q_err = g.q()
    wsteps, wsec = warm_up(g, warmup, warmup_s, err_s, restarts)
    if restarts:
        restarts.before(steps)
    # timed region: the device drained on both sides, K steps queued back to
    # back; HIP events on the solver's stream bracket the same K steps
    g.synchronize()
    t0 = time.perf_counter()
    ms_events = g.time_steps(steps)
    g.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    if restarts:
        restarts.before(nprof)
    prof, prof_total = g.profile_steps(nprof)
    return g, info, {"wall_ms": wall_ms, "event_ms": ms_events, "warmup_steps": wsteps,
                     "warmup_s": wsec + err_s, "error_run_steps": err_steps, "setup_s": setup_s, "prof": prof,
                     "prof_total_ms": prof_total, "restarts": restarts.count if restarts else 0}, q_err


def rel_errors(case, q, qref, D):
    from lfm_public_amd import cases
    gam = cases.gamma_of(case.params)
    R = cases.gas_constant(case.params["mol_weight"])

    def prim(qq):
        rho = qq[0]
        U = qq[1:D + 1] / rho
        p = (qq[D + 1] - 0.5 * rho * (U * U).sum(0)) * (gam - 1)
        return p, U, p / (rho * R)

    p, U, T = prim(q)
    pr, Ur, Tr = prim(qref)
    inf = {"p": case.params["p_inf"], "T": case.params["T_inf"]}
    out = {}
    for name, a, b in (("p", p, pr), ("T", T, Tr)):
        out[name] = {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                     "pert_l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b - inf[name]), 1e-300))}
    for k in range(D):
        uinf = case.params["U_inf"][k]
        out[f"U{k}"] = {"rel_l2": float(np.linalg.norm(U[k] - Ur[k]) / max(np.linalg.norm(Ur[k]), 1e-300)),
                        "pert_l2": float(np.linalg.norm(U[k] - Ur[k]) / max(np.linalg.norm(Ur[k] - uinf), 1e-300))}
    # the velocity vector as one field (a component that is itself a small
    # perturbation, like a spanwise velocity, is ill-conditioned on its own)
    uinf = np.asarray(case.params["U_inf"][:D], np.float64)[:, None]
    out["U"] = {"rel_l2": float(np.linalg.norm(U - Ur) / np.linalg.norm(Ur)),
                "pert_l2": float(np.linalg.norm(U - Ur) / max(np.linalg.norm(Ur - uinf), 1e-300))}
    return out


def cpu_baseline(case, budget_s, threads, what="the full mesh"):
    """The oracle (the C restatement of the reference's time step, one
    thread per solve) on the same case family, run as the reference runs on
    a CPU node: one independent single-rank solve per core, all cores busy,
    a bounded number of steps; the aggregate rate over the cores."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    insts = [None] * threads

    def make(i):
        insts[i] = O.Oracle(case.mesh, case.params, case.p, case.T, case.U, case.alpha)

    def run(i, n):
        insts[i].run(n)

    def parallel(fn, *a):
        ts = [threading.Thread(target=fn, args=(i,) + a) for i in range(threads)]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return time.perf_counter() - t0

    parallel(make)
    t1 = parallel(run, 1)  # warm-up step (page faults, caches)
    n = max(1, min(10, int(budget_s / max(t1, 1e-3)) - 1))
    wall = parallel(run, n)
    insts.clear()
    return {"value": threads * case.n_cells * n / wall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference time step, FP64) on {what} ({case.n_cells} cells): "
                      f"{threads} independent single-thread solves on {threads} cores, {n} RK5 steps each after 1 "
                      f"warm-up step, {wall:.1f} s; aggregate cell-updates/s (the reference itself needs MPI and "
                      f"OpenFOAM, absent here)"}


def traffic_of(key, kernel):
    """Measured HBM bytes per launch of one kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload."""
    f = os.path.join(ROOT, "profiles", f"{key}_traffic.json")
    try:
        with open(f) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None, None
    k = t.get("kernels", {}).get(kernel)
    if t.get("workload") != key or not k:
        return None, None
    src = (f"profiles/{key}_traffic.json: rocprofv3 FETCH_SIZE x{t['fetch_factor']:.3f} + WRITE_SIZE "
           f"x{t['write_factor']:.3f} per launch of {kernel} (factors from a 256 MiB device copy); "
           f"{t.get('note', '')}".rstrip("; "))
    return int(k["fetch_bytes"] + k["write_bytes"]), src


def roofline(case, nf, path_prm, nstages, prof, nprof, ms_step, w, traffic_key):
    """Roofline of the path's dominant kernel: its algorithmic bytes per
    launch over its mean launch duration (HIP events on the solver's stream,
    over the profiled steps), plus the step-level figure."""
    D = case.params["dim"]
    nc = case.n_cells
    B = stage_bytes(nc, nf, D, w, w)
    plan = launch_plan(path_prm, nstages)
    per_launch = {g: prof[g] / (n * nprof) for g, (n, _) in plan.items() if n > 0}
    per_step = {g: prof[g] / nprof for g in plan}
    dom = max(per_step, key=per_step.get)
    n_launch, kname = plan[dom]
    one_stage = kname.startswith("stage_fused")
    out = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "dominant_kernel": kname,
           "launches_per_step": {plan[g][1]: plan[g][0] for g in plan},
           "us_per_launch_events": {plan[g][1]: round(v * 1e3, 2) for g, v in per_launch.items()},
           "algorithmic_bytes_per_stage": int(B), "algorithmic_bytes_per_step": int(nstages * B),
           "step_achieved": round(nstages * B / (ms_step * 1e-3) / 1e9, 1),
           "step_frac": round(nstages * B / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if one_stage:
        t = per_launch[dom] * 1e-3
        ach = B / t / 1e9
        out.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                    "algorithmic_bytes_per_launch": int(B),
                    "achieved_basis": f"one {kname} launch = one RK stage of every cell ({B} algorithmic bytes: "
                                      f"Nc*[4(D+2)w+2w] + Nf*[(2D+1)w+8], w the word size) / its mean launch time "
                                      f"{per_launch[dom] * 1e3:.1f} us (HIP events, {nprof} steps)"})
    else:
        # three launches per stage: the per-stage bytes over the stage's time
        t_stage = sum(per_step[g] for g in ("grad_vis", "face_flux", "cell_update", "cell_update_cfl")
                      if g in per_step) / nstages * 1e-3
        ach = B / t_stage / 1e9
        out.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                    "achieved_basis": "one RK stage (grad_vis, face_flux, cell_update launches) of per-stage "
                                      "algorithmic bytes over the stage's mean event time"})
    traffic, src = traffic_of(traffic_key, kname)
    out["traffic"] = traffic
    if traffic:
        out["traffic_source"] = src
        out["traffic_over_algorithmic"] = round(traffic / B, 3) if one_stage else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configuration's own step count)")
    ap.add_argument("--warmup", type=int, default=20, help="minimum warm-up steps")
    ap.add_argument("--warmup-s", type=float, default=2.0, help="minimum warm-up seconds per path")
    ap.add_argument("--workload", default="box3d", choices=sorted(OWN_STEPS))
    ap.add_argument("--solver", type=int, default=2, help="cylinder2d scheme: 2 (M2AUSM + minmod) or 1 (M2)")
    ap.add_argument("--renumber", default="none", choices=("none", "morton", "rcm", "column"),
                    help="renumber the cells before the run (renumberMesh-style preprocessing)")
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps after which the FP32 error against FP64 is measured (default: the "
                         "configuration's own step count; 0 skips)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help=f"CPU baseline cores (default: every visible core, at most {BOX_CPU_SHARE})")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--secondary", default="cylinder2d",
                    help="a second workload timed briefly after the headline ('none' skips)")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests of the serving stream run after the timed region (0: none)")
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--fp32-only", action="store_true",
                    help="time the FP32 path alone (profiling runs; no FP64 line, no error)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = OWN_STEPS[args.workload]
    if args.error_steps is None:
        args.error_steps = OWN_STEPS[args.workload]
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by spatial "
              f"decomposition but only the single-GPU partition is built and measured.", file=sys.stderr)
        sys.exit(2)

    from lfm_public_amd import capi
    t0 = time.perf_counter()
    case, desc = build_case(args.workload, args.solver, args.renumber)
    D = case.params["dim"]
    nc = case.n_cells
    log(f"workload {case.name}: {nc} cells, built in {time.perf_counter() - t0:.1f} s")
    nprof = max(1, min(args.steps, 10))
    prm64 = dict(case.params, precision=64)
    prm32 = dict(case.params, precision=32)
    nstages = int(case.params["rk_order"])  # rkOrder 3 runs 3 stages of the 4-stage table
    traffic_key = case.name

    # ---- FP64 -------------------------------------------------------------
    n_err = 0 if (args.fp64_only or args.fp32_only) else args.error_steps
    have64 = not args.fp32_only
    q64 = None
    if have64:
        g64, info64, t64, q64 = time_path(case, capi.PREC_FP64, args.steps, args.warmup, args.warmup_s, nprof,
                                          n_err, limit=OWN_STEPS[args.workload])
        nf = info64.n_faces
        ms64 = t64["wall_ms"] / args.steps
        rate64 = nc * args.steps / (t64["wall_ms"] * 1e-3)
        log(f"FP64: {ms64:.3f} ms/step ({t64['event_ms'] / args.steps:.3f} by events), {rate64:.3e} "
            f"cell-updates/s")
        bad64 = g64.n_nonfinite()
        g64.close()
        if bad64:
            raise SystemExit(f"FP64 state went non-finite ({bad64} values): the workload's dt is unstable")

    # ---- FP32 deviation path ---------------------------------------------
    have32 = False
    err = None
    if not args.fp64_only:
        try:
            g32, info32, t32, q32 = time_path(case, capi.PREC_FP32P, args.steps, args.warmup, args.warmup_s,
                                              nprof, n_err, limit=OWN_STEPS[args.workload])
            have32 = True
        except capi.LfmError as e:
            log(f"FP32-perturbation path unavailable: {e}")
    if have32:
        nf = info32.n_faces
        ms32 = t32["wall_ms"] / args.steps
        rate32 = nc * args.steps / (t32["wall_ms"] * 1e-3)
        log(f"FP32P: {ms32:.3f} ms/step ({t32['event_ms'] / args.steps:.3f} by events), {rate32:.3e} "
            f"cell-updates/s")
        if q32 is not None:
            err = rel_errors(case, q32, q64, D)
            log(f"FP32P vs FP64 after {n_err} steps: p' {err['p']['pert_l2']:.2e}, T "
                f"{err['T']['pert_l2']:.2e}, U {err['U']['pert_l2']:.2e} (perturbation-normalised)")
        bad32 = g32.n_nonfinite()
        g32.close()
        if bad32:
            raise SystemExit(f"FP32P state went non-finite ({bad32} values)")
        del q32
    del q64

    rl64 = roofline(case, nf, prm64, nstages, t64["prof"], nprof, ms64, 8, traffic_key) if have64 else None
    if have32:
        value, ms_step, dtype = rate32, ms32, "f32"
        rl = roofline(case, nf, prm32, nstages, t32["prof"], nprof, ms32, 4, traffic_key)
    else:
        value, ms_step, dtype = rate64, ms64, "f64"
        rl = rl64

    cpu = None
    visible = len(os.sched_getaffinity(0))
    if not args.no_cpu:
        threads = args.cpu_threads or min(BOX_CPU_SHARE, visible)
        cc, what = cpu_sample_case(args.workload, case)
        # one oracle instance holds ~0.7 kB (2D) / ~1.2 kB (3D) per cell
        per = cc.n_cells * (700 if D == 2 else 1200)
        try:
            avail = int([l for l in open("/proc/meminfo") if l.startswith("MemAvailable")][0].split()[1]) * 1024
            threads = max(1, min(threads, int(0.5 * avail // per)))
        except (OSError, IndexError, ValueError):
            pass
        log(f"CPU baseline: {threads} threads on {cc.n_cells} cells")
        cpu = cpu_baseline(cc, args.cpu_budget, threads, what)
        cpu["host_cpus_visible"] = visible
        cpu["note"] = (f"{threads} of {visible} visible cores: the host share of a one-GPU box is "
                       f"{BOX_CPU_SHARE} cores")

    def path_rec(t, rate, ms, info):
        return {"cell_updates_per_s": rate, "ms_per_step": ms, "ms_per_step_events": t["event_ms"] / args.steps,
                "warmup_steps": t["warmup_steps"], "warmup_s": round(t["warmup_s"], 2),
                "restarts_from_initial_state": t["restarts"],
                "setup_s": round(t["setup_s"], 1), "device_bytes": int(info.device_bytes),
                "profiled_ms_per_step": t["prof_total_ms"] / nprof}

    out = {
        "metric": "cell-updates/s FP64 vs FP32-perturbation at 1e-5 rel err; HBM GB/s; p99",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"{case.name}: {desc}; {nc} cells, {nf} faces",
                   "cells": nc, "faces": nf, "rk_stages": nstages,
                   "precision_paths": (["fp64"] if have64 else []) + (["fp32p"] if have32 else [])},
        "fp64": dict(path_rec(t64, rate64, ms64, info64), roofline=rl64) if have64 else None,
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    if have32:
        out["fp32p"] = dict(path_rec(t32, rate32, ms32, info32), speedup_vs_fp64=rate32 / rate64 if have64 else None,
                            error_vs_fp64=err, error_steps=n_err if err else None)
    if args.secondary != "none" and args.secondary != args.workload and not args.fp64_only:
        out["secondary"] = secondary_run(args.secondary, args.solver)
    if have32 and args.serve_requests > 0:
        out["serving"] = serving_run(args.serve_requests, args.serve_steps)
    print(json.dumps(out), flush=True)


def secondary_run(workload, solver, steps=200):
    """A second configuration timed briefly (FP32P and FP64 ms per step)."""
    from lfm_public_amd import capi
    case, desc = build_case(workload, solver)
    log(f"secondary {case.name}: {case.n_cells} cells")
    rec = {"workload": f"{case.name}: {desc}", "cells": case.n_cells, "steps": steps}
    for name, prec in (("fp32p", capi.PREC_FP32P), ("fp64", capi.PREC_FP64)):
        g, info, t, _ = time_path(case, prec, steps, 20, 1.0, 1, limit=OWN_STEPS[workload])
        g.close()
        ms = t["wall_ms"] / steps
        rec[name] = {"ms_per_step": ms, "cell_updates_per_s": case.n_cells / (ms * 1e-3)}
    rec["speedup_fp32p_vs_fp64"] = rec["fp64"]["ms_per_step"] / rec["fp32p"]["ms_per_step"]
    log(f"secondary: FP32P {rec['fp32p']['ms_per_step']:.3f} ms/step, FP64 {rec['fp64']['ms_per_step']:.3f}")
    return rec


def serving_run(n_req, nsteps, budget_gib=2.0):
    """A short serving stream after the timed region: seeded requests of
    2D vortex / cylinder cases, 16k-1M cells, Poisson arrivals offering 80 %
    of 1.5e9 cell-updates/s, FP32P, admitted under an HBM budget and stepped
    in shared graph-replayed launches; time-to-solution p50 / p99."""
    from lfm_public_amd import serving
    log(f"serving: {n_req} requests x {nsteps} steps")
    reqs = serving.make_stream(n_req, nsteps=nsteps)
    srv = serving.Server(int(budget_gib * (1 << 30)))
    try:
        res = srv.run(reqs)
    finally:
        srv.close()
    log(f"serving: p50 {res['tts_p50_s']:.3f} s, p99 {res['tts_p99_s']:.3f} s, "
        f"{res['cell_updates_per_s']:.3e} cell-updates/s")
    return {"requests": n_req, "steps_per_request": nsteps, "hbm_budget_gib": budget_gib,
            "tts_p50_s": res["tts_p50_s"], "tts_p99_s": res["tts_p99_s"], "tts_max_s": res["tts_max_s"],
            "cell_updates_per_s": res["cell_updates_per_s"], "wall_s": res["wall_s"],
            "max_batch_members": res["max_batch_members"], "all_finite": res["all_finite"],
            "arrivals": "Poisson at 80 % of 1.5e9 cell-updates/s"}


if __name__ == "__main__":
    main()
