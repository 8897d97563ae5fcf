#!/usr/bin/env python3
"""Benchmark of the MI355X time step on the headline workload.

Workload (default): 2D unstructured annulus around a unit cylinder,
1,048,576 cells (256 rings of quads, 384 rings of triangles, 10 % seeded
node jitter), uniform inflow M = 0.3 with 1e-4 pressure noise, M2 scheme
with minmod AUSM pressure dissipation (solver 2), viscous terms, sponge of
5 diameters from the outlet, RK5, constant dt.  Synthetic mesh and fields.

A "step" is one full time step (5 RK stages) of every cell; the metric is
cell-updates/s = cells x steps / s with the state resident in HBM.  The
FP32-perturbation path is the headline value when it is built; the FP64
path is timed beside it and its error is reported.

Prints ONE JSON line (rank 0, one GPU).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(nc, nf, D, w_s, w_g):
    """Algorithmic bytes of one RK stage: q and dq read + written, 1/V and
    sigma per cell; S, d, w and the two cell indices per face."""
    return nc * (4 * (D + 2) * w_s + 2 * w_g) + nf * ((2 * D + 1) * w_g + 8)


def build_case(args):
    """The workload and a one-line description of it."""
    from lfm_public_amd import cases
    w = args.workload
    if w == "cylinder2d":
        n = args.n_theta
        c = cases.cylinder_2d(n_theta=n, n_quad=n // 4, n_tri=3 * n // 8, solver=args.solver)
        desc = (f"2D mixed quad/tri annulus around a unit cylinder (10 % node jitter), M=0.3, 1e-4 p noise, "
                f"solver {args.solver} ({'M2 + minmod AUSM dissipation' if args.solver == 2 else 'M2, minmodExists'})"
                f", viscous, sponge Ls=5D, RK5")
    elif w == "vortex2d":
        c = cases.isentropic_vortex(nx=256)
        desc = "2D 256x256 quad box, isentropic vortex (strength 5) at M=0.5, M1, inviscid, RK5, CFL 0.5"
    elif w == "ogrid3d":
        c = cases.cylinder_3d(nr=256, nth=256, nz=64, solver=1, minmod=1)
        desc = ("3D cylinder O-grid 256x256x64 hexes, spanwise cyclic, M=0.2, mu from Re=3900, 1 % spanwise "
                "perturbation, M2 + minmodExists, sponge Ls=13, RK5")
    elif w == "box3d":
        c = cases.acoustic_box(n=256)
        desc = "3D 256^3 hex box, Gaussian pulse 1e-3 p_inf (half-width 8 cells) in M=0.1 flow, M1, inviscid, RK5"
    else:
        raise SystemExit(f"unknown workload {w}")
    if args.renumber != "none":
        from lfm_public_amd import renumber
        c = renumber.renumber_case(c, args.renumber)
        desc += f"; cells renumbered ({args.renumber})"
    return c, desc


def cpu_sample_case(args, case):
    """The CPU baseline's sample: the workload itself up to ~1.5M cells, else
    the same case family at a size one CPU solve finishes steps of in seconds."""
    from lfm_public_amd import cases
    if case.n_cells <= 1_500_000:
        return case, "the full mesh"
    if args.workload == "ogrid3d":
        c = cases.cylinder_3d(nr=64, nth=64, nz=32, solver=1, minmod=1)
    else:
        c = cases.acoustic_box(n=64)
    return c, f"the same case family at {c.n_cells} cells ({c.name})"


def time_path(case, precision, steps, warmup):
    from lfm_public_amd import capi
    g = capi.GpuSolver.from_case(case, precision=precision)
    info = g.info()
    g.run(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ms = g.time_steps(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    prof, prof_total = g.profile_steps(max(1, min(steps, 10)))
    return g, info, ms, wall, prof, prof_total


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
    thread per solve) on the same mesh, run as the reference runs on a CPU
    node: one independent single-rank solve per core, all cores busy, a
    bounded number of steps; the aggregate rate over the cores."""
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
    return {"value": threads * case.n_cells * n / wall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference time step, FP64) on {what} ({case.n_cells} cells): "
                      f"{threads} independent single-thread solves on {threads} cores, {n} RK5 steps each after 1 "
                      f"warm-up step, {wall:.1f} s; aggregate cell-updates/s (the reference itself needs MPI and "
                      f"OpenFOAM, absent here)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configuration's own step count)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cylinder2d")
    ap.add_argument("--n-theta", type=int, default=1024)
    ap.add_argument("--solver", type=int, default=2, help="cylinder2d scheme: 2 (M2AUSM + minmod) or 1 (M2)")
    ap.add_argument("--renumber", default="none", choices=("none", "morton", "rcm"),
                    help="renumber the cells before the run (renumberMesh-style preprocessing)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline cores (default: up to 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests of the serving stream run after the timed region (0: none)")
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"cylinder2d": 2000, "vortex2d": 1000, "ogrid3d": 500, "box3d": 1000}[args.workload]
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by spatial "
              f"decomposition but only the single-GPU partition is built and measured.", file=sys.stderr)
        sys.exit(2)

    from lfm_public_amd import capi
    case, desc = build_case(args)
    D = case.params["dim"]
    nc = case.n_cells
    print(f"[bench] workload {case.name}: {nc} cells", file=sys.stderr, flush=True)

    g64, info64, ms64, wall64, prof64, prof64_tot = time_path(case, capi.PREC_FP64, args.steps, args.warmup)
    nf = info64.n_faces
    rate64 = nc * args.steps / (ms64 * 1e-3)
    print(f"[bench] FP64: {ms64 / args.steps:.3f} ms/step, {rate64:.3e} cell-updates/s", file=sys.stderr,
          flush=True)
    q64 = g64.q()
    bad64 = g64.n_nonfinite()
    g64.close()
    if bad64:
        raise SystemExit(f"FP64 state went non-finite ({bad64} values): the workload's dt is unstable")

    have32 = False
    res32 = None
    if not args.fp64_only:
        try:
            g32, info32, ms32, wall32, prof32, prof32_tot = time_path(case, capi.PREC_FP32P, args.steps,
                                                                      args.warmup)
            have32 = True
        except capi.LfmError as e:
            print(f"[bench] FP32-perturbation path unavailable: {e}", file=sys.stderr)
    if have32:
        rate32 = nc * args.steps / (ms32 * 1e-3)
        # both paths advanced warmup + steps + profiled steps from the same state
        q32 = g32.q()
        err = rel_errors(case, q32, q64, D)
        g32.close()
        print(f"[bench] FP32P: {ms32 / args.steps:.3f} ms/step, {rate32:.3e} cell-updates/s", file=sys.stderr,
              flush=True)

    # Roofline of the time step.  The unit the algorithmic byte count is
    # defined on is one RK stage (q, dq read and written; 1/V, sigma; S, d, w
    # and two indices per face); a step is 5 of them, launched as
    # 5 x (grad_vis, face_flux, cell_update), the last update also forming
    # the CFL/dt block extrema, + the final CFL/dt reduction.  The
    # achieved rate divides the step's algorithmic bytes by the timed ms/step
    # (inputs resident in HBM).  HBM traffic per step comes from the
    # rocprofv3 FETCH_SIZE/WRITE_SIZE passes committed under profiles/ for
    # this workload (calibrated on a plain copy), when they exist.
    nprof = max(1, min(args.steps, 10))
    traffic_key = case.name + (f"_solver{args.solver}" if args.workload == "cylinder2d" and args.solver != 2 else "")

    def traffic_of(path):
        f = os.path.join(ROOT, "profiles", f"{traffic_key}_traffic.json")
        try:
            with open(f) as fh:
                t = json.load(fh)
        except (OSError, ValueError):
            return None, None
        if t.get("workload") != traffic_key:
            return None, None
        st = t.get("per_step", {}).get(path)
        if not st:
            return None, None
        src = (f"profiles/{traffic_key}_traffic.json: rocprofv3 FETCH_SIZE x{t['fetch_factor']:.3f} + "
               f"WRITE_SIZE x{t['write_factor']:.3f} (factors from a 256 MiB device copy), summed over "
               + ", ".join(f"{n} x {k}" for k, n in st["kernels"].items()))
        return int(st["total"]), src

    def kernel_desc(path):
        prm = case.params
        if path == "fp32p" and prm["mu"] == 0.0 and prm.get("laminar", 1) and prm["solver"] < 2:
            return ("time step: stage 1 as face_flux + cell_update, stages 2-5 as one stage_fused launch each "
                    "(the face fluxes formed in the update's ordered gather, the state double-buffered); the "
                    "updates refresh the boundary ghosts and the last one forms the CFL/dt block extrema, then "
                    "cfl_dt_final")
        return ("time step: 5 x (grad_vis, face_flux, cell_update); the updates refresh the boundary "
                "ghosts and the last one forms the CFL/dt block extrema, then cfl_dt_final")

    def roof(ms_step, prof, w, path):
        B = stage_bytes(nc, nf, D, w, w)
        B_step = 5 * B
        ach = B_step / (ms_step * 1e-3) / 1e9
        traffic, src = traffic_of(path)
        per = {k: round(v / (5 * nprof) * 1e3, 2) for k, v in prof.items() if k not in ("cfl_dt", "bc_fill")}
        out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
               "kernel": kernel_desc(path),
               "algorithmic_bytes_per_stage": int(B), "algorithmic_bytes_per_step": int(B_step),
               "us_per_launch_events": per,
               "dominant_kernel": max(("grad_vis", "face_flux", "cell_update"), key=lambda k: prof[k])}
        if traffic:
            out["traffic_source"] = src
            out["traffic_GBps"] = round(traffic / (ms_step * 1e-3) / 1e9, 1)
            out["traffic_frac_of_peak"] = round(traffic / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        return out

    if have32:
        value, ms_step, dtype = rate32, ms32 / args.steps, "f32"
        rl = roof(ms_step, prof32, 4, "fp32p")
    else:
        value, ms_step, dtype = rate64, ms64 / args.steps, "f64"
        rl = roof(ms_step, prof64, 8, "fp64")

    threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
    cpu = None
    if not args.no_cpu:
        cc, what = cpu_sample_case(args, case)
        # one oracle instance holds ~0.7 kB (2D) / ~1.2 kB (3D) per cell
        per = cc.n_cells * (700 if D == 2 else 1200)
        try:
            avail = int([l for l in open("/proc/meminfo") if l.startswith("MemAvailable")][0].split()[1]) * 1024
            threads = max(1, min(threads, int(0.5 * avail // per)))
        except (OSError, IndexError, ValueError):
            pass
        cpu = cpu_baseline(cc, args.cpu_budget, threads, what)
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
                   "cells": nc, "faces": nf, "rk_stages": 5, "precision_paths": ["fp64"] + (["fp32p"] if have32 else [])},
        "fp64": {"cell_updates_per_s": rate64, "ms_per_step": ms64 / args.steps,
                 "roofline": roof(ms64 / args.steps, prof64, 8, "fp64")},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    if have32:
        out["fp32p"] = {"cell_updates_per_s": rate32, "ms_per_step": ms32 / args.steps,
                        "speedup_vs_fp64": rate32 / rate64, "error_vs_fp64": err,
                        "error_steps": args.warmup + args.steps + nprof}
    if have32 and args.serve_requests > 0:
        out["serving"] = serving_run(args.serve_requests, args.serve_steps)
    print(json.dumps(out))


def serving_run(n_req, nsteps, budget_gib=2.0):
    """A short serving stream after the timed region: seeded requests of
    2D vortex / cylinder cases, 16k-1M cells, Poisson arrivals offering 80 %
    of 1.5e9 cell-updates/s, FP32P, admitted under an HBM budget and stepped
    in shared graph-replayed launches; time-to-solution p50 / p99."""
    from lfm_public_amd import serving
    print(f"[bench] serving: {n_req} requests x {nsteps} steps", file=sys.stderr, flush=True)
    reqs = serving.make_stream(n_req, nsteps=nsteps)
    srv = serving.Server(int(budget_gib * (1 << 30)))
    try:
        res = srv.run(reqs)
    finally:
        srv.close()
    print(f"[bench] serving: p50 {res['tts_p50_s']:.3f} s, p99 {res['tts_p99_s']:.3f} s, "
          f"{res['cell_updates_per_s']:.3e} cell-updates/s", file=sys.stderr, flush=True)
    return {"requests": n_req, "steps_per_request": nsteps, "hbm_budget_gib": budget_gib,
            "tts_p50_s": res["tts_p50_s"], "tts_p99_s": res["tts_p99_s"], "tts_max_s": res["tts_max_s"],
            "cell_updates_per_s": res["cell_updates_per_s"], "wall_s": res["wall_s"],
            "max_batch_members": res["max_batch_members"], "all_finite": res["all_finite"],
            "arrivals": "Poisson at 80 % of 1.5e9 cell-updates/s"}


if __name__ == "__main__":
    main()
