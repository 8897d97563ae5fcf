"""lfm_solve on the MI355X: run an OpenFOAM case directory.

Mirrors info/lfm_solve.cpp + Mesh<P,D>::solve (src/mesh_solver.cpp:173-853)
for what the time step needs: read the case (undecomposed, or every
processor<N>/ of a decomposed one), create the resident solver(s), advance
from startTime to endTime, print step / time / deltaT / CFL (and the
residuals when lfm/post/haveResiduals is on) every printInfoFreq steps, and
write p, U, T, rho, E into time directories at the writeInterval.  With
lfm/post/haveAverage the running pAvg / pRMS fields are written beside them;
with lfm/post/haveForces the wall forces recorded every saveForcesStep steps
are appended to output/postProcess_Forces_<k> (one file per wall patch,
mesh_solver.cpp:333-358, 1081-1123; cfd_v0.cpp:3255-3280).

    python -m lfm_public_amd.solve CASE [--steps N] [--precision 64|32]
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

from . import capi, foam


def _steps_between(start, end, dt):
    return max(0, int(round((end - start) / dt)))


_XYZ = "xyz"


def _forces_header(path: str, dim: int):
    """output_Forces(firstRun), mesh_solver.cpp:1104-1117."""
    cols = f"{'time':>25}" + "".join(f"{'Fp,':>25}{_XYZ[k]}" for k in range(dim))
    cols += "".join(f"{'Fv,':>25}{_XYZ[k]}" for k in range(dim))
    cols += "".join(f"{'Fp,':>25}{_XYZ[k]} + Fv,{_XYZ[k]}" for k in range(dim))
    with open(path, "w") as f:
        f.write(cols + "\n")


def _forces_append(path: str, times, F):
    """postProcForcesOutput, cfd_v0.cpp:3255-3272: time, Fp, Fv, Fp + Fv."""
    with open(path, "a") as f:
        for t, (fp, fv) in zip(times, F):
            vals = [t, *fp, *fv, *(fp + fv)]
            f.write("".join(f"{v:25.16e}" for v in vals) + "\n")


def run_case(case_dir: str, steps=None, precision=None, quiet=False, write=True) -> dict:
    ranks = foam.ranks_of(case_dir)
    cases = [foam.FoamCase(case_dir, r) for r in ranks] if ranks else [foam.FoamCase(case_dir)]
    c0 = cases[0]
    prm = c0.params
    dim = prm["dim"]
    solvers = [c.solver(precision) for c in cases]
    group = None
    if ranks:
        capi.connect(solvers)
        group = capi.PartGroup(solvers)
    info_freq = int(c0.dict_scalar("system/controlDict", "lfm/post", "printInfoFreq", 1)) or 1
    post = c0.post
    forces_files = []
    if post.have_average or post.have_forces:
        for s in solvers:
            s.post_configure(cfg=post)
    if post.have_forces:
        walls = solvers[0].walls()
        out_dir = os.path.join(case_dir, "output")
        os.makedirs(os.path.join(out_dir, "Times"), exist_ok=True)
        forces_files = [os.path.join(out_dir, f"postProcess_Forces_{k + 1}") for k in range(len(walls))]
        if write and forces_files and not os.path.exists(forces_files[0]):
            for fn in forces_files:
                _forces_header(fn, dim)
        for s in solvers:  # the start-up postProcForces (mesh_solver.cpp:427-428)
            s.post_forces(c0.start)

    def flush_forces():
        rows = [s.forces_read() for s in solvers]
        times = rows[0][0]
        F = sum(r[1] for r in rows)  # MPI_Allreduce SUM over ranks
        if write:
            for k, fn in enumerate(forces_files):
                _forces_append(fn, times, F[:, k])
        return times, F
    forces_log = []
    n_total = steps if steps is not None else _steps_between(c0.start, c0.end, c0.delta_t)
    adjust = bool(prm["adjust_dt"])
    t = c0.start
    written = []
    next_write = c0.write_interval if c0.write_control == 0 else c0.start + c0.write_interval
    t0 = time.perf_counter()
    done = 0
    log = (lambda s: None) if quiet else (lambda s: print(s, flush=True))
    log(f"lfm_solve: {case_dir}: {sum(c.mesh_arrays['n_cells'] for c in cases)} cells, "
        f"{len(cases)} partition(s), {n_total} steps, precision {solvers[0].info().precision}")
    while done < n_total:
        # chunk to the next print / write point (one step at a time when dt adapts)
        if adjust:
            chunk = 1
        else:
            chunk = min(n_total - done, info_freq - done % info_freq)
            if c0.write_control == 0 and write:
                chunk = min(chunk, int(next_write) - done)
        chunk = max(1, chunk)
        dt_used = solvers[0].info().dt
        if group is not None:
            group.run(chunk)
            group.synchronize()
        else:
            solvers[0].run(chunk)
            solvers[0].synchronize()
        done += chunk
        t += chunk * dt_used
        if done % info_freq == 0 or done == n_total:
            infos = [s.info() for s in solvers]
            cfl = max(i.cfl for i in infos)
            line = f"step {done}  time {t:.9g}  deltaT {infos[0].dt:.6g}"
            line += f"  CFL {cfl:.6g}" if not adjust else ""
            if prm["have_res"]:
                res = np.sum([s.residual() for s in solvers], axis=0)
                line += "  RES " + " ".join(f"{math.sqrt(max(r, 0.0)):.6e}" for r in res)
            log(line)
        due = (c0.write_control == 0 and done >= next_write) or (c0.write_control == 1 and t >= next_write - 1e-12)
        if write and (due or done == n_total):
            for c, s in zip(cases, solvers):
                c.write_time(t, s.q(), dim)
                if post.have_average:
                    n = done - post.average_start
                    pavg, prms = s.average_fields(n) if n > 0 else (np.zeros(s.n_cells), np.zeros(s.n_cells))
                    c.write_scalar(t, "pAvg", pavg, "[1 -1 -2 0 0 0 0]")
                    c.write_scalar(t, "pRMS", prms, "[1 -1 -2 0 0 0 0]")
            if post.have_forces:
                forces_log.append(flush_forces())
            written.append(t)
            if due:
                next_write += c0.write_interval
    wall = time.perf_counter() - t0
    ncells = sum(c.mesh_arrays["n_cells"] for c in cases)
    if post.have_forces:
        forces_log.append(flush_forces())
    out = dict(steps=done, time=t, wall_s=wall, cell_updates_per_s=ncells * done / max(wall, 1e-12),
               written=written, q=[s.q() for s in solvers])
    if post.have_forces:
        out["forces_time"] = np.concatenate([x[0] for x in forces_log])
        out["forces"] = np.concatenate([x[1] for x in forces_log])
    if group is not None:
        group.close()
    for s in solvers:
        s.close()
    for c in cases:
        c.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="run an OpenFOAM case with the MI355X time step")
    ap.add_argument("case")
    ap.add_argument("--steps", type=int, default=None, help="steps to run (default: startTime to endTime)")
    ap.add_argument("--precision", type=int, default=None, choices=(32, 64),
                    help="override lfm/doublePrecision: 64 FP64, 32 FP32 deviations")
    ap.add_argument("--no-write", action="store_true")
    a = ap.parse_args(argv)
    r = run_case(a.case, a.steps, a.precision, write=not a.no_write)
    print(f"lfm_solve: {r['steps']} steps in {r['wall_s']:.2f} s ({r['cell_updates_per_s']:.3e} cell-updates/s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
