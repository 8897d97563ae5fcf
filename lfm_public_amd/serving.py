"""Continuous-batching service for many independent solves on one GPU.

Requests (independent 2D cases of uneven size) arrive over time.  The loop

* admits arrived requests first-come-first-served while the resident device
  memory stays under an HBM budget (a request waits while it does not fit);
* builds each admitted request (mesh, host set-up, upload) on a worker
  thread while the GPU keeps stepping the resident ones;
* keeps every admitted state resident and steps all residents of one scheme
  class with shared launches (lfm_batch: one kernel launch per stage piece
  covers all of them, one RK step captured as a hipGraph);
* re-forms the batches at step boundaries, advancing in chunks that end
  exactly when the next resident finishes, so no request over-runs its step
  count and finished ones free their memory at once;
* reports p50/p99 time-to-solution (arrival -> result on the host) and the
  aggregate cell-updates/s.

Each request's result is independent of the mix: the batched kernels run the
same per-cell arithmetic as a lone solver (checked by
tests/test_gpu_serving.py).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import dataclasses
import json
import math
import threading
import time
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from . import capi, cases, meshgen


@dataclasses.dataclass
class Request:
    rid: int
    kind: str            # "vortex" | "cylinder"
    n: int               # mesh size parameter (cells ~ n^2)
    mach: float
    seed: int
    nsteps: int
    arrival: float       # seconds after the start of the run
    strength: float = 5.0

    @property
    def cls(self) -> Tuple[int, int, int]:
        # (dim, solver, minmod): vortices run M1, cylinders M2 + minmod AUSM
        return (2, 0, 0) if self.kind == "vortex" else (2, 2, 1)

    @property
    def n_cells(self) -> int:
        return self.n * self.n


def est_device_bytes(n_cells: int, kind: Optional[str] = None) -> int:
    """Resident FP32-perturbation footprint of a 2D solve (measured ~330 B/cell);
    a vortex request (inviscid M1) also holds, once batched, the one-launch
    stages' second state buffer and face slots (~80 B/cell).  Without a kind,
    the larger."""
    per = 360 if kind == "cylinder" else 440
    return int(per * n_cells + (1 << 20))


_mesh_cache: Dict[Tuple[str, int], object] = {}
_mesh_lock = threading.Lock()


def build_case(req: Request) -> cases.Case:
    """The request's case: a Shu vortex on an n x n box or a cylinder annulus
    with n azimuthal cells (n^2 cells, 1/4 quad rings, 3/8 triangle rings)."""
    key = (req.kind, req.n)
    with _mesh_lock:
        mesh = _mesh_cache.get(key)
    if req.kind == "vortex":
        c = cases.isentropic_vortex(nx=req.n, mach=req.mach, seed=req.seed, strength=req.strength, solver=0,
                                    mesh=mesh)
        # the request convects its vortex ~10 units over its steps (it starts
        # 6-8 units upstream of the centre and stays inside the box), at a CFL
        # of 0.5 on the initial field's fastest wave, whichever dt is smaller
        g = cases.gamma_of(c.params)
        R = cases.gas_constant(c.params["mol_weight"])
        wave = float((np.linalg.norm(c.U, axis=1) + np.sqrt(g * R * c.T)).max())
        dx = 20.0 / req.n
        U_inf = c.params["U_inf"][0]
        prm = dict(c.params)
        prm["dt"] = min(0.5 * dx / wave, 10.0 / (req.nsteps * U_inf))
        c = cases.Case(c.name, c.mesh, prm, c.p, c.T, c.U, c.alpha)
    else:
        # one mesh per size (seed 202): requests of one size differ in flow, not
        # in mesh; dt for twice the free-stream speed (the shoulders' speed)
        c = cases.cylinder_2d(n_theta=req.n, n_quad=req.n // 4, n_tri=3 * req.n // 8, mach=req.mach, seed=req.seed,
                              solver=2, minmod=1, mesh=mesh, speed_factor=2.0)
    with _mesh_lock:
        _mesh_cache.setdefault(key, c.mesh)
    return c


def make_stream(n_req=256, seed=505, cells_lo=16_384, cells_hi=1_048_576, nsteps=2000, rate=None,
                load=0.8, cell_rate_est=1.5e9) -> List[Request]:
    """Seeded request stream: 50/50 vortex/cylinder, cell counts log-uniform
    in [cells_lo, cells_hi] (n rounded to a multiple of 16), Mach uniform in
    [0.1, 0.7], random vortex strengths, Poisson arrivals.  Without an
    explicit rate, arrivals offer ``load`` of the throughput ``cell_rate_est``."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi), n_req))
    n = np.maximum(16, (np.sqrt(cells) / 16).round().astype(int) * 16)
    kinds = np.where(rng.uniform(size=n_req) < 0.5, "vortex", "cylinder")
    mach = rng.uniform(0.1, 0.7, n_req)
    strength = rng.uniform(1.0, 5.0, n_req)
    work = float((n.astype(float) ** 2).sum() * nsteps)
    if rate is None:
        rate = n_req / (work / (load * cell_rate_est))
    gaps = rng.exponential(1.0 / rate, n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    return [Request(i, str(kinds[i]), int(n[i]), float(mach[i]), int(seed * 1000 + i), nsteps, float(arrivals[i]),
                    float(strength[i])) for i in range(n_req)]


@dataclasses.dataclass
class Resident:
    req: Request
    solver: capi.GpuSolver
    remaining: int
    nbytes: int
    admitted: float
    started: float


class Server:
    def __init__(self, hbm_budget: int, chunk_max: int = 64, builders: int = 4, keep_results: bool = False,
                 time_fn: Callable[[], float] = time.perf_counter):
        self.budget = hbm_budget
        self.chunk_max = chunk_max
        self.pool = cf.ThreadPoolExecutor(max_workers=builders)
        self.keep_results = keep_results
        self.now = time_fn
        self.batches: Dict[Tuple[int, int, int], capi.Batch] = {}
        self.results: Dict[int, np.ndarray] = {}

    def _build(self, req: Request):
        c = build_case(req)
        return capi.GpuSolver.from_case(c, precision=capi.PREC_FP32P)

    def run(self, requests: List[Request], log: Optional[Callable[[str], None]] = None) -> dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: List[Request] = []
        building: Dict[cf.Future, Tuple[Request, int, float]] = {}
        active: Dict[Tuple[int, int, int], List[Resident]] = {}
        dirty = set()
        reserved = 0
        done: Dict[int, dict] = {}
        t0 = self.now()
        busy_cells_steps = 0
        peak_resident = 0
        max_members = 0
        while len(done) < len(requests):
            now = self.now() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admission: FIFO under the HBM budget (head-of-line waits)
            while waiting:
                r = waiting[0]
                est = est_device_bytes(r.n_cells, r.kind)
                if reserved + est > self.budget and (reserved > 0 or building):
                    break
                waiting.pop(0)
                reserved += est
                building[self.pool.submit(self._build, r)] = (r, est, now)
            # collect finished builds
            for fut in [f for f in building if f.done()]:
                r, est, t_adm = building.pop(fut)
                s = fut.result()
                # the estimate stays reserved until the batch has allocated
                # what it needs (the refresh below), unless the solve already
                # holds more
                nbytes = max(s.info().device_bytes, est)
                reserved += nbytes - est
                active.setdefault(r.cls, []).append(Resident(r, s, r.nsteps, nbytes, t_adm, self.now() - t0))
                dirty.add(r.cls)
            peak_resident = max(peak_resident, reserved)
            for cls in list(dirty):
                members = active.get(cls, [])
                b = self.batches.get(cls)
                if b is None:
                    b = self.batches[cls] = capi.Batch()
                b.set([m.solver for m in members])
                max_members = max(max_members, len(members))
                # joining a batch may allocate the one-launch stages' buffers
                for m in members:
                    nb = m.solver.info().device_bytes
                    reserved += nb - m.nbytes
                    m.nbytes = nb
            peak_resident = max(peak_resident, reserved)
            dirty.clear()
            residents = [m for ms in active.values() for m in ms]
            if not residents:
                # idle: wait for the next arrival or build
                if building:
                    cf.wait(list(building), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                elif pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - (self.now() - t0))))
                continue
            chunk = min(self.chunk_max, min(m.remaining for m in residents))
            for cls, ms in active.items():
                if ms:
                    self.batches[cls].run(chunk)
            for cls, ms in active.items():
                if ms:
                    self.batches[cls].synchronize()
            busy_cells_steps += chunk * sum(m.req.n_cells for m in residents)
            tnow = self.now() - t0
            for cls, ms in active.items():
                keep = []
                for m in ms:
                    m.remaining -= chunk
                    if m.remaining > 0:
                        keep.append(m)
                        continue
                    q = m.solver.q()
                    finite = bool(np.isfinite(q).all())
                    if self.keep_results:
                        self.results[m.req.rid] = q
                    m.solver.close()
                    reserved -= m.nbytes
                    done[m.req.rid] = dict(rid=m.req.rid, kind=m.req.kind, cells=m.req.n_cells, steps=m.req.nsteps,
                                           arrival=m.req.arrival, admitted=m.admitted, started=m.started,
                                           finished=self.now() - t0, finite=finite)
                    if log:
                        log(f"[serve] t={tnow:7.2f}s done req {m.req.rid} ({m.req.kind}, {m.req.n_cells} cells) "
                            f"tts={done[m.req.rid]['finished'] - m.req.arrival:.2f}s residents={len(residents)}")
                if len(keep) != len(ms):
                    dirty.add(cls)
                active[cls] = keep
        wall = self.now() - t0
        tts = np.array([d["finished"] - d["arrival"] for d in done.values()])
        total = sum(d["cells"] * d["steps"] for d in done.values())
        return dict(n_requests=len(done), wall_s=wall, cell_updates=total, cell_updates_per_s=total / wall,
                    tts_p50_s=float(np.percentile(tts, 50)), tts_p99_s=float(np.percentile(tts, 99)),
                    tts_max_s=float(tts.max()), peak_resident_bytes=int(peak_resident), hbm_budget=self.budget,
                    max_batch_members=max_members, all_finite=all(d["finite"] for d in done.values()),
                    n_nonfinite=sum(not d["finite"] for d in done.values()),
                    requests=sorted(done.values(), key=lambda d: d["rid"]))

    def close(self):
        for b in self.batches.values():
            b.close()
        self.pool.shutdown(wait=True)


def main():
    ap = argparse.ArgumentParser(description="continuous-batching serving run")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--cells-hi", type=int, default=1_048_576)
    ap.add_argument("--budget-gib", type=float, default=2.0)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--rate-est", type=float, default=1.5e9)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    reqs = make_stream(args.requests, nsteps=args.steps, cells_hi=args.cells_hi, load=args.load,
                       cell_rate_est=args.rate_est)
    srv = Server(int(args.budget_gib * (1 << 30)))
    res = srv.run(reqs, log=lambda s: print(s, flush=True))
    srv.close()
    summary = {k: v for k, v in res.items() if k != "requests"}
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
