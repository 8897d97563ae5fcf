"""Continuous-batching service (needs an MI355X).

Each request's result must be the same alone and in the mix: the batched
launches run the same per-cell arithmetic as a lone solver, so the FP32
states are compared bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(req):
    from lfm_public_amd import capi, serving
    s = capi.GpuSolver.from_case(serving.build_case(req), precision=capi.PREC_FP32P)
    s.run(req.nsteps)
    q = s.q()
    s.close()
    return q


def test_batch_members_match_lone_runs():
    """Shared launches: solvers of two scheme classes and uneven sizes, stepped
    together by lfm_batch with members joining and leaving between chunks,
    end bit-identical to the same solvers stepped alone."""
    from lfm_public_amd import capi, serving
    reqs = [serving.Request(0, "vortex", 32, 0.3, 21, 60, 0.0, 2.0),
            serving.Request(1, "vortex", 48, 0.6, 22, 45, 0.0, 4.0),
            serving.Request(2, "vortex", 16, 0.2, 23, 30, 0.0, 1.0),
            serving.Request(3, "cylinder", 32, 0.25, 24, 50, 0.0),
            serving.Request(4, "cylinder", 48, 0.5, 25, 35, 0.0)]
    solvers = {r.rid: capi.GpuSolver.from_case(serving.build_case(r), precision=capi.PREC_FP32P) for r in reqs}
    left = {r.rid: r.nsteps for r in reqs}
    batches = {}
    # r2 and r4 join after 10 steps; everyone leaves when done
    joined = {0, 1, 3}
    t = 0
    most = 0
    while any(left.values()):
        if t >= 10:
            joined |= {2, 4}
        by_cls = {}
        for r in reqs:
            if r.rid in joined and left[r.rid] > 0:
                by_cls.setdefault(r.cls, []).append(r.rid)
        chunk = min(min(left[i] for ids in by_cls.values() for i in ids), 10 - t if t < 10 else 1 << 30)
        for cls, ids in by_cls.items():
            b = batches.setdefault(cls, capi.Batch())
            b.set([solvers[i] for i in ids])
            b.run(chunk)
        for b in batches.values():
            b.synchronize()
        for ids in by_cls.values():
            for i in ids:
                left[i] -= chunk
        t += chunk
        most = max(most, sum(len(ids) for ids in by_cls.values()))
    assert most == 5
    for r in reqs:
        qa = _alone(r)
        qm = solvers[r.rid].q()
        assert np.array_equal(qa, qm), (r.rid, np.abs(qa - qm).max())
    for b in batches.values():
        b.close()


def test_results_same_alone_and_in_the_mix():
    from lfm_public_amd import serving
    reqs = [serving.Request(0, "vortex", 32, 0.3, 11, 40, 0.0, 2.0),
            serving.Request(1, "cylinder", 32, 0.25, 12, 25, 0.0),
            serving.Request(2, "vortex", 48, 0.6, 13, 30, 0.05, 4.0),
            serving.Request(3, "cylinder", 48, 0.5, 14, 35, 0.10),
            serving.Request(4, "vortex", 16, 0.1, 15, 10, 0.12, 1.0)]
    print("[serving] mix of 5 requests", flush=True)
    srv = serving.Server(hbm_budget=1 << 30, chunk_max=7, keep_results=True)
    res = srv.run(reqs)
    srv.close()
    assert res["n_requests"] == len(reqs) and res["all_finite"]
    for r in reqs:
        print(f"[serving] alone {r.rid}", flush=True)
        qa = _alone(r)
        qm = srv.results[r.rid]
        assert np.array_equal(qa, qm), (r.rid, np.abs(qa - qm).max())


def test_budget_forces_queueing_and_all_complete():
    from lfm_public_amd import serving
    reqs = serving.make_stream(n_req=8, seed=7, cells_lo=1024, cells_hi=16384, nsteps=20, rate=1e6)
    budget = serving.est_device_bytes(16384) + 1
    srv = serving.Server(hbm_budget=budget, chunk_max=5)
    res = srv.run(reqs)
    srv.close()
    assert res["n_requests"] == 8 and res["all_finite"]
    assert res["peak_resident_bytes"] <= budget + (1 << 22)
    assert res["tts_p99_s"] >= res["tts_p50_s"] > 0


def test_serving_stream_at_size():
    """The serving configuration's own stream (seed 505: 50/50 vortex and
    cylinder requests, 16 k - 1 M cells log-uniform, Mach 0.1-0.7) at its
    size range: its first 24 requests, 40 RK5 steps each, under a 2 GiB HBM
    budget (the requests need ~2.8 GB together, so admission queues).  The largest and
    smallest member of each scheme class end bit-identical to their lone
    runs, and the resident device bytes never exceed the budget."""
    from lfm_public_amd import serving
    reqs = serving.make_stream(256, seed=505, nsteps=40)[:24]
    cells = [r.n_cells for r in reqs]
    assert min(cells) <= 20_000 and max(cells) >= 1_000_000
    budget = 2 << 30
    assert sum(serving.est_device_bytes(c) for c in cells) > budget
    print(f"[serving] stream of seed 505: {len(reqs)} requests, {min(cells)}-{max(cells)} cells", flush=True)
    srv = serving.Server(hbm_budget=budget, chunk_max=16, keep_results=True)
    res = srv.run(reqs)
    srv.close()
    assert res["n_requests"] == len(reqs)
    assert res["peak_resident_bytes"] <= budget, res["peak_resident_bytes"]
    # how many requests of a class share launches depends on how build times
    # and steps overlap on the box; the shared launches at size are checked
    # deterministically below
    print(f"[serving] at most {res['max_batch_members']} members per batch", flush=True)
    picks = []
    for kind in ("vortex", "cylinder"):
        ks = [r for r in reqs if r.kind == kind]
        picks += [max(ks, key=lambda r: r.n_cells), min(ks, key=lambda r: r.n_cells)]
    for r in picks:
        print(f"[serving] alone {r.rid} ({r.kind}, {r.n_cells} cells)", flush=True)
        qa = _alone(r)
        qm = srv.results[r.rid]
        assert np.all(np.isfinite(qa))
        assert np.array_equal(qa, qm), (r.rid, np.abs(qa - qm).max())


def test_serving_sizes_in_shared_launches():
    """The largest, smallest and a middle request of each scheme class of the
    serving stream (seed 505) stepped together in one batch per class (40
    RK5 steps, chunks of 7) end bit-identical to their lone runs."""
    from lfm_public_amd import capi, serving
    reqs = serving.make_stream(256, seed=505, nsteps=40)[:24]
    picks = []
    for kind in ("vortex", "cylinder"):
        ks = sorted((r for r in reqs if r.kind == kind), key=lambda r: r.n_cells)
        picks.append([ks[0], ks[len(ks) // 2], ks[-1]])
    for group in picks:
        print(f"[serving] batch of {[r.n_cells for r in group]} cells ({group[0].kind})", flush=True)
        solvers = [capi.GpuSolver.from_case(serving.build_case(r), precision=capi.PREC_FP32P) for r in group]
        b = capi.Batch()
        b.set(solvers)
        left = group[0].nsteps
        while left > 0:
            n = min(7, left)
            b.run(n)
            left -= n
        b.synchronize()
        b.close()
        for r, s in zip(group, solvers):
            qa = _alone(r)
            qm = s.q()
            s.close()
            assert np.all(np.isfinite(qm))
            assert np.array_equal(qa, qm), (r.rid, np.abs(qa - qm).max())
