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
    assert res["max_batch_members"] >= 2
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
