"""Parity at the benchmark workloads' full sizes (needs an MI355X).

FP64 against the oracle on the same inputs, at sizes the oracle steps in
seconds: the 256^2 isentropic vortex (65,536 cells, M1, 40 RK5 steps), the
1,048,576-cell mixed quad/triangle annulus (M2AUSM + minmod, viscous,
sponge, 2 steps) and the 4,194,304-cell 3D cylinder O-grid (M2 +
minmodExists, viscous, spanwise cyclic, 1 step).  Bar: relative L2 <= 1e-10
on p, U and T (the FP64 kernels are expected bit-identical).

FP32P against FP64 on the full-size annulus after 200 steps.  Bar:
perturbation-normalised relative L2 <= 1e-5 on p, U and T.
"""
import numpy as np
import pytest

from conftest import pert_l2, primitives, rel_l2
from lfm_public_amd import capi, cases

pytestmark = pytest.mark.gpu


def _workload(name):
    if name == "vortex256":
        return cases.isentropic_vortex(nx=256), 40
    if name == "annulus1M":
        n = 1024
        return cases.cylinder_2d(n_theta=n, n_quad=n // 4, n_tri=3 * n // 8, solver=2), 2
    return cases.cylinder_3d(nr=256, nth=256, nz=64, solver=1, minmod=1), 1


def _prims(case, q):
    g, R = cases.gamma_of(case.params), cases.gas_constant(case.params["mol_weight"])
    return primitives(q, case.params["dim"], g, R)


@pytest.mark.parametrize("name", ["vortex256", "annulus1M", "ogrid4M"])
def test_fp64_full_size_matches_oracle(name, oracle_lib):
    import oracle as O
    case, nsteps = _workload(name)
    print(f"[fullsize] {name}: {case.n_cells} cells, {nsteps} steps", flush=True)
    g = capi.GpuSolver.from_case(case)
    o = O.Oracle(case.mesh, case.params, case.p, case.T, case.U, case.alpha)
    g.run(nsteps)
    o.run(nsteps)
    qg, qo = g.q(), o.q()
    g.close()
    assert np.all(np.isfinite(qg))
    pg, Ug, Tg = _prims(case, qg)
    po, Uo, To = _prims(case, qo)
    errs = {"p": rel_l2(pg, po), "T": rel_l2(Tg, To), "U": rel_l2(Ug, Uo)}
    print(name, {k: f"{v:.2e}" for k, v in errs.items()}, "bit-identical:", bool(np.array_equal(qg, qo)),
          flush=True)
    for k, v in errs.items():
        assert v <= 1e-10, (k, v)


def test_fp32p_full_size_annulus_within_bar():
    case, _ = _workload("annulus1M")
    print(f"[fullsize] FP32P vs FP64, {case.n_cells} cells, 200 steps", flush=True)
    g64 = capi.GpuSolver.from_case(case, precision=capi.PREC_FP64)
    g32 = capi.GpuSolver.from_case(case, precision=capi.PREC_FP32P)
    g64.run(200)
    g32.run(200)
    q64, q32 = g64.q(), g32.q()
    g64.close()
    g32.close()
    p6, U6, T6 = _prims(case, q64)
    p3, U3, T3 = _prims(case, q32)
    prm = case.params
    errs = {"p": pert_l2(p3, p6, prm["p_inf"]), "T": pert_l2(T3, T6, prm["T_inf"])}
    for k in range(2):
        errs[f"U{k}"] = pert_l2(U3[k], U6[k], prm["U_inf"][k])
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-5, (k, v)
