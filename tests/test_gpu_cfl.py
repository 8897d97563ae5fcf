"""The end-of-step CFL / dt reduction (cfd_v0.cpp:2887-2942, called by
Mesh::solve after the RK stages, mesh_solver.cpp:705-736) as the last stage's
update forms it equals the stand-alone compute_cfl / compute_dt of the same
state bit for bit: both take the min / max over the same per-cell values,
and min / max are exact in any order.  Covers FP64 and FP32P, 2D and 3D,
the adjusted-dt mode, the one-stage scheme whose only stage also forms the
residual (the reduction then keeps its own launch), rkOrder 0 (no stage: the
state stays, the reduction runs), partition groups and batched solvers.
"""
import numpy as np
import pytest

import small_cases as sc
from lfm_public_amd import capi, cases, meshgen

pytestmark = pytest.mark.gpu

PRECS = {"fp64": capi.PREC_FP64, "fp32p": capi.PREC_FP32P}
CASES = {
    "vortex_M1_rk3_adjust_res": lambda: sc.PARITY_CASES["vortex_M1_rk3_adjust_res"](),
    "cylinder_M2AUSM_sponge": lambda: sc.PARITY_CASES["cylinder_M2AUSM_sponge"](),
    "ogrid_M2AUSM_cyclic": lambda: sc.PARITY_CASES["ogrid_M2AUSM_cyclic"](),
    "box_M1_adjust": lambda: sc.with_params(sc.PARITY_CASES["box_M1_visc"](), adjust_dt=1, max_co=0.3),
    "vortex_rk1_res_adjust": lambda: sc.vortex(16, solver=0, mu=0.02, rk_order=1, adjust_dt=1, max_co=0.2,
                                               have_res=1),
    # rkOrder 0: Mesh::solve runs no stage (mesh_solver.cpp:500); the state
    # stays and only the end-of-step reduction runs
    "vortex_rk0": lambda: sc.vortex(8, solver=0, rk_order=0),
    # 2D cells of 5 and 6 faces
    "poly_M2AUSM_visc_adjust": lambda: sc.PARITY_CASES["poly_M2AUSM_visc_adjust"](),
    # 3D cells of 5 faces
    "prism_M2_visc_adjust": lambda: sc.PARITY_CASES["prism_M2_visc_adjust"](),
}


def _check_end_of_step(g, c):
    info = g.info()
    if c.params.get("adjust_dt"):
        assert g.compute_dt(c.params["max_co"]) == info.dt
    else:
        assert info.cfl > 0
        assert g.compute_cfl(info.dt) == info.cfl


@pytest.mark.parametrize("prec", sorted(PRECS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_step_reduction_equals_standalone(name, prec):
    print(f"[cfl] {name} {prec}", flush=True)
    c = CASES[name]()
    g = capi.GpuSolver.from_case(c, precision=PRECS[prec])
    q0 = g.q()
    for n in (1, 3):
        g.run(n)
        g.synchronize()
        _check_end_of_step(g, c)
    if c.params["rk_order"] == 0:
        assert np.array_equal(g.q(), q0)
        assert g.info().steps_done == 4
    g.close()


def test_step_reduction_partition_group():
    base = sc.with_params(sc.PARITY_CASES["cylinder_M2AUSM_sponge"](), adjust_dt=1, max_co=0.3, have_res=0)
    pcs = cases.partition(base, meshgen.slab_ranks(base.mesh, 2, axis=0))
    gp = [capi.GpuSolver.from_case(pc) for pc in pcs]
    capi.connect(gp)
    grp = capi.PartGroup(gp)
    grp.run(3)
    grp.synchronize()
    dts = [g.compute_dt(base.params["max_co"]) for g in gp]
    # the group's deltaT is the minimum over its partitions
    assert all(g.info().dt == min(dts) for g in gp)
    grp.close()
    for g in gp:
        g.close()


def test_step_reduction_batch():
    cs = [sc.vortex(16, solver=2, minmod=1, mu=0.02), sc.vortex(24, solver=2, minmod=1, mu=0.02)]
    gs = [capi.GpuSolver.from_case(c, precision=capi.PREC_FP32P) for c in cs]
    b = capi.Batch()
    b.set(gs)
    b.run(3)
    b.synchronize()
    for g in gs:
        info = g.info()
        assert info.cfl > 0
        assert g.compute_cfl(info.dt) == info.cfl
    b.close()
    for g in gs:
        g.close()
