"""Partitions on one MI355X: device halo pack/unpack and the partition group
against the oracle's partitioned run (needs a GPU).

Bars: FP64 partitions bit-identical to the oracle's partitioned run (same
arithmetic, same accumulation order, halo payloads copied bit for bit);
partitioned vs serial within 1e-12 where the reference's payloads are
complete (all but solver 2 with minmodExists); FP32P partitions within the
FP32 path's 1e-5 of FP64.
"""
import numpy as np
import pytest

import small_cases as sc
from conftest import DeviceBuffer, pert_l2, primitives
from lfm_public_amd import capi, cases, meshgen

import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 6


def _split(name, nparts):
    c = sc.PARITY_CASES[name]()
    ranks = meshgen.slab_ranks(c.mesh, nparts, axis=0)
    return c, ranks, cases.partition(c, ranks)


def _gpu_parts(pcs, precision=capi.PREC_FP64):
    parts = [capi.GpuSolver.from_case(pc, precision=precision) for pc in pcs]
    capi.connect(parts)
    return parts


def _gather(c, ranks, per_rank_q):
    out = np.zeros((per_rank_q[0].shape[0], c.n_cells))
    for (lm, cells), q in zip(meshgen.decompose(c.mesh, ranks), per_rank_q):
        out[:, cells] = q
    return out


@pytest.mark.parametrize("nparts", [2, 3])
@pytest.mark.parametrize("name", sorted(sc.PARITY_CASES))
def test_partition_group_matches_partitioned_oracle(name, nparts):
    print(f"[partition] {name} x{nparts}", flush=True)
    c, ranks, pcs = _split(name, nparts)
    op = [O.Oracle(pc.mesh, pc.params, pc.p, pc.T, pc.U, pc.alpha) for pc in pcs]
    O.connect(op)
    O.run_partitioned(op, STEPS)
    gp = _gpu_parts(pcs)
    grp = capi.PartGroup(gp)
    grp.run(STEPS)
    grp.synchronize()
    for g, o in zip(gp, op):
        qg, qo = g.q(), o.q()
        assert np.array_equal(qg, qo), np.abs(qg - qo).max()
    if c.params.get("adjust_dt"):
        assert all(g.info().dt == op[0].dt for g in gp)
    grp.close()


def test_device_pack_matches_oracle_pack():
    """The device payloads are the oracle's send buffers, value for value."""
    c, ranks, pcs = _split("cylinder_M2AUSM_sponge", 3)
    op = [O.Oracle(pc.mesh, pc.params, pc.p, pc.T, pc.U, pc.alpha) for pc in pcs]
    O.connect(op)
    O.run_partitioned(op, 3)
    gp = _gpu_parts(pcs)
    grp = capi.PartGroup(gp)
    grp.run(3)
    grp.synchronize()
    for rank, (g, o) in enumerate(zip(gp, op)):
        for nb in g.halo_neighbours():
            for kind in (capi.HALO_SOLVARS, capi.HALO_VISCOUS, capi.HALO_PACKED):
                sb, _ = g.halo_sizes(nb, kind)
                buf = DeviceBuffer(sb)
                g.halo_pack(nb, kind, buf.ptr)
                g.synchronize()
                ref = o.halo_pack(nb, kind)
                assert np.array_equal(buf.numpy()[: sb // 8], ref), (rank, nb, kind)
                buf.close()


def test_per_stage_calls_with_explicit_exchange_match_the_group():
    """The ISolver-style entry points with lfm_halo_pack/unpack between them
    (what an MPI binding of mpi_communication/mpi_wait would call) give the
    group's result."""
    c, ranks, pcs = _split("vortex_M2_visc", 3)
    gp = _gpu_parts(pcs)
    ref = _gpu_parts(pcs)
    grp = capi.PartGroup(ref)
    grp.run(2)
    grp.synchronize()
    bufs = {}
    for p, g in enumerate(gp):
        for nb in g.halo_neighbours():
            sb, _ = g.halo_sizes(nb, capi.HALO_PACKED)
            bufs[(p, nb)] = DeviceBuffer(sb)

    def exchange(kind):
        for p, g in enumerate(gp):
            for nb in g.halo_neighbours():
                g.halo_pack(nb, kind, bufs[(p, nb)].ptr)
            g.synchronize()
        for q, g in enumerate(gp):
            for nb in g.halo_neighbours():
                g.halo_unpack(nb, kind, bufs[(nb, q)].ptr)
            g.synchronize()

    dt = c.params["dt"]
    exchange(capi.HALO_SOLVARS)
    for _ in range(2):
        for g in gp:
            g.prepare_for_timestep()
        for rk in range(c.params["rk_order"]):
            for g in gp:
                g.prepare_for_RKstep(rk)
                g.set_boundary_conditions()
                g.calc_VIS()
            exchange(capi.HALO_VISCOUS)
            for g in gp:
                g.one_rk_step(rk, dt)
            exchange(capi.HALO_SOLVARS)
    for g, r in zip(gp, ref):
        assert np.array_equal(g.q(), r.q())


@pytest.mark.parametrize("name", ["vortex_M1_visc", "cylinder_M2_minmod", "box_M1_visc", "ogrid_M1_cyclic_rk2",
                                  "vortex_M1_rk3_adjust_res"])
def test_partitioned_matches_serial(name):
    c, ranks, pcs = _split(name, 3)
    ser = capi.GpuSolver.from_case(c)
    ser.run(STEPS)
    gp = _gpu_parts(pcs)
    grp = capi.PartGroup(gp)
    grp.run(STEPS)
    grp.synchronize()
    qs = ser.q()
    qp = _gather(c, ranks, [g.q() for g in gp])
    err = float((np.abs(qp - qs).max(axis=1) / np.abs(qs).max(axis=1)).max())
    print(name, "partitioned vs serial", err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["cylinder_M2AUSM_sponge", "box_M2AUSM_visc", "vortex_M1_visc"])
def test_fp32p_partitions_match_fp64_partitions(name):
    c, ranks, pcs = _split(name, 3)
    g64 = _gpu_parts(pcs, capi.PREC_FP64)
    g32 = _gpu_parts(pcs, capi.PREC_FP32P)
    for gp in (g64, g32):
        grp = capi.PartGroup(gp)
        grp.run(20)
        grp.synchronize()
        grp.close()
    q64 = _gather(c, ranks, [g.q() for g in g64])
    q32 = _gather(c, ranks, [g.q() for g in g32])
    g, R = cases.gamma_of(c.params), cases.gas_constant(c.params["mol_weight"])
    D = c.params["dim"]
    p3, U3, T3 = primitives(q32, D, g, R)
    p6, U6, T6 = primitives(q64, D, g, R)
    errs = [pert_l2(p3, p6, c.params["p_inf"]), pert_l2(T3, T6, c.params["T_inf"])]
    errs += [pert_l2(U3[k], U6[k], c.params["U_inf"][k]) for k in range(D)]
    print(name, "fp32p vs fp64 (partitioned)", errs, flush=True)
    assert max(errs) <= 1e-5


def test_partition_guards():
    c, ranks, pcs = _split("vortex_M1_euler", 2)
    gp = [capi.GpuSolver.from_case(pc) for pc in pcs]
    # not imported yet: no stepping, no group
    with pytest.raises(capi.LfmError) as e:
        capi.PartGroup(gp)
    assert e.value.code == capi.LFM_ESTATE
    with pytest.raises(capi.LfmError) as e:
        gp[0].calc_VIS()
    assert e.value.code == capi.LFM_ESTATE
    capi.connect(gp)
    # whole steps on one partition would skip the exchange
    with pytest.raises(capi.LfmError) as e:
        gp[0].run(1)
    assert e.value.code == capi.LFM_ESTATE
    # out of order ranks
    with pytest.raises(capi.LfmError) as e:
        capi.PartGroup(gp[::-1])
    assert e.value.code == capi.LFM_EINVAL
    with pytest.raises(capi.LfmError):
        gp[0].halo_export(5)
    g32 = [capi.GpuSolver.from_case(pc, precision=capi.PREC_FP32P) for pc in pcs]
    b = capi.Batch()
    with pytest.raises(capi.LfmError) as e:
        b.set([g32[0]])
    assert e.value.code == capi.LFM_EINVAL
    b.close()
