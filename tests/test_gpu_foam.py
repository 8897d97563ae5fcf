"""OpenFOAM case directories driven on the GPU (needs an MI355X): a solver
built from a case directory steps exactly like one built from the same case
in memory, and the lfm_solve driver runs undecomposed and decomposed cases
to endTime with time-directory output."""
import os

import numpy as np
import pytest

import small_cases as sc
from lfm_public_amd import capi, cases, foam, meshgen, solve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cylinder_M2AUSM_sponge", "ogrid_M2AUSM_cyclic", "vortex_M1_rk3_adjust_res"])
def test_case_directory_solver_matches_in_memory(tmp_path, name):
    c = sc.PARITY_CASES[name]()
    d = str(tmp_path / name)
    foam.write_case(d, c, end_steps=5)
    fc = foam.FoamCase(d)
    g1 = fc.solver()
    g2 = capi.GpuSolver.from_case(c)
    g1.run(5)
    g2.run(5)
    assert np.array_equal(g1.q(), g2.q())
    fc.close()


def test_solve_driver_serial_and_decomposed(tmp_path):
    c = sc.PARITY_CASES["vortex_M2_visc"]()
    ranks = meshgen.slab_ranks(c.mesh, 3)
    ds, dd = str(tmp_path / "serial"), str(tmp_path / "decomposed")
    foam.write_case(ds, c, end_steps=12, write_interval=6)
    foam.write_case(dd, c, end_steps=12, write_interval=6, cell_rank=ranks)
    rs = solve.run_case(ds, quiet=True)
    rd = solve.run_case(dd, quiet=True)
    assert rs["steps"] == rd["steps"] == 12
    assert rs["written"] == pytest.approx([6 * c.params["dt"], 12 * c.params["dt"]])
    # serial driver = in-memory solver
    g = capi.GpuSolver.from_case(c)
    g.run(12)
    assert np.array_equal(rs["q"][0], g.q())
    # decomposed driver = in-memory partition group, and close to serial
    pcs = cases.partition(c, ranks)
    gp = [capi.GpuSolver.from_case(pc) for pc in pcs]
    capi.connect(gp)
    grp = capi.PartGroup(gp)
    grp.run(12)
    grp.synchronize()
    for qd, gi in zip(rd["q"], gp):
        assert np.array_equal(qd, gi.q())
    # the written time directory holds p of the final state
    tname = repr(12 * c.params["dt"])
    assert os.path.isdir(os.path.join(ds, "%.12g" % (12 * c.params["dt"]))) or os.path.isdir(os.path.join(ds, tname))
