"""The case reader on the reference's own example inputs.

tests/golden/reference_examples/ holds the dictionaries of the reference's
seven example cases (examples/*/system/{controlDict,fvSchemes},
examples/*/constant/{thermophysicalProperties,spongeDict,
turbulenceProperties}; copied by tests/golden/make_reference_dicts.py).
Each is placed over a case directory with a mesh and initial fields of the
example's dimension (the examples' polyMesh lists are not in the reference
repository), opened through lfm_case_open, and the values the reference's
CInputReader::initialize (src/inputReader.cpp:13-64), dictReaderOF and
runTimeManagerOF take from those files are checked against the values the
files state (transcribed below).  The examples' gas is the "normalised" one
their thermophysicalProperties describe: speed of sound 1 at T = 1 and
gamma = 7/5, which pins the molWeight / Cp reading.
"""
import os
import shutil

import pytest

import small_cases as sc
from lfm_public_amd import cases, foam

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_examples")
FILES = ["system/controlDict", "system/fvSchemes", "constant/thermophysicalProperties", "constant/spongeDict",
         "constant/turbulenceProperties"]

_2D = dict(dim=2, rk_order=5, minmod=0, precision=64, laminar=1, adjust_dt=0, max_co=1.0, have_res=1, cp=2.5,
           mu=0.0018667, pr=0.75, mol_weight=11640.3, p_inf=1.0, T_inf=1.0, U_inf=(0.2, 0.0, 0.0), mach=0.2,
           start=0.0, end=100.0, write_interval=500, have_average=1, have_forces=0, forces_every=1000)
EXPECTED = {
    "SBS_antiphase": dict(_2D, solver=0, dt=1.5e-2, Ls=40.0),
    "SBS_inphase": dict(_2D, solver=0, dt=1.5e-2, Ls=40.0),
    "cylinder_vortex": dict(_2D, solver=0, dt=3e-3, Ls=15.0),
    "cylinder_vortex_unstructured": dict(_2D, solver=1, dt=2e-3, Ls=15.0, mol_weight=11640.258, end=2.4e3,
                                         write_interval=100000, have_average=0, have_forces=1, forces_every=1),
    "tandem_vortex": dict(_2D, solver=0, dt=9e-3, Ls=40.0),
    "3D_Cylinder_Re3900_S": dict(_2D, dim=3, solver=1, dt=2e-3, Ls=13.0, mu=7.17948717948718e-05, have_res=0,
                                 end=1.0, write_interval=10000, have_average=0, have_forces=0,
                                 forces_every=1000000),
    "3D_Cylinder_Re3900_L": dict(_2D, dim=3, solver=1, dt=2e-3, Ls=13.0, mu=7.17948717948718e-05, have_res=0,
                                 end=1.0, write_interval=10000, have_average=0, have_forces=0,
                                 forces_every=1000000),
}
PARAM_KEYS = ["dim", "solver", "rk_order", "minmod", "precision", "laminar", "adjust_dt", "max_co", "have_res",
              "dt", "cp", "mu", "pr", "mol_weight", "p_inf", "T_inf", "U_inf", "Ls", "mach"]


def test_fixtures_present():
    assert sorted(os.listdir(GOLDEN)) == sorted(EXPECTED)
    for name in EXPECTED:
        for f in FILES:
            assert os.path.isfile(os.path.join(GOLDEN, name, f)), (name, f)


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_example_dictionaries(tmp_path, name):
    exp = EXPECTED[name]
    case = sc.vortex(6) if exp["dim"] == 2 else sc.box(4)
    d = str(tmp_path)
    foam.write_case(d, case, end_steps=10, write_interval=5)
    for f in FILES:
        shutil.copyfile(os.path.join(GOLDEN, name, f), os.path.join(d, f))
    fc = foam.FoamCase(d)
    prm = fc.params
    got = {k: prm[k] for k in PARAM_KEYS}
    want = {k: exp[k] for k in PARAM_KEYS}
    assert got == want
    # spongeDict's k is absent from the examples: the reader's default
    assert prm["k_sponge"] == pytest.approx(0.434294481903252, rel=1e-15)
    assert (fc.start, fc.end, fc.delta_t, fc.write_interval) == (exp["start"], exp["end"], exp["dt"],
                                                                  exp["write_interval"])
    assert fc.write_control == 0  # writeControl timeStep
    post = fc.post
    assert (post.have_average, post.have_forces, post.forces_every) == (exp["have_average"], exp["have_forces"],
                                                                        exp["forces_every"])
    # the normalised gas: c = sqrt(gamma R T) = 1 at T = 1, gamma = 7/5
    assert cases.gamma_of(prm) == pytest.approx(1.4, rel=2e-5)
    assert cases.sound_speed(prm) == pytest.approx(1.0, rel=2e-5)
    fc.close()
