"""The oracle's pins catch plausible mistakes (CPU only).

Each case below plants one mistake in a copy of oracle/lfm_oracle.c -- a
flipped sign, swapped sides, a dropped term, a transposed index, a wrong
closed form -- builds it with the oracle's own flags into a temporary
library, and runs the oracle's pin tests against that build
(LFM_ORACLE_LIB).  Every mutant must make at least one of them fail; the
unmodified source must pass them.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lfm_oracle.c")
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle.py", "tests/test_oracle_les.py"]

# (name, text in lfm_oracle.c, replacement)
MUTANTS = [
    ("AUSM pu sign", "double pu = -0.75 *", "double pu = 0.75 *"),
    ("AUSM M+/M- swapped in p_half", "pN * (p5Pos(MN, alpha) - p5Neg(MN, alpha))",
     "pN * (p5Pos(MP, alpha) - p5Neg(MP, alpha))"),
    ("sensor of one cell only", "double tha = 0.5 * (cth + nth);", "double tha = cth;"),
    ("sensor without rotation", "sqrt(cdiv * cdiv + ccurl + 4e-2)", "sqrt(cdiv * cdiv + 4e-2)"),
    ("neighbour side with the cell's gradient", "double pN = interp_minmod(D, cp_, np_, nb->p_grad,",
     "double pN = interp_minmod(D, cp_, np_, cell->p_grad,"),
    ("outlet density kept", "rhoExt = rhoInt * pExt / pInt;", "rhoExt = rhoInt;"),
    ("inlet pressure from the freestream", "pExt = pInt;", "pExt = o->p_inf;"),
    ("sponge without (1 - M^2)", "(3.0 * (1.0 - prm->mach * prm->mach)", "(3.0 * (1.0 - 0.0 * prm->mach)"),
    ("dTdx sign", "vars->dTdx[i] += fT * cs[i];", "vars->dTdx[i] -= fT * cs[i];"),
    ("rhoU gradient transposed", "vars->rhoU_grad[j][i] += rhoU[j] * cs[i];",
     "vars->rhoU_grad[i][j] += rhoU[j] * cs[i];"),
    ("speed-of-sound gradient of Rpsi", "vars->c_grad[i] += cf * cs[i];", "vars->c_grad[i] += Rp * cs[i];"),
    ("laplacian of T without the K term", "double lapT = k_ * (dm * (aT - cT) * dmi + dotD(D, K, dTdx));",
     "double lapT = k_ * (dm * (aT - cT) * dmi);"),
    ("laplacian of T reversed", "double lapT = k_ * (dm * (aT - cT) * dmi", "double lapT = k_ * (dm * (cT - aT) * dmi"),
    ("boundary face without the one-sided dTdx", "dTdx[i] = (aT - cT) * dn[i] * dmi;", "dTdx[i] = dTdx[i];"),
    ("laplacian of U with the stress tensor", "dotD(D, K, dudx[i]));", "dotD(D, K, tau[i]));"),
    ("wall ghost keeps the momentum", "for (int k = 0; k < D; k++) gh->q[k + 1] = -in->q[k + 1];",
     "for (int k = 0; k < D; k++) gh->q[k + 1] = in->q[k + 1];"),
    ("RK5 weight B1 off in the 4th digit", "o->Bk[0] = 0.1496590219993;", "o->Bk[0] = 0.1497590219993;"),
    ("sponge relaxing away from the freestream", "cell->dq[0] += dt * sg * (o->rho_inf - cell->q[0]);",
     "cell->dq[0] -= dt * sg * (o->rho_inf - cell->q[0]);"),
    ("minmod limiter capped at 2", "double m1 = (1.0 < r) ? 1.0 : r;", "double m1 = (2.0 < r) ? 2.0 : r;"),
    ("compute_dt without the factor 2", "double dt = max_co * rho * 2.0 * vol / sum;",
     "double dt = max_co * rho * vol / sum;"),
    ("boundary mirror distance halved", "G->d[s][k] = 2.0 * fabs(dPS) * G->S[s][k] / S2;",
     "G->d[s][k] = 1.0 * fabs(dPS) * G->S[s][k] / S2;"),
]


def _run_pins(lib):
    env = dict(os.environ, LFM_ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
                           *PINS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


def _build(src_text, out):
    c = out + ".c"
    with open(c, "w") as f:
        f.write(src_text)
    subprocess.check_call(["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-I", os.path.join(ROOT, "oracle"), "-o",
                           out, c, "-lm"])


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


def test_unmodified_oracle_passes(source, tmp_path):
    lib = str(tmp_path / "liblfm_oracle_ok.so")
    _build(source, lib)
    r = _run_pins(lib)
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(source, tmp_path, name, old, new):
    assert source.count(old) >= 1, f"mutation site not found: {old}"
    lib = str(tmp_path / "liblfm_oracle_mut.so")
    _build(source.replace(old, new, 1), lib)
    r = _run_pins(lib)
    assert r.returncode != 0, f"{name}: no pin failed\n{r.stdout[-1500:]}"
