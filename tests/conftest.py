"""Shared test set-up: markers, import paths, oracle/library builds."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def primitives(q, dim, gamma, R):
    """p, U (dim comps), T from conservatives (D+2, n)."""
    rho = q[0]
    U = q[1:dim + 1] / rho
    p = (q[dim + 1] - 0.5 * rho * (U * U).sum(0)) * (gamma - 1.0)
    T = p / (rho * R)
    return p, U, T


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))


def pert_l2(x, ref, inf):
    """Relative L2 normalised by the perturbation ||ref - inf||."""
    den = np.linalg.norm(np.asarray(ref) - inf)
    return float(np.linalg.norm(np.asarray(x) - np.asarray(ref)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle as O
    O.build()
    return O
