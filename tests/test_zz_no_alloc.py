"""lfm_check_finite_noalloc: the finiteness count without any allocation
(its per-block partials in the library's static device scratch), against
the allocating lfm_check_finite, through the C ABI (needs an MI355X).

Two calls in a row, the second on a smaller solver: the first call leaves
partials in scratch slots the second does not use, and the second must not
read them."""
import ctypes as C

import numpy as np
import pytest

import small_cases as sc
from lfm_public_amd import capi

pytestmark = pytest.mark.gpu


def _noalloc(g):
    f = capi.load().lfm_check_finite_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    f.restype = C.c_int
    n = C.c_int64(-1)
    capi._check(f(g.h, C.byref(n)))
    return n.value


def _alloc(g):
    f = capi.load().lfm_check_finite
    f.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    f.restype = C.c_int
    n = C.c_int64(-1)
    capi._check(f(g.h, C.byref(n)))
    return n.value


def _plant(g, cells, comps, vals):
    q = g.q()
    for c, k, v in zip(cells, comps, vals):
        q[k, c] = v
    g.set_field(capi.FIELD_Q, q)
    return len(set(zip(cells, comps)))


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32P])
def test_noalloc_count_equals_allocating_count(precision):
    print(f"[noalloc] precision {precision}", flush=True)
    rng = np.random.default_rng(precision)
    big = capi.GpuSolver.from_case(sc.seeded_perturbation(sc.box(24)), precision=precision)
    small = capi.GpuSolver.from_case(sc.vortex(12), precision=precision)
    assert _noalloc(big) == _alloc(big) == 0
    nb, nsm = big.n_cells, small.n_cells
    assert nb > 40 * 256 and nsm <= 256  # many scratch slots, then one
    # non-finite values spread over the big mesh, up to its last block
    cells = list(rng.choice(nb, 37, replace=False)) + [nb - 1, nb - 300]
    comps = list(rng.integers(0, 5, len(cells)))
    vals = [np.nan if i % 3 else (np.inf if i % 2 else -np.inf) for i in range(len(cells))]
    want_big = _plant(big, cells, comps, vals)
    want_small = _plant(small, [3, 77, 143], [0, 3, 1], [np.nan, np.inf, np.nan])
    # two calls in a row, the second on the smaller solver
    a1, n1 = _alloc(big), _noalloc(big)
    a2, n2 = _alloc(small), _noalloc(small)
    print(f"[noalloc] big {n1}/{a1} (planted {want_big}), small {n2}/{a2} (planted {want_small})", flush=True)
    assert n1 == a1 == want_big
    assert n2 == a2 == want_small
    # and back: the big solver's own slots are rewritten, not accumulated
    assert _noalloc(big) == want_big
    big.close()
    small.close()
