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


class DeviceBuffer:
    """Device memory from the HIP runtime liblfm_hip.so itself loaded (a
    torch build bundles a runtime of its own, which does not share the device
    with it inside one process)."""

    _hip = None

    @classmethod
    def _runtime(cls):
        if cls._hip is None:
            import ctypes as C
            from lfm_public_amd import capi
            capi.load()
            path = None
            with open("/proc/self/maps") as f:
                for line in f:
                    if "libamdhip64.so" in line:
                        path = line.split()[-1]
                        break
            assert path, "liblfm_hip.so did not load a HIP runtime"
            hip = C.CDLL(path)
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipFree.argtypes = [C.c_void_p]
            hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = hip
        return cls._hip

    def __init__(self, nbytes: int):
        import ctypes as C
        hip = self._runtime()
        self.nbytes = max(int(nbytes), 8)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p.value
        assert hip.hipMemset(self.ptr, 0, self.nbytes) == 0
        assert hip.hipDeviceSynchronize() == 0  # the library's streams do not wait for the null stream

    def numpy(self, dtype=np.float64) -> np.ndarray:
        out = np.zeros(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self._runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        if self.ptr:
            self._runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
