"""ctypes binding of the C ABI in include/lfm_hip.h (liblfm_hip.so).

The HIP library is the only compute path: if it is missing or fails to
load, every call raises :class:`LfmError` (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import meshgen

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblfm_hip.so")

LFM_OK, LFM_EINVAL, LFM_EDEVICE, LFM_ENOMEM, LFM_ESTATE, LFM_ENONFINITE = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_FP32P = 64, 32
FIELD_Q, FIELD_DQ, FIELD_DUDX, FIELD_DTDX, FIELD_TAUMC, FIELD_SIGMAU, FIELD_RHOGRAD, FIELD_PGRAD, FIELD_UGRAD, \
    FIELD_RES = range(10)
FIELD_PRIM = 16

# every entry point declared in include/lfm_hip.h
EXPORTS = ("lfm_patch_role", "lfm_solver_create", "lfm_solver_destroy", "lfm_last_error", "lfm_get_info",
           "lfm_prepare_for_timestep", "lfm_prepare_for_rkstep", "lfm_set_boundary_conditions",
           "lfm_calc_gradients", "lfm_calc_vis", "lfm_one_rk_step", "lfm_compute_dt", "lfm_compute_cfl", "lfm_run",
           "lfm_synchronize", "lfm_time_steps", "lfm_get_field", "lfm_set_field", "lfm_get_residual",
           "lfm_check_finite", "lfm_check_finite_noalloc", "lfm_profile_steps", "lfm_conservation", "lfm_batch_create", "lfm_batch_destroy",
           "lfm_batch_set", "lfm_batch_run", "lfm_batch_synchronize", "lfm_solver_create_part",
           "lfm_halo_neighbours", "lfm_halo_export", "lfm_halo_import", "lfm_halo_sizes", "lfm_halo_pack",
           "lfm_halo_unpack", "lfm_part_group_create", "lfm_part_group_run", "lfm_part_group_synchronize",
           "lfm_part_group_destroy", "lfm_case_open", "lfm_case_close", "lfm_case_mesh", "lfm_case_params",
           "lfm_case_fields", "lfm_case_time", "lfm_case_dict_scalar", "lfm_case_write_time",
           "lfm_case_post", "lfm_case_write_scalar", "lfm_post_configure", "lfm_post_average",
           "lfm_post_average_fields", "lfm_post_walls", "lfm_post_forces", "lfm_post_forces_read")
HALO_SOLVARS, HALO_VISCOUS, HALO_PACKED = 1, 2, 3


class LfmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lfm error {code}: {msg}")
        self.code = code


class PolyMeshDesc(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("points", C.c_void_p), ("n_faces", C.c_int32),
                ("face_off", C.c_void_p), ("face_pts", C.c_void_p), ("owner", C.c_void_p),
                ("n_internal", C.c_int32), ("neighbour", C.c_void_p), ("n_cells", C.c_int32),
                ("n_patches", C.c_int32), ("patch_start", C.c_void_p), ("patch_size", C.c_void_p),
                ("patch_role", C.c_void_p), ("patch_nbr", C.c_void_p), ("cell_submesh", C.c_void_p)]


class PartitionDesc(C.Structure):
    _fields_ = [("rank", C.c_int32), ("face_key", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("dim", C.c_int32), ("solver", C.c_int32), ("rk_order", C.c_int32), ("minmod", C.c_int32),
                ("adjust_dt", C.c_int32), ("have_res", C.c_int32), ("max_co", C.c_double), ("dt", C.c_double),
                ("mol_weight", C.c_double), ("cp", C.c_double), ("mu", C.c_double), ("pr", C.c_double),
                ("p_inf", C.c_double), ("T_inf", C.c_double), ("U_inf", C.c_double * 3),
                ("Ls", C.c_double), ("k_sponge", C.c_double), ("mach", C.c_double),
                ("precision", C.c_int32), ("laminar", C.c_int32)]


class Post(C.Structure):
    _fields_ = [("have_average", C.c_int32), ("average_start", C.c_int32), ("have_forces", C.c_int32),
                ("forces_every", C.c_int32), ("forces_own_face", C.c_int32), ("time", C.c_double)]


class Info(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n_cells", C.c_int32), ("n_cells_sm0", C.c_int32), ("n_faces", C.c_int32),
                ("n_ghosts", C.c_int32), ("max_faces_per_cell", C.c_int32), ("precision", C.c_int32),
                ("solver", C.c_int32), ("device_bytes", C.c_int64), ("dt", C.c_double), ("cfl", C.c_double),
                ("steps_done", C.c_int64)]


_lib = None


def load() -> C.CDLL:
    """Load liblfm_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LfmError(LFM_EINVAL, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    L.lfm_last_error.restype = C.c_char_p
    L.lfm_patch_role.argtypes = [C.c_char_p, C.c_char_p]
    L.lfm_solver_create.argtypes = [C.POINTER(PolyMeshDesc), C.POINTER(Params), vp, vp, vp, vp, C.POINTER(vp)]
    L.lfm_solver_create_part.argtypes = [C.POINTER(PolyMeshDesc), C.POINTER(PartitionDesc), C.POINTER(Params), vp,
                                         vp, vp, vp, C.POINTER(vp)]
    L.lfm_halo_neighbours.argtypes = [vp, C.POINTER(i32), vp]
    L.lfm_halo_export.argtypes = [vp, i32, C.POINTER(i32), vp, vp, vp]
    L.lfm_halo_import.argtypes = [vp, i32, i32, vp, vp, vp]
    L.lfm_halo_sizes.argtypes = [vp, i32, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.lfm_halo_pack.argtypes = [vp, i32, C.c_int, vp]
    L.lfm_halo_unpack.argtypes = [vp, i32, C.c_int, vp]
    L.lfm_part_group_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.lfm_part_group_run.argtypes = [vp, C.c_int]
    L.lfm_part_group_synchronize.argtypes = [vp]
    L.lfm_part_group_destroy.argtypes = [vp]
    L.lfm_solver_destroy.argtypes = [vp]
    L.lfm_get_info.argtypes = [vp, C.POINTER(Info)]
    for n in ("lfm_prepare_for_timestep", "lfm_set_boundary_conditions", "lfm_calc_gradients", "lfm_calc_vis",
              "lfm_synchronize"):
        getattr(L, n).argtypes = [vp]
    L.lfm_prepare_for_rkstep.argtypes = [vp, C.c_int]
    L.lfm_one_rk_step.argtypes = [vp, C.c_int, dbl, vp]
    L.lfm_compute_dt.argtypes = [vp, dbl, C.POINTER(dbl)]
    L.lfm_compute_cfl.argtypes = [vp, dbl, C.POINTER(dbl)]
    L.lfm_run.argtypes = [vp, C.c_int]
    L.lfm_time_steps.argtypes = [vp, C.c_int, C.POINTER(dbl)]
    L.lfm_profile_steps.argtypes = [vp, C.c_int, vp, C.POINTER(dbl)]
    L.lfm_get_field.argtypes = [vp, C.c_int, vp]
    L.lfm_set_field.argtypes = [vp, C.c_int, vp]
    L.lfm_get_residual.argtypes = [vp, vp]
    L.lfm_check_finite.argtypes = [vp, C.POINTER(C.c_int64)]
    L.lfm_conservation.argtypes = [vp, vp]
    L.lfm_batch_create.argtypes = [C.POINTER(vp)]
    L.lfm_batch_destroy.argtypes = [vp]
    L.lfm_batch_set.argtypes = [vp, vp, C.c_int]
    L.lfm_batch_run.argtypes = [vp, C.c_int]
    L.lfm_batch_synchronize.argtypes = [vp]
    L.lfm_post_configure.argtypes = [vp, C.POINTER(Post)]
    L.lfm_post_average.argtypes = [vp, i32]
    L.lfm_post_average_fields.argtypes = [vp, i32, vp, vp]
    L.lfm_post_walls.argtypes = [vp, C.POINTER(i32), vp]
    L.lfm_post_forces.argtypes = [vp, dbl]
    L.lfm_post_forces_read.argtypes = [vp, i32, vp, C.POINTER(i32)]
    _lib = L
    return L


def _check(rc: int):
    if rc != LFM_OK:
        raise LfmError(rc, load().lfm_last_error().decode(errors="replace"))


def patch_role(ptype: str, name: str) -> int:
    return load().lfm_patch_role(ptype.encode(), name.encode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class GpuSolver:
    """One partition resident on the current HIP device (both submeshes).

    Mirrors the ISolver time-step methods (api/cfdv0_solver.h:47-104) plus a
    fused ``run`` of whole Mesh::solve iterations."""

    def __init__(self, mesh: "meshgen.PolyMesh", params: dict, p, T, U, alpha=None, precision: int = PREC_FP64,
                 cell_submesh: Optional[Sequence[int]] = None, roles: Optional[Sequence[int]] = None):
        L = load()
        keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a

        start, size, role, nbr = mesh.patch_arrays()
        if roles is not None:
            role = np.asarray(roles, np.int32)
        m = PolyMeshDesc()
        m.n_points = mesh.points.shape[0]
        m.points = _ptr(arr(mesh.points, np.float64))
        m.n_faces = mesh.n_faces
        m.face_off = _ptr(arr(mesh.face_off, np.int32))
        m.face_pts = _ptr(arr(mesh.face_pts, np.int32))
        m.owner = _ptr(arr(mesh.owner, np.int32))
        m.n_internal = mesh.n_internal
        m.neighbour = _ptr(arr(mesh.neighbour, np.int32))
        m.n_cells = mesh.n_cells
        m.n_patches = len(mesh.patches)
        m.patch_start = _ptr(arr(start, np.int32))
        m.patch_size = _ptr(arr(size, np.int32))
        m.patch_role = _ptr(arr(role, np.int32))
        m.patch_nbr = _ptr(arr(nbr, np.int32))
        m.cell_submesh = _ptr(arr(cell_submesh, np.int32)) if cell_submesh is not None else None
        pr = Params()
        for k, v in params.items():
            if k == "U_inf":
                for i in range(3):
                    pr.U_inf[i] = float(v[i])
            elif hasattr(pr, k):
                setattr(pr, k, v)
        pr.precision = precision
        pr.laminar = int(params.get("laminar", 1))
        self.dim = pr.dim
        self.n_cells = mesh.n_cells
        nc = mesh.n_cells
        pa = arr(p, np.float64)
        Ta = arr(T, np.float64)
        Ua = arr(np.asarray(U, np.float64).reshape(nc, 3), np.float64)
        aa = arr(alpha, np.float64) if alpha is not None else None
        h = C.c_void_p()
        self.h = None
        if mesh.face_key is not None:
            part = PartitionDesc()
            part.rank = int(mesh.rank)
            part.face_key = _ptr(arr(mesh.face_key, np.int32))
            _check(L.lfm_solver_create_part(C.byref(m), C.byref(part), C.byref(pr), _ptr(pa), _ptr(Ta), _ptr(Ua),
                                            _ptr(aa), C.byref(h)))
        else:
            _check(L.lfm_solver_create(C.byref(m), C.byref(pr), _ptr(pa), _ptr(Ta), _ptr(Ua), _ptr(aa),
                                       C.byref(h)))
        self.h = h.value

    @classmethod
    def from_case(cls, case, precision: int = PREC_FP64, **kw):
        return cls(case.mesh, case.params, case.p, case.T, case.U, case.alpha, precision=precision, **kw)

    @classmethod
    def from_descriptors(cls, m: PolyMeshDesc, part: Optional[PartitionDesc], prm: Params, fields) -> "GpuSolver":
        """A solver from C-ABI descriptors (as lfm_case_mesh / lfm_case_params
        fill them); ``fields`` has p, T, U (n x 3) and alpha arrays."""
        L = load()
        self = cls.__new__(cls)
        self.h = None
        self.dim = prm.dim
        self.n_cells = m.n_cells
        keep = [np.ascontiguousarray(fields.p, np.float64), np.ascontiguousarray(fields.T, np.float64),
                np.ascontiguousarray(np.asarray(fields.U, np.float64).reshape(-1, 3)),
                None if fields.alpha is None else np.ascontiguousarray(fields.alpha, np.float64)]
        h = C.c_void_p()
        if part is not None:
            _check(L.lfm_solver_create_part(C.byref(m), C.byref(part), C.byref(prm), _ptr(keep[0]), _ptr(keep[1]),
                                            _ptr(keep[2]), _ptr(keep[3]), C.byref(h)))
        else:
            _check(L.lfm_solver_create(C.byref(m), C.byref(prm), _ptr(keep[0]), _ptr(keep[1]), _ptr(keep[2]),
                                       _ptr(keep[3]), C.byref(h)))
        self.h = h.value
        return self

    def close(self):
        if getattr(self, "h", None):
            load().lfm_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ISolver mirror -------------------------------------------------------
    def prepare_for_timestep(self):
        _check(load().lfm_prepare_for_timestep(self.h))

    def prepare_for_RKstep(self, rk: int):
        _check(load().lfm_prepare_for_rkstep(self.h, rk))

    def set_boundary_conditions(self):
        _check(load().lfm_set_boundary_conditions(self.h))

    def calc_gradients(self):
        _check(load().lfm_calc_gradients(self.h))

    def calc_VIS(self):
        _check(load().lfm_calc_vis(self.h))

    def one_rk_step(self, rk: int, dt: float, res: Optional[np.ndarray] = None):
        _check(load().lfm_one_rk_step(self.h, rk, dt, _ptr(res)))

    def compute_dt(self, max_co: float) -> float:
        out = C.c_double()
        _check(load().lfm_compute_dt(self.h, max_co, C.byref(out)))
        return out.value

    def compute_cfl(self, dt: float) -> float:
        out = C.c_double()
        _check(load().lfm_compute_cfl(self.h, dt, C.byref(out)))
        return out.value

    # fused path -------------------------------------------------------------
    def run(self, nsteps: int):
        _check(load().lfm_run(self.h, nsteps))

    def synchronize(self):
        _check(load().lfm_synchronize(self.h))

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(load().lfm_time_steps(self.h, nsteps, C.byref(ms)))
        return ms.value

    PROFILE_GROUPS = ("bc_fill", "grad_vis", "face_flux", "cell_update", "cfl_dt", "cell_update_cfl")

    def profile_steps(self, nsteps: int):
        """(ms per kernel group summed over nsteps, total ms) from HIP events."""
        per = np.zeros(8)
        tot = C.c_double()
        _check(load().lfm_profile_steps(self.h, nsteps, _ptr(per), C.byref(tot)))
        return dict(zip(self.PROFILE_GROUPS, per[:6].tolist())), tot.value

    def info(self) -> Info:
        i = Info()
        _check(load().lfm_get_info(self.h, C.byref(i)))
        return i

    def ncomp(self, which: int) -> int:
        D = self.dim
        return {FIELD_Q: D + 2, FIELD_DQ: D + 2, FIELD_RES: D + 2, FIELD_PRIM: D + 2, FIELD_DUDX: D * D,
                FIELD_TAUMC: D * D, FIELD_UGRAD: D * D}.get(which, D)

    def field(self, which: int) -> np.ndarray:
        out = np.zeros((self.ncomp(which), self.n_cells))
        _check(load().lfm_get_field(self.h, which, _ptr(out)))
        return out

    def q(self) -> np.ndarray:
        return self.field(FIELD_Q)

    def set_field(self, which: int, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float64)
        _check(load().lfm_set_field(self.h, which, _ptr(v)))

    def residual(self) -> np.ndarray:
        out = np.zeros(5)
        _check(load().lfm_get_residual(self.h, _ptr(out)))
        return out[: self.dim + 2]

    def conservation(self) -> np.ndarray:
        """FP64 integrals sum_c V_c q_c (mass, momentum, energy)."""
        out = np.zeros(5)
        _check(load().lfm_conservation(self.h, _ptr(out)))
        return out[: self.dim + 2]

    def n_nonfinite(self) -> int:
        n = C.c_int64()
        _check(load().lfm_check_finite(self.h, C.byref(n)))
        return n.value

    # post-processing ------------------------------------------------------
    def post_configure(self, have_average=False, average_start=0, have_forces=False, forces_every=1,
                       forces_own_face=False, time=0.0, cfg: Optional[Post] = None):
        if cfg is None:
            cfg = Post(int(have_average), int(average_start), int(have_forces), int(forces_every),
                       int(forces_own_face), float(time))
        _check(load().lfm_post_configure(self.h, C.byref(cfg)))

    def post_average(self, n: int):
        """postProcAverage(n)."""
        _check(load().lfm_post_average(self.h, int(n)))

    def average_fields(self, n: int):
        """(pAvg, pRMS) in original cell order (updateAverageField(n))."""
        pavg = np.zeros(self.n_cells)
        prms = np.zeros(self.n_cells)
        _check(load().lfm_post_average_fields(self.h, int(n), _ptr(pavg), _ptr(prms)))
        return pavg, prms

    def walls(self) -> np.ndarray:
        n = C.c_int32()
        _check(load().lfm_post_walls(self.h, C.byref(n), None))
        out = np.zeros(max(n.value, 1), np.int32)
        _check(load().lfm_post_walls(self.h, C.byref(n), _ptr(out)))
        return out[: n.value]

    def post_forces(self, time: float):
        """postProcForces(time): record one row now."""
        _check(load().lfm_post_forces(self.h, float(time)))

    def forces_read(self):
        """Drain the recorded rows: (times [n], forces [n, n_walls, 2, D])."""
        L = load()
        n = C.c_int32()
        _check(L.lfm_post_forces_read(self.h, 0, None, C.byref(n)))
        nw = len(self.walls())
        width = 1 + 2 * self.dim * nw
        rows = np.zeros((max(n.value, 1), width))
        _check(L.lfm_post_forces_read(self.h, n.value, _ptr(rows), C.byref(n)))
        rows = rows[: n.value]
        return rows[:, 0].copy(), rows[:, 1:].reshape(n.value, nw, 2, self.dim)

    # halo (processor patches) ----------------------------------------------
    def halo_neighbours(self):
        n = C.c_int32()
        _check(load().lfm_halo_neighbours(self.h, C.byref(n), None))
        out = np.zeros(max(1, n.value), np.int32)
        _check(load().lfm_halo_neighbours(self.h, C.byref(n), _ptr(out)))
        return out[: n.value].tolist()

    def halo_export(self, rank: int):
        """Set-up message for neighbour ``rank``: (keys, cells, disp[n, 3])."""
        n = C.c_int32()
        _check(load().lfm_halo_export(self.h, rank, C.byref(n), None, None, None))
        keys = np.zeros(n.value, np.int32)
        cells = np.zeros(n.value, np.int32)
        disp = np.zeros((n.value, 3))
        _check(load().lfm_halo_export(self.h, rank, C.byref(n), _ptr(keys), _ptr(cells), _ptr(disp)))
        return keys, cells, disp

    def halo_import(self, rank: int, msg):
        keys, cells, disp = (np.ascontiguousarray(msg[0], np.int32), np.ascontiguousarray(msg[1], np.int32),
                             np.ascontiguousarray(msg[2], np.float64))
        _check(load().lfm_halo_import(self.h, rank, len(keys), _ptr(keys), _ptr(cells), _ptr(disp)))

    def halo_sizes(self, rank: int, kind: int):
        sb, rb = C.c_int64(), C.c_int64()
        _check(load().lfm_halo_sizes(self.h, rank, kind, C.byref(sb), C.byref(rb)))
        return sb.value, rb.value

    def halo_pack(self, rank: int, kind: int, dev_ptr: int):
        _check(load().lfm_halo_pack(self.h, rank, kind, dev_ptr))

    def halo_unpack(self, rank: int, kind: int, dev_ptr: int):
        _check(load().lfm_halo_unpack(self.h, rank, kind, dev_ptr))


def connect(parts: Sequence["GpuSolver"]):
    """Exchange the halo set-up messages between the partitions of one case
    (the reference does this over MPI at start-up)."""
    for q, sq in enumerate(parts):
        for r in sq.halo_neighbours():
            sq.halo_import(r, parts[r].halo_export(q))


class PartGroup:
    """Partitions of one case resident on this device, stepped together with
    device-side halo exchange (lfm_part_group_*)."""

    def __init__(self, parts: Sequence[GpuSolver]):
        arr = (C.c_void_p * len(parts))(*[p.h for p in parts])
        h = C.c_void_p()
        _check(load().lfm_part_group_create(arr, len(parts), C.byref(h)))
        self.h = h.value
        self.parts = list(parts)

    def run(self, nsteps: int):
        _check(load().lfm_part_group_run(self.h, nsteps))

    def synchronize(self):
        _check(load().lfm_part_group_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            load().lfm_part_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Shared launches over resident FP32-perturbation solvers (lfm_batch_*)."""

    def __init__(self):
        h = C.c_void_p()
        _check(load().lfm_batch_create(C.byref(h)))
        self.h = h.value
        self.members = []

    def set(self, solvers):
        arr = (C.c_void_p * max(1, len(solvers)))(*[s.h for s in solvers])
        _check(load().lfm_batch_set(self.h, arr, len(solvers)))
        self.members = list(solvers)

    def run(self, nsteps: int):
        _check(load().lfm_batch_run(self.h, nsteps))

    def synchronize(self):
        _check(load().lfm_batch_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            load().lfm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
