"""ctypes binding of the CPU oracle (oracle/liblfm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LFM_ORACLE_LIB names another build of the oracle (the mutation checks of
# tests/test_oracle_mutants.py load deliberately broken copies through it)
LIB_PATH = os.environ.get("LFM_ORACLE_LIB") or os.path.join(HERE, "liblfm_oracle.so")

F_Q, F_DQ, F_DUDX, F_DTDX, F_TAUMC, F_SIGMAU, F_RHOGRAD, F_PGRAD, F_UGRAD, F_RES, F_VOLINV, F_SIGMA, \
    F_CENTRE, F_VOLUME, F_SUBMESH, F_RHOUGRAD, F_RHOEGRAD, F_RPSIGRAD, F_CGRAD = range(19)


class _Mesh(C.Structure):
    _fields_ = [("n_points", C.c_int), ("points", C.c_void_p), ("n_faces", C.c_int),
                ("face_off", C.c_void_p), ("face_pts", C.c_void_p), ("owner", C.c_void_p),
                ("n_internal", C.c_int), ("neighbour", C.c_void_p), ("n_cells", C.c_int),
                ("n_patches", C.c_int), ("patch_start", C.c_void_p), ("patch_size", C.c_void_p),
                ("patch_role", C.c_void_p), ("patch_nbr", C.c_void_p), ("cell_submesh", C.c_void_p),
                ("face_key", C.c_void_p), ("rank", C.c_int)]


class _Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("solver", C.c_int), ("rk_order", C.c_int), ("minmod", C.c_int),
                ("adjust_dt", C.c_int), ("have_res", C.c_int), ("max_co", C.c_double), ("dt", C.c_double),
                ("mol_weight", C.c_double), ("cp", C.c_double), ("mu", C.c_double), ("pr", C.c_double),
                ("p_inf", C.c_double), ("T_inf", C.c_double), ("U_inf", C.c_double * 3),
                ("Ls", C.c_double), ("k_sponge", C.c_double), ("mach", C.c_double), ("laminar", C.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (ISO C99, no FMA contraction)."""
    src = os.path.join(HERE, "lfm_oracle.c")
    if os.environ.get("LFM_ORACLE_LIB"):
        return LIB_PATH
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.lfmo_create.restype = C.c_void_p
        L.lfmo_create.argtypes = [C.POINTER(_Mesh), C.POINTER(_Params), C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_char_p, C.c_int]
        L.lfmo_destroy.argtypes = [C.c_void_p]
        L.lfmo_run.argtypes = [C.c_void_p, C.c_int]
        L.lfmo_stage.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.lfmo_get_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.lfmo_set_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.lfmo_ncomp.argtypes = [C.c_void_p, C.c_int]
        L.lfmo_get_dt.restype = C.c_double
        L.lfmo_get_dt.argtypes = [C.c_void_p]
        L.lfmo_get_cfl.restype = C.c_double
        L.lfmo_get_cfl.argtypes = [C.c_void_p]
        L.lfmo_get_res.argtypes = [C.c_void_p, C.c_void_p]
        L.lfmo_compute_dt.restype = C.c_double
        L.lfmo_compute_dt.argtypes = [C.c_void_p, C.c_double]
        L.lfmo_compute_cfl.restype = C.c_double
        L.lfmo_compute_cfl.argtypes = [C.c_void_p, C.c_double]
        L.lfmo_get_slots.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.lfmo_n_ghost.argtypes = [C.c_void_p]
        L.lfmo_get_ghosts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lfmo_minmod.restype = C.c_double
        L.lfmo_minmod.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.lfmo_p5.argtypes = [C.c_double, C.c_double, C.c_void_p]
        L.lfmo_halo_nbrs.argtypes = [C.c_void_p, C.c_void_p]
        L.lfmo_halo_export.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lfmo_halo_import.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lfmo_halo_pack.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.lfmo_halo_unpack.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.lfmo_stage_phase.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.lfmo_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.lfmo_set_boundary_conditions.argtypes = [C.c_void_p]
        L.lfmo_average.argtypes = [C.c_void_p, C.c_int]
        L.lfmo_average_fields.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.lfmo_walls.argtypes = [C.c_void_p, C.c_void_p]
        L.lfmo_forces.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.lfmo_face_rhs.argtypes = [C.POINTER(_Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _params(params: dict) -> _Params:
    pr = _Params()
    pr.laminar = 1
    for k, v in params.items():
        if k == "U_inf":
            for i in range(3):
                pr.U_inf[i] = float(v[i])
        elif hasattr(pr, k):
            setattr(pr, k, v)
    return pr


def minmod(D, cphi, aphi, grad, d, wl, flux) -> float:
    g = np.ascontiguousarray(grad, np.float64)
    dd = np.ascontiguousarray(d, np.float64)
    return lib().lfmo_minmod(D, cphi, aphi, _ptr(g), _ptr(dd), wl, flux)


def p5(M, alpha):
    out = np.zeros(2)
    lib().lfmo_p5(M, alpha, _ptr(out))
    return out[0], out[1]


def face_rhs(params: dict, qP, qN, S, d, w, ghost=False, visP=None, visN=None) -> np.ndarray:
    D = params["dim"]
    nvis = 3 * D * D + 4 * D
    vp = np.zeros(nvis) if visP is None else np.ascontiguousarray(visP, np.float64)
    vn = np.zeros(nvis) if visN is None else np.ascontiguousarray(visN, np.float64)
    a = [np.ascontiguousarray(x, np.float64) for x in (qP, qN, S, d)]
    rhs = np.zeros(D + 2)
    pr = _params(params)
    lib().lfmo_face_rhs(C.byref(pr), _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), w, int(ghost), _ptr(vp),
                        _ptr(vn), _ptr(rhs))
    return rhs


class Oracle:
    """One oracle solve: mirrors Mesh<P,D>::initialize/initializeSolver + solve."""

    def __init__(self, mesh, params: dict, p, T, U, alpha=None, cell_submesh=None):
        L = lib()
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            self._keep.append(a)
            return a

        start, size, role, nbr = mesh.patch_arrays()
        m = _Mesh()
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
        fk = getattr(mesh, "face_key", None)
        m.face_key = _ptr(arr(fk, np.int32)) if fk is not None else None
        m.rank = int(getattr(mesh, "rank", -1))
        pr = _Params()
        pr.laminar = 1
        for k, v in params.items():
            if k == "U_inf":
                for i in range(3):
                    pr.U_inf[i] = float(v[i])
            elif hasattr(pr, k):
                setattr(pr, k, v)
        self.dim = pr.dim
        self.n_cells = mesh.n_cells
        nc = mesh.n_cells
        p = arr(p, np.float64)
        T = arr(T, np.float64)
        U = arr(np.asarray(U, np.float64).reshape(nc, 3), np.float64)
        a = arr(alpha if alpha is not None else np.full(nc, 1e300), np.float64)
        err = C.create_string_buffer(256)
        self.h = L.lfmo_create(C.byref(m), C.byref(pr), _ptr(p), _ptr(T), _ptr(U), _ptr(a), err, 256)
        if not self.h:
            raise ValueError("oracle: " + err.value.decode())
        self.params = dict(params)

    def __del__(self):
        if getattr(self, "h", None):
            lib().lfmo_destroy(self.h)
            self.h = None

    def run(self, nsteps: int):
        lib().lfmo_run(self.h, nsteps)

    def stage(self, rk: int, begin_step: bool):
        lib().lfmo_stage(self.h, rk, int(begin_step))

    def field(self, which: int) -> np.ndarray:
        nk = lib().lfmo_ncomp(self.h, which)
        out = np.zeros((nk, self.n_cells))
        lib().lfmo_get_field(self.h, which, _ptr(out))
        return out

    def q(self) -> np.ndarray:
        return self.field(F_Q)

    def set_field(self, which: int, values: np.ndarray):
        """Overwrite q (F_Q) or dq (F_DQ), [(D+2), cells] in original order."""
        v = np.ascontiguousarray(values, np.float64)
        assert v.shape == (self.dim + 2, self.n_cells)
        if lib().lfmo_set_field(self.h, which, _ptr(v)) != 0:
            raise ValueError("only q and dq can be set")

    @property
    def dt(self) -> float:
        return lib().lfmo_get_dt(self.h)

    @property
    def cfl(self) -> float:
        return lib().lfmo_get_cfl(self.h)

    def res(self) -> np.ndarray:
        out = np.zeros(5)
        lib().lfmo_get_res(self.h, _ptr(out))
        return out[: self.dim + 2]

    def compute_dt(self, max_co: float) -> float:
        return lib().lfmo_compute_dt(self.h, max_co)

    def compute_cfl(self, dt: float) -> float:
        return lib().lfmo_compute_cfl(self.h, dt)

    def slots(self):
        nc = self.n_cells
        nface = np.zeros(nc, np.int32)
        nvalid = np.zeros(nc, np.int32)
        nbr = np.zeros(nc * 6, np.int32)
        S = np.zeros(nc * 18)
        d = np.zeros(nc * 18)
        w = np.zeros(nc * 6)
        lib().lfmo_get_slots(self.h, _ptr(nface), _ptr(nvalid), _ptr(nbr), _ptr(S), _ptr(d), _ptr(w))
        return nface, nvalid, nbr.reshape(nc, 6), S.reshape(nc, 6, 3), d.reshape(nc, 6, 3), w.reshape(nc, 6)

    @property
    def n_ghost(self) -> int:
        return lib().lfmo_n_ghost(self.h)

    def ghosts(self):
        """Boundary ghosts: (q [(D+2), ng], interior cell id [ng], role [ng])."""
        ng = self.n_ghost
        q = np.zeros((self.dim + 2, max(ng, 1)))
        cell = np.zeros(max(ng, 1), np.int32)
        role = np.zeros(max(ng, 1), np.int32)
        lib().lfmo_get_ghosts(self.h, _ptr(q), _ptr(cell), _ptr(role))
        return q[:, :ng], cell[:ng], role[:ng]

    # post-processing ---------------------------------------------------------
    def set_boundary_conditions(self):
        lib().lfmo_set_boundary_conditions(self.h)

    def average(self, n: int):
        """postProcAverage(n)."""
        lib().lfmo_average(self.h, int(n))

    def average_fields(self, n: int):
        pavg = np.zeros(self.n_cells)
        prms = np.zeros(self.n_cells)
        if lib().lfmo_average_fields(self.h, int(n), _ptr(pavg), _ptr(prms)) != 0:
            raise ValueError("no averages accumulated")
        return pavg, prms

    def walls(self):
        n = lib().lfmo_walls(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().lfmo_walls(self.h, _ptr(out))
        return out[:n]

    def forces(self, own_face: bool = False) -> np.ndarray:
        """postProcForces: [n_walls, 2, D] (pressure, viscous)."""
        n = len(self.walls())
        out = np.zeros(max(n, 1) * 2 * self.dim)
        lib().lfmo_forces(self.h, int(own_face), _ptr(out))
        return out[: n * 2 * self.dim].reshape(n, 2, self.dim)

    # partitions --------------------------------------------------------------
    def halo_neighbours(self):
        n = lib().lfmo_halo_nbrs(self.h, None)
        out = np.zeros(max(1, n), np.int32)
        lib().lfmo_halo_nbrs(self.h, _ptr(out))
        return out[:n].tolist()

    def halo_export(self, rank: int):
        n = lib().lfmo_halo_export(self.h, rank, None, None, None)
        keys = np.zeros(n, np.int32)
        cells = np.zeros(n, np.int32)
        disp = np.zeros((n, 3))
        lib().lfmo_halo_export(self.h, rank, _ptr(keys), _ptr(cells), _ptr(disp))
        return keys, cells, disp

    def halo_import(self, rank: int, msg):
        keys, cells, disp = (np.ascontiguousarray(msg[0], np.int32), np.ascontiguousarray(msg[1], np.int32),
                             np.ascontiguousarray(msg[2], np.float64))
        rc = lib().lfmo_halo_import(self.h, rank, len(keys), _ptr(keys), _ptr(cells), _ptr(disp))
        if rc != 0:
            raise ValueError(f"oracle halo import from rank {rank} failed ({rc})")

    def halo_pack(self, rank: int, kind: int) -> np.ndarray:
        buf = np.zeros(64 + self.n_cells * 48)
        n = lib().lfmo_halo_pack(self.h, rank, kind, _ptr(buf))
        return buf[:n].copy()

    def halo_unpack(self, rank: int, kind: int, buf: np.ndarray):
        b = np.ascontiguousarray(buf, np.float64)
        if lib().lfmo_halo_unpack(self.h, rank, kind, _ptr(b)) < 0:
            raise ValueError("oracle halo unpack before import")

    def stage_phase(self, rk: int, begin_step: bool, phase: int):
        lib().lfmo_stage_phase(self.h, rk, int(begin_step), phase)

    def set_dt(self, dt: float):
        lib().lfmo_set_dt(self.h, dt)


def connect(parts):
    """The start-up exchange of neighbour centres between partitions."""
    for q, o in enumerate(parts):
        for r in o.halo_neighbours():
            o.halo_import(r, parts[r].halo_export(q))


def _exchange(parts, kind):
    bufs = {(p, r): o.halo_pack(r, kind) for p, o in enumerate(parts) for r in o.halo_neighbours()}
    for q, o in enumerate(parts):
        for r in o.halo_neighbours():
            o.halo_unpack(r, kind, bufs[(r, q)])


def run_partitioned(parts, nsteps: int, with_grads: bool = False):
    """Mesh::solve over ranks (src/mesh_solver.cpp:409-419, 496-717), ranks
    interleaved phase by phase: solvars exchanged before every stage's
    boundary conditions, viscous tensors after calc_VIS, dt reduced with
    MIN over ranks when adjustTimeStep is on.  ``with_grads`` also sends the
    M2AUSM minmod gradients with the viscous payload (not a reference
    payload)."""
    _exchange(parts, 1)
    p0 = parts[0].params
    nrk = int(p0["rk_order"])
    for _ in range(nsteps):
        for rk in range(nrk):
            for o in parts:
                o.stage_phase(rk, rk == 0, 0)
            _exchange(parts, 6 if with_grads else 2)
            for o in parts:
                o.stage_phase(rk, rk == 0, 1)
            _exchange(parts, 1)
        if p0.get("adjust_dt"):
            dt = min(o.compute_dt(p0["max_co"]) for o in parts)
            for o in parts:
                o.set_dt(dt)
        else:
            for o in parts:
                o.compute_cfl(o.dt)
