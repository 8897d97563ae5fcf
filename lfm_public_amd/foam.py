"""OpenFOAM case directories: writing one from a synthetic Case (ASCII
polyMesh, the dictionaries the reference reads, 0/ fields, optionally
decomposed into processor<N>/ like decomposePar) and reading one through the
library's C reader (include/lfm_case.h).

The dictionary keys are CInputReader's (src/inputReader.cpp:13-64) plus the
Foam::Time controls (startTime, endTime, deltaT, writeControl,
writeInterval, adjustTimeStep, maxCo).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import capi, meshgen
from .cases import Case

_HEADER = "FoamFile\n{{\n    version     2.0;\n    format      ascii;\n    class       {cls};\n" \
          "    location    \"{loc}\";\n    object      {obj};\n}}\n\n"


def _r(x: float) -> str:
    return repr(float(x))


def _write(path: str, text: str):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


_BIN_HEADER = _HEADER.replace("format      ascii;", "format      binary;\n    arch        \"LSB;label=32;scalar=64\";")


def _bin_list(a: np.ndarray) -> bytes:
    """OpenFOAM binary list payload: N(<raw little-endian bytes>)."""
    return f"{a.shape[0]}\n(".encode() + np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes() + b")\n"


def _write_polymesh_binary(pm: str, mesh: meshgen.PolyMesh):
    loc = "constant/polyMesh"
    with open(os.path.join(pm, "points"), "wb") as f:
        f.write(_BIN_HEADER.format(cls="vectorField", loc=loc, obj="points").encode())
        pts = np.ascontiguousarray(mesh.points, np.float64)
        f.write(f"{pts.shape[0]}\n(".encode() + pts.astype("<f8").tobytes() + b")\n")
    with open(os.path.join(pm, "faces"), "wb") as f:
        f.write(_BIN_HEADER.format(cls="faceCompactList", loc=loc, obj="faces").encode())
        f.write(_bin_list(np.asarray(mesh.face_off, np.int32)))
        f.write(_bin_list(np.asarray(mesh.face_pts, np.int32)))
    for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
        with open(os.path.join(pm, name), "wb") as f:
            f.write(_BIN_HEADER.format(cls="labelList", loc=loc, obj=name).encode())
            f.write(_bin_list(np.asarray(arr, np.int32)))


def write_polymesh(d: str, mesh: meshgen.PolyMesh, binary: bool = False):
    """constant/polyMesh/{points,faces,owner,neighbour,boundary} (ASCII, or
    OpenFOAM's binary format for the point, face and label lists)."""
    pm = os.path.join(d, "constant", "polyMesh")
    os.makedirs(pm, exist_ok=True)
    loc = "constant/polyMesh"
    if binary:
        _write_polymesh_binary(pm, mesh)
    else:
        with open(os.path.join(pm, "points"), "w") as f:
            f.write(_HEADER.format(cls="vectorField", loc=loc, obj="points"))
            f.write(f"{mesh.points.shape[0]}\n(\n")
            f.write("".join(f"({_r(x)} {_r(y)} {_r(z)})\n" for x, y, z in mesh.points))
            f.write(")\n")
        with open(os.path.join(pm, "faces"), "w") as f:
            f.write(_HEADER.format(cls="faceList", loc=loc, obj="faces"))
            f.write(f"{mesh.n_faces}\n(\n")
            fo, fp = mesh.face_off, mesh.face_pts
            f.write("".join(f"{fo[i + 1] - fo[i]}({' '.join(map(str, fp[fo[i]:fo[i + 1]]))})\n"
                            for i in range(mesh.n_faces)))
            f.write(")\n")
        note = f"nPoints: {mesh.points.shape[0]} nCells: {mesh.n_cells} nFaces: {mesh.n_faces} " \
               f"nInternalFaces: {mesh.n_internal}"
        for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
            with open(os.path.join(pm, name), "w") as f:
                f.write(_HEADER.format(cls="labelList", loc=loc, obj=name).replace(
                    "    object", f"    note        \"{note}\";\n    object"))
                f.write(f"{arr.shape[0]}\n(\n" + "\n".join(map(str, arr.tolist())) + "\n)\n")
    with open(os.path.join(pm, "boundary"), "w") as f:
        f.write(_HEADER.format(cls="polyBoundaryMesh", loc=loc, obj="boundary"))
        f.write(f"{len(mesh.patches)}\n(\n")
        for p in mesh.patches:
            f.write(f"    {p.name}\n    {{\n        type            {p.type};\n")
            if p.type == "processor":
                f.write(f"        inGroups        List<word> 1(processor);\n        matchTolerance  0.0001;\n"
                        f"        myProcNo        {mesh.rank};\n        neighbProcNo    {p.neighb_proc};\n")
            if p.type == "cyclic" and p.neighbour:
                f.write(f"        neighbourPatch  {p.neighbour};\n")
            f.write(f"        nFaces          {p.size};\n        startFace       {p.start};\n    }}\n")
        f.write(")\n")
    if mesh.face_proc is not None:
        if binary:
            with open(os.path.join(pm, "faceProcAddressing"), "wb") as f:
                f.write(_BIN_HEADER.format(cls="labelList", loc=loc, obj="faceProcAddressing").encode())
                f.write(_bin_list(np.asarray(mesh.face_proc, np.int32)))
        else:
            with open(os.path.join(pm, "faceProcAddressing"), "w") as f:
                f.write(_HEADER.format(cls="labelList", loc=loc, obj="faceProcAddressing"))
                f.write(f"{mesh.face_proc.shape[0]}\n(\n" + "\n".join(map(str, mesh.face_proc.tolist())) +
                        "\n)\n")


def _boundary_block(mesh: meshgen.PolyMesh) -> str:
    out = "boundaryField\n{\n"
    for p in mesh.patches:
        t = p.type if p.type in ("empty", "cyclic", "processor") else "zeroGradient"
        out += f"    {p.name}\n    {{\n        type            {t};\n    }}\n"
    return out + "}\n"


def _write_fields_binary(t: str, time_name: str, bnd: str, p, T, U, alpha=None):
    def field(name, cls, dims, v, kind):
        with open(os.path.join(t, name), "wb") as f:
            f.write(_BIN_HEADER.format(cls=cls, loc=time_name, obj=name).encode())
            f.write(f"dimensions      {dims};\n\ninternalField   nonuniform List<{kind}> ".encode())
            a = np.ascontiguousarray(v, np.float64)
            f.write(f"{a.shape[0]}\n(".encode() + a.astype("<f8").tobytes() + b")\n;\n\n" + bnd.encode())
    field("p", "volScalarField", "[1 -1 -2 0 0 0 0]", p, "scalar")
    field("T", "volScalarField", "[0 0 0 1 0 0 0]", T, "scalar")
    if alpha is not None:
        field("alpha", "volScalarField", "[0 1 0 0 0 0 0]", alpha, "scalar")
    field("U", "volVectorField", "[0 1 -1 0 0 0 0]", np.asarray(U, np.float64).reshape(-1, 3), "vector")


def write_fields(d: str, time_name: str, mesh: meshgen.PolyMesh, p, T, U, alpha=None, binary: bool = False):
    t = os.path.join(d, time_name)
    os.makedirs(t, exist_ok=True)
    bnd = _boundary_block(mesh)
    if binary:
        _write_fields_binary(t, time_name, bnd, p, T, U, alpha)
        return

    def scalar(name, dims, v):
        _write(os.path.join(t, name), _HEADER.format(cls="volScalarField", loc=time_name, obj=name) +
               f"dimensions      {dims};\n\ninternalField   nonuniform List<scalar> \n{len(v)}\n(\n" +
               "".join(_r(x) + "\n" for x in v) + ")\n;\n\n" + bnd)

    scalar("p", "[1 -1 -2 0 0 0 0]", p)
    scalar("T", "[0 0 0 1 0 0 0]", T)
    if alpha is not None:
        scalar("alpha", "[0 1 0 0 0 0 0]", alpha)
    Ua = np.asarray(U, np.float64).reshape(-1, 3)
    _write(os.path.join(t, "U"), _HEADER.format(cls="volVectorField", loc=time_name, obj="U") +
           f"dimensions      [0 1 -1 0 0 0 0];\n\ninternalField   nonuniform List<vector> \n{Ua.shape[0]}\n(\n" +
           "".join(f"({_r(a)} {_r(b)} {_r(c)})\n" for a, b, c in Ua) + ")\n;\n\n" + bnd)


def write_dicts(d: str, params: dict, end_time: float, write_interval: int = 100, extra_control: str = "",
                post: Optional[dict] = None):
    """system/ and constant/ dictionaries; ``post`` holds lfm/post keys
    (haveAverage, tStartAverage, haveForces, saveForcesStep, ...)."""
    prm = params
    yes = lambda b: "true" if b else "false"  # noqa: E731
    post_keys = "".join(f"        {k} {yes(v) if isinstance(v, bool) else _r(v) if isinstance(v, float) else v};\n"
                        for k, v in (post or {}).items())
    _write(os.path.join(d, "system", "controlDict"),
           _HEADER.format(cls="dictionary", loc="system", obj="controlDict") +
           "lfm\n{\n    commType 2;\n    haloCommType 1;\n"
           f"    doublePrecision {yes(prm.get('precision', 64) == 64)};\n"
           f"    post\n    {{\n        haveResiduals {yes(prm.get('have_res', 0))};\n"
           "        printInfoFreq 10;\n" + post_keys + "    }\n}\n\n"
           f"startFrom       startTime;\nstartTime       0;\nstopAt          endTime;\n"
           f"endTime         {_r(end_time)};\ndeltaT          {_r(prm['dt'])};\nwriteControl    timeStep;\n"
           f"writeInterval   {int(write_interval)};\nwriteFormat     ascii;\nwritePrecision  17;\n"
           f"timeFormat      general;\nadjustTimeStep  {'yes' if prm.get('adjust_dt') else 'no'};\n"
           f"maxCo           {_r(prm.get('max_co', 1.0))};\n" + extra_control)
    _write(os.path.join(d, "system", "fvSchemes"),
           _HEADER.format(cls="dictionary", loc="system", obj="fvSchemes") +
           f"lfm\n{{\n    solver          {prm['solver']};\n    dimension       {prm['dim']};\n"
           f"    rkOrder         {prm['rk_order']};\n    minmodExists    {yes(prm.get('minmod', 0))};\n"
           "    constantTimeStep true;\n}\n")
    U = prm["U_inf"]
    _write(os.path.join(d, "constant", "spongeDict"),
           _HEADER.format(cls="dictionary", loc="constant", obj="spongeDict") +
           f"Ls      {_r(prm.get('Ls', 0.0))};\nM       {_r(prm.get('mach', 0.0))};\n"
           f"k       {_r(prm.get('k_sponge', 0.434294481903252))};\npinf    {_r(prm['p_inf'])};\n"
           f"Tinf    {_r(prm['T_inf'])};\nUinf_x  {_r(U[0])};\nUinf_y  {_r(U[1])};\nUinf_z  {_r(U[2])};\n")
    _write(os.path.join(d, "constant", "thermophysicalProperties"),
           _HEADER.format(cls="dictionary", loc="constant", obj="thermophysicalProperties") +
           "thermoType\n{\n    type hePsiThermo;\n    mixture pureMixture;\n}\n\n"
           f"mixture\n{{\n    specie\n    {{\n        nMoles 1;\n        molWeight {_r(prm['mol_weight'])};\n    }}\n"
           f"    thermodynamics\n    {{\n        Cp {_r(prm['cp'])};\n        Hf 0;\n    }}\n"
           f"    transport\n    {{\n        mu {_r(prm['mu'])};\n        Pr {_r(prm['pr'])};\n    }}\n}}\n")
    _write(os.path.join(d, "constant", "turbulenceProperties"),
           _HEADER.format(cls="dictionary", loc="constant", obj="turbulenceProperties") +
           f"simulationType  {'laminar' if prm.get('laminar', 1) else 'LES'};\n")


def write_case(d: str, case: Case, end_steps: int = 100, write_interval: int = 100,
               cell_rank: Optional[np.ndarray] = None, post: Optional[dict] = None, binary: bool = False):
    """The case as an OpenFOAM directory; with cell_rank, decomposed into
    processor<N>/ (polyMesh with processor patches and faceProcAddressing,
    0/ fields) beside the root's system/ and constant/ dictionaries.
    ``binary`` writes the mesh lists and the fields in OpenFOAM's binary
    format."""
    write_dicts(d, case.params, end_steps * case.params["dt"], write_interval, post=post)
    if cell_rank is None:
        write_polymesh(d, case.mesh, binary)
        write_fields(d, "0", case.mesh, case.p, case.T, case.U, case.alpha, binary)
        return
    for lm, cells in meshgen.decompose(case.mesh, cell_rank):
        pd = os.path.join(d, f"processor{lm.rank}")
        write_polymesh(pd, lm, binary)
        write_fields(pd, "0", lm, case.p[cells], case.T[cells], case.U[cells],
                     None if case.alpha is None else case.alpha[cells], binary)


# ---------------------------------------------------------------------------
# reading through the library
# ---------------------------------------------------------------------------
def _lib():
    L = capi.load()
    if not getattr(L, "_case_bound", False):
        vp = C.c_void_p
        L.lfm_case_open.argtypes = [C.c_char_p, C.c_int32, C.POINTER(vp)]
        L.lfm_case_close.argtypes = [vp]
        L.lfm_case_mesh.argtypes = [vp, C.POINTER(capi.PolyMeshDesc), C.POINTER(capi.PartitionDesc)]
        L.lfm_case_params.argtypes = [vp, C.POINTER(capi.Params)]
        L.lfm_case_fields.argtypes = [vp] + [C.POINTER(C.POINTER(C.c_double))] * 4
        L.lfm_case_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.lfm_case_dict_scalar.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_int32,
                                           C.POINTER(C.c_double)]
        L.lfm_case_write_time.argtypes = [vp, C.c_double, vp, C.c_int32]
        L.lfm_case_post.argtypes = [vp, C.POINTER(capi.Post)]
        L.lfm_case_write_scalar.argtypes = [vp, C.c_double, C.c_char_p, C.c_char_p, vp]
        L._case_bound = True
    return L


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ct = {np.int32: C.c_int32, np.float64: C.c_double}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


class FoamCase:
    """An OpenFOAM case read by lfm_case_open (rank < 0: undecomposed)."""

    def __init__(self, case_dir: str, rank: int = -1):
        L = _lib()
        h = C.c_void_p()
        capi._check(L.lfm_case_open(case_dir.encode(), rank, C.byref(h)))
        self.h = h.value
        self.dir = case_dir
        self.rank = rank
        m = capi.PolyMeshDesc()
        part = capi.PartitionDesc()
        capi._check(L.lfm_case_mesh(self.h, C.byref(m), C.byref(part)))
        self.desc, self.part = m, part
        prm = capi.Params()
        capi._check(L.lfm_case_params(self.h, C.byref(prm)))
        self.cparams = prm
        nf, ni, nc, npch = m.n_faces, m.n_internal, m.n_cells, m.n_patches
        face_off = _arr(m.face_off, nf + 1, np.int32)
        self.mesh_arrays = dict(
            points=_arr(m.points, 3 * m.n_points, np.float64).reshape(-1, 3), face_off=face_off,
            face_pts=_arr(m.face_pts, int(face_off[-1]), np.int32), owner=_arr(m.owner, nf, np.int32),
            neighbour=_arr(m.neighbour, ni, np.int32), n_cells=nc,
            patch_start=_arr(m.patch_start, npch, np.int32), patch_size=_arr(m.patch_size, npch, np.int32),
            patch_role=_arr(m.patch_role, npch, np.int32), patch_nbr=_arr(m.patch_nbr, npch, np.int32),
            cell_submesh=_arr(m.cell_submesh, nc, np.int32) if m.cell_submesh else None,
            face_key=_arr(part.face_key, nf - ni, np.int32) if part.face_key else None)
        ptrs = [C.POINTER(C.c_double)() for _ in range(4)]
        capi._check(L.lfm_case_fields(self.h, *[C.byref(x) for x in ptrs]))
        self.p = _arr(ptrs[0], nc, np.float64)
        self.T = _arr(ptrs[1], nc, np.float64)
        self.U = _arr(ptrs[2], 3 * nc, np.float64).reshape(nc, 3)
        self.alpha = _arr(ptrs[3], nc, np.float64) if ptrs[3] else None
        t = [C.c_double() for _ in range(4)]
        wc = C.c_int32()
        capi._check(L.lfm_case_time(self.h, *[C.byref(x) for x in t], C.byref(wc)))
        self.start, self.end, self.delta_t, self.write_interval = [x.value for x in t]
        self.write_control = wc.value

    @property
    def params(self) -> dict:
        p = self.cparams
        d = {k: getattr(p, k) for k, _ in capi.Params._fields_ if k != "U_inf"}
        d["U_inf"] = tuple(p.U_inf)
        return d

    def dict_scalar(self, dict_path: str, subdict: str, key: str, default: Optional[float] = None) -> float:
        out = C.c_double()
        capi._check(_lib().lfm_case_dict_scalar(self.h, dict_path.encode(), subdict.encode(), key.encode(),
                                                0.0 if default is None else default, int(default is not None),
                                                C.byref(out)))
        return out.value

    def solver(self, precision: Optional[int] = None) -> capi.GpuSolver:
        """A resident solver for this case (the case's precision unless given)."""
        prm = capi.Params()
        C.memmove(C.byref(prm), C.byref(self.cparams), C.sizeof(prm))
        if precision is not None:
            prm.precision = precision
        return capi.GpuSolver.from_descriptors(self.desc, self.part if self.rank >= 0 else None, prm, self)

    def write_time(self, time: float, q: np.ndarray, dim: int):
        qa = np.ascontiguousarray(q, np.float64)
        capi._check(_lib().lfm_case_write_time(self.h, time, qa.ctypes.data, dim))

    @property
    def post(self) -> capi.Post:
        """lfm/post controls (haveAverage, tStartAverage, haveForces,
        saveForcesStep) for the start time."""
        out = capi.Post()
        capi._check(_lib().lfm_case_post(self.h, C.byref(out)))
        return out

    def write_scalar(self, time: float, name: str, values: np.ndarray, dimensions: str = ""):
        v = np.ascontiguousarray(values, np.float64)
        capi._check(_lib().lfm_case_write_scalar(self.h, time, name.encode(), dimensions.encode(), v.ctypes.data))

    def close(self):
        if getattr(self, "h", None):
            _lib().lfm_case_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ranks_of(case_dir: str) -> List[int]:
    """processor<N> ranks of a decomposed case (empty when undecomposed)."""
    out = []
    for n in os.listdir(case_dir):
        if n.startswith("processor") and n[9:].isdigit():
            out.append(int(n[9:]))
    return sorted(out)
