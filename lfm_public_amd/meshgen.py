"""Synthetic OpenFOAM-style polyMesh generators for the solver's workloads.

Every generator returns a :class:`PolyMesh` with the same arrays an
OpenFOAM ``constant/polyMesh`` directory holds (points, faces, owner,
neighbour, boundary), in OpenFOAM's conventions:

* internal faces first, in upper-triangular order (owner ascending, then
  neighbour ascending), ``owner < neighbour``;
* boundary faces grouped per patch, in patch order;
* face points ordered so the right-hand normal points out of the owner.

2D meshes are one cell thick in z with an ``empty`` front/back patch, as
OpenFOAM stores them (the reference masks those faces,
polyMeshReaderOF/polyMeshReaderOF.cpp:47-59).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence

import numpy as np

# Boundary roles, resolved from patch type and name with the rule of
# init_boundary_conditions (src/cfd_v0.cpp:973-1005).
BC_UNSET, BC_WALL, BC_INLET, BC_OUTLET, BC_EMPTY, BC_CYCLIC, BC_PROCESSOR = 0, 1, 2, 3, 4, 5, 6
INLET_NAMES = ("inlet", "Inlet", "Inflow", "inflow")
OUTLET_NAMES = ("outlet", "Outlet", "Outflow", "outflow")


def patch_role(ptype: str, name: str) -> int:
    """Role of a patch: wall by type, inlet/outlet by name on type ``patch``."""
    if ptype == "empty":
        return BC_EMPTY
    if ptype == "cyclic":
        return BC_CYCLIC
    if ptype == "processor":
        return BC_PROCESSOR
    if ptype == "wall":
        return BC_WALL
    if ptype == "patch":
        if name in INLET_NAMES:
            return BC_INLET
        if name in OUTLET_NAMES:
            return BC_OUTLET
    return BC_UNSET


@dataclasses.dataclass
class Patch:
    name: str
    type: str
    start: int
    size: int
    neighbour: Optional[str] = None  # cyclic partner
    neighb_proc: int = -1            # processor patch: neighbour rank


@dataclasses.dataclass
class PolyMesh:
    points: np.ndarray      # (Np, 3) float64
    face_off: np.ndarray    # (Nf+1,) int32
    face_pts: np.ndarray    # int32
    owner: np.ndarray       # (Nf,) int32
    neighbour: np.ndarray   # (Ni,) int32
    n_cells: int
    patches: List[Patch]
    dim: int = 2
    # decomposed case: this partition's rank and, per boundary face, the
    # global face index shared by both sides of a processor face (-1 else)
    rank: int = -1
    face_key: Optional[np.ndarray] = None
    # decomposePar's faceProcAddressing: +-(global face + 1), negative where
    # the face is reversed (None for an undecomposed mesh)
    face_proc: Optional[np.ndarray] = None

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_internal(self) -> int:
        return int(self.neighbour.shape[0])

    def patch_arrays(self):
        names = [p.name for p in self.patches]
        start = np.array([p.start for p in self.patches], np.int32)
        size = np.array([p.size for p in self.patches], np.int32)
        role = np.array([patch_role(p.type, p.name) for p in self.patches], np.int32)
        nbr = np.array([p.neighb_proc if p.type == "processor" else (names.index(p.neighbour) if p.neighbour else -1)
                        for p in self.patches], np.int32)
        return start, size, role, nbr

    def face_points(self, f: int) -> np.ndarray:
        return self.face_pts[self.face_off[f]:self.face_off[f + 1]]


def _assemble(points, faces_list: Sequence[np.ndarray], owner, neighbour, n_internal,
              patch_specs, n_cells, dim) -> PolyMesh:
    """faces_list: list of (n, k) point arrays, concatenated in final order."""
    sizes = []
    flat = []
    for arr in faces_list:
        if arr.shape[0] == 0:
            continue
        sizes.append(np.full(arr.shape[0], arr.shape[1], np.int64))
        flat.append(arr.reshape(-1))
    sizes = np.concatenate(sizes)
    face_off = np.zeros(sizes.shape[0] + 1, np.int64)
    np.cumsum(sizes, out=face_off[1:])
    face_pts = np.concatenate(flat).astype(np.int32)
    patches = []
    start = n_internal
    for name, ptype, size, nbr in patch_specs:
        patches.append(Patch(name, ptype, start, int(size), nbr))
        start += int(size)
    assert start == owner.shape[0]
    return PolyMesh(np.ascontiguousarray(points, np.float64), face_off.astype(np.int32), face_pts,
                    owner.astype(np.int32), neighbour[:n_internal].astype(np.int32), n_cells, patches, dim)


# --------------------------------------------------------------------------
# 2D: extrude a planar mixed quad/triangle mesh one cell in z
# --------------------------------------------------------------------------
def extrude_2d(nodes_xy: np.ndarray, cells: np.ndarray, classify: Callable[[np.ndarray, np.ndarray], np.ndarray],
               patch_names: Sequence[tuple], thickness: float = 1.0) -> PolyMesh:
    """Build a one-cell-thick polyMesh from planar cells.

    cells: (Nc, K) int array of CCW node ids, K >= 4, padded with ``-1`` after
    each cell's last node (triangles, quads, and polygons of up to K nodes).
    classify(mid_xy, edge_dir) -> patch index per boundary edge (into patch_names,
    which lists (name, type) pairs; an ``empty`` frontAndBack patch is appended).
    """
    nn = nodes_xy.shape[0]
    nc, K = cells.shape
    nv = (cells >= 0).sum(axis=1)
    # edges a->b in CCW order of each cell
    a_list, b_list, c_list = [], [], []
    for k in range(K):
        a = cells[:, k]
        b = np.where(k + 1 < nv, cells[:, (k + 1) % K], cells[:, 0])
        use = k < nv
        a_list.append(a[use]); b_list.append(b[use]); c_list.append(np.nonzero(use)[0])
    ea = np.concatenate(a_list).astype(np.int64)
    eb = np.concatenate(b_list).astype(np.int64)
    ec = np.concatenate(c_list).astype(np.int64)
    key = np.minimum(ea, eb) * nn + np.maximum(ea, eb)
    order = np.lexsort((ec, key))
    key_s, ea_s, eb_s, ec_s = key[order], ea[order], eb[order], ec[order]
    same_next = np.zeros(key_s.shape[0], bool)
    same_next[:-1] = key_s[1:] == key_s[:-1]
    first = np.nonzero(same_next)[0]          # interior pairs (first, first+1); first has the lower cell id
    is_pair_second = np.zeros_like(same_next)
    is_pair_second[first + 1] = True
    bnd = np.nonzero(~same_next & ~is_pair_second)[0]

    # interior faces: owner = lower cell, oriented along the owner's CCW edge
    own_i = ec_s[first]
    nei_i = ec_s[first + 1]
    fa_i, fb_i = ea_s[first], eb_s[first]
    srt = np.lexsort((nei_i, own_i))
    own_i, nei_i, fa_i, fb_i = own_i[srt], nei_i[srt], fa_i[srt], fb_i[srt]

    # boundary edges
    own_b, fa_b, fb_b = ec_s[bnd], ea_s[bnd], eb_s[bnd]
    mid = 0.5 * (nodes_xy[fa_b] + nodes_xy[fb_b])
    pid = classify(mid, nodes_xy[fb_b] - nodes_xy[fa_b])
    srt = np.lexsort((own_b, pid))
    own_b, fa_b, fb_b, pid = own_b[srt], fa_b[srt], fb_b[srt], pid[srt]

    pts = np.zeros((2 * nn, 3))
    pts[:nn, :2] = nodes_xy
    pts[nn:, :2] = nodes_xy
    pts[nn:, 2] = thickness

    def side(fa, fb):
        return np.stack([fa, fb, fb + nn, fa + nn], axis=1)

    faces = [side(fa_i, fb_i)]
    owner = [own_i]
    specs = []
    for p, (name, ptype) in enumerate(patch_names):
        sel = pid == p
        faces.append(side(fa_b[sel], fb_b[sel]))
        owner.append(own_b[sel])
        specs.append((name, ptype, int(sel.sum()), None))
    # front/back: bottom faces reversed (outward -z), top faces CCW (outward
    # +z), one block per node count (OpenFOAM only requires patch-contiguity)
    groups = [(cells[nv == m][:, :m], np.nonzero(nv == m)[0]) for m in sorted(set(nv.tolist()), reverse=True)]
    faces += [g[:, ::-1] for g, _ in groups] + [g + nn for g, _ in groups]
    owner += [i for _, i in groups] + [i for _, i in groups]
    specs.append(("frontAndBack", "empty", 2 * nc, None))
    owner = np.concatenate(owner)
    return _assemble(pts, faces, owner, nei_i, own_i.shape[0], specs, nc, 2)


def cartesian_2d(nx: int, ny: int, x0=-10.0, x1=10.0, y0=-10.0, y1=10.0,
                 names=(("inlet", "patch"), ("outlet", "patch"), ("Inflow", "patch"), ("inflow", "patch")),
                 thickness=1.0) -> PolyMesh:
    """Uniform quads on [x0,x1]x[y0,y1]: patches x-, x+, y-, y+ in that order."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = j * (nx + 1) + i
    cells = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=-1).reshape(-1, 4)
    tol = 1e-9 * max(x1 - x0, y1 - y0)

    def classify(mid, _dir):
        pid = np.full(mid.shape[0], 0)
        pid[np.abs(mid[:, 0] - x0) < tol] = 0
        pid[np.abs(mid[:, 0] - x1) < tol] = 1
        pid[np.abs(mid[:, 1] - y0) < tol] = 2
        pid[np.abs(mid[:, 1] - y1) < tol] = 3
        return pid

    return extrude_2d(nodes, cells, classify, names, thickness)


def polygon_strips_2d(nx: int, ny: int, h: float = 1.0, x0: float = 0.0, y0: float = 0.0,
                      names=(("inlet", "patch"), ("outlet", "patch"), ("Inflow", "patch"), ("inflow", "patch")),
                      thickness=1.0) -> PolyMesh:
    """Columns of alternating coarse and fine cells on [x0, x0+nx h] x [y0, y0+ny h]:
    even columns hold ny cells of height h, odd columns 2 ny cells of height
    h/2, so a coarse cell meets two fine cells on each side it shares with a
    fine column: 6 faces inside (hexagons with collinear edges), 5 in the
    first column (its boundary side is one face) and in a last coarse column.
    Patches x-, x+, y-, y+ in that order."""
    ys = y0 + 0.5 * h * np.arange(2 * ny + 1)          # every vertical line carries the fine spacing
    xs = x0 + h * np.arange(nx + 1)
    # node (i, k) on vertical line i at ys[k]; the boundary lines of coarse
    # columns with no fine neighbour on that side keep only the coarse nodes
    keep = np.ones((nx + 1, 2 * ny + 1), bool)
    for i in (0, nx):
        col = 0 if i == 0 else nx - 1
        if col % 2 == 0:
            keep[i, 1::2] = False
    idx = -np.ones(keep.shape, np.int64)
    idx[keep] = np.arange(int(keep.sum()))
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    nodes = np.stack([X[keep], Y[keep]], axis=1)
    cells = []
    for i in range(nx):
        if i % 2 == 0:
            for j in range(ny):
                k0, km, k1 = 2 * j, 2 * j + 1, 2 * j + 2
                ring = [idx[i, k0], idx[i + 1, k0], idx[i + 1, km], idx[i + 1, k1], idx[i, k1], idx[i, km]]
                cells.append([v for v in ring if v >= 0] + [-1] * sum(v < 0 for v in ring))
        else:
            for k in range(2 * ny):
                cells.append([idx[i, k], idx[i + 1, k], idx[i + 1, k + 1], idx[i, k + 1], -1, -1])
    cells = np.array(cells, np.int64)
    x1, y1 = xs[-1], ys[-1]
    tol = 1e-9 * h

    def classify(mid, _dir):
        pid = np.full(mid.shape[0], 0)
        pid[np.abs(mid[:, 0] - x0) < tol] = 0
        pid[np.abs(mid[:, 0] - x1) < tol] = 1
        pid[np.abs(mid[:, 1] - y0) < tol] = 2
        pid[np.abs(mid[:, 1] - y1) < tol] = 3
        return pid

    return extrude_2d(nodes, cells, classify, names, thickness)


def annulus_mixed(n_theta: int, n_quad: int, n_tri: int, r_in=0.5, r_out=None, jitter=0.1,
                  seed=202, growth=None) -> PolyMesh:
    """Annulus around a cylinder of radius r_in: ``n_quad`` rings of quads next
    to the wall, then ``n_tri`` rings split into triangles; nodes jittered by up
    to ``jitter`` of the local spacing (seeded).  Patches: cylinder (wall),
    inlet (outer, x<0), outlet (outer, x>=0)."""
    nr = n_quad + n_tri
    dth = 2 * math.pi / n_theta
    if growth is None:
        growth = 1.0 + dth  # near-square cells
    dr0 = r_in * dth
    radii = [r_in]
    for k in range(nr):
        radii.append(radii[-1] + dr0 * growth ** k)
    radii = np.array(radii)
    if r_out is not None:
        radii = r_in + (radii - r_in) * (r_out - r_in) / (radii[-1] - r_in)
    th = np.arange(n_theta) * dth
    R, TH = np.meshgrid(radii, th, indexing="ij")
    rng = np.random.default_rng(seed)
    dr = np.gradient(radii)[:, None] * np.ones_like(TH)
    jr = np.zeros_like(R)
    jt = np.zeros_like(R)
    jr[1:-1] = rng.uniform(-jitter, jitter, size=(nr - 1, n_theta)) * dr[1:-1]
    jt[1:-1] = rng.uniform(-jitter, jitter, size=(nr - 1, n_theta)) * dth
    R = R + jr
    TH = TH + jt
    nodes = np.stack([(R * np.cos(TH)).ravel(), (R * np.sin(TH)).ravel()], axis=1)
    k, i = np.meshgrid(np.arange(nr), np.arange(n_theta), indexing="ij")
    a = k * n_theta + i
    b = k * n_theta + (i + 1) % n_theta
    c = (k + 1) * n_theta + (i + 1) % n_theta
    d = (k + 1) * n_theta + i
    quads = np.stack([a, d, c, b], axis=-1)       # CCW: out in r, then back in theta
    cells = []
    for ring in range(nr):
        qa = quads[ring]
        if ring < n_quad:
            cells.append(qa)
        else:
            # split each quad into two triangles, alternating diagonal
            alt = (np.arange(n_theta) + ring) % 2 == 0
            t1 = np.where(alt[:, None], qa[:, [0, 1, 2]], qa[:, [0, 1, 3]])
            t2 = np.where(alt[:, None], qa[:, [0, 2, 3]], qa[:, [1, 2, 3]])
            tt = np.stack([t1, t2], axis=1).reshape(-1, 3)
            cells.append(np.concatenate([tt, -np.ones((tt.shape[0], 1), np.int64)], axis=1))
    cells = np.concatenate(cells).astype(np.int64)
    rmid = 0.5 * (radii[0] + radii[1])

    def classify(mid, _dir):
        rr = np.hypot(mid[:, 0], mid[:, 1])
        pid = np.where(rr < rmid, 0, np.where(mid[:, 0] < 0, 1, 2))
        return pid

    return extrude_2d(nodes, cells, classify, (("cylinder", "wall"), ("inlet", "patch"), ("outlet", "patch")),
                      thickness=1.0)


# --------------------------------------------------------------------------
# 3D structured hex blocks
# --------------------------------------------------------------------------
def _hex_block(P: np.ndarray, ni: int, nj: int, nk: int, periodic_j: bool,
               patch_of_side: dict, cyclic_k: Optional[tuple]) -> PolyMesh:
    """Hex mesh from a point lattice P[(k, j, i)] of shape (nk+1, nj', ni+1, 3)
    where nj' = nj if periodic_j (j wraps conformally) else nj+1.
    Cell id = (k*nj + j)*ni + i.  patch_of_side maps 'i0','i1','j0','j1','k0','k1'
    to (name, type); cyclic_k = (name0, name1) makes k0/k1 a cyclic pair."""
    njp = nj if periodic_j else nj + 1
    pts = P.reshape(-1, 3)

    def pid(k, j, i):
        return (k * njp + (j % njp if periodic_j else j)) * (ni + 1) + i

    def cid(k, j, i):
        return (k * nj + j) * ni + i

    K, J, I = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    faces_int, own_int, nei_int = [], [], []
    # i-faces between (i-1) and i: normal +i, points (i, j..j+1, k..k+1)
    def ifaces(k, j, i):
        return np.stack([pid(k, j, i), pid(k, j + 1, i), pid(k + 1, j + 1, i), pid(k + 1, j, i)], -1)

    def jfaces(k, j, i):
        return np.stack([pid(k, j, i), pid(k + 1, j, i), pid(k + 1, j, i + 1), pid(k, j, i + 1)], -1)

    def kfaces(k, j, i):
        return np.stack([pid(k, j, i), pid(k, j, i + 1), pid(k, j + 1, i + 1), pid(k, j + 1, i)], -1)

    # interior i-faces
    sel = I > 0
    faces_int.append(ifaces(K[sel], J[sel], I[sel])); own_int.append(cid(K[sel], J[sel], I[sel] - 1))
    nei_int.append(cid(K[sel], J[sel], I[sel]))
    # interior j-faces (with wrap)
    if periodic_j:
        jj = J
        ownj = cid(K, (J - 1) % nj, I)
        f = jfaces(K, jj, I)
        o = ownj.ravel(); n = cid(K, J, I).ravel(); f = f.reshape(-1, 4)
        swap = o > n
        # orientation: normal from (j-1) to j is +j; if owner is the larger id, flip
        f[swap] = f[swap][:, ::-1]
        o2 = np.where(swap, n, o); n2 = np.where(swap, o, n)
        faces_int.append(f); own_int.append(o2); nei_int.append(n2)
    else:
        sel = J > 0
        faces_int.append(jfaces(K[sel], J[sel], I[sel])); own_int.append(cid(K[sel], J[sel] - 1, I[sel]))
        nei_int.append(cid(K[sel], J[sel], I[sel]))
    sel = K > 0
    faces_int.append(kfaces(K[sel], J[sel], I[sel])); own_int.append(cid(K[sel] - 1, J[sel], I[sel]))
    nei_int.append(cid(K[sel], J[sel], I[sel]))
    F = np.concatenate([f.reshape(-1, 4) for f in faces_int])
    O = np.concatenate([o.ravel() for o in own_int])
    N = np.concatenate([n.ravel() for n in nei_int])
    srt = np.lexsort((N, O))
    F, O, N = F[srt], O[srt], N[srt]

    faces = [F]
    owner = [O]
    specs = []
    sides = []
    kk, jj = np.meshgrid(np.arange(nk), np.arange(nj), indexing="ij")
    sides.append(("i0", ifaces(kk, jj, 0)[..., ::-1], cid(kk, jj, 0)))
    sides.append(("i1", ifaces(kk, jj, ni), cid(kk, jj, ni - 1)))
    if not periodic_j:
        kk2, ii2 = np.meshgrid(np.arange(nk), np.arange(ni), indexing="ij")
        sides.append(("j0", jfaces(kk2, 0, ii2)[..., ::-1], cid(kk2, 0, ii2)))
        sides.append(("j1", jfaces(kk2, nj, ii2), cid(kk2, nj - 1, ii2)))
    jj3, ii3 = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    sides.append(("k0", kfaces(0, jj3, ii3)[..., ::-1], cid(0, jj3, ii3)))
    sides.append(("k1", kfaces(nk, jj3, ii3), cid(nk - 1, jj3, ii3)))
    names = {}
    for key, f, o in sides:
        name, ptype = patch_of_side[key]
        nbr = None
        if cyclic_k is not None and key in ("k0", "k1"):
            ptype = "cyclic"
            nbr = cyclic_k[1] if key == "k0" else cyclic_k[0]
            name = cyclic_k[0] if key == "k0" else cyclic_k[1]
        f = f.reshape(-1, 4)
        o = o.ravel()
        if name in names:
            # merge into an existing patch of the same name (must be adjacent)
            idx = names[name]
            faces[idx] = np.concatenate([faces[idx], f])
            owner[idx] = np.concatenate([owner[idx], o])
            specs[idx - 1] = (name, ptype, specs[idx - 1][2] + f.shape[0], nbr)
            continue
        names[name] = len(faces)
        faces.append(f)
        owner.append(o)
        specs.append((name, ptype, f.shape[0], nbr))
    own_all = np.concatenate(owner)
    return _assemble(pts, faces, own_all, N, O.shape[0], specs, ni * nj * nk, 3)


def extrude_layers_3d(nodes_xy: np.ndarray, cells: np.ndarray, nz: int, dz: float,
                      classify: Callable[[np.ndarray, np.ndarray], np.ndarray],
                      patch_names: Sequence[tuple], bottom=("Outflow", "patch"),
                      top=("outflow", "patch")) -> PolyMesh:
    """A 3D mesh of nz layers of thickness dz over a planar mesh of triangles
    and quads (cells as in extrude_2d): triangular prisms (5 faces) and hexes
    (6).  Cell id = layer * n2 + planar cell.  Side patches come from
    classify (as in extrude_2d); the bottom and top faces form the last two
    patches."""
    nn = nodes_xy.shape[0]
    n2, K = cells.shape
    nv = (cells >= 0).sum(axis=1)
    a_l, b_l, c_l = [], [], []
    for k in range(K):
        a = cells[:, k]
        b = np.where(k + 1 < nv, cells[:, (k + 1) % K], cells[:, 0])
        use = k < nv
        a_l.append(a[use]); b_l.append(b[use]); c_l.append(np.nonzero(use)[0])
    ea, eb, ec = (np.concatenate(x).astype(np.int64) for x in (a_l, b_l, c_l))
    key = np.minimum(ea, eb) * nn + np.maximum(ea, eb)
    order = np.lexsort((ec, key))
    key_s, ea_s, eb_s, ec_s = key[order], ea[order], eb[order], ec[order]
    same = np.zeros(key_s.shape[0], bool)
    same[:-1] = key_s[1:] == key_s[:-1]
    first = np.nonzero(same)[0]
    second = np.zeros_like(same)
    second[first + 1] = True
    bnd = np.nonzero(~same & ~second)[0]
    # planar interior edges (owner = lower cell, along its CCW edge) and
    # boundary edges with their patches
    ie_o, ie_n, ie_a, ie_b = ec_s[first], ec_s[first + 1], ea_s[first], eb_s[first]
    be_o, be_a, be_b = ec_s[bnd], ea_s[bnd], eb_s[bnd]
    pid = classify(0.5 * (nodes_xy[be_a] + nodes_xy[be_b]), nodes_xy[be_b] - nodes_xy[be_a])

    pts = np.zeros(((nz + 1) * nn, 3))
    for k in range(nz + 1):
        pts[k * nn:(k + 1) * nn, :2] = nodes_xy
        pts[k * nn:(k + 1) * nn, 2] = k * dz

    def side(a, b, k):
        return np.stack([a + k * nn, b + k * nn, b + (k + 1) * nn, a + (k + 1) * nn], axis=1)

    # internal faces: side faces of every layer and the faces between
    # layers, as (owner, neighbour, points) then sorted upper-triangular
    own, nei, fpts = [], [], []
    for k in range(nz):
        own.append(k * n2 + ie_o); nei.append(k * n2 + ie_n)
        fpts += [list(r) for r in side(ie_a, ie_b, k)]
    for k in range(nz - 1):
        for i in range(n2):
            own.append(np.array([k * n2 + i])); nei.append(np.array([(k + 1) * n2 + i]))
            fpts.append(list(cells[i, :nv[i]] + (k + 1) * nn))
    own = np.concatenate(own); nei = np.concatenate(nei)
    srt = np.lexsort((nei, own))
    faces = [np.array(fpts[f], np.int64) for f in srt]
    owner = [own[srt]]
    n_internal = len(faces)
    specs = []
    for p, (name, ptype) in enumerate(patch_names):
        sel = np.nonzero(pid == p)[0]
        cnt = 0
        for k in range(nz):
            for e in sel:
                faces.append(side(be_a[e:e + 1], be_b[e:e + 1], k)[0])
            owner.append(k * n2 + be_o[sel])
            cnt += sel.shape[0]
        specs.append((name, ptype, cnt, None))
    for (name, ptype), k, rev in ((bottom, 0, True), (top, nz - 1, False)):
        lvl = 0 if rev else nz
        for i in range(n2):
            ring = cells[i, :nv[i]] + lvl * nn
            faces.append(ring[::-1] if rev else ring)
        owner.append(k * n2 + np.arange(n2))
        specs.append((name, ptype, n2, None))
    owner = np.concatenate(owner)
    sizes = np.array([len(f) for f in faces], np.int64)
    face_off = np.zeros(len(faces) + 1, np.int64)
    np.cumsum(sizes, out=face_off[1:])
    patches, start = [], n_internal
    for name, ptype, size, nbr in specs:
        patches.append(Patch(name, ptype, start, int(size), nbr))
        start += int(size)
    return PolyMesh(np.ascontiguousarray(pts), face_off.astype(np.int32),
                    np.concatenate(faces).astype(np.int32), owner.astype(np.int32),
                    nei[srt].astype(np.int32), nz * n2, patches, 3)


def prism_hex_3d(nx: int, ny: int, nz: int, h: float = 1.0) -> PolyMesh:
    """nz layers over an nx x ny planar grid whose left half is quads and
    right half triangles (each quad split along a diagonal): hexes and
    triangular prisms.  Patches inlet (x-), outlet (x+), Inflow (y-), inflow
    (y+), Outflow (bottom), outflow (top)."""
    xs = np.arange(nx + 1) * h
    ys = np.arange(ny + 1) * h
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    cells = []
    for j in range(ny):
        for i in range(nx):
            n0 = j * (nx + 1) + i
            q = [n0, n0 + 1, n0 + nx + 2, n0 + nx + 1]
            if i < nx // 2:
                cells.append(q)
            else:
                cells.append([q[0], q[1], q[2], -1])
                cells.append([q[0], q[2], q[3], -1])
    cells = np.array(cells, np.int64)
    x1, y1 = xs[-1], ys[-1]
    tol = 1e-9 * h

    def classify(mid, _dir):
        pid = np.full(mid.shape[0], 0)
        pid[np.abs(mid[:, 0] - x1) < tol] = 1
        pid[np.abs(mid[:, 1]) < tol] = 2
        pid[np.abs(mid[:, 1] - y1) < tol] = 3
        return pid

    return extrude_layers_3d(nodes, cells, nz, h, classify,
                             (("inlet", "patch"), ("outlet", "patch"), ("Inflow", "patch"), ("inflow", "patch")))


def box_3d(n: int, h: float = 1.0, nx=None, ny=None, nz=None) -> PolyMesh:
    """Uniform hex box with spacing h, patches inlet (x-) / outlet (x+) /
    Inflow (y-) / inflow (y+) / Outflow (z-) / outflow (z+)."""
    nx = nx or n; ny = ny or n; nz = nz or n
    x = np.arange(nx + 1) * h - 0.5 * nx * h
    y = np.arange(ny + 1) * h - 0.5 * ny * h
    z = np.arange(nz + 1) * h - 0.5 * nz * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    P = np.stack([X, Y, Z], -1)
    sides = {"i0": ("inlet", "patch"), "i1": ("outlet", "patch"), "j0": ("Inflow", "patch"),
             "j1": ("inflow", "patch"), "k0": ("Outflow", "patch"), "k1": ("outflow", "patch")}
    return _hex_block(P, nx, ny, nz, False, sides, None)


def ogrid_3d(nr: int, nth: int, nz: int, r_in=0.5, r_out=15.0, span=math.pi) -> PolyMesh:
    """Cylinder O-grid: i radial (geometric stretching), j azimuthal (wraps
    conformally), k spanwise with a cyclic front/back pair.  The outer ring is
    split into inlet (x<0) and outlet (x>=0) by face centre."""
    ratio = (r_out / r_in) ** (1.0 / nr)
    r = r_in * ratio ** np.arange(nr + 1)
    th = np.arange(nth) * 2 * math.pi / nth
    z = np.linspace(0.0, span, nz + 1)
    Z, TH, R = np.meshgrid(z, th, r, indexing="ij")
    P = np.stack([R * np.cos(TH), R * np.sin(TH), Z], -1)
    sides = {"i0": ("cylinder", "wall"), "i1": ("farOuter", "patch"), "k0": ("back", "cyclic"),
             "k1": ("front", "cyclic")}
    m = _hex_block(P, nr, nth, nz, True, sides, ("back", "front"))
    # split the outer patch into inlet/outlet by the sign of x of the face centre
    outer = next(p for p in m.patches if p.name == "farOuter")
    fidx = np.arange(outer.start, outer.start + outer.size)
    fp = m.face_pts[(m.face_off[fidx][:, None] + np.arange(4)[None, :])]
    xc = m.points[fp, 0].mean(axis=1)
    order = np.concatenate([fidx[xc < 0], fidx[xc >= 0]])
    nin = int((xc < 0).sum())
    # rewrite faces of the outer patch in the new order
    fp_new = m.face_pts[(m.face_off[order][:, None] + np.arange(4)[None, :])]
    own_new = m.owner[order]
    base = m.face_off[outer.start]
    m.face_pts[base:base + 4 * outer.size] = fp_new.reshape(-1)
    m.owner[outer.start:outer.start + outer.size] = own_new
    i = m.patches.index(outer)
    m.patches[i:i + 1] = [Patch("inlet", "patch", outer.start, nin),
                          Patch("outlet", "patch", outer.start + nin, outer.size - nin)]
    return m


# --------------------------------------------------------------------------
# initial fields
# --------------------------------------------------------------------------
def cell_centres_approx(m: PolyMesh) -> np.ndarray:
    """Mean of face centres (point means) per cell: used only to place initial
    fields, never by the solver's geometry."""
    nf = m.n_faces
    cnt = np.diff(m.face_off)
    fc = np.zeros((nf, 3))
    if np.all(cnt == cnt[0]):
        k = int(cnt[0])
        fc = m.points[m.face_pts.reshape(-1, k)].mean(axis=1)
    else:
        idx = np.repeat(np.arange(nf), cnt)
        np.add.at(fc, idx, m.points[m.face_pts])
        fc /= cnt[:, None]
    nc = m.n_cells
    own = np.asarray(m.owner)
    nei = np.asarray(m.neighbour)
    ni = m.n_internal
    nfc = np.bincount(own, minlength=nc) + np.bincount(nei, minlength=nc)
    cc = np.empty((nc, 3))
    for k in range(3):
        cc[:, k] = np.bincount(own, fc[:, k], minlength=nc) + np.bincount(nei, fc[:ni, k], minlength=nc)
    return cc / nfc[:, None]


def decompose(mesh: PolyMesh, cell_rank: np.ndarray):
    """Split a case into partitions the way decomposePar lays them out.

    Returns, per rank, ``(local mesh, cell addressing)``.  Local cells keep
    their global order; internal faces keep theirs; every original patch is
    kept (possibly empty), restricted to the rank's faces; then one
    ``processor`` patch per neighbour rank (ascending), faces in global face
    order, owned by the local cell (the higher rank sees the face reversed).
    ``face_key`` gives each processor face its global face index.  A cyclic
    face and its partner must fall on the same rank.
    """
    cell_rank = np.asarray(cell_rank, np.int64)
    assert cell_rank.shape[0] == mesh.n_cells
    nr = int(cell_rank.max()) + 1
    nf, ni = mesh.n_faces, mesh.n_internal
    owner = mesh.owner.astype(np.int64)
    nbr = mesh.neighbour.astype(np.int64)
    ro = cell_rank[owner]
    rn = cell_rank[nbr]
    start, size, role, pnbr = mesh.patch_arrays()
    fpatch = np.full(nf, -1, np.int64)
    for i, p in enumerate(mesh.patches):
        fpatch[p.start:p.start + p.size] = i
    for i, p in enumerate(mesh.patches):
        if role[i] == BC_CYCLIC:
            tw = np.arange(p.size) + int(start[pnbr[i]])
            if np.any(ro[p.start:p.start + p.size] != ro[tw]):
                raise ValueError(f"cyclic patch {p.name} straddles partitions")
    out = []
    for r in range(nr):
        cells = np.nonzero(cell_rank == r)[0]
        g2l = np.full(mesh.n_cells, -1, np.int64)
        g2l[cells] = np.arange(cells.shape[0])
        fint = np.nonzero((ro[:ni] == r) & (rn == r))[0]
        bpat = []  # (name, type, faces, neighbour name, neighb_proc)
        for i, p in enumerate(mesh.patches):
            fs = np.arange(p.start, p.start + p.size)
            bpat.append((p.name, p.type, fs[ro[fs] == r], p.neighbour, -1, False))
        cut = np.nonzero(ro[:ni] != rn)[0]
        nbrs = sorted(set(rn[cut][ro[cut] == r].tolist()) | set(ro[cut][rn[cut] == r].tolist()))
        for q in nbrs:
            fs = cut[((ro[cut] == r) & (rn[cut] == q)) | ((rn[cut] == r) & (ro[cut] == q))]
            bpat.append((f"procBoundary{r}to{q}", "processor", np.sort(fs), None, q, True))
        allf = [fint] + [b[2] for b in bpat]
        faces = np.concatenate(allf)
        # local owner / neighbour and point lists (processor faces seen from r)
        lown = np.empty(faces.shape[0], np.int64)
        pts_list = []
        k = 0
        flip_all = np.zeros(faces.shape[0], bool)
        for idx, fs in enumerate(allf):
            is_proc = idx > 0 and bpat[idx - 1][5]
            for f in fs:
                if is_proc and int(ro[f]) != r:
                    lown[k] = g2l[nbr[f]]
                    flip_all[k] = True
                else:
                    lown[k] = g2l[owner[f]]
                k += 1
        lnbr = g2l[nbr[fint]]
        used = np.zeros(mesh.points.shape[0], bool)
        for f in faces:
            used[mesh.face_points(int(f))] = True
        pmap = np.full(mesh.points.shape[0], -1, np.int64)
        pmap[used] = np.arange(int(used.sum()))
        offs = [0]
        flat = []
        for j, f in enumerate(faces):
            fp = pmap[mesh.face_points(int(f))]
            if flip_all[j]:
                fp = fp[::-1]
            flat.append(fp)
            offs.append(offs[-1] + fp.shape[0])
        patches = []
        s0 = fint.shape[0]
        for (name, ptype, fs, pn, q, is_proc) in bpat:
            patches.append(Patch(name, ptype, s0, int(fs.shape[0]), pn, q))
            s0 += int(fs.shape[0])
        key = np.full(faces.shape[0] - fint.shape[0], -1, np.int32)
        off = fint.shape[0]
        for (name, ptype, fs, pn, q, is_proc) in bpat:
            if is_proc:
                key[off - fint.shape[0]:off - fint.shape[0] + fs.shape[0]] = fs
            off += fs.shape[0]
        fproc = (faces.astype(np.int64) + 1) * np.where(flip_all, -1, 1)
        lm = PolyMesh(np.ascontiguousarray(mesh.points[used]), np.asarray(offs, np.int32),
                      np.concatenate(flat).astype(np.int32), lown.astype(np.int32), lnbr.astype(np.int32),
                      int(cells.shape[0]), patches, mesh.dim, rank=r, face_key=key,
                      face_proc=fproc.astype(np.int32))
        out.append((lm, cells))
    return out


def slab_ranks(mesh: PolyMesh, n: int, axis: int = 0) -> np.ndarray:
    """Cell -> rank by equal-count slabs along a coordinate axis."""
    cc = cell_centres_approx(mesh)[:, axis]
    order = np.argsort(cc, kind="stable")
    rank = np.empty(mesh.n_cells, np.int64)
    rank[order] = np.arange(mesh.n_cells) * n // mesh.n_cells
    return rank
