"""Cell renumbering of a polyMesh (the preprocessing renumberMesh does for the
reference's cases: the examples use RCM, examples/*/constant/renumberMeshDict,
and the repository ships an hpath method, hpathRenumber/hpathRenumber.C).

The time step visits cells in their numbering (boundary submesh first), so
the numbering sets both the reference's accumulation order and the memory
locality of the device gathers.  A renumbered mesh is a different input: the
oracle and the device path agree on it bit for bit, and it agrees with the
original numbering to rounding.

* ``rcm_order``: reverse Cuthill-McKee on the cell graph (scipy).
* ``morton_order``: cells sorted along a Z-order curve of their centres; in
  3D it keeps most face neighbours within a short index distance, which is
  what the per-XCD L2 needs (an x-fastest box has its z neighbours n^2
  cells away).
* ``column_order``: cells grouped into columns of tile x tile cell widths
  in the first two coordinates, each column numbered layer by layer along
  the third (x fastest within a layer).  Neighbours across layers are then
  tile^2 cells apart instead of a whole plane, while runs of consecutive
  cells along x stay contiguous (coalesced gathers); 2D meshes are tiled in
  x only.  Cell coordinates are the centres in units of the median distance
  between face neighbours, so the ordering applies to any mesh.
"""
from __future__ import annotations

import numpy as np

from . import meshgen
from .cases import Case


def cell_graph(mesh: meshgen.PolyMesh):
    import scipy.sparse as sp
    o = mesh.owner[: mesh.n_internal].astype(np.int64)
    n = mesh.neighbour.astype(np.int64)
    nc = mesh.n_cells
    data = np.ones(2 * o.shape[0], np.int8)
    return sp.csr_matrix((data, (np.concatenate([o, n]), np.concatenate([n, o]))), shape=(nc, nc))


def rcm_order(mesh: meshgen.PolyMesh) -> np.ndarray:
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    return np.asarray(reverse_cuthill_mckee(cell_graph(mesh), symmetric_mode=True), np.int64)


def _spread(v: np.ndarray, dim: int) -> np.ndarray:
    """Insert dim-1 zero bits between the bits of v (up to 21 bits)."""
    v = v.astype(np.uint64)
    out = np.zeros_like(v)
    for b in range(21):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim)
    return out


def morton_order(mesh: meshgen.PolyMesh, bits: int = 10) -> np.ndarray:
    cc = meshgen.cell_centres_approx(mesh)
    dim = mesh.dim
    lo = cc[:, :dim].min(axis=0)
    hi = cc[:, :dim].max(axis=0)
    scale = np.where(hi > lo, (2 ** bits - 1) / np.where(hi > lo, hi - lo, 1.0), 0.0)
    q = np.clip(((cc[:, :dim] - lo) * scale).round(), 0, 2 ** bits - 1).astype(np.uint64)
    code = np.zeros(mesh.n_cells, np.uint64)
    for k in range(dim):
        code |= _spread(q[:, k], dim) << np.uint64(k)
    return np.argsort(code, kind="stable").astype(np.int64)


def column_order(mesh: meshgen.PolyMesh, tile: int = 64) -> np.ndarray:
    cc = meshgen.cell_centres_approx(mesh)
    dim = mesh.dim
    lo = cc.min(axis=0)
    # the median distance between face-neighbour centres: the cell width
    ni = mesh.n_internal
    h = float(np.median(np.linalg.norm(cc[mesh.neighbour] - cc[mesh.owner[:ni]], axis=1))) if ni else 1.0
    idx = np.floor((cc - lo) / h + 0.5).astype(np.int64)  # integer cell coordinates
    if dim == 2:
        keys = (idx[:, 0], idx[:, 1], idx[:, 0] // tile)
    else:
        keys = (idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 0] // tile, idx[:, 1] // tile)
    return np.lexsort(keys).astype(np.int64)


def renumber_mesh(mesh: meshgen.PolyMesh, order: np.ndarray) -> meshgen.PolyMesh:
    """New cell i is old cell order[i].  Internal faces are re-sorted into
    OpenFOAM's upper-triangular order (owner ascending, then neighbour) and
    flipped where the new owner would exceed the new neighbour; boundary
    faces keep their patches and order (cyclic pairs stay paired)."""
    order = np.asarray(order, np.int64)
    nc, ni = mesh.n_cells, mesh.n_internal
    assert np.array_equal(np.sort(order), np.arange(nc)), "order must be a permutation of the cells"
    new = np.empty(nc, np.int64)
    new[order] = np.arange(nc)
    o = new[mesh.owner.astype(np.int64)]
    n = new[mesh.neighbour.astype(np.int64)]
    flip = o[:ni] > n
    lo = np.where(flip, n, o[:ni])
    hi = np.where(flip, o[:ni], n)
    perm = np.lexsort((hi, lo))
    # new face list: permuted internal faces (flipped where needed), then the
    # boundary faces unchanged; point lists gathered without a Python loop
    fo = mesh.face_off.astype(np.int64)
    fl = np.diff(fo)
    src_face = np.concatenate([perm, np.arange(ni, mesh.n_faces)])
    rev = np.concatenate([flip[perm], np.zeros(mesh.n_faces - ni, bool)])
    L = fl[src_face]
    face_off = np.zeros(src_face.shape[0] + 1, np.int64)
    np.cumsum(L, out=face_off[1:])
    elem_face = np.repeat(np.arange(src_face.shape[0]), L)
    k = np.arange(face_off[-1]) - face_off[:-1][elem_face]
    kk = np.where(rev[elem_face], L[elem_face] - 1 - k, k)
    face_pts = mesh.face_pts.astype(np.int64)[fo[src_face][elem_face] + kk]
    owner = np.concatenate([lo[perm], o[ni:]])
    return meshgen.PolyMesh(mesh.points.copy(), face_off.astype(np.int32), face_pts.astype(np.int32),
                            owner.astype(np.int32), hi[perm].astype(np.int32), nc,
                            [meshgen.Patch(p.name, p.type, p.start, p.size, p.neighbour, p.neighb_proc)
                             for p in mesh.patches], mesh.dim)


def renumber_case(case: Case, method: str = "morton") -> Case:
    if method == "none":
        return case
    order = {"morton": morton_order, "rcm": rcm_order, "column": column_order}[method](case.mesh)
    m = renumber_mesh(case.mesh, order)
    alpha = None if case.alpha is None else case.alpha[order]
    return Case(f"{case.name}_{method}", m, case.params, case.p[order], case.T[order], case.U[order], alpha)
