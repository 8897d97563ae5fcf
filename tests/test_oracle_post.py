"""The oracle's post-processing (postProcAverage / updateAverageField,
cfd_v0.cpp:2966-2970, 3083-3103; postProcForces, cfd_v0.cpp:3169-3233)
against closed forms (CPU).

Averages: after one accumulation pAvg is the cell pressure and pRMS is 0;
after two, pAvg = (p1 + p2)/2 and pRMS = |p2 - p1| / (2 sqrt 2), since the
running sum adds (p_n - mean_n)^2 with the current p already in the mean.

Forces on a channel with walls at y = 0 and y = Ly, at start-up (ghosts
from set_boundary_conditions): a uniform excess pressure dp gives
Fp = -dp Lx t (bottom) and +dp Lx t (top) along y; the shear flow
U = (g y, 0) gives the one-sided wall gradient dU/dy = (U_ghost - U_cell)/h
with U_ghost = -U_cell, hence Fv_x = mu g Lx t on the bottom wall and
2 mu g (Ly - h/2) Lx t / h on the top wall (both drag the walls along +x).

The reference looks a wall face up by the slot it had before reorder_faces
in the arrays as they are after it; on meshes whose boundary faces keep
their slot (every face before them valid) that is the face itself, and on
an O-grid whose cyclic patches are numbered before the wall it is another
face for the cells at the cyclic end (checked against the oracle's own
post-reorder slots and the cells() face order).
"""
import math

import numpy as np
import pytest

import small_cases as sc
from conftest import primitives
from lfm_public_amd import cases, meshgen

import oracle as O


def _oracle(case):
    return O.Oracle(case.mesh, case.params, case.p, case.T, case.U, case.alpha)


def test_average_closed_form():
    c = sc.vortex(16, solver=0)
    o = _oracle(c)
    g = cases.gamma_of(c.params)
    R = cases.gas_constant(c.params["mol_weight"])
    o.average(0)  # n <= 0 adds nothing (steps before tStartAverage)
    o.average(-3)
    o.run(1)
    p1 = primitives(o.q(), 2, g, R)[0]
    o.average(1)
    a1, r1 = o.average_fields(1)
    assert np.allclose(a1, p1, rtol=1e-15, atol=0)
    assert np.all(r1 == 0.0)
    o.run(1)
    p2 = primitives(o.q(), 2, g, R)[0]
    assert np.abs(p2 - p1).max() > 1e-6 * p1.max()
    o.average(2)
    a2, r2 = o.average_fields(2)
    assert np.allclose(a2, 0.5 * (p1 + p2), rtol=1e-14, atol=0)
    assert np.allclose(r2, np.abs(p2 - p1) / (2 * math.sqrt(2)), rtol=1e-6, atol=1e-9 * p1.max())


@pytest.mark.parametrize("own", [0, 1])
def test_forces_pressure_closed_form(own):
    c = sc.channel(dp=250.0)
    o = _oracle(c)
    o.set_boundary_conditions()
    walls = [c.mesh.patches[i].name for i in o.walls()]
    assert walls == ["bottom", "top"]
    F = o.forces(own_face=bool(own))
    A = 2.0 * 0.5
    assert np.allclose(F[0, 0], [0.0, -250.0 * A], rtol=1e-12, atol=1e-9)
    assert np.allclose(F[1, 0], [0.0, 250.0 * A], rtol=1e-12, atol=1e-9)
    assert np.allclose(F[:, 1], 0.0, atol=1e-15)


@pytest.mark.parametrize("own", [0, 1])
def test_forces_shear_closed_form(own):
    g, mu, Ly, ny = 40.0, 1e-3, 1.5, 6
    c = sc.channel(ny=ny, Ly=Ly, mu=mu, shear=g)
    o = _oracle(c)
    o.set_boundary_conditions()
    F = o.forces(own_face=bool(own))
    h, A = Ly / ny, 2.0 * 0.5
    assert np.allclose(F[0, 1], [mu * g * A, 0.0], rtol=1e-12, atol=1e-15)
    assert np.allclose(F[1, 1], [2 * mu * g * (Ly - h / 2) / h * A, 0.0], rtol=1e-12, atol=1e-15)
    # p = p_inf up to the rounding of the conservative round trip
    assert np.abs(F[:, 0]).max() < 1e-9 * c.params["p_inf"] * A


def _cells_faces(mesh):
    """cells(): per cell, owned faces ascending, then neighbour faces
    ascending, without faces of empty patches."""
    empty = np.zeros(mesh.n_faces, bool)
    for p in mesh.patches:
        if p.type == "empty":
            empty[p.start:p.start + p.size] = True
    out = [[] for _ in range(mesh.n_cells)]
    for f in range(mesh.n_faces):
        if not empty[f]:
            out[mesh.owner[f]].append(f)
    for f in range(mesh.n_internal):
        if not empty[f]:
            out[mesh.neighbour[f]].append(f)
    return out


def test_forces_slot_lookup_after_reorder():
    base = sc.ogrid(4, 8, 3, solver=1, minmod=0)
    dp = 0.01
    for order, quirk in ((["cylinder", "inlet", "outlet", "back", "front"], False),
                         (["back", "front", "cylinder", "inlet", "outlet"], True)):
        mesh = sc.reorder_patches(base.mesh, order)
        nc = mesh.n_cells
        p = np.full(nc, base.params["p_inf"] + dp)
        c = cases.Case("ogrid", mesh, base.params, p, base.T, np.zeros((nc, 3)), base.alpha)
        o = _oracle(c)
        o.set_boundary_conditions()
        Fref = o.forces(own_face=False)
        Fown = o.forces(own_face=True)
        # the closed wall: sum of its own S vanishes
        assert np.abs(Fown[0, 0]).max() < 1e-12
        assert np.abs(Fref[0, 1]).max() == 0.0 and np.abs(Fown[0, 1]).max() == 0.0
        # brute force: slot index of each wall face in its cell's face list,
        # read from the oracle's post-reorder slot arrays
        nface, nvalid, nbr, S, d, w = o.slots()
        cf = _cells_faces(mesh)
        wall = next(p for p in mesh.patches if p.name == "cylinder")
        tot = np.zeros(3)
        moved = 0
        for f in range(wall.start, wall.start + wall.size):
            cell = mesh.owner[f]
            k = cf[cell].index(f)
            tot += S[cell, k]
            moved += nbr[cell, k] >= 0  # a real cell: not the wall face's ghost
        assert np.allclose(Fref[0, 0], dp * tot, rtol=1e-12, atol=1e-14)
        if quirk:
            assert moved > 0 and np.abs(Fref[0, 0]).max() > 1e-6
        else:
            assert moved == 0 and np.array_equal(Fref, Fown)
