"""Case set-up for the solver's workloads: meshes, gas and scheme parameters,
initial p/T/U fields and the sponge distance field ``alpha``.

Parameters use the dictionary keys of the reference (src/inputReader.cpp:13-64):
``solver``/``rkOrder``/``minmodExists`` (fvSchemes lfm{}), ``molWeight``/``Cp``/
``mu``/``Pr`` (thermophysicalProperties), ``pinf``/``Tinf``/``Uinf_*``/``Ls``/``M``/
``k`` (spongeDict), ``deltaT``/``adjustTimeStep``/``maxCo`` (controlDict).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import meshgen

R_UNIVERSAL = 8.31447  # api/cfdv0_solver.h:145

AIR = dict(mol_weight=28.9647, pr=0.72)
# "normalised" gas of the reference examples: c = 1 at T = 1, gamma = 7/5
# (examples/cylinder_vortex/constant/thermophysicalProperties)
NORM_GAS = dict(mol_weight=11640.3, cp=2.5, pr=0.75)


def gas_constant(mol_weight: float) -> float:
    return R_UNIVERSAL / mol_weight * 1000


def base_params(dim=2, solver=0, rk_order=5, minmod=0, mu=0.0, dt=1e-3, adjust_dt=0, max_co=1.0,
                have_res=0, p_inf=1e5, T_inf=300.0, U_inf=(0.0, 0.0, 0.0), Ls=0.0, k_sponge=4.34294481903252e-01,
                mach=0.0, mol_weight=28.9647, cp=None, pr=0.72):
    R = gas_constant(mol_weight)
    if cp is None:
        cp = 3.5 * R  # gamma = 1.4 exactly by cp/(cp-R)
    return dict(dim=dim, solver=solver, rk_order=rk_order, minmod=minmod, adjust_dt=adjust_dt, have_res=have_res,
                max_co=max_co, dt=dt, mol_weight=mol_weight, cp=cp, mu=mu, pr=pr, p_inf=p_inf, T_inf=T_inf,
                U_inf=tuple(U_inf), Ls=Ls, k_sponge=k_sponge, mach=mach)


def gamma_of(params) -> float:
    R = gas_constant(params["mol_weight"])
    return params["cp"] / (params["cp"] - R)


def sound_speed(params) -> float:
    return math.sqrt(gamma_of(params) * gas_constant(params["mol_weight"]) * params["T_inf"])


class Case:
    """mesh + params + initial fields (original cell order)."""

    def __init__(self, name, mesh, params, p, T, U, alpha=None):
        self.name = name
        self.mesh = mesh
        self.params = params
        self.p = np.ascontiguousarray(p, np.float64)
        self.T = np.ascontiguousarray(T, np.float64)
        self.U = np.ascontiguousarray(np.asarray(U, np.float64).reshape(-1, 3))
        self.alpha = None if alpha is None else np.ascontiguousarray(alpha, np.float64)

    @property
    def n_cells(self):
        return self.mesh.n_cells


def isentropic_vortex(nx=256, ny=None, L=10.0, strength=5.0, radius=1.0, mach=0.5, seed=101, solver=0,
                      steps_cfl=0.5, rk_order=5, minmod=0, mu=0.0, thickness=1.0, mesh=None) -> Case:
    """Shu isentropic vortex convected at ``mach`` in a [-L, L]^2 box of air
    (p = 1e5 Pa, T = 300 K, gamma = 1.4), constant dt at CFL ``steps_cfl``."""
    ny = ny or nx
    if mesh is None:
        mesh = meshgen.cartesian_2d(nx, ny, -L, L, -L, L, thickness=thickness)
    prm = base_params(dim=2, solver=solver, rk_order=rk_order, minmod=minmod, mu=mu)
    g = gamma_of(prm)
    R = gas_constant(prm["mol_weight"])
    c_inf = math.sqrt(g * R * prm["T_inf"])
    U_inf = mach * c_inf
    prm["U_inf"] = (U_inf, 0.0, 0.0)
    prm["mach"] = mach
    dx = 2 * L / nx
    prm["dt"] = steps_cfl * dx / (U_inf + c_inf)
    rng = np.random.default_rng(seed)
    xc0 = rng.uniform(-0.8 * L, -0.6 * L)
    yc0 = rng.uniform(-0.2 * L, 0.2 * L)
    cc = meshgen.cell_centres_approx(mesh)
    x = (cc[:, 0] - xc0) / radius
    y = (cc[:, 1] - yc0) / radius
    r2 = x * x + y * y
    us = math.sqrt(R * prm["T_inf"])  # velocity scale: c_inf = sqrt(gamma) * us
    f = strength / (2 * math.pi) * np.exp(0.5 * (1 - r2))
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = U_inf - us * f * y
    U[:, 1] = us * f * x
    Tn = 1.0 - (g - 1) * strength ** 2 / (8 * g * math.pi ** 2) * np.exp(1 - r2)
    T = prm["T_inf"] * Tn
    p = prm["p_inf"] * Tn ** (g / (g - 1))
    return Case(f"vortex{nx}x{ny}", mesh, prm, p, T, U)


def cylinder_2d(n_theta=512, n_quad=None, n_tri=None, mach=0.3, noise=1e-4, seed=202, solver=2, minmod=1,
                Ls_diam=5.0, Re=150.0, r_out=None, jitter=0.1, rk_order=5, cfl=0.5, mesh=None,
                speed_factor=1.0) -> Case:
    """Unstructured annulus (quads next to the wall, triangles outside) around a
    unit-diameter cylinder, uniform inflow at ``mach`` with seeded pressure noise,
    sponge of length ``Ls_diam`` diameters measured from the outlet patch."""
    n_quad = n_quad if n_quad is not None else n_theta // 8
    n_tri = n_tri if n_tri is not None else n_theta // 8
    if mesh is None:
        mesh = meshgen.annulus_mixed(n_theta, n_quad, n_tri, r_in=0.5, r_out=r_out, jitter=jitter, seed=seed)
    prm = base_params(dim=2, solver=solver, rk_order=rk_order, minmod=minmod, p_inf=1.0, T_inf=1.0,
                      mol_weight=NORM_GAS["mol_weight"], cp=NORM_GAS["cp"], pr=NORM_GAS["pr"])
    c_inf = sound_speed(prm)
    U_inf = mach * c_inf
    prm["U_inf"] = (U_inf, 0.0, 0.0)
    prm["mach"] = mach
    prm["mu"] = U_inf * 1.0 / Re  # rho_inf = p/(R T) = 1/R; mu chosen for Re on D=1 with rho_inf
    rho_inf = prm["p_inf"] / (gas_constant(prm["mol_weight"]) * prm["T_inf"])
    prm["mu"] *= rho_inf
    rng = np.random.default_rng(seed)
    nc = mesh.n_cells
    p = prm["p_inf"] * (1.0 + noise * rng.uniform(-1.0, 1.0, nc))
    T = np.full(nc, prm["T_inf"])
    U = np.zeros((nc, 3))
    U[:, 0] = U_inf
    cc = meshgen.cell_centres_approx(mesh)
    # distance to the outlet patch (outer boundary, x >= 0)
    r = np.hypot(cc[:, 0], cc[:, 1])
    pts = mesh.points
    r_outer = np.hypot(pts[:, 0], pts[:, 1]).max()
    theta = np.arctan2(cc[:, 1], cc[:, 0])
    # nearest point of the outlet arc (|theta| <= pi/2 on the outer circle)
    th_c = np.clip(theta, -math.pi / 2, math.pi / 2)
    nx_, ny_ = r_outer * np.cos(th_c), r_outer * np.sin(th_c)
    alpha = np.hypot(cc[:, 0] - nx_, cc[:, 1] - ny_)
    Ls = Ls_diam * 1.0
    prm["Ls"] = Ls
    # constant dt from the smallest (jittered) wall cell: convective CFL and
    # the explicit viscous limit nu dt / h^2 <= 1/4
    h_min = (1.0 - jitter) * 0.5 * 2 * math.pi / n_theta
    nu = prm["mu"] / rho_inf
    # (speed_factor scales U_inf to the fastest flow expected: 2 for the
    # potential-flow speed at the cylinder's shoulders)
    prm["dt"] = cfl * min(h_min / (speed_factor * U_inf + c_inf), 0.25 * h_min * h_min / nu)
    del r
    return Case(f"cylinder2d_{nc}", mesh, prm, p, T, U, alpha)


def acoustic_box(n=64, mach=0.1, amp=1e-3, halfwidth_cells=8.0, seed=404, solver=0, rk_order=5, cfl=0.5,
                 mu=0.0) -> Case:
    """Gaussian acoustic pulse in a uniform flow at ``mach`` in an n^3 unit-spaced
    hex box; the pulse centre is jittered by the seed."""
    mesh = meshgen.box_3d(n, 1.0)
    prm = base_params(dim=3, solver=solver, rk_order=rk_order, mu=mu)
    g = gamma_of(prm)
    c_inf = sound_speed(prm)
    U_inf = mach * c_inf
    prm["U_inf"] = (U_inf, 0.0, 0.0)
    prm["mach"] = mach
    prm["dt"] = cfl * 1.0 / (U_inf + c_inf)
    rng = np.random.default_rng(seed)
    c0 = rng.uniform(-0.5, 0.5, 3)
    cc = meshgen.cell_centres_approx(mesh)
    r2 = ((cc - c0) ** 2).sum(1)
    pp = amp * prm["p_inf"] * np.exp(-math.log(2.0) * r2 / halfwidth_cells ** 2)
    p = prm["p_inf"] + pp
    # isentropic: T/T_inf = (p/p_inf)^((g-1)/g)
    T = prm["T_inf"] * (p / prm["p_inf"]) ** ((g - 1) / g)
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = U_inf
    return Case(f"box{n}", mesh, prm, p, T, U)


def cylinder_3d(nr=64, nth=64, nz=16, mach=0.2, Re=3900.0, pert=0.01, seed=303, solver=2, minmod=1, Ls=13.0,
                r_out=15.0, rk_order=5, cfl=0.5) -> Case:
    """3D cylinder O-grid with spanwise cyclic pair; mu from Re on D = 1;
    1 % spanwise velocity perturbation with seeded phases; sponge from the
    outlet side of the outer boundary."""
    span = math.pi
    mesh = meshgen.ogrid_3d(nr, nth, nz, 0.5, r_out, span)
    prm = base_params(dim=3, solver=solver, rk_order=rk_order, minmod=minmod, p_inf=1.0, T_inf=1.0,
                      mol_weight=NORM_GAS["mol_weight"], cp=NORM_GAS["cp"], pr=NORM_GAS["pr"])
    c_inf = sound_speed(prm)
    U_inf = mach * c_inf
    rho_inf = prm["p_inf"] / (gas_constant(prm["mol_weight"]) * prm["T_inf"])
    prm["U_inf"] = (U_inf, 0.0, 0.0)
    prm["mach"] = mach
    prm["mu"] = rho_inf * U_inf * 1.0 / Re
    rng = np.random.default_rng(seed)
    cc = meshgen.cell_centres_approx(mesh)
    ph = rng.uniform(0, 2 * math.pi, 4)
    kz = 2 * math.pi / span
    w = sum(np.sin((m + 1) * kz * cc[:, 2] + ph[m]) for m in range(4)) / 4.0
    nc = mesh.n_cells
    U = np.zeros((nc, 3))
    U[:, 0] = U_inf
    U[:, 2] = pert * U_inf * w
    p = np.full(nc, prm["p_inf"])
    T = np.full(nc, prm["T_inf"])
    r = np.hypot(cc[:, 0], cc[:, 1])
    theta = np.arctan2(cc[:, 1], cc[:, 0])
    th_c = np.clip(theta, -math.pi / 2, math.pi / 2)
    alpha = np.hypot(cc[:, 0] - r_out * np.cos(th_c), cc[:, 1] - r_out * np.sin(th_c))
    prm["Ls"] = Ls
    h_min = 0.5 * (0.5 * ((r_out / 0.5) ** (1.0 / nr) - 1.0))
    prm["dt"] = cfl * min(h_min, 0.5 * 2 * math.pi / nth, span / nz) / (U_inf + c_inf)
    del r
    return Case(f"cyl3d_{nr}x{nth}x{nz}", mesh, prm, p, T, U, alpha)


def partition(case: Case, cell_rank) -> List[Case]:
    """The case decomposed like decomposePar (meshgen.decompose): one Case
    per rank with its local mesh (processor patches, face keys) and the
    initial fields of its cells."""
    out = []
    for lm, cells in meshgen.decompose(case.mesh, cell_rank):
        alpha = None if case.alpha is None else case.alpha[cells]
        out.append(Case(f"{case.name}_rank{lm.rank}", lm, case.params, case.p[cells], case.T[cells], case.U[cells],
                        alpha))
    return out
