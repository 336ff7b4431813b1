"""Oracle viscous flux (the intent of RhoSolver::updateViscid, rhoSolver/RhoSolver.cpp:371-429), pinned
by closed forms and physics, CPU only.

- Green-Gauss gradients of u and T are exact for linear fields on the interior cells of a regular mesh;
- the face stress for a linear velocity field is mu (grad u + grad u^T) + lam (div u) I, and the energy
  flux for a fluid at rest with a linear temperature is kappa grad T, on faces between such cells;
- the residual is the inviscid one minus the gathered viscous flux;
- a sinusoidal shear wave in a periodic box decays at the rate nu k^2 (Stokes' first problem).
"""
import math

import numpy as np
import pytest

import mstcfd as M
from helpers import GAMMA, cons
from oracle import pyoracle as O

MU, KAPPA, CV = 0.03, 0.05, 2.5


def regular_mesh():
    return M.Mesh.gen_tri(12, 10, 0.0, 1.2, 0.0, 1.0, 0.0, 1, (M.BC_WALL,) * 4)


def interior_cells(raw):
    nc = int(raw["n_cells"])
    touch = np.zeros(nc, bool)
    touch[raw["c0"][raw["c1"] < 0]] = True
    # cells whose own faces and whose neighbours' faces are all interior
    nb = [[] for _ in range(nc)]
    for f in np.nonzero(raw["c1"] >= 0)[0]:
        nb[raw["c0"][f]].append(raw["c1"][f])
        nb[raw["c1"][f]].append(raw["c0"][f])
    deep = np.array([not touch[c] and not any(touch[o] for o in nb[c]) for c in range(nc)])
    return ~touch, deep


def cfg(mu=MU, kappa=KAPPA, **kw):
    return O.make_config(mu=mu, kappa=kappa, cv=CV, **kw)


def test_velocity_gradient_and_stress_closed_forms():
    m = regular_mesh()
    raw = m.raw()
    om = O.OracleMesh(raw)
    xc = om.geometry()["ccen"]
    A = np.array([[0.3, -0.7], [0.45, 0.2]])  # du_i/dx_d
    rho = 1.2
    vel = 0.1 + xc @ A.T
    Q = cons(np.full(len(xc), rho), vel, np.full(len(xc), 1.0))
    c = cfg()
    Fv, gw = om.viscous_flux(c, Q)
    inner, deep = interior_cells(raw)
    assert np.allclose(gw[inner, :2, :], A[None], rtol=0, atol=1e-12)
    div = np.trace(A)
    tau = MU * (A + A.T) + c.lam * div * np.eye(2)
    faces = np.nonzero((raw["c1"] >= 0) & inner[raw["c0"]] & inner[np.maximum(raw["c1"], 0)])[0]
    assert len(faces) > 50
    for d in range(2):
        assert np.allclose(Fv[faces, d, 1:3], tau[:, d][None], rtol=0, atol=1e-12)
        assert np.all(Fv[:, d, 0] == 0)


def test_heat_flux_closed_form():
    m = regular_mesh()
    raw = m.raw()
    om = O.OracleMesh(raw)
    xc = om.geometry()["ccen"]
    rho, e = 0.9, np.array([0.8, -0.35])
    E = 2.0 + xc @ e  # fluid at rest: T = E / rho / cv, linear
    Q = np.column_stack([np.full(len(xc), rho), np.zeros(len(xc)), np.zeros(len(xc)), E])
    Fv, gw = om.viscous_flux(cfg(), Q)
    inner, _ = interior_cells(raw)
    assert np.allclose(gw[inner, 2, :], e[None] / rho / CV, rtol=0, atol=1e-12)
    assert np.allclose(gw[inner, :2, :], 0, atol=1e-15)
    faces = np.nonzero((raw["c1"] >= 0) & inner[raw["c0"]] & inner[np.maximum(raw["c1"], 0)])[0]
    for d in range(2):
        assert np.allclose(Fv[faces, d, 3], KAPPA * e[d] / rho / CV, rtol=1e-12, atol=0)
        assert np.allclose(Fv[faces, d, 1:3], 0, atol=1e-15)


@pytest.mark.parametrize("kind", ["tri", "periodic", "tet"])
def test_residual_is_inviscid_minus_gathered_viscous_flux(kind):
    if kind == "tri":
        m = M.Mesh.gen_tri(9, 7, 0, 2, 0, 1, 0.2, 3, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY))
    elif kind == "periodic":
        m = M.Mesh.gen_tri(8, 8, -1, 1, -1, 1, 0.15, 5, (M.BC_PERIODIC,) * 4)
    else:
        m = M.Mesh.gen_tet(3, 3, 2, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11, (M.BC_WALL,) * 6)
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    dim = int(raw["dim"])
    rng = np.random.default_rng(3)
    n = int(raw["n_cells"])
    Q = cons(rng.uniform(0.8, 1.2, n), rng.uniform(-0.3, 0.3, (n, dim)), rng.uniform(0.8, 1.2, n))
    kw = dict(dim=dim, order=2, limiter=O.LIM_VANLEER)
    c = cfg(**kw)
    R1, _ = om.residual(c, Q)
    Fv, _ = om.viscous_flux(c, Q)
    S = geo["S"] * geo["dac"][:, None]

    def gather(F):
        R = np.zeros((n, dim + 2))
        for f in range(len(raw["c0"])):
            g = np.einsum("d,dk->k", S[f], F[f])
            R[raw["c0"][f]] += g
            if raw["c1"][f] >= 0:
                R[raw["c1"][f]] -= g
        return R

    # the convective part of a viscous configuration (its walls are no-slip), gathered in file order
    R0 = gather(om.face_flux(c, Q))
    Rv = gather(Fv)
    scale = np.abs(Rv).max()
    assert scale > 0
    assert np.abs(R1 - (R0 - Rv)).max() <= 1e-12 * max(scale, np.abs(R0).max())


def test_noslip_wall_closed_forms():
    """A linear shear u = a y over a wall at y = 0 (periodic in x, outlet on top), regular mesh.  On a
    no-slip wall the face velocity is zero, which a linear profile through the wall also gives, so the
    Green-Gauss gradient stays exact in the wall cells: du/dy = a.  The wall face carries the shear
    stress mu a in x-momentum, no normal stress and no work (the face temperature is the cell's).  Its
    convective flux pairs the cell with a ghost of negated momentum (rhoSolver/RhoSolver.cpp:160-163),
    not the inviscid mirror image."""
    a = 0.6
    m = M.Mesh.gen_tri(12, 10, 0.0, 1.2, 0.0, 1.0, 0.0, 1, (M.BC_PERIODIC, M.BC_PERIODIC, M.BC_WALL, M.BC_OUTLET))
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    xc = geo["ccen"]
    n = len(xc)
    Q = cons(np.full(n, 1.1), np.column_stack([a * xc[:, 1], np.zeros(n)]), np.full(n, 0.9))
    c = cfg(compat=O.COMPAT_REFERENCE & ~O.COMPAT_GRAD_LASTINT)  # the quirk gives one face a one-sided value
    Fv, gw = om.viscous_flux(c, Q)
    top = np.zeros(n, bool)
    top[raw["c0"][raw["ftype"] == M.BC_OUTLET]] = True
    wall_faces = np.nonzero(raw["ftype"] == M.BC_WALL)[0]
    wall_cells = raw["c0"][wall_faces]
    expect = np.array([[0.0, a], [0.0, 0.0]])
    assert np.allclose(gw[~top, :2, :], expect[None], rtol=0, atol=1e-12)
    assert np.allclose(gw[wall_cells, :2, :], expect[None], rtol=0, atol=1e-12)
    for f in wall_faces:
        # energy: no work at the wall (u_f = 0), only the conduction of the cell's T gradient
        cw = raw["c0"][f]
        assert np.allclose(Fv[f, 1], [0.0, MU * a, 0.0, KAPPA * gw[cw, 2, 1]], rtol=0, atol=1e-13), Fv[f, 1]
    # convective wall flux: Riemann problem of the cell and its negated-momentum ghost
    F = om.face_flux(c, Q)
    flags = om.lrflags(c.compat)
    for f in wall_faces:
        q = Q[raw["c0"][f]]
        g = q.copy()
        g[1:3] = -q[1:3]
        for d in range(2):
            L, R = (q, g) if flags[f, d] else (g, q)
            Fe = O.roe_flux(L, R, d, c)
            assert np.abs(F[f, d] - Fe).max() <= 1e-14 * np.abs(Fe).max()


def shear_wave(om, ccen, mu, steps, dt, order=2):
    k = 2 * math.pi
    u0 = 0.01
    n = len(ccen)
    vel = np.column_stack([u0 * np.sin(k * ccen[:, 1]), np.zeros(n)])
    Q = cons(np.ones(n), vel, np.full(n, 1 / GAMMA))
    c = O.make_config(order=order, limiter=O.LIM_VANLEER, rk=2, mu=mu, kappa=0.0, cv=CV)
    basis = np.sin(k * ccen[:, 1])
    a0 = (Q[:, 1] * basis).sum() / (basis * basis).sum()
    for _ in range(steps):
        Q, _ = om.rk_step(c, Q, dt)
    a1 = (Q[:, 1] / Q[:, 0] * basis).sum() / (basis * basis).sum()
    return -math.log(a1 / a0) / (k * k * steps * dt)


def test_shear_wave_decays_at_the_viscous_rate():
    """u = U sin(2 pi y) in a periodic unit box: the amplitude decays as exp(-nu k^2 t), nu = mu/rho.
    The inviscid run shows how much of the decay is numerical dissipation."""
    m = M.Mesh.gen_tri(24, 24, 0, 1, 0, 1, 0.0, 1, (M.BC_PERIODIC,) * 4)
    om = O.OracleMesh(m.raw())
    ccen = om.geometry()["ccen"]
    mu = 0.02
    nu_eff = shear_wave(om, ccen, mu, 120, 0.0025)
    nu_num = shear_wave(om, ccen, 0.0, 120, 0.0025)
    assert abs(nu_eff - mu) <= 0.12 * mu, nu_eff
    assert nu_num < 0.15 * mu, nu_num


def test_face_gradient_weights_follow_eta0():
    """On a jittered mesh (eta0 != 1/2), the stress of an interior face is built from
    eta0 grad W_c0 + (1 - eta0) grad W_c1 (RhoSolver.cpp:392-394), eta0 = d(c1,f) / (d(c0,f) + d(c1,f))."""
    m = M.Mesh.gen_tri(10, 8, 0, 1, 0, 1, 0.2, 9, (M.BC_WALL,) * 4)
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    rng = np.random.default_rng(5)
    n = int(raw["n_cells"])
    Q = cons(rng.uniform(0.8, 1.2, n), rng.uniform(-0.3, 0.3, (n, 2)), rng.uniform(0.8, 1.2, n))
    c = cfg(compat=0)
    Fv, gw = om.viscous_flux(c, Q)
    faces = np.nonzero(raw["c1"] >= 0)[0]
    eta = geo["eta"][faces]
    assert np.abs(eta - 0.5).max() > 0.05
    g = eta[:, None, None] * gw[raw["c0"][faces]] + (1 - eta)[:, None, None] * gw[raw["c1"][faces]]
    div = g[:, 0, 0] + g[:, 1, 1]
    for d in range(2):
        for i in range(2):
            tau = MU * (g[:, i, d] + g[:, d, i]) + (c.lam * div if i == d else 0)
            assert np.allclose(Fv[faces, d, 1 + i], tau, rtol=1e-12, atol=1e-15)


def test_batches_refuse_viscous_cases():
    """Batches run inviscid cases only; the refusal comes before any device work."""
    m = M.Mesh.gen_tri(8, 4, 0, 1, 0, 0.1, 0.1, 1, (M.BC_WALL,) * 4)
    with pytest.raises(M.MstError, match="inviscid"):
        M.Batch([m], [M.make_config(cfl=0.5, mu=0.01, kappa=0.0, cv=2.5)])
