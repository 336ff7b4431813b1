"""GPU viscous flux (the intent of RhoSolver::updateViscid, rhoSolver/RhoSolver.cpp:371-429) against the
oracle: residuals to 1e-12 of the flux scale, states after RK steps and LU-SGS iterations to 1e-12
relative, and the shear-wave decay rate."""
import math

import numpy as np
import pytest

import mstcfd as M
from helpers import GAMMA, cons, quad_grid, rel_state_error
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

VISC = dict(mu=0.03, kappa=0.05, cv=2.5)


def mild(n, dim, seed):
    rng = np.random.default_rng(seed)
    return cons(rng.uniform(0.85, 1.2, n), rng.uniform(-0.4, 0.4, (n, dim)), rng.uniform(0.85, 1.2, n))


def inlet(dim):
    return list(cons(np.array([1.2]), np.array([[0.6] + [0.1] * (dim - 1)]), np.array([1.1]))[0])


CASES = {
    "tri_roe1": (lambda: M.Mesh.gen_tri(30, 20, 0, 3, 0, 2, 0.15, 7,
                                        (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY)), M.ROE, 1, M.LIM_NONE),
    "periodic_roe2": (lambda: M.Mesh.gen_tri(24, 24, -1, 1, -1, 1, 0.15, 3, (M.BC_PERIODIC,) * 4), M.ROE, 2,
                      M.LIM_VANLEER),
    "quad_ausm2": (lambda: M.Mesh.from_raw(quad_grid(20, bc=M.BC_WALL, h=0.05, jitter=0.2, seed=4)), M.AUSM, 2,
                   M.LIM_MINMOD),
    "tet_ausm1": (lambda: M.Mesh.gen_tet(6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11,
                                         (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY,
                                          M.BC_SYMMETRY)), M.AUSM, 1, M.LIM_NONE),
    "tet_roe2": (lambda: M.Mesh.gen_tet(5, 4, 4, (0, 0, 0), (1, 0.8, 0.7), 0.1, 5, (M.BC_WALL,) * 6), M.ROE, 2,
                 M.LIM_VANLEER),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_viscous_residual_and_steps_match_oracle(name):
    mk, scheme, order, lim = CASES[name]
    m = mk()
    raw = m.raw()
    dim = m.dim
    kw = dict(dim=dim, scheme=scheme, order=order, limiter=lim, q_inlet=inlet(dim), **VISC)
    print(f"viscous {name}: {m.n_cells} cells", flush=True)
    om = O.OracleMesh(raw)
    Q0 = mild(m.n_cells, dim, 3)
    s = M.Solver(m, M.make_config(**kw))
    s.set_state(Q0)
    Rg = s.residual()
    Rr, Sr = om.residual(O.make_config(**kw), Q0)
    Rv = Rr - om.residual(O.make_config(**dict(kw, mu=0.0, kappa=0.0)), Q0)[0]
    assert np.abs(Rv).max() > 1e-3 * np.abs(Rr).max()  # the viscous part is not negligible
    den = Sr + np.abs(Rv) + 1e-300
    assert float((np.abs(Rg - Rr) / den).max()) <= 1e-12
    # explicit RK2 steps with a fixed time step
    s.set_config(M.make_config(**dict(kw, rk=2)))
    s.set_state(Q0)
    dt = 2e-4
    s.set_dt(dt)
    s.step(5)
    Qr = Q0.copy()
    oc = O.make_config(**dict(kw, rk=2))
    for _ in range(5):
        Qr, _ = om.rk_step(oc, Qr, dt)
    assert rel_state_error(s.get_state(), Qr) <= 1e-12
    s.close()


def test_viscous_lusgs_matches_oracle():
    mk, scheme, order, lim = CASES["tri_roe1"]
    m = mk()
    kw = dict(scheme=scheme, order=order, limiter=lim, q_inlet=inlet(2), **VISC)
    s = M.Solver(m, M.make_config(**kw))
    om = O.OracleMesh(m.raw())
    Q0 = mild(m.n_cells, 2, 4)
    s.set_state(Q0)
    s.set_steady_driver(M.STEADY_LUSGS, 5.0, 1.0)
    col, _ = s.lusgs_colours()
    Qr = Q0.copy()
    for _ in range(3):
        s.lusgs_step(1)
        Qr, _ = om.lusgs_step(O.make_config(**kw), Qr, col, 5.0, 1.0)
    assert rel_state_error(s.get_state(), Qr) <= 1e-12
    s.close()


def test_shear_wave_decay_rate():
    """u = U sin(2 pi y) in a periodic unit box (48 x 48 x 2 triangles): amplitude ~ exp(-nu k^2 t)."""
    m = M.Mesh.gen_tri(48, 48, 0, 1, 0, 1, 0.0, 1, (M.BC_PERIODIC,) * 4)
    cc = m.cell_centres()
    k, mu = 2 * math.pi, 0.02
    n = m.n_cells
    Q = cons(np.ones(n), np.column_stack([0.01 * np.sin(k * cc[:, 1]), np.zeros(n)]), np.full(n, 1 / GAMMA))
    basis = np.sin(k * cc[:, 1])
    rates = {}
    for mu_ in (mu, 0.0):
        s = M.Solver(m, M.make_config(order=2, limiter=M.LIM_VANLEER, rk=2, mu=mu_, kappa=0.0, cv=2.5))
        s.set_state(Q)
        s.set_dt(0.0025)
        s.step(120)
        Qe = s.get_state()
        s.close()
        a0 = (Q[:, 1] * basis).sum()
        a1 = (Qe[:, 1] / Qe[:, 0] * basis).sum()
        rates[mu_] = -math.log(a1 / a0) / (k * k * 120 * 0.0025)
    assert abs(rates[mu] - mu) <= 0.05 * mu, rates
    assert rates[0.0] < 0.05 * mu, rates
