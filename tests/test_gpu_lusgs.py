"""GPU LU-SGS steady driver (mst_lusgs_step / mst_run_to_steady) against the oracle.

The device colouring is checked as integers (a proper colouring, the preprocessing's); the oracle then
runs the same colours, so the states after each iteration must agree to 1e-12 relative
(the order of the cells inside a colour does not change the result, only the rounding of sums).
"""
import os

import numpy as np
import pytest

import mstcfd as M
from helpers import cons, periodic_in_x, quad_grid, rel_state_error
from mstcfd import ramp
from oracle import pyoracle as O
from test_oracle_lusgs import incidences

pytestmark = pytest.mark.gpu

TOL = 1e-12


def mild_state(n, dim, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.85, 1.2, n)
    p = rng.uniform(0.85, 1.2, n)
    vel = rng.uniform(-0.4, 0.4, (n, dim))
    return cons(rho, vel, p)


def inlet(dim):
    return list(cons(np.array([1.2]), np.array([[0.6] + [0.1] * (dim - 1)]), np.array([1.1]))[0])


CASES = {
    # name: (mesh, scheme, order, limiter)
    "tri_roe": (lambda: M.Mesh.gen_tri(40, 30, 0, 4, 0, 2, 0.1, 7,
                                       (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY)), M.ROE, 1, M.LIM_NONE),
    "periodic_ausm2": (lambda: M.Mesh.gen_tri(30, 30, -10, 10, -10, 10, 0.1, 3, (M.BC_PERIODIC,) * 4), M.AUSM, 2,
                       M.LIM_VANLEER),
    "tet_ausm": (lambda: M.Mesh.gen_tet(6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11,
                                        (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY,
                                         M.BC_SYMMETRY)), M.AUSM, 1, M.LIM_NONE),
    "tet_roe2": (lambda: M.Mesh.gen_tet(5, 4, 4, (0, 0, 0), (1, 0.8, 0.7), 0.1, 5,
                                        (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY,
                                         M.BC_SYMMETRY)), M.ROE, 2, M.LIM_MINMOD),
    # 3D Roe with the residual computed in the forward sweeps, and a 3D periodic box (outlet and
    # periodic faces are where a per-slot loop over face kinds once compiled wrongly)
    "tet_roe1": (lambda: M.Mesh.gen_tet(6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.15, 12,
                                        (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY,
                                         M.BC_SYMMETRY)), M.ROE, 1, M.LIM_NONE),
    "tet_periodic_ausm1": (lambda: M.Mesh.from_raw(periodic_in_x(M.Mesh.gen_tet(
        6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.0, 13, (M.BC_WALL,) * 6).raw())), M.AUSM, 1, M.LIM_NONE),
    # four-slot 2D cells (two colours)
    "quad_roe1": (lambda: M.Mesh.from_raw(quad_grid(24, bc=M.BC_WALL, h=0.05, jitter=0.2, seed=4)), M.ROE, 1,
                  M.LIM_NONE),
    # an unstructured Fluent mesh of the reference (not bipartite: three colours after the Kempe pass)
    "stair_un_ausm1": (lambda: M.Mesh.from_raw(fixture("2d-stair-un-5-tri")), M.AUSM, 1, M.LIM_NONE),
    "stair_un_roe2": (lambda: M.Mesh.from_raw(fixture("2d-stair-un-5-tri")), M.ROE, 2, M.LIM_VANLEER),
}


def fixture(name):
    d = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"msh_{name}.npz")))
    for k in ("dim", "n_cells", "n_int"):
        d[k] = int(d[k])
    return d


@pytest.mark.parametrize("name", sorted(CASES))
def test_lusgs_iterations_match_oracle(name):
    mk, scheme, order, lim = CASES[name]
    m = mk()
    raw = m.raw()
    dim = int(raw["dim"])
    nc = int(raw["n_cells"])
    kw = dict(dim=dim, scheme=scheme, order=order, limiter=lim, q_inlet=inlet(dim))
    cfg, ocfg = M.make_config(**kw), O.make_config(**kw)
    print(f"case {name}: {nc} cells", flush=True)
    s = M.Solver(m, cfg)
    om = O.OracleMesh(raw)
    Q0 = mild_state(nc, dim, seed=3)

    # colours: proper, and the preprocessing's colouring (checked on the CPU in test_oracle_lusgs)
    col, ncol = s.lusgs_colours()
    if name.startswith("stair_un"):
        assert ncol == 3
    inc = incidences(raw, om.geometry())
    assert col.min() == 0 and col.max() == ncol - 1
    assert all(col[o] != col[c] for c in range(nc) for (_, o, _) in inc[c] if o >= 0 and o != c)
    assert np.array_equal(col, M.Prep(m).cell_colours()[0])

    for cfl, omega in ((1e4, 1.0), (3.0, 1.5)):
        s.set_state(Q0)
        s.set_steady_driver(M.STEADY_LUSGS, cfl, omega)
        Qr = Q0.copy()
        for it in range(4):
            s.lusgs_step(1)
            Qr, res = om.lusgs_step(ocfg, Qr, col, cfl, omega)
            Qg = s.get_state()
            err = rel_state_error(Qg, Qr)
            assert err <= TOL, (name, cfl, it, err)
            rg = s.residual_norms()
            assert np.allclose(rg, res, rtol=1e-12, atol=0), (name, it, rg, res)
        # one call of four iterations: only its last one reduces norms, and at order 1 the others
        # compute R inside the forward sweep, colour 0's backward sweep of each fused with the next
        # one's forward sweep; the state must equal the one-by-one run bit for bit
        Qg = s.get_state()
        s.set_state(Q0)
        s.lusgs_step(4)
        assert np.array_equal(s.get_state(), Qg), name
        assert np.allclose(s.residual_norms(), res, rtol=1e-12, atol=0)
    s.close()


def test_lusgs_ramp_converges_to_the_explicit_steady_state():
    mach, theta = 2.5, 10.0
    m = ramp.mesh(theta, 60, 36)
    q0 = ramp.freestream(m, mach)
    out = {}
    for driver in (M.STEADY_EXPLICIT, M.STEADY_LUSGS):
        s = M.Solver(m, ramp.config(m, mach))
        s.set_state(q0)
        s.set_steady_driver(driver, 1e4, 1.0)
        steps, drop, _ = s.run_to_steady(1e-10, 20000, 50)
        print(f"driver {driver}: {steps} iterations, drop {drop:.2e}", flush=True)
        assert drop <= 1e-10
        out[driver] = (steps, s.get_state())
        s.close()
    assert out[M.STEADY_EXPLICIT][0] >= 3 * out[M.STEADY_LUSGS][0]
    Qe, Ql = out[M.STEADY_EXPLICIT][1], out[M.STEADY_LUSGS][1]
    assert np.abs(Qe - Ql).max() <= 1e-7 * np.abs(Qe).max()


def test_lusgs_rejects_bad_settings():
    m = M.Mesh.gen_tri(8, 6, 0, 1, 0, 1, 0.1, 1, (M.BC_WALL,) * 4)
    s = M.Solver(m, M.make_config())
    for args in ((M.STEADY_LUSGS, 0.0, 1.0), (M.STEADY_LUSGS, 5.0, 0.0), (M.STEADY_LUSGS, float("inf"), 1.0),
                 (7, 5.0, 1.0)):
        with pytest.raises(M.MstError):
            s.set_steady_driver(*args)
    with pytest.raises(M.MstError):
        s.lusgs_step(-1)
    s.close()


def test_lusgs_mixed_precision_follows_fp64():
    """Mixed form: the residual comes from the FP32 flux kernels (k_flux + k_update), the sweeps stay
    FP64.  After 20 iterations the state stays within the mixed form's rounding of the FP64 one."""
    m = M.Mesh.gen_tri(60, 40, 0, 3, 0, 2, 0.1, 9, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY))
    nc = m.n_cells
    Q0 = mild_state(nc, 2, seed=8)
    out = {}
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        s = M.Solver(m, M.make_config(scheme=M.ROE, order=2, limiter=M.LIM_MINMOD, q_inlet=inlet(2), precision=prec))
        s.set_state(Q0)
        s.set_steady_driver(M.STEADY_LUSGS, 50.0, 1.0)
        s.lusgs_step(20)
        out[prec] = (s.get_state(), s.residual_norms())
        s.close()
    (Qd, rd), (Qm, rm) = out[M.PREC_FP64], out[M.PREC_MIXED]
    assert np.abs(Qm - Qd).max() <= 1e-5 * np.abs(Qd).max()
    assert np.allclose(rm, rd, rtol=1e-4)
