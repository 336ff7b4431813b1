"""Mixed-precision form: FP32 gradient/limiter/flux arithmetic, FP32 face fluxes, compensated FP64
gather and FP64 update.  Checked against the FP64 oracle and the plain FP64 GPU path."""
import math

import numpy as np
import pytest

import mstcfd as M
from exact import vortex_exact
from helpers import cons, flux_scale, normal_sums, residual_parity_error, smooth_state
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def pulse_state(ccen, amp=0.1, sigma=0.1, mach=0.8, gamma=1.4):
    """3D Gaussian density/pressure pulse at the box centre on a Mach-0.8 x-flow (rho=1, p=1/gamma)."""
    r2 = ((ccen - 0.5) ** 2).sum(1)
    b = amp * np.exp(-r2 / (2 * sigma ** 2))
    rho = 1 + b
    p = (1 + b) / gamma
    vel = np.zeros((ccen.shape[0], 3))
    vel[:, 0] = mach
    return cons(rho, vel, p)


def field_errors(Qm, Qd):
    """relative L2 (||dQ_k|| / ||Q_k||) and Linf (max|dQ_k| / max|Q_k|) per field; the momentum
    components share the norm of the momentum field"""
    nv = Qd.shape[1]
    d = Qm - Qd
    n2 = np.sqrt((Qd ** 2).sum(0))
    n2[1:nv - 1] = np.sqrt((Qd[:, 1:nv - 1] ** 2).sum())
    ninf = np.abs(Qd).max(0)
    ninf[1:nv - 1] = ninf[1:nv - 1].max()
    return np.sqrt((d ** 2).sum(0)) / n2, np.abs(d).max(0) / ninf


@pytest.mark.parametrize("scheme,order,lim", [(M.ROE, 2, M.LIM_VANLEER), (M.AUSM, 2, M.LIM_MINMOD), (M.ROE, 1, 0)])
def test_mixed_residual_close_to_oracle(scheme, order, lim):
    m = M.Mesh.gen_tri(60, 50, -10, 10, -10, 10, 0.1, 3, (M.BC_PERIODIC,) * 4)
    om = O.OracleMesh(m.raw())
    geo = om.geometry()
    Q = smooth_state(geo["ccen"], seed=3, mach=0.8)
    kw = dict(scheme=scheme, order=order, limiter=lim)
    s = M.Solver(m, M.make_config(precision=M.PREC_MIXED, **kw))
    s.set_state(Q)
    Rg = s.residual()
    Rr, Sr = om.residual(O.make_config(**kw), Q)
    err = residual_parity_error(Rg, Rr, Sr, flux_scale(Q) * normal_sums(m.raw(), geo["S"])[:, None])
    print(f"mixed residual scheme={scheme} order={order} lim={lim}: err/S = {err:.2e}", flush=True)
    assert 1e-9 < err <= 2e-6  # FP32 flux arithmetic: well above FP64 rounding, far below 1e-5


@pytest.mark.parametrize("scheme,lim", [(M.ROE, M.LIM_MINMOD), (M.AUSM, M.LIM_VANLEER)])
def test_mixed_3d_residual_close_to_oracle(scheme, lim):
    """Tetrahedra with every boundary kind, order 2: the FP32 3D gradient reads the difference-form
    weights folded into its normals, (float)(w n), against the oracle's FP64 eta-weighted face sums."""
    bc = (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY, M.BC_SYMMETRY)
    m = M.Mesh.gen_tet(10, 8, 6, (0, 0, 0), (1, 0.8, 0.7), 0.15, 21, bc)
    om = O.OracleMesh(m.raw())
    geo = om.geometry()
    Q = smooth_state(geo["ccen"], seed=5, mach=0.7)
    qin = list(cons([1.2], [[0.6, 0.1, 0.1]], [1.1])[0])
    kw = dict(scheme=scheme, order=2, limiter=lim, q_inlet=qin, dim=3)
    s = M.Solver(m, M.make_config(precision=M.PREC_MIXED, **kw))
    s.set_state(Q)
    Rg = s.residual()
    Rr, Sr = om.residual(O.make_config(**kw), Q)
    err = residual_parity_error(Rg, Rr, Sr, flux_scale(Q) * normal_sums(m.raw(), geo["S"])[:, None])
    print(f"mixed 3D residual scheme={scheme} lim={lim}: err/S = {err:.2e}", flush=True)
    assert 1e-9 < err <= 2e-6


def test_mixed_conserves_exactly_and_tracks_fp64():
    """Each FP32 face flux enters both cells with opposite signs in FP64, so periodic totals stay
    conserved to FP64 rounding; the state follows the FP64 run to FP32 accuracy."""
    m = M.Mesh.gen_tri(200, 200, -10, 10, -10, 10, 0.1, 12345, (M.BC_PERIODIC,) * 4)
    kw = dict(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER, rk=3)
    s64 = M.Solver(m, M.make_config(**kw))
    s32 = M.Solver(m, M.make_config(precision=M.PREC_MIXED, **kw))
    geo = s64.prep.geometry()
    Q0 = vortex_exact(geo["ccen"])
    for s in (s64, s32):
        s.set_state(Q0)
    for _ in range(50):
        dt = s64.compute_dt(0.5)
        for s in (s64, s32):
            s.set_dt(dt)
            s.step(1)
    Qd, Qm = s64.get_state(), s32.get_state()
    V = geo["vol"][:, None]
    tot0, tot = (V * Q0).sum(0), (V * Qm).sum(0)
    assert (np.abs(tot - tot0) <= 1e-13 * np.abs(V * Q0).sum(0)).all()
    l2, linf = field_errors(Qm, Qd)
    print(f"vortex 80k cells, 50 RK3 steps: mixed vs FP64 rel L2 {l2}, Linf {linf}", flush=True)
    assert l2[0] <= 1e-6 and linf[0] < 2e-6


def test_mixed_pulse3d_density_error():
    """C3 construction at reduced size (20^3 hexes -> 48,000 tets): AUSM + minmod, RK2, 200 steps."""
    bc = (M.BC_INLET, M.BC_OUTLET) + (M.BC_SYMMETRY,) * 4
    m = M.Mesh.gen_tet(20, 20, 20, (0, 0, 0), (1, 1, 1), 0.15, 12345, bc)
    qin = list(cons([1.0], [[0.8, 0, 0]], [1 / 1.4])[0])
    kw = dict(scheme=M.AUSM, order=2, limiter=M.LIM_MINMOD, rk=2, q_inlet=qin, dim=3)
    s64 = M.Solver(m, M.make_config(**kw))
    s32 = M.Solver(m, M.make_config(precision=M.PREC_MIXED, **kw))
    Q0 = pulse_state(s64.prep.geometry()["ccen"])
    for s in (s64, s32):
        s.set_state(Q0)
    for _ in range(200):
        dt = s64.compute_dt(0.5)
        for s in (s64, s32):
            s.set_dt(dt)
            s.step(1)
    l2, linf = field_errors(s32.get_state(), s64.get_state())
    print(f"pulse3d 48k tets, 200 RK2 steps: mixed vs FP64 rel L2 {l2}, Linf {linf}", flush=True)
    # relative density error target (L2 norm); the pointwise error stays within a few FP32 ulps x steps
    assert l2[0] <= 1e-6
    assert linf[0] <= 2e-5


@pytest.mark.parametrize("dim,rk", [(2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_mixed_step_is_the_ssp_combination_of_its_residuals(dim, rk):
    """One mixed step equals the SSP-RK combination of the mixed residuals of its stage states,
    Q_{s+1} = A_s Q^n + B_s (Q_s - dt/V R(Q_s)), to FP64 rounding: the intermediate stages store the
    next stage's A Q^n + B Q_s in place of their state, which must not change what a stage computes.
    A wrong coefficient or a Q^n taken from the wrong buffer moves the state by O(dt R)."""
    if dim == 2:
        m = M.Mesh.gen_tri(40, 40, -10, 10, -10, 10, 0.1, 7, (M.BC_PERIODIC,) * 4)
        kw = dict(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER, rk=rk)
    else:
        bc = (M.BC_INLET, M.BC_OUTLET) + (M.BC_SYMMETRY,) * 4
        m = M.Mesh.gen_tet(8, 8, 8, (0, 0, 0), (1, 1, 1), 0.15, 12345, bc)
        qin = list(cons([1.0], [[0.8, 0, 0]], [1 / 1.4])[0])
        kw = dict(scheme=M.AUSM, order=2, limiter=M.LIM_MINMOD, rk=rk, q_inlet=qin, dim=3)
    s = M.Solver(m, M.make_config(precision=M.PREC_MIXED, **kw))
    geo = s.prep.geometry()
    Q0 = vortex_exact(geo["ccen"]) if dim == 2 else pulse_state(geo["ccen"])
    dtv = None
    A, B = {1: ([0.0], [1.0]), 2: ([0.0, 0.5], [1.0, 0.5]), 3: ([0.0, 0.75, 1 / 3], [1.0, 0.25, 2 / 3])}[rk]
    Q = Q0
    for st in range(rk):
        s.set_state(Q)
        if dtv is None:
            dt = s.compute_dt(0.5)
            dtv = dt / geo["vol"][:, None]
        R = s.residual()
        Q = A[st] * Q0 + B[st] * (Q - dtv * R)
    s.set_state(Q0)
    s.set_dt(dt)
    s.step(1)
    Qg = s.get_state()
    scale = np.abs(Q0).max(0)
    scale[1:-1] = scale[1:-1].max()  # the momentum components share the momentum scale
    err = (np.abs(Qg - Q).max(0) / scale).max()
    moved = (np.abs(Q - Q0).max(0) / scale).max()
    print(f"mixed dim={dim} rk={rk}: step vs SSP combination {err:.2e} (the step moves the state {moved:.2e})",
          flush=True)
    assert moved > 1e-4
    assert err <= 1e-14


def test_mixed_freestream_and_switching():
    m = M.Mesh.gen_tri(40, 40, -1, 1, -1, 1, 0.2, 4, (M.BC_PERIODIC,) * 4)
    s = M.Solver(m, M.make_config(order=2, limiter=M.LIM_MINMOD, precision=M.PREC_MIXED))
    q = cons([1.2], [[0.4, -0.3]], [0.9])[0]
    s.set_state(np.tile(q, (m.n_cells, 1)))
    R = s.residual()
    assert np.abs(R).max() <= 1e-6 * 2.0 * (2 / 40) * 3  # FP32 normals: closure to ~1e-7
    # switching the precision of a live context keeps the state and agrees with a fresh FP64 context
    s.set_config(M.make_config(order=2, limiter=M.LIM_MINMOD))
    R64 = s.residual()
    assert np.abs(R64).max() <= 1e-13
