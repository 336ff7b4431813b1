"""GPU physics checks at the workloads' real sizes: exact solutions, invariants, failure detection."""
import math

import numpy as np
import pytest

import mstcfd as M
from exact import sod_exact, vortex_exact
from helpers import cons, rel_state_error, sod_state
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def advance_to(s, t_end, cfl):
    """Time-accurate march: dt from the device CFL reduction, last step clipped to t_end."""
    cfg = s.cfg
    fixed = M.make_config(scheme=cfg.scheme, order=cfg.order, limiter=cfg.limiter, rk=cfg.rk, compat=cfg.compat,
                          cfl=0.0, q_inlet=list(cfg.q_inlet))
    s.set_config(fixed)
    t, n = 0.0, 0
    while t < t_end - 1e-12:
        dt = min(s.compute_dt(cfl), t_end - t)
        s.set_dt(dt)
        s.step(1)
        t += dt
        n += 1
    return n


def test_sod_c1_matches_exact_and_oracle():
    """C1: [0,1]x[0,0.01], 2000x2 quads -> 8000 jittered triangles, Roe + minmod, RK2, CFL 0.5, t = 0.2.

    GPU and oracle march in lockstep with the same dt.  Short horizons agree to rounding; over the
    ~5000 steps of the full run the limiter's switching at the shock amplifies rounding, so the end
    states are compared against a bound two orders below the discretisation error."""
    m = M.Mesh.gen_tri(2000, 2, 0, 1, 0, 0.01, 0.1, 12345, (M.BC_WALL,) * 4)
    kw = dict(scheme=M.ROE, order=2, limiter=M.LIM_MINMOD, rk=2)
    s = M.Solver(m, M.make_config(**kw))
    geo = s.prep.geometry()
    Q0 = sod_state(geo["ccen"])
    s.set_state(Q0)
    om = O.OracleMesh(m.raw())
    co = O.make_config(cfl=0.5, **kw)
    Qr, t, n = Q0.copy(), 0.0, 0
    errs = {}
    while t < 0.2 - 1e-12:
        dt = min(om.cfl_dt(co, Qr), 0.2 - t)
        Qr, _ = om.rk_step(co, Qr, dt)
        s.set_dt(dt)
        s.step(1)
        t += dt
        n += 1
        if n in (1, 10, 100, 1000):
            errs[n] = rel_state_error(s.get_state(), Qr)
    Q = s.get_state()
    errs[n] = rel_state_error(Q, Qr)
    ex = sod_exact(geo["ccen"][:, 0], 0.2)
    l1 = np.abs(Q[:, 0] - ex[:, 0]).mean()
    print(f"Sod C1: {n} steps, L1(rho) vs exact = {l1:.3e}; GPU vs oracle by step: "
          + ", ".join(f"{k}: {v:.1e}" for k, v in errs.items()), flush=True)
    assert l1 < 3e-3
    assert errs[1] <= 1e-14 and errs[10] <= 1e-13
    assert errs[n] <= 1e-2 * l1


@pytest.mark.parametrize("n", [250, 500, 1000])
def test_isentropic_vortex_one_period(n, record={}):
    """C2 construction (n=1000 is the full 2M-cell case): Roe + van Leer, RK3, one convection period;
    the exact solution is the initial field."""
    m = M.Mesh.gen_tri(n, n, -10, 10, -10, 10, 0.1, 12345, (M.BC_PERIODIC,) * 4)
    s = M.Solver(m, M.make_config(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER, rk=3))
    geo = s.prep.geometry()
    Q0 = vortex_exact(geo["ccen"])
    s.set_state(Q0)
    period = 20.0 / (0.5 * math.sqrt(1.4))
    steps = advance_to(s, period, 0.5)
    Q = s.get_state()
    V = geo["vol"]
    l2 = math.sqrt((V * (Q[:, 0] - Q0[:, 0]) ** 2).sum() / V.sum())
    mass = abs((V * Q[:, 0]).sum() - (V * Q0[:, 0]).sum()) / (V * Q0[:, 0]).sum()
    record[n] = l2
    print(f"vortex {2 * n * n} cells: {steps} steps, L2(rho) = {l2:.3e}, mass drift {mass:.1e}", flush=True)
    assert mass < 1e-11  # ~4e4 steps of sums over up to 2M cells
    assert l2 < 0.05
    if n // 2 in record:
        # Green-Gauss with the reference's distance-weighted face values is not second-order
        # consistent on jittered meshes, so the observed order lies between 1 and 2
        order = math.log(record[n // 2] / l2, 2)
        print(f"vortex observed order {order:.2f}", flush=True)
        assert order > 0.9


def test_full_size_invariants():
    """C2 mesh (2M cells): freestream preserved, periodic conservation exact to roundoff."""
    m = M.Mesh.gen_tri(1000, 1000, -10, 10, -10, 10, 0.1, 12345, (M.BC_PERIODIC,) * 4)
    for scheme in (M.ROE, M.AUSM):
        s = M.Solver(m, M.make_config(scheme=scheme, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.5))
        q = cons([1.0], [[0.7, -0.2]], [1.0])[0]
        s.set_state(np.tile(q, (m.n_cells, 1)))
        R = s.residual()
        assert np.abs(R).max() <= 1e-14 * 3.0 * 0.02
        geo = s.prep.geometry()
        Q0 = vortex_exact(geo["ccen"])
        s.set_state(Q0)
        s.step(10)
        Q = s.get_state()
        V = geo["vol"]
        tot0, tot = (V[:, None] * Q0).sum(0), (V[:, None] * Q).sum(0)
        assert (np.abs(tot - tot0) <= 1e-12 * np.abs(V[:, None] * Q0).sum(0)).all()
        s.close()


def test_nonphysical_state_is_reported():
    """The update kernel flags rho <= 0, p <= 0 or NaN; the C ABI returns MST_E_STATE."""
    m = M.Mesh.gen_tri(200, 2, 0, 1, 0, 0.01, 0.1, 12345, (M.BC_WALL,) * 4)
    for bad in ("nan", "negative_p"):
        s = M.Solver(m, M.make_config(order=2, limiter=M.LIM_MINMOD, rk=2, cfl=0.5))
        Q = sod_state(s.prep.geometry()["ccen"])
        if bad == "nan":
            Q[17, 3] = np.nan
        else:
            Q[17, 3] = 0.5 * (Q[17, 1] ** 2 + Q[17, 2] ** 2) / Q[17, 0] - 1.0
        s.set_state(Q)
        with pytest.raises(M.MstError, match="non-physical"):
            s.step(1)
        # a clean state afterwards runs again: the flag is cleared once reported
        s.set_state(sod_state(s.prep.geometry()["ccen"]))
        s.step(2)
        assert np.isfinite(s.get_state()).all()
