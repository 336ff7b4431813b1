"""GPU parity: the HIP residual path (through the C ABI) against the CPU oracle.

Bar: per cell and field |R_gpu - R_ref| <= 1e-12 * S_ref with S = sum_f sum_d |n_d||F_d| (the
flux scale of the cell's own terms), and <= 1e-12 relative on updated states.
"""
import numpy as np
import pytest

import mstcfd as M
from helpers import (brick_grid, quad_grid, cons, flip_faces, flux_scale, jumpy_state, normal_sums, periodic_in_x,
                     rel_state_error, residual_parity_error, smooth_state, sod_state)
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
ROE, AUSM = M.ROE, M.AUSM


def both_cfg(dim=2, **kw):
    return M.make_config(dim=dim, **kw), O.make_config(dim=dim, **kw)


_MESHES = {}


def mesh(name):
    if name not in _MESHES:
        if name == "strip_wall":
            m = M.Mesh.gen_tri(200, 4, 0, 1, 0, 0.02, 0.1, 12345, (M.BC_WALL,) * 4)
        elif name == "allbc":
            m = M.Mesh.gen_tri(40, 30, 0, 4, 0, 2, 0.1, 7, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY))
        elif name == "periodic":
            m = M.Mesh.gen_tri(40, 40, -10, 10, -10, 10, 0.1, 3, (M.BC_PERIODIC,) * 4)
        elif name == "quad":  # 4 faces per cell: the 4-slot kernels
            m = M.Mesh.from_raw(quad_grid(40, bc=M.BC_WALL, h=0.05, jitter=0.2, seed=4))
        elif name == "tet":
            m = M.Mesh.gen_tet(6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11,
                               (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY, M.BC_SYMMETRY))
        elif name == "brick":  # six-faced cells: the 8-slot kernels
            m = M.Mesh.from_raw(brick_grid(24, 20, h=0.05, jitter=0.1, seed=5, bc=M.BC_WALL))
        elif name == "tet_periodic":  # periodic in x, walls and symmetry planes around
            bc = (M.BC_WALL, M.BC_WALL, M.BC_SYMMETRY, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY)
            box = M.Mesh.gen_tet(6, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.0, 11, bc)
            m = M.Mesh.from_raw(periodic_in_x(box.raw()))
        else:
            raise KeyError(name)
        _MESHES[name] = (m, O.OracleMesh(m.raw()), M.Prep(m).geometry())
    return _MESHES[name]


def state(name, kind, geo, dim, order=1):
    if kind == "sod":  # piecewise constant: quiescent regions where every face flux of a field is roundoff
        return sod_state(geo["ccen"], x_split=0.5 * (geo["ccen"][:, 0].min() + geo["ccen"][:, 0].max()))
    if kind == "smooth":
        return smooth_state(geo["ccen"], seed=5, mach=1.3)
    if order == 1:
        return jumpy_state(geo["vol"].shape[0], dim=dim, seed=9)
    # second order: cell-to-cell jumps small enough that reconstructed states stay physical
    return mild_jumps(geo["vol"].shape[0], dim, seed=9)


def mild_jumps(n, dim, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.85, 1.2, n)
    p = rng.uniform(0.85, 1.2, n)
    vel = rng.uniform(-0.4, 0.4, (n, dim))
    return cons(rho, vel, p)


def inlet(dim):
    q = cons(np.array([1.2]), np.array([[0.6] + [0.1] * (dim - 1)]), np.array([1.1]))[0]
    return list(q)


def test_riemann_batch_matches_oracle():
    rng = np.random.default_rng(1)
    for dim in (2, 3):
        n = 4000
        QL = jumpy_state(n, dim, seed=2)
        QR = jumpy_state(n, dim, seed=3)
        # near-equal pairs exercise the Harten fix and the subsonic AUSM branch
        QR[: n // 4] = QL[: n // 4] * (1 + 1e-3 * rng.standard_normal((n // 4, dim + 2)))
        dirs = rng.integers(0, dim, n).astype(np.int32)
        for scheme in (ROE, AUSM):
            for compat in (O.COMPAT_REFERENCE, 0):
                cg, co = both_cfg(dim=dim, scheme=scheme, compat=compat)
                Fg = M.riemann_flux(cg, QL, QR, dirs, dim=dim)
                fn = O.roe_flux if scheme == ROE else O.ausm_flux
                Fr = np.array([fn(QL[i], QR[i], int(dirs[i]), co, dim=dim) for i in range(n)])
                scale = np.abs(Fr).max(1) + np.abs(QL).max(1) + np.abs(QR).max(1)
                err = (np.abs(Fg - Fr).max(1) / scale).max()
                print(f"riemann dim={dim} scheme={scheme} compat={compat}: max scaled err {err:.2e}", flush=True)
                assert err <= TOL


CASES = [
    ("strip_wall", ROE, 1, 0), ("strip_wall", AUSM, 1, 0),
    ("allbc", ROE, 1, 0), ("allbc", AUSM, 1, 0),
    ("periodic", ROE, 1, 0), ("periodic", AUSM, 1, 0),
    ("strip_wall", ROE, 2, M.LIM_MINMOD), ("allbc", ROE, 2, M.LIM_NONE), ("allbc", AUSM, 2, M.LIM_MINMOD),
    ("allbc", ROE, 2, M.LIM_VANLEER), ("periodic", ROE, 2, M.LIM_VANLEER), ("periodic", AUSM, 2, M.LIM_NONE),
    ("tet", AUSM, 1, 0), ("tet", ROE, 1, 0), ("tet", AUSM, 2, M.LIM_MINMOD), ("tet", ROE, 2, M.LIM_VANLEER),
    ("tet_periodic", ROE, 1, 0), ("tet_periodic", AUSM, 2, M.LIM_MINMOD),
    ("brick", ROE, 1, 0), ("brick", AUSM, 2, M.LIM_VANLEER),
]


@pytest.mark.parametrize("mname,scheme,order,lim", CASES)
@pytest.mark.parametrize("kind", ["smooth", "jumpy", "sod"])
def test_residual_parity(mname, scheme, order, lim, kind):
    m, om, geo = mesh(mname)
    dim = m.dim
    cg, co = both_cfg(dim=dim, scheme=scheme, order=order, limiter=lim, q_inlet=inlet(dim))
    Q = state(mname, kind, geo, dim, order)
    s = M.Solver(m, cg)
    s.set_state(Q)
    Rg = s.residual()
    Rr, Sr = om.residual(co, Q)
    err = residual_parity_error(Rg, Rr, Sr, flux_scale(Q) * normal_sums(m.raw(), geo["S"])[:, None])
    print(f"{mname} scheme={scheme} order={order} lim={lim} {kind}: residual err/S = {err:.2e}", flush=True)
    assert err <= TOL
    s.close()


@pytest.mark.parametrize("mname,lim", [("allbc", M.LIM_MINMOD), ("periodic", M.LIM_VANLEER), ("tet", M.LIM_MINMOD),
                                       ("strip_wall", M.LIM_NONE), ("tet_periodic", M.LIM_VANLEER),
                                       ("brick", M.LIM_MINMOD)])
def test_gradient_and_limiter_parity(mname, lim):
    m, om, geo = mesh(mname)
    dim = m.dim
    cg, co = both_cfg(dim=dim, order=2, limiter=lim)
    Q = smooth_state(geo["ccen"], seed=4)
    s = M.Solver(m, cg)
    s.set_state(Q)
    gg, phg = s.gradient()
    gr = om.gradient(co, Q)
    phr = om.limiter(co, Q, gr)
    scale = np.abs(gr).max(axis=(0, 2), keepdims=True)  # per-field gradient scale
    gerr = float((np.abs(gg - gr) / scale).max())
    perr = float(np.abs(phg - phr).max())
    print(f"{mname} lim={lim}: grad err {gerr:.2e}, phi err {perr:.2e}", flush=True)
    assert gerr <= 1e-11
    assert perr <= 1e-9


STRATEGIES = [M.STRATEGY_FACE, M.STRATEGY_CELL, M.STRATEGY_PATCH, M.STRATEGY_FUSED]


def strategies_for(m, cg):
    """The strategies a mesh takes: the patch kernels hold cells of at most four faces, and the library
    refuses them for others (checked here)."""
    raw = m.raw()
    if np.bincount(np.concatenate([raw["c0"], raw["c1"][raw["c1"] >= 0]])).max() <= 4:
        return STRATEGIES
    s = M.Solver(m, cg)
    for st in (M.STRATEGY_PATCH, M.STRATEGY_FUSED):
        with pytest.raises(M.MstError, match="at most 4 faces per cell"):
            s.strategy = st
    return [M.STRATEGY_FACE, M.STRATEGY_CELL]


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("mname,scheme", [("allbc", ROE), ("strip_wall", AUSM), ("tet", AUSM)])
def test_euler_solve_matches_oracle(mname, scheme, strategy):
    """RhoSolver::solve: Q_new = Q_old - DT/V R, then updateNewToOld."""
    m, om, geo = mesh(mname)
    dim = m.dim
    cg, co = both_cfg(dim=dim, scheme=scheme, q_inlet=inlet(dim))
    Q = smooth_state(geo["ccen"], seed=2, mach=0.5)
    dt = 1.0 / 4000.0 if mname != "allbc" else 1e-3
    s = M.Solver(m, cg)
    s.strategy = strategy
    s.set_state(Q)
    s.set_dt(dt)
    s.solve()
    Qg = s.get_new_state()
    Qr = om.euler(co, Q, dt)
    assert rel_state_error(Qg, Qr) <= TOL
    s.update_new_to_old()
    assert np.array_equal(s.get_state(), Qg)


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("mname,scheme,order,lim,rk", [
    ("strip_wall", ROE, 2, M.LIM_MINMOD, 2), ("periodic", ROE, 2, M.LIM_VANLEER, 3),
    ("allbc", AUSM, 2, M.LIM_MINMOD, 2), ("tet", AUSM, 2, M.LIM_MINMOD, 2), ("allbc", ROE, 1, 0, 1),
    ("tet", ROE, 2, M.LIM_VANLEER, 3), ("quad", ROE, 2, M.LIM_VANLEER, 3), ("quad", AUSM, 2, M.LIM_MINMOD, 2)])
def test_rk_steps_match_oracle(mname, scheme, order, lim, rk, strategy):
    m, om, geo = mesh(mname)
    dim = m.dim
    cg, co = both_cfg(dim=dim, scheme=scheme, order=order, limiter=lim, rk=rk, cfl=0.4, q_inlet=inlet(dim))
    Q = smooth_state(geo["ccen"], seed=3, mach=0.6)
    s = M.Solver(m, cg)
    s.strategy = strategy
    s.set_state(Q)
    Qr = Q.copy()
    for it in range(10):
        dt_r = om.cfl_dt(co, Qr)
        dt_g = s.compute_dt(0.4)
        assert abs(dt_g - dt_r) <= 1e-13 * dt_r
        Qr, res = om.rk_step(co, Qr, dt_r)
        s.step(1)
        ng = s.residual_norms()
        assert np.allclose(ng, res, rtol=1e-10, atol=0)
    Qg = s.get_state()
    err = rel_state_error(Qg, Qr)
    print(f"{mname} rk{rk} strategy {strategy}: 10-step state err {err:.2e}", flush=True)
    assert err <= TOL


def test_branch2_orientation_compat():
    """Faces whose area vector points out of c0 (every bundled Fluent mesh): physical L/R gives the
    same residual as the unflipped mesh bit for bit; the literal reference flags match the oracle."""
    m, om, geo = mesh("allbc")
    raw2, sel = flip_faces(m.raw(), 0.5, seed=4)
    m2 = M.Mesh.from_raw(raw2)
    om2 = O.OracleMesh(raw2)
    Q = smooth_state(geo["ccen"], seed=7, mach=0.7)
    for scheme in (ROE, AUSM):
        cg, co = both_cfg(scheme=scheme, q_inlet=inlet(2))
        s1, s2 = M.Solver(m, cg), M.Solver(m2, cg)
        s1.set_state(Q)
        s2.set_state(Q)
        R1, R2 = s1.residual(), s2.residual()
        assert residual_parity_error(R2, *om2.residual(co, Q)) <= TOL
        assert np.abs(R1 - R2).max() <= 1e-12 * np.abs(R1).max()
        cg2, co2 = both_cfg(scheme=scheme, q_inlet=inlet(2), compat=O.COMPAT_REFERENCE | O.COMPAT_BRANCH2_LR)
        s2.set_config(cg2)
        Rg = s2.residual()
        Rr, Sr = om2.residual(co2, Q)
        assert residual_parity_error(Rg, Rr, Sr) <= TOL
        assert np.abs(Rg - R1).max() > 1e-6  # the literal flags really do differ


def test_monitor_maxrel_matches_reference_rule():
    m, om, geo = mesh("strip_wall")
    cg, co = both_cfg()
    Q = smooth_state(geo["ccen"], seed=8, mach=0.3)
    s = M.Solver(m, cg)
    s.set_state(Q)
    s.set_dt(1e-4)
    s.solve()
    got = s.monitor_maxrel()
    ref = O.monitor_maxrel(s.get_new_state(), s.get_state())
    assert np.allclose(got, ref, rtol=0, atol=0)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
@pytest.mark.parametrize("mname,scheme,lim", [("periodic", ROE, M.LIM_VANLEER), ("allbc", AUSM, M.LIM_MINMOD),
                                              ("tet", ROE, M.LIM_MINMOD), ("quad", AUSM, M.LIM_VANLEER),
                                              ("tet_periodic", AUSM, M.LIM_MINMOD), ("brick", ROE, M.LIM_MINMOD)])
def test_cell_strategy_matches_face_strategy(mname, scheme, lim, precision):
    """The fused cell kernel solves each interior face twice, the patch kernel each face of a patch
    once (a border face in both patches), always from the same memory inputs with the same
    instructions, so both reproduce the face split bit for bit (same fluxes, same summation order)."""
    m, om, geo = mesh(mname)
    dim = m.dim
    cg = M.make_config(dim=dim, scheme=scheme, order=2, limiter=lim, rk=3, cfl=0.45, q_inlet=inlet(dim),
                       precision=precision)
    Q = smooth_state(geo["ccen"], seed=6, mach=0.7)
    out = []
    strats = strategies_for(m, cg)
    for strategy in strats:
        s = M.Solver(m, cg)
        s.strategy = strategy
        assert s.strategy == strategy
        s.set_state(Q)
        s.step(20)
        out.append((s.get_state(), s.residual_norms()))
    for k in range(1, len(strats)):
        err = rel_state_error(out[k][0], out[0][0])
        print(f"{mname} prec={precision}: strategy {strats[k]} vs face after 20 RK3 steps {err:.2e}",
              flush=True)
        assert np.array_equal(out[k][0], out[0][0])
        assert np.array_equal(out[k][1], out[0][1])


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
@pytest.mark.parametrize("mname,scheme", [("periodic", ROE), ("allbc", AUSM), ("tet", ROE), ("quad", AUSM),
                                          ("tet_periodic", ROE), ("brick", AUSM)])
def test_first_order_strategies_match(mname, scheme, precision):
    """Order 1: the patch kernel stages the cell states of its faces in LDS; every strategy still
    reproduces the face split bit for bit."""
    m, om, geo = mesh(mname)
    dim = m.dim
    cg = M.make_config(dim=dim, scheme=scheme, order=1, rk=2, cfl=0.45, q_inlet=inlet(dim), precision=precision)
    Q = smooth_state(geo["ccen"], seed=8, mach=0.7)
    out = []
    strats = strategies_for(m, cg)
    for strategy in strats:
        s = M.Solver(m, cg)
        s.strategy = strategy
        s.set_state(Q)
        s.step(20)
        out.append((s.get_state(), s.residual_norms()))
    for k in range(1, len(strats)):
        assert np.array_equal(out[k][0], out[0][0]), (mname, precision, strats[k])
        assert np.array_equal(out[k][1], out[0][1])


def test_fused_follows_reconfiguration():
    """The fused kernel reads its halo cells' geometry from a copy packed in halo order.  A
    configuration change that moves the incidence weights (the last-interior-face quirk) or switches
    the precision refreshes that copy, so a reconfigured context steps exactly as a fresh one."""
    m, om, geo = mesh("allbc")
    Q = smooth_state(geo["ccen"], seed=6, mach=0.7)
    base = dict(scheme=AUSM, order=2, limiter=M.LIM_MINMOD, rk=3, cfl=0.45, q_inlet=inlet(2))
    cfgs = [M.make_config(compat=0, **base), M.make_config(**base), M.make_config(precision=M.PREC_MIXED, **base),
            M.make_config(compat=0, precision=M.PREC_MIXED, **base), M.make_config(compat=0, **base)]
    s = M.Solver(m, cfgs[0])
    s.strategy = M.STRATEGY_FUSED
    prev = None
    for cg in cfgs[1:]:
        s.set_config(cg)
        s.set_state(Q)
        s.step(5)
        f = M.Solver(m, cg)
        f.strategy = M.STRATEGY_FUSED
        f.set_state(Q)
        f.step(5)
        got = s.get_state()
        assert np.array_equal(got, f.get_state())
        if prev is not None:
            assert not np.array_equal(got, prev)
        prev = got
    # AUTO frees the copy when it keeps another strategy; fused again gathers it anew
    s.strategy = M.STRATEGY_AUTO
    s.set_state(Q)
    s.step(1)
    s.strategy = M.STRATEGY_FUSED
    s.set_state(Q)
    s.step(5)
    assert np.array_equal(s.get_state(), prev)


def test_auto_strategy_resolves_and_matches():
    """AUTO times both strategies on the first stage (outputs to a spare buffer) and keeps one; the
    run equals an explicitly chosen strategy bit for bit."""
    m, om, geo = mesh("periodic")
    cg = M.make_config(scheme=ROE, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.45)
    Q = smooth_state(geo["ccen"], seed=6, mach=0.7)
    s = M.Solver(m, cg)
    assert s.strategy == M.STRATEGY_AUTO
    s.set_state(Q)
    s.step(5)
    assert s.strategy in STRATEGIES
    f = M.Solver(m, cg)
    f.strategy = M.STRATEGY_FACE
    f.set_state(Q)
    f.step(5)
    assert np.array_equal(s.get_state(), f.get_state())
    # a configuration change measures again
    s.set_config(M.make_config(scheme=AUSM, order=2, limiter=M.LIM_MINMOD, rk=2, cfl=0.4))
    assert s.strategy == M.STRATEGY_AUTO


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_kernel_timing_counts_and_leaves_results_unchanged(strategy):
    """Per-kernel timing (events carried by the dispatches) counts one interval per launch, each
    positive, their sum below the wall time of the steps, and changes no result bit."""
    import time
    m, om, geo = mesh("periodic")
    cg = M.make_config(scheme=ROE, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.45)
    Q = smooth_state(geo["ccen"], seed=8, mach=0.6)
    plain = M.Solver(m, cg)
    plain.strategy = strategy
    plain.set_state(Q)
    plain.step(4)
    s = M.Solver(m, cg)
    s.strategy = strategy
    s.set_state(Q)
    s.sync()
    s.timing_reset()
    s.timing(True)
    t0 = time.perf_counter()
    s.step(4)
    s.sync()
    wall_ms = 1e3 * (time.perf_counter() - t0)
    s.timing(False)
    assert np.array_equal(s.get_state(), plain.get_state())
    face, cell, patch, fused = (strategy == st for st in STRATEGIES)
    want = {M.K_GRAD: 12 * (not fused), M.K_FLUX: 12 * face, M.K_UPDATE: 12 * face, M.K_CELL: 12 * cell,
            M.K_PATCH: 12 * patch, M.K_FUSED: 12 * fused, M.K_DT: 1}  # later steps: dt from the last stage
    total = 0.0
    for k, n_want in want.items():
        n, ms = s.timing_get(k)
        assert n == n_want, (M.KERNEL_NAMES[k], n)
        assert (ms > 0) == (n > 0), (M.KERNEL_NAMES[k], ms)
        total += ms
    assert total < wall_ms
    s.timing_reset()
    assert s.timing_get(M.K_GRAD) == (0, 0.0)
    # sampled: one launch in 5 of each class (counted from the enable), the same results
    s.set_state(Q)
    s.timing(5)
    s.step(4)
    s.sync()
    s.timing(False)
    assert np.array_equal(s.get_state(), plain.get_state())
    for k, n_want in want.items():
        n, ms = s.timing_get(k)
        assert n == (n_want + 4) // 5, (M.KERNEL_NAMES[k], n)
        assert (ms > 0) == (n > 0)
    with pytest.raises(M.MstError, match="timing"):
        s.timing(-1)


@pytest.mark.parametrize("mname", ["allbc", "tet"])
def test_rcm_cell_order_all_strategies_match_oracle(mname):
    """Under the reverse Cuthill-McKee order the patches are strips, not compact: every strategy must
    still match the oracle, and the strategies one another bit for bit."""
    old = M.get_cell_order()
    M.set_cell_order(M.ORDER_RCM)
    try:
        m, om, geo = mesh(mname)
        dim = m.dim
        cg, co = both_cfg(dim=dim, scheme=ROE, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.4, q_inlet=inlet(dim))
        Q = smooth_state(geo["ccen"], seed=12, mach=0.6)
        Qr = Q.copy()
        for _ in range(5):
            Qr, _res = om.rk_step(co, Qr, om.cfl_dt(co, Qr))
        out = []
        for strategy in STRATEGIES:
            s = M.Solver(m, cg)
            s.strategy = strategy
            s.set_state(Q)
            s.step(5)
            out.append(s.get_state())
            s.close()
    finally:
        M.set_cell_order(old)
    assert rel_state_error(out[0], Qr) <= TOL
    for k in range(1, len(out)):
        assert np.array_equal(out[k], out[0]), STRATEGIES[k]


@pytest.mark.parametrize("dim", [2, 3])
def test_tiny_meshes_all_strategies_match_oracle(dim):
    """One partial patch of a few cells (4 triangles, 6 tetrahedra), walls or inlet/outlet around:
    tiny face lists and halos, most faces on the boundary."""
    if dim == 2:
        m = M.Mesh.gen_tri(2, 1, 0, 1, 0, 0.5, 0.1, 21, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY))
    else:
        m = M.Mesh.gen_tet(1, 1, 1, (0, 0, 0), (1, 1, 1), 0.1, 21,
                           (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY, M.BC_WALL, M.BC_WALL))
    om = O.OracleMesh(m.raw())
    geo = M.Prep(m).geometry()
    cg, co = both_cfg(dim=dim, scheme=ROE, order=2, limiter=M.LIM_MINMOD, rk=2, cfl=0.4, q_inlet=inlet(dim))
    Q = smooth_state(geo["ccen"], seed=4, mach=0.5)
    Qr = Q.copy()
    for _ in range(3):
        Qr, _res = om.rk_step(co, Qr, om.cfl_dt(co, Qr))
    out = []
    for strategy in STRATEGIES:
        s = M.Solver(m, cg)
        s.strategy = strategy
        s.set_state(Q)
        s.step(3)
        out.append(s.get_state())
        s.close()
    print(f"dim {dim}: {m.n_cells} cells, err vs oracle {rel_state_error(out[0], Qr):.2e}", flush=True)
    assert rel_state_error(out[0], Qr) <= TOL
    for k in range(1, len(out)):
        assert np.array_equal(out[k], out[0]), STRATEGIES[k]
