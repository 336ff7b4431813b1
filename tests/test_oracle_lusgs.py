"""Oracle LU-SGS steady driver, pinned without the GPU.

The oracle's sweeps apply (D+L) D^-1 (D+U) matrix-free.  Here the same operator is assembled as a
dense matrix from its definition -- face spectral radii, the scalar diagonal and the Rusanov-split
off-diagonal blocks 1/2 (A_n(Q_j) - omega r_f I), with A_n the Jacobian of the textbook Euler flux
taken by complex-step differentiation -- and the linear system is solved directly.  For a residual
of size eps the sweeps must return that solution to O(eps): a wrong sign, a swapped L/U test, a
missing D^-1 between the sweeps or a normal of the wrong cell fails by O(1).
"""
import numpy as np
import pytest

import mstcfd as M
from helpers import GAMMA, jumpy_state
from mstcfd import ramp
from oracle import pyoracle as O


def incidences(raw, geo):
    """per cell: list of (file face, neighbour or -1, outward normal of the cell)"""
    c0, c1, ft, pa = raw["c0"], raw["c1"], raw["ftype"], raw["partner"]
    S = geo["S"] * geo["dac"][:, None]
    inc = [[] for _ in range(int(raw["n_cells"]))]
    for f in range(len(c0)):
        if c1[f] >= 0:
            inc[c0[f]].append((f, int(c1[f]), S[f]))
            inc[c1[f]].append((f, int(c0[f]), -S[f]))
        elif ft[f] == M.BC_PERIODIC and pa[f] >= 0:
            inc[c0[f]].append((f, int(c0[pa[f]]), S[f]))
        else:
            inc[c0[f]].append((f, -1, S[f]))
    return inc


def greedy_colours(inc, order=None, seed=None):
    n = len(inc)
    order = np.arange(n) if order is None else order
    col = -np.ones(n, np.int32)
    for c in order:
        used = {int(col[o]) for (_, o, _) in inc[c] if o >= 0 and o != c and col[o] >= 0}
        k = 0
        while k in used:
            k += 1
        col[c] = k
    return col


def euler_flux(q, n):
    rho, m, E = q[0], q[1:-1], q[-1]
    un = np.dot(m, n) / rho
    p = (GAMMA - 1) * (E - 0.5 * np.dot(m, m) / rho)
    return np.concatenate([[rho * un], m * un + p * n, [(E + p) * un]])


def jacobian(q, n):
    nv = q.size
    A = np.zeros((nv, nv))
    h = 1e-30
    for k in range(nv):
        z = q.astype(complex)
        z[k] += 1j * h
        A[:, k] = euler_flux(z, n).imag / h
    return A


def radius(q, n):
    rho, m, E = q[0], q[1:-1], q[-1]
    p = (GAMMA - 1) * (E - 0.5 * np.dot(m, m) / rho)
    return abs(np.dot(m, n) / rho) + np.sqrt(GAMMA * p / rho) * np.linalg.norm(n)


def dense_operator(raw, geo, Q, col, cfl, omega):
    inc = incidences(raw, geo)
    nc, nv = Q.shape
    rf = {}
    for c in range(nc):
        for (f, o, n) in inc[c]:
            r = radius(Q[c], n) if o < 0 else max(radius(Q[c], n), radius(Q[o], n))
            rf[(c, f)] = r
    Dm = np.zeros((nc * nv, nc * nv))
    L = np.zeros_like(Dm)
    U = np.zeros_like(Dm)
    for c in range(nc):
        d = (1 / cfl + omega / 2) * sum(rf[(c, f)] for (f, _, _) in inc[c])
        Dm[c * nv:(c + 1) * nv, c * nv:(c + 1) * nv] = d * np.eye(nv)
        for (f, o, n) in inc[c]:
            if o < 0 or o == c or col[o] == col[c]:
                continue
            blk = 0.5 * (jacobian(Q[o], n) - omega * rf[(c, f)] * np.eye(nv))
            tgt = L if col[o] < col[c] else U
            tgt[c * nv:(c + 1) * nv, o * nv:(o + 1) * nv] += blk
    return (Dm + L) @ np.linalg.inv(Dm) @ (Dm + U), inc


MESHES = {
    "allbc": lambda: M.Mesh.gen_tri(7, 5, 0, 4, 0, 2, 0.1, 7, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY)),
    "periodic": lambda: M.Mesh.gen_tri(6, 6, -1, 1, -1, 1, 0.1, 3, (M.BC_PERIODIC,) * 4),
    "tet": lambda: M.Mesh.gen_tet(3, 2, 2, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11,
                                  (M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_WALL, M.BC_SYMMETRY, M.BC_SYMMETRY)),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("cfl,omega", [(5.0, 1.0), (1e4, 1.5)])
def test_sweeps_solve_the_factorised_system(name, cfl, omega):
    m = MESHES[name]()
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    dim = int(raw["dim"])
    nc, nv = int(raw["n_cells"]), dim + 2
    Q = jumpy_state(nc, dim=dim, seed=4, vmax=0.8)
    cfg = O.make_config(dim=dim)
    rng = np.random.default_rng(2)
    for colouring in ("file", "shuffled"):
        inc = incidences(raw, geo)
        order = np.arange(nc) if colouring == "file" else rng.permutation(nc)
        col = greedy_colours(inc, order)
        # a proper colouring: no two neighbours alike
        assert all(col[o] != col[c] for c in range(nc) for (_, o, _) in inc[c] if o >= 0 and o != c)
        A, _ = dense_operator(raw, geo, Q, col, cfl, omega)
        eps = 1e-8
        R = eps * rng.standard_normal((nc, nv))
        x = np.linalg.solve(A, -R.reshape(-1)).reshape(nc, nv)
        dQ = om.lusgs_sweep(cfg, Q, R, col, cfl, omega)
        assert np.abs(dQ - x).max() <= 1e-5 * np.abs(x).max()


def test_small_cfl_is_a_local_time_step():
    """CFL -> 0: D dominates and dQ -> -CFL V_i/sum_f r_f * R_i/V_i (explicit local time stepping)."""
    m = MESHES["allbc"]()
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    nc = int(raw["n_cells"])
    Q = jumpy_state(nc, dim=2, seed=6, vmax=0.8)
    inc = incidences(raw, geo)
    col = greedy_colours(inc)
    R, _ = om.residual(O.make_config(), Q)
    cfl = 1e-9
    dQ = om.lusgs_sweep(O.make_config(), Q, R, col, cfl, 1.0)
    rsum = np.array([sum(radius(Q[c], n) if o < 0 else max(radius(Q[c], n), radius(Q[o], n))
                         for (_, o, n) in inc[c]) for c in range(nc)])
    ref = -cfl * R / rsum[:, None]
    assert np.abs(dQ - ref).max() <= 1e-7 * np.abs(ref).max()


def test_uniform_flow_is_a_fixed_point():
    m = MESHES["periodic"]()
    raw = m.raw()
    om = O.OracleMesh(raw)
    nc = int(raw["n_cells"])
    q = np.array([1.0, 0.3, -0.2, 2.5])
    Q = np.tile(q, (nc, 1))
    col = greedy_colours(incidences(raw, om.geometry()))
    Q1, res = om.lusgs_step(O.make_config(), Q, col, 1e4, 1.0)
    assert res.max() < 1e-13
    assert np.abs(Q1 - Q).max() < 1e-13


def test_ramp_steady_state_matches_explicit_and_takes_fewer_iterations():
    """Both drivers converge to R(Q) = 0: the same steady state.  The implicit one needs several
    times fewer iterations than the explicit SSP-RK2 steps (each of which evaluates R twice)."""
    mach, theta = 2.5, 10.0
    m = ramp.mesh(theta, 30, 18)
    raw = m.raw()
    om = O.OracleMesh(raw)
    c = ramp.config(m, mach)
    cfg = O.make_config(scheme=O.ROE, order=1, rk=2, cfl=0.8, q_inlet=[c.q_inlet[i] for i in range(4)])
    col = greedy_colours(incidences(raw, om.geometry()))
    q0 = ramp.freestream(m, mach)
    tol = 1e-10

    Q, ref, it_lu = q0.copy(), None, None
    for it in range(3000):
        Q, res = om.lusgs_step(cfg, Q, col, 1e4, 1.0)
        ref = res.copy() if ref is None else ref
        if (res / ref).max() <= tol:
            it_lu = it + 1
            break
    assert it_lu is not None
    Q_lu = Q

    Q, ref, it_rk = q0.copy(), None, None
    for it in range(20000):
        Q, res = om.rk_step(cfg, Q, om.cfl_dt(cfg, Q))
        ref = res.copy() if ref is None else ref
        if (res / ref).max() <= tol:
            it_rk = it + 1
            break
    assert it_rk is not None
    assert it_rk >= 3 * it_lu
    assert np.abs(Q_lu - Q).max() <= 1e-7 * np.abs(Q).max()


def _neighbour_sets(raw, geo):
    return [sorted({o for (_, o, _) in cell if o >= 0 and o != c}) for c, cell in enumerate(incidences(raw, geo))]


def _bipartite(nb):
    col = -np.ones(len(nb), int)
    for s in range(len(nb)):
        if col[s] >= 0:
            continue
        col[s] = 0
        stack = [s]
        while stack:
            c = stack.pop()
            for o in nb[c]:
                if col[o] < 0:
                    col[o] = 1 - col[c]
                    stack.append(o)
                elif col[o] == col[c]:
                    return False
    return True


COLOUR_MESHES = {
    "tri": lambda: M.Mesh.gen_tri(30, 20, 0, 3, 0, 2, 0.2, 5, (M.BC_WALL,) * 4),
    "periodic": lambda: M.Mesh.gen_tri(12, 10, -1, 1, -1, 1, 0.1, 3, (M.BC_PERIODIC,) * 4),
    "tet": lambda: M.Mesh.gen_tet(5, 4, 3, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11, (M.BC_WALL,) * 6),
    "stair-un-tri": lambda: M.Mesh.from_raw(_fixture("2d-stair-un-5-tri")),
    "stairW": lambda: M.Mesh.from_raw(_fixture("2d-stairW-1")),
}


def _fixture(name):
    import os
    d = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"msh_{name}.npz")))
    for k in ("dim", "n_cells", "n_int"):
        d[k] = int(d[k])
    return d


@pytest.mark.parametrize("name", sorted(COLOUR_MESHES))
def test_sweep_colouring(name):
    """The LU-SGS colouring: proper; two colours exactly when the cell graph is bipartite; otherwise
    no more colours than greedy in device order (the Kempe pass only removes colours)."""
    m = COLOUR_MESHES[name]()
    raw = m.raw()
    p = M.Prep(m)
    col, ncol = p.cell_colours()
    nb = _neighbour_sets(raw, O.OracleMesh(raw).geometry())
    assert col.min() == 0 and col.max() == ncol - 1 and len(np.unique(col)) == ncol
    assert all(col[o] != col[c] for c in range(len(nb)) for o in nb[c])
    order_dev = np.argsort(p.permutation()[0])
    greedy = greedy_colours(incidences(raw, O.OracleMesh(raw).geometry()), order_dev)
    if _bipartite(nb):
        assert ncol == 2
    else:
        assert 3 <= ncol <= greedy.max() + 1
    assert np.array_equal(col, p.cell_colours()[0])  # deterministic
