"""Pins of the oracle's boundary ghosts, its 1D reduction and its 3D residual (CPU only).

Each check is stated independently of the oracle's own formulas:
- wall and symmetry ghosts on slanted walls, built from the wall's angle by a rotation into the wall
  frame (rhoSolver/RhoSolver.cpp:148-159, 187-198; second order :285-301, 319-336);
- mirror symmetry of the wall Riemann problem on axis-aligned walls (mass and energy fluxes vanish);
- uniform flow along a straight slanted wall is a steady state;
- the first-order residual on a y-uniform strip, summed over a column of cells, is the finite-volume
  difference of a 1D Roe or AUSM+ flux written from the textbooks (Roe 1981 in the wave-strength form
  of Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., 11.2.2, with Harten's
  entropy fix at the reference's absolute delta; Liou, J. Comput. Phys. 129, 364-382, 1996), with
  reflecting ghosts at the end walls (rhoSolver/RhoSolver.cpp:45-65, 90-233);
- in 3D, where the reference has no working semantics (work/FUNCTION.cpp:4,9,13,
  rhoSolver/SolverRoe.cpp:98-103): freestream preservation and conservation on jittered tetrahedra,
  the same 1D reduction over x-slabs, and symmetry of the residual under permutations of the axes
  (the Kuhn split of a cube is invariant under them).
tests/test_oracle_planted.py plants mistakes in the oracle and checks that these tests catch them."""
import itertools
import math

import numpy as np
import pytest

import mstcfd as M
from helpers import cons, periodic_in_x, prim, smooth_state
from oracle import pyoracle as O

G = 1.4
DELTA = 0.125  # the reference's absolute Harten delta (rhoSolver/SolverRoe.cpp:114-123)
NO_QUIRKS = 0  # consistent AUSM+, no last-interior-face quirk


# ------------------------------------------------------------------------ textbook 1D fluxes
def _prim1(q):
    rho, m, E = q
    u = m / rho
    p = (G - 1) * (E - 0.5 * rho * u * u)
    return rho, u, p, (E + p) / rho


def _euler1(q):
    rho, u, p, H = _prim1(q)
    return np.array([rho * u, rho * u * u + p, rho * u * H])


def roe1d(qL, qR):
    """Roe's flux in Toro's wave-strength form (11.2.2), |lambda| with Harten's fix at delta."""
    rL, uL, pL, HL = _prim1(qL)
    rR, uR, pR, HR = _prim1(qR)
    wL, wR = math.sqrt(rL), math.sqrt(rR)
    u = (wL * uL + wR * uR) / (wL + wR)
    H = (wL * HL + wR * HR) / (wL + wR)
    a = math.sqrt((G - 1) * (H - 0.5 * u * u))
    d = np.asarray(qR, float) - np.asarray(qL, float)
    a2 = (G - 1) / (a * a) * (d[0] * (H - u * u) + u * d[1] - d[2])
    a1 = (d[0] * (u + a) - d[1] - a * a2) / (2 * a)
    a3 = d[0] - (a1 + a2)
    fix = lambda lam: abs(lam) if abs(lam) > DELTA else (lam * lam + DELTA * DELTA) / (2 * DELTA)
    K = [np.array([1, u - a, H - u * a]), np.array([1, u, 0.5 * u * u]), np.array([1, u + a, H + u * a])]
    diss = fix(u - a) * a1 * K[0] + fix(u) * a2 * K[1] + fix(u + a) * a3 * K[2]
    return 0.5 * (_euler1(qL) + _euler1(qR)) - 0.5 * diss


def ausm1d(qL, qR):
    """AUSM+ (Liou 1996, beta = 1/8, alpha = 3/16) with the interface speed min(a~_L, a~_R),
    a~ = a*^2 / max(a*, |u|).  For |u| < a* this is Liou's a~ exactly (the reference writes |U| with
    the full speed, rhoSolver/SolverAusm.cpp:10-11,20-21; the tests keep |u| < a*)."""
    rL, uL, pL, HL = _prim1(qL)
    rR, uR, pR, HR = _prim1(qR)
    sL = math.sqrt(2 * (G - 1) / (G + 1) * HL)
    sR = math.sqrt(2 * (G - 1) / (G + 1) * HR)
    a = min(sL * sL / max(sL, abs(uL)), sR * sR / max(sR, abs(uR)))
    ML, MR = uL / a, uR / a

    def mp(Mv, s):  # M^{+-}(M) and P^{+-}(M), s = +1 / -1
        if abs(Mv) <= 1:
            m = s * 0.25 * (Mv + s) ** 2 + s * 0.125 * (Mv * Mv - 1) ** 2
            P = 0.25 * (Mv + s) ** 2 * (2 - s * Mv) + s * 0.1875 * Mv * (Mv * Mv - 1) ** 2
        else:
            m = 0.5 * (Mv + s * abs(Mv))
            P = 0.5 * (1 + s * math.copysign(1, Mv))
        return m, P

    mLp, PLp = mp(ML, 1)
    mRm, PRm = mp(MR, -1)
    m12 = mLp + mRm
    p12 = PLp * pL + PRm * pR
    psiL = np.array([rL, rL * uL, rL * HL])
    psiR = np.array([rR, rR * uR, rR * HR])
    F = a * (max(m12, 0) * psiL + min(m12, 0) * psiR)
    F[1] += p12
    return F


def flux1d(scheme, qL, qR):
    return roe1d(qL, qR) if scheme == O.ROE else ausm1d(qL, qR)


# ------------------------------------------------------------------------ meshes
def rotated_channel(alpha, bc, nx=14, ny=8, jitter=0.2, seed=21):
    """gen_tri on [0,2]x[0,1] rotated by alpha: straight walls at angle alpha (boundary nodes only
    slide along their sides, so the sides stay straight)."""
    raw = M.Mesh.gen_tri(nx, ny, 0, 2, 0, 1, jitter, seed, bc).raw()
    c, s = math.cos(alpha), math.sin(alpha)
    x, y = raw["xyz"][:, 0].copy(), raw["xyz"][:, 1].copy()
    raw = dict(raw)
    raw["xyz"] = np.ascontiguousarray(np.stack([c * x - s * y, s * x + c * y], 1))
    return raw


def side_of(raw, zones_bc):
    """Face index ranges of the four generated sides (x0, x1, y0, y1), in file order."""
    nint = int(raw["n_int"])
    ny = int((raw["ftype"][nint:] == zones_bc[0]).sum()) if zones_bc[0] == zones_bc[1] else None
    return nint, ny


def recon(Q, grad, phi, ccen, fcen, c, f):
    """Limited linear reconstruction Q_c + phi_c o (grad_c (x_f - x_c)) (rhoSolver/RhoSolver.cpp:250)."""
    return Q[c] + phi[c] * (grad[c] @ (fcen[f] - ccen[c]))


# ------------------------------------------------------------------------ wall / symmetry ghosts
@pytest.mark.parametrize("bc", [O.WALL, O.SYMMETRY])
@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
@pytest.mark.parametrize("order", [1, 2])
def test_wall_ghost_on_slanted_walls(bc, scheme, order):
    """Wall faces at angle alpha: the ghost keeps the tangential momentum and negates the normal
    one, with the unit normal taken from the wall's angle (not from the face's area vector).
    Order 2: a wall reflects the reconstructed interior state (RhoSolver.cpp:285-301); a symmetry
    face pairs the unreconstructed cell state with a ghost whose density and energy are
    reconstructed and whose momentum is the cell's, reflected (:319-336)."""
    alpha = 0.37
    raw = rotated_channel(alpha, (O.INLET, O.OUTLET, bc, bc))
    om = O.OracleMesh(raw)
    geo = om.geometry()
    cfg = O.make_config(scheme=scheme, order=order, limiter=O.LIM_VANLEER if order == 2 else O.LIM_NONE)
    Q = smooth_state(geo["ccen"], seed=5, mach=0.7)
    F = om.face_flux(cfg, Q)
    flags = om.lrflags(cfg.compat)
    if order == 2:
        grad = om.gradient(cfg, Q)
        phi = om.limiter(cfg, Q, grad)
    t_hat = np.array([math.cos(alpha), math.sin(alpha)])
    n_hat = np.array([-math.sin(alpha), math.cos(alpha)])

    def mirror(q):
        mt, mn = q[1:3] @ t_hat, q[1:3] @ n_hat
        g = q.copy()
        g[1:3] = mt * t_hat - mn * n_hat
        return g

    point = O.roe_flux if scheme == O.ROE else O.ausm_flux
    walls = np.nonzero(raw["ftype"] == bc)[0]
    assert len(walls) == 2 * 14
    worst = 0.0
    for f in walls:
        c = raw["c0"][f]
        qc = Q[c]
        if order == 1:
            inner, ghost = qc, mirror(qc)
        elif bc == O.WALL:
            inner = recon(Q, grad, phi, geo["ccen"], geo["fcen"], c, f)
            ghost = mirror(inner)
        else:
            inner = qc
            ghost = recon(Q, grad, phi, geo["ccen"], geo["fcen"], c, f)
            ghost[1:3] = mirror(qc)[1:3]
        for d in range(2):
            L, R = (inner, ghost) if flags[f, d] else (ghost, inner)
            Fe = point(L, R, d, cfg)
            worst = max(worst, np.abs(F[f, d] - Fe).max() / np.abs(Fe).max())
    assert worst <= 1e-13, worst


@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
@pytest.mark.parametrize("bc", [O.WALL, O.SYMMETRY])
def test_wall_mirror_fluxes_vanish(scheme, bc):
    """Axis-aligned walls with subsonic flow into or out of them: the ghost is the mirror image of
    the cell, so the Riemann problem along the wall normal is mirror-symmetric and its mass and
    energy fluxes vanish (Roe exactly; AUSM+ for |M| <= 1, including the reference's quirks)."""
    m = M.Mesh.gen_tri(10, 6, 0, 1.5, 0, 1, 0.15, 3, (bc,) * 4)
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    cfg = O.make_config(scheme=scheme)
    Q = smooth_state(geo["ccen"], seed=8, mach=0.6)
    F = om.face_flux(cfg, Q)
    for f in np.nonzero(raw["ftype"] == bc)[0]:
        S = geo["S"][f]
        d = 0 if abs(S[0]) > abs(S[1]) else 1
        assert S[1 - d] == 0.0
        scale = np.abs(F[f, d]).max()
        assert abs(F[f, d, 0]) <= 1e-14 * scale and abs(F[f, d, 3]) <= 1e-14 * scale, (f, F[f, d])
        # the momentum flux along the normal is the pressure-like wall force, nonzero
        assert abs(F[f, d, 1 + d]) > 0.1 * scale


@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
@pytest.mark.parametrize("bc", [O.WALL, O.SYMMETRY])
@pytest.mark.parametrize("order", [1, 2])
def test_uniform_flow_along_a_slanted_wall_is_steady(scheme, bc, order):
    alpha = -0.52
    raw = rotated_channel(alpha, (O.INLET, O.OUTLET, bc, bc))
    om = O.OracleMesh(raw)
    vel = 0.8 * np.array([[math.cos(alpha), math.sin(alpha)]])
    q = cons([1.3], vel, [0.9])[0]
    cfg = O.make_config(scheme=scheme, order=order, limiter=O.LIM_MINMOD if order == 2 else O.LIM_NONE,
                        q_inlet=list(q))
    R, S = om.residual(cfg, np.tile(q, (raw["n_cells"], 1)))
    assert (np.abs(R) <= 1e-14 * S.max()).all(), np.abs(R).max() / S.max()


# ------------------------------------------------------------------------ 1D reduction
def profile1d(x, subsonic=True):
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) + 0.4 * (x > 0.55)
    p = 1.0 + 0.5 * np.cos(3 * np.pi * x) * (x < 0.7) + 0.2
    a = np.sqrt(G * p / rho)
    u = (0.45 if subsonic else 1.4) * a * np.sin(2.5 * np.pi * x + 0.3)
    return rho, u, p


@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
def test_one_dimensional_reduction_on_a_strip(scheme):
    """First order on a 2-row strip of right triangles (no jitter), data constant in each column
    with v = 0, walls all round: the residual summed over a column equals
    H (F^_{i+1/2} - F^_{i-1/2}) for rho, rho u, E (H the strip height, F^ the 1D flux, with ghosts
    (rho, -rho u, E) at the end walls) and zero for rho v.  Each triangle's rho v residual is zero too."""
    nx, H = 40, 0.05
    m = M.Mesh.gen_tri(nx, 2, 0, 1, 0, H, 0.0, 1, (O.WALL,) * 4)
    raw = m.raw()
    om = O.OracleMesh(raw)
    geo = om.geometry()
    col = np.floor(geo["ccen"][:, 0] * nx).astype(int)
    assert np.bincount(col, minlength=nx).tolist() == [4] * nx
    xc = (np.arange(nx) + 0.5) / nx
    rho, u, p = profile1d(xc)
    q1 = np.stack([rho, rho * u, p / (G - 1) + 0.5 * rho * u * u], 1)
    Q = cons(rho[col], np.stack([u[col], np.zeros(len(col))], 1), p[col])
    cfg = O.make_config(scheme=scheme, eor=0.0, compat=NO_QUIRKS)
    R, S = om.residual(cfg, Q)
    ghostL = q1[0] * np.array([1, -1, 1])
    ghostR = q1[-1] * np.array([1, -1, 1])
    Fh = [flux1d(scheme, ghostL, q1[0])]
    Fh += [flux1d(scheme, q1[i], q1[i + 1]) for i in range(nx - 1)]
    Fh += [flux1d(scheme, q1[-1], ghostR)]
    Fh = np.array(Fh)
    scale = H * np.abs(Fh).max()
    for i in range(nx):
        Rc = R[col == i].sum(0)
        expect = H * (Fh[i + 1] - Fh[i])
        assert np.abs(Rc[[0, 1, 3]] - expect).max() <= 1e-12 * scale, (i, Rc, expect)
        assert abs(Rc[2]) <= 1e-12 * scale
    assert np.abs(R[:, 2]).max() <= 1e-12 * scale


# ------------------------------------------------------------------------ 3D
@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
@pytest.mark.parametrize("order,lim", [(1, O.LIM_NONE), (2, O.LIM_MINMOD), (2, O.LIM_VANLEER)])
def test_3d_freestream_and_conservation(scheme, order, lim):
    m = M.Mesh.gen_tet(5, 4, 6, (0, 0, 0), (1.0, 0.8, 1.2), 0.2, 7, (O.INLET,) * 6)
    om = O.OracleMesh(m.raw())
    q = cons([1.1], [[0.5, -0.35, 0.25]], [0.8])[0]
    cfg = O.make_config(scheme=scheme, order=order, limiter=lim, q_inlet=list(q), dim=3)
    R, S = om.residual(cfg, np.tile(q, (m.n_cells, 1)))
    assert (np.abs(R) <= 1e-14 * S.max()).all(), np.abs(R).max() / S.max()
    # conservation: sum_c R_c = sum over boundary faces of n_out(c0) . F_f, every boundary type
    bc = (O.INLET, O.OUTLET, O.WALL, O.SYMMETRY, O.WALL, O.OUTLET)
    m2 = M.Mesh.gen_tet(5, 4, 6, (0, 0, 0), (1.0, 0.8, 1.2), 0.2, 8, bc)
    raw2 = m2.raw()
    om2 = O.OracleMesh(raw2)
    geo = om2.geometry()
    Q = smooth_state(geo["ccen"], seed=4, mach=0.7)
    R, S = om2.residual(cfg, Q)
    F = om2.face_flux(cfg, Q)
    b = np.nonzero(raw2["c1"] < 0)[0]
    nout = geo["dac"][b, None] * geo["S"][b]
    boundary = np.einsum("fd,fdk->k", nout, F[b])
    assert np.abs(R.sum(0) - boundary).max() <= 1e-13 * S.sum(0).max()
    # and it is not trivially zero
    assert np.abs(boundary).max() > 1e-3 * S.sum(0).max()


@pytest.mark.parametrize("scheme,order,lim", [(O.ROE, 1, 0), (O.AUSM, 2, O.LIM_MINMOD), (O.ROE, 2, O.LIM_VANLEER)])
def test_3d_periodic_pairs(scheme, order, lim):
    """An unjittered box periodic in x (h_x = 1/8, exact in binary), walls around: uniform flow along
    x is steady; the mass and energy residuals sum to zero (a wall's ghost mirrors its face state, so
    the wall carries neither, and a periodic pair's fluxes cancel; a second-order symmetry face would
    not do, since its ghost pairs the raw interior state with the reflected reconstructed one); and
    translating the state by one cell layer along x, across the periodic seam, translates the
    residual.  A partner face wired to the wrong cell, or a c1 side that keeps its own face centre for
    the gradient's offset, breaks the translation at the seam."""
    n = 8
    bc = (O.WALL,) * 6
    box = M.Mesh.gen_tet(n, 5, 4, (0, 0, 0), (1, 0.8, 0.7), 0.0, 3, bc)
    raw = periodic_in_x(box.raw())
    om = O.OracleMesh(raw)
    geo = om.geometry()
    cfg = O.make_config(scheme=scheme, order=order, limiter=lim, dim=3, compat=NO_QUIRKS)
    q = cons([1.1], [[0.6, 0.0, 0.0]], [0.9])[0]
    R, S = om.residual(cfg, np.tile(q, (box.n_cells, 1)))
    assert (np.abs(R) <= 1e-14 * S.max()).all(), np.abs(R).max() / S.max()
    Q = smooth_state(geo["ccen"], seed=12, mach=0.6)
    R, S = om.residual(cfg, Q)
    tot = R.sum(0)
    assert abs(tot[0]) <= 1e-13 * S.sum(0)[0] and abs(tot[4]) <= 1e-13 * S.sum(0)[4], tot / S.sum(0)
    # translation by one layer: cell c moves to the cell whose centre is c's shifted by h_x (mod 1)
    key = lambda x: tuple(np.round(x, 9))
    where = {key(x): i for i, x in enumerate(geo["ccen"])}
    shifted = geo["ccen"].copy()
    shifted[:, 0] = np.mod(shifted[:, 0] + 1.0 / n, 1.0)
    to = np.array([where[key(x)] for x in shifted])
    assert len(set(to.tolist())) == box.n_cells
    Qt = np.empty_like(Q)
    Qt[to] = Q
    Rt, _ = om.residual(cfg, Qt)
    assert np.abs(Rt[to] - R).max() <= 1e-13 * S.max()
    # across the seam the residual is not trivially small
    assert np.abs(R).max() > 1e-3 * S.max()


@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
def test_3d_one_dimensional_reduction(scheme):
    """x-slabs of an unjittered Kuhn-split box (h = 1/8, exact in binary), data constant per slab with
    v = w = 0, walls at x = 0, 1 and symmetry at the sides: each slab's summed residual is the 1D
    flux difference over the unit cross-section."""
    n = 8
    bc = (O.WALL, O.WALL, O.SYMMETRY, O.SYMMETRY, O.WALL, O.WALL)
    m = M.Mesh.gen_tet(n, n, n, (0, 0, 0), (1, 1, 1), 0.0, 1, bc)
    om = O.OracleMesh(m.raw())
    geo = om.geometry()
    slab = np.floor(geo["ccen"][:, 0] * n).astype(int)
    assert np.bincount(slab, minlength=n).tolist() == [6 * n * n] * n
    xc = (np.arange(n) + 0.5) / n
    rho, u, p = profile1d(xc)
    q1 = np.stack([rho, rho * u, p / (G - 1) + 0.5 * rho * u * u], 1)
    vel = np.zeros((m.n_cells, 3))
    vel[:, 0] = u[slab]
    Q = cons(rho[slab], vel, p[slab])
    cfg = O.make_config(scheme=scheme, eor=0.0, compat=NO_QUIRKS, dim=3)
    R, S = om.residual(cfg, Q)
    flip = np.array([1, -1, 1])
    Fh = np.array([flux1d(scheme, q1[0] * flip, q1[0])] + [flux1d(scheme, q1[i], q1[i + 1]) for i in range(n - 1)]
                  + [flux1d(scheme, q1[-1], q1[-1] * flip)])
    scale = np.abs(Fh).max()
    for i in range(n):
        Rs = R[slab == i].sum(0)
        assert np.abs(Rs[[0, 1, 4]] - (Fh[i + 1] - Fh[i])).max() <= 1e-12 * scale, (i, Rs)
        assert np.abs(Rs[[2, 3]]).max() <= 1e-12 * scale


def _pulse3d(ccen):
    """A field symmetric under every permutation of the axes about (1/2, 1/2, 1/2): radial velocity,
    density and pressure depending on r and on x + y + z."""
    d = ccen - 0.5
    r2 = (d ** 2).sum(1)
    g = np.exp(-r2 / (2 * 0.2 ** 2))
    s = d.sum(1)
    rho = 1 + 0.3 * g * (1 + 0.5 * s)
    p = 1 + 0.4 * g - 0.2 * s * g
    vel = 1.2 * d * g[:, None] * (1 + 0.3 * s)[:, None]
    return cons(rho, vel, p)


@pytest.mark.parametrize("scheme", [O.ROE, O.AUSM])
@pytest.mark.parametrize("order,lim", [(1, O.LIM_NONE), (2, O.LIM_MINMOD)])
def test_3d_residual_is_symmetric_under_axis_permutations(scheme, order, lim):
    """The Kuhn split maps onto itself under any permutation of the axes, so with a permutation-
    symmetric field and the same boundary type on every side the residual must permute with the
    axes: R_{pi(c)} = P R_c.  A 3D term that treats one axis differently (a kinetic energy over two
    components, a misplaced eigenvector entry, an index slip in the face normals) breaks it.
    The last-interior-face quirk is off: it singles out one face."""
    n = 6
    for bc in (O.WALL, O.SYMMETRY):
        m = M.Mesh.gen_tet(n, n, n, (0, 0, 0), (1, 1, 1), 0.0, 1, (bc,) * 6)
        om = O.OracleMesh(m.raw())
        geo = om.geometry()
        Q = _pulse3d(geo["ccen"])
        compat = O.COMPAT_REFERENCE & ~O.COMPAT_GRAD_LASTINT
        cfg = O.make_config(scheme=scheme, order=order, limiter=lim, compat=compat, dim=3)
        R, S = om.residual(cfg, Q)
        key = {tuple(k): i for i, k in enumerate(np.rint(geo["ccen"] * 4 * n * 6).astype(np.int64))}
        scale = S.max()
        for perm in itertools.permutations(range(3)):
            if perm == (0, 1, 2):
                continue
            P = list(perm)
            img = np.array([key[tuple(k)] for k in np.rint(geo["ccen"][:, P] * 4 * n * 6).astype(np.int64)])
            # cell img[c] has centre (x_P0, x_P1, x_P2) of cell c; its momentum component j is c's P[j]
            Rp = R[img]
            expect = R.copy()
            expect[:, 1:4] = R[:, [1 + P[0], 1 + P[1], 1 + P[2]]]
            err = np.abs(Rp - expect).max() / scale
            assert err <= 1e-13, (bc, perm, err)
        assert np.abs(R[:, 1:4]).max() > 1e-3 * scale  # the pulse drives every momentum component
