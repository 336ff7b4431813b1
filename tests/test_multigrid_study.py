"""Would a learned start pay under a multigrid driver?  A CPU study (numpy), not a product path.

Under the two drivers of the library (explicit SSP-RK2, LU-SGS) a start near the steady state saves
little: the time to a 1e-8 residual is set by the slowest-decaying error, which the prediction
leaves too (tests/test_gpu_warmstart.py).  A multigrid driver removes smooth errors on coarse grids,
so its asymptotic convergence is fast, and the nonlinear transient from the freestream could then
dominate, which a good start skips.  This study measures that on the steady compression ramp
(first-order Roe, 3,000 triangles) with an agglomeration FAS multigrid: greedy agglomeration of
cells with their neighbours, coarse faces carrying the summed normals of the fine faces between two
agglomerates, volume-weighted restriction of the state, summed restriction of residuals, the FAS
forcing term, injection of the coarse correction, and a three-stage scheme with local time steps as
the smoother on every level (coefficients 0.1481, 0.4, 1 at CFL 1.5, van Leer, Tai and Powell,
AIAA 89-1933, for first-order upwind schemes).  Its starts: the freestream; the steady state
perturbed by 1e-3 and 1e-5 per cell and field; and the steady state of the same ramp at a Mach
number 0.05 higher, a start as close as the network's prediction (1-3 % relative L2 error against
the network's 1-2 %).  The network itself needs the GPU; these starts are its CPU stand-ins."""
import numpy as np
import pytest

import mstcfd as M
from mstcfd import ramp as RP

G = 1.4
ALPHA = (0.1481, 0.4, 1.0)
CFL = 1.5
KIND_WALL, KIND_SYM, KIND_INLET = 3, 7, 10  # Fluent zone types of the raw mesh


def prim(Q):
    r = Q[:, 0]
    u, v = Q[:, 1] / r, Q[:, 2] / r
    return r, u, v, (G - 1) * (Q[:, 3] - 0.5 * r * (u * u + v * v))


def euler_flux(Q, n):
    r, u, v, p = prim(Q)
    un = u * n[:, 0] + v * n[:, 1]
    return np.stack([r * un, r * un * u + p * n[:, 0], r * un * v + p * n[:, 1], un * (Q[:, 3] + p)], 1)


def roe(QL, QR, n):
    """Roe's flux through the unit normal n (textbook wave form, Harten's fix at a tenth of a)."""
    rL, uL, vL, pL = prim(QL)
    rR, uR, vR, pR = prim(QR)
    HL, HR = (QL[:, 3] + pL) / rL, (QR[:, 3] + pR) / rR
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    w = sL + sR
    u, v, H = (sL * uL + sR * uR) / w, (sL * vL + sR * vR) / w, (sL * HL + sR * HR) / w
    q2 = u * u + v * v
    a = np.sqrt((G - 1) * (H - 0.5 * q2))
    nx, ny = n[:, 0], n[:, 1]
    un, r = u * nx + v * ny, sL * sR
    dp, dr = pR - pL, rR - rL
    dun = (uR - uL) * nx + (vR - vL) * ny
    dut, dvt = (uR - uL) - dun * nx, (vR - vL) - dun * ny
    eps = 0.1 * a
    fix = lambda lam: np.where(lam < eps, (lam * lam + eps * eps) / (2 * eps), lam)
    l1, l2, l3 = fix(np.abs(un - a)), np.abs(un), fix(np.abs(un + a))
    a1, a3 = (dp - r * a * dun) / (2 * a * a), (dp + r * a * dun) / (2 * a * a)
    a2 = dr - dp / (a * a)
    D = np.stack([l1 * a1 + l2 * a2 + l3 * a3,
                  l1 * a1 * (u - a * nx) + l2 * (a2 * u + r * dut) + l3 * a3 * (u + a * nx),
                  l1 * a1 * (v - a * ny) + l2 * (a2 * v + r * dvt) + l3 * a3 * (v + a * ny),
                  l1 * a1 * (H - a * un) + l2 * (a2 * 0.5 * q2 + r * (u * dut + v * dvt)) + l3 * a3 * (H + a * un)], 1)
    return 0.5 * (euler_flux(QL, n) + euler_flux(QR, n)) - 0.5 * D


class Level:
    """Cells (volumes) and faces (c0, c1 or -1, area vector S outward of c0, boundary zone type)."""

    def __init__(self, c0, c1, S, kind, vol):
        self.nc, self.c0, self.c1, self.S, self.kind, self.vol = len(vol), c0, c1, S, kind, vol
        self.area = np.linalg.norm(S, axis=1)
        self.n = S / self.area[:, None]
        self.inner = c1 >= 0

    @staticmethod
    def from_mesh(m):
        raw = m.raw()
        xyz, fn = raw["xyz"], raw["fnodes"]
        c0, c1 = raw["c0"].astype(np.int64), raw["c1"].astype(np.int64)
        x0, x1 = xyz[fn[:, 0]], xyz[fn[:, 1]]
        S = np.stack([-(x0[:, 1] - x1[:, 1]), x0[:, 0] - x1[:, 0]], 1)
        fc = 0.5 * (x0 + x1)
        nc = int(raw["n_cells"])
        cc, cnt = np.zeros((nc, 2)), np.zeros(nc)
        for c in (c0, c1):
            ok = c >= 0
            np.add.at(cc, c[ok], fc[ok])
            np.add.at(cnt, c[ok], 1)
        cc /= cnt[:, None]
        S[((fc - cc[c0]) * S).sum(1) < 0] *= -1
        vol = np.zeros(nc)  # divergence theorem
        np.add.at(vol, c0, 0.5 * (fc * S).sum(1))
        np.add.at(vol, c1[c1 >= 0], -0.5 * (fc[c1 >= 0] * S[c1 >= 0]).sum(1))
        return Level(c0, c1, S, raw["ftype"].copy(), vol)

    def residual(self, Q, qin):
        QL = Q[self.c0]
        QR = QL.copy()  # outlet ghost: the interior state
        QR[self.inner] = Q[self.c1[self.inner]]
        b = ~self.inner
        k, nb, g = self.kind[b], self.n[b], QR[b]
        wall = (k == KIND_WALL) | (k == KIND_SYM)
        mn = g[:, 1] * nb[:, 0] + g[:, 2] * nb[:, 1]
        g[wall, 1] -= 2 * mn[wall] * nb[wall, 0]
        g[wall, 2] -= 2 * mn[wall] * nb[wall, 1]
        g[k == KIND_INLET] = qin
        QR[b] = g
        F = roe(QL, QR, self.n) * self.area[:, None]
        R = np.zeros_like(Q)
        np.add.at(R, self.c0, F)
        np.add.at(R, self.c1[self.inner], -F[self.inner])
        return R

    def local_dt(self, Q):
        _, u, v, p = prim(Q)
        a = np.sqrt(G * p / Q[:, 0])
        lam = np.zeros(self.nc)
        for c, mask in ((self.c0, np.ones_like(self.inner)), (self.c1, self.inner)):
            cc = c[mask]
            np.add.at(lam, cc, np.abs(u[cc] * self.S[mask, 0] + v[cc] * self.S[mask, 1]) + a[cc] * self.area[mask])
        return CFL * self.vol / lam

    def norm(self, Q, qin):
        return np.sqrt(((self.residual(Q, qin) / self.vol[:, None]) ** 2).mean(0))


def agglomerate(L):
    """Each unassigned cell in order takes its unassigned neighbours; a cell left alone joins a
    neighbour's agglomerate.  Returns the coarse level and the agglomerate of every fine cell."""
    nbrs = [[] for _ in range(L.nc)]
    for f in np.nonzero(L.inner)[0]:
        nbrs[L.c0[f]].append(L.c1[f])
        nbrs[L.c1[f]].append(L.c0[f])
    agg, na = -np.ones(L.nc, np.int64), 0
    for c in range(L.nc):
        if agg[c] >= 0:
            continue
        agg[c] = na
        for o in nbrs[c]:
            if agg[o] < 0:
                agg[o] = na
        na += 1
    size = np.bincount(agg, minlength=na)
    for c in range(L.nc):
        if size[agg[c]] == 1 and nbrs[c]:
            size[agg[c]] -= 1
            agg[c] = agg[nbrs[c][0]]
            size[agg[c]] += 1
    _, agg = np.unique(agg, return_inverse=True)
    na = int(agg.max()) + 1
    A0 = agg[L.c0]
    A1 = np.where(L.inner, agg[np.maximum(L.c1, 0)], -1)
    cut = L.inner & (A0 != A1)
    lo, hi, S = np.minimum(A0[cut], A1[cut]), np.maximum(A0[cut], A1[cut]), L.S[cut].copy()
    S[A0[cut] > A1[cut]] *= -1  # outward of the lower-numbered agglomerate
    keys, inv = np.unique(lo * na + hi, return_inverse=True)
    Sc = np.zeros((len(keys), 2))
    np.add.at(Sc, inv, S)
    b = ~L.inner
    C = Level(np.concatenate([keys // na, A0[b]]), np.concatenate([keys % na, -np.ones(b.sum(), np.int64)]),
              np.concatenate([Sc, L.S[b]]), np.concatenate([np.full(len(keys), 2), L.kind[b]]),
              np.bincount(agg, weights=L.vol, minlength=na))
    return C, agg


def restrict_state(L, C, agg, Q):
    out = np.zeros((C.nc, Q.shape[1]))
    np.add.at(out, agg, Q * L.vol[:, None])
    return out / C.vol[:, None]


def restrict_sum(C, agg, R):
    out = np.zeros((C.nc, R.shape[1]))
    np.add.at(out, agg, R)
    return out


def smooth(L, Q, P, qin, sweeps):
    for _ in range(sweeps):
        dt = (L.local_dt(Q) / L.vol)[:, None]
        Q0 = Q
        for al in ALPHA:
            Q = Q0 - al * dt * (L.residual(Q, qin) + P)
    return Q


def cycle(levels, aggs, l, Q, P, qin, gamma):
    """One FAS cycle on level l (gamma 1: V, 2: W), one smoothing sweep before and after."""
    L = levels[l]
    Q = smooth(L, Q, P, qin, 1)
    if l + 1 < len(levels):
        C, agg = levels[l + 1], aggs[l]
        Qc = restrict_state(L, C, agg, Q)
        Pc = restrict_sum(C, agg, L.residual(Q, qin) + P) - C.residual(Qc, qin)
        Qc0 = Qc
        for _ in range(gamma):
            Qc = cycle(levels, aggs, l + 1, Qc, Pc, qin, gamma)
        Q = smooth(L, Q + (Qc - Qc0)[agg], P, qin, 1)
    return Q


def solve(levels, aggs, Q, qin, ref, tol, gamma=2, max_cycles=3000):
    """Cycles (or, with one level, smoothing sweeps) until every field's L2(R/V) <= tol * ref."""
    P = np.zeros_like(Q)
    for it in range(1, max_cycles + 1):
        Q = cycle(levels, aggs, 0, Q, P, qin, gamma)
        if (levels[0].norm(Q, qin) / ref).max() <= tol:
            return it, Q
    raise AssertionError(f"no convergence to {tol} in {max_cycles} cycles")


@pytest.fixture(scope="module")
def ramp():
    mach, theta = 2.2, 10.0
    m = RP.mesh(theta, 60, 25)
    levels, aggs = [Level.from_mesh(m)], []
    for _ in range(3):
        C, agg = agglomerate(levels[-1])
        levels.append(C)
        aggs.append(agg)
    qin = RP.freestream(m, mach)[0]
    Qfs = np.tile(qin, (levels[0].nc, 1))
    return dict(m=m, mach=mach, levels=levels, aggs=aggs, qin=qin, Qfs=Qfs, ref=levels[0].norm(Qfs, qin))


def test_flux_and_agglomeration_invariants(ramp):
    L, levels, aggs, qin = ramp["levels"][0], ramp["levels"], ramp["aggs"], ramp["qin"]
    rng = np.random.default_rng(1)
    Q = ramp["Qfs"] * (1 + 0.05 * rng.random(ramp["Qfs"].shape))
    n = L.n
    assert np.allclose(roe(Q[L.c0], Q[L.c0], n), euler_flux(Q[L.c0], n), rtol=0, atol=1e-13)  # consistency
    assert [C.nc for C in levels] == sorted({C.nc for C in levels}, reverse=True)
    for l, agg in enumerate(aggs):
        F, C = levels[l], levels[l + 1]
        assert np.isclose(C.vol.sum(), F.vol.sum(), rtol=1e-14)
        # every coarse cell is closed: its outward area vectors sum to zero
        closure = np.zeros((C.nc, 2))
        np.add.at(closure, C.c0, C.S)
        np.add.at(closure, C.c1[C.inner], -C.S[C.inner])
        assert np.abs(closure).max() <= 1e-13
        # a uniform state restricts to itself, and restricted residuals keep the fine totals
        u = np.tile(qin, (F.nc, 1))
        assert np.allclose(restrict_state(F, C, agg, u), qin, rtol=1e-14)
        R = rng.standard_normal((F.nc, 4))
        assert np.allclose(restrict_sum(C, agg, R).sum(0), R.sum(0), rtol=1e-12)


def test_fas_reaches_the_fine_grid_solution(ramp):
    """FAS only accelerates: its fixed point is the fine grid's own steady state, reached with less
    work than by single-grid sweeps (a W-cycle over these four levels costs about eight fine sweeps:
    two on the fine level, and the coarse levels' visits doubling as the cells halve)."""
    levels, aggs, qin, Qfs, ref = ramp["levels"], ramp["aggs"], ramp["qin"], ramp["Qfs"], ramp["ref"]
    n_sg, Q_sg = solve(levels[:1], [], Qfs, qin, ref, 1e-10)
    n_mg, Q_mg = solve(levels, aggs, Qfs, qin, ref, 1e-10)
    print(f"to 1e-10: {n_sg} single-grid sweeps, {n_mg} W-cycles on {[C.nc for C in levels]} cells", flush=True)
    sc = np.abs(Q_sg).max(0)
    assert np.abs((Q_mg - Q_sg) / sc).max() <= 1e-6
    assert n_mg * 8 < n_sg


def test_near_starts_save_little_under_multigrid(ramp):
    levels, aggs, qin, Qfs, ref, m = ramp["levels"], ramp["aggs"], ramp["qin"], ramp["Qfs"], ramp["ref"], ramp["m"]
    n_cold, Qs = solve(levels, aggs, Qfs, qin, ref, 1e-8)
    _, Qs = solve(levels, aggs, Qs, qin, ref, 1e-12)
    rng = np.random.default_rng(7)
    starts = {f"steady x (1 + {eps:g} noise)": Qs * (1 + eps * rng.standard_normal(Qs.shape)) for eps in (1e-3, 1e-5)}
    qn = RP.freestream(m, ramp["mach"] + 0.05)[0]
    Qn = np.tile(qn, (levels[0].nc, 1))
    _, starts["steady state at M + 0.05"] = solve(levels, aggs, Qn, qn, levels[0].norm(Qn, qn), 1e-10)
    sc = np.abs(Qs).max(0)
    err = lambda Q: float(np.sqrt((((Q - Qs) / sc) ** 2).mean(0)).mean())
    lines = [f"freestream: {n_cold} W-cycles (error {err(Qfs):.3f})"]
    for name, Q0 in starts.items():
        n, _ = solve(levels, aggs, Q0, qin, ref, 1e-8)
        lines.append(f"{name}: {n} W-cycles (error {err(Q0):.1e}), {n_cold / n:.2f}x")
        assert n_cold / n < 3
    print("; ".join(lines), flush=True)
