"""Learned warm start of steady supersonic compression-ramp runs (configuration C5).

The solver generates steady solutions of ramp cases (inflow Mach number M, ramp angle theta) with the
batched steady driver; a 4 x 128 tanh network (x, y, M, theta) -> encoded state is trained on them by
the library's kernels; a new case then starts from the network's prediction instead of the uniform
freestream.  The freestream start is the fallback (a prediction the solver cannot advance) and the
judge: both runs must bring every field's L2(R/V) to the same absolute level, tol times the
freestream run's first-step residual.

Encoding of the state (decoded on the device by mst_ctx_warm_start):
    y = (log(rho/rho_inf), u/|u_inf|, v/|u_inf|, log(p/p_inf)),  rho_inf = 1, p_inf = 1/gamma, |u_inf| = M.
"""
import math
import time

import numpy as np

from . import (BC_INLET, BC_OUTLET, BC_WALL, INIT_FREESTREAM, LIM_NONE, MLP, ROE, STEADY_EXPLICIT, STEADY_LUSGS,
               Batch, MstError, Mesh, Solver, make_config)

GAMMA = 1.4
X0, X1, HEIGHT = -0.5, 1.5, 1.2
BC = (BC_INLET, BC_OUTLET, BC_WALL, BC_OUTLET)  # left, right, bottom (ramp wall), top
IN_SHIFT = np.array([0.5, 0.6, 2.25, 12.5], np.float32)
IN_SCALE = np.array([1.0, 1 / 0.6, 1 / 0.75, 1 / 7.5], np.float32)


def sample_cases(n, seed):
    """(M, theta in degrees): M ~ U[1.5, 3], theta ~ U[5, 20]."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(1.5, 3.0)), float(rng.uniform(5.0, 20.0))) for _ in range(n)]


def mesh(theta, nx, ny, seed=2024):
    return Mesh.gen_ramp(nx, ny, theta, x0=X0, x1=X1, height=HEIGHT, jitter=0.1, seed=seed, bc=BC)


def freestream(m, mach):
    return m.init_state(INIT_FREESTREAM, [GAMMA, mach, 0.0])


def config(m, mach, cfl=0.8, **kw):
    """First-order Roe, SSP-RK2, CFL time step; the inlet ghost is the freestream."""
    qin = freestream(m, mach)[0].tolist()
    return make_config(scheme=ROE, order=1, limiter=LIM_NONE, rk=2, cfl=cfl, q_inlet=qin, **kw)


def encode(Q, mach):
    rho = Q[:, 0]
    u = Q[:, 1] / rho
    v = Q[:, 2] / rho
    p = (GAMMA - 1) * (Q[:, 3] - 0.5 * rho * (u * u + v * v))
    return np.stack([np.log(rho), u / mach, v / mach, np.log(GAMMA * p)], axis=1).astype(np.float32)


def features(m, mach, theta):
    c = m.cell_centres()
    X = np.empty((m.n_cells, 4), np.float32)
    X[:, :2] = c
    X[:, 2] = mach
    X[:, 3] = theta
    return X


def generate(cases, nx, ny, tol=1e-8, max_steps=20000, check_every=500, autotune_cache=None, driver="lusgs",
             lu_max=4000, lu_check=50):
    """Steady solutions of the cases, solved together in one batch; returns (X, T, info).
    driver "lusgs": the batched LU-SGS driver (CFL 1e4), falling back to the explicit one if a case
    leaves the admissible states; "explicit": SSP-RK2 with the CFL time step."""
    t0 = time.perf_counter()
    meshes = [mesh(th, nx, ny) for (_, th) in cases]
    b = Batch(meshes, [config(m, M) for m, (M, _) in zip(meshes, cases)])

    def reset():
        for i, (m, (M, _)) in enumerate(zip(meshes, cases)):
            b.set_state(i, freestream(m, M))

    reset()
    t_setup = time.perf_counter() - t0
    if autotune_cache:
        b.autotune(autotune_cache, steps=3)
    t1 = time.perf_counter()
    used = driver
    if driver == "lusgs":
        try:
            b.set_steady_driver(STEADY_LUSGS, 1e4, 1.0)
            steps, drop, _ = b.run_to_steady(tol, lu_max, lu_check)
        except MstError:
            used = "explicit (LU-SGS left the admissible states)"
            b.set_steady_driver(STEADY_EXPLICIT)
            reset()
            steps, drop, _ = b.run_to_steady(tol, max_steps, check_every)
    else:
        steps, drop, _ = b.run_to_steady(tol, max_steps, check_every)
    t_solve = time.perf_counter() - t1
    X = np.concatenate([features(m, M, th) for m, (M, th) in zip(meshes, cases)])
    T = np.concatenate([encode(b.get_state(i), M) for i, (M, _) in enumerate(cases)])
    b.close()
    info = dict(cases=len(cases), cells_per_case=meshes[0].n_cells, samples=int(X.shape[0]), setup_s=t_setup,
                solve_s=t_solve, converged=int((drop <= tol).sum()), driver=used,
                max_steps=lu_max if used == "lusgs" else max_steps, median_steps=float(np.median(steps)))
    return X, T, info


def train(X, T, steps=(20000, 8000, 4000), lrs=(1e-3, 3e-4, 1e-4), batch=8192, width=128, hidden=4, seed=2024):
    """4 x 128 tanh network, Adam with a step schedule; returns (net, info)."""
    net = MLP(4, width, hidden, 4, seed=seed)
    mu, sd = T.mean(0), T.std(0) + 1e-6
    net.set_scaling(IN_SHIFT, IN_SCALE, mu, sd)
    net.set_data(X, T)
    t0 = time.perf_counter()
    losses = []
    for n, lr in zip(steps, lrs):
        losses.append(net.train(n, batch=batch, lr=lr))
    info = dict(train_s=time.perf_counter() - t0, steps=int(sum(steps)), batch=batch, last_batch_loss=losses[-1])
    return net, info


def initial_error(s, m, mach, theta, net, Q_steady):
    """Relative L2 distance (per field, scaled by the field's max) of the freestream and of the
    network's prediction from a converged state."""
    sc = np.abs(Q_steady).max(0)
    fs = freestream(m, mach)
    s.warm_start(net, [mach, theta])
    Qn = s.get_state()
    e = lambda Q: np.sqrt((((Q - Q_steady) / sc) ** 2).mean(0))
    return e(fs), e(Qn)


def initial_residual(s, m, mach, theta, net):
    """L2(R/V) per field of the freestream and of the network's prediction: the level each start's
    residual has to fall from.  A driver whose residual falls at a fixed rate per iteration needs
    log(start / target) / log(rate) iterations, so a start saves the fraction
    log(freestream / prediction) / log(freestream / target) of them at most."""
    V = s.prep.geometry()["vol"][:, None]
    norm = lambda R: np.sqrt(((R / V) ** 2).mean(0))
    s.set_state(freestream(m, mach))
    r_fs = norm(s.residual())
    s.warm_start(net, [mach, theta])
    r_net = norm(s.residual())
    return r_fs, r_net


def time_to_steady(s, m, mach, theta, net=None, ref=None, tol=1e-8, max_steps=60000, check_every=200):
    """One run of solver s (mesh m) from the freestream (net None) or from the network's prediction;
    returns steps, seconds (prediction included), final ratio and the reference norms.  A prediction
    the solver cannot advance falls back to the freestream start (its steps and time counted)."""
    s.set_state(freestream(m, mach))
    s.sync()
    t0 = time.perf_counter()
    used = "freestream"
    lost = 0
    if net is not None:
        try:
            s.warm_start(net, [mach, theta])
            used = "network"
        except MstError:
            s.set_state(freestream(m, mach))
    try:
        steps, ratio, r = s.run_to_steady(tol, max_steps, check_every, ref)
    except MstError:
        if used != "network":
            raise
        lost = max_steps
        s.set_state(freestream(m, mach))
        used = "fallback"
        steps, ratio, r = s.run_to_steady(tol, max_steps, check_every, ref)
    s.sync()
    wall = time.perf_counter() - t0
    return dict(steps=int(steps) + lost, seconds=wall, ratio=float(ratio), ref=r, start=used, converged=ratio <= tol)


def evaluate(net, cases, nx, ny, tol=1e-8, max_steps=60000, check_every=200, lusgs=None, lu_max=20000,
             lu_check=50):
    """Per held-out case: the explicit driver from the freestream (cold) and from the prediction (warm);
    with lusgs = (cfl, omega) also the LU-SGS driver from both starts.  Every run goes to the same
    absolute level: tol x the explicit cold run's first-step norms."""
    rows = []
    for (M, th) in cases:
        m = mesh(th, nx, ny)
        s = Solver(m, config(m, M))
        cold = time_to_steady(s, m, M, th, None, None, tol, max_steps, check_every)
        e_fs, e_net = initial_error(s, m, M, th, net, s.get_state())
        r_fs, r_net = initial_residual(s, m, M, th, net)
        warm = time_to_steady(s, m, M, th, net, cold["ref"], tol, max_steps, check_every)
        row = dict(mach=M, theta=th, cold_steps=cold["steps"], cold_s=cold["seconds"], cold_ratio=cold["ratio"],
                   warm_steps=warm["steps"], warm_s=warm["seconds"], warm_ratio=warm["ratio"],
                   warm_start=warm["start"], converged=bool(cold["converged"] and warm["converged"]),
                   speedup=cold["seconds"] / warm["seconds"],
                   freestream_error=e_fs.tolist(), network_error=e_net.tolist(),
                   freestream_residual=r_fs.tolist(), network_residual=r_net.tolist())
        if lusgs is not None:
            s.set_steady_driver(STEADY_LUSGS, *lusgs)
            s.lusgs_colours()  # the colouring is built once per mesh, outside the timed runs
            lc = time_to_steady(s, m, M, th, None, cold["ref"], tol, lu_max, lu_check)
            lw = time_to_steady(s, m, M, th, net, cold["ref"], tol, lu_max, lu_check)
            s.set_steady_driver(STEADY_EXPLICIT)
            row.update(lu_cold_steps=lc["steps"], lu_cold_s=lc["seconds"], lu_cold_ratio=lc["ratio"],
                       lu_warm_steps=lw["steps"], lu_warm_s=lw["seconds"], lu_warm_ratio=lw["ratio"],
                       lu_warm_start=lw["start"], lu_converged=bool(lc["converged"] and lw["converged"]),
                       lu_speedup=lc["seconds"] / lw["seconds"], driver_speedup=cold["seconds"] / lc["seconds"],
                       combined_speedup=cold["seconds"] / lw["seconds"])
        s.close()
        rows.append(row)
    return rows


def summary(rows):
    ok = [r for r in rows if r["converged"]]
    sp = np.array([r["speedup"] for r in ok]) if ok else np.array([np.nan])
    st = np.array([r["cold_steps"] / r["warm_steps"] for r in ok]) if ok else np.array([np.nan])
    gain = np.array([np.mean(r["freestream_error"]) / np.mean(r["network_error"]) for r in rows])
    # largest over the fields of the prediction's initial L2(R/V) over the freestream's
    rres = np.array([max(b / a for a, b in zip(r["freestream_residual"], r["network_residual"])) for r in rows])
    out = dict(cases=len(rows), converged=len(ok), geomean_speedup=float(np.exp(np.log(sp).mean())),
               geomean_initial_error_reduction=float(np.exp(np.log(gain).mean())),
               geomean_initial_residual_ratio=float(np.exp(np.log(rres).mean())),
               median_speedup=float(np.median(sp)), geomean_step_ratio=float(np.exp(np.log(st).mean())),
               total_cold_s=float(sum(r["cold_s"] for r in ok)), total_warm_s=float(sum(r["warm_s"] for r in ok)))
    lu = [r for r in rows if r.get("lu_converged") and r["converged"]]
    if lu:
        gm = lambda k: float(np.exp(np.mean([np.log(r[k]) for r in lu])))
        out.update(lusgs_converged=len(lu), lusgs_geomean_warm_speedup=gm("lu_speedup"),
                   lusgs_geomean_driver_speedup=gm("driver_speedup"),
                   lusgs_geomean_combined_speedup=gm("combined_speedup"),
                   lusgs_geomean_iteration_ratio=float(np.exp(np.mean([np.log(r["lu_cold_steps"] / r["lu_warm_steps"])
                                                                        for r in lu]))),
                   lusgs_total_cold_s=float(sum(r["lu_cold_s"] for r in lu)),
                   lusgs_total_warm_s=float(sum(r["lu_warm_s"] for r in lu)))
    return out
