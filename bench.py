"""Benchmark of the MI355X residual path: face-flux updates/s and achieved HBM bandwidth, and the
time to a 1e-8 residual drop of the steady ramp, warm against cold.

Default run (one JSON line), every configuration measured in the same process:
- headline (top-level keys): 2D isentropic vortex (strength 5, Mach 0.5 freestream, gamma 1.4) on
  [-10,10]^2, 1000x1000 quads split into 2,000,000 jittered triangles (+-0.1 h, seed 12345), periodic
  on all sides; Roe + van Leer, SSP-RK3, CFL 0.5 (device-side time step), FP64.  One "step" is one RK3
  time step = 3 residual evaluations (gradient+limiter, face flux, gather+update each) + the CFL
  reduction.  face-updates/s = N_faces * 3 * steps / wall time, inputs resident in HBM.  Every kernel
  strategy (face, cell, patch, fused) is timed on the same context; `value` is the fastest.
- configs.pulse3d: the 3D pulse, 160^3 hexes x 6 jittered tets (24,576,000 cells), AUSM + minmod,
  SSP-RK2, FP64 and the mixed form (FP32 flux and limiter, FP64 accumulation and update) with their
  ratio and per-field L2 / Linf differences, the dominant kernel's roofline and a CPU sample;
- configs.ramp: the steady supersonic compression ramp (600x250x2 = 300,000 triangles): time to a
  1e-8 drop with the explicit and LU-SGS drivers from the freestream (LU-SGS sweeps' roofline, the
  oracle's LU-SGS as CPU baseline), and the learned warm start (256 solver-generated training cases, a
  4 x 128 network trained on the GPU) against the freestream start on held-out cases.
--headline-only skips the two configs (profiler passes).  Other single workloads: --workload pulse3d,
steady, ramp, batch, mlp (the warm-start network's kernels), dropin (the C++ drop-in time loop).  Only one GPU is used: --gpus N>1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mst-cfd_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def build_workload(M, name):
    if name == "vortex2d":
        mesh = M.Mesh.gen_tri(1000, 1000, -10.0, 10.0, -10.0, 10.0, 0.1, 12345, (M.BC_PERIODIC,) * 4)
        cfg = dict(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.5)
        desc = dict(workload="isentropic vortex 2D, 1000x1000x2 jittered triangles, periodic",
                    cells=mesh.n_cells, faces=mesh.n_faces, scheme="Roe", limiter="van Leer", rk=3, cfl=0.5)
        return mesh, cfg, desc, lambda: mesh.init_state(M.INIT_VORTEX, [1.4, 0.5, 5.0, 0.0, 0.0])
    if name == "pulse3d":
        bc = (M.BC_INLET, M.BC_OUTLET) + (M.BC_SYMMETRY,) * 4
        mesh = M.Mesh.gen_tet(160, 160, 160, (0, 0, 0), (1, 1, 1), 0.15, 12345, bc)
        E = (1 / 1.4) / 0.4 + 0.5 * 0.8 * 0.8
        cfg = dict(scheme=M.AUSM, order=2, limiter=M.LIM_MINMOD, rk=2, cfl=0.5, dim=3,
                   q_inlet=[1.0, 0.8, 0.0, 0.0, E])
        desc = dict(workload="3D Gaussian pulse in a Mach-0.8 box flow, 160^3 hexes x 6 jittered tets",
                    cells=mesh.n_cells, faces=mesh.n_faces, scheme="AUSM", limiter="minmod", rk=2, cfl=0.5)
        return mesh, cfg, desc, lambda: mesh.init_state(M.INIT_PULSE, [1.4, 0.1, 0.1, 0.5, 0.5, 0.5, 0.8])
    raise SystemExit(f"unknown workload {name}")


# HBM bytes per launch of a workload's dominant kernel from the committed rocprofv3 counter passes
# (FETCH_SIZE and WRITE_SIZE scaled by a plain copy kernel of known size); None where not profiled
TRAFFIC_PROFILES = {("vortex2d", st, k): ("profiles/hbm_traffic_vortex2d.md", name)
                    for st, k, name in (("face", "k_grad", "k_grad<2, 3, 2, double>"),
                                        ("face", "k_flux", "k_flux<2, 0, true, double>"),
                                        ("cell", "k_grad", "k_grad<2, 3, 2, double>"),
                                        ("cell", "k_cell", "k_cell<2, 0, true, double, 3>"),
                                        ("patch", "k_grad", "k_grad<2, 3, 2, double>"),
                                        ("patch", "k_patch", "k_patch<2, 0, true, double, 3, true>"),
                                        ("fused", "k_fused", "k_fused<2, 0, 2, double, 3>"))}
TRAFFIC_PROFILES[("steady", "lusgs", "k_lu_sweep")] = ("profiles/hbm_traffic_lusgs_ramp300k.md", "k_lu_sweep")
TRAFFIC_PROFILES.update({("pulse3d", "patch", k): ("profiles/hbm_traffic_pulse3d.md", name)
                         for k, name in (("k_grad", "k_grad<3, 4, 1, double>"),
                                         ("k_patch", "k_patch<3, 1, true, double, 4, true>"))})


def committed_traffic(workload, strategy, kernel):
    entry = TRAFFIC_PROFILES.get((workload, strategy, kernel))
    if not entry:
        return None, None
    path, name = entry
    try:
        for line in open(os.path.join(ROOT, path)):
            if line.startswith(f"traffic_bytes_per_launch {name}:"):
                return float(line.split(":")[1]), path
    except OSError:
        pass
    return None, None


# One launch in TIMING_STRIDE of each kernel class carries HIP events inside the timed region: an
# event-carrying dispatch costs about 3 us of throughput (12 us per step of the 2M-triangle vortex when
# every launch carries them).  5 is prime to the 2 and 3 stages of a step, so the sampled launches
# cover every stage alike.
TIMING_STRIDE = 5


def measure(M, solver, steps, warmup):
    """Run warmup steps, then time exactly `steps` steps (stream-synchronised on both sides), with HIP
    events recorded on the solver's stream by one launch in TIMING_STRIDE of each kernel class:
    `launches` and `total_ms` are over the timed launches, `avg_us` their mean."""
    solver.step(warmup)
    solver.sync()
    solver.timing_reset()
    solver.timing(TIMING_STRIDE)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(steps)
    solver.sync()
    t1 = time.perf_counter()
    solver.timing(False)
    kern = {}
    for k, name in M.KERNEL_NAMES.items():
        n, ms = solver.timing_get(k)
        if n:
            b = solver.algorithmic_bytes(k)
            kern[name] = dict(launches=n, avg_us=1e3 * ms / n, total_ms=ms, alg_bytes=b,
                              gbs=b / (1e-3 * ms / n) / 1e9, timed_one_in=TIMING_STRIDE)
    return t1 - t0, kern


def field_errors(Qm, Qd):
    """Per-field relative difference of two states: L2 = ||dQ_k||_2 / ||Q_k||_2 and
    Linf = max|dQ_k| / max|Q_k|; the momentum components share the norm of the momentum field."""
    nv = Qd.shape[1]
    d = Qm - Qd
    n2 = np.sqrt((Qd ** 2).sum(0))
    n2[1:nv - 1] = np.sqrt((Qd[:, 1:nv - 1] ** 2).sum())
    ninf = np.abs(Qd).max(0)
    ninf[1:nv - 1] = ninf[1:nv - 1].max()
    return (np.sqrt((d ** 2).sum(0)) / n2).tolist(), (np.abs(d).max(0) / ninf).tolist()


def cpu_baseline(mesh, cfg_kw, Q, threads, budget_s=10.0):
    """Oracle (the CPU restatement of the reference path, OpenMP) on one RK3 step of the same mesh."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as O

    O.set_threads(threads)
    om = O.OracleMesh(mesh.raw())
    co = O.make_config(**cfg_kw)
    Qc = Q
    nsteps, t = 0, 0.0
    while nsteps < 1 or (t < budget_s and nsteps < 200):
        dt = om.cfl_dt(co, Qc)
        t0 = time.perf_counter()
        Qc, _ = om.rk_step(co, Qc, dt)
        t += time.perf_counter() - t0
        nsteps += 1
    return dict(value=mesh.n_faces * cfg_kw["rk"] * nsteps / t, unit="face-updates/s", cores=threads, kind="port",
                sample=f"{nsteps} RK{cfg_kw['rk']} steps ({cfg_kw['rk']} residual evaluations each) of the same "
                       f"{mesh.n_cells}-cell workload, oracle/mst_oracle.c with OpenMP, {t:.2f} s")


def host_cpu():
    """The host the CPU baseline ran on: processor model, logical CPUs and those this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = None
    return dict(cpu_model=model, nproc=os.cpu_count(), usable_cpus=usable)


def batch_cases(M, n=512, seed=777):
    """512 Sod/Riemann variants on [0,1] x [0,0.01] strips of 2 x nx quads (N_c = 4 nx uniform in
    [500, 50000]), rho_R = 0.125, p_R = 0.1, rho_L/rho_R ~ U[1, 10], Roe or AUSM+ with probability
    1/2; p_L/p_R = 10^U[log10 2, 2] for Roe cases and 10^U[log10 2, log10 20] for AUSM+ cases (above
    a ratio of about 25 AUSM+ loses positivity in the first steps on these strips, at any CFL number
    tried down to 0.05, in the oracle as on the device); walls all round, minmod, SSP-RK2, CFL 0.5."""
    rng = np.random.default_rng(seed)
    meshes, cfgs, Q0 = [], [], []
    for i in range(n):
        nc = rng.uniform(500, 50000)
        nx = max(2, int(round(nc / 4)))
        roe = rng.uniform() < 0.5
        hi = 2.0 if roe else math.log10(20.0)
        pr = 10 ** (math.log10(2) + rng.uniform() * (hi - math.log10(2)))
        dr = rng.uniform(1, 10)
        m = M.Mesh.gen_tri(nx, 2, 0.0, 1.0, 0.0, 0.01, 0.1, seed * 1000 + i)
        meshes.append(m)
        cfgs.append(dict(scheme=M.ROE if roe else M.AUSM, order=2, limiter=M.LIM_MINMOD, rk=2, cfl=0.5,
                         compat=M.COMPAT_REFERENCE if roe else M.COMPAT_GRAD_LASTINT))
        Q0.append(m.init_state(M.INIT_RIEMANN, [1.4, 0, 0.5, 0.125 * dr, 0, 0, 0.1 * pr, 0.125, 0, 0, 0.1]))
    return meshes, cfgs, Q0


def bench_batch(M, args):
    """Batch throughput: all cases advance one SSP-RK2 step per graph replay."""
    t0 = time.perf_counter()
    meshes, cfgs, Q0 = batch_cases(M)
    cache = os.environ.get("MSTCFD_TUNE_CACHE") or os.path.join(
        os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache"), "mstcfd", "batch_tune.json")
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    out_forms = {}
    for form, prec in (("fp64", M.PREC_FP64), ("mixed", M.PREC_MIXED)):
        b = M.Batch(meshes, [M.make_config(precision=prec, **c) for c in cfgs])
        for i, q in enumerate(Q0):
            b.set_state(i, q)
        t_setup = time.perf_counter() - t0
        tune_ms = b.autotune(cache, force=args.retune, steps=5)
        block, tiles = b.get_tuning()
        res = {}
        for graph in (True, False):
            for i, q in enumerate(Q0):  # both timings advance the same steps from the initial field
                b.set_state(i, q)
            b.set_graph(graph)
            b.step(args.warmup)  # mst_batch_step returns after the stream has drained
            t1 = time.perf_counter()
            b.step(args.steps)
            wall = time.perf_counter() - t1
            res["graph" if graph else "direct"] = wall
        n_items = b.counts()
        finite = all(np.isfinite(b.get_state(i)).all() for i in range(0, len(meshes), 37))
        faces = b.n_faces
        out_forms[form] = dict(value=faces * 2 * args.steps / res["graph"], ms_per_step=1e3 * res["graph"] / args.steps,
                               direct_ms_per_step=1e3 * res["direct"] / args.steps, block=block, tiles=tiles,
                               strategy={M.STRATEGY_FACE: "face", M.STRATEGY_CELL: "cell", M.STRATEGY_PATCH: "patch",
                                         M.STRATEGY_FUSED: "fused"}.get(b.strategy, b.strategy),
                               autotune_ms_per_step=tune_ms, cell_items=n_items[3], face_items=n_items[4],
                               state_finite=bool(finite))
        b.close()
    v = out_forms["fp64"]
    out = dict(metric="batch face-flux updates/s (512 uneven Riemann/Sod cases, graph-replayed steps)",
               value=v["value"], unit="face-updates/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=v["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload="512 Sod/Riemann variants, N_c ~ U[500,50000] triangles, Roe/AUSM per case",
                           cases=len(meshes), cells=int(sum(m.n_cells for m in meshes)),
                           faces=int(sum(m.n_faces for m in meshes)), limiter="minmod", rk=2, cfl=0.5,
                           setup_s=t_setup, tune_cache=cache),
               fp64=v, mixed=out_forms["mixed"], mixed_speedup=out_forms["mixed"]["value"] / v["value"])
    print(json.dumps(out))


def bench_ramp(M, args):
    """Learned warm start (compression ramp): 256 solver-generated steady cases (75,000 triangles
    each, solved in one batch), a 4 x 128 network trained on the GPU, then 16 held-out cases on the
    300,000-triangle ramp from the freestream and from the network's prediction to the same
    residual level (1e-8 x the freestream run's first-step L2(R/V), every field), with the explicit
    SSP-RK2 driver and with the LU-SGS driver.  value: the warm start's gain under LU-SGS; the
    summary also holds the explicit driver's warm-start gain, the LU-SGS driver's gain over the
    explicit one (both from the freestream) and the two combined."""
    from mstcfd import ramp as RP
    t0 = time.perf_counter()
    train_cases = RP.sample_cases(256, 2024)
    X, T, gen = RP.generate(train_cases, 300, 125, max_steps=20000, check_every=500)
    net, tr = RP.train(X, T)
    test_cases = RP.sample_cases(16, 2025)
    rows = RP.evaluate(net, test_cases, 600, 250, max_steps=40000, lusgs=(1e4, 1.0))
    sm = RP.summary(rows)
    out = dict(metric="time-to-1e-8 residual drop, freestream start / learned warm start (LU-SGS steady driver)",
               value=sm.get("lusgs_geomean_warm_speedup"), unit="x (geometric mean over held-out cases that converge)",
               n_gpus=1, steps=None, warmup=None, ms_per_step=None, higher_is_better=True, scaling="strong",
               vs_baseline=None, dtype="f64 solver / f32 network", data="synthetic (solver-generated)",
               config=dict(workload="steady supersonic compression ramp, 600x250x2 = 300,000 triangles, first-order Roe; "
                                    "LU-SGS (CFL 1e4, omega 1) and SSP-RK2 (CFL 0.8) steady drivers; "
                                    "M ~ U[1.5,3], theta ~ U[5,20] deg",
                           training=gen, network=dict(tr, layers="4-128-128-128-128-4 tanh"),
                           wall_s=time.perf_counter() - t0),
               summary=sm, cases=rows)
    print(json.dumps(out))


def bench_dropin(M, args):
    """The drop-in loop: the reference's time loop (setDT, solve, the max |New-Old|/Old monitor,
    updateNewToOld; time/Time.cpp:54-80) in C++ over include/mstcfd.hpp, on 1000 x 1000 x 2 = 2,000,000
    triangles (first-order Roe, forward Euler, fixed DT), with the monitor over the host mirrors (the
    reference's loop as written: the new state crosses PCIe every step) and reduced on the device.
    The PCIe-inclusive rate is reported here; it is never a headline value."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "mst-cfd_amd", "lib", "rhosolver_driver")
    res = {}
    with tempfile.TemporaryDirectory() as td:
        for mon in ("host", "device"):
            r = subprocess.run([exe, os.path.join(td, "q.bin"), str(args.steps), "1e-5", mon, "1000", "1000"],
                               capture_output=True, text=True, timeout=600)
            if r.returncode:
                raise SystemExit(f"rhosolver_driver failed: {r.stdout} {r.stderr}")
            ms = float(r.stdout.split("ms_per_step")[1].split()[0])
            res[mon] = dict(ms_per_step=ms, state_bytes_read_back_per_step=(2_000_000 * 32 if mon == "host" else 32))
    out = dict(metric="drop-in time loop (C++ RhoSolverGPU, the reference's Time::goNextTimeStep) ms per step",
               value=res["device"]["ms_per_step"], unit="ms/step", n_gpus=1, steps=args.steps, warmup=0,
               ms_per_step=res["device"]["ms_per_step"], higher_is_better=False, scaling="strong", vs_baseline=None,
               dtype="f64", data="synthetic",
               config=dict(workload="Sod strip, 1000x1000x2 = 2,000,000 triangles, first-order Roe, forward Euler"),
               monitors=res, host_over_device=res["host"]["ms_per_step"] / res["device"]["ms_per_step"])
    print(json.dumps(out))


def bench_mlp(M, args):
    """The warm-start network's kernels (4 -> 128 x 4 -> 4, tanh, FP32): `steps` Adam steps on batches of
    8192 rows of a 2M-row synthetic training set, each kernel timed by start/stop HIP events carried by
    its own dispatch (the library's timing; no profiler).  GEMM rates are 2 M N K per launch over the
    launch's duration, against MI355X's dense FP32 matrix peak (157.3 TFLOP/s)."""
    rng = np.random.default_rng(2024)
    n, batch = 2_000_000, 8192
    X = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 2], np.cos(X[:, 1]) + X[:, 3], X[:, 0] * X[:, 1], np.tanh(3 * X[:, 2])],
                 axis=1).astype(np.float32)
    net = M.MLP(4, 128, 4, 4, seed=2024)
    net.set_data(X, T)
    net.train(args.warmup, batch=batch, lr=1e-3, graph=True)
    t0 = time.perf_counter()
    net.train(args.steps, batch=batch, lr=1e-3, graph=True)  # graph replay: the training rate
    wall = time.perf_counter() - t0
    net.timing(True)
    net.timing_reset()
    net.train(args.steps, batch=batch, lr=1e-3, graph=True)  # timed: direct launches
    kern = {}
    gemm_ms = gemm_flop = 0.0
    for k, name in M.MLP_KERNEL_NAMES.items():
        cnt, ms, fl = net.timing_get(k)
        if not cnt:
            continue
        kern[name] = dict(launches=cnt, avg_us=1e3 * ms / cnt, total_ms=ms,
                          tflops=(fl / (ms * 1e-3) / 1e12) if fl else None)
        if fl:
            gemm_ms += ms
            gemm_flop += fl
    net.timing(False)
    net.close()
    ach = gemm_flop / (gemm_ms * 1e-3) / 1e12
    out = dict(metric="learned warm start network: training steps/s (graph replay) and GEMM TFLOP/s",
               value=args.steps / wall, unit="training steps/s (batch 8192)", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="strong",
               vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload="4-128-128-128-128-4 tanh network, Adam, batch 8192 of 2,000,000 rows"),
               roofline=dict(bound="mfma", kernel="k_gemm (all GEMMs of a training step)", achieved=ach,
                             peak=157.3, unit="TFLOP/s", frac=ach / 157.3, traffic=None),
               kernels=kern, timed_ms_per_step=sum(v["total_ms"] for v in kern.values()) / args.steps)
    print(json.dumps(out))


def bench_steady(M, args, emit=True):
    """Steady supersonic compression ramp (the C5 mesh: 600x250x2 = 300,000 triangles; M 2.51,
    theta 9.3 deg, first-order Roe) from the freestream to 1e-8 x the first-step L2(R/V) of every
    field, with the explicit SSP-RK2 driver (CFL 0.8, device time step) and with the LU-SGS driver
    (pseudo-time CFL 1e4, omega 1).  value: the LU-SGS time to that level (inputs resident in HBM,
    the colouring built before the timed run); the line also holds the explicit time, the
    iteration counts, the per-kernel times of LU-SGS iterations and the sweeps' roofline."""
    from mstcfd import ramp as RP
    mach, theta, tol = 2.51, 9.3, 1e-8
    copy = None
    if args.calibrate_copy:
        M.copy_probe(args.calibrate_copy)  # first launch pays the code-object load
        ms = min(M.copy_probe(args.calibrate_copy) for _ in range(3))
        copy = dict(bytes=args.calibrate_copy, ms=ms, gbs=2 * args.calibrate_copy / (ms * 1e-3) / 1e9)
    m = RP.mesh(theta, 600, 250)
    s = M.Solver(m, RP.config(m, mach))
    q0 = RP.freestream(m, mach)
    runs = {}
    ref = None
    for name, drv in (("explicit", M.STEADY_EXPLICIT), ("lusgs", M.STEADY_LUSGS)):
        s.set_steady_driver(drv, 1e4, 1.0)
        if drv == M.STEADY_LUSGS:
            _, ncol = s.lusgs_colours()
        s.set_state(q0)
        s.sync()
        t0 = time.perf_counter()
        cap = args.max_steps or (60000 if drv == M.STEADY_EXPLICIT else 20000)
        steps, drop, r = s.run_to_steady(tol, cap, min(cap, 200 if drv == M.STEADY_EXPLICIT else 50), ref)
        s.sync()
        runs[name] = dict(seconds=time.perf_counter() - t0, iterations=int(steps), drop=float(drop),
                          converged=bool(drop <= tol))
        ref = r if ref is None else ref
    # per-kernel times of LU-SGS iterations (dispatch-carried HIP events on the solver's stream)
    s.set_state(q0)
    s.lusgs_step(args.warmup)
    s.sync()
    s.timing_reset()
    s.timing(True)
    s.lusgs_step(args.steps)
    s.sync()
    s.timing(False)
    kern = {}
    for k, name in M.KERNEL_NAMES.items():
        n, ms = s.timing_get(k)
        if n:
            kern[name] = dict(launches=n, avg_us=1e3 * ms / n, per_iteration_us=1e3 * ms / args.steps)
            if k == M.K_LUSGS:  # the residual kernels run in residual-only mode here
                b = s.algorithmic_bytes(k)
                kern[name].update(alg_bytes=b, gbs=b / (1e-3 * ms / n) / 1e9)
    sw = kern["k_lu_sweep"]
    ach = sw["alg_bytes"] / (sw["avg_us"] * 1e-6) / 1e9
    traffic, tsrc = committed_traffic("steady", "lusgs", "k_lu_sweep")
    lu, ex = runs["lusgs"], runs["explicit"]
    out = dict(metric="time-to-1e-8 residual drop from the freestream (steady ramp, LU-SGS driver)",
               value=lu["seconds"], unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=1e3 * lu["seconds"] / max(lu["iterations"], 1), higher_is_better=False,
               scaling="strong", vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="steady supersonic compression ramp, 600x250x2 = 300,000 triangles, "
                                    f"M {mach}, theta {theta} deg, first-order Roe; LU-SGS CFL 1e4, omega 1, "
                                    f"{ncol} colours; explicit SSP-RK2 CFL 0.8 alongside",
                           cells=m.n_cells, faces=m.n_faces),
               drivers=runs, speedup_vs_explicit=ex["seconds"] / lu["seconds"],
               roofline=dict(bound="hbm", kernel="k_lu_sweep", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=ach / HBM_PEAK_GBS, traffic=traffic, traffic_source=tsrc, alg_bytes=sw["alg_bytes"]),
               kernels=kern)
    if copy:
        out["copy_probe"] = copy
    if args.max_steps:
        out["note"] = f"to-steady runs capped at {args.max_steps} iterations (profiler pass): value is not a time to 1e-8"
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = lusgs_cpu_baseline(m, mach, s.lusgs_colours()[0], lu["iterations"])
        out["cpu_baseline"].update(host_cpu())
    s.close()
    if emit:
        print(json.dumps(out))
    return out


def lusgs_cpu_baseline(m, mach, colour, iterations, budget_s=10.0):
    """The oracle's LU-SGS iteration (OpenMP residual, sequential colour sweeps) on the same mesh and
    colours: a bounded sample of iterations, scaled to the GPU run's iteration count."""
    sys.path.insert(0, ROOT)
    from mstcfd import ramp as RP
    from oracle import pyoracle as O

    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    om = O.OracleMesh(m.raw())
    c = RP.config(m, mach)
    co = O.make_config(scheme=O.ROE, order=1, q_inlet=[c.q_inlet[i] for i in range(4)])
    Q = RP.freestream(m, mach)
    n, t = 0, 0.0
    while n < 1 or (t < budget_s and n < 100):
        t0 = time.perf_counter()
        Q, _ = om.lusgs_step(co, Q, colour, 1e4, 1.0)
        t += time.perf_counter() - t0
        n += 1
    return dict(value=t / n * iterations, unit="s", cores=threads, kind="port",
                sample=f"{n} LU-SGS iterations of the same 300,000-cell case ({t:.2f} s, oracle/mst_oracle.c, "
                       f"OpenMP residual, sequential sweeps), scaled to the GPU run's {iterations} iterations")


def bench_unsteady(M, args, workload, copy=None, cpu_sample_mesh=None):
    """One unsteady workload on one context: every strategy timed in FP64 and in the mixed form, the
    mixed form's state difference after compare_steps steps with the same time steps, the dominant
    kernel's roofline and the CPU baseline (on cpu_sample_mesh() when given: a smaller mesh of the same
    construction, for a mesh too large for a bounded CPU sample)."""
    mesh, cfg_kw, desc, init = build_workload(M, workload)
    t_setup = time.perf_counter()
    solver = M.Solver(mesh, M.make_config(**cfg_kw))
    t_setup = time.perf_counter() - t_setup
    Q0 = init()
    stages = cfg_kw["rk"]
    names = {"face": M.STRATEGY_FACE, "cell": M.STRATEGY_CELL, "patch": M.STRATEGY_PATCH, "fused": M.STRATEGY_FUSED}
    tried = tuple(names) if args.strategy == "auto" else (args.strategy,)
    runs = {}
    for st in tried:  # each strategy on the same context from the same initial field
        solver.strategy = names[st]
        solver.set_state(Q0)
        wall, kern = measure(M, solver, args.steps, args.warmup)
        dom = max(kern, key=lambda k: kern[k]["total_ms"])
        runs[st] = dict(wall=wall, kernels=kern, dominant=dom)
    best = max(runs, key=lambda st: 1.0 / runs[st]["wall"])
    wall, kern, dom = runs[best]["wall"], runs[best]["kernels"], runs[best]["dominant"]
    solver.strategy = names[best]
    value = mesh.n_faces * stages * args.steps / wall
    desc = dict(desc, strategy=best, cell_order=args.order, context_setup_s=t_setup)

    def roofline(st):
        r = runs[st]
        d = r["kernels"][r["dominant"]]
        ach = d["alg_bytes"] / (d["avg_us"] * 1e-6) / 1e9
        traffic, src = committed_traffic(workload, st, r["dominant"])
        return dict(bound="hbm", kernel=r["dominant"], achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=ach / HBM_PEAK_GBS, traffic=traffic, traffic_source=src, alg_bytes=d["alg_bytes"])

    out = dict(
        metric="face-flux updates/s (FP64 residual path); %HBM peak of the dominant kernel",
        value=value, unit="face-updates/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic", config=desc, roofline=roofline(best), kernels=kern,
        strategies={st: dict(value=mesh.n_faces * stages * args.steps / runs[st]["wall"],
                             ms_per_step=1e3 * runs[st]["wall"] / args.steps, roofline=roofline(st),
                             kernels=runs[st]["kernels"]) for st in runs},
    )
    if copy:
        out["copy_probe"] = copy
    if not args.no_mixed:
        # mixed form on the same context: throughput, then the state difference after `compare_steps`
        # steps of both forms from the same initial field with the same dt sequence
        # the mixed form times every strategy too: the fastest one differs from the FP64 form's
        solver.set_config(M.make_config(precision=M.PREC_MIXED, **cfg_kw))
        runs_m = {}
        for st in tried:
            solver.strategy = names[st]
            solver.set_state(Q0)
            runs_m[st] = measure(M, solver, args.steps, args.warmup)
        best_m = min(runs_m, key=lambda st: runs_m[st][0])
        wall_m, kern_m = runs_m[best_m]
        value_m = mesh.n_faces * stages * args.steps / wall_m
        nsteps = args.compare_steps
        fixed = dict(cfg_kw, cfl=0.0)
        solver.set_config(M.make_config(**fixed))
        solver.strategy = names[best]
        solver.set_state(Q0)
        dts = []
        for _ in range(nsteps):
            dts.append(solver.compute_dt(cfg_kw["cfl"]))
            solver.set_dt(dts[-1])
            solver.step(1)
        Qd = solver.get_state()
        solver.set_config(M.make_config(precision=M.PREC_MIXED, **fixed))
        solver.strategy = names[best_m]
        solver.set_state(Q0)
        for dt in dts:
            solver.set_dt(dt)
            solver.step(1)
        l2, linf = field_errors(solver.get_state(), Qd)
        out["mixed"] = dict(value=value_m, ms_per_step=1e3 * wall_m / args.steps, speedup=value_m / value,
                            dtype="fp32 flux+limiter / fp64 gather+update", strategy=best_m, kernels=kern_m,
                            strategies={st: 1e3 * runs_m[st][0] / args.steps for st in runs_m},
                            compare_steps=nsteps, rel_l2_per_field=l2, rel_linf_per_field=linf)
        solver.set_config(M.make_config(**cfg_kw))
        solver.strategy = names[best]
    # the fused-minimum lower bound: per stage each face's two cell indices and normal and
    # each cell's state in, volume and state out, once (2D: 24 N_f + 72 N_c bytes)
    dim = mesh.dim
    lb = mesh.n_faces * (8 + 8 * dim) + mesh.n_cells * (2 * 8 * (dim + 2) + 8)
    out["fused_minimum"] = dict(bytes_per_stage=lb, achieved_gbs=lb * stages * args.steps / wall / 1e9,
                                frac=lb * stages * args.steps / wall / 1e9 / HBM_PEAK_GBS)
    Qf = solver.get_state()
    out["state_finite"] = bool(np.isfinite(Qf).all())
    solver.close()
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        if cpu_sample_mesh is not None:
            del Qf
            cmesh, cinit = cpu_sample_mesh()
            cQ0 = cinit()
        else:
            cmesh, cQ0 = mesh, Q0
        out["cpu_baseline"] = cpu_baseline(cmesh, cfg_kw, cQ0, threads)
        out["cpu_baseline"].update(host_cpu())
        if threads != 8:  # the reference's own OpenMP team size (NUM_CPU_THREADS, include/CONST.h:7)
            t8 = cpu_baseline(cmesh, cfg_kw, cQ0, 8, budget_s=5.0)
            out["cpu_baseline"]["threads8"] = dict(value=t8["value"], sample=t8["sample"])
    return out


def pulse_sample_mesh(M):
    """The pulse workload's construction on 64^3 hexes x 6 tets (1,572,864 cells): the CPU sample of
    the 24.6M-cell configuration (the oracle's unit is face-updates/s, so a bounded sample of the
    same mesh construction and scheme measures the same rate)."""
    bc = (M.BC_INLET, M.BC_OUTLET) + (M.BC_SYMMETRY,) * 4
    mesh = M.Mesh.gen_tet(64, 64, 64, (0, 0, 0), (1, 1, 1), 0.15, 12345, bc)
    return mesh, lambda: mesh.init_state(M.INIT_PULSE, [1.4, 0.1, 0.1, 0.5, 0.5, 0.5, 0.8])


def ramp_section(M, args, held_out=8):
    """C5 in the default line: the steady runs of bench_steady (one case, explicit and LU-SGS from the
    freestream, with the sweeps' roofline and the oracle's LU-SGS as CPU baseline), then the learned
    warm start: 256 training cases solved in one batch, the network trained on the GPU, `held_out`
    cases of the 300,000-triangle ramp from the freestream and from the prediction to the same
    absolute residual level."""
    from mstcfd import ramp as RP
    steady = bench_steady(M, args, emit=False)
    t0 = time.perf_counter()
    X, T, gen = RP.generate(RP.sample_cases(256, 2024), 300, 125, max_steps=20000, check_every=500)
    net, tr = RP.train(X, T)
    rows = RP.evaluate(net, RP.sample_cases(held_out, 2025), 600, 250, max_steps=40000, lusgs=(1e4, 1.0))
    net.close()
    sm = RP.summary(rows)
    lu = steady["drivers"]["lusgs"]
    return dict(metric="time-to-1e-8 residual drop (L2 of R/V, every field), steady ramp",
                config=steady["config"], unit="s",
                cold=dict(explicit_s=steady["drivers"]["explicit"]["seconds"], lusgs_s=lu["seconds"],
                          explicit_iterations=steady["drivers"]["explicit"]["iterations"],
                          lusgs_iterations=lu["iterations"], lusgs_speedup_vs_explicit=steady["speedup_vs_explicit"]),
                roofline=steady["roofline"], kernels=steady["kernels"], cpu_baseline=steady.get("cpu_baseline"),
                warm_start=dict(value=sm.get("lusgs_geomean_warm_speedup"),
                                unit="x, freestream time / warm-start time (geometric mean over held-out cases that "
                                     "converge), LU-SGS driver",
                                explicit_driver=sm.get("geomean_speedup"), summary=sm,
                                training=gen, network=dict(tr, layers="4-128-128-128-128-4 tanh"),
                                held_out=held_out, wall_s=time.perf_counter() - t0, cases=rows))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="vortex2d")
    ap.add_argument("--headline-only", action="store_true",
                    help="default workload without the pulse3d and ramp configurations (profiler passes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--retune", action="store_true", help="batch workload: ignore the tuning cache")
    ap.add_argument("--strategy", choices=("auto", "face", "cell", "patch", "fused"), default="auto",
                    help="face: flux kernel + gather/update kernel; cell: fused cell kernel; patch: fused patch "
                         "kernel (face fluxes of a 256-cell patch in LDS); fused: the patch kernel with the "
                         "gradients and face states in LDS too; auto: time each and report the fastest")
    ap.add_argument("--order", choices=("hilbert", "rcm"), default="hilbert",
                    help="cell order of the device layout (Hilbert-curve patches or reverse Cuthill-McKee)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison")
    ap.add_argument("--max-steps", type=int, default=0,
                    help="steady workload: cap the to-steady runs at this many steps/iterations (profiler passes; "
                         "the line then carries a note and no time to 1e-8); 0: no cap")
    ap.add_argument("--compare-steps", type=int, default=200,
                    help="steps of the FP64 vs mixed state comparison")
    ap.add_argument("--held-out", type=int, default=8, help="ramp cases of the warm-start comparison in the default line")
    ap.add_argument("--calibrate-copy", type=int, default=268435456, metavar="BYTES",
                    help="first launch one plain copy kernel of BYTES (the achievable-bandwidth reference and the "
                         "scale for hardware byte counters; 0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the multi-GPU path is a "
              f"design (mesh partition + halo exchange), not built in this tree.", file=sys.stderr)
        sys.exit(2)
    import mstcfd as M

    M.set_cell_order(M.ORDER_HILBERT if args.order == "hilbert" else M.ORDER_RCM)
    if args.workload == "batch":
        return bench_batch(M, args)
    if args.workload == "ramp":
        return bench_ramp(M, args)
    if args.workload == "steady":
        return bench_steady(M, args)
    if args.workload == "mlp":
        return bench_mlp(M, args)
    if args.workload == "dropin":
        return bench_dropin(M, args)
    t0 = time.perf_counter()
    copy = None
    if args.calibrate_copy:
        M.copy_probe(args.calibrate_copy)  # first launch pays the code-object load
        ms = min(M.copy_probe(args.calibrate_copy) for _ in range(3))
        copy = dict(bytes=args.calibrate_copy, ms=ms, gbs=2 * args.calibrate_copy / (ms * 1e-3) / 1e9)
    if args.workload == "pulse3d":
        print(json.dumps(bench_unsteady(M, args, "pulse3d", copy, lambda: pulse_sample_mesh(M))))
        return
    out = bench_unsteady(M, args, args.workload, copy)
    if args.workload == "vortex2d" and not args.headline_only:
        print(f"bench: vortex2d done at {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        c3 = bench_unsteady(M, args, "pulse3d", None, lambda: pulse_sample_mesh(M))
        print(f"bench: pulse3d done at {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        c5 = ramp_section(M, args, args.held_out)
        print(f"bench: ramp done at {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        out["configs"] = dict(pulse3d=c3, ramp=c5)
    out["wall_s"] = time.perf_counter() - t0
    print(json.dumps(out))


if __name__ == "__main__":
    main()
