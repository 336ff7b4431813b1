"""Network kernels (forward, backward, Adam, graph-replayed training) against the CPU reference
network, and the decoding of the learned warm start."""
import numpy as np
import pytest

import mlp_ref as R
import mstcfd as M

pytestmark = pytest.mark.gpu


def make(width=128, hidden=4, seed=3):
    net = M.MLP(4, width, hidden, 4, seed=seed)
    rng = np.random.default_rng(seed)
    sc = tuple(np.asarray(a, np.float32) for a in (rng.uniform(-0.5, 0.5, 4), rng.uniform(0.5, 2, 4),
                                                     rng.uniform(-0.5, 0.5, 4), rng.uniform(0.5, 2, 4)))
    net.set_scaling(*sc)
    return net, sc


def data(n, seed=4):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 2], np.cos(X[:, 1]) + X[:, 3], X[:, 0] * X[:, 1], np.tanh(3 * X[:, 2])],
                 axis=1).astype(np.float32)
    return X, T


def test_forward_and_gradient_match_cpu_reference():
    net, sc = make()
    X, T = data(3000)
    p = net.get_params().astype(np.float64)
    layers = net.layers()
    _, Yr = R.forward(p, layers, X.astype(np.float64), [a.astype(np.float64) for a in sc])
    Y = net.forward(X)
    assert np.abs(Y - Yr).max() <= 2e-5 * np.abs(Yr).max()
    loss_r, g_r = R.loss_and_grad(p, layers, X.astype(np.float64), T.astype(np.float64),
                                  [a.astype(np.float64) for a in sc])
    g, loss = net.gradient(X, T)
    rel = np.linalg.norm(g - g_r) / np.linalg.norm(g_r)
    print(f"MLP 4x128: loss {loss:.6e} vs {loss_r:.6e}; gradient rel. error {rel:.2e}", flush=True)
    assert abs(loss - loss_r) <= 1e-5 * loss_r
    assert rel <= 1e-4
    # per layer, so that no small block hides behind a large one
    o = 0
    for (n_o, n_i) in layers:
        for size in (n_o * n_i, n_o):
            blk = slice(o, o + size)
            assert np.linalg.norm(g[blk] - g_r[blk]) <= 1e-4 * np.linalg.norm(g_r[blk]) + 1e-9
            o += size


def test_full_batch_adam_matches_cpu_reference_and_graph_replay():
    X, T = data(500, seed=5)
    out = []
    for graph in (False, True):
        net, sc = make(seed=6)
        p0 = net.get_params().astype(np.float64)
        net.set_data(X, T)
        net.train(5, batch=500, lr=1e-3, graph=graph)
        out.append(net.get_params())
    assert np.array_equal(out[0], out[1])
    layers = net.layers()
    scd = [a.astype(np.float64) for a in sc]
    p, m, v = p0, np.zeros_like(p0), np.zeros_like(p0)
    for t in range(1, 6):
        _, g = R.loss_and_grad(p, layers, X.astype(np.float64), T.astype(np.float64), scd)
        p, m, v = R.adam(p, g, m, v, t, lr=1e-3)
    # Adam's early steps move every parameter by ~lr: compare the displacement
    d_gpu, d_ref = out[0] - p0, p - p0
    rel = np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref)
    print(f"5 full-batch Adam steps: displacement rel. error {rel:.2e}", flush=True)
    assert rel <= 2e-3


def test_graph_is_rebuilt_when_its_buffers_change():
    """The captured training step points at the training set and the split buffer: a forward pass, a
    larger gradient call (which grows the split buffer) and a new training set between training
    calls must not leave it replaying into freed memory.  Graph replay and direct launches of the
    same call sequence end on the same parameters bit for bit."""
    X1, T1 = data(6000, seed=7)
    X2, T2 = data(9000, seed=8)
    out = []
    for graph in (False, True):
        net, _ = make(seed=9)
        net.forward(data(200000, seed=10)[0])
        net.set_data(X1, T1)
        net.train(10, batch=4096, lr=1e-3, graph=graph)
        net.gradient(*data(100000, seed=11))
        net.train(10, batch=4096, lr=1e-3, graph=graph)
        net.set_data(X2, T2)
        net.train(10, batch=4096, lr=1e-3, graph=graph)
        out.append(net.get_params())
        net.close()
    assert np.array_equal(out[0], out[1])


def test_kernel_timing_counts_and_leaves_results_unchanged():
    """Per-dispatch timing of the network's kernels: the launches and GEMM operation counts of a step
    follow from the layer shapes, and a timed run (direct launches) ends on the graph run's
    parameters bit for bit."""
    X, T = data(20000, seed=12)
    out = []
    for timed in (False, True):
        net, _ = make(seed=13)
        net.set_data(X, T)
        if timed:
            net.timing(True)
        net.train(6, batch=2048, lr=1e-3, graph=True)
        out.append(net.get_params())
        if timed:
            layers = net.layers()  # (n_out, n_in) per layer
            fwd = sum(2.0 * 2048 * o * i for (o, i) in layers)
            n, ms, fl = net.timing_get(M.MLP_K_GEMM_FWD)
            assert n == 6 * len(layers) and ms > 0 and fl == 6 * fwd
            n, ms, fl = net.timing_get(M.MLP_K_GEMM_WGRAD)
            assert n == 6 * len(layers) and fl == 6 * fwd
            n, ms, fl = net.timing_get(M.MLP_K_GEMM_DGRAD)
            assert n == 6 * (len(layers) - 1) and fl == 6 * (fwd - 2.0 * 2048 * layers[0][0] * layers[0][1])
            n, ms, fl = net.timing_get(M.MLP_K_OTHER)
            assert n == 6 * (2 * len(layers) + 4) and fl == 0
            net.timing_reset()
            assert net.timing_get(M.MLP_K_GEMM_FWD)[0] == 0
        net.close()
    assert np.array_equal(out[0], out[1])


def test_training_fits_a_smooth_field():
    X, T = data(200000, seed=7)
    net, _ = make(seed=8)
    net.set_scaling(np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))
    net.set_data(X, T)
    l0 = float(((net.forward(X) - T) ** 2).mean())
    net.train(3000, batch=4096, lr=1e-3)
    l1 = float(((net.forward(X) - T) ** 2).mean())
    print(f"fit: loss {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 <= 1e-2 * l0


def test_warm_start_decoding():
    """A network whose output layer is constant: rho = exp(y0), u = M y1, v = M y2, p = exp(y3)/gamma."""
    m = M.Mesh.gen_ramp(40, 16, 10.0, jitter=0.1)
    net = M.MLP(4, 16, 1, 4, seed=1)
    p = np.zeros(net.n_params, np.float32)
    y = np.array([0.1, 0.9, -0.05, 0.3], np.float32)
    p[-4:] = y  # output biases
    net.set_params(p)
    s = M.Solver(m, M.make_config(cfl=0.5))
    s.warm_start(net, [2.0, 10.0])
    Q = s.get_state()
    rho, u, v, pr = np.exp(0.1), 2.0 * 0.9, 2.0 * -0.05, np.exp(0.3) / 1.4
    exp = np.array([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)])
    assert np.allclose(Q, exp[None, :], rtol=1e-6)
    p[-4] = np.nan
    net.set_params(p)
    with pytest.raises(M.MstError, match="non-physical"):
        s.warm_start(net, [2.0, 10.0])
    with pytest.raises(M.MstError):
        s.warm_start(net, [2.0])  # 2 + 1 features != 4 inputs
