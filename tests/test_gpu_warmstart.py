"""Learned warm start end to end at small size: the solver generates steady ramp solutions in one
batch, the network is trained on them by the library's kernels, and held-out cases start from its
prediction.  The freestream start sets the target: both runs must reach the same absolute residual."""
import numpy as np
import pytest

import mstcfd as M
from mstcfd import ramp as RP

pytestmark = pytest.mark.gpu


def test_warm_start_pipeline_small():
    X, T, gen = RP.generate(RP.sample_cases(24, 2024), 60, 25, max_steps=10000, check_every=500)
    assert gen["converged"] >= 20 and X.shape == (24 * 3000, 4) and np.isfinite(T).all()
    net, tr = RP.train(X, T, steps=(3000, 1000, 500), batch=4096)
    rows = RP.evaluate(net, [(2.2, 8.0), (2.6, 12.0), (1.9, 10.0)], 120, 50, max_steps=30000)
    sm = RP.summary(rows)
    print(f"warm start: {sm}", flush=True)
    for r in rows:
        assert r["warm_start"] == "network" and r["converged"]
        # what the network gains: the predicted field is much closer to the steady state
        assert np.mean(r["network_error"]) * 10 < np.mean(r["freestream_error"])
        # and what that buys in steps to the same absolute residual: within a factor 2 either way (the
        # convergence is bound by propagation, test_a_start_near_the_steady_state_saves_little)
        assert 0.5 <= r["cold_steps"] / r["warm_steps"] <= 2.0, r


@pytest.mark.parametrize("driver", ["explicit", "lusgs"])
def test_a_start_near_the_steady_state_saves_little(driver):
    """Why no initial field can cut the time to a 1e-8 residual threefold on the ramp.  The steady
    state itself, perturbed by 1e-3 or 1e-5 (relative, per cell and field), is a start a hundred to
    ten thousand times closer than the network's prediction can be.  From it, both drivers still need
    most of the freestream start's steps to reach the same absolute level: the residual falls at a rate
    set by how fast perturbations leave the domain, not by their size."""
    mach, theta = 2.2, 10.0
    m = RP.mesh(theta, 120, 50)
    s = M.Solver(m, RP.config(m, mach))
    every = 10
    if driver == "lusgs":
        s.set_steady_driver(M.STEADY_LUSGS, 1e4, 1.0)
        s.lusgs_colours()
    s.set_state(RP.freestream(m, mach))
    n_cold, drop, ref = s.run_to_steady(1e-8, 60000, every)
    assert drop <= 1e-8
    # a steady state well below the target level
    _, drop, _ = s.run_to_steady(1e-11, 200000, every, ref)
    assert drop <= 1e-11
    Qs = s.get_state()
    rng = np.random.default_rng(7)
    steps = {}
    for eps in (1e-3, 1e-5):
        s.set_state(Qs * (1 + eps * rng.standard_normal(Qs.shape)))
        n, drop, _ = s.run_to_steady(1e-8, 60000, every, ref)
        assert drop <= 1e-8
        steps[eps] = n
    s.close()
    print(f"{driver}: freestream {n_cold} steps; steady state perturbed by 1e-3: {steps[1e-3]}, by 1e-5: "
          f"{steps[1e-5]}", flush=True)
    assert n_cold / steps[1e-3] < 3 and n_cold / steps[1e-5] < 3
