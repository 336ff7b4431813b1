"""Learned warm start end to end at small size: the solver generates steady ramp solutions in one
batch, the network is trained on them by the library's kernels, and held-out cases start from its
prediction.  The freestream start sets the target: both runs must reach the same absolute residual."""
import numpy as np
import pytest

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
