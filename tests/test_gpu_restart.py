"""Restart from a checkpoint (mst_checkpoint_write / mst_checkpoint_read) continues a run bit for bit."""
import numpy as np
import pytest

import mstcfd as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("driver", ["explicit", "lusgs"])
def test_checkpoint_restart_is_bitwise(driver, tmp_path):
    m = M.Mesh.gen_tri(40, 40, -10, 10, -10, 10, 0.1, 3, (M.BC_PERIODIC,) * 4)
    cfg = M.make_config(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER, rk=3, cfl=0.5)
    Q0 = m.init_state(M.INIT_VORTEX, [1.4, 0.5, 5.0, 0.0, 0.0])
    print(f"restart {driver}: {m.n_cells} cells", flush=True)

    def run(Q, n):
        s = M.Solver(m, cfg)
        s.set_state(Q)
        if driver == "lusgs":
            s.set_steady_driver(M.STEADY_LUSGS, 5.0, 1.0)
            s.lusgs_step(n)
        else:
            s.step(n)
        out = s.get_state()
        s.close()
        return out

    straight = run(Q0, 10)
    path = tmp_path / "mid.ckpt"
    m.write_checkpoint(path, run(Q0, 4), step=4, time=0.0)
    Q, step, _ = m.read_checkpoint(path)
    assert step == 4
    resumed = run(Q, 10 - step)
    assert np.array_equal(resumed, straight)
