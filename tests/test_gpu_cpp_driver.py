"""The C++ host class (include/mstcfd.hpp, RhoSolverGPU with RhoSolver's method set) driven by the
reference's own time loop (tests/cpp/rhosolver_driver.cpp: setDT, solve, monitor, updateNewToOld) on
the Sod strip: the final state matches the oracle's forward-Euler steps to 1e-12 relative."""
import os
import subprocess

import numpy as np
import pytest

import mstcfd as M
from helpers import rel_state_error
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_rhosolver_time_loop_matches_oracle(tmp_path):
    exe = os.path.join(ROOT, "mst-cfd_amd", "lib", "rhosolver_driver")
    out = tmp_path / "state.bin"
    nsteps, dt = 20, 1e-4
    print("C++ driver: Sod strip, 20 steps", flush=True)
    r = subprocess.run([exe, str(out), str(nsteps), repr(dt)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok 1600 cells" in r.stdout
    m = M.Mesh.gen_tri(200, 4, 0.0, 1.0, 0.0, 0.02, 0.1, 12345, (M.BC_WALL,) * 4)
    Q = m.init_state(M.INIT_REFERENCE, [1.0, 0.0, 0.0, 2.5])
    om = O.OracleMesh(m.raw())
    cfg = O.make_config(q_inlet=[1.0, 0.0, 0.0, 2.5])
    for _ in range(nsteps):
        Q = om.euler(cfg, Q, dt)
    got = np.fromfile(out, dtype=np.float64).reshape(m.n_cells, 4)
    assert rel_state_error(got, Q) <= 1e-12
    assert np.abs(got - m.init_state(M.INIT_REFERENCE, [1.0, 0.0, 0.0, 2.5])).max() > 1e-3  # it moved
