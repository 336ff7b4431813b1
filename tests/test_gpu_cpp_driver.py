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


def _run(exe, out, nsteps, dt, monitor):
    r = subprocess.run([exe, str(out), str(nsteps), repr(dt), monitor], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok 1600 cells" in r.stdout
    mon = [float(v) for v in r.stdout.split("monitor")[1].split()[:4]]
    return np.fromfile(out, dtype=np.float64).reshape(-1, 4), mon


def test_cpp_rhosolver_time_loop_matches_oracle(tmp_path):
    """The monitor over the host mirrors (the reference's loop) and the device-side monitor give the
    same maxima bit for bit and the same final state."""
    exe = os.path.join(ROOT, "mst-cfd_amd", "lib", "rhosolver_driver")
    nsteps, dt = 20, 1e-4
    print("C++ driver: Sod strip, 20 steps, host and device monitors", flush=True)
    got, mon_h = _run(exe, tmp_path / "host.bin", nsteps, dt, "host")
    got_d, mon_d = _run(exe, tmp_path / "device.bin", nsteps, dt, "device")
    m = M.Mesh.gen_tri(200, 4, 0.0, 1.0, 0.0, 0.02, 0.1, 12345, (M.BC_WALL,) * 4)
    Q = m.init_state(M.INIT_REFERENCE, [1.0, 0.0, 0.0, 2.5])
    om = O.OracleMesh(m.raw())
    cfg = O.make_config(q_inlet=[1.0, 0.0, 0.0, 2.5])
    for _ in range(nsteps - 1):
        Q = om.euler(cfg, Q, dt)
    Qlast = om.euler(cfg, Q, dt)
    assert rel_state_error(got, Qlast) <= 1e-12
    assert np.array_equal(got, got_d)
    assert mon_h == mon_d
    # the reference's monitor of the last step, on the oracle's states
    assert np.allclose(mon_h, O.monitor_maxrel(Qlast, Q), rtol=1e-9, atol=0)
    assert np.abs(got - m.init_state(M.INIT_REFERENCE, [1.0, 0.0, 0.0, 2.5])).max() > 1e-3  # it moved
