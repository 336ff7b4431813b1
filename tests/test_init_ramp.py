"""Ramp generator (configuration C5) and initial conditions (Time::initialization, time/Time.cpp:13-36,
plus the synthetic workloads' fields).  CPU only."""
import math

import numpy as np
import pytest

import mstcfd as M
from exact import vortex_exact
from helpers import GAMMA, cons, prim
from oracle import pyoracle as O


def centres(m):
    return O.OracleMesh(m.raw()).geometry()["ccen"]


@pytest.mark.parametrize("theta", [0.0, 5.0, 20.0])
def test_ramp_geometry_closed_forms(theta):
    x0, x1, H = -0.5, 1.5, 1.2
    m = M.Mesh.gen_ramp(80, 30, theta, x0=x0, x1=x1, height=H, jitter=0.15, seed=2024)
    raw = m.raw()
    geo = O.OracleMesh(raw).geometry()
    t = math.tan(math.radians(theta))
    # area = channel minus the ramp triangle; every cell positive
    assert geo["vol"].min() > 0
    assert abs(geo["vol"].sum() - ((x1 - x0) * H - 0.5 * x1 * x1 * t)) <= 1e-12 * (x1 - x0) * H
    xy = raw["xyz"]
    assert np.any(np.all(xy == 0.0, axis=1))  # the corner is a node
    # bottom wall: every node on y = max(x, 0) tan(theta); area vectors of the wall faces point into
    # the domain at theta + 90 degrees downstream of the corner
    zones = {int(z[3]): z for z in m.zones()}
    w = zones[M.BC_WALL]
    f = np.arange(w[1], w[2])
    nodes = raw["fnodes"][f]
    for k in range(2):
        p = xy[nodes[:, k]]
        assert np.allclose(p[:, 1], np.maximum(p[:, 0], 0) * t, atol=1e-14)
    S = geo["S"][f]
    fx = geo["fcen"][f][:, 0]
    ang = np.degrees(np.arctan2(S[:, 1], S[:, 0]))
    assert np.allclose(ang[fx > 0], 90 + theta, atol=1e-9)
    assert np.allclose(ang[fx < 0], 90, atol=1e-9)
    # every face points into its c0 cell (the orientation the generators guarantee)
    c0 = raw["c0"]
    d = geo["ccen"][c0] - geo["fcen"]
    assert np.all((d * geo["S"]).sum(1) > 0)


def test_ramp_rejects_bad_geometry():
    with pytest.raises(M.MstError):
        M.Mesh.gen_ramp(40, 10, 50.0)
    with pytest.raises(M.MstError):
        M.Mesh.gen_ramp(40, 10, 30.0, x1=3.0, height=1.0)  # ramp above the top


def test_reference_initialisation_rule():
    """Time.cpp:26-36: cells with centre x > 0.5 get rho*0.125 and rho*E*0.1, momentum untouched."""
    m = M.Mesh.gen_tri(40, 3, 0, 1, 0, 0.02, 0.1, 12345)
    cc = centres(m)
    q0 = np.array([1.0, 0.3, -0.2, 2.5])
    Q = m.init_state(M.INIT_REFERENCE, q0)
    right = cc[:, 0] > 0.5
    assert 0 < right.sum() < m.n_cells
    assert np.array_equal(Q[~right], np.tile(q0, ((~right).sum(), 1)))
    exp = np.tile(q0, (right.sum(), 1))
    exp[:, 0] *= 0.125
    exp[:, 3] *= 0.1
    assert np.array_equal(Q[right], exp)
    # with u = 0 and p = 1 this is Sod's problem: right state rho = 0.125, p = 0.1
    Q = m.init_state(M.INIT_REFERENCE, [1.0, 0.0, 0.0, 1.0 / (GAMMA - 1)])
    r, u, p = prim(Q[right])
    assert np.allclose(r, 0.125, rtol=0, atol=1e-15) and np.allclose(p, 0.1, rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_riemann_freestream(dim):
    m = M.Mesh.gen_tri(20, 6, 0, 1, 0, 0.3, 0.1, 3) if dim == 2 else M.Mesh.gen_tet(6, 3, 3, jitter=0.1, seed=4)
    cc = centres(m)
    uL, uR = [0.3, -0.1, 0.2][:dim], [0.0, 0.05, -0.4][:dim]
    Q = m.init_state(M.INIT_UNIFORM, [1.4, 1.3] + uL + [0.7])
    r, u, p = prim(Q)
    assert np.allclose(r, 1.3) and np.allclose(u, uL) and np.allclose(p, 0.7, rtol=1e-14)
    Q = m.init_state(M.INIT_RIEMANN, [1.4, 0, 0.4, 1.0] + uL + [1.0, 0.125] + uR + [0.1])
    left = cc[:, 0] <= 0.4
    assert np.allclose(Q[left], cons(np.ones(left.sum()), np.tile(uL, (left.sum(), 1)), np.ones(left.sum())))
    n = (~left).sum()
    assert np.allclose(Q[~left], cons(np.full(n, 0.125), np.tile(uR, (n, 1)), np.full(n, 0.1)))
    Q = m.init_state(M.INIT_FREESTREAM, [1.4, 2.5, 10.0])
    r, u, p = prim(Q)
    a = np.sqrt(1.4 * p / r)
    assert np.allclose(np.sqrt((u ** 2).sum(1)) / a, 2.5, rtol=1e-14)
    assert np.allclose(np.degrees(np.arctan2(u[:, 1], u[:, 0])), 10.0)
    with pytest.raises(M.MstError):
        m.init_state(M.INIT_RIEMANN, [1.4, 5, 0.4])


def test_vortex_and_pulse_fields():
    m = M.Mesh.gen_tri(30, 30, -10, 10, -10, 10, 0.1, 12345, (M.BC_PERIODIC,) * 4)
    cc = centres(m)
    Q = m.init_state(M.INIT_VORTEX, [1.4, 0.5, 5.0, 0.0, 0.0])
    assert np.allclose(Q, vortex_exact(cc), rtol=1e-13, atol=1e-14)
    r, u, p = prim(Q)
    assert np.allclose(p / r ** 1.4, 1.0, rtol=1e-13)  # isentropic: p / rho^gamma = p_inf / rho_inf^gamma
    t = M.Mesh.gen_tet(5, 5, 5, jitter=0.15, seed=1)
    ct = centres(t)
    Q = t.init_state(M.INIT_PULSE, [1.4, 0.1, 0.1, 0.5, 0.5, 0.5, 0.8])
    b = 0.1 * np.exp(-((ct - 0.5) ** 2).sum(1) / (2 * 0.1 ** 2))
    r, u, p = prim(Q)
    assert np.allclose(r, 1 + b, rtol=1e-15) and np.allclose(p, (1 + b) / 1.4, rtol=1e-14)
    assert np.allclose(u[:, 0], 0.8) and np.allclose(u[:, 1:], 0)
