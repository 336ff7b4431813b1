"""The allocation-free monitor (mst_monitor_maxrel_noalloc) against the context's monitor (GPU).

It reads caller-owned device arrays and writes its maxima to a device array; the context's
mst_monitor_maxrel reduces the same two states.  A maximum does not depend on the order of the
cells, so the two agree bit for bit.  Two calls in a row share the scratch array: the second, smaller
one (a 3D mesh, five fields) must not see the first one's partials.  The test's own device buffers
come from the HIP runtime the library itself is linked against (hipMalloc through ctypes): a second
runtime in the process (torch's bundled one) would not see the library's device."""
import ctypes as C
import os

import numpy as np
import pytest

import mstcfd as M
from helpers import smooth_state
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The libamdhip64 that libmstcfd.so loaded (its path from this process's mappings)."""
    M.lib()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if os.path.basename(path).startswith("libamdhip64.so"):
            H = C.CDLL(path)
            H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            H.hipFree.argtypes = [C.c_void_p]
            return H
    raise RuntimeError("libamdhip64 is not loaded")


class Dev:
    """A device copy of a float64 array (the test's buffers; the entry point allocates nothing)."""

    def __init__(self, H, a):
        self.H, self.shape, self.nbytes = H, a.shape, a.nbytes
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), a.nbytes) == 0
        a = np.ascontiguousarray(a, np.float64)
        assert H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def _stepped(mesh, cfg):
    s = M.Solver(mesh, cfg)
    Q = smooth_state(s.prep.geometry()["ccen"], seed=3, mach=0.6)
    s.set_state(Q)
    s.set_dt(0.5 * s.compute_dt(0.5))
    s.solve()  # New = one Euler step of Old; both stay in the context
    return s


def _noalloc(H, Qnew, Qold):
    qn, qo = Dev(H, Qnew), Dev(H, Qold)
    out = Dev(H, np.full(Qnew.shape[1], -1.0))
    rc = M.lib().mst_monitor_maxrel_noalloc(qn.p, qo.p, Qnew.shape[0], Qnew.shape[1], out.p, None)
    assert rc == 0, M.last_error()
    got = out.get()  # hipMemcpy on the null stream waits for the launches
    for d in (qn, qo, out):
        d.free()
    return got


def test_noalloc_monitor_matches_the_context_monitor():
    cases = [
        (M.Mesh.gen_tri(400, 300, -10, 10, -10, 10, 0.1, 5, (M.BC_PERIODIC,) * 4),
         M.make_config(scheme=M.ROE, order=2, limiter=M.LIM_VANLEER)),
        (M.Mesh.gen_tet(10, 8, 6, (0, 0, 0), (1, 0.8, 0.6), 0.1, 2, (M.BC_WALL,) * 6),
         M.make_config(dim=3, scheme=M.AUSM, order=2, limiter=M.LIM_MINMOD)),
    ]
    H = _hip()
    for m, cfg in cases:  # the second call is smaller (and has five fields)
        s = _stepped(m, cfg)
        Qold, Qnew = s.get_state(), s.get_new_state()
        ref = s.monitor_maxrel()
        got = _noalloc(H, Qnew, Qold)
        print(f"{m.n_cells} cells, nv {s.nv}: context {ref}, allocation-free {got}", flush=True)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, O.monitor_maxrel(Qnew, Qold))
        assert (got > 0).all()
        s.close()


def test_noalloc_monitor_rejects_bad_arguments():
    H = _hip()
    q, out = Dev(H, np.zeros((10, 4))), Dev(H, np.zeros(4))
    L = M.lib()
    assert L.mst_monitor_maxrel_noalloc(None, q.p, 10, 4, out.p, None) == -1
    assert L.mst_monitor_maxrel_noalloc(q.p, q.p, 0, 4, out.p, None) == -1
    assert L.mst_monitor_maxrel_noalloc(q.p, q.p, 10, 3, out.p, None) == -1
    assert L.mst_monitor_maxrel_noalloc(q.p, q.p, (1 << 40) + 1, 4, out.p, None) == -1
    assert "MST_NOALLOC_MAX_CELLS" in M.last_error()
    misaligned = C.c_void_p(q.p.value + 8)
    assert L.mst_monitor_maxrel_noalloc(misaligned, q.p, 9, 4, out.p, None) == -1
    assert "aligned" in M.last_error()
    q.free()
    out.free()
