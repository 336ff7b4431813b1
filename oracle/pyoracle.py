"""ctypes binding of the CPU oracle (oracle/libmst_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, where it is the checker (or the timed CPU baseline).  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

INTERIOR, WALL, OUTLET, SYMMETRY, INLET, PERIODIC = 2, 3, 5, 7, 10, 12
ROE, AUSM = 0, 1
LIM_NONE, LIM_MINMOD, LIM_VANLEER = 0, 1, 2
COMPAT_BRANCH2_LR, COMPAT_GRAD_LASTINT, COMPAT_AUSM_LOCAL_A, COMPAT_AUSM_P_UNSCALED = 1, 2, 4, 8
COMPAT_AUSM_ONESIDED = 16
COMPAT_REFERENCE = COMPAT_GRAD_LASTINT | COMPAT_AUSM_LOCAL_A | COMPAT_AUSM_P_UNSCALED | COMPAT_AUSM_ONESIDED


class Config(C.Structure):
    """Mirrors orc_config / mst_config (same field layout)."""

    _fields_ = [
        ("scheme", C.c_int32),
        ("order", C.c_int32),
        ("limiter", C.c_int32),
        ("rk", C.c_int32),
        ("compat", C.c_int32),
        ("precision", C.c_int32),
        ("gamma", C.c_double),
        ("eor", C.c_double),
        ("entropy_delta", C.c_double),
        ("cfl", C.c_double),
        ("q_inlet", C.c_double * 5),
        ("mu", C.c_double),
        ("lam", C.c_double),
        ("kappa", C.c_double),
        ("cv", C.c_double),
    ]


def reference_inlet(dim=2):
    """Inlet state of include/CONST.h:70-83 (rho=1, u=v=0, E = rho*(T*CV + 0.5|u|^2), T=1/286.32)."""
    E = 1 * (1 / 286.32 * 715.8 + 0.5 * (0 * 0 + 0 * 0))
    q = [1.0] + [0.0] * dim + [E]
    return q + [0.0] * (5 - len(q))


def make_config(scheme=ROE, order=1, limiter=LIM_NONE, rk=1, compat=COMPAT_REFERENCE, gamma=1.4, eor=1e-10,
                entropy_delta=0.125, cfl=0.0, q_inlet=None, dim=2, mu=0.0, lam=None, kappa=0.0, cv=715.8):
    c = Config()
    c.scheme, c.order, c.limiter, c.rk, c.compat = scheme, order, limiter, rk, compat
    c.gamma, c.eor, c.entropy_delta, c.cfl = gamma, eor, entropy_delta, cfl
    c.mu, c.lam, c.kappa, c.cv = mu, (-0.666667 * mu if lam is None else lam), kappa, cv
    q = reference_inlet(dim) if q_inlet is None else list(q_inlet) + [0.0] * (5 - len(q_inlet))
    for i in range(5):
        c.q_inlet[i] = q[i]
    return c


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    # MST_ORACLE_LIB selects another build of the same source (tests/test_oracle_planted.py loads
    # builds with a planted mistake this way)
    path = os.environ.get("MST_ORACLE_LIB") or os.path.join(_HERE, "libmst_oracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    L.orc_mesh_create.restype = C.c_void_p
    L.orc_mesh_create.argtypes = [C.c_int, C.c_int64, F64P, C.c_int64, C.c_int64, C.c_int, I32P, I32P, I32P, I32P,
                                  I32P, C.c_int64]
    L.orc_mesh_destroy.argtypes = [C.c_void_p]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = C.c_int
    L.orc_mesh_geometry.argtypes = [C.c_void_p, F64P, F64P, F64P, I32P, F64P, F64P]
    L.orc_mesh_lrflags.argtypes = [C.c_void_p, C.c_int, U8P]
    L.orc_pressure.argtypes = [F64P, C.c_int, C.c_double]
    L.orc_pressure.restype = C.c_double
    L.orc_enthalpy.argtypes = [F64P, C.c_int, C.c_double]
    L.orc_enthalpy.restype = C.c_double
    for fn in (L.orc_roe_flux, L.orc_ausm_flux):
        fn.argtypes = [F64P, F64P, C.c_int, C.c_int, C.POINTER(Config), F64P]
    L.orc_roe_eigvecs.argtypes = [F64P, F64P, C.c_int, C.c_int, C.POINTER(Config), F64P, F64P, F64P]
    L.orc_invert.argtypes = [C.c_int, F64P, F64P]
    L.orc_gradient.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P]
    L.orc_limiter.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P, F64P]
    L.orc_face_flux.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P]
    L.orc_residual.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P, F64P]
    L.orc_euler.argtypes = [C.c_void_p, C.POINTER(Config), F64P, C.c_double, F64P]
    L.orc_cfl_dt.argtypes = [C.c_void_p, C.POINTER(Config), F64P]
    L.orc_cfl_dt.restype = C.c_double
    L.orc_rk_step.argtypes = [C.c_void_p, C.POINTER(Config), F64P, C.c_double, F64P]
    L.orc_monitor_maxrel.argtypes = [C.c_int64, C.c_int, F64P, F64P, F64P]
    L.orc_viscous_flux.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P, F64P]
    L.orc_lusgs_sweep.argtypes = [C.c_void_p, C.POINTER(Config), F64P, F64P, I32P, C.c_double, C.c_double, F64P]
    L.orc_lusgs_step.argtypes = [C.c_void_p, C.POINTER(Config), F64P, I32P, C.c_double, C.c_double, F64P]
    _LIB = L
    return L


def set_threads(n):
    return lib().orc_set_threads(int(n))


class OracleMesh:
    """Oracle view of a raw (file-order) mesh: dict with dim, xyz, fnodes, c0, c1, ftype, partner, n_cells, n_int."""

    def __init__(self, raw):
        L = lib()
        self.raw = raw
        self.dim = int(raw["dim"])
        self.nv = self.dim + 2
        self.n_cells = int(raw["n_cells"])
        self.n_faces = int(raw["c0"].shape[0])
        xyz = np.ascontiguousarray(raw["xyz"], dtype=np.float64).reshape(-1)
        fn = np.ascontiguousarray(raw["fnodes"], dtype=np.int32)
        fnn = fn.shape[1]
        self._keep = [xyz, fn]
        self.h = L.orc_mesh_create(self.dim, xyz.size // self.dim, xyz, self.n_cells, self.n_faces, fnn,
                                   fn.reshape(-1), np.ascontiguousarray(raw["c0"], dtype=np.int32),
                                   np.ascontiguousarray(raw["c1"], dtype=np.int32),
                                   np.ascontiguousarray(raw["ftype"], dtype=np.int32),
                                   np.ascontiguousarray(raw["partner"], dtype=np.int32), int(raw["n_int"]))
        if not self.h:
            raise RuntimeError("orc_mesh_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def geometry(self):
        d, nf, nc = self.dim, self.n_faces, self.n_cells
        out = dict(fcen=np.zeros((nf, d)), S=np.zeros((nf, d)), eta=np.zeros(nf), dac=np.zeros(nf, np.int32),
                   ccen=np.zeros((nc, d)), vol=np.zeros(nc))
        lib().orc_mesh_geometry(self.h, out["fcen"].reshape(-1), out["S"].reshape(-1), out["eta"], out["dac"],
                                out["ccen"].reshape(-1), out["vol"])
        return out

    def lrflags(self, compat):
        fl = np.zeros((self.n_faces, self.dim), np.uint8)
        lib().orc_mesh_lrflags(self.h, int(compat), fl.reshape(-1))
        return fl

    def gradient(self, cfg, Q):
        g = np.zeros((self.n_cells, self.nv, self.dim))
        lib().orc_gradient(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1), g.reshape(-1))
        return g

    def limiter(self, cfg, Q, grad):
        phi = np.zeros((self.n_cells, self.nv))
        lib().orc_limiter(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1),
                          np.ascontiguousarray(grad, np.float64).reshape(-1), phi.reshape(-1))
        return phi

    def face_flux(self, cfg, Q):
        F = np.zeros((self.n_faces, self.dim, self.nv))
        lib().orc_face_flux(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1), F.reshape(-1))
        return F

    def residual(self, cfg, Q):
        R = np.zeros((self.n_cells, self.nv))
        S = np.zeros((self.n_cells, self.nv))
        lib().orc_residual(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1), R.reshape(-1),
                           S.reshape(-1))
        return R, S

    def euler(self, cfg, Q, dt):
        out = np.zeros((self.n_cells, self.nv))
        lib().orc_euler(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1), float(dt),
                        out.reshape(-1))
        return out

    def cfl_dt(self, cfg, Q):
        return lib().orc_cfl_dt(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1))

    def rk_step(self, cfg, Q, dt):
        Qc = np.array(Q, dtype=np.float64, copy=True).reshape(self.n_cells, self.nv)
        res = np.zeros(self.nv)
        lib().orc_rk_step(self.h, C.byref(cfg), Qc.reshape(-1), float(dt), res)
        return Qc, res

    def viscous_flux(self, cfg, Q):
        """(Fv [n_faces][dim][nv], Green-Gauss gradients of (u, T) [n_cells][dim+1][dim])"""
        Fv = np.zeros((self.n_faces, self.dim, self.nv))
        gw = np.zeros((self.n_cells, self.dim + 1, self.dim))
        if lib().orc_viscous_flux(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1),
                                  Fv.reshape(-1), gw.reshape(-1)):
            raise MemoryError("orc_viscous_flux")
        return Fv, gw

    def lusgs_sweep(self, cfg, Q, R, colour, cfl, omega):
        dQ = np.zeros((self.n_cells, self.nv))
        rc = lib().orc_lusgs_sweep(self.h, C.byref(cfg), np.ascontiguousarray(Q, np.float64).reshape(-1),
                                   np.ascontiguousarray(R, np.float64).reshape(-1),
                                   np.ascontiguousarray(colour, np.int32), float(cfl), float(omega), dQ.reshape(-1))
        if rc:
            raise ValueError("orc_lusgs_sweep: bad colouring, cfl or omega")
        return dQ

    def lusgs_step(self, cfg, Q, colour, cfl, omega):
        Qc = np.array(Q, dtype=np.float64, copy=True).reshape(self.n_cells, self.nv)
        res = np.zeros(self.nv)
        rc = lib().orc_lusgs_step(self.h, C.byref(cfg), Qc.reshape(-1), np.ascontiguousarray(colour, np.int32),
                                  float(cfl), float(omega), res)
        if rc:
            raise ValueError("orc_lusgs_step: bad colouring, cfl or omega")
        return Qc, res


def roe_flux(QL, QR, d, cfg, dim=2):
    F = np.zeros(dim + 2)
    lib().orc_roe_flux(np.ascontiguousarray(QL, np.float64), np.ascontiguousarray(QR, np.float64), d, dim,
                       C.byref(cfg), F)
    return F


def ausm_flux(QL, QR, d, cfg, dim=2):
    F = np.zeros(dim + 2)
    lib().orc_ausm_flux(np.ascontiguousarray(QL, np.float64), np.ascontiguousarray(QR, np.float64), d, dim,
                        C.byref(cfg), F)
    return F


def roe_eigvecs(QL, QR, d, cfg, dim=2):
    nv = dim + 2
    Rm, Ri, lam = np.zeros(nv * nv), np.zeros(nv * nv), np.zeros(nv)
    lib().orc_roe_eigvecs(np.ascontiguousarray(QL, np.float64), np.ascontiguousarray(QR, np.float64), d, dim,
                          C.byref(cfg), Rm, Ri, lam)
    return Rm.reshape(nv, nv), Ri.reshape(nv, nv), lam


def pressure(Q, dim=2, gamma=1.4):
    return lib().orc_pressure(np.ascontiguousarray(Q, np.float64), dim, gamma)


def enthalpy(Q, dim=2, gamma=1.4):
    return lib().orc_enthalpy(np.ascontiguousarray(Q, np.float64), dim, gamma)


def monitor_maxrel(Qnew, Qold):
    Qn = np.ascontiguousarray(Qnew, np.float64)
    Qo = np.ascontiguousarray(Qold, np.float64)
    nv = Qn.shape[1]
    out = np.zeros(nv)
    lib().orc_monitor_maxrel(Qn.shape[0], nv, Qn.reshape(-1), Qo.reshape(-1), out)
    return out
