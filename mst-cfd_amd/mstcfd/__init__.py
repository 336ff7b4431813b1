"""Python binding (ctypes) of libmstcfd.so, the MI355X residual path of MST-CFD.

The library is the product; this module only marshals numpy arrays across its C ABI
(include/mstcfd.h).  There is no Python or CPU fallback for any compute entry point: if the
library is missing the import-time loader raises, and every solver call runs the HIP kernels.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MSTCFD_LIB selects another build of the library (e.g. one with different launch bounds)
LIB_PATH = os.environ.get("MSTCFD_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libmstcfd.so")

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

BC_INTERIOR, BC_WALL, BC_OUTLET, BC_SYMMETRY, BC_INLET, BC_PERIODIC = 2, 3, 5, 7, 10, 12
ROE, AUSM = 0, 1
LIM_NONE, LIM_MINMOD, LIM_VANLEER = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
COMPAT_BRANCH2_LR, COMPAT_GRAD_LASTINT, COMPAT_AUSM_LOCAL_A, COMPAT_AUSM_P_UNSCALED = 1, 2, 4, 8
COMPAT_AUSM_ONESIDED = 16
COMPAT_REFERENCE = COMPAT_GRAD_LASTINT | COMPAT_AUSM_LOCAL_A | COMPAT_AUSM_P_UNSCALED | COMPAT_AUSM_ONESIDED
K_GRAD, K_FLUX, K_UPDATE, K_DT, K_CELL, K_PATCH, K_FUSED, K_LUSGS, K_VISC = 0, 1, 2, 3, 4, 5, 6, 7, 8
KERNEL_NAMES = {K_GRAD: "k_grad", K_FLUX: "k_flux", K_UPDATE: "k_update", K_DT: "k_cfl", K_CELL: "k_cell",
                K_PATCH: "k_patch", K_FUSED: "k_fused", K_LUSGS: "k_lu_sweep", K_VISC: "k_visc"}
STRATEGY_FACE, STRATEGY_CELL, STRATEGY_AUTO, STRATEGY_PATCH, STRATEGY_FUSED = 0, 1, 2, 3, 4
INIT_REFERENCE, INIT_UNIFORM, INIT_RIEMANN, INIT_VORTEX, INIT_PULSE, INIT_FREESTREAM = 0, 1, 2, 3, 4, 5
BATCH_BUCKETS = 4
ORDER_HILBERT, ORDER_RCM = 0, 1
STEADY_EXPLICIT, STEADY_LUSGS = 0, 1
MLP_K_GEMM_FWD, MLP_K_GEMM_WGRAD, MLP_K_GEMM_DGRAD, MLP_K_OTHER = 0, 1, 2, 3
MLP_KERNEL_NAMES = {MLP_K_GEMM_FWD: "k_gemm forward", MLP_K_GEMM_WGRAD: "k_gemm weight gradient",
                    MLP_K_GEMM_DGRAD: "k_gemm activation gradient",
                    MLP_K_OTHER: "gather, loss gradient, split sums, bias gradients, Adam"}

# public exported symbols of include/mstcfd.h (checked by the CPU tests)
EXPORTS = [
    "mst_last_error", "mst_mesh_read_msh", "mst_mesh_write_msh", "mst_mesh_from_arrays", "mst_mesh_counts",
    "mst_mesh_arrays", "mst_mesh_zones", "mst_mesh_destroy", "mst_mesh_gen_tri", "mst_mesh_gen_tet",
    "mst_prep_create", "mst_prep_destroy", "mst_prep_counts", "mst_prep_permutation", "mst_prep_geometry",
    "mst_prep_tables", "mst_ctx_create", "mst_ctx_destroy", "mst_ctx_set_config", "mst_ctx_prep", "mst_set_state",
    "mst_get_state", "mst_get_new_state", "mst_set_dt", "mst_face_flux", "mst_residual_update", "mst_solve",
    "mst_update_new_to_old", "mst_step", "mst_compute_dt", "mst_residual_norms", "mst_monitor_maxrel",
    "mst_get_residual", "mst_get_face_flux", "mst_get_gradient", "mst_riemann_flux", "mst_timing_enable",
    "mst_timing_get", "mst_timing_reset", "mst_sync", "mst_algorithmic_bytes", "mst_copy_probe",
    "mst_set_strategy", "mst_get_strategy", "mst_mesh_gen_ramp", "mst_mesh_init_state",
    "mst_batch_create", "mst_batch_destroy", "mst_batch_counts", "mst_batch_case_counts", "mst_batch_set_state",
    "mst_batch_get_state", "mst_batch_step", "mst_batch_residual_norms", "mst_batch_dt", "mst_batch_set_active",
    "mst_batch_set_tuning", "mst_batch_get_tuning", "mst_batch_set_graph", "mst_batch_autotune",
    "mst_mlp_create", "mst_mlp_destroy", "mst_mlp_num_params", "mst_mlp_get_params", "mst_mlp_set_params",
    "mst_mlp_set_scaling", "mst_mlp_set_data", "mst_mlp_reset_optimizer", "mst_mlp_train", "mst_mlp_forward",
    "mst_mlp_gradient", "mst_ctx_warm_start", "mst_run_to_steady", "mst_batch_run_to_steady",
    "mst_mesh_cell_centres", "mst_batch_set_strategy", "mst_batch_get_strategy", "mst_set_cell_order",
    "mst_get_cell_order", "mst_set_steady_driver", "mst_lusgs_step", "mst_lusgs_colours", "mst_prep_cell_colours",
    "mst_write_tecplot", "mst_checkpoint_write", "mst_checkpoint_read", "mst_batch_set_steady_driver",
    "mst_batch_lusgs_step", "mst_mesh_write_cache", "mst_mesh_read_cache", "mst_monitor_maxrel_noalloc",
    "mst_mlp_timing_enable", "mst_mlp_timing_get", "mst_mlp_timing_reset",
]


class MstError(RuntimeError):
    pass


class Config(C.Structure):
    """mst_config (include/mstcfd.h)."""

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
    """Inlet state of include/CONST.h:70-83 (rho=1, u=0, E = rho*(T*CV + |u|^2/2), T = 1/286.32, CV = 715.8)."""
    E = 1 * (1 / 286.32 * 715.8 + 0.5 * (0 * 0 + 0 * 0))
    return [1.0] + [0.0] * dim + [E]


def make_config(scheme=ROE, order=1, limiter=LIM_NONE, rk=1, compat=COMPAT_REFERENCE, gamma=1.4, eor=1e-10,
                entropy_delta=0.125, cfl=0.0, q_inlet=None, dim=2, precision=PREC_FP64, mu=0.0, lam=None, kappa=0.0,
                cv=715.8):
    """mu / lam / kappa / cv: viscous terms (mu = kappa = 0: inviscid, the reference's FLAGVISCID 0);
    lam defaults to -0.666667 mu (include/CONST.h:47)."""
    c = Config()
    c.scheme, c.order, c.limiter, c.rk, c.compat, c.precision = scheme, order, limiter, rk, compat, precision
    c.gamma, c.eor, c.entropy_delta, c.cfl = gamma, eor, entropy_delta, cfl
    c.mu, c.lam, c.kappa, c.cv = mu, (-0.666667 * mu if lam is None else lam), kappa, cv
    q = list(reference_inlet(dim) if q_inlet is None else q_inlet)
    q = q + [0.0] * (5 - len(q))
    for i in range(5):
        c.q_inlet[i] = float(q[i])
    return c


_LIB = None


def lib():
    """Load libmstcfd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MstError(f"libmstcfd.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, pp = C.c_void_p, C.POINTER(C.c_void_p)
    sig = {
        "mst_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
        "mst_mesh_read_msh": (C.c_int, [C.c_char_p, pp]),
        "mst_mesh_read_cache": (C.c_int, [C.c_char_p, pp]),
        "mst_mesh_write_cache": (C.c_int, [vp, C.c_char_p]),
        "mst_mesh_write_msh": (C.c_int, [vp, C.c_char_p]),
        "mst_mesh_from_arrays": (C.c_int, [C.c_int, C.c_int64, F64P, C.c_int64, C.c_int64, C.c_int, I32P, I32P, I32P,
                                           I32P, I32P, pp]),
        "mst_mesh_counts": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mst_mesh_arrays": (C.c_int, [vp, F64P, I32P, I32P, I32P, I32P, I32P]),
        "mst_mesh_zones": (C.c_int, [vp, I32P]),
        "mst_mesh_destroy": (None, [vp]),
        "mst_mesh_gen_tri": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_uint64, I32P, pp]),
        "mst_mesh_gen_tet": (C.c_int, [C.c_int, C.c_int, C.c_int, F64P, F64P, C.c_double, C.c_uint64, I32P, pp]),
        "mst_prep_create": (C.c_int, [vp, pp]),
        "mst_set_cell_order": (C.c_int, [C.c_int]),
        "mst_get_cell_order": (C.c_int, [C.POINTER(C.c_int)]),
        "mst_prep_destroy": (None, [vp]),
        "mst_prep_counts": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mst_prep_permutation": (C.c_int, [vp, I32P, I32P, I32P]),
        "mst_prep_geometry": (C.c_int, [vp, F64P, F64P, F64P, I32P, F64P, F64P]),
        "mst_prep_tables": (C.c_int, [vp, I32P, I32P, I32P, I32P]),
        "mst_ctx_create": (C.c_int, [vp, C.POINTER(Config), pp]),
        "mst_ctx_destroy": (C.c_int, [vp]),
        "mst_ctx_set_config": (C.c_int, [vp, C.POINTER(Config)]),
        "mst_set_strategy": (C.c_int, [vp, C.c_int]),
        "mst_get_strategy": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "mst_ctx_prep": (vp, [vp]),
        "mst_set_state": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_get_state": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_get_new_state": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_set_dt": (C.c_int, [vp, C.c_double]),
        "mst_face_flux": (C.c_int, [vp]),
        "mst_residual_update": (C.c_int, [vp]),
        "mst_solve": (C.c_int, [vp]),
        "mst_update_new_to_old": (C.c_int, [vp]),
        "mst_step": (C.c_int, [vp, C.c_int]),
        "mst_compute_dt": (C.c_int, [vp, C.c_double, C.POINTER(C.c_double)]),
        "mst_residual_norms": (C.c_int, [vp, F64P]),
        "mst_monitor_maxrel": (C.c_int, [vp, F64P]),
        "mst_monitor_maxrel_noalloc": (C.c_int, [vp, vp, C.c_int64, C.c_int, vp, vp]),
        "mst_get_residual": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_get_face_flux": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_get_gradient": (C.c_int, [vp, F64P, F64P, C.c_int64]),
        "mst_riemann_flux": (C.c_int, [C.POINTER(Config), C.c_int, C.c_int64, F64P, F64P, I32P, F64P]),
        "mst_timing_enable": (C.c_int, [vp, C.c_int]),
        "mst_timing_get": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "mst_timing_reset": (C.c_int, [vp]),
        "mst_sync": (C.c_int, [vp]),
        "mst_algorithmic_bytes": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double)]),
        "mst_copy_probe": (C.c_int, [C.c_int64, C.POINTER(C.c_double)]),
        "mst_mesh_gen_ramp": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_uint64, I32P, pp]),
        "mst_mesh_init_state": (C.c_int, [vp, C.c_int, F64P, C.c_int, F64P, C.c_int64]),
        "mst_batch_create": (C.c_int, [C.c_int, pp, C.POINTER(Config), pp]),
        "mst_batch_destroy": (C.c_int, [vp]),
        "mst_batch_counts": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mst_batch_case_counts": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "mst_batch_set_state": (C.c_int, [vp, C.c_int, F64P, C.c_int64]),
        "mst_batch_get_state": (C.c_int, [vp, C.c_int, F64P, C.c_int64]),
        "mst_batch_step": (C.c_int, [vp, C.c_int]),
        "mst_batch_residual_norms": (C.c_int, [vp, F64P]),
        "mst_batch_dt": (C.c_int, [vp, F64P]),
        "mst_batch_set_active": (C.c_int, [vp, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")]),
        "mst_batch_set_tuning": (C.c_int, [vp, C.c_int, I32P]),
        "mst_batch_get_tuning": (C.c_int, [vp, C.POINTER(C.c_int), I32P]),
        "mst_batch_set_graph": (C.c_int, [vp, C.c_int]),
        "mst_batch_set_strategy": (C.c_int, [vp, C.c_int]),
        "mst_batch_get_strategy": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "mst_batch_autotune": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "mst_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, pp]),
        "mst_mlp_destroy": (C.c_int, [vp]),
        "mst_mlp_num_params": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mst_mlp_get_params": (C.c_int, [vp, F32P, C.c_int64]),
        "mst_mlp_set_params": (C.c_int, [vp, F32P, C.c_int64]),
        "mst_mlp_set_scaling": (C.c_int, [vp, F32P, F32P, F32P, F32P]),
        "mst_mlp_set_data": (C.c_int, [vp, F32P, F32P, C.c_int64]),
        "mst_mlp_reset_optimizer": (C.c_int, [vp]),
        "mst_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double)]),
        "mst_mlp_forward": (C.c_int, [vp, F32P, C.c_int64, F32P]),
        "mst_mlp_gradient": (C.c_int, [vp, F32P, F32P, C.c_int64, F32P, C.POINTER(C.c_double)]),
        "mst_mlp_timing_enable": (C.c_int, [vp, C.c_int]),
        "mst_mlp_timing_get": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
        "mst_mlp_timing_reset": (C.c_int, [vp]),
        "mst_ctx_warm_start": (C.c_int, [vp, vp, F64P, C.c_int]),
        "mst_run_to_steady": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, F64P, C.POINTER(C.c_int),
                                        C.POINTER(C.c_double)]),
        "mst_batch_run_to_steady": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, F64P, I32P, F64P]),
        "mst_mesh_cell_centres": (C.c_int, [vp, F64P, C.c_int64]),
        "mst_set_steady_driver": (C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
        "mst_lusgs_step": (C.c_int, [vp, C.c_int]),
        "mst_lusgs_colours": (C.c_int, [vp, I32P, C.c_int64, C.POINTER(C.c_int)]),
        "mst_prep_cell_colours": (C.c_int, [vp, I32P, C.POINTER(C.c_int)]),
        "mst_write_tecplot": (C.c_int, [vp, F64P, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_char_p,
                                        C.c_char_p]),
        "mst_checkpoint_write": (C.c_int, [C.c_char_p, vp, F64P, C.c_int64, C.c_int64, C.c_double]),
        "mst_batch_set_steady_driver": (C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
        "mst_batch_lusgs_step": (C.c_int, [vp, C.c_int]),
        "mst_checkpoint_read": (C.c_int, [C.c_char_p, vp, F64P, C.c_int64, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def last_error():
    buf = C.create_string_buffer(1024)
    lib().mst_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise MstError(f"{what} failed ({rc}): {last_error()}")


class Mesh:
    """Host mesh in file order (the reference's MshBlock)."""

    def __init__(self, handle):
        self.h = handle
        c = (C.c_int64 * 7)()
        _check(lib().mst_mesh_counts(self.h, c), "mst_mesh_counts")
        self.dim, self.fnn, self.n_nodes, self.n_cells, self.n_faces, self.n_int, self.n_zones = [int(v) for v in c]
        self.nv = self.dim + 2

    def __del__(self):
        if getattr(self, "h", None):
            lib().mst_mesh_destroy(self.h)
            self.h = None

    @staticmethod
    def _new(fn, *args):
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)), fn.__name__)
        return Mesh(h)

    @classmethod
    def gen_tri(cls, nx, ny, x0, x1, y0, y1, jitter=0.0, seed=12345, bc=(BC_WALL,) * 4):
        return cls._new(lib().mst_mesh_gen_tri, int(nx), int(ny), float(x0), float(x1), float(y0), float(y1),
                        float(jitter), int(seed), np.asarray(bc, np.int32))

    @classmethod
    def gen_tet(cls, nx, ny, nz, lo=(0, 0, 0), hi=(1, 1, 1), jitter=0.0, seed=12345, bc=(BC_WALL,) * 6):
        return cls._new(lib().mst_mesh_gen_tet, int(nx), int(ny), int(nz), np.asarray(lo, np.float64),
                        np.asarray(hi, np.float64), float(jitter), int(seed), np.asarray(bc, np.int32))

    @classmethod
    def gen_ramp(cls, nx, ny, theta_deg, x0=-0.5, x1=1.5, height=1.2, jitter=0.0, seed=2024,
                 bc=(BC_INLET, BC_OUTLET, BC_WALL, BC_OUTLET)):
        """Compression ramp: wall corner at x = 0, bc = (left, right, bottom, top)."""
        return cls._new(lib().mst_mesh_gen_ramp, int(nx), int(ny), float(x0), float(x1), float(height),
                        float(theta_deg), float(jitter), int(seed), np.asarray(bc, np.int32))

    def cell_centres(self):
        c = np.zeros((self.n_cells, self.dim))
        _check(lib().mst_mesh_cell_centres(self.h, c.reshape(-1), self.n_cells), "mst_mesh_cell_centres")
        return c

    def init_state(self, kind, params):
        """Initial field on the cell centres, file order (mst_mesh_init_state)."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1)
        Q = np.zeros((self.n_cells, self.nv))
        _check(lib().mst_mesh_init_state(self.h, int(kind), p if p.size else np.zeros(1), int(p.size),
                                         Q.reshape(-1), self.n_cells), "mst_mesh_init_state")
        return Q

    @classmethod
    def read_msh(cls, path):
        return cls._new(lib().mst_mesh_read_msh, str(path).encode())

    @classmethod
    def read_cache(cls, path):
        """A mesh from its binary cache (mst_mesh_write_cache)."""
        return cls._new(lib().mst_mesh_read_cache, str(path).encode())

    def write_cache(self, path):
        _check(lib().mst_mesh_write_cache(self.h, str(path).encode()), "mst_mesh_write_cache")

    @classmethod
    def from_raw(cls, raw):
        xyz = np.ascontiguousarray(raw["xyz"], np.float64).reshape(-1)
        fn = np.ascontiguousarray(raw["fnodes"], np.int32)
        dim = int(raw["dim"])
        return cls._new(lib().mst_mesh_from_arrays, dim, xyz.size // dim, xyz, int(raw["n_cells"]),
                        int(fn.shape[0]), int(fn.shape[1]), fn.reshape(-1), np.ascontiguousarray(raw["c0"], np.int32),
                        np.ascontiguousarray(raw["c1"], np.int32), np.ascontiguousarray(raw["ftype"], np.int32),
                        np.ascontiguousarray(raw["partner"], np.int32))

    def write_tecplot(self, Q, path, gamma=1.4, cv=715.8, wall_noslip=False, title="mst-cfd"):
        """Tecplot ASCII file of state Q (file order): nodal X Y [Z] rho u v [w] T p Ma + the cells' nodes."""
        Q = np.ascontiguousarray(Q, np.float64).reshape(self.n_cells, self.nv)
        _check(lib().mst_write_tecplot(self.h, Q.reshape(-1), self.n_cells, float(gamma), float(cv),
                                       1 if wall_noslip else 0, title.encode(), str(path).encode()),
               "mst_write_tecplot")

    def write_checkpoint(self, path, Q, step=0, time=0.0):
        Q = np.ascontiguousarray(Q, np.float64).reshape(self.n_cells, self.nv)
        _check(lib().mst_checkpoint_write(str(path).encode(), self.h, Q.reshape(-1), self.n_cells, int(step),
                                          float(time)), "mst_checkpoint_write")

    def read_checkpoint(self, path):
        """(state, step, time) of a checkpoint written for this mesh."""
        Q = np.zeros((self.n_cells, self.nv))
        st, t = C.c_int64(), C.c_double()
        _check(lib().mst_checkpoint_read(str(path).encode(), self.h, Q.reshape(-1), self.n_cells, C.byref(st),
                                         C.byref(t)), "mst_checkpoint_read")
        return Q, st.value, t.value

    def write_msh(self, path):
        _check(lib().mst_mesh_write_msh(self.h, str(path).encode()), "mst_mesh_write_msh")

    def raw(self):
        """File-order arrays (the oracle's input)."""
        xyz = np.zeros((self.n_nodes, self.dim))
        fnodes = np.zeros((self.n_faces, self.fnn), np.int32)
        c0 = np.zeros(self.n_faces, np.int32)
        c1 = np.zeros(self.n_faces, np.int32)
        ft = np.zeros(self.n_faces, np.int32)
        pa = np.zeros(self.n_faces, np.int32)
        _check(lib().mst_mesh_arrays(self.h, xyz.reshape(-1), fnodes.reshape(-1), c0, c1, ft, pa), "mst_mesh_arrays")
        return dict(dim=self.dim, xyz=xyz, fnodes=fnodes, c0=c0, c1=c1, ftype=ft, partner=pa, n_cells=self.n_cells,
                    n_int=self.n_int)

    def zones(self):
        z = np.zeros((max(self.n_zones, 1), 4), np.int32)
        _check(lib().mst_mesh_zones(self.h, z.reshape(-1)), "mst_mesh_zones")
        return z[: self.n_zones]


class Prep:
    """Host preprocessing (geometry, cell order, face order, colouring) without a device."""

    def __init__(self, mesh=None, handle=None, owner=None):
        if handle is None:
            h = C.c_void_p()
            _check(lib().mst_prep_create(mesh.h, C.byref(h)), "mst_prep_create")
            self.h, self._own = h, True
        else:
            self.h, self._own, self._owner = handle, False, owner
        c = (C.c_int64 * 8)()
        _check(lib().mst_prep_counts(self.h, c), "mst_prep_counts")
        (self.dim, self.nv, self.n_cells, self.n_faces, self.maxf, self.ncolours, self.n_interior,
         self.n_periodic) = [int(v) for v in c]

    def __del__(self):
        if getattr(self, "_own", False) and getattr(self, "h", None):
            lib().mst_prep_destroy(self.h)
            self.h = None

    def permutation(self):
        cn = np.zeros(self.n_cells, np.int32)
        fnw = np.zeros(self.n_faces, np.int32)
        col = np.zeros(self.n_faces, np.int32)
        _check(lib().mst_prep_permutation(self.h, cn, fnw, col), "mst_prep_permutation")
        return cn, fnw, col

    def cell_colours(self):
        """(colour per file-order cell, number of colours) of the LU-SGS sweeps."""
        col = np.zeros(self.n_cells, np.int32)
        n = C.c_int()
        _check(lib().mst_prep_cell_colours(self.h, col, C.byref(n)), "mst_prep_cell_colours")
        return col, n.value

    def geometry(self):
        d, nf, nc = self.dim, self.n_faces, self.n_cells
        out = dict(fcen=np.zeros((nf, d)), S=np.zeros((nf, d)), eta=np.zeros(nf), dac=np.zeros(nf, np.int32),
                   ccen=np.zeros((nc, d)), vol=np.zeros(nc))
        _check(lib().mst_prep_geometry(self.h, out["fcen"].reshape(-1), out["S"].reshape(-1), out["eta"], out["dac"],
                                       out["ccen"].reshape(-1), out["vol"]), "mst_prep_geometry")
        return out

    def tables(self):
        fc0 = np.zeros(self.n_faces, np.int32)
        fc1 = np.zeros(self.n_faces, np.int32)
        kind = np.zeros(self.n_faces, np.int32)
        ell = np.zeros(self.maxf * self.n_cells, np.int32)
        _check(lib().mst_prep_tables(self.h, fc0, fc1, kind, ell), "mst_prep_tables")
        return fc0, fc1, kind, ell.reshape(self.maxf, self.n_cells)


class Solver:
    """Device context: the GPU RhoSolver (rhoSolver/RhoSolver.h:14-24) over the C ABI."""

    def __init__(self, mesh, cfg):
        h = C.c_void_p()
        _check(lib().mst_ctx_create(mesh.h, C.byref(cfg), C.byref(h)), "mst_ctx_create")
        self.h = h
        self.cfg = cfg
        self.mesh = mesh
        self.prep = Prep(handle=lib().mst_ctx_prep(self.h), owner=self)
        self.n_cells, self.n_faces, self.nv, self.dim = self.prep.n_cells, self.prep.n_faces, self.prep.nv, self.prep.dim

    def close(self):
        if getattr(self, "h", None):
            lib().mst_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_config(self, cfg):
        _check(lib().mst_ctx_set_config(self.h, C.byref(cfg)), "mst_ctx_set_config")
        self.cfg = cfg

    def run_to_steady(self, tol=1e-8, max_steps=60000, check_every=200, ref=None):
        """Steps until every field's L2(R/V) / ref <= tol; ref None: the first step's norms.
        Returns (steps, final ratio, ref)."""
        r = np.zeros(self.nv) if ref is None else np.ascontiguousarray(ref, np.float64).copy()
        n, d = C.c_int(), C.c_double()
        _check(lib().mst_run_to_steady(self.h, float(tol), int(max_steps), int(check_every), r, C.byref(n),
                                       C.byref(d)), "mst_run_to_steady")
        return n.value, d.value, r

    def set_steady_driver(self, driver, cfl=1e4, omega=1.0):
        """MST_STEADY_EXPLICIT (mst_step) or MST_STEADY_LUSGS (implicit, pseudo-time CFL cfl) for run_to_steady."""
        _check(lib().mst_set_steady_driver(self.h, int(driver), float(cfl), float(omega)), "mst_set_steady_driver")

    def lusgs_step(self, n=1):
        """n LU-SGS iterations (Old -> New, swapped)."""
        _check(lib().mst_lusgs_step(self.h, int(n)), "mst_lusgs_step")

    def lusgs_colours(self):
        """(colour per cell in file order, number of colours) of the LU-SGS sweeps."""
        col = np.zeros(self.n_cells, np.int32)
        n = C.c_int()
        _check(lib().mst_lusgs_colours(self.h, col, self.n_cells, C.byref(n)), "mst_lusgs_colours")
        return col, n.value

    def warm_start(self, mlp, cond):
        """Old state := the network's decoded prediction at the cell centres (mst_ctx_warm_start)."""
        c = np.ascontiguousarray(cond, np.float64).reshape(-1)
        _check(lib().mst_ctx_warm_start(self.h, mlp.h, c, c.size), "mst_ctx_warm_start")

    @property
    def strategy(self):
        v = C.c_int()
        _check(lib().mst_get_strategy(self.h, C.byref(v)), "mst_get_strategy")
        return v.value

    @strategy.setter
    def strategy(self, value):
        _check(lib().mst_set_strategy(self.h, int(value)), "mst_set_strategy")

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, np.float64).reshape(self.n_cells, self.nv)
        _check(lib().mst_set_state(self.h, Q.reshape(-1), self.n_cells), "mst_set_state")

    def get_state(self):
        Q = np.zeros((self.n_cells, self.nv))
        _check(lib().mst_get_state(self.h, Q.reshape(-1), self.n_cells), "mst_get_state")
        return Q

    def get_new_state(self):
        Q = np.zeros((self.n_cells, self.nv))
        _check(lib().mst_get_new_state(self.h, Q.reshape(-1), self.n_cells), "mst_get_new_state")
        return Q

    def set_dt(self, dt):
        _check(lib().mst_set_dt(self.h, float(dt)), "mst_set_dt")

    def face_flux(self):
        _check(lib().mst_face_flux(self.h), "mst_face_flux")

    def residual_update(self):
        _check(lib().mst_residual_update(self.h), "mst_residual_update")

    def solve(self):
        _check(lib().mst_solve(self.h), "mst_solve")

    def update_new_to_old(self):
        _check(lib().mst_update_new_to_old(self.h), "mst_update_new_to_old")

    def step(self, n=1):
        _check(lib().mst_step(self.h, int(n)), "mst_step")

    def compute_dt(self, cfl):
        dt = C.c_double()
        _check(lib().mst_compute_dt(self.h, float(cfl), C.byref(dt)), "mst_compute_dt")
        return dt.value

    def residual_norms(self):
        out = np.zeros(self.nv)
        _check(lib().mst_residual_norms(self.h, out), "mst_residual_norms")
        return out

    def monitor_maxrel(self):
        out = np.zeros(self.nv)
        _check(lib().mst_monitor_maxrel(self.h, out), "mst_monitor_maxrel")
        return out

    def residual(self):
        R = np.zeros((self.n_cells, self.nv))
        _check(lib().mst_get_residual(self.h, R.reshape(-1), self.n_cells), "mst_get_residual")
        return R

    def face_flux_values(self):
        G = np.zeros((self.n_faces, self.nv))
        _check(lib().mst_get_face_flux(self.h, G.reshape(-1), self.n_faces), "mst_get_face_flux")
        return G

    def gradient(self):
        g = np.zeros((self.n_cells, self.nv, self.dim))
        phi = np.zeros((self.n_cells, self.nv))
        _check(lib().mst_get_gradient(self.h, g.reshape(-1), phi.reshape(-1), self.n_cells), "mst_get_gradient")
        return g, phi

    def timing(self, on=True):
        """True / 1: HIP events on every launch; an int n > 1: on one launch in n of each kernel class."""
        _check(lib().mst_timing_enable(self.h, int(on)), "mst_timing_enable")

    def timing_get(self, kernel):
        n, ms = C.c_int64(), C.c_double()
        _check(lib().mst_timing_get(self.h, int(kernel), C.byref(n), C.byref(ms)), "mst_timing_get")
        return n.value, ms.value

    def timing_reset(self):
        _check(lib().mst_timing_reset(self.h), "mst_timing_reset")

    def sync(self):
        _check(lib().mst_sync(self.h), "mst_sync")

    def algorithmic_bytes(self, kernel):
        b = C.c_double()
        _check(lib().mst_algorithmic_bytes(self.h, int(kernel), C.byref(b)), "mst_algorithmic_bytes")
        return b.value


def set_cell_order(order):
    """Cell order of every later Prep / Solver / Batch creation in this process (ORDER_HILBERT or ORDER_RCM)."""
    _check(lib().mst_set_cell_order(int(order)), "mst_set_cell_order")


def get_cell_order():
    o = C.c_int()
    _check(lib().mst_get_cell_order(C.byref(o)), "mst_get_cell_order")
    return o.value


def copy_probe(nbytes):
    """One launch of a plain device copy of nbytes; returns its duration in ms."""
    ms = C.c_double()
    _check(lib().mst_copy_probe(int(nbytes), C.byref(ms)), "mst_copy_probe")
    return ms.value


def riemann_flux(cfg, QL, QR, dirs, dim=2):
    """Batched device Riemann flux: F[i] = scheme.set(QL[i], QR[i]).solverAll(dirs[i])."""
    QL = np.ascontiguousarray(QL, np.float64).reshape(-1, dim + 2)
    QR = np.ascontiguousarray(QR, np.float64).reshape(-1, dim + 2)
    d = np.ascontiguousarray(dirs, np.int32).reshape(-1)
    F = np.zeros_like(QL)
    _check(lib().mst_riemann_flux(C.byref(cfg), int(dim), QL.shape[0], QL.reshape(-1), QR.reshape(-1), d,
                                  F.reshape(-1)), "mst_riemann_flux")
    return F


class Batch:
    """Many independent cases advanced together (persistent kernels, graph-replayed RK steps)."""

    def __init__(self, meshes, cfgs):
        self.meshes = list(meshes)
        n = len(self.meshes)
        arr = (C.c_void_p * n)(*[m.h for m in self.meshes])
        cf = (Config * n)(*cfgs)
        h = C.c_void_p()
        _check(lib().mst_batch_create(n, C.cast(arr, C.POINTER(C.c_void_p)), cf, C.byref(h)), "mst_batch_create")
        self.h = h
        c = (C.c_int64 * 6)()
        _check(lib().mst_batch_counts(self.h, c), "mst_batch_counts")
        self.n_cases, self.n_cells, self.n_faces, self.n_cell_items, self.n_face_items, self.dim = [int(v) for v in c]
        self.nv = self.dim + 2

    def close(self):
        if getattr(self, "h", None):
            lib().mst_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def counts(self):
        c = (C.c_int64 * 6)()
        _check(lib().mst_batch_counts(self.h, c), "mst_batch_counts")
        return [int(v) for v in c]

    def set_state(self, i, Q):
        Q = np.ascontiguousarray(Q, np.float64)
        _check(lib().mst_batch_set_state(self.h, int(i), Q.reshape(-1), Q.shape[0]), "mst_batch_set_state")

    def get_state(self, i):
        m = self.meshes[i]
        Q = np.zeros((m.n_cells, self.nv))
        _check(lib().mst_batch_get_state(self.h, int(i), Q.reshape(-1), m.n_cells), "mst_batch_get_state")
        return Q

    def step(self, n=1):
        _check(lib().mst_batch_step(self.h, int(n)), "mst_batch_step")

    def residual_norms(self):
        out = np.zeros((self.n_cases, self.nv))
        _check(lib().mst_batch_residual_norms(self.h, out.reshape(-1)), "mst_batch_residual_norms")
        return out

    def dt(self):
        out = np.zeros(self.n_cases)
        _check(lib().mst_batch_dt(self.h, out), "mst_batch_dt")
        return out

    def set_active(self, active):
        a = np.ascontiguousarray(np.asarray(active, bool).astype(np.uint8))
        _check(lib().mst_batch_set_active(self.h, a), "mst_batch_set_active")

    def set_steady_driver(self, driver, cfl=1e4, omega=1.0):
        """STEADY_EXPLICIT or STEADY_LUSGS (FP64 batches) for run_to_steady."""
        _check(lib().mst_batch_set_steady_driver(self.h, int(driver), float(cfl), float(omega)),
               "mst_batch_set_steady_driver")

    def lusgs_step(self, n=1):
        """n LU-SGS iterations of every active case."""
        _check(lib().mst_batch_lusgs_step(self.h, int(n)), "mst_batch_lusgs_step")

    def run_to_steady(self, tol=1e-8, max_steps=60000, check_every=500, ref=None):
        """Per case: steps taken, final ratio, reference norms (converged cases are frozen)."""
        r = np.zeros((self.n_cases, self.nv)) if ref is None else np.ascontiguousarray(ref, np.float64).copy()
        steps = np.zeros(self.n_cases, np.int32)
        drop = np.zeros(self.n_cases)
        _check(lib().mst_batch_run_to_steady(self.h, float(tol), int(max_steps), int(check_every), r.reshape(-1),
                                             steps, drop), "mst_batch_run_to_steady")
        return steps, drop, r

    def set_tuning(self, block, tiles):
        _check(lib().mst_batch_set_tuning(self.h, int(block), np.asarray(tiles, np.int32)), "mst_batch_set_tuning")

    def get_tuning(self):
        b = C.c_int()
        t = np.zeros(BATCH_BUCKETS, np.int32)
        _check(lib().mst_batch_get_tuning(self.h, C.byref(b), t), "mst_batch_get_tuning")
        return b.value, t.tolist()

    def set_graph(self, on):
        _check(lib().mst_batch_set_graph(self.h, 1 if on else 0), "mst_batch_set_graph")

    @property
    def strategy(self):
        v = C.c_int()
        _check(lib().mst_batch_get_strategy(self.h, C.byref(v)), "mst_batch_get_strategy")
        return v.value

    @strategy.setter
    def strategy(self, value):
        _check(lib().mst_batch_set_strategy(self.h, int(value)), "mst_batch_set_strategy")

    def autotune(self, cache_path=None, force=False, steps=5):
        ms = C.c_double()
        path = str(cache_path).encode() if cache_path else None
        _check(lib().mst_batch_autotune(self.h, path, 1 if force else 0, int(steps), C.byref(ms)),
               "mst_batch_autotune")
        return ms.value


class MLP:
    """Fully connected tanh network trained and evaluated by the library's kernels (FP32)."""

    def __init__(self, n_in=4, width=128, hidden=4, n_out=4, seed=2024):
        h = C.c_void_p()
        _check(lib().mst_mlp_create(int(n_in), int(width), int(hidden), int(n_out), int(seed), C.byref(h)),
               "mst_mlp_create")
        self.h = h
        self.n_in, self.width, self.hidden, self.n_out = n_in, width, hidden, n_out
        n = C.c_int64()
        _check(lib().mst_mlp_num_params(self.h, C.byref(n)), "mst_mlp_num_params")
        self.n_params = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().mst_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def layers(self):
        """[(W, b)] shapes in parameter order."""
        dims = [self.n_in] + [self.width] * self.hidden + [self.n_out]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib().mst_mlp_get_params(self.h, p, p.size), "mst_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().mst_mlp_set_params(self.h, p, p.size), "mst_mlp_set_params")

    def set_scaling(self, in_shift, in_scale, out_shift, out_scale):
        f = lambda a: np.ascontiguousarray(a, np.float32)
        _check(lib().mst_mlp_set_scaling(self.h, f(in_shift), f(in_scale), f(out_shift), f(out_scale)),
               "mst_mlp_set_scaling")

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        _check(lib().mst_mlp_set_data(self.h, X.reshape(-1), T.reshape(-1), X.shape[0]), "mst_mlp_set_data")

    def reset_optimizer(self):
        _check(lib().mst_mlp_reset_optimizer(self.h), "mst_mlp_reset_optimizer")

    def train(self, steps, batch=8192, lr=1e-3, graph=True):
        loss = C.c_double()
        _check(lib().mst_mlp_train(self.h, int(steps), int(batch), float(lr), 1 if graph else 0, C.byref(loss)),
               "mst_mlp_train")
        return loss.value

    def timing(self, on=True):
        """Per-dispatch timing of the network's kernels (training launches directly while it is on)."""
        _check(lib().mst_mlp_timing_enable(self.h, 1 if on else 0), "mst_mlp_timing_enable")

    def timing_get(self, kclass):
        """(launches, total ms, total floating-point operations) of a kernel class (MLP_K_*)."""
        n, ms, fl = C.c_int64(), C.c_double(), C.c_double()
        _check(lib().mst_mlp_timing_get(self.h, int(kclass), C.byref(n), C.byref(ms), C.byref(fl)),
               "mst_mlp_timing_get")
        return n.value, ms.value, fl.value

    def timing_reset(self):
        _check(lib().mst_mlp_timing_reset(self.h), "mst_mlp_timing_reset")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], self.n_out), np.float32)
        _check(lib().mst_mlp_forward(self.h, X.reshape(-1), X.shape[0], Y.reshape(-1)), "mst_mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(lib().mst_mlp_gradient(self.h, X.reshape(-1), T.reshape(-1), X.shape[0], g, C.byref(loss)),
               "mst_mlp_gradient")
        return g, loss.value
