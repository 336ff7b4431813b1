// solver.hip -- device context behind include/mstcfd.h: SoA device arrays, launch sequences of the
// residual path (kernels.hpp), state transfer in the reference's AllData layout, timing.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "lusgs.hpp"
#include "viscous.hpp"
#include "mlp.hpp"
#include "mstcfd.h"
#include "prep.hpp"

namespace mst {
const std::string &last_error();
}

using namespace mstk;

namespace {

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess)                                                                            \
            return mst::fail(MST_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

int check_config(const mst_config *c) {
    if (!c) return mst::fail(MST_E_ARG, "null config");
    if (c->scheme != MST_SCHEME_ROE && c->scheme != MST_SCHEME_AUSM) return mst::fail(MST_E_ARG, "bad scheme");
    if (c->order != 1 && c->order != 2) return mst::fail(MST_E_ARG, "order must be 1 or 2");
    if (c->limiter < MST_LIM_NONE || c->limiter > MST_LIM_VANLEER) return mst::fail(MST_E_ARG, "bad limiter");
    if (c->rk < 1 || c->rk > 3) return mst::fail(MST_E_ARG, "rk must be 1, 2 or 3");
    if (c->precision != MST_PREC_FP64 && c->precision != MST_PREC_MIXED) return mst::fail(MST_E_ARG, "bad precision");
    if (!(c->gamma > 1) || !(c->entropy_delta > 0) || c->eor < 0 || c->cfl < 0)
        return mst::fail(MST_E_ARG, "bad gamma / entropy_delta / eor / cfl");
    if (!(c->q_inlet[0] > 0)) return mst::fail(MST_E_ARG, "inlet density must be positive");
    if (!(c->mu >= 0) || !(c->kappa >= 0) || !std::isfinite(c->mu) || !std::isfinite(c->kappa) ||
        !std::isfinite(c->lam))
        return mst::fail(MST_E_ARG, "viscous coefficients must be finite, mu and kappa >= 0");
    if ((c->mu > 0 || c->kappa > 0) && !(c->cv > 0)) return mst::fail(MST_E_ARG, "viscous runs need cv > 0");
    return MST_OK;
}

}  // namespace

// ====================================================================================== context
struct mst_ctx {
    mst_prep prep;
    mst_config cfg;
    Params<double> p64;
    Params<float> p32;
    bool mixed = false;
    int strategy = MST_STRATEGY_AUTO;
    int resolved = -1;  // strategy in use (AUTO: measured at the first stage after a configuration)
    hipStream_t stream = nullptr;
    int device = 0;
    int dim = 2, nv = 4, maxf = 3;
    int64_t nc = 0, nf = 0, nblk_c = 0, nblk_f = 0;
    // FP64 arrays (new order)
    int32_t *fc0 = nullptr, *fc1 = nullptr, *finfo = nullptr, *ell = nullptr;
    double *fn = nullptr, *fxf = nullptr, *pxf = nullptr, *feta = nullptr;
    double *cvol = nullptr, *cxc = nullptr, *cshalf = nullptr;
    double *Q[4] = {nullptr, nullptr, nullptr, nullptr};
    int iold = 0, inew = 1;
    double *G = nullptr, *rec = nullptr, *graw = nullptr, *phi = nullptr, *R = nullptr;
    int32_t *frec0 = nullptr, *frec1 = nullptr;  // incidence records of each face's c0 / c1 side
    // patch strategy: per block of BLOCK cells its flux-carrying faces, per cell slot the local index
    int32_t *pstart = nullptr, *pface = nullptr;
    int16_t *ploc = nullptr;
    uint32_t *pfi = nullptr;  // pface with the faces' kinds and L/R masks packed in (null: 2^26 faces or more)
    int max_pf = 0;        // most faces of one patch (sizes the kernel's LDS)
    int64_t n_psolve = 0;  // Riemann problems per patch launch
    bool patch_ok = false;
    bool fused_ok = false;  // the fused kernel's LDS (states and staged normals) fits 64 KiB
    // fused strategy: each patch's halo cells (across its border faces) and their slots' local faces
    int32_t *hstart = nullptr, *hcell = nullptr;
    int16_t *hloc = nullptr;
    int64_t n_halo = 0;
    HaloGeo hg;  // the halo cells' incidence geometry in halo order, filled on the fused kernel's first use
    bool hg_dirty = true;  // hg does not hold the current incidence weights
    // mixed form: FP32 shadows of the states, FP32 geometry, offsets, face fluxes, limited gradients
    float *Q32[4] = {nullptr, nullptr, nullptr, nullptr};
    bool shadow_ok[4] = {false, false, false, false};
    float *rec32 = nullptr;  // FP32 reconstructed face states (order 2)
    // incidence layout of the gradient kernel (slot-major over the padded ELL slots)
    int slots = 3;
    int64_t n_inc = 0;  // non-empty incidences
    int32_t *inc_code = nullptr;
    double *inc_w = nullptr, *inc_n = nullptr, *inc_dx = nullptr;
    float *inc_w32 = nullptr, *inc_n32 = nullptr, *inc_dx32 = nullptr;
    float *ivol32 = nullptr;  // (float) 1/V, the FP32 gradient's volume input
    double *scal = nullptr;  // [0] dt, [1..5] l2 norms, [6..10] maxrel
    double *part = nullptr;  // per-block partials
    double *fin_mid = nullptr;  // first-level results of the two-level finish
    bool dt_valid = false;   // scal[0] holds the CFL time step of the current Old state
    int *bad = nullptr;
    bool viscous = false;     // mu or kappa > 0: the face strategy with the viscous kernels
    double *gw = nullptr;     // viscous: per cell Green-Gauss gradients of (u, T), [nc][dim + 1][dim]
    // LU-SGS steady driver (built on first use): cell colours, the cells of each colour in ascending
    // order (colour k: lu_cells[lu_start[k] .. lu_start[k + 1])), face spectral radii, the update
    int32_t *lu_cells = nullptr, *lu_nb = nullptr, *lu_face = nullptr;
    std::vector<int64_t> lu_start;
    std::vector<int32_t> lu_colour;  // per cell (host copy)
    double *lu_n = nullptr, *lu_r = nullptr, *lu_D = nullptr, *lu_dQ = nullptr;
    int steady = MST_STEADY_EXPLICIT;
    double lu_cfl = 1e4, lu_omega = 1.0;
    bool lu_inline = false;  // the configuration lets the forward sweeps compute the residual (all but a
                             // call's last iteration do)
    bool timing = false;
    int timing_stride = 1;           // one launch in timing_stride of each kernel class carries events
    int64_t tick[MST_K_COUNT] = {};  // launches per kernel class since timing was enabled
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[MST_K_COUNT];
    std::vector<hipEvent_t> pool;
};

namespace {

hipEvent_t take_event(mst_ctx *x) {
    hipEvent_t e;
    if (!x->pool.empty()) {
        e = x->pool.back();
        x->pool.pop_back();
        return e;
    }
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// A timed launch hands its two events to the dispatch itself (hipExtLaunchKernelGGL): they take the
// kernel's own start and end timestamps, with no marker packets between the kernels of a step, and
// the interval is the kernel alone.  Such a dispatch still costs about 3 us of throughput (12 us per
// step of four launches on the 2M-triangle vortex), so timing may sample one launch in n of each class.
template <typename K, typename... Args>
int launch_shm(mst_ctx *x, int kclass, K kernel, int64_t nthreads, size_t shm, Args... args) {
    if (nthreads <= 0) return MST_OK;
    const int64_t grid = (nthreads + BLOCK - 1) / BLOCK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (x->timing && kclass >= 0 && x->tick[kclass]++ % x->timing_stride == 0) {
        e0 = take_event(x);
        e1 = take_event(x);
        if (!e0 || !e1) return mst::fail(MST_E_HIP, "hipEventCreate failed");
        hipExtLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), (uint32_t)shm, x->stream, e0, e1, 0,
                              args...);
        HIPCHK(hipGetLastError());
        x->ev[kclass].push_back({e0, e1});
        return MST_OK;
    }
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), (uint32_t)shm, x->stream, args...);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

template <typename K, typename... Args>
int launch(mst_ctx *x, int kclass, K kernel, int64_t nthreads, Args... args) {
    return launch_shm(x, kclass, kernel, nthreads, 0, args...);
}

// FP32 shadow of state buffer i (mixed form), made on demand after a host write
int ensure_shadow(mst_ctx *x, int i) {
    if (!x->mixed || x->shadow_ok[i]) return MST_OK;
    int rc = launch(x, -1, k_shadow, x->nc * x->nv, x->nc * x->nv, (const double *)x->Q[i], x->Q32[i]);
    if (rc) return rc;
    x->shadow_ok[i] = true;
    return MST_OK;
}

template <typename T>
GradArgs<T> grad_args(const mst_ctx *x, int isrc, bool debug) {
    GradArgs<T> a;
    a.nc = x->nc;
    a.inc_code = x->inc_code;
    if constexpr (std::is_same<T, float>::value) {
        a.inc_w = x->inc_w32;
        a.inc_n = x->inc_n32;
        a.inc_dx = x->inc_dx32;
        a.Q = x->Q32[isrc];
        a.rec = x->rec32;
        a.vol = x->ivol32;
    } else {
        a.inc_w = x->inc_w;
        a.inc_n = x->inc_n;
        a.inc_dx = x->inc_dx;
        a.Q = x->Q[isrc];
        a.rec = x->rec;
        a.vol = x->cvol;
    }
    a.graw = debug ? x->graw : nullptr;
    a.phi = debug ? x->phi : nullptr;
    return a;
}

template <int DIM, typename T>
int run_grad_t(mst_ctx *x, int isrc, bool debug) {
    const GradArgs<T> a = grad_args<T>(x, isrc, debug);
    const int lim = x->cfg.limiter;
#define MST_GRAD(MF)                                                                                    \
    if (lim == MST_LIM_NONE) return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_NONE, T>, x->nc, a);     \
    if (lim == MST_LIM_MINMOD) return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_MINMOD, T>, x->nc, a); \
    return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_VANLEER, T>, x->nc, a);
    if (x->slots == 3) { MST_GRAD(3) }
    if (x->slots == 4) { MST_GRAD(4) }
    MST_GRAD(8)
#undef MST_GRAD
}

int run_grad(mst_ctx *x, int isrc, bool debug) {
    int rc = ensure_shadow(x, isrc);
    if (rc) return rc;
    if (x->mixed) return x->dim == 2 ? run_grad_t<2, float>(x, isrc, debug) : run_grad_t<3, float>(x, isrc, debug);
    return x->dim == 2 ? run_grad_t<2, double>(x, isrc, debug) : run_grad_t<3, double>(x, isrc, debug);
}

template <typename T>
FluxArgs<T> flux_args(const mst_ctx *x, int isrc) {
    FluxArgs<T> a;
    a.nf = x->nf;
    a.fc0 = x->fc0;
    a.fc1 = x->fc1;
    a.finfo = x->finfo;
    a.fn = x->fn;
    a.frec0 = x->frec0;
    a.frec1 = x->frec1;
    a.G = x->G;
    if constexpr (std::is_same<T, float>::value) {
        a.Q = x->Q32[isrc];
        a.rec = x->rec32;
        a.P = x->p32;
    } else {
        a.Q = x->Q[isrc];
        a.rec = x->rec;
        a.P = x->p64;
    }
    return a;
}

template <int DIM, typename T>
int run_flux_t(mst_ctx *x, int isrc) {
    const FluxArgs<T> a = flux_args<T>(x, isrc);
    const bool second = x->cfg.order >= 2;
    if (x->cfg.scheme == MST_SCHEME_ROE)
        return second ? launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_ROE, true, T>, x->nf, a)
                      : launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_ROE, false, T>, x->nf, a);
    return second ? launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_AUSM, true, T>, x->nf, a)
                  : launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_AUSM, false, T>, x->nf, a);
}

// viscous terms of state isrc subtracted from the face fluxes (after the flux kernel)
int run_viscous(mst_ctx *x, int isrc) {
    ViscArgs a;
    a.nc = x->nc;
    a.nf = x->nf;
    a.inc_code = x->inc_code;
    a.ell = x->ell;
    a.finfo = x->finfo;
    a.inc_w = x->inc_w;
    a.inc_n = x->inc_n;
    a.vol = x->cvol;
    a.fc0 = x->fc0;
    a.fc1 = x->fc1;
    a.feta = x->feta;
    a.fn = x->fn;
    a.Q = x->Q[isrc];
    a.gw = x->gw;
    a.G = x->G;
    a.mu = x->cfg.mu;
    a.lam = x->cfg.lam;
    a.kappa = x->cfg.kappa;
    a.cv = x->cfg.cv;
    int rc;
    if (x->dim == 2)
        rc = x->slots == 3 ? launch(x, MST_K_VISC, k_visc_grad<2, 3>, x->nc, a)
                           : (x->slots == 4 ? launch(x, MST_K_VISC, k_visc_grad<2, 4>, x->nc, a)
                                            : launch(x, MST_K_VISC, k_visc_grad<2, 8>, x->nc, a));
    else
        rc = x->slots == 4 ? launch(x, MST_K_VISC, k_visc_grad<3, 4>, x->nc, a)
                           : launch(x, MST_K_VISC, k_visc_grad<3, 8>, x->nc, a);
    if (rc) return rc;
    return x->dim == 2 ? launch(x, MST_K_VISC, k_visc_flux<2>, x->nf, a) : launch(x, MST_K_VISC, k_visc_flux<3>, x->nf, a);
}

// residual evaluation up to the face fluxes: gradient + limiter (order 2), the flux kernel, and the
// viscous terms when the configuration has them
int run_face_flux(mst_ctx *x, int isrc) {
    int rc = ensure_shadow(x, isrc);
    if (rc) return rc;
    if (x->cfg.order >= 2) {
        rc = run_grad(x, isrc, false);
        if (rc) return rc;
    }
    if (x->mixed)
        rc = x->dim == 2 ? run_flux_t<2, float>(x, isrc) : run_flux_t<3, float>(x, isrc);
    else
        rc = x->dim == 2 ? run_flux_t<2, double>(x, isrc) : run_flux_t<3, double>(x, isrc);
    if (rc || !x->viscous) return rc;
    return run_viscous(x, isrc);
}

UpdArgs upd_args(const mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    UpdArgs a;
    a.nc = x->nc;
    a.ell = x->ell;
    a.G = x->G;
    a.Qs = is >= 0 ? x->Q[is] : nullptr;
    a.Qn = in >= 0 ? x->Q[in] : nullptr;
    a.vol = x->cvol;
    a.sh = x->cshalf;
    a.dt = x->scal;
    set_stage(a, sg);
    a.maxf = x->maxf;
    a.l2 = (mode & UPD_L2) != 0;
    a.dtred = (mode & UPD_DT) != 0;
    a.check = (mode & UPD_CHECK) != 0;
    a.gamma = x->cfg.gamma;
    a.cfl = x->cfg.cfl;
    a.Qout = iout >= 0 ? x->Q[iout] : nullptr;
    a.Qout32 = (iout >= 0 && x->mixed) ? x->Q32[iout] : nullptr;
    a.R = x->R;
    a.part_l2 = x->part;
    a.part_dt = x->part + x->nblk_c * x->nv;
    a.bad = x->bad;
    return a;
}

template <int NV, bool GT>
int run_update_t(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    const UpdArgs a = upd_args(x, mode, is, in, sg, iout);
#define MST_UPD(MF)                                                                                             \
    switch (mode & ~UPD_CHECK) {                                                                                             \
        case UPD_RESIDUAL: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_RESIDUAL, GT>, x->nc, a);        \
        case UPD_L2: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_L2, GT>, x->nc, a);                    \
        case UPD_STAGE: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE, GT>, x->nc, a);              \
        case UPD_STAGE | UPD_L2: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_L2, GT>, x->nc, a); \
        case UPD_STAGE | UPD_DT: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_DT, GT>, x->nc, a); \
        default:                                                                                                \
            return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_L2 | UPD_DT, GT>, x->nc, a);        \
    }
    if (x->maxf <= 3) { MST_UPD(3) }
    if (x->maxf <= 4) { MST_UPD(4) }
    MST_UPD(8)
#undef MST_UPD
}

int run_update(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    int rc;
    if (x->mixed)
        rc = x->nv == 4 ? run_update_t<4, true>(x, mode, is, in, sg, iout)
                        : run_update_t<5, true>(x, mode, is, in, sg, iout);
    else
        rc = x->nv == 4 ? run_update_t<4, false>(x, mode, is, in, sg, iout)
                        : run_update_t<5, false>(x, mode, is, in, sg, iout);
    if (rc == MST_OK && iout >= 0) x->shadow_ok[iout] = x->mixed;
    return rc;
}

// cell strategy: gradient + limiter (order 2), then the fused flux + gather + stage update
template <int DIM, typename T>
int run_cell_t(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    const FluxArgs<T> fa = flux_args<T>(x, is);
    const UpdArgs a = upd_args(x, mode, is, in, sg, iout);
    const bool second = x->cfg.order >= 2;
    const bool roe = x->cfg.scheme == MST_SCHEME_ROE;
    // slot count of the padded slot-major table: 3 (triangles), 4 (quadrilaterals, tetrahedra), 8
#define MST_CELL(MF)                                                                                            \
    if (roe)                                                                                                    \
        return second ? launch(x, MST_K_CELL, k_cell<DIM, MST_SCHEME_ROE, true, T, MF>, x->nc, fa, a)           \
                      : launch(x, MST_K_CELL, k_cell<DIM, MST_SCHEME_ROE, false, T, MF>, x->nc, fa, a);         \
    return second ? launch(x, MST_K_CELL, k_cell<DIM, MST_SCHEME_AUSM, true, T, MF>, x->nc, fa, a)              \
                  : launch(x, MST_K_CELL, k_cell<DIM, MST_SCHEME_AUSM, false, T, MF>, x->nc, fa, a);
    if (x->slots == 3) { MST_CELL(3) }
    if (x->slots == 4) { MST_CELL(4) }
    MST_CELL(8)
#undef MST_CELL
}

// patch strategy: gradient + limiter (order 2), then the patch kernel (faces into LDS, sum, update)
template <int DIM, typename T>
int run_patch_t(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    if (!x->patch_ok) return mst::fail(MST_E_UNSUPPORTED, "no patch tables for this mesh");
    const FluxArgs<T> fa = flux_args<T>(x, is);
    const UpdArgs a = upd_args(x, mode, is, in, sg, iout);
    PatchArgs pa;
    pa.start = x->pstart;
    pa.face = x->pface;
    pa.loc = x->ploc;
    pa.fi = x->pfi;
    const size_t shm = (size_t)x->max_pf * x->nv * sizeof(double);
    const bool second = x->cfg.order >= 2;
    const bool roe = x->cfg.scheme == MST_SCHEME_ROE;
    if constexpr (std::is_same<T, float>::value) {
        if (second) {  // the mixed form stages its face states in LDS from the records (k_patch_rec)
            HaloArgs ha{};
            ha.start = x->hstart;
            ha.cell = x->hcell;
            ha.loc = x->hloc;
#define MST_PATCH_REC(MF)                                                                                      \
    return roe ? launch_shm(x, MST_K_PATCH, k_patch_rec<DIM, MST_SCHEME_ROE, MF>, x->nc, shm, fa, a, pa, ha) \
               : launch_shm(x, MST_K_PATCH, k_patch_rec<DIM, MST_SCHEME_AUSM, MF>, x->nc, shm, fa, a, pa, ha);
            if (x->slots == 3) { MST_PATCH_REC(3) }
            MST_PATCH_REC(4)
#undef MST_PATCH_REC
        }
    }
#define MST_PATCH(MF)                                                                                          \
    if (roe)                                                                                                   \
        return second ? launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_ROE, true, T, MF>, x->nc, shm, fa, a, pa) \
                      : launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_ROE, false, T, MF>, x->nc, shm, fa, a, pa); \
    return second ? launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_AUSM, true, T, MF>, x->nc, shm, fa, a, pa)    \
                  : launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_AUSM, false, T, MF>, x->nc, shm, fa, a, pa);
    if (x->slots == 3) { MST_PATCH(3) }
    MST_PATCH(4)
#undef MST_PATCH
}

// the patch kernel's residual-only form (FP64): R of state is, and its L2 partials when mode has UPD_L2
template <int DIM>
int run_patch_residual_t(mst_ctx *x, int mode, int is) {
    const FluxArgs<double> fa = flux_args<double>(x, is);
    const UpdArgs a = upd_args(x, mode, is, -1, Stage{0, 0}, -1);
    PatchArgs pa;
    pa.start = x->pstart;
    pa.face = x->pface;
    pa.loc = x->ploc;
    pa.fi = x->pfi;
    const size_t shm = (size_t)x->max_pf * x->nv * sizeof(double);
    const bool second = x->cfg.order >= 2;
    const bool roe = x->cfg.scheme == MST_SCHEME_ROE;
#define MST_PATCH_R(MF)                                                                                          \
    if (roe)                                                                                                     \
        return second ? launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_ROE, true, double, MF, false>, x->nc, shm, fa, a, pa) \
                      : launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_ROE, false, double, MF, false>, x->nc, shm, fa, a, pa); \
    return second ? launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_AUSM, true, double, MF, false>, x->nc, shm, fa, a, pa)    \
                  : launch_shm(x, MST_K_PATCH, k_patch<DIM, MST_SCHEME_AUSM, false, double, MF, false>, x->nc, shm, fa, a, pa);
    if (x->slots == 3) { MST_PATCH_R(3) }
    MST_PATCH_R(4)
#undef MST_PATCH_R
}

int ensure_halo_geo(mst_ctx *x);

// fused strategy (order 2): gradients, face states, fluxes, sums and update of a patch in one kernel
template <int DIM, typename T>
int run_fused_t(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    if (x->cfg.order < 2) return run_patch_t<DIM, T>(x, mode, is, in, sg, iout);
    if (!x->fused_ok) return mst::fail(MST_E_UNSUPPORTED, "fused strategy: a patch's face states exceed 64 KiB of LDS");
    int rc = ensure_halo_geo(x);
    if (rc) return rc;
    const GradArgs<T> ga = grad_args<T>(x, is, false);
    const FluxArgs<T> fa = flux_args<T>(x, is);
    const UpdArgs a = upd_args(x, mode, is, in, sg, iout);
    PatchArgs pa;
    pa.start = x->pstart;
    pa.face = x->pface;
    pa.loc = x->ploc;
    pa.fi = x->pfi;
    HaloArgs ha;
    ha.start = x->hstart;
    ha.cell = x->hcell;
    ha.loc = x->hloc;
    x->hg.args<T>(ha);
    // the face states and fluxes (and in 2D the faces' normals), and afterwards the block reductions
    // (NV doubles per wave)
    const size_t shm = std::max(fused_lds_bytes(DIM, x->nv, x->max_pf, std::is_same<T, float>::value,
                                                fused_lds_normals<DIM>()),
                                sizeof(double) * x->nv * (BLOCK / 64));
    const bool roe = x->cfg.scheme == MST_SCHEME_ROE;
    const int lim = x->cfg.limiter;
#define MST_FUSED_L(SC, LM, MF) \
    return launch_shm(x, MST_K_FUSED, k_fused<DIM, SC, LM, T, MF>, x->nc, shm, ga, fa, a, pa, ha);
#define MST_FUSED_S(SC, MF)                                                                   \
    if (lim == MST_LIM_NONE) { MST_FUSED_L(SC, MST_LIM_NONE, MF) }                            \
    if (lim == MST_LIM_MINMOD) { MST_FUSED_L(SC, MST_LIM_MINMOD, MF) }                        \
    MST_FUSED_L(SC, MST_LIM_VANLEER, MF)
#define MST_FUSED(MF)                          \
    if (roe) { MST_FUSED_S(MST_SCHEME_ROE, MF) } \
    MST_FUSED_S(MST_SCHEME_AUSM, MF)
    if (x->slots == 3) { MST_FUSED(3) }
    MST_FUSED(4)
#undef MST_FUSED
#undef MST_FUSED_S
#undef MST_FUSED_L
}

int resolve_strategy(mst_ctx *x);

// one residual evaluation + stage update (mode includes UPD_STAGE) with the context's strategy
int run_stage(mst_ctx *x, int mode, int is, int in, const Stage &sg, int iout) {
    if (x->resolved < 0) {
        int rc = resolve_strategy(x);
        if (rc) return rc;
    }
    if (x->resolved == MST_STRATEGY_FACE || x->viscous) {  // the viscous kernels follow the flux kernel
        int rc = run_face_flux(x, is);
        if (rc) return rc;
        return run_update(x, mode, is, in, sg, iout);
    }
    int rc = ensure_shadow(x, is);
    if (rc) return rc;
    if (x->resolved == MST_STRATEGY_FUSED) {
        if (x->mixed)
            rc = x->dim == 2 ? run_fused_t<2, float>(x, mode, is, in, sg, iout)
                             : run_fused_t<3, float>(x, mode, is, in, sg, iout);
        else
            rc = x->dim == 2 ? run_fused_t<2, double>(x, mode, is, in, sg, iout)
                             : run_fused_t<3, double>(x, mode, is, in, sg, iout);
        if (rc == MST_OK && iout >= 0) x->shadow_ok[iout] = x->mixed;
        return rc;
    }
    if (x->cfg.order >= 2) {
        rc = run_grad(x, is, false);
        if (rc) return rc;
    }
    if (x->resolved == MST_STRATEGY_PATCH) {
        if (x->mixed)
            rc = x->dim == 2 ? run_patch_t<2, float>(x, mode, is, in, sg, iout)
                             : run_patch_t<3, float>(x, mode, is, in, sg, iout);
        else
            rc = x->dim == 2 ? run_patch_t<2, double>(x, mode, is, in, sg, iout)
                             : run_patch_t<3, double>(x, mode, is, in, sg, iout);
    } else if (x->mixed) {
        rc = x->dim == 2 ? run_cell_t<2, float>(x, mode, is, in, sg, iout)
                         : run_cell_t<3, float>(x, mode, is, in, sg, iout);
    } else {
        rc = x->dim == 2 ? run_cell_t<2, double>(x, mode, is, in, sg, iout)
                         : run_cell_t<3, double>(x, mode, is, in, sg, iout);
    }
    if (rc == MST_OK && iout >= 0) x->shadow_ok[iout] = x->mixed;
    return rc;
}

// finish of f's partials: one block when they are few, else a first level of FIN_CHUNK-partial
// blocks into fin_mid and one block over those (fixed partition: bit-reproducible)
int launch_finish(mst_ctx *x, FinArgs f) {
    f.chunk = f.nblk;
    f.mid = nullptr;
    // sums keep one fixed two-level order above 2 FIN_CHUNK partials (bit-reproducible norms); a lone
    // minimum (the CFL time step) does not depend on the order: one block, one launch, up to
    // 16 FIN_CHUNK partials
    if (f.n == 1 && f.kind[0] == 1 && f.post[0] == 0 && f.nblk <= 16 * FIN_CHUNK) {
        hipLaunchKernelGGL(k_finish_min<0>, dim3(1), dim3(FIN), 0, x->stream, f.src[0], f.nblk, f.dst[0]);
        HIPCHK(hipGetLastError());
        return MST_OK;
    }
    if (f.nblk > 2 * FIN_CHUNK) {
        const int64_t g = (f.nblk + FIN_CHUNK - 1) / FIN_CHUNK;
        FinArgs a = f;
        a.chunk = FIN_CHUNK;
        a.mid = x->fin_mid;
        hipLaunchKernelGGL(k_finish, dim3((uint32_t)g), dim3(FIN), 0, x->stream, a);
        HIPCHK(hipGetLastError());
        for (int k = 0; k < f.n; k++) f.src[k] = x->fin_mid + k * g;
        f.nblk = g;
        f.chunk = g;
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN), 0, x->stream, f);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

// one finish: L2 norms (nl2 fields) and/or the minimum time step, from the per-block partials
int run_finish(mst_ctx *x, int nl2, int want_dt) {
    FinArgs f;
    memset(&f, 0, sizeof(f));
    const int64_t nb = x->nblk_c;
    f.nblk = nb;
    f.count = (double)x->nc;
    for (int k = 0; k < nl2; k++) {
        f.src[f.n] = x->part + k * nb;
        f.kind[f.n] = 0;
        f.dst[f.n] = x->scal + 1 + k;
        f.post[f.n] = 1;
        f.n++;
    }
    if (want_dt) {
        f.src[f.n] = x->part + x->nv * nb;
        f.kind[f.n] = 1;
        f.dst[f.n] = x->scal;
        f.post[f.n] = 0;
        f.n++;
    }
    return launch_finish(x, f);
}

int run_cfl(mst_ctx *x, const double *Qsrc, double cfl) {
    double *pdt = x->part + x->nblk_c * x->nv;
    int rc = x->dim == 2 ? launch(x, MST_K_DT, k_cfl<2>, x->nc, x->nc, Qsrc, (const double *)x->cvol,
                                  (const double *)x->cshalf, x->cfg.gamma, cfl, pdt)
                         : launch(x, MST_K_DT, k_cfl<3>, x->nc, x->nc, Qsrc, (const double *)x->cvol,
                                  (const double *)x->cshalf, x->cfg.gamma, cfl, pdt);
    if (rc) return rc;
    return run_finish(x, 0, 1);
}

template <typename T>
int dalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
    return MST_OK;
}
template <typename T>
int dupload(T **p, const std::vector<T> &v) {
    int rc = dalloc(p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return MST_OK;
}

// device arrays of the mixed form (allocated the first time a mixed configuration is set)
int alloc_mixed(mst_ctx *x) {
    if (x->rec32) return MST_OK;
    for (int i = 0; i < 4; i++) {
        int rc = dalloc(&x->Q32[i], x->nc * x->nv);
        if (rc) return rc;
        x->shadow_ok[i] = false;
    }
    return dalloc(&x->rec32, (int64_t)x->slots * x->nc * x->nv);
}

// incidence arrays from the padded ELL table: code, signed normal, offset (FP64, and FP32 copies
// in the mixed form); the weights follow the configuration (upload_incidence_weights)
int build_incidence(mst_ctx *x) {
    const mst::Prepared &P = x->prep.P;
    const int dim = x->dim;
    const int64_t n = (int64_t)x->slots * x->nc;
    std::vector<int32_t> code(n, -2);
    std::vector<double> nrm(n * dim, 0.0), dx(n * dim, 0.0);
    for (int s = 0; s < x->slots; s++)
        for (int64_t c = 0; c < x->nc; c++) {
            const int64_t i = (int64_t)s * x->nc + c;
            if (mst::incidence(P, s, c, code[i], &nrm[i * dim], &dx[i * dim]) >= 0) x->n_inc++;
        }
    // records of each face's sides: the incidence (s, c) whose table entry is the face
    std::vector<int32_t> r0(x->nf, -1), r1(x->nf, -1);
    for (int s = 0; s < P.maxf; s++)
        for (int64_t c = 0; c < x->nc; c++) {
            const int32_t e = P.ell[(int64_t)s * x->nc + c];
            if (e < 0) continue;
            ((e & 1) ? r1 : r0)[e >> 1] = (int32_t)((int64_t)s * x->nc + c);
        }
    int rc = dupload(&x->inc_code, code);
    if (!rc) rc = dupload(&x->inc_n, nrm);
    if (!rc) rc = dupload(&x->inc_dx, dx);
    if (!rc) rc = dalloc(&x->inc_w, n);
    if (!rc) rc = dupload(&x->frec0, r0);
    if (!rc) rc = dupload(&x->frec1, r1);
    return rc;
}

// Patch tables: for each block of BLOCK consecutive cells, the faces that carry a flux into one of its
// cells (ascending; a face joining two of its cells once), and for each cell slot (local index << 1) |
// is_c1, or -1 for an empty slot or a zero-flux face; and each patch's halo for the fused kernel.
// The kernels keep a patch's face states and fluxes in LDS, so the strategies need at most 4 slots and
// 16 * NV bytes per face of the largest patch within 64 KiB.
int build_patches(mst_ctx *x) {
    const mst::Prepared &P = x->prep.P;
    x->patch_ok = x->fused_ok = false;
    if (x->slots > 4) return MST_OK;
    const int64_t nc = x->nc, nb = (nc + BLOCK - 1) / BLOCK;
    std::vector<int32_t> start(nb + 1, 0), face, list, hstart(nb + 1, 0), hcell, halo;
    std::vector<int16_t> loc((size_t)x->slots * nc, -1), hloc;
    face.reserve((size_t)(nc * 2));
    size_t max_pf = 0;
    for (int64_t b = 0; b < nb; b++) {
        const int64_t ca = b * BLOCK, cz = std::min(nc, ca + BLOCK);
        list.clear();
        for (int s = 0; s < P.maxf; s++)
            for (int64_t c = ca; c < cz; c++) {
                const int32_t e = P.ell[(int64_t)s * nc + c];
                if (e >= 0 && P.fkind[e >> 1] < mst::KIND_NONE) list.push_back(e >> 1);
            }
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        max_pf = std::max(max_pf, list.size());
        for (int s = 0; s < P.maxf; s++)
            for (int64_t c = ca; c < cz; c++) {
                const int32_t e = P.ell[(int64_t)s * nc + c];
                if (e < 0 || P.fkind[e >> 1] >= mst::KIND_NONE) continue;
                const int64_t l = std::lower_bound(list.begin(), list.end(), e >> 1) - list.begin();
                loc[(size_t)s * nc + c] = (int16_t)((l << 1) | (e & 1));
            }
        start[b] = (int32_t)face.size();
        face.insert(face.end(), list.begin(), list.end());
        // halo: the other cell of every patch face with one cell outside the patch, and for each of its
        // slots the local index of the slot's face when the face is one of the patch's
        halo.clear();
        for (const int32_t f : list) {
            if (P.fkind[f] != mst::KIND_INTERIOR && P.fkind[f] != mst::KIND_PERIODIC) continue;
            for (const int32_t c : {P.fc0[f], P.fc1[f]})
                if (c < ca || c >= cz) halo.push_back(c);
        }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        hstart[b] = (int32_t)hcell.size();
        for (const int32_t c : halo) {
            hcell.push_back(c);
            for (int s = 0; s < x->slots; s++) {
                const int32_t e = s < P.maxf ? P.ell[(int64_t)s * nc + c] : -1;
                int16_t v = -1;
                if (e >= 0 && P.fkind[e >> 1] < mst::KIND_NONE) {
                    const auto it = std::lower_bound(list.begin(), list.end(), e >> 1);
                    if (it != list.end() && *it == (e >> 1)) v = (int16_t)(((it - list.begin()) << 1) | (e & 1));
                }
                hloc.push_back(v);
            }
        }
    }
    start[nb] = (int32_t)face.size();
    hstart[nb] = (int32_t)hcell.size();
    // the fused kernel keeps 16 * NV bytes per patch face in LDS (the patch kernel half of that)
    if (max_pf * x->nv * 16 > 65536) return MST_OK;
    int rc = dupload(&x->pstart, start);
    if (!rc) rc = dupload(&x->pface, face);
    if (!rc && x->nf < ((int64_t)1 << FI_SHIFT)) rc = dalloc(&x->pfi, face.size());  // packed by upload_config
    if (!rc) rc = dupload(&x->ploc, loc);
    if (!rc) rc = dupload(&x->hstart, hstart);
    if (!rc) rc = dupload(&x->hcell, hcell);
    if (!rc) rc = dupload(&x->hloc, hloc);
    if (rc) return rc;
    x->n_halo = (int64_t)hcell.size();
    x->max_pf = (int)max_pf;
    x->n_psolve = (int64_t)face.size();
    x->patch_ok = true;
    x->fused_ok = fused_lds_bytes(x->dim, x->nv, x->max_pf, false, x->dim == 2) <= 65536;
    return MST_OK;
}

// per incidence: FP64 eta of the face (with the last-interior-face quirk), FP32 the difference-form
// weight (1 - eta on the c0 side, eta on the c1 side)
int upload_incidence_weights(mst_ctx *x, const std::vector<double> &eta) {
    const mst::Prepared &P = x->prep.P;
    const int64_t n = (int64_t)x->slots * x->nc;
    std::vector<double> w(n, 0.0);
    std::vector<float> w32(x->mixed ? n : 0, 0.f);
    for (int s = 0; s < P.maxf; s++)
        for (int64_t c = 0; c < x->nc; c++) {
            const int32_t e = P.ell[(int64_t)s * x->nc + c];
            if (e < 0) continue;
            const int64_t i = (int64_t)s * x->nc + c;
            w[i] = eta[e >> 1];
            if (x->mixed) {
                const float ef = (float)eta[e >> 1];
                w32[i] = (e & 1) ? ef : 1.0f - ef;
            }
        }
    HIPCHK(hipMemcpy(x->inc_w, w.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    if (x->mixed) {
        if (!x->inc_w32) {
            const int dim = x->dim;
            int rc = dalloc(&x->inc_w32, n);
            if (rc) return rc;
            std::vector<double> h(n * dim);
            std::vector<float> f(n * dim);
            for (auto pr : {std::make_pair(x->inc_n, &x->inc_n32), std::make_pair(x->inc_dx, &x->inc_dx32)}) {
                if (dim == 3 && pr.second == &x->inc_n32) {  // weighted normals, filled below
                    rc = dalloc(&x->inc_n32, n * dim);
                    if (rc) return rc;
                    continue;
                }
                HIPCHK(hipMemcpy(h.data(), pr.first, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
                for (size_t k = 0; k < h.size(); k++) f[k] = (float)h[k];
                rc = dupload(pr.second, f);
                if (rc) return rc;
            }
            rc = dalloc(&x->ivol32, x->nc);
            if (rc) return rc;
            hipLaunchKernelGGL(k_recip_vol<0>, dim3((uint32_t)((x->nc + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, x->stream,
                               x->nc, (const double *)x->cvol, x->ivol32);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpy(x->inc_w32, w32.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        if (x->dim == 3) {  // the FP32 3D gradient reads the weights folded into its normals
            hipLaunchKernelGGL(k_weighted_normals<3>, dim3((uint32_t)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, x->stream,
                               n, (const int32_t *)x->inc_code, (const double *)x->inc_w, (const double *)x->inc_n,
                               x->inc_n32);
            HIPCHK(hipGetLastError());
        }
    }
    x->hg_dirty = true;
    return MST_OK;
}

// the fused kernel's packed halo geometry, gathered on the context's stream when the weights changed
int ensure_halo_geo(mst_ctx *x) {
    if (!x->hg_dirty || !x->patch_ok || x->n_halo == 0) return MST_OK;
    const IncView<double> f64{x->inc_code, x->inc_w, x->inc_n, x->inc_dx, x->cvol, x->nc};
    const IncView<float> f32{x->mixed ? x->inc_code : nullptr, x->inc_w32, x->inc_n32, x->inc_dx32, x->ivol32, x->nc};
    HIPCHK(x->dim == 2 ? x->hg.fill<2>(x->slots, x->nc, x->n_halo, x->hcell, f64, f32, x->stream)
                       : x->hg.fill<3>(x->slots, x->nc, x->n_halo, x->hcell, f64, f32, x->stream));
    x->hg_dirty = false;
    return MST_OK;
}

int upload_config(mst_ctx *x) {
    std::vector<int32_t> mask;
    mst::lr_masks(x->prep.P, x->cfg.compat, mask);
    std::vector<int32_t> info(x->nf);
    for (int64_t i = 0; i < x->nf; i++) info[i] = x->prep.P.fkind[i] | (mask[i] << 4);
    HIPCHK(hipMemcpy(x->finfo, info.data(), sizeof(int32_t) * x->nf, hipMemcpyHostToDevice));
    if (x->pfi && x->n_psolve > 0) {  // the patch face entries with the new kinds and masks
        const int64_t n = x->n_psolve;
        hipLaunchKernelGGL(k_pack_faces<0>, dim3((uint32_t)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, x->stream, n,
                           (const int32_t *)x->pface, (const int32_t *)x->finfo, x->pfi);
        HIPCHK(hipGetLastError());
    }
    std::vector<double> eta = x->prep.P.feta;
    if ((x->cfg.compat & MST_COMPAT_GRAD_LASTINT) && x->prep.P.quirk_face >= 0) eta[x->prep.P.quirk_face] = 1.0;
    HIPCHK(hipMemcpy(x->feta, eta.data(), sizeof(double) * x->nf, hipMemcpyHostToDevice));
    x->p64 = make_params<double>(x->cfg);
    x->p32 = make_params<float>(x->cfg);
    x->mixed = x->cfg.precision == MST_PREC_MIXED;
    x->viscous = x->cfg.mu > 0 || x->cfg.kappa > 0;
    if (x->viscous && !x->gw) {
        int rc = dalloc(&x->gw, (size_t)(x->nc * (x->dim + 1) * x->dim));
        if (rc) return rc;
    }
    if (x->mixed) {
        int rc = alloc_mixed(x);
        if (rc) return rc;
        for (int i = 0; i < 4; i++) x->shadow_ok[i] = false;
    }
    return upload_incidence_weights(x, eta);
}

void release(mst_ctx *x) {
    if (!x) return;
    void *bufs[] = {x->fc0, x->fc1, x->finfo, x->ell, x->fn, x->fxf, x->pxf, x->feta, x->cvol, x->cxc,
                    x->cshalf, x->Q[0], x->Q[1], x->Q[2], x->Q[3], x->G, x->rec, x->graw, x->phi, x->R,
                    x->scal, x->part, x->fin_mid, x->bad, x->Q32[0], x->Q32[1], x->Q32[2], x->Q32[3], x->rec32,
                    x->frec0, x->frec1, x->inc_code, x->inc_w, x->inc_n, x->inc_dx, x->inc_w32, x->inc_n32,
                    x->inc_dx32, x->ivol32, x->pstart, x->pface, x->pfi, x->ploc, x->hstart, x->hcell, x->hloc,
                    x->lu_cells, x->lu_nb, x->lu_face, x->lu_n, x->lu_r, x->lu_D, x->lu_dQ, x->gw};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    x->hg.release();
    for (int k = 0; k < MST_K_COUNT; k++)
        for (auto &pr : x->ev[k]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto e : x->pool) (void)hipEventDestroy(e);
    if (x->stream) (void)hipStreamDestroy(x->stream);
    delete x;
}

int permute_in(const mst_ctx *x, const double *Q, std::vector<double> &buf) {
    const int nv = x->nv;
    buf.resize(x->nc * nv);
    const auto &old_of_new = x->prep.P.cell_old_of_new;
    for (int64_t i = 0; i < x->nc; i++)
        for (int k = 0; k < nv; k++) buf[i * nv + k] = Q[(int64_t)old_of_new[i] * nv + k];
    return MST_OK;
}

int fetch_cells(mst_ctx *x, const double *dsrc, int width, double *out) {
    std::vector<double> buf(x->nc * width);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(buf.data(), dsrc, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    const auto &new_of_old = x->prep.P.cell_new_of_old;
    for (int64_t c = 0; c < x->nc; c++)
        for (int k = 0; k < width; k++) out[c * width + k] = buf[(int64_t)new_of_old[c] * width + k];
    return MST_OK;
}

int check_bad(mst_ctx *x) {
    int h = 0;
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(&h, x->bad, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        HIPCHK(hipMemset(x->bad, 0, sizeof(int)));
        return mst::fail(MST_E_STATE, "non-physical state (rho<=0, p<=0 or NaN, or no positive time step) produced by the update");
    }
    return MST_OK;
}

// LU-SGS: the cell colouring (prep.cpp, lusgs_colours), the cells sorted by colour (ascending inside
// a colour), and per (slot, position) the neighbour record and the cell's outward normal.  Built once
// per context, on the first implicit iteration.
int build_colours(mst_ctx *x) {
    if (x->lu_cells) return MST_OK;
    if (x->nc >= ((int64_t)1 << 29)) return mst::fail(MST_E_UNSUPPORTED, "LU-SGS records hold cell indices below 2^29");
    const auto &P = x->prep.P;
    const int64_t nc = x->nc;
    const int dim = x->dim;
    std::vector<int32_t> col;
    const int ncol = lusgs_colours(P, col);
    std::vector<int64_t> start(ncol + 1, 0);
    for (int64_t c = 0; c < nc; c++) start[col[c] + 1]++;
    for (int k = 0; k < ncol; k++) start[k + 1] += start[k];
    std::vector<int32_t> cells(nc);
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t c = 0; c < nc; c++) cells[fill[col[c]]++] = (int32_t)c;
    const int64_t slots = x->slots;
    std::vector<int32_t> nb(slots * nc, -2), fcode(slots * nc, -1);
    std::vector<double> nrm(slots * nc * dim, 0.0);
    for (int64_t t = 0; t < nc; t++) {
        const int64_t c = cells[t];
        for (int s = 0; s < P.maxf; s++) {
            const int32_t e = P.ell[(int64_t)s * nc + c];
            if (e < 0) break;
            const int32_t f = e >> 1;
            const bool c1side = e & 1;
            const int32_t o = c1side ? P.fc0[f] : P.fc1[f];
            const int64_t i = (int64_t)s * nc + t;
            for (int d = 0; d < dim; d++) nrm[i * dim + d] = c1side ? -P.fn[(int64_t)f * dim + d] : P.fn[(int64_t)f * dim + d];
            nb[i] = o < 0 ? -1 : ((o << 2) | (col[o] < col[c] ? 1 : 0) | (col[o] > col[c] ? 2 : 0));
            fcode[i] = e;
        }
    }
    int rc = dupload(&x->lu_cells, cells);
    if (rc == MST_OK) rc = dupload(&x->lu_nb, nb);
    if (rc == MST_OK) rc = dupload(&x->lu_face, fcode);
    if (rc == MST_OK) rc = dupload(&x->lu_n, nrm);
    if (rc == MST_OK) rc = dalloc(&x->lu_r, (size_t)(slots * nc));
    if (rc == MST_OK) rc = dalloc(&x->lu_D, (size_t)nc);
    if (rc == MST_OK) rc = dalloc(&x->lu_dQ, (size_t)(nc * x->nv));
    if (rc == MST_OK && !x->R) rc = dalloc(&x->R, (size_t)(nc * x->nv));
    if (rc) return rc;
    x->lu_start = start;
    x->lu_colour = std::move(col);
    return MST_OK;
}

// forward launches; RS >= 0: the first-order residual computed by the sweep (scheme RS)
template <int DIM, int MAXF, int RS>
int lu_forward(mst_ctx *x, const LuArgs &a) {
    const int ncol = (int)x->lu_start.size() - 1;
    int rc = MST_OK;
    for (int k = 0; k < ncol && rc == MST_OK; k++) {
        const int64_t t0 = x->lu_start[k], n = x->lu_start[k + 1] - t0;
        rc = k == ncol - 1 ? launch(x, MST_K_LUSGS, k_lu_sweep<DIM, true, true, MAXF, RS>, n, a, t0, n)
                           : launch(x, MST_K_LUSGS, k_lu_sweep<DIM, true, false, MAXF, RS>, n, a, t0, n);
    }
    return rc;
}

template <int DIM, int MAXF>
int lu_sweeps(mst_ctx *x, const LuArgs &a, bool inline_res) {
    const int ncol = (int)x->lu_start.size() - 1;
    int rc;
    if constexpr (MAXF <= 4) {
        if (inline_res)
            rc = x->cfg.scheme == MST_SCHEME_ROE ? lu_forward<DIM, MAXF, MST_SCHEME_ROE>(x, a)
                                                 : lu_forward<DIM, MAXF, MST_SCHEME_AUSM>(x, a);
        else
            rc = lu_forward<DIM, MAXF, -1>(x, a);
    } else {
        rc = lu_forward<DIM, MAXF, -1>(x, a);
    }
    for (int k = ncol - 2; k >= 0 && rc == MST_OK; k--) {
        const int64_t t0 = x->lu_start[k], n = x->lu_start[k + 1] - t0;
        rc = launch(x, MST_K_LUSGS, k_lu_sweep<DIM, false, false, MAXF>, n, a, t0, n);
    }
    return rc;
}

// the sweeps' arguments for the current Old (the iteration's start) and New buffers
LuArgs lu_args(const mst_ctx *x) {
    LuArgs a;
    memset(&a, 0, sizeof(a));  // the batch-only fields (pcase, gam, active, G) stay null: single context
    a.nc = x->nc;
    a.cells = x->lu_cells;
    a.nb = x->lu_nb;
    a.n = x->lu_n;
    a.r = x->lu_r;
    a.D = x->lu_D;
    a.Q = x->Q[x->iold];
    a.R = x->R;
    a.dQ = x->lu_dQ;
    a.Qout = x->Q[x->inew];
    a.bad = x->bad;
    a.gamma = x->cfg.gamma;
    a.omega = x->lu_omega;
    a.dscale = 1.0 / x->lu_cfl + 0.5 * x->lu_omega;
    a.face = x->lu_face;
    a.finfo = x->finfo;
    a.P = x->p64;
    return a;
}

// one LU-SGS iteration: residual of Old (+ its L2 norms when l2 is set), the two sweeps into New
int lu_iteration(mst_ctx *x, bool l2) {
    int rc = MST_OK;
    const int mode = l2 ? UPD_L2 : UPD_RESIDUAL;
    // order 1, FP64, inviscid, cells of <= 4 faces: the forward sweeps compute the residual themselves
    // (the norms of a call's last iteration still come from the residual kernel)
    x->lu_inline = x->cfg.order == 1 && !x->mixed && !x->viscous && x->slots <= 4;
    const bool inline_res = !l2 && x->lu_inline;
    if (inline_res) {
        // nothing to launch before the sweeps
    } else if (x->patch_ok && !x->mixed && !x->viscous) {  // one launch: the patch kernel's residual-only form (same R)
        if (x->cfg.order >= 2) {
            rc = run_grad(x, x->iold, false);
            if (rc) return rc;
        }
        rc = x->dim == 2 ? run_patch_residual_t<2>(x, mode, x->iold) : run_patch_residual_t<3>(x, mode, x->iold);
    } else {
        rc = run_face_flux(x, x->iold);
        if (rc) return rc;
        rc = run_update(x, mode, x->iold, -1, Stage{0, 0}, -1);
    }
    if (rc) return rc;
    if (l2) {
        rc = run_finish(x, x->nv, 0);
        if (rc) return rc;
    }
    const LuArgs a = lu_args(x);
    // the records are padded to the slot count of the device table (3, 4 or 8)
    if (x->dim == 2) {
        if (x->slots == 3) return lu_sweeps<2, 3>(x, a, inline_res);
        if (x->slots == 4) return lu_sweeps<2, 4>(x, a, inline_res);
        return lu_sweeps<2, 8>(x, a, inline_res);
    }
    if (x->slots == 3) return lu_sweeps<3, 3>(x, a, inline_res);
    if (x->slots == 4) return lu_sweeps<3, 4>(x, a, inline_res);
    return lu_sweeps<3, 8>(x, a, inline_res);
}

// n >= 1 consecutive iterations with the residual computed in the forward sweeps (lu_inline, two
// colours or more): colour 0's backward sweep of each iteration but the last runs fused with the next
// iteration's forward sweep of colour 0 (k_lu_sweep_bf), 2 K - 2 launches per iteration for K colours
// instead of 2 K - 1, with the same results.  Each iteration ends with the Old/New swap.
template <int DIM, int MAXF, int RS>
int lu_inline_run(mst_ctx *x, int n) {
    const int ncol = (int)x->lu_start.size() - 1;
    const int64_t n0 = x->lu_start[1] - x->lu_start[0];
    int rc = MST_OK;
    for (int it = 0; it < n && rc == MST_OK; it++) {
        const LuArgs a = lu_args(x);
        for (int k = it == 0 ? 0 : 1; k < ncol && rc == MST_OK; k++) {  // colour 0: done by the fused launch
            const int64_t t0 = x->lu_start[k], m = x->lu_start[k + 1] - t0;
            rc = k == ncol - 1 ? launch(x, MST_K_LUSGS, k_lu_sweep<DIM, true, true, MAXF, RS>, m, a, t0, m)
                               : launch(x, MST_K_LUSGS, k_lu_sweep<DIM, true, false, MAXF, RS>, m, a, t0, m);
        }
        for (int k = ncol - 2; k >= 1 && rc == MST_OK; k--) {
            const int64_t t0 = x->lu_start[k], m = x->lu_start[k + 1] - t0;
            rc = launch(x, MST_K_LUSGS, k_lu_sweep<DIM, false, false, MAXF>, m, a, t0, m);
        }
        if (rc) break;
        rc = it < n - 1 ? launch(x, MST_K_LUSGS, k_lu_sweep_bf<DIM, MAXF, RS>, n0, a, x->lu_start[0], n0)
                        : launch(x, MST_K_LUSGS, k_lu_sweep<DIM, false, false, MAXF>, n0, a, x->lu_start[0], n0);
        x->shadow_ok[x->inew] = false;
        std::swap(x->iold, x->inew);
    }
    return rc;
}

template <int DIM, int MAXF>
int lu_inline_run_s(mst_ctx *x, int n) {
    return x->cfg.scheme == MST_SCHEME_ROE ? lu_inline_run<DIM, MAXF, MST_SCHEME_ROE>(x, n)
                                           : lu_inline_run<DIM, MAXF, MST_SCHEME_AUSM>(x, n);
}

int lu_inline_iterations(mst_ctx *x, int n) {
    if (x->dim == 2) return x->slots == 3 ? lu_inline_run_s<2, 3>(x, n) : lu_inline_run_s<2, 4>(x, n);
    return x->slots == 3 ? lu_inline_run_s<3, 3>(x, n) : lu_inline_run_s<3, 4>(x, n);
}


// AUTO: time one stage of each strategy on this mesh and configuration (outputs go to a spare state
// buffer; the Old and New states are untouched) and keep the faster
int resolve_strategy(mst_ctx *x) {
    if (x->strategy != MST_STRATEGY_AUTO) {
        x->resolved = x->strategy;
        return MST_OK;
    }
    int spare = 0;
    while (spare == x->iold || spare == x->inew) spare++;
    const bool timing = x->timing;
    x->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = INFINITY;
    int pick = MST_STRATEGY_FACE, rc = MST_OK;
    for (int s : {MST_STRATEGY_FACE, MST_STRATEGY_CELL, MST_STRATEGY_PATCH, MST_STRATEGY_FUSED}) {
        if ((s == MST_STRATEGY_PATCH || s == MST_STRATEGY_FUSED) && !x->patch_ok) continue;
        if (s == MST_STRATEGY_FUSED && !x->fused_ok) continue;
        x->resolved = s;
        rc = run_stage(x, UPD_STAGE, x->iold, x->iold, Stage{}, spare);  // warm-up
        if (rc) break;
        if (hipEventRecord(e0, x->stream) != hipSuccess) break;
        for (int r = 0; r < 3 && !rc; r++) rc = run_stage(x, UPD_STAGE, x->iold, x->iold, Stage{}, spare);
        if (rc || hipEventRecord(e1, x->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
            best = ms;
            pick = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    x->timing = timing;
    x->resolved = pick;
    // the trial stages may have flagged their (unchecked) outputs, e.g. with no time step set yet
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemset(x->bad, 0, sizeof(int)));
    if (pick != MST_STRATEGY_FUSED) {  // the fused trial's halo copy (0.64 N_c entries in 3D) is not kept
        x->hg.release();
        x->hg_dirty = true;
    }
    return rc;
}

}  // namespace

#define CTX_GUARD(x)                                           \
    if (!(x)) return mst::fail(MST_E_ARG, "null context");

// ------------------------------------------------------------------ allocation-free monitor
// The monitor over caller-owned device states, for callers that cannot allocate.  Its only scratch,
// one partial maximum per block and field, lives in the file-scope array below: NOALLOC_GRID blocks
// stride over any number of cells, so the array never grows.  Each device holds its own copy; its
// address is looked up once per device (hipGetSymbolAddress, with that device current) and kept.
// Every block writes its partials before the finish reads them, so nothing needs clearing between
// calls; calls that overlap on one device share the array and race.
namespace {
constexpr int NOALLOC_GRID = 2048;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_GRID * 5];
std::mutex noalloc_mu;
double *noalloc_scratch[NOALLOC_MAX_DEVICES];

int noalloc_scratch_ptr(double **out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return mst::fail(MST_E_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (!noalloc_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return mst::fail(MST_E_HIP, "hipGetSymbolAddress returned no address for the monitor scratch");
        noalloc_scratch[dev] = static_cast<double *>(p);
    }
    *out = noalloc_scratch[dev];
    return MST_OK;
}
}  // namespace

extern "C" {

int mst_ctx_create(const mst_mesh *m, const mst_config *cfg, mst_ctx **out) {
    if (!m || !out) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return mst::fail(MST_E_HIP, "no HIP device available: the residual path runs only on the GPU");
    mst_ctx *x = new (std::nothrow) mst_ctx();
    if (!x) return mst::fail(MST_E_ARG, "out of host memory");
    try {
        rc = mst::prepare(*m, x->prep.P);
    } catch (const std::bad_alloc &) {
        rc = mst::fail(MST_E_ARG, "out of host memory");
    }
    if (rc) {
        delete x;
        return rc;
    }
    if (x->prep.P.bad_cells > 0) {
        rc = mst::fail(MST_E_MESH, std::to_string(x->prep.P.bad_cells) + " cells have non-positive or NaN volume "
                                   "(first: file cell " + std::to_string(x->prep.P.first_bad_cell) + ")");
        delete x;
        return rc;
    }
    const mst::Prepared &P = x->prep.P;
    x->cfg = *cfg;
    x->dim = P.dim;
    x->nv = P.nv;
    x->maxf = P.maxf;
    x->nc = P.nc;
    x->nf = P.nf;
    x->nblk_c = (x->nc + BLOCK - 1) / BLOCK;
    x->nblk_f = (x->nf + BLOCK - 1) / BLOCK;
#define TRY(e)            \
    do {                  \
        int _r = (e);     \
        if (_r) {         \
            release(x);   \
            return _r;    \
        }                 \
    } while (0)
    if (hipGetDevice(&x->device) != hipSuccess) {
        delete x;
        return mst::fail(MST_E_HIP, "hipGetDevice failed");
    }
    if (hipStreamCreate(&x->stream) != hipSuccess) {
        delete x;
        return mst::fail(MST_E_HIP, "hipStreamCreate failed");
    }
    TRY(dupload(&x->fc0, P.fc0));
    TRY(dupload(&x->fc1, P.fc1));
    TRY(dalloc(&x->finfo, x->nf));
    {
        // the gather loops are instantiated for 3, 4 or 8 slots: pad the slot-major table with empty
        // rows up to that count so an unrolled loop over a full cell never reads past it
        const int64_t slots = x->maxf <= 3 ? 3 : (x->maxf <= 4 ? 4 : 8);
        x->slots = (int)slots;
        if (slots * x->nc >= (int64_t)INT32_MAX) {  // incidence records are indexed in 32 bits
            release(x);
            return mst::fail(MST_E_UNSUPPORTED, "mesh too large: cells x face slots must stay below 2^31");
        }
        std::vector<int32_t> ell(P.ell);
        ell.resize(slots * x->nc, -1);
        TRY(dupload(&x->ell, ell));
    }

    TRY(dupload(&x->fn, P.fn));
    TRY(dupload(&x->fxf, P.fxf));
    TRY(dupload(&x->pxf, P.pxf));
    TRY(dalloc(&x->feta, x->nf));
    TRY(dupload(&x->cvol, P.cvol));
    TRY(dupload(&x->cxc, P.cxc));
    TRY(dupload(&x->cshalf, P.cshalf));
    for (int i = 0; i < 4; i++) TRY(dalloc(&x->Q[i], x->nc * x->nv));
    TRY(dalloc(&x->G, x->nf * x->nv));
    TRY(dalloc(&x->rec, (int64_t)x->slots * x->nc * x->nv));
    TRY(dalloc(&x->scal, 16));
    TRY(dalloc(&x->part, std::max(x->nblk_c, x->nblk_f) * (x->nv + 1) + 16));
    TRY(dalloc(&x->fin_mid, 12 * (x->nblk_c / FIN_CHUNK + 2)));
    TRY(dalloc(&x->bad, 1));
    if (hipMemset(x->bad, 0, sizeof(int)) != hipSuccess || hipMemset(x->scal, 0, 16 * sizeof(double)) != hipSuccess ||
        hipMemset(x->rec, 0, sizeof(double) * x->slots * x->nc * x->nv) != hipSuccess) {
        release(x);
        return mst::fail(MST_E_HIP, "hipMemset failed");
    }
    TRY(build_incidence(x));
    TRY(build_patches(x));
    TRY(upload_config(x));
#undef TRY
    *out = x;
    return MST_OK;
}

int mst_ctx_destroy(mst_ctx *ctx) {
    if (!ctx) return MST_OK;
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    return MST_OK;
}

int mst_ctx_set_config(mst_ctx *ctx, const mst_config *cfg) {
    CTX_GUARD(ctx);
    int rc = check_config(cfg);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cfg = *cfg;
    ctx->dt_valid = false;
    if (ctx->strategy == MST_STRATEGY_AUTO) ctx->resolved = -1;  // the best strategy depends on the configuration
    return upload_config(ctx);
}

const mst_prep *mst_ctx_prep(const mst_ctx *ctx) { return ctx ? &ctx->prep : nullptr; }

int mst_set_state(mst_ctx *ctx, const double *Q, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (!Q || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "state size does not match the mesh");
    std::vector<double> buf;
    permute_in(ctx, Q, buf);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->Q[ctx->iold], buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    ctx->shadow_ok[ctx->iold] = false;
    ctx->dt_valid = false;
    return MST_OK;
}

int mst_get_state(mst_ctx *ctx, double *Q, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (!Q || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "state size does not match the mesh");
    return fetch_cells(ctx, ctx->Q[ctx->iold], ctx->nv, Q);
}

int mst_get_new_state(mst_ctx *ctx, double *Q, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (!Q || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "state size does not match the mesh");
    return fetch_cells(ctx, ctx->Q[ctx->inew], ctx->nv, Q);
}

int mst_set_dt(mst_ctx *ctx, double dt) {
    CTX_GUARD(ctx);
    if (!(dt > 0)) return mst::fail(MST_E_ARG, "dt must be positive");
    HIPCHK(hipMemcpyAsync(ctx->scal, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->dt_valid = false;
    return MST_OK;
}

int mst_face_flux(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    return run_face_flux(ctx, ctx->iold);
}

int mst_residual_update(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    int rc = run_update(ctx, UPD_STAGE | UPD_CHECK, ctx->iold, ctx->iold, Stage{}, ctx->inew);
    if (rc) return rc;
    return check_bad(ctx);
}

int mst_solve(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    int rc = run_stage(ctx, UPD_STAGE | UPD_CHECK, ctx->iold, ctx->iold, Stage{}, ctx->inew);
    if (rc) return rc;
    return check_bad(ctx);
}

int mst_set_strategy(mst_ctx *ctx, int strategy) {
    CTX_GUARD(ctx);
    if (strategy != MST_STRATEGY_FACE && strategy != MST_STRATEGY_CELL && strategy != MST_STRATEGY_PATCH &&
        strategy != MST_STRATEGY_FUSED && strategy != MST_STRATEGY_AUTO)
        return mst::fail(MST_E_ARG, "bad strategy");
    if ((strategy == MST_STRATEGY_PATCH || strategy == MST_STRATEGY_FUSED) && !ctx->patch_ok)
        return mst::fail(MST_E_UNSUPPORTED, "patch strategy: needs at most 4 faces per cell and 64 KiB of fluxes per patch");
    if (strategy == MST_STRATEGY_FUSED && !ctx->fused_ok)
        return mst::fail(MST_E_UNSUPPORTED, "fused strategy: a patch's face states exceed 64 KiB of LDS");
    ctx->strategy = strategy;
    ctx->resolved = strategy == MST_STRATEGY_AUTO ? -1 : strategy;
    return MST_OK;
}

int mst_get_strategy(const mst_ctx *ctx, int *strategy) {
    CTX_GUARD(ctx);
    if (!strategy) return mst::fail(MST_E_ARG, "null argument");
    *strategy = ctx->resolved >= 0 ? ctx->resolved : ctx->strategy;
    return MST_OK;
}

int mst_update_new_to_old(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    std::swap(ctx->iold, ctx->inew);
    return MST_OK;
}

int mst_step(mst_ctx *ctx, int nsteps) {
    CTX_GUARD(ctx);
    if (nsteps < 0) return mst::fail(MST_E_ARG, "nsteps must be >= 0");
    static const double A2[2] = {0, 0.5}, B2[2] = {1, 0.5};
    static const double A3[3] = {0, 0.75, 1.0 / 3.0}, B3[3] = {1, 0.25, 2.0 / 3.0};
    const int ns = ctx->cfg.rk;
    const double *A = ns == 3 ? A3 : A2, *B = ns == 3 ? B3 : B2;
    const bool cfl = ctx->cfg.cfl > 0;
    for (int it = 0; it < nsteps; it++) {
        int rc;
        double *Qn = ctx->Q[ctx->iold];
        if (cfl && !ctx->dt_valid) {
            rc = run_cfl(ctx, Qn, ctx->cfg.cfl);
            if (rc) return rc;
        }
        // stage buffers: the two spare buffers alternate, the last stage lands in "new"; the first
        // stage of the call's last step reduces |R/V| (the norms are a monitor: one reduction per
        // call), the last stage of every step the CFL time step of the new state (next step's dt)
        int spare[2], w = 0;
        for (int i = 0; i < 4; i++)
            if (i != ctx->iold && i != ctx->inew) spare[w++] = i;
        const bool l2 = it == nsteps - 1;
        // mixed form: a stage that has a successor stores the successor's combination A Q^n + B Q_s in
        // place of its FP64 state (the FP32 kernels read only the shadow), so no later stage reads Q^n
        // again (SSP-RK2: none; RK3: the middle stage).  The viscous kernels read the FP64 stage state.
        const bool pre = ctx->mixed && !ctx->viscous;
        int src = ctx->iold;
        for (int s = 0; s < ns; s++) {
            const int dst = (s == ns - 1) ? ctx->inew : spare[s & 1];
            const int mode = UPD_STAGE | ((l2 && s == 0) ? UPD_L2 : 0) | ((cfl && s == ns - 1) ? UPD_DT : 0) |
                             (s == ns - 1 ? UPD_CHECK : 0);
            Stage sg;
            sg.ca = A[s];
            sg.cb = B[s];
            sg.stage0 = s == 0;
            sg.pre = pre;
            sg.wa = pre && s < ns - 1 ? A[s + 1] : 0;
            sg.wb = pre && s < ns - 1 ? B[s + 1] : 0;
            rc = run_stage(ctx, mode, src, ctx->iold, sg, dst);
            if (rc) return rc;
            src = dst;
        }
        if (l2 || cfl) {
            rc = run_finish(ctx, l2 ? ctx->nv : 0, cfl ? 1 : 0);
            if (rc) return rc;
        }
        ctx->dt_valid = cfl;
        std::swap(ctx->iold, ctx->inew);
    }
    return check_bad(ctx);
}

int mst_run_to_steady(mst_ctx *ctx, double tol, int max_steps, int check_every, double *ref, int *steps_done,
                      double *drop) {
    CTX_GUARD(ctx);
    if (!ref || !(tol > 0) || max_steps < 1 || check_every < 1) return mst::fail(MST_E_ARG, "bad arguments");
    const bool lu = ctx->steady == MST_STEADY_LUSGS;
    if (!lu && !(ctx->cfg.cfl > 0))
        return mst::fail(MST_E_ARG, "explicit steady runs advance with the CFL time step (cfl > 0)");
    const int nv = ctx->nv;
    double r[5];
    int steps = 0;
    if (!(ref[0] > 0)) {  // reference: the residual of the first step
        int rc = lu ? mst_lusgs_step(ctx, 1) : mst_step(ctx, 1);
        if (rc) return rc;
        rc = mst_residual_norms(ctx, ref);
        if (rc) return rc;
        steps = 1;
    }
    double ratio = INFINITY;
    while (steps < max_steps) {
        const int k = std::min(check_every, max_steps - steps);
        int rc = lu ? mst_lusgs_step(ctx, k) : mst_step(ctx, k);
        if (rc) return rc;
        steps += k;
        rc = mst_residual_norms(ctx, r);
        if (rc) return rc;
        ratio = 0;
        for (int i = 0; i < nv; i++) ratio = std::max(ratio, ref[i] > 0 ? r[i] / ref[i] : 0.0);
        if (ratio <= tol) break;
    }
    if (steps_done) *steps_done = steps;
    if (drop) *drop = ratio;
    return MST_OK;
}

int mst_set_steady_driver(mst_ctx *ctx, int driver, double cfl, double omega) {
    CTX_GUARD(ctx);
    if (driver != MST_STEADY_EXPLICIT && driver != MST_STEADY_LUSGS)
        return mst::fail(MST_E_ARG, "driver must be MST_STEADY_EXPLICIT or MST_STEADY_LUSGS");
    if (driver == MST_STEADY_LUSGS && (!(cfl > 0) || !(omega > 0) || !std::isfinite(cfl)))
        return mst::fail(MST_E_ARG, "LU-SGS needs a finite cfl > 0 and omega > 0");
    ctx->steady = driver;
    if (driver == MST_STEADY_LUSGS) {
        ctx->lu_cfl = cfl;
        ctx->lu_omega = omega;
    }
    return MST_OK;
}

int mst_lusgs_step(mst_ctx *ctx, int n_iter) {
    CTX_GUARD(ctx);
    if (n_iter < 0) return mst::fail(MST_E_ARG, "n_iter must be >= 0");
    int rc = build_colours(ctx);
    if (rc) return rc;
    ctx->lu_inline = ctx->cfg.order == 1 && !ctx->mixed && !ctx->viscous && ctx->slots <= 4;
    const bool fuse = ctx->lu_inline && ctx->lu_start.size() >= 3;  // two colours or more
    for (int it = 0; it < n_iter;) {
        if (fuse && it < n_iter - 1) {  // the iterations before the call's last, colour 0's sweeps fused
            const int m = n_iter - 1 - it;
            rc = lu_inline_iterations(ctx, m);
            if (rc) return rc;
            it += m;
            continue;
        }
        rc = lu_iteration(ctx, it == n_iter - 1);  // norms of the call's last iteration (a monitor)
        if (rc) return rc;
        ctx->shadow_ok[ctx->inew] = false;
        std::swap(ctx->iold, ctx->inew);
        it++;
    }
    if (n_iter > 0) ctx->dt_valid = false;
    return check_bad(ctx);
}

int mst_lusgs_colours(mst_ctx *ctx, int32_t *colour, int64_t n_cells, int *n_colours) {
    CTX_GUARD(ctx);
    if (!colour || !n_colours || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "size does not match the mesh");
    int rc = build_colours(ctx);
    if (rc) return rc;
    const auto &new_of_old = ctx->prep.P.cell_new_of_old;
    for (int64_t c = 0; c < ctx->nc; c++) colour[c] = ctx->lu_colour[new_of_old[c]];
    *n_colours = (int)ctx->lu_start.size() - 1;
    return MST_OK;
}

int mst_compute_dt(mst_ctx *ctx, double cfl, double *dt) {
    CTX_GUARD(ctx);
    if (!(cfl > 0) || !dt) return mst::fail(MST_E_ARG, "cfl must be positive");
    int rc = run_cfl(ctx, ctx->Q[ctx->iold], cfl);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(dt, ctx->scal, sizeof(double), hipMemcpyDeviceToHost));
    ctx->dt_valid = cfl == ctx->cfg.cfl;
    if (!(*dt > 0)) return mst::fail(MST_E_STATE, "non-physical state: no positive CFL time step");
    return MST_OK;
}

int mst_residual_norms(mst_ctx *ctx, double *out) {
    CTX_GUARD(ctx);
    if (!out) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, ctx->scal + 1, sizeof(double) * ctx->nv, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_monitor_maxrel(mst_ctx *ctx, double *out) {
    CTX_GUARD(ctx);
    if (!out) return mst::fail(MST_E_ARG, "null argument");
    int rc = ctx->nv == 4 ? launch(ctx, -1, k_maxrel<4>, ctx->nc, ctx->nc, (const double *)ctx->Q[ctx->inew],
                                   (const double *)ctx->Q[ctx->iold], ctx->part)
                          : launch(ctx, -1, k_maxrel<5>, ctx->nc, ctx->nc, (const double *)ctx->Q[ctx->inew],
                                   (const double *)ctx->Q[ctx->iold], ctx->part);
    if (rc) return rc;
    FinArgs f;
    memset(&f, 0, sizeof(f));
    f.nblk = ctx->nblk_c;
    f.count = 1.0;
    for (int k = 0; k < ctx->nv; k++) {
        f.src[k] = ctx->part + k * ctx->nblk_c;
        f.kind[k] = 2;
        f.dst[k] = ctx->scal + 6 + k;
    }
    f.n = ctx->nv;
    rc = launch_finish(ctx, f);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, ctx->scal + 6, sizeof(double) * ctx->nv, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_monitor_maxrel_noalloc(const double *qnew, const double *qold, int64_t n_cells, int nv, double *out,
                               void *stream) {
    if (!qnew || !qold || !out) return mst::fail(MST_E_ARG, "null device pointer");
    if (nv != 4 && nv != 5) return mst::fail(MST_E_ARG, "nv must be 4 (2D) or 5 (3D)");
    if (n_cells < 1 || n_cells > MST_NOALLOC_MAX_CELLS)
        return mst::fail(MST_E_ARG, "n_cells must be in [1, MST_NOALLOC_MAX_CELLS]");
    if (nv == 4 && (((uintptr_t)qnew | (uintptr_t)qold) & 15u))
        return mst::fail(MST_E_ARG, "2D states must be 16-byte aligned (records are loaded as 16-byte pairs)");
    double *part = nullptr;
    int rc = noalloc_scratch_ptr(&part);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t grid = std::min<int64_t>((n_cells + BLOCK - 1) / BLOCK, NOALLOC_GRID);
    if (nv == 4)
        hipLaunchKernelGGL(k_maxrel<4>, dim3((uint32_t)grid), dim3(BLOCK), 0, st, n_cells, qnew, qold, part);
    else
        hipLaunchKernelGGL(k_maxrel<5>, dim3((uint32_t)grid), dim3(BLOCK), 0, st, n_cells, qnew, qold, part);
    HIPCHK(hipGetLastError());
    FinArgs f;
    memset(&f, 0, sizeof(f));
    f.nblk = grid;  // <= 2 * FIN_CHUNK: one finish block
    f.chunk = grid;
    f.count = 1.0;
    f.n = nv;
    for (int k = 0; k < nv; k++) {
        f.src[k] = part + k * grid;
        f.kind[k] = 2;
        f.dst[k] = out + k;
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN), 0, st, f);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

int mst_ctx_warm_start(mst_ctx *ctx, mst_mlp *mlp, const double *cond, int n_cond) {
    CTX_GUARD(ctx);
    if (!mlp || n_cond < 1 || n_cond > 8 || !cond) return mst::fail(MST_E_ARG, "bad arguments");
    int n_in = 0, n_out = 0;
    mst::mlp_dims(mlp, &n_in, &n_out);
    if (n_in != ctx->dim + n_cond || n_out != ctx->nv)
        return mst::fail(MST_E_ARG, "network sizes do not match (dim + conditions) -> nv");
    if (!(cond[0] > 0)) return mst::fail(MST_E_ARG, "the first condition (freestream Mach number) must be positive");
    WarmArgs w;
    w.nc = ctx->nc;
    w.cxc = ctx->cxc;
    w.n_cond = n_cond;
    for (int k = 0; k < 8; k++) w.cond[k] = k < n_cond ? (float)cond[k] : 0.f;
    w.mach = cond[0];
    w.gamma = ctx->cfg.gamma;
    w.Q = ctx->Q[ctx->iold];
    w.bad = ctx->bad;
    float *feat = nullptr, *out = nullptr;
    HIPCHK(hipMalloc((void **)&feat, sizeof(float) * ctx->nc * n_in));
    if (hipMalloc((void **)&out, sizeof(float) * ctx->nc * n_out) != hipSuccess) {
        (void)hipFree(feat);
        return mst::fail(MST_E_HIP, "hipMalloc failed");
    }
    w.feat = feat;
    w.Y = out;
    int rc = ctx->dim == 2 ? launch(ctx, -1, k_warm_features<2>, ctx->nc, w) : launch(ctx, -1, k_warm_features<3>, ctx->nc, w);
    if (!rc) rc = mst::mlp_forward_device(mlp, feat, ctx->nc, out, ctx->stream);
    if (!rc)
        rc = ctx->dim == 2 ? launch(ctx, -1, k_warm_decode<2>, ctx->nc, w) : launch(ctx, -1, k_warm_decode<3>, ctx->nc, w);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(feat);
    (void)hipFree(out);
    if (rc) return rc;
    if (e != hipSuccess) return mst::fail(MST_E_HIP, hipGetErrorString(e));
    for (int i = 0; i < 4; i++) ctx->shadow_ok[i] = false;
    ctx->dt_valid = false;
    return check_bad(ctx);
}

int mst_get_residual(mst_ctx *ctx, double *R, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (!R || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "size does not match the mesh");
    if (!ctx->R) {
        int rc = dalloc(&ctx->R, ctx->nc * ctx->nv);
        if (rc) return rc;
    }
    int rc = run_face_flux(ctx, ctx->iold);
    if (rc) return rc;
    rc = run_update(ctx, UPD_RESIDUAL, -1, -1, Stage{0, 0, 0}, -1);
    if (rc) return rc;
    return fetch_cells(ctx, ctx->R, ctx->nv, R);
}

int mst_get_face_flux(mst_ctx *ctx, double *G, int64_t n_faces) {
    CTX_GUARD(ctx);
    if (!G || n_faces != ctx->nf) return mst::fail(MST_E_ARG, "size does not match the mesh");
    const int nv = ctx->nv;
    std::vector<double> buf(ctx->nf * nv);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(buf.data(), ctx->G, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    const auto &new_of_old = ctx->prep.P.face_new_of_old;
    const auto &merged = ctx->prep.P.fmerged_into;
    for (int64_t f = 0; f < ctx->nf; f++) {
        const int64_t i = new_of_old[f];
        for (int k = 0; k < nv; k++)  // the merged half of a periodic pair carries minus its pair's flux
            G[f * nv + k] = merged[i] >= 0 ? -buf[(int64_t)merged[i] * nv + k] : buf[i * nv + k];
    }
    return MST_OK;
}

int mst_get_gradient(mst_ctx *ctx, double *grad, double *phi, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (n_cells != ctx->nc) return mst::fail(MST_E_ARG, "size does not match the mesh");
    if (!ctx->graw) {
        int rc = dalloc(&ctx->graw, ctx->nc * ctx->nv * ctx->dim);
        if (rc) return rc;
        rc = dalloc(&ctx->phi, ctx->nc * ctx->nv);
        if (rc) return rc;
    }
    int rc = run_grad(ctx, ctx->iold, true);
    if (rc) return rc;
    if (grad) {
        rc = fetch_cells(ctx, ctx->graw, ctx->nv * ctx->dim, grad);
        if (rc) return rc;
    }
    if (phi) return fetch_cells(ctx, ctx->phi, ctx->nv, phi);
    return MST_OK;
}

int mst_riemann_flux(const mst_config *cfg, int dim, int64_t n, const double *QL, const double *QR,
                     const int32_t *dir, double *F) {
    int rc = check_config(cfg);
    if (rc) return rc;
    if ((dim != 2 && dim != 3) || n < 0 || (n > 0 && (!QL || !QR || !dir || !F)))
        return mst::fail(MST_E_ARG, "bad arguments");
    for (int64_t i = 0; i < n; i++)
        if (dir[i] < 0 || dir[i] >= dim) return mst::fail(MST_E_ARG, "direction out of range");
    if (n == 0) return MST_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mst::fail(MST_E_HIP, "no HIP device available");
    const int nv = dim + 2;
    double *dL = nullptr, *dR = nullptr, *dF = nullptr;
    int32_t *dd = nullptr;
    HIPCHK(hipMalloc((void **)&dL, sizeof(double) * n * nv));
    HIPCHK(hipMalloc((void **)&dR, sizeof(double) * n * nv));
    HIPCHK(hipMalloc((void **)&dF, sizeof(double) * n * nv));
    HIPCHK(hipMalloc((void **)&dd, sizeof(int32_t) * n));
    HIPCHK(hipMemcpy(dL, QL, sizeof(double) * n * nv, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dR, QR, sizeof(double) * n * nv, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dd, dir, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    Params<double> P = make_params<double>(*cfg);
    const uint32_t grid = (uint32_t)((n + BLOCK - 1) / BLOCK);
    if (dim == 2) {
        if (cfg->scheme == MST_SCHEME_ROE)
            hipLaunchKernelGGL((k_riemann<2, MST_SCHEME_ROE>), dim3(grid), dim3(BLOCK), 0, 0, n, dL, dR, dd, dF, P);
        else
            hipLaunchKernelGGL((k_riemann<2, MST_SCHEME_AUSM>), dim3(grid), dim3(BLOCK), 0, 0, n, dL, dR, dd, dF, P);
    } else {
        if (cfg->scheme == MST_SCHEME_ROE)
            hipLaunchKernelGGL((k_riemann<3, MST_SCHEME_ROE>), dim3(grid), dim3(BLOCK), 0, 0, n, dL, dR, dd, dF, P);
        else
            hipLaunchKernelGGL((k_riemann<3, MST_SCHEME_AUSM>), dim3(grid), dim3(BLOCK), 0, 0, n, dL, dR, dd, dF, P);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(F, dF, sizeof(double) * n * nv, hipMemcpyDeviceToHost);
    (void)hipFree(dL);
    (void)hipFree(dR);
    (void)hipFree(dF);
    (void)hipFree(dd);
    if (e != hipSuccess) return mst::fail(MST_E_HIP, std::string("riemann kernel: ") + hipGetErrorString(e));
    return MST_OK;
}

int mst_copy_probe(int64_t nbytes, double *ms) {
    if (nbytes < 16 || nbytes % 16) return mst::fail(MST_E_ARG, "nbytes must be a positive multiple of 16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mst::fail(MST_E_HIP, "no HIP device available");
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, nbytes));
    HIPCHK(hipMalloc((void **)&b, nbytes));
    HIPCHK(hipMemset(a, 0, nbytes));
    const int64_t n2 = nbytes / 16;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)((n2 + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, 0, n2,
                       (const double2 *)a, b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return MST_OK;
}

int mst_timing_enable(mst_ctx *ctx, int on) {
    CTX_GUARD(ctx);
    if (on < 0) return mst::fail(MST_E_ARG, "timing: on must be >= 0");
    ctx->timing = on != 0;
    ctx->timing_stride = on > 1 ? on : 1;
    for (int k = 0; k < MST_K_COUNT; k++) ctx->tick[k] = 0;
    return MST_OK;
}

int mst_timing_get(mst_ctx *ctx, int kernel, int64_t *count, double *total_ms) {
    CTX_GUARD(ctx);
    if (kernel < 0 || kernel >= MST_K_COUNT || !count || !total_ms) return mst::fail(MST_E_ARG, "bad kernel id");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double tot = 0;
    for (auto &pr : ctx->ev[kernel]) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
    }
    *count = (int64_t)ctx->ev[kernel].size();
    *total_ms = tot;
    return MST_OK;
}

int mst_timing_reset(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < MST_K_COUNT; k++) {
        for (auto &pr : ctx->ev[k]) {
            ctx->pool.push_back(pr.first);
            ctx->pool.push_back(pr.second);
        }
        ctx->ev[k].clear();
    }
    return MST_OK;
}

int mst_sync(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MST_OK;
}

// Compulsory HBM traffic of one launch: every array the kernel must touch, once.  r is the size of
// the arithmetic real of the gradient/flux kernels (8 FP64, 4 mixed: FP32 state shadow, normals,
// offsets, limited gradients); face fluxes, the stage states, volumes and centres stay FP64.
int mst_algorithmic_bytes(const mst_ctx *ctx, int kernel, double *bytes) {
    CTX_GUARD(ctx);
    if (!bytes) return mst::fail(MST_E_ARG, "null argument");
    const double nc = (double)ctx->nc, nf = (double)ctx->nf, D = ctx->dim, NV = ctx->nv;
    const bool mixed = ctx->mixed, second = ctx->cfg.order >= 2;
    const double r = mixed ? 4 : 8;
    const double ell = 4.0 * ctx->maxf * nc;
    const double w_inc = mixed && ctx->dim == 3 ? 0 : r;  // incidence weight bytes of the gradient
    const double stages = ctx->cfg.rk;
    // faces: c0, c1, info, FP64 normal; cells: the arithmetic state; order 2: face-cell offsets
    // (FP64: face centres, partner centres and cell centres) and limited gradients
    // order 2: the side records (each read once), the record indices, the raw c0 state of boundary faces
    const double ninc = (double)ctx->n_inc, nbnd = (double)(ctx->nf - ctx->prep.P.n_interior);
    double flux_in = nf * (12 + 8 * D) + nc * r * NV;
    if (second) flux_in = nf * (12 + 8 * D + 8) + ninc * r * NV + nbnd * r * NV;
    // update: cell faces, stage state, volume, new state (+ FP32 shadow), Q^n on stages after the
    // first (mixed form: pre-combined stage states, Q^n only on RK3's middle stage), the CFL half-sums
    // on the last stage when the time step is reduced there
    const double qn_stages = mixed && !ctx->viscous ? std::max(stages - 2, 0.0) : stages - 1;
    double upd = ell + nc * (8 * NV + 8 + 8 * NV) + nc * 8 * NV * qn_stages / stages;
    if (mixed) upd += nc * 4 * NV;
    if (ctx->cfg.cfl > 0) upd += nc * 8 * D / stages;
    switch (kernel) {
        case MST_K_FLUX:
            *bytes = flux_in + nf * 8 * NV;
            break;
        case MST_K_GRAD:  // incidences: code, weight (not FP32 3D: folded into the normal), normal, offset,
                          // record out; cells: state, volume
            *bytes = 4.0 * ctx->slots * nc + ninc * (w_inc + 2 * r * D + r * NV) + nc * (r * NV + r);
            break;
        case MST_K_UPDATE:
            *bytes = upd + nf * 8 * NV;
            break;
        case MST_K_CELL:  // flux inputs + update, no face-flux array; FP64 reads the stage state once
            *bytes = flux_in + upd - (mixed ? 0 : (second ? nbnd : nc) * 8 * NV);
            break;
        case MST_K_PATCH:  // as the cell kernel, with the patch face lists and local slots for the table
            // (a face's kind comes with its packed list entry when there is one)
            *bytes = flux_in + upd - (mixed ? 0 : (second ? nbnd : nc) * 8 * NV) - ell +
                     4.0 * (double)ctx->n_psolve + 2.0 * ctx->slots * nc - (ctx->pfi ? 4.0 * nf : 0.0);
            break;
        case MST_K_FUSED: {  // gradient inputs + per-face inputs + update, each once; no records
            // per incidence: code, weight, offset, and in 3D the normal (2D takes it from its face's)
            const double inc_geo = ctx->dim == 2 ? r + r * D : w_inc + 2 * r * D;
            const double grad_in = 4.0 * ctx->slots * nc + ninc * inc_geo + nc * (r * NV + r);
            const double flux_f = nf * ((ctx->pfi ? 0 : 4) + 8 * D) + 4.0 * (double)ctx->n_psolve + nbnd * 4;
            *bytes = grad_in + flux_f + upd - ell + 2.0 * ctx->slots * nc - (mixed ? 0 : nc * 8 * NV);
            break;
        }
        case MST_K_DT:
            *bytes = nc * (8 * NV + 8 + 8 * D);
            break;
        case MST_K_VISC: {  // the gradient kernel and the face kernel, averaged over the two launches
            const double nw = D + 1;
            const double grad = 4.0 * ctx->slots * nc + ninc * (8 + 8 * D) + nc * (8 * NV + 8 + 8 * nw * D);
            const double face = nf * (8 + 8 + 8 * D + 2 * 8 * NV) + nc * (8 * NV + 8 * nw * D);
            *bytes = 0.5 * (grad + face);
            break;
        }
        case MST_K_LUSGS: {  // one iteration's sweeps, each array once per sweep, over their launches
            if (ctx->lu_start.size() < 2) return mst::fail(MST_E_ARG, "no LU-SGS iteration has run");
            // 2 K - 1 launches for K colours; 2 K - 2 where colour 0's sweeps of consecutive
            // iterations run fused (the iterations of a call before its last)
            const double K = (double)(ctx->lu_start.size() - 1);
            const double S = ctx->slots, launches = (ctx->lu_inline && K >= 2) ? 2.0 * K - 2 : 2.0 * K - 1;
            // forward: positions, residual, states, neighbour updates, update out, D out; per slot the
            // record (neighbour code, normal) and r out.  backward: positions, D, states, updates in
            // and out, new state; per slot the record and r.
            double fwd = nc * (4 + 8 * NV * 4 + 8) + S * nc * (4 + 8 * D + 8);
            // forward sweeps that compute the residual read each slot's face code and face info
            // instead of the residual array
            if (ctx->lu_inline) fwd += S * nc * 8 - nc * 8 * NV;
            const double bwd = nc * (4 + 8 + 8 * NV * 4) + S * nc * (4 + 8 * D + 8);
            *bytes = (fwd + bwd) / launches;
            break;
        }
        default:
            return mst::fail(MST_E_ARG, "bad kernel id");
    }
    return MST_OK;
}

}  // extern "C"
