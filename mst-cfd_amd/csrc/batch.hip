// batch.hip -- many independent cases advanced together (no reference counterpart: the reference
// runs one mesh per process, work/Work.cpp:53-89).
//
// Layout: the cases' prepared meshes are concatenated into one set of SoA arrays (global cell and
// face indices); per-case parameters (scheme, gamma, inlet state, CFL number) and the per-case time
// step live in small device arrays.  A step is a fixed sequence of persistent kernels (grid = CU
// count x resident blocks) that pull (case, tile) work items from a device counter: gradient +
// limiter over cell items, fluxes over face items (Roe or AUSM chosen per item's case), gather +
// stage update over cell items with per-item partial reductions, then one block per case reduces
// its items' partials into the case's residual norms and next time step.  A whole RK step is
// captured once into a hipGraph per state-buffer parity and replayed.  Tile sizes (per case-size
// bucket) and the block size come from a measured search whose results are cached in a JSON file.
//
// Counters: each persistent kernel pulls from its own counter and, at its start, zeroes the other
// counters (the kernels before it in the stream have finished), so a replayed graph needs no
// memset nodes.  Results do not depend on the tiling: every cell and face is computed by the same
// code as the single-case context, and the time step is a minimum (order-independent).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "lusgs.hpp"
#include "mstcfd.h"
#include "prep.hpp"

using namespace mstk;

namespace {

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess)                                                                            \
            return mst::fail(MST_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

struct Item {
    int32_t cs, begin, end, pad;
};

constexpr int NCNT = 4;
enum { CNT_GRAD = 0, CNT_FLUX = 1, CNT_UPD = 2, CNT_CFL = 3 };
constexpr int FIN_B = 256;

template <typename T>
struct BArgs {
    GradArgs<T> g;
    FluxArgs<T> f;
    UpdArgs u;
    const Params<T> *P;
    const int32_t *scheme;
    const double *dts, *gam, *cfl;
    const Item *items;
    int n_items, own;
    int *cnt;
    double *part_l2;  // [field][item]
    double *part_dt;  // [item]
};

// next work item of this block (one atomic per block, broadcast through LDS)
__device__ __forceinline__ int pull(int *cnt, int own) {
    __shared__ int s_item;
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(cnt + own, 1);
    __syncthreads();
    return s_item;
}

__device__ __forceinline__ void reset_others(int *cnt, int own) {
    if (blockIdx.x == 0 && (int)threadIdx.x < NCNT && (int)threadIdx.x != own) cnt[threadIdx.x] = 0;
}

template <int DIM, int MAXF, int LIM, typename T, int NT>
__global__ __launch_bounds__(NT) void kb_grad(const BArgs<T> a) {
    reset_others(a.cnt, a.own);
    for (;;) {
        const int it = pull(a.cnt, a.own);
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        for (int64_t c = I.begin + (int)threadIdx.x; c < I.end; c += NT) grad_cell<DIM, MAXF, LIM, T>(a.g, c);
    }
}

template <int DIM, bool SECOND, typename T, int NT>
__global__ __launch_bounds__(NT) void kb_flux(const BArgs<T> a) {
    constexpr int NV = DIM + 2;
    reset_others(a.cnt, a.own);
    for (;;) {
        const int it = pull(a.cnt, a.own);
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        FluxArgs<T> fa = a.f;
        fa.P = a.P[I.cs];
        const bool roe = a.scheme[I.cs] == MST_SCHEME_ROE;
        for (int64_t f = I.begin + (int)threadIdx.x; f < I.end; f += NT) {
            const int info = fa.finfo[f];
            const int kind = info & 15, mask = info >> 4;
            double g[NV];
            if (kind >= mst::KIND_NONE) {
#pragma unroll
                for (int k = 0; k < NV; k++) g[k] = 0;
            } else if (roe) {
                face_flux<DIM, MST_SCHEME_ROE, SECOND, T>(fa, f, kind, mask, g);
            } else {
                face_flux<DIM, MST_SCHEME_AUSM, SECOND, T>(fa, f, kind, mask, g);
            }
            stq<NV>(fa.G + f * NV, g);
        }
    }
}

template <int NV, int MAXF, typename T, int NT>
__global__ __launch_bounds__(NT) void kb_update(const BArgs<T> a) {
    constexpr bool F32 = std::is_same<T, float>::value;
    reset_others(a.cnt, a.own);
    const bool L2 = a.u.l2 != 0, DT = a.u.dtred != 0;
    for (;;) {
        const int it = pull(a.cnt, a.own);
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        UpdArgs ua = a.u;
        ua.bad = a.u.bad + I.cs;  // per-case flag
        ua.dt = a.dts + I.cs;
        ua.gamma = a.gam[I.cs];
        ua.cfl = a.cfl[I.cs];
        double red[NV + 1];
#pragma unroll
        for (int k = 0; k < NV; k++) red[k] = 0;
        red[NV] = INFINITY;
        for (int64_t c = I.begin + (int)threadIdx.x; c < I.end; c += NT) {
            double r[NV], cmp[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) r[k] = cmp[k] = 0;
#pragma unroll
            for (int s = 0; s < MAXF; s++) {
                const int32_t e = ua.ell[(int64_t)s * ua.nc + c];
                if (e < 0) break;
                double gf[NV];
                ldq<NV>(ua.G + (int64_t)(e >> 1) * NV, gf);
                accumulate<NV, F32>(r, cmp, gf, e & 1);
            }
            if constexpr (F32) {
#pragma unroll
                for (int k = 0; k < NV; k++) r[k] += cmp[k];
            }
            stage_cell<NV, F32>(ua, c, r, red, L2, DT);
        }
        if (L2) {
            double v[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) v[k] = red[k];
            block_partial<NV, false, NT>(v, a.part_l2 + it, a.n_items);
        }
        if (DT) {
            double v[1] = {red[NV]};
            block_partial<1, true, NT>(v, a.part_dt + it, 1);
        }
    }
}

// cell strategy: each lane solves its cell's faces (scheme per case), sums and updates the stage
template <int DIM, bool SECOND, typename T, int NT, int MAXF>
__global__ __launch_bounds__(NT) void kb_cell(const BArgs<T> a) {
    constexpr int NV = DIM + 2;
    constexpr bool F32 = std::is_same<T, float>::value;
    reset_others(a.cnt, a.own);
    const bool L2 = a.u.l2 != 0, DT = a.u.dtred != 0;
    for (;;) {
        const int it = pull(a.cnt, a.own);
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        FluxArgs<T> fa = a.f;
        fa.P = a.P[I.cs];
        const bool roe = a.scheme[I.cs] == MST_SCHEME_ROE;
        UpdArgs ua = a.u;
        ua.bad = a.u.bad + I.cs;
        ua.dt = a.dts + I.cs;
        ua.gamma = a.gam[I.cs];
        ua.cfl = a.cfl[I.cs];
        double red[NV + 1];
#pragma unroll
        for (int k = 0; k < NV; k++) red[k] = 0;
        red[NV] = INFINITY;
        for (int64_t c = I.begin + (int)threadIdx.x; c < I.end; c += NT) {
            double r[NV], cmp[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) r[k] = cmp[k] = 0;
            cell_fluxes<DIM, SCHEME_RUNTIME, SECOND, T, MAXF>(fa, ua.ell, ua.nc, roe, c, r, cmp);
            if constexpr (F32) {
#pragma unroll
                for (int k = 0; k < NV; k++) r[k] += cmp[k];
            }
            stage_cell<NV, F32>(ua, c, r, red, L2, DT);
        }
        if (L2) {
            double v[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) v[k] = red[k];
            block_partial<NV, false, NT>(v, a.part_l2 + it, a.n_items);
        }
        if (DT) {
            double v[1] = {red[NV]};
            block_partial<1, true, NT>(v, a.part_dt + it, 1);
        }
    }
}

// patch strategy: one item is one patch (BLOCK consecutive cells of a case, a compact patch in the
// Hilbert order): its faces solved once into LDS (scheme per case), then sums in slot order and the
// stage update per cell, as the single-case patch kernel does
template <int DIM, bool SECOND, typename T, int MAXF>
__global__ __launch_bounds__(BLOCK) void kb_patch(const BArgs<T> a, const PatchArgs p) {
    constexpr int NV = DIM + 2;
    constexpr bool F32 = std::is_same<T, float>::value;
    extern __shared__ double gsh[];  // NV x patch faces, component-major
    reset_others(a.cnt, a.own);
    const bool L2 = a.u.l2 != 0, DT = a.u.dtred != 0;
    for (;;) {
        const int it = pull(a.cnt, a.own);  // (its barrier also ends the previous item's LDS reads)
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        FluxArgs<T> fa = a.f;
        fa.P = a.P[I.cs];
        const bool roe = a.scheme[I.cs] == MST_SCHEME_ROE;
        const int32_t f0 = p.start[I.pad];
        const int npf = p.start[I.pad + 1] - f0;
        for (int i = threadIdx.x; i < npf; i += BLOCK) {
            int64_t f;
            int info;
            patch_face(p, fa.finfo, f0 + i, f, info);
            const int kind = info & 15;
            double n64[DIM];
            T X[NV], Y[NV];
            face_load<DIM, SECOND, T>(fa, f, kind, n64, X, Y);
            double g[NV];
            face_solve<DIM, SCHEME_RUNTIME, T>(fa.P, roe, kind, info >> 4, n64, X, Y, g);
#pragma unroll
            for (int k = 0; k < NV; k++) gsh[k * npf + i] = g[k];
        }
        __syncthreads();
        UpdArgs ua = a.u;
        ua.bad = a.u.bad + I.cs;
        ua.dt = a.dts + I.cs;
        ua.gamma = a.gam[I.cs];
        ua.cfl = a.cfl[I.cs];
        double red[NV + 1];
#pragma unroll
        for (int k = 0; k < NV; k++) red[k] = 0;
        red[NV] = INFINITY;
        const int64_t c = I.begin + (int)threadIdx.x;
        if (c < I.end) {
            int cl[MAXF];
#pragma unroll
            for (int s = 0; s < MAXF; s++) cl[s] = p.loc[(int64_t)s * ua.nc + c];
            double r[NV];
            patch_sum<NV, MAXF, F32>(gsh, npf, cl, r);
            stage_cell<NV, F32>(ua, c, r, red, L2, DT);
        }
        if (L2) {
            double v[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) v[k] = red[k];
            block_partial<NV, false, BLOCK>(v, a.part_l2 + it, a.n_items);
        }
        if (DT) {
            double v[1] = {red[NV]};
            block_partial<1, true, BLOCK>(v, a.part_dt + it, 1);
        }
    }
}

// fused strategy (order 2): one item is one patch, whose cells' and halo's gradients, face states and
// fluxes stay in the block (fused_patch), then sums in slot order and the stage update per cell
// FP32 triangles: 5 resident waves per SIMD (at most 96 VGPRs, 4 registers spilled); elsewhere the
// compiler chooses.
template <int DIM, typename T, int MAXF>
constexpr int kb_fused_waves() {
    return (sizeof(T) == 4 && DIM == 2 && MAXF == 3) ? 5 : 1;
}

template <int DIM, int LIM, typename T, int MAXF>
__global__ __launch_bounds__(BLOCK, (kb_fused_waves<DIM, T, MAXF>())) void kb_fused(const BArgs<T> a, const PatchArgs p,
                                                                                     const HaloArgs h) {
    constexpr int NV = DIM + 2;
    constexpr bool F32 = std::is_same<T, float>::value;
    extern __shared__ double sh[];
    reset_others(a.cnt, a.own);
    const bool L2 = a.u.l2 != 0, DT = a.u.dtred != 0;
    for (;;) {
        const int it = pull(a.cnt, a.own);  // (its barrier also ends the previous item's LDS reads)
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        FluxArgs<T> fa = a.f;
        fa.P = a.P[I.cs];
        const bool roe = a.scheme[I.cs] == MST_SCHEME_ROE;
        const int32_t f0 = p.start[I.pad];
        FusedLds<T> L(sh, p.start[I.pad + 1] - f0, NV);
        int cl[MAXF];
        // (no early fetch of the update's inputs here: with both schemes inlined it would cost a wave)
        fused_patch<DIM, SCHEME_RUNTIME, LIM, T, MAXF, fused_lds_normals<DIM>()>(a.g, fa, roe, p, h, I.pad, I.begin,
                                                                                I.end, a.u.nc, f0, L, cl, []() {});
        UpdArgs ua = a.u;
        ua.bad = a.u.bad + I.cs;
        ua.dt = a.dts + I.cs;
        ua.gamma = a.gam[I.cs];
        ua.cfl = a.cfl[I.cs];
        double red[NV + 1];
#pragma unroll
        for (int k = 0; k < NV; k++) red[k] = 0;
        red[NV] = INFINITY;
        const int64_t c = I.begin + (int)threadIdx.x;
        if (c < I.end) {
            double r[NV];
            patch_sum<NV, MAXF, F32>(L.g, L.npf, cl, r);
            stage_cell<NV, F32>(ua, c, r, red, L2, DT);
        }
        if (L2) {
            double v[NV];
#pragma unroll
            for (int k = 0; k < NV; k++) v[k] = red[k];
            block_partial<NV, false, BLOCK>(v, a.part_l2 + it, a.n_items);
        }
        if (DT) {
            double v[1] = {red[NV]};
            block_partial<1, true, BLOCK>(v, a.part_dt + it, 1);
        }
    }
}

template <int DIM, typename T, int NT>
__global__ __launch_bounds__(NT) void kb_cfl(const BArgs<T> a, const double *__restrict__ Q) {
    constexpr int NV = DIM + 2;
    reset_others(a.cnt, a.own);
    for (;;) {
        const int it = pull(a.cnt, a.own);
        if (it >= a.n_items) break;
        const Item I = a.items[it];
        const double gam = a.gam[I.cs], cfl = a.cfl[I.cs];
        double v[1] = {INFINITY};
        for (int64_t c = I.begin + (int)threadIdx.x; c < I.end; c += NT) {
            double q[NV];
            ldq<NV>(Q + c * NV, q);
            v[0] = fmin(v[0], local_dt<DIM>(q, a.u.sh + c * DIM, a.u.vol[c], gam, cfl));
        }
        block_partial<1, true, NT>(v, a.part_dt + it, 1);
    }
}

// one block per case: its items' partials in item order -> L2 norms of R/V and the minimum time step
struct BFin {
    int *cnt;                  // work counters, zeroed here (the finish ends every step)
    const int32_t *ci0, *ci1;  // cell-item range of each case (empty: inactive)
    const double *part_l2, *part_dt;
    const double *ncells;
    double *l2, *dts;
    int n_items, nv, want_l2, want_dt;
};

__global__ __launch_bounds__(FIN_B) void kb_finish(const BFin a) {
    if (blockIdx.x == 0 && (int)threadIdx.x < NCNT) a.cnt[threadIdx.x] = 0;
    const int cs = blockIdx.x;
    const int i0 = a.ci0[cs], i1 = a.ci1[cs];
    if (i0 >= i1) return;
    __shared__ double s[6][FIN_B];
    double acc[6];
    for (int k = 0; k < 5; k++) acc[k] = 0;
    acc[5] = INFINITY;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += FIN_B) {
        if (a.want_l2)
            for (int k = 0; k < a.nv; k++) acc[k] += a.part_l2[(int64_t)k * a.n_items + i];
        if (a.want_dt) acc[5] = fmin(acc[5], a.part_dt[i]);
    }
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = FIN_B / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 5; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
            s[5][threadIdx.x] = fmin(s[5][threadIdx.x], s[5][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (a.want_l2)
            for (int k = 0; k < a.nv; k++) a.l2[cs * a.nv + k] = sqrt(s[k][0] / a.ncells[cs]);
        if (a.want_dt) a.dts[cs] = s[5][0];
    }
}

__global__ __launch_bounds__(BLOCK) void kb_shadow(int64_t n, const double *__restrict__ Q, float *__restrict__ Q32) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) Q32[i] = (float)Q[i];
}

// LU-SGS (implicit steady driver) over the whole batch: every active case's faces in one launch (the
// case's scheme and parameters per face), then the colour sweeps of lusgs.hpp over the cases' merged
// colour lists, gathering each cell's residual from these fluxes.
template <int DIM, bool SECOND>
__global__ __launch_bounds__(BLOCK) void kb_lu_flux(const FluxArgs<double> f0, const Params<double> *__restrict__ P,
                                                    const int32_t *__restrict__ scheme,
                                                    const int32_t *__restrict__ fcase,
                                                    const uint8_t *__restrict__ active) {
    constexpr int NV = DIM + 2;
    const int64_t f = block_id() * BLOCK + threadIdx.x;
    if (f >= f0.nf) return;
    const int32_t cs = fcase[f];
    if (!active[cs]) return;
    FluxArgs<double> fa = f0;
    fa.P = P[cs];
    const int info = fa.finfo[f];
    const int kind = info & 15, mask = info >> 4;
    double g[NV];
    if (kind >= mst::KIND_NONE) {
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] = 0;
    } else if (scheme[cs] == MST_SCHEME_ROE) {
        face_flux<DIM, MST_SCHEME_ROE, SECOND, double>(fa, f, kind, mask, g);
    } else {
        face_flux<DIM, MST_SCHEME_AUSM, SECOND, double>(fa, f, kind, mask, g);
    }
    stq<NV>(fa.G + f * NV, g);
}

// per active case (one block each): L2 over its cells of R/V, R gathered from G in slot order
template <int NV, int MAXF>
__global__ __launch_bounds__(BLOCK) void kb_lu_norms(int64_t NC, const int32_t *__restrict__ ell,
                                                     const double *__restrict__ G, const double *__restrict__ vol,
                                                     const int64_t *__restrict__ cstart,
                                                     const uint8_t *__restrict__ active, double *__restrict__ l2) {
    const int cs = blockIdx.x;
    if (!active[cs]) return;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0;
    for (int64_t c = cstart[cs] + threadIdx.x; c < cstart[cs + 1]; c += BLOCK) {
        double r[NV], cmp[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) r[k] = cmp[k] = 0;
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            const int32_t e = ell[(int64_t)s * NC + c];
            if (e < 0) break;
            double gf[NV];
            ldq<NV>(G + (int64_t)(e >> 1) * NV, gf);
            accumulate<NV, false>(r, cmp, gf, e & 1);
        }
        const double iV = frcp(vol[c]);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const double x = r[k] * iV;
            acc[k] += x * x;
        }
    }
    __shared__ double sh[NV];
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = acc[k];
    block_partial<NV, false>(v, sh, 1);
    __syncthreads();
    if ((int)threadIdx.x < NV) l2[cs * NV + threadIdx.x] = sqrt(sh[threadIdx.x] / (double)(cstart[cs + 1] - cstart[cs]));
}

int bucket_of(int64_t nc) {
    if (nc < 2048) return 0;
    if (nc < 8192) return 1;
    if (nc < 32768) return 2;
    return 3;
}

}  // namespace

struct mst_batch {
    int n = 0, dim = 2, nv = 4, maxf = 3, slots = 3, order = 1, limiter = 0, rk = 1;
    bool mixed = false;
    std::vector<mst_config> cfgs;
    std::vector<mst::Prepared> preps;
    std::vector<int64_t> cstart, fstart;
    int64_t NC = 0, NF = 0;
    hipStream_t stream = nullptr;
    int ncu = 0;
    int32_t *fc0 = nullptr, *fc1 = nullptr, *finfo = nullptr, *ell = nullptr;
    double *fn = nullptr, *fxf = nullptr, *feta = nullptr, *cvol = nullptr, *cxc = nullptr, *cshalf = nullptr;
    double *Q[4] = {nullptr, nullptr, nullptr, nullptr};
    double *G = nullptr, *rec = nullptr;  // face fluxes; reconstructed face states (order 2)
    int32_t *frec0 = nullptr, *frec1 = nullptr;  // incidence records of each face's c0 / c1 side
    float *Q32[4] = {nullptr, nullptr, nullptr, nullptr};
    float *rec32 = nullptr;
    int32_t *inc_code = nullptr;  // gradient incidence layout (slot-major, stride NC)
    double *inc_w = nullptr, *inc_n = nullptr, *inc_dx = nullptr;
    float *inc_w32 = nullptr, *inc_n32 = nullptr, *inc_dx32 = nullptr;
    float *ivol32 = nullptr;  // (float) 1/V, the FP32 gradient's volume input
    Params<double> *P64 = nullptr;
    Params<float> *P32 = nullptr;
    int32_t *scheme = nullptr;
    double *dts = nullptr, *gam = nullptr, *cfl = nullptr, *l2 = nullptr, *ncells = nullptr;
    // work lists
    Item *citems = nullptr, *fitems = nullptr;
    int n_citems = 0, n_fitems = 0, cap_citems = 0, cap_fitems = 0;
    int32_t *ci0 = nullptr, *ci1 = nullptr;
    double *part = nullptr;
    int *cnt = nullptr, *bad = nullptr;
    int iold = 0, inew = 1;
    bool dt_valid = false;
    bool shadow_ok[4] = {false, false, false, false};
    int block = 256;
    int tiles[MST_BATCH_BUCKETS] = {1024, 2048, 4096, 8192};
    int strategy = MST_STRATEGY_CELL;  // FACE: flux + gather/update kernels; CELL: fused cell kernel;
                                       // PATCH: patch kernel over BLOCK-cell items
    // patch tables (global indices): per patch its flux-carrying faces, per cell slot the local index
    int32_t *pstart = nullptr, *pface = nullptr;
    uint32_t *pfi = nullptr;  // pface with the faces' kinds and L/R masks packed in (null: 2^26 faces or more)
    int16_t *ploc = nullptr;
    std::vector<int32_t> patch_base;  // first patch of each case
    int max_pf = 0;
    bool patch_ok = false;
    // fused strategy: each patch's halo cells and their slots' local faces
    int32_t *hstart = nullptr, *hcell = nullptr;
    int16_t *hloc = nullptr;
    bool fused_ok = false;
    HaloGeo hg;  // the halo cells' incidence geometry in halo order
    std::vector<uint8_t> active;
    bool use_graph = true;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    bool graphs_valid = false;
    // LU-SGS steady driver (built on first use): positions sorted by colour across the cases, per
    // (slot, position) records, per position its case, per face its case, device active flags
    int steady = MST_STEADY_EXPLICIT;
    double lu_cfl = 1e4, lu_omega = 1.0;
    int32_t *lu_cells = nullptr, *lu_nb = nullptr, *lu_face = nullptr, *lu_pcase = nullptr, *lu_fcase = nullptr;
    double *lu_n = nullptr, *lu_r = nullptr, *lu_D = nullptr, *lu_dQ = nullptr;
    int64_t *lu_cstart = nullptr;
    uint8_t *lu_active = nullptr;
    std::vector<int64_t> lu_start;
};

namespace {

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

void drop_graphs(mst_batch *b) {
    for (auto &g : b->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    b->graphs_valid = false;
}

void release(mst_batch *b) {
    if (!b) return;
    drop_graphs(b);
    void *bufs[] = {b->fc0, b->fc1, b->finfo, b->ell, b->fn, b->fxf, b->feta, b->cvol, b->cxc, b->cshalf,
                    b->Q[0], b->Q[1], b->Q[2], b->Q[3], b->G, b->rec, b->Q32[0], b->Q32[1], b->Q32[2], b->Q32[3],
                    b->frec0, b->frec1, b->rec32, b->P64, b->P32, b->scheme, b->dts, b->gam,
                    b->cfl, b->l2, b->ncells, b->citems, b->fitems, b->ci0, b->ci1, b->part, b->cnt, b->bad,
                    b->inc_code, b->inc_w, b->inc_n, b->inc_dx, b->inc_w32, b->inc_n32, b->inc_dx32, b->ivol32,
                    b->pstart, b->pface, b->pfi, b->ploc, b->hstart, b->hcell, b->hloc, b->lu_cells, b->lu_nb,
                    b->lu_face, b->lu_pcase, b->lu_fcase, b->lu_n, b->lu_r, b->lu_D, b->lu_dQ, b->lu_cstart,
                    b->lu_active};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    b->hg.release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Patch tables of every case (global indices): each case's cells in runs of BLOCK (its patches), per
// patch the faces that carry a flux into one of its cells (ascending), per cell slot the face's local
// index in its patch.  The patch strategy needs at most 4 slots and each patch's fluxes (NV doubles
// per face) within 64 KiB of LDS.
int build_patches(mst_batch *b) {
    b->patch_ok = b->fused_ok = false;
    b->patch_base.assign(b->n + 1, 0);
    if (b->slots > 4) return MST_OK;
    std::vector<int32_t> start, face, list, hstart, hcell, halo;
    std::vector<int16_t> loc((size_t)b->slots * b->NC, -1), hloc;
    size_t max_pf = 0;
    for (int i = 0; i < b->n; i++) {
        const mst::Prepared &P = b->preps[i];
        const int64_t cs = b->cstart[i], fs = b->fstart[i];
        b->patch_base[i] = (int32_t)start.size();
        for (int64_t ca = 0; ca < P.nc; ca += BLOCK) {
            const int64_t cz = std::min<int64_t>(P.nc, ca + BLOCK);
            list.clear();
            for (int s = 0; s < P.maxf; s++)
                for (int64_t c = ca; c < cz; c++) {
                    const int32_t e = P.ell[(int64_t)s * P.nc + c];
                    if (e >= 0 && P.fkind[e >> 1] < mst::KIND_NONE) list.push_back(e >> 1);
                }
            std::sort(list.begin(), list.end());
            list.erase(std::unique(list.begin(), list.end()), list.end());
            max_pf = std::max(max_pf, list.size());
            for (int s = 0; s < P.maxf; s++)
                for (int64_t c = ca; c < cz; c++) {
                    const int32_t e = P.ell[(int64_t)s * P.nc + c];
                    if (e < 0 || P.fkind[e >> 1] >= mst::KIND_NONE) continue;
                    const int64_t l = std::lower_bound(list.begin(), list.end(), e >> 1) - list.begin();
                    loc[(size_t)s * b->NC + cs + c] = (int16_t)((l << 1) | (e & 1));
                }
            start.push_back((int32_t)face.size());
            for (const int32_t f : list) face.push_back((int32_t)(f + fs));
            // halo: the other cell of every patch face with one cell outside the patch, and for each of
            // its slots the local index of the slot's face when the face is one of the patch's
            halo.clear();
            for (const int32_t f : list) {
                if (P.fkind[f] != mst::KIND_INTERIOR && P.fkind[f] != mst::KIND_PERIODIC) continue;
                for (const int32_t c : {P.fc0[f], P.fc1[f]})
                    if (c < ca || c >= cz) halo.push_back(c);
            }
            std::sort(halo.begin(), halo.end());
            halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
            hstart.push_back((int32_t)hcell.size());
            for (const int32_t c : halo) {
                hcell.push_back((int32_t)(c + cs));
                for (int s = 0; s < b->slots; s++) {
                    const int32_t e = s < P.maxf ? P.ell[(int64_t)s * P.nc + c] : -1;
                    int16_t v = -1;
                    if (e >= 0 && P.fkind[e >> 1] < mst::KIND_NONE) {
                        const auto it = std::lower_bound(list.begin(), list.end(), e >> 1);
                        if (it != list.end() && *it == (e >> 1)) v = (int16_t)(((it - list.begin()) << 1) | (e & 1));
                    }
                    hloc.push_back(v);
                }
            }
        }
    }
    b->patch_base[b->n] = (int32_t)start.size();
    start.push_back((int32_t)face.size());
    hstart.push_back((int32_t)hcell.size());
    if (max_pf * b->nv * sizeof(double) > 65536) return MST_OK;
    int rc = dupload(&b->pstart, start);
    if (!rc) rc = dupload(&b->pface, face);
    if (!rc && b->NF < ((int64_t)1 << FI_SHIFT) && !face.empty()) {  // finfo is final for a batch
        rc = dalloc(&b->pfi, face.size());
        if (!rc) {
            const int64_t n = (int64_t)face.size();
            hipLaunchKernelGGL(k_pack_faces<0>, dim3((uint32_t)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, b->stream, n,
                               (const int32_t *)b->pface, (const int32_t *)b->finfo, b->pfi);
            HIPCHK(hipGetLastError());
        }
    }
    if (!rc) rc = dupload(&b->ploc, loc);
    if (!rc) rc = dupload(&b->hstart, hstart);
    if (!rc) rc = dupload(&b->hcell, hcell);
    if (!rc) rc = dupload(&b->hloc, hloc);
    if (rc) return rc;
    b->max_pf = (int)max_pf;
    b->patch_ok = true;
    // the fused kernel keeps states and fluxes (2D: and the faces' normals) in LDS
    b->fused_ok = fused_lds_bytes(b->dim, b->nv, (int)max_pf, false, b->dim == 2) <= 65536;
    if (b->fused_ok && !hcell.empty()) {
        const IncView<double> f64{b->inc_code, b->inc_w, b->inc_n, b->inc_dx, b->cvol, b->NC};
        const IncView<float> f32{b->mixed ? b->inc_code : nullptr, b->inc_w32, b->inc_n32, b->inc_dx32, b->ivol32, b->NC};
        const int64_t nh = (int64_t)hcell.size();
        HIPCHK(b->dim == 2 ? b->hg.fill<2>(b->slots, b->NC, nh, b->hcell, f64, f32, b->stream)
                          : b->hg.fill<3>(b->slots, b->NC, nh, b->hcell, f64, f32, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return MST_OK;
}

// work items of the active cases for the current tiles; grows the device lists as needed
int build_items(mst_batch *b) {
    std::vector<Item> ci, fi;
    std::vector<int32_t> c0(b->n, 0), c1(b->n, 0);
    for (int i = 0; i < b->n; i++) {
        const int64_t nc = b->cstart[i + 1] - b->cstart[i], nf = b->fstart[i + 1] - b->fstart[i];
        c0[i] = c1[i] = (int32_t)ci.size();
        if (!b->active[i]) continue;
        if (b->strategy == MST_STRATEGY_PATCH || b->strategy == MST_STRATEGY_FUSED) {  // one item per patch
            for (int64_t s = 0, k = 0; s < nc; s += BLOCK, k++)
                ci.push_back({i, (int32_t)(b->cstart[i] + s), (int32_t)(b->cstart[i] + std::min<int64_t>(nc, s + BLOCK)),
                              (int32_t)(b->patch_base[i] + k)});
            c1[i] = (int32_t)ci.size();
            continue;
        }
        const int64_t tc = b->tiles[bucket_of(nc)];
        const int64_t tf = std::max<int64_t>(64, ((tc * nf / std::max<int64_t>(nc, 1)) + 63) / 64 * 64);
        for (int64_t s = 0; s < nc; s += tc)
            ci.push_back({i, (int32_t)(b->cstart[i] + s), (int32_t)(b->cstart[i] + std::min(nc, s + tc)), 0});
        for (int64_t s = 0; s < nf; s += tf)
            fi.push_back({i, (int32_t)(b->fstart[i] + s), (int32_t)(b->fstart[i] + std::min(nf, s + tf)), 0});
        c1[i] = (int32_t)ci.size();
    }
    if ((int)ci.size() > b->cap_citems) {
        if (b->citems) (void)hipFree(b->citems);
        if (b->part) (void)hipFree(b->part);
        b->citems = nullptr;
        b->part = nullptr;
        b->cap_citems = (int)ci.size();
        int rc = dalloc(&b->citems, ci.size());
        if (!rc) rc = dalloc(&b->part, ci.size() * (b->nv + 1));
        if (rc) return rc;
    }
    if ((int)fi.size() > b->cap_fitems) {
        if (b->fitems) (void)hipFree(b->fitems);
        b->fitems = nullptr;
        b->cap_fitems = (int)fi.size();
        int rc = dalloc(&b->fitems, fi.size());
        if (rc) return rc;
    }
    b->n_citems = (int)ci.size();
    b->n_fitems = (int)fi.size();
    HIPCHK(hipStreamSynchronize(b->stream));
    if (!ci.empty()) HIPCHK(hipMemcpy(b->citems, ci.data(), sizeof(Item) * ci.size(), hipMemcpyHostToDevice));
    if (!fi.empty()) HIPCHK(hipMemcpy(b->fitems, fi.data(), sizeof(Item) * fi.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->ci0, c0.data(), sizeof(int32_t) * b->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->ci1, c1.data(), sizeof(int32_t) * b->n, hipMemcpyHostToDevice));
    drop_graphs(b);
    return MST_OK;
}

template <typename T>
BArgs<T> base_args(const mst_batch *b, int isrc) {
    BArgs<T> a;
    memset(&a, 0, sizeof(a));
    GradArgs<T> &g = a.g;
    g.nc = b->NC;
    g.inc_code = b->inc_code;
    g.graw = nullptr;
    g.phi = nullptr;
    FluxArgs<T> &f = a.f;
    f.nf = b->NF;
    f.fc0 = b->fc0;
    f.fc1 = b->fc1;
    f.finfo = b->finfo;
    f.fn = b->fn;
    f.frec0 = b->frec0;
    f.frec1 = b->frec1;
    f.G = b->G;
    if constexpr (std::is_same<T, float>::value) {
        g.inc_w = b->inc_w32;
        g.inc_n = b->inc_n32;
        g.inc_dx = b->inc_dx32;
        g.Q = b->Q32[isrc];
        g.rec = b->rec32;
        g.vol = b->ivol32;
        f.Q = b->Q32[isrc];
        f.rec = b->rec32;
        a.P = b->P32;
    } else {
        g.inc_w = b->inc_w;
        g.inc_n = b->inc_n;
        g.inc_dx = b->inc_dx;
        g.Q = b->Q[isrc];
        g.rec = b->rec;
        g.vol = b->cvol;
        f.Q = b->Q[isrc];
        f.rec = b->rec;
        a.P = b->P64;
    }
    UpdArgs &u = a.u;
    u.nc = b->NC;
    u.ell = b->ell;
    u.G = b->G;
    u.vol = b->cvol;
    u.sh = b->cshalf;
    u.maxf = b->maxf;
    u.bad = b->bad;
    a.scheme = b->scheme;
    a.dts = b->dts;
    a.gam = b->gam;
    a.cfl = b->cfl;
    a.cnt = b->cnt;
    a.part_l2 = b->part;
    a.part_dt = b->part + (int64_t)b->nv * std::max(b->n_citems, 1);
    return a;
}

template <typename K>
int persistent_grid(const mst_batch *b, K kernel, int nt, int n_items, int *grid, size_t shm = 0) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, shm));
    *grid = std::max(1, std::min(n_items, std::max(per_cu, 1) * b->ncu));
    return MST_OK;
}

template <typename K, typename... A>
int launch_persistent_shm(mst_batch *b, K kernel, int nt, size_t shm, int n_items, const A &...args) {
    if (n_items <= 0) return MST_OK;
    int grid = 0;
    int rc = persistent_grid(b, kernel, nt, n_items, &grid, shm);
    if (rc) return rc;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(nt), (uint32_t)shm, b->stream, args...);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

template <typename K, typename... A>
int launch_persistent(mst_batch *b, K kernel, int nt, int n_items, const A &...args) {
    return launch_persistent_shm(b, kernel, nt, 0, n_items, args...);
}

template <int DIM, typename T, int NT>
int run_grad(mst_batch *b, int isrc) {
    BArgs<T> a = base_args<T>(b, isrc);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = CNT_GRAD;
#define MST_BG(MF)                                                                                    \
    if (b->limiter == MST_LIM_NONE) return launch_persistent(b, kb_grad<DIM, MF, MST_LIM_NONE, T, NT>, NT, a.n_items, a); \
    if (b->limiter == MST_LIM_MINMOD) return launch_persistent(b, kb_grad<DIM, MF, MST_LIM_MINMOD, T, NT>, NT, a.n_items, a); \
    return launch_persistent(b, kb_grad<DIM, MF, MST_LIM_VANLEER, T, NT>, NT, a.n_items, a);
    if (b->slots == 3) { MST_BG(3) }
    if (b->slots == 4) { MST_BG(4) }
    MST_BG(8)
#undef MST_BG
}

template <int DIM, typename T, int NT>
int run_flux(mst_batch *b, int isrc) {
    BArgs<T> a = base_args<T>(b, isrc);
    a.items = b->fitems;
    a.n_items = b->n_fitems;
    a.own = CNT_FLUX;
    if (b->order >= 2) return launch_persistent(b, kb_flux<DIM, true, T, NT>, NT, a.n_items, a);
    return launch_persistent(b, kb_flux<DIM, false, T, NT>, NT, a.n_items, a);
}

template <int DIM, typename T, int NT>
int run_update(mst_batch *b, int is, int in, const Stage &sg, int iout, bool l2, bool dt,
               bool check) {
    constexpr int NV = DIM + 2;
    BArgs<T> a = base_args<T>(b, is);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = CNT_UPD;
    a.u.Qs = b->Q[is];
    a.u.Qn = b->Q[in];
    a.u.Qout = b->Q[iout];
    a.u.Qout32 = b->mixed ? b->Q32[iout] : nullptr;
    set_stage(a.u, sg);
    a.u.l2 = l2;
    a.u.dtred = dt;
    a.u.check = check;
    if (b->slots == 3) return launch_persistent(b, kb_update<NV, 3, T, NT>, NT, a.n_items, a);
    if (b->slots == 4) return launch_persistent(b, kb_update<NV, 4, T, NT>, NT, a.n_items, a);
    return launch_persistent(b, kb_update<NV, 8, T, NT>, NT, a.n_items, a);
}

// stage s of a step pulls from counter UPD (even s) or FLUX (odd s): without a gradient kernel between
// them, consecutive cell kernels must not share a counter (each zeroes only the others at its start)
template <int DIM, typename T, int NT>
int run_cell(mst_batch *b, int stage, int is, int in, const Stage &sg, int iout, bool l2, bool dt,
             bool check) {
    BArgs<T> a = base_args<T>(b, is);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = (stage & 1) ? CNT_FLUX : CNT_UPD;
    a.u.Qs = b->Q[is];
    a.u.Qn = b->Q[in];
    a.u.Qout = b->Q[iout];
    a.u.Qout32 = b->mixed ? b->Q32[iout] : nullptr;
    set_stage(a.u, sg);
    a.u.l2 = l2;
    a.u.dtred = dt;
    a.u.check = check;
#define MST_BCELL(MF)                                                                                          \
    if (b->order >= 2) return launch_persistent(b, kb_cell<DIM, true, T, NT, MF>, NT, a.n_items, a);          \
    return launch_persistent(b, kb_cell<DIM, false, T, NT, MF>, NT, a.n_items, a);
    if (b->slots == 3) { MST_BCELL(3) }
    if (b->slots == 4) { MST_BCELL(4) }
    MST_BCELL(8)
#undef MST_BCELL
}

template <int DIM, typename T>
int run_patch(mst_batch *b, int stage, int is, int in, const Stage &sg, int iout, bool l2, bool dt,
              bool check) {
    BArgs<T> a = base_args<T>(b, is);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = (stage & 1) ? CNT_FLUX : CNT_UPD;
    a.u.Qs = b->Q[is];
    a.u.Qn = b->Q[in];
    a.u.Qout = b->Q[iout];
    a.u.Qout32 = b->mixed ? b->Q32[iout] : nullptr;
    set_stage(a.u, sg);
    a.u.l2 = l2;
    a.u.dtred = dt;
    a.u.check = check;
    PatchArgs pa;
    pa.start = b->pstart;
    pa.face = b->pface;
    pa.loc = b->ploc;
    pa.fi = b->pfi;
    const size_t shm = (size_t)b->max_pf * b->nv * sizeof(double);
#define MST_BPATCH(MF)                                                                                        \
    if (b->order >= 2) return launch_persistent_shm(b, kb_patch<DIM, true, T, MF>, BLOCK, shm, a.n_items, a, pa); \
    return launch_persistent_shm(b, kb_patch<DIM, false, T, MF>, BLOCK, shm, a.n_items, a, pa);
    if (b->slots == 3) { MST_BPATCH(3) }
    MST_BPATCH(4)
#undef MST_BPATCH
}

template <int DIM, typename T>
int run_fused(mst_batch *b, int stage, int is, int in, const Stage &sg, int iout, bool l2, bool dt,
              bool check) {
    BArgs<T> a = base_args<T>(b, is);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = (stage & 1) ? CNT_FLUX : CNT_UPD;
    a.u.Qs = b->Q[is];
    a.u.Qn = b->Q[in];
    a.u.Qout = b->Q[iout];
    a.u.Qout32 = b->mixed ? b->Q32[iout] : nullptr;
    set_stage(a.u, sg);
    a.u.l2 = l2;
    a.u.dtred = dt;
    a.u.check = check;
    PatchArgs pa;
    pa.start = b->pstart;
    pa.face = b->pface;
    pa.loc = b->ploc;
    pa.fi = b->pfi;
    HaloArgs ha;
    ha.start = b->hstart;
    ha.cell = b->hcell;
    ha.loc = b->hloc;
    b->hg.args<T>(ha);
    const size_t shm = fused_lds_bytes(DIM, b->nv, b->max_pf, std::is_same<T, float>::value, fused_lds_normals<DIM>());
#define MST_BFUSED(LM, MF) return launch_persistent_shm(b, kb_fused<DIM, LM, T, MF>, BLOCK, shm, a.n_items, a, pa, ha);
#define MST_BFUSED_L(MF)                                          \
    if (b->limiter == MST_LIM_NONE) { MST_BFUSED(MST_LIM_NONE, MF) } \
    if (b->limiter == MST_LIM_MINMOD) { MST_BFUSED(MST_LIM_MINMOD, MF) } \
    MST_BFUSED(MST_LIM_VANLEER, MF)
    if (b->slots == 3) { MST_BFUSED_L(3) }
    MST_BFUSED_L(4)
#undef MST_BFUSED_L
#undef MST_BFUSED
}

int run_finish(mst_batch *b, bool l2, bool dt) {
    BFin f;
    f.cnt = b->cnt;
    f.ci0 = b->ci0;
    f.ci1 = b->ci1;
    f.part_l2 = b->part;
    f.part_dt = b->part + (int64_t)b->nv * std::max(b->n_citems, 1);
    f.ncells = b->ncells;
    f.l2 = b->l2;
    f.dts = b->dts;
    f.n_items = b->n_citems;
    f.nv = b->nv;
    f.want_l2 = l2;
    f.want_dt = dt;
    hipLaunchKernelGGL(kb_finish, dim3(b->n), dim3(FIN_B), 0, b->stream, f);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

template <int DIM, typename T, int NT>
int run_cfl_t(mst_batch *b) {
    BArgs<T> a = base_args<T>(b, b->iold);
    a.items = b->citems;
    a.n_items = b->n_citems;
    a.own = CNT_CFL;
    int rc = launch_persistent(b, kb_cfl<DIM, T, NT>, NT, a.n_items, a, (const double *)b->Q[b->iold]);
    if (rc) return rc;
    return run_finish(b, false, true);
}

// one RK step from buffer iold to inew with the spare buffers between stages
template <int DIM, typename T, int NT>
int enqueue_step_t(mst_batch *b, int iold, int inew) {
    static const double A2[2] = {0, 0.5}, B2[2] = {1, 0.5};
    static const double A3[3] = {0, 0.75, 1.0 / 3.0}, B3[3] = {1, 0.25, 2.0 / 3.0};
    const int ns = b->rk;
    const double *A = ns == 3 ? A3 : A2, *B = ns == 3 ? B3 : B2;
    int spare[2], w = 0;
    for (int i = 0; i < 4; i++)
        if (i != iold && i != inew) spare[w++] = i;
    // mixed form: a stage that has a successor stores the successor's combination A Q^n + B Q_s in place
    // of its FP64 state (the FP32 kernels read only the shadow), so no later stage reads Q^n again
    // (SSP-RK2: none; RK3: the middle stage)
    const bool pre = b->mixed;
    int src = iold;
    for (int s = 0; s < ns; s++) {
        const int dst = (s == ns - 1) ? inew : spare[s & 1];
        Stage sg;
        sg.ca = A[s];
        sg.cb = B[s];
        sg.stage0 = s == 0;
        sg.pre = pre;
        sg.wa = pre && s < ns - 1 ? A[s + 1] : 0;
        sg.wb = pre && s < ns - 1 ? B[s + 1] : 0;
        int rc;
        if (b->strategy == MST_STRATEGY_FUSED && b->order >= 2) {
            rc = run_fused<DIM, T>(b, s, src, iold, sg, dst, s == 0, s == ns - 1, s == ns - 1);
            if (rc) return rc;
            src = dst;
            continue;
        }
        if (b->order >= 2) {
            rc = run_grad<DIM, T, NT>(b, src);
            if (rc) return rc;
        }
        if (b->strategy == MST_STRATEGY_PATCH || b->strategy == MST_STRATEGY_FUSED) {  // FUSED at order 1: PATCH
            rc = run_patch<DIM, T>(b, s, src, iold, sg, dst, s == 0, s == ns - 1, s == ns - 1);
            if (rc) return rc;
        } else if (b->strategy == MST_STRATEGY_CELL) {
            rc = run_cell<DIM, T, NT>(b, s, src, iold, sg, dst, s == 0, s == ns - 1, s == ns - 1);
            if (rc) return rc;
        } else {
            rc = run_flux<DIM, T, NT>(b, src);
            if (rc) return rc;
            rc = run_update<DIM, T, NT>(b, src, iold, sg, dst, s == 0, s == ns - 1, s == ns - 1);
            if (rc) return rc;
        }
        src = dst;
    }
    return run_finish(b, true, true);
}

template <int DIM, typename T>
int enqueue_step_nt(mst_batch *b, int iold, int inew) {
    if (b->strategy == MST_STRATEGY_PATCH || b->strategy == MST_STRATEGY_FUSED)
        return enqueue_step_t<DIM, T, BLOCK>(b, iold, inew);  // a patch is BLOCK cells
    if (b->block == 64) return enqueue_step_t<DIM, T, 64>(b, iold, inew);
    if (b->block == 128) return enqueue_step_t<DIM, T, 128>(b, iold, inew);
    return enqueue_step_t<DIM, T, 256>(b, iold, inew);
}

int enqueue_step(mst_batch *b, int iold, int inew) {
    if (b->mixed) return b->dim == 2 ? enqueue_step_nt<2, float>(b, iold, inew) : enqueue_step_nt<3, float>(b, iold, inew);
    return b->dim == 2 ? enqueue_step_nt<2, double>(b, iold, inew) : enqueue_step_nt<3, double>(b, iold, inew);
}

int run_cfl(mst_batch *b) {
    if (b->dim == 2) return b->mixed ? run_cfl_t<2, float, 256>(b) : run_cfl_t<2, double, 256>(b);
    return b->mixed ? run_cfl_t<3, float, 256>(b) : run_cfl_t<3, double, 256>(b);
}

int ensure_shadow(mst_batch *b, int i) {
    if (!b->mixed || b->shadow_ok[i]) return MST_OK;
    const int64_t n = b->NC * b->nv;
    hipLaunchKernelGGL(kb_shadow, dim3((uint32_t)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, b->stream, n,
                       (const double *)b->Q[i], b->Q32[i]);
    HIPCHK(hipGetLastError());
    b->shadow_ok[i] = true;
    return MST_OK;
}

int build_graphs(mst_batch *b) {
    drop_graphs(b);
    for (int par = 0; par < 2; par++) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(b, par, 1 - par);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return mst::fail(MST_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->gexec[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return mst::fail(MST_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    b->graphs_valid = true;
    return MST_OK;
}

// LU-SGS tables of the batch: each case's colouring (prep.cpp, lusgs_colours), the positions of
// colour k = every case's colour-k cells (case by case, ascending); records as the single context's
int build_lu(mst_batch *b) {
    if (b->lu_cells) return MST_OK;
    const int64_t NC = b->NC, S = b->slots;
    const int dim = b->dim;
    if (NC >= ((int64_t)1 << 29)) return mst::fail(MST_E_UNSUPPORTED, "LU-SGS records hold cell indices below 2^29");
    std::vector<int32_t> col(NC);
    int ncol = 0;
    for (int i = 0; i < b->n; i++) {
        std::vector<int32_t> ci;
        ncol = std::max(ncol, mst::lusgs_colours(b->preps[i], ci));
        for (size_t c = 0; c < ci.size(); c++) col[b->cstart[i] + c] = ci[c];
    }
    std::vector<int64_t> start(ncol + 1, 0);
    for (int64_t c = 0; c < NC; c++) start[col[c] + 1]++;
    for (int k = 0; k < ncol; k++) start[k + 1] += start[k];
    std::vector<int32_t> cells(NC), pcase(NC), fcase(b->NF);
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < b->n; i++)
        for (int64_t c = b->cstart[i]; c < b->cstart[i + 1]; c++) {
            const int64_t t = fill[col[c]]++;
            cells[t] = (int32_t)c;
            pcase[t] = i;
        }
    for (int i = 0; i < b->n; i++)
        for (int64_t f = b->fstart[i]; f < b->fstart[i + 1]; f++) fcase[f] = i;
    std::vector<int32_t> nb(S * NC, -2), fcode(S * NC, -1);
    std::vector<double> nrm(S * NC * dim, 0.0);
    for (int64_t t = 0; t < NC; t++) {
        const int i = pcase[t];
        const mst::Prepared &P = b->preps[i];
        const int64_t c0 = b->cstart[i], f0 = b->fstart[i];
        const int64_t lc = cells[t] - c0;
        for (int s = 0; s < P.maxf; s++) {
            const int32_t e = P.ell[(int64_t)s * P.nc + lc];
            if (e < 0) break;
            const int32_t f = e >> 1;
            const bool c1side = e & 1;
            const int32_t lo = c1side ? P.fc0[f] : P.fc1[f];
            const int64_t idx = (int64_t)s * NC + t;
            for (int d = 0; d < dim; d++)
                nrm[idx * dim + d] = c1side ? -P.fn[(int64_t)f * dim + d] : P.fn[(int64_t)f * dim + d];
            if (lo < 0) {
                nb[idx] = -1;
            } else {
                const int64_t o = c0 + lo;
                nb[idx] = (int32_t)((o << 2) | (col[o] < col[cells[t]] ? 1 : 0) | (col[o] > col[cells[t]] ? 2 : 0));
            }
            fcode[idx] = (int32_t)(((f0 + f) << 1) | (c1side ? 1 : 0));
        }
    }
    std::vector<int64_t> cst(b->cstart.begin(), b->cstart.end());
    int rc = dupload(&b->lu_cells, cells);
    if (rc == MST_OK) rc = dupload(&b->lu_nb, nb);
    if (rc == MST_OK) rc = dupload(&b->lu_face, fcode);
    if (rc == MST_OK) rc = dupload(&b->lu_pcase, pcase);
    if (rc == MST_OK) rc = dupload(&b->lu_fcase, fcase);
    if (rc == MST_OK) rc = dupload(&b->lu_n, nrm);
    if (rc == MST_OK) rc = dupload(&b->lu_cstart, cst);
    if (rc == MST_OK) rc = dalloc(&b->lu_r, (size_t)(S * NC));
    if (rc == MST_OK) rc = dalloc(&b->lu_D, (size_t)NC);
    if (rc == MST_OK) rc = dalloc(&b->lu_dQ, (size_t)(NC * b->nv));
    if (rc == MST_OK) rc = dalloc(&b->lu_active, (size_t)b->n);
    if (rc) return rc;
    b->lu_start = start;
    return MST_OK;
}

template <int DIM, int MAXF>
int lu_sweeps_b(mst_batch *b, const LuArgs &a) {
    const int ncol = (int)b->lu_start.size() - 1;
    for (int k = 0; k < ncol; k++) {
        const int64_t t0 = b->lu_start[k], n = b->lu_start[k + 1] - t0;
        if (n <= 0) continue;
        const uint32_t grid = (uint32_t)((n + BLOCK - 1) / BLOCK);
        if (k == ncol - 1)
            hipLaunchKernelGGL((k_lu_sweep<DIM, true, true, MAXF, LU_R_GATHER>), dim3(grid), dim3(BLOCK), 0, b->stream, a, t0, n);
        else
            hipLaunchKernelGGL((k_lu_sweep<DIM, true, false, MAXF, LU_R_GATHER>), dim3(grid), dim3(BLOCK), 0, b->stream, a, t0, n);
        HIPCHK(hipGetLastError());
    }
    for (int k = ncol - 2; k >= 0; k--) {
        const int64_t t0 = b->lu_start[k], n = b->lu_start[k + 1] - t0;
        if (n <= 0) continue;
        const uint32_t grid = (uint32_t)((n + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL((k_lu_sweep<DIM, false, false, MAXF>), dim3(grid), dim3(BLOCK), 0, b->stream, a, t0, n);
        HIPCHK(hipGetLastError());
    }
    return MST_OK;
}

// one LU-SGS iteration of every active case: gradients (order 2), fluxes, sweeps into New, swap
template <int DIM>
int lu_iteration_b(mst_batch *b, bool l2) {
    constexpr int NV = DIM + 2;
    const int iold = b->iold, inew = b->inew;
    if (b->order >= 2) {
        HIPCHK(hipMemsetAsync(b->cnt, 0, sizeof(int) * NCNT, b->stream));  // direct launches, no graph
        int rc = run_grad<DIM, double, BLOCK>(b, iold);
        if (rc) return rc;
    }
    BArgs<double> ba = base_args<double>(b, iold);
    const uint32_t fgrid = (uint32_t)((b->NF + BLOCK - 1) / BLOCK);
    if (b->order >= 2)
        hipLaunchKernelGGL((kb_lu_flux<DIM, true>), dim3(fgrid), dim3(BLOCK), 0, b->stream, ba.f, (const Params<double> *)b->P64,
                           (const int32_t *)b->scheme, (const int32_t *)b->lu_fcase, (const uint8_t *)b->lu_active);
    else
        hipLaunchKernelGGL((kb_lu_flux<DIM, false>), dim3(fgrid), dim3(BLOCK), 0, b->stream, ba.f, (const Params<double> *)b->P64,
                           (const int32_t *)b->scheme, (const int32_t *)b->lu_fcase, (const uint8_t *)b->lu_active);
    HIPCHK(hipGetLastError());
    if (l2) {
        if (b->slots == 3)
            hipLaunchKernelGGL((kb_lu_norms<NV, 3>), dim3(b->n), dim3(BLOCK), 0, b->stream, b->NC, (const int32_t *)b->ell,
                               (const double *)b->G, (const double *)b->cvol, (const int64_t *)b->lu_cstart,
                               (const uint8_t *)b->lu_active, b->l2);
        else
            hipLaunchKernelGGL((kb_lu_norms<NV, 4>), dim3(b->n), dim3(BLOCK), 0, b->stream, b->NC, (const int32_t *)b->ell,
                               (const double *)b->G, (const double *)b->cvol, (const int64_t *)b->lu_cstart,
                               (const uint8_t *)b->lu_active, b->l2);
        HIPCHK(hipGetLastError());
    }
    LuArgs a;
    memset(&a, 0, sizeof(a));
    a.nc = b->NC;
    a.cells = b->lu_cells;
    a.nb = b->lu_nb;
    a.n = b->lu_n;
    a.r = b->lu_r;
    a.D = b->lu_D;
    a.Q = b->Q[iold];
    a.dQ = b->lu_dQ;
    a.Qout = b->Q[inew];
    a.bad = b->bad;
    a.dscale = 1.0 / b->lu_cfl + 0.5 * b->lu_omega;
    a.omega = b->lu_omega;
    a.face = b->lu_face;
    a.pcase = b->lu_pcase;
    a.gam = b->gam;
    a.active = b->lu_active;
    a.G = b->G;
    return b->slots == 3 ? lu_sweeps_b<DIM, 3>(b, a) : lu_sweeps_b<DIM, 4>(b, a);
}

int check_bad(mst_batch *b) {
    std::vector<int> h(b->n, 0);
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(h.data(), b->bad, sizeof(int) * b->n, hipMemcpyDeviceToHost));
    std::string list;
    int count = 0;
    for (int i = 0; i < b->n; i++)
        if (h[i]) {
            if (count < 32) list += (count ? ", " : "") + std::to_string(i);
            count++;
        }
    if (count) {
        HIPCHK(hipMemset(b->bad, 0, sizeof(int) * b->n));
        return mst::fail(MST_E_STATE, "non-physical state (rho<=0, p<=0 or NaN, or no positive time step) produced by the update in " +
                                          std::to_string(count) + " case(s): " + list + (count > 32 ? ", ..." : ""));
    }
    return MST_OK;
}

// the steps proper: iold/inew alternate; a graph per parity or direct launches
int do_steps(mst_batch *b, int nsteps) {
    if (b->n_citems == 0) return MST_OK;
    int rc = ensure_shadow(b, b->iold);
    if (rc) return rc;
    if (!b->dt_valid) {
        rc = run_cfl(b);
        if (rc) return rc;
        b->dt_valid = true;
    }
    if (b->use_graph && !b->graphs_valid) {
        rc = build_graphs(b);
        if (rc) return rc;
    }
    for (int it = 0; it < nsteps; it++) {
        if (b->use_graph) {
            HIPCHK(hipGraphLaunch(b->gexec[b->iold], b->stream));
        } else {
            rc = enqueue_step(b, b->iold, b->inew);
            if (rc) return rc;
        }
        std::swap(b->iold, b->inew);
    }
    return MST_OK;
}

std::string signature(const mst_batch *b) {
    int hist[MST_BATCH_BUCKETS] = {0, 0, 0, 0};
    int64_t cells = 0;
    for (int i = 0; i < b->n; i++) {
        if (!b->active[i]) continue;
        const int64_t nc = b->cstart[i + 1] - b->cstart[i];
        hist[bucket_of(nc)]++;
        cells += nc;
    }
    std::ostringstream s;
    s << "dim" << b->dim << "-o" << b->order << "-l" << b->limiter << "-rk" << b->rk << "-p" << (b->mixed ? 1 : 0)
      << "-h" << hist[0] << "." << hist[1] << "." << hist[2] << "." << hist[3] << "-c" << (cells >> 16);
    return s.str();
}

// cache file: one JSON object, one entry per line
//   "key": {"block": B, "tiles": [t0, t1, t2, t3], "strategy": S, "ms": x}
bool cache_lookup(const std::string &path, const std::string &key, int *block, int *tiles, int *strategy, double *ms) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line, want = "\"" + key + "\"";
    while (std::getline(in, line)) {
        if (line.find(want) == std::string::npos) continue;
        const char *p = strstr(line.c_str(), "{\"block\":");
        if (!p) continue;
        if (sscanf(p, "{\"block\": %d, \"tiles\": [%d, %d, %d, %d], \"strategy\": %d, \"ms\": %lf}", block,
                   &tiles[0], &tiles[1], &tiles[2], &tiles[3], strategy, ms) == 7)
            return true;
    }
    return false;
}

void cache_store(const std::string &path, const std::string &key, int block, const int *tiles, int strategy,
                 double ms) {
    std::vector<std::string> lines;
    {
        std::ifstream in(path);
        std::string line, want = "\"" + key + "\"";
        while (in && std::getline(in, line))
            if (line.find("{\"block\":") != std::string::npos && line.find(want) == std::string::npos) {
                while (!line.empty() && (line.back() == ',' || line.back() == ' ')) line.pop_back();
                lines.push_back(line);
            }
    }
    char buf[256];
    snprintf(buf, sizeof(buf), "  \"%s\": {\"block\": %d, \"tiles\": [%d, %d, %d, %d], \"strategy\": %d, \"ms\": %.6f}",
             key.c_str(), block, tiles[0], tiles[1], tiles[2], tiles[3], strategy, ms);
    lines.push_back(buf);
    std::ofstream out(path, std::ios::trunc);
    out << "{\n";
    for (size_t i = 0; i < lines.size(); i++) out << lines[i] << (i + 1 < lines.size() ? ",\n" : "\n");
    out << "}\n";
}

// time nsteps steps with the current tiles/block (graph replay), after 2 warm-up steps
int time_steps(mst_batch *b, int nsteps, double *ms) {
    int rc = build_items(b);
    if (rc) return rc;
    rc = do_steps(b, 2);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, b->stream));
    rc = do_steps(b, nsteps);
    if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / nsteps;
    return MST_OK;
}

int check_cfg(const mst_config *c) {
    if (c->scheme != MST_SCHEME_ROE && c->scheme != MST_SCHEME_AUSM) return mst::fail(MST_E_ARG, "bad scheme");
    if (c->order != 1 && c->order != 2) return mst::fail(MST_E_ARG, "order must be 1 or 2");
    if (c->limiter < MST_LIM_NONE || c->limiter > MST_LIM_VANLEER) return mst::fail(MST_E_ARG, "bad limiter");
    if (c->rk < 1 || c->rk > 3) return mst::fail(MST_E_ARG, "rk must be 1, 2 or 3");
    if (c->precision != MST_PREC_FP64 && c->precision != MST_PREC_MIXED) return mst::fail(MST_E_ARG, "bad precision");
    if (!(c->gamma > 1) || !(c->entropy_delta > 0) || c->eor < 0 || !(c->cfl > 0))
        return mst::fail(MST_E_ARG, "bad gamma / entropy_delta / eor, or cfl not positive (batched cases "
                                    "advance with their own CFL time step)");
    if (!(c->q_inlet[0] > 0)) return mst::fail(MST_E_ARG, "inlet density must be positive");
    return MST_OK;
}

}  // namespace

#define B_GUARD(b) \
    if (!(b)) return mst::fail(MST_E_ARG, "null batch");

extern "C" {

int mst_batch_create(int n_cases, const mst_mesh *const *meshes, const mst_config *cfgs, mst_batch **out) {
    if (!out || n_cases <= 0 || !meshes || !cfgs) return mst::fail(MST_E_ARG, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n_cases; i++) {
        if (!meshes[i]) return mst::fail(MST_E_ARG, "null mesh");
        int rc = check_cfg(&cfgs[i]);
        if (rc) return rc;
        if (cfgs[i].mu != 0 || cfgs[i].kappa != 0)
            return mst::fail(MST_E_UNSUPPORTED, "batches run inviscid cases only (mu = kappa = 0)");
        const mst_config &a = cfgs[0], &c = cfgs[i];
        if (c.order != a.order || c.limiter != a.limiter || c.rk != a.rk || c.precision != a.precision ||
            meshes[i]->dim != meshes[0]->dim)
            return mst::fail(MST_E_ARG, "batched cases must share dim, order, limiter, rk and precision");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return mst::fail(MST_E_HIP, "no HIP device available: the batched path runs only on the GPU");
    mst_batch *b = new (std::nothrow) mst_batch();
    if (!b) return mst::fail(MST_E_ARG, "out of host memory");
    b->n = n_cases;
    b->cfgs.assign(cfgs, cfgs + n_cases);
    b->preps.resize(n_cases);
    b->dim = meshes[0]->dim;
    b->nv = b->dim + 2;
    b->order = cfgs[0].order;
    b->limiter = cfgs[0].limiter;
    b->rk = cfgs[0].rk;
    b->mixed = cfgs[0].precision == MST_PREC_MIXED;
    b->cstart.assign(n_cases + 1, 0);
    b->fstart.assign(n_cases + 1, 0);
    b->active.assign(n_cases, 1);
    b->maxf = 0;
    int rc = MST_OK;
    try {
        for (int i = 0; i < n_cases && !rc; i++) {
            rc = mst::prepare(*meshes[i], b->preps[i]);
            if (rc) break;
            const mst::Prepared &P = b->preps[i];
            if (P.bad_cells > 0) rc = mst::fail(MST_E_MESH, "case " + std::to_string(i) + ": degenerate cells");
            else if (P.pcount > 0) rc = mst::fail(MST_E_UNSUPPORTED, "periodic faces are not supported in batches");
            b->cstart[i + 1] = b->cstart[i] + P.nc;
            b->fstart[i + 1] = b->fstart[i] + P.nf;
            b->maxf = std::max(b->maxf, P.maxf);
        }
    } catch (const std::bad_alloc &) {
        rc = mst::fail(MST_E_ARG, "out of host memory");
    }
    if (!rc && (b->cstart[n_cases] >= INT32_MAX / 2 || b->fstart[n_cases] >= INT32_MAX / 2))
        rc = mst::fail(MST_E_UNSUPPORTED, "batch too large for 32-bit face indices");
    if (rc) {
        delete b;
        return rc;
    }
    b->NC = b->cstart[n_cases];
    b->NF = b->fstart[n_cases];
    b->slots = b->maxf <= 3 ? 3 : (b->maxf <= 4 ? 4 : 8);
    if ((int64_t)b->slots * b->NC >= (int64_t)INT32_MAX) {  // incidence records are indexed in 32 bits
        delete b;
        return mst::fail(MST_E_UNSUPPORTED, "batch too large: cells x face slots must stay below 2^31");
    }
    const int dim = b->dim, nv = b->nv;
    // concatenated arrays (global indices)
    std::vector<int32_t> fc0(b->NF), fc1(b->NF), finfo(b->NF), ell((size_t)b->slots * b->NC, -1), scheme(n_cases);
    std::vector<double> fn(b->NF * dim), fxf(b->NF * dim), feta(b->NF), cvol(b->NC), cxc(b->NC * dim),
        cshalf(b->NC * dim), gam(n_cases), cflv(n_cases), ncells(n_cases);
    std::vector<int32_t> inc_code((size_t)b->slots * b->NC, -2), frec0(b->NF, -1), frec1(b->NF, -1);
    std::vector<double> inc_w((size_t)b->slots * b->NC, 0.0), inc_n((size_t)b->slots * b->NC * dim, 0.0),
        inc_dx((size_t)b->slots * b->NC * dim, 0.0);
    std::vector<float> inc_w32(b->mixed ? (size_t)b->slots * b->NC : 0, 0.f);
    std::vector<float> inc_n32(b->mixed ? (size_t)b->slots * b->NC * dim : 0, 0.f);
    std::vector<Params<double>> p64(n_cases);
    std::vector<Params<float>> p32(n_cases);
    for (int i = 0; i < n_cases; i++) {
        const mst::Prepared &P = b->preps[i];
        const mst_config &c = cfgs[i];
        const int64_t cs = b->cstart[i], fs = b->fstart[i];
        std::vector<int32_t> mask;
        mst::lr_masks(P, c.compat, mask);
        for (int64_t f = 0; f < P.nf; f++) {
            fc0[fs + f] = (int32_t)(P.fc0[f] + cs);
            fc1[fs + f] = P.fc1[f] >= 0 ? (int32_t)(P.fc1[f] + cs) : -1;
            finfo[fs + f] = P.fkind[f] | (mask[f] << 4);
            feta[fs + f] = P.feta[f];
            for (int d = 0; d < dim; d++) {
                fn[(fs + f) * dim + d] = P.fn[f * dim + d];
                fxf[(fs + f) * dim + d] = P.fxf[f * dim + d];
            }
        }
        if ((c.compat & MST_COMPAT_GRAD_LASTINT) && P.quirk_face >= 0) feta[fs + P.quirk_face] = 1.0;
        for (int64_t k = 0; k < P.nc; k++) {
            cvol[cs + k] = P.cvol[k];
            for (int d = 0; d < dim; d++) {
                cxc[(cs + k) * dim + d] = P.cxc[k * dim + d];
                cshalf[(cs + k) * dim + d] = P.cshalf[k * dim + d];
            }
            for (int s = 0; s < P.maxf; s++) {
                const int32_t e = P.ell[(int64_t)s * P.nc + k];
                if (e >= 0) ell[(size_t)s * b->NC + cs + k] = (int32_t)((((int64_t)(e >> 1) + fs) << 1) | (e & 1));
            }
        }
        for (int sl = 0; sl < b->slots; sl++)
            for (int64_t k = 0; k < P.nc; k++) {
                const size_t gi = (size_t)sl * b->NC + cs + k;
                int32_t code;
                double nn[3], dd[3];
                const int64_t f = mst::incidence(P, sl, k, code, nn, dd);
                inc_code[gi] = code >= 0 ? (int32_t)((((int64_t)(code >> 1) + cs) << 1) | (code & 1)) : code;
                if (f < 0) continue;
                const int32_t ent = P.ell[(int64_t)sl * P.nc + k];
                ((ent & 1) ? frec1 : frec0)[fs + f] = (int32_t)gi;
                const double e = feta[fs + f];
                inc_w[gi] = e;
                for (int d = 0; d < dim; d++) {
                    inc_n[gi * dim + d] = nn[d];
                    inc_dx[gi * dim + d] = dd[d];
                }
                if (b->mixed) {
                    inc_w32[gi] = (code >= 0 && (code & 1)) ? (float)e : 1.0f - (float)e;
                    // the FP32 3D gradient's weighted normals (k_weighted_normals: the same arithmetic)
                    const double w = dim == 3 && code >= 0 ? ((code & 1) ? e : 1.0 - e) : 1.0;
                    for (int d = 0; d < dim; d++) inc_n32[gi * dim + d] = (float)(w * nn[d]);
                }
            }
        p64[i] = make_params<double>(c);
        p32[i] = make_params<float>(c);
        scheme[i] = c.scheme;
        gam[i] = c.gamma;
        cflv[i] = c.cfl;
        ncells[i] = (double)P.nc;
    }
#define TRY(e)            \
    do {                  \
        int _r = (e);     \
        if (_r) {         \
            release(b);   \
            return _r;    \
        }                 \
    } while (0)
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return mst::fail(MST_E_HIP, "hipStreamCreate failed");
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        release(b);
        return mst::fail(MST_E_HIP, "device query failed");
    }
    b->ncu = prop.multiProcessorCount;
    TRY(dupload(&b->fc0, fc0));
    TRY(dupload(&b->fc1, fc1));
    TRY(dupload(&b->finfo, finfo));
    TRY(dupload(&b->ell, ell));
    TRY(dupload(&b->fn, fn));
    TRY(dupload(&b->fxf, fxf));
    TRY(dupload(&b->feta, feta));
    TRY(dupload(&b->cvol, cvol));
    TRY(dupload(&b->cxc, cxc));
    TRY(dupload(&b->cshalf, cshalf));
    TRY(dupload(&b->inc_code, inc_code));
    TRY(dupload(&b->inc_w, inc_w));
    TRY(dupload(&b->inc_n, inc_n));
    TRY(dupload(&b->inc_dx, inc_dx));
    if (b->mixed) {
        std::vector<float> d32(inc_dx.begin(), inc_dx.end());
        TRY(dupload(&b->inc_w32, inc_w32));
        TRY(dupload(&b->inc_n32, inc_n32));
        TRY(dupload(&b->inc_dx32, d32));
        TRY(dalloc(&b->ivol32, b->NC));
        hipLaunchKernelGGL(k_recip_vol<0>, dim3((uint32_t)((b->NC + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, b->stream,
                           b->NC, (const double *)b->cvol, b->ivol32);
        if (hipGetLastError() != hipSuccess) {
            release(b);
            return mst::fail(MST_E_HIP, "k_recip_vol launch failed");
        }
    }
    TRY(dupload(&b->P64, p64));
    TRY(dupload(&b->P32, p32));
    TRY(dupload(&b->scheme, scheme));
    TRY(dupload(&b->gam, gam));
    TRY(dupload(&b->cfl, cflv));
    TRY(dupload(&b->ncells, ncells));
    TRY(dalloc(&b->dts, n_cases));
    TRY(dalloc(&b->l2, (size_t)n_cases * nv));
    TRY(dalloc(&b->ci0, n_cases));
    TRY(dalloc(&b->ci1, n_cases));
    for (int i = 0; i < 4; i++) TRY(dalloc(&b->Q[i], b->NC * nv));
    TRY(dalloc(&b->G, b->NF * nv));
    TRY(dalloc(&b->cnt, NCNT));
    TRY(dalloc(&b->bad, n_cases));
    TRY(dupload(&b->frec0, frec0));
    TRY(dupload(&b->frec1, frec1));
    if (b->order >= 2 && !b->mixed) TRY(dalloc(&b->rec, (size_t)b->slots * b->NC * nv));
    if (b->mixed) {
        for (int i = 0; i < 4; i++) TRY(dalloc(&b->Q32[i], b->NC * nv));
        if (b->order >= 2) TRY(dalloc(&b->rec32, (size_t)b->slots * b->NC * nv));
    }
    if (hipMemset(b->cnt, 0, sizeof(int) * NCNT) != hipSuccess || hipMemset(b->bad, 0, sizeof(int) * n_cases) != hipSuccess ||
        hipMemset(b->l2, 0, sizeof(double) * n_cases * nv) != hipSuccess ||
        hipMemset(b->dts, 0, sizeof(double) * n_cases) != hipSuccess) {
        release(b);
        return mst::fail(MST_E_HIP, "hipMemset failed");
    }
    TRY(build_patches(b));
    TRY(build_items(b));
#undef TRY
    *out = b;
    return MST_OK;
}

int mst_batch_destroy(mst_batch *b) {
    if (!b) return MST_OK;
    (void)hipStreamSynchronize(b->stream);
    release(b);
    return MST_OK;
}

int mst_batch_counts(const mst_batch *b, int64_t *out) {
    B_GUARD(b);
    if (!out) return mst::fail(MST_E_ARG, "null argument");
    out[0] = b->n;
    out[1] = b->NC;
    out[2] = b->NF;
    out[3] = b->n_citems;
    out[4] = b->n_fitems;
    out[5] = b->dim;
    return MST_OK;
}

int mst_batch_case_counts(const mst_batch *b, int case_id, int64_t *n_cells, int64_t *n_faces) {
    B_GUARD(b);
    if (case_id < 0 || case_id >= b->n || !n_cells || !n_faces) return mst::fail(MST_E_ARG, "bad case");
    *n_cells = b->cstart[case_id + 1] - b->cstart[case_id];
    *n_faces = b->fstart[case_id + 1] - b->fstart[case_id];
    return MST_OK;
}

int mst_batch_set_state(mst_batch *b, int case_id, const double *Q, int64_t n_cells) {
    B_GUARD(b);
    if (case_id < 0 || case_id >= b->n || !Q) return mst::fail(MST_E_ARG, "bad case");
    const mst::Prepared &P = b->preps[case_id];
    if (n_cells != P.nc) return mst::fail(MST_E_ARG, "state size does not match the case's mesh");
    const int nv = b->nv;
    std::vector<double> buf(P.nc * nv);
    for (int64_t i = 0; i < P.nc; i++)
        for (int k = 0; k < nv; k++) buf[i * nv + k] = Q[(int64_t)P.cell_old_of_new[i] * nv + k];
    HIPCHK(hipStreamSynchronize(b->stream));
    // an inactive case keeps its state in every buffer (the buffers rotate under it)
    for (int i = 0; i < 4; i++)
        if (i == b->iold || !b->active[case_id])
            HIPCHK(hipMemcpy(b->Q[i] + b->cstart[case_id] * nv, buf.data(), sizeof(double) * buf.size(),
                             hipMemcpyHostToDevice));
    for (int i = 0; i < 4; i++) b->shadow_ok[i] = false;
    b->dt_valid = false;
    return MST_OK;
}

int mst_batch_get_state(mst_batch *b, int case_id, double *Q, int64_t n_cells) {
    B_GUARD(b);
    if (case_id < 0 || case_id >= b->n || !Q) return mst::fail(MST_E_ARG, "bad case");
    const mst::Prepared &P = b->preps[case_id];
    if (n_cells != P.nc) return mst::fail(MST_E_ARG, "state size does not match the case's mesh");
    const int nv = b->nv;
    std::vector<double> buf(P.nc * nv);
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(buf.data(), b->Q[b->iold] + b->cstart[case_id] * nv, sizeof(double) * buf.size(),
                     hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < P.nc; c++)
        for (int k = 0; k < nv; k++) Q[c * nv + k] = buf[(int64_t)P.cell_new_of_old[c] * nv + k];
    return MST_OK;
}

int mst_batch_step(mst_batch *b, int nsteps) {
    B_GUARD(b);
    if (nsteps < 0) return mst::fail(MST_E_ARG, "nsteps must be >= 0");
    int rc = do_steps(b, nsteps);
    if (rc) return rc;
    return check_bad(b);
}

int mst_batch_residual_norms(mst_batch *b, double *out) {
    B_GUARD(b);
    if (!out) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(out, b->l2, sizeof(double) * b->n * b->nv, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_batch_dt(mst_batch *b, double *out) {
    B_GUARD(b);
    if (!out) return mst::fail(MST_E_ARG, "null argument");
    if (!b->dt_valid) {
        int rc = ensure_shadow(b, b->iold);
        if (!rc) rc = run_cfl(b);
        if (rc) return rc;
        b->dt_valid = true;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(out, b->dts, sizeof(double) * b->n, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_batch_set_active(mst_batch *b, const uint8_t *active) {
    B_GUARD(b);
    if (!active) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(b->stream));
    const int nv = b->nv;
    for (int i = 0; i < b->n; i++) {
        const uint8_t a = active[i] ? 1 : 0;
        if (a == b->active[i]) continue;
        if (!a) {  // freeze: the current state goes to every buffer
            const size_t off = b->cstart[i] * nv, bytes = sizeof(double) * (b->cstart[i + 1] - b->cstart[i]) * nv;
            for (int k = 0; k < 4; k++)
                if (k != b->iold) HIPCHK(hipMemcpy(b->Q[k] + off, b->Q[b->iold] + off, bytes, hipMemcpyDeviceToDevice));
            for (int k = 0; k < 4; k++) b->shadow_ok[k] = false;
        }
        b->active[i] = a;
    }
    return build_items(b);
}

int mst_batch_set_steady_driver(mst_batch *b, int driver, double cfl, double omega) {
    B_GUARD(b);
    if (driver != MST_STEADY_EXPLICIT && driver != MST_STEADY_LUSGS)
        return mst::fail(MST_E_ARG, "driver must be MST_STEADY_EXPLICIT or MST_STEADY_LUSGS");
    if (driver == MST_STEADY_LUSGS) {
        if (!(cfl > 0) || !(omega > 0) || !std::isfinite(cfl))
            return mst::fail(MST_E_ARG, "LU-SGS needs a finite cfl > 0 and omega > 0");
        if (b->mixed) return mst::fail(MST_E_UNSUPPORTED, "batched LU-SGS runs FP64 batches");
        if (b->slots > 4) return mst::fail(MST_E_UNSUPPORTED, "batched LU-SGS covers cells of at most 4 faces");
        b->lu_cfl = cfl;
        b->lu_omega = omega;
    }
    b->steady = driver;
    return MST_OK;
}

int mst_batch_lusgs_step(mst_batch *b, int n_iter) {
    B_GUARD(b);
    if (n_iter < 0) return mst::fail(MST_E_ARG, "n_iter must be >= 0");
    if (b->steady != MST_STEADY_LUSGS) return mst::fail(MST_E_ARG, "select the LU-SGS driver first (mst_batch_set_steady_driver)");
    int rc = build_lu(b);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(b->lu_active, b->active.data(), b->n, hipMemcpyHostToDevice));
    for (int it = 0; it < n_iter; it++) {
        rc = b->dim == 2 ? lu_iteration_b<2>(b, it == n_iter - 1) : lu_iteration_b<3>(b, it == n_iter - 1);
        if (rc) return rc;
        b->shadow_ok[b->inew] = false;
        std::swap(b->iold, b->inew);
    }
    if (n_iter > 0) b->dt_valid = false;
    return check_bad(b);
}

int mst_batch_run_to_steady(mst_batch *b, double tol, int max_steps, int check_every, double *ref, int32_t *steps,
                            double *drop) {
    B_GUARD(b);
    if (!ref || !steps || !drop || !(tol > 0) || max_steps < 1 || check_every < 1)
        return mst::fail(MST_E_ARG, "bad arguments");
    const int nv = b->nv, n = b->n;
    const std::vector<uint8_t> active0 = b->active;
    std::vector<uint8_t> act = active0;
    std::vector<double> r((size_t)n * nv);
    for (int i = 0; i < n; i++) {
        steps[i] = 0;
        drop[i] = INFINITY;
    }
    int done = 0;
    int rc = MST_OK;
    // cases without a reference take the residual of their first step
    bool need_ref = false;
    for (int i = 0; i < n; i++) need_ref |= act[i] && !(ref[(size_t)i * nv] > 0);
    const bool lu = b->steady == MST_STEADY_LUSGS;
    if (need_ref) {
        rc = lu ? mst_batch_lusgs_step(b, 1) : mst_batch_step(b, 1);
        if (!rc) rc = mst_batch_residual_norms(b, r.data());
        if (rc) return rc;
        for (int i = 0; i < n; i++) {
            if (!act[i]) continue;
            steps[i] = 1;
            if (!(ref[(size_t)i * nv] > 0))
                for (int k = 0; k < nv; k++) ref[(size_t)i * nv + k] = r[(size_t)i * nv + k];
        }
        done = 1;
    }
    while (done < max_steps) {
        int nact = 0;
        for (int i = 0; i < n; i++) nact += act[i];
        if (!nact) break;
        const int k = std::min(check_every, max_steps - done);
        rc = lu ? mst_batch_lusgs_step(b, k) : mst_batch_step(b, k);
        if (rc) break;
        done += k;
        rc = mst_batch_residual_norms(b, r.data());
        if (rc) break;
        bool changed = false;
        for (int i = 0; i < n; i++) {
            if (!act[i]) continue;
            steps[i] += k;
            double q = 0;
            for (int c = 0; c < nv; c++)
                q = std::max(q, ref[(size_t)i * nv + c] > 0 ? r[(size_t)i * nv + c] / ref[(size_t)i * nv + c] : 0.0);
            drop[i] = q;
            if (q <= tol) {  // converged: freeze
                act[i] = 0;
                changed = true;
            }
        }
        if (changed) {
            rc = mst_batch_set_active(b, act.data());
            if (rc) break;
        }
    }
    int rc2 = mst_batch_set_active(b, active0.data());
    return rc ? rc : rc2;
}

int mst_batch_set_tuning(mst_batch *b, int block, const int32_t *tiles) {
    B_GUARD(b);
    if (block != 64 && block != 128 && block != 256) return mst::fail(MST_E_ARG, "block must be 64, 128 or 256");
    if (!tiles) return mst::fail(MST_E_ARG, "null argument");
    for (int k = 0; k < MST_BATCH_BUCKETS; k++)
        if (tiles[k] < 64 || tiles[k] > (1 << 24)) return mst::fail(MST_E_ARG, "tile out of range [64, 2^24]");
    b->block = block;
    for (int k = 0; k < MST_BATCH_BUCKETS; k++) b->tiles[k] = tiles[k];
    return build_items(b);
}

int mst_batch_get_tuning(const mst_batch *b, int *block, int32_t *tiles) {
    B_GUARD(b);
    if (!block || !tiles) return mst::fail(MST_E_ARG, "null argument");
    *block = b->block;
    for (int k = 0; k < MST_BATCH_BUCKETS; k++) tiles[k] = b->tiles[k];
    return MST_OK;
}

int mst_batch_set_strategy(mst_batch *b, int strategy) {
    B_GUARD(b);
    if (strategy != MST_STRATEGY_FACE && strategy != MST_STRATEGY_CELL && strategy != MST_STRATEGY_PATCH &&
        strategy != MST_STRATEGY_FUSED)
        return mst::fail(MST_E_ARG, "batch strategy must be MST_STRATEGY_FACE, _CELL, _PATCH or _FUSED");
    if ((strategy == MST_STRATEGY_PATCH && !b->patch_ok) || (strategy == MST_STRATEGY_FUSED && !b->fused_ok))
        return mst::fail(MST_E_UNSUPPORTED, "patch strategies: need at most 4 faces per cell and a patch's face data within 64 KiB");
    if (strategy != b->strategy) {
        auto patched = [](int st) { return st == MST_STRATEGY_PATCH || st == MST_STRATEGY_FUSED; };
        const bool items = patched(strategy) != patched(b->strategy);
        b->strategy = strategy;
        drop_graphs(b);
        if (items) return build_items(b);  // patch items replace the tiles
    }
    return MST_OK;
}

int mst_batch_get_strategy(const mst_batch *b, int *strategy) {
    B_GUARD(b);
    if (!strategy) return mst::fail(MST_E_ARG, "null argument");
    *strategy = b->strategy;
    return MST_OK;
}

int mst_batch_set_graph(mst_batch *b, int on) {
    B_GUARD(b);
    b->use_graph = on != 0;
    return MST_OK;
}

int mst_batch_autotune(mst_batch *b, const char *cache_path, int force, int steps, double *ms_per_step) {
    B_GUARD(b);
    if (steps < 1) return mst::fail(MST_E_ARG, "steps must be >= 1");
    const std::string key = signature(b);
    int tiles[MST_BATCH_BUCKETS];
    int block = 0, strat = MST_STRATEGY_CELL;
    double ms = 0;
    if (!force && cache_path && cache_lookup(cache_path, key, &block, tiles, &strat, &ms)) {
        int rc = mst_batch_set_strategy(b, strat);
        if (!rc) rc = mst_batch_set_tuning(b, block, tiles);
        if (rc) return rc;
        if (ms_per_step) *ms_per_step = ms;
        return MST_OK;
    }
    // the search advances the solution: keep the state and restore it afterwards
    const size_t bytes = sizeof(double) * b->NC * b->nv;
    double *save = nullptr;
    HIPCHK(hipMalloc((void **)&save, bytes));
    const int iold0 = b->iold, inew0 = b->inew;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(save, b->Q[b->iold], bytes, hipMemcpyDeviceToDevice));
    int hist[MST_BATCH_BUCKETS] = {0, 0, 0, 0};
    for (int i = 0; i < b->n; i++)
        if (b->active[i]) hist[bucket_of(b->cstart[i + 1] - b->cstart[i])]++;
    static const int cand_tiles[] = {256, 512, 1024, 2048, 4096, 8192, 16384};
    static const int cand_blocks[] = {64, 128, 256};
    int best_block = b->block, best_strat = b->strategy, best_tiles[MST_BATCH_BUCKETS];
    for (int k = 0; k < MST_BATCH_BUCKETS; k++) best_tiles[k] = b->tiles[k];
    double best = INFINITY;
    int rc = MST_OK;
    auto trial = [&](int strategy, int blk, const int *t) -> int {
        // every trial starts from the saved state, so all trials time the same steps
        HIPCHK(hipStreamSynchronize(b->stream));
        b->iold = iold0;
        b->inew = inew0;
        HIPCHK(hipMemcpy(b->Q[b->iold], save, bytes, hipMemcpyDeviceToDevice));
        for (int k = 0; k < 4; k++) b->shadow_ok[k] = false;
        b->dt_valid = false;
        b->block = blk;
        b->strategy = strategy;
        for (int k = 0; k < MST_BATCH_BUCKETS; k++) b->tiles[k] = t[k];
        double m = 0;
        int r = time_steps(b, steps, &m);
        if (r) return r;
        if (m < best) {
            best = m;
            best_block = blk;
            best_strat = strategy;
            for (int k = 0; k < MST_BATCH_BUCKETS; k++) best_tiles[k] = t[k];
        }
        return MST_OK;
    };
    // coordinate search: strategy x block size, then each populated bucket's tile, then the block
    // size again
    for (int pass = 0; pass < 2 && !rc; pass++) {
        for (int st : {MST_STRATEGY_FACE, MST_STRATEGY_CELL, MST_STRATEGY_PATCH, MST_STRATEGY_FUSED}) {
            if (pass == 1 && st != best_strat) continue;
            if (st == MST_STRATEGY_PATCH || st == MST_STRATEGY_FUSED) {  // fixed shape: one BLOCK-cell patch per item
                const bool ok = st == MST_STRATEGY_PATCH ? b->patch_ok : (b->fused_ok && b->order >= 2);
                if (pass == 0 && ok) rc = trial(st, BLOCK, best_tiles);
                if (rc) break;
                continue;
            }
            for (int blk : cand_blocks) {
                if (pass == 1 && blk == best_block) continue;
                rc = trial(st, blk, best_tiles);
                if (rc) break;
            }
            if (rc) break;
        }
        if (pass == 1 || rc) break;
        if (best_strat == MST_STRATEGY_PATCH || best_strat == MST_STRATEGY_FUSED) break;  // tiles do not apply
        for (int k = 0; k < MST_BATCH_BUCKETS && !rc; k++) {
            if (!hist[k]) continue;
            for (int t : cand_tiles) {
                if (t == best_tiles[k]) continue;
                int tt[MST_BATCH_BUCKETS];
                for (int j = 0; j < MST_BATCH_BUCKETS; j++) tt[j] = best_tiles[j];
                tt[k] = t;
                rc = trial(best_strat, best_block, tt);
                if (rc) break;
            }
        }
    }
    // restore the state and apply the best shape
    if (hipStreamSynchronize(b->stream) == hipSuccess) {
        b->iold = iold0;
        b->inew = inew0;
        (void)hipMemcpy(b->Q[b->iold], save, bytes, hipMemcpyDeviceToDevice);
    }
    (void)hipFree(save);
    for (int k = 0; k < 4; k++) b->shadow_ok[k] = false;
    b->dt_valid = false;
    if (rc) return rc;
    int rc2 = check_bad(b);
    if (rc2) return rc2;
    rc = mst_batch_set_strategy(b, best_strat);
    if (!rc) rc = mst_batch_set_tuning(b, best_block, best_tiles);
    if (rc) return rc;
    if (cache_path) cache_store(cache_path, key, best_block, best_tiles, best_strat, best);
    if (ms_per_step) *ms_per_step = best;
    return MST_OK;
}

}  // extern "C"
