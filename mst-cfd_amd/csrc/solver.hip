// solver.hip -- gfx950 kernels of the density-based explicit residual path and the device context
// behind include/mstcfd.h.
//
// Per residual evaluation (one RK stage) the device runs, all on the context's stream:
//   k_grad    (order 2)  one lane per cell: Green-Gauss gradient + limiter -> limited gradient
//   k_flux               one lane per face: reconstruct, ghost state, dimension-split Roe/AUSM,
//                        G_f = sum_d n_d F_d  (c0 outward normal)
//   k_update             one lane per cell: R_c = sum +-G_f over its faces in a fixed order (no float
//                        atomics -> bit-reproducible), then the Euler / SSP-RK stage update
// plus k_cfl (device CFL time step) and block-partial reductions finished by one-block kernels.
// Memory layout (HBM): cell states as nv-double records (32 B in 2D = two 16 B loads per gather),
// face streams as separate coalesced arrays, cell->face incidences slot-major (coalesced per slot).
// Workgroups are remapped so that each XCD (own L2) walks a contiguous range of the RCM order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mstcfd.h"
#include "prep.hpp"

namespace mst {
const std::string &last_error();
}

namespace {

constexpr int BLOCK = 256;

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess)                                                                            \
            return mst::fail(MST_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

struct DevParams {
    double gamma, gm1, eor, delta, dd, ihd;  // dd = delta^2, ihd = 0.5/delta (Harten fix)
    double astar_c;                   // 2 (gamma-1) / (gamma+1)   (SolverAusm.cpp:6-7)
    int compat;
    double qin[5];
};

// ---------------------------------------------------------------------------------- device helpers
// Each XCD receives every 8th workgroup; remap so that XCD x processes one contiguous slice of
// the (RCM-ordered) index space and its L2 sees neighbouring cells.
__device__ __forceinline__ int64_t block_id() {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t q = G >> 3, r = G & 7u, x = b & 7u, s = b >> 3;
    return (int64_t)(x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

template <int NV>
__device__ __forceinline__ void ldq(const double *__restrict__ p, double *q) {
    if constexpr (NV == 4) {
        const double2 a = reinterpret_cast<const double2 *>(p)[0];
        const double2 b = reinterpret_cast<const double2 *>(p)[1];
        q[0] = a.x; q[1] = a.y; q[2] = b.x; q[3] = b.y;
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) q[k] = p[k];
    }
}
template <int NV>
__device__ __forceinline__ void stq(double *__restrict__ p, const double *q) {
    if constexpr (NV == 4) {
        reinterpret_cast<double2 *>(p)[0] = make_double2(q[0], q[1]);
        reinterpret_cast<double2 *>(p)[1] = make_double2(q[2], q[3]);
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) p[k] = q[k];
    }
}

// Q + lg . dx  (RhoSolver.cpp:250,254 with the limiter factor folded into lg)
template <int DIM>
__device__ __forceinline__ void recon(double *q, const double *__restrict__ lg, const double *dx) {
    constexpr int NV = DIM + 2;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = lg[k * DIM + 0] * dx[0] + lg[k * DIM + 1] * dx[1];
        if constexpr (DIM == 3) s += lg[k * DIM + 2] * dx[2];
        q[k] = q[k] + s;
    }
}

// wall/symmetry ghost momentum m - 2 n^ (n^ . m) from src, n^ = n/|n| (RhoSolver.cpp:150-158)
template <int DIM>
__device__ __forceinline__ void reflect(double *dst, const double *src, const double *n) {
    double nn = n[0] * n[0] + n[1] * n[1];
    if constexpr (DIM == 3) nn += n[2] * n[2];
    const double nrm = sqrt(nn);
    double nh[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) nh[d] = n[d] / nrm;
    double s = nh[0] * src[1] + nh[1] * src[2];
    if constexpr (DIM == 3) s += nh[2] * src[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) dst[1 + d] = src[1 + d] - 2 * nh[d] * s;
}

// 1/x from the hardware reciprocal estimate refined by two Newton steps (the same seed the
// compiler's IEEE division expands from, without the scale/fixup stages): within 1 ulp for the
// positive, normal arguments it is used on (densities, speeds, volumes)
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double hfix(double x, const DevParams &P) {  // SolverRoe.cpp:114-123
    return x > P.delta ? x : (x * x + P.dd) * P.ihd;
}

// Roe, dimension split (SolverRoe.cpp:3-16,70-112).  The Roe average is computed once per face
// (it is symmetric in L/R); per axis d the dissipation R|L|R^-1 (Q_R-Q_L) uses the analytic wave
// strengths of the reference's eigenvector matrix (same columns as SolverRoe.cpp:86,92, exact
// inverse including the (H - q^2/2) denominator the explicit 4x4 inverse produces), and flips
// sign when the ghost/c1 side is the Riemann left state.
template <int DIM>
__device__ __forceinline__ void roe_face(const double *A, const double *B, const double *n, int mask,
                                         const DevParams &P, double *g) {
    constexpr int NV = DIM + 2;
    const double iA = frcp(A[0]), iB = frcp(B[0]);
    double mA2 = 0, mB2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        mA2 += A[1 + k] * A[1 + k];
        mB2 += B[1 + k] * B[1 + k];
    }
    const double pA = (A[NV - 1] - 0.5 * mA2 * iA) * P.gm1;
    const double pB = (B[NV - 1] - 0.5 * mB2 * iB) * P.gm1;
    const double HA = (A[NV - 1] + pA) * iA, HB = (B[NV - 1] + pB) * iB;
    const double D = sqrt(fabs(B[0] * iA));
    const double i1D = frcp(1.0 + D);
    double u[DIM], q2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        u[k] = (A[1 + k] * iA + D * (B[1 + k] * iB)) * i1D;
        q2 += u[k] * u[k];
    }
    const double Ht = (HA + D * HB) * i1D;
    const double h2 = 0.5 * q2;
    const double a = sqrt(fabs(P.gm1 * (Ht - h2)));
    const double ia = frcp(a), iden = 1.0 / (Ht - h2);
    const double dr = B[0] - A[0], dE = B[NV - 1] - A[NV - 1];
    double beta[DIM], udm = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        const double dm = B[1 + k] - A[1 + k];
        udm += u[k] * dm;
        beta[k] = dm - u[k] * dr;
    }
    const double s = (dE - udm + h2 * dr) * iden;  // alpha(-) + alpha(+)
    const double a0 = dr - s;                      // entropy wave strength
    const double ieA = frcp(A[0] + P.eor), ieB = frcp(B[0] + P.eor);
#pragma unroll
    for (int k = 0; k < NV; k++) g[k] = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double bd = beta[d] * ia;
        const double ap = 0.5 * (s + bd), am = 0.5 * (s - bd);
        const double lm = hfix(fabs(u[d] - a), P), l0 = hfix(fabs(u[d]), P), lp = hfix(fabs(u[d] + a), P);
        const double wm = lm * am, wp = lp * ap, w0 = l0 * a0;
        const double sr = wm + w0 + wp;
        double Dv[NV];
        Dv[0] = sr;
        double esh = 0;
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            if (k == d) {
                Dv[1 + k] = u[k] * sr + a * (wp - wm);
            } else {
                const double t = l0 * beta[k];
                Dv[1 + k] = u[k] * sr + t;
                esh += u[k] * t;
            }
        }
        Dv[NV - 1] = Ht * (wm + wp) + u[d] * a * (wp - wm) + h2 * w0 + esh;
        const double sg = ((mask >> d) & 1) ? 0.5 : -0.5;
        const double mAd = A[1 + d], mBd = B[1 + d];
        double F[NV];
        F[0] = 0.5 * (mAd + mBd) - sg * Dv[0];
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            double fa = A[1 + k] * mAd * ieA, fb = B[1 + k] * mBd * ieB;
            if (k == d) {
                fa += pA;
                fb += pB;
            }
            F[1 + k] = 0.5 * (fa + fb) - sg * Dv[1 + k];
        }
        F[NV - 1] = 0.5 * (HA * mAd + HB * mBd) - sg * Dv[NV - 1];
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] += n[d] * F[k];
    }
}

// AUSM+ with the reference's quirks (SolverAusm.cpp:3-26,53-63,105-143)
template <int DIM>
__device__ __forceinline__ void ausm_face(const double *A, const double *B, const double *n, int mask,
                                          const DevParams &P, double *g) {
    constexpr int NV = DIM + 2;
    const double iA = frcp(A[0]), iB = frcp(B[0]);
    double mA2 = 0, mB2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        mA2 += A[1 + k] * A[1 + k];
        mB2 += B[1 + k] * B[1 + k];
    }
    const double pA = (A[NV - 1] - 0.5 * mA2 * iA) * P.gm1;
    const double pB = (B[NV - 1] - 0.5 * mB2 * iB) * P.gm1;
    const double HA = (A[NV - 1] + pA) * iA, HB = (B[NV - 1] + pB) * iB;
    const double asA = sqrt(HA * P.astar_c), asB = sqrt(HB * P.astar_c);
    const double UA = sqrt(mA2 * iA * iA), UB = sqrt(mB2 * iB * iB);
    const double tA = asA * asA * frcp(fmax(asA, UA)), tB = asB * asB * frcp(fmax(asB, UB));
    const double aF = fmin(tA, tB);
    const double iaF = frcp(aF);
    double aA = aF, aB = aF;
    if (P.compat & MST_COMPAT_AUSM_LOCAL_A) {
        aA = sqrt(P.gamma * pA * iA);
        aB = sqrt(P.gamma * pB * iB);
    }
    const bool unscaled = (P.compat & MST_COMPAT_AUSM_P_UNSCALED) != 0;
#pragma unroll
    for (int k = 0; k < NV; k++) g[k] = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const bool left_is_A = (mask >> d) & 1;
        const double *L = left_is_A ? A : B, *R = left_is_A ? B : A;
        const double iL = left_is_A ? iA : iB, iR = left_is_A ? iB : iA;
        const double pL = left_is_A ? pA : pB, pR = left_is_A ? pB : pA;
        const double aL = left_is_A ? aA : aB, aR = left_is_A ? aB : aA;
        const double mL = L[1 + d] * iL * iaF, mR = R[1 + d] * iR * iaF;
        double Mp, Mm, Pp, Pm;
        const double qL = mL * mL - 1, qR = mR * mR - 1;
        if (mL <= 1) {
            Mp = 0.25 * (mL + 1) * (mL + 1) + 0.125 * qL * qL;
            const double poly = 0.25 * (mL + 1) * (mL + 1) * (2 - mL);
            const double corr = 0.1875 * mL * qL * qL;
            Pp = unscaled ? pL * poly + corr : pL * (poly + corr);
        } else {
            Mp = 0.5 * (mL + fabs(mL));
            Pp = pL * 0.5 * (mL + fabs(mL)) / mL;
        }
        if (mR <= 1) {
            Mm = -0.25 * (mR - 1) * (mR - 1) - 0.125 * qR * qR;
            const double poly = 0.25 * (mR - 1) * (mR - 1) * (2 + mR);
            const double corr = 0.1875 * mR * qR * qR;
            Pm = unscaled ? pR * poly - corr : pR * (poly - corr);
        } else {
            Mm = 0.5 * (mR - fabs(mR));
            Pm = pR * 0.5 * (mR - fabs(mR)) / mR;
        }
        const double Mf = Mm + Mp, pf = Pm + Pp, aMf = fabs(Mf);
        double F[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const double psL = (k == NV - 1) ? L[k] + pL : L[k];
            const double psR = (k == NV - 1) ? R[k] + pR : R[k];
            F[k] = 0.5 * (Mf * (aL * psL + aR * psR) - aMf * (aR * psR - aL * psL));
        }
        F[1 + d] += pf;
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] += n[d] * F[k];
    }
}

template <int DIM, int SCHEME>
__device__ __forceinline__ void riemann_face(const double *A, const double *B, const double *n, int mask,
                                             const DevParams &P, double *g) {
    if constexpr (SCHEME == MST_SCHEME_ROE)
        roe_face<DIM>(A, B, n, mask, P, g);
    else
        ausm_face<DIM>(A, B, n, mask, P, g);
}

// ---------------------------------------------------------------------------------- face flux
struct FluxArgs {
    int64_t nf, pstart;
    const int32_t *__restrict__ fc0;
    const int32_t *__restrict__ fc1;
    const int32_t *__restrict__ finfo;  // kind | (L/R mask << 4)
    const double *__restrict__ fn;
    const double *__restrict__ fxf;
    const double *__restrict__ pxf;
    const double *__restrict__ Q;
    const double *__restrict__ cxc;
    const double *__restrict__ lg;
    double *__restrict__ G;
    DevParams P;
};

// RhoSolver::updateFaceFlux (RhoSolver.cpp:90-233) / updateFaceFlux2ndOrder (:234-369), one lane per face
template <int DIM, int SCHEME, bool SECOND>
__global__ __launch_bounds__(BLOCK) void k_flux(const FluxArgs a) {
    constexpr int NV = DIM + 2;
    const int64_t f = block_id() * BLOCK + threadIdx.x;
    if (f >= a.nf) return;
    const int info = a.finfo[f];
    const int kind = info & 15, mask = info >> 4;
    double g[NV];
    if (kind == mst::KIND_NONE) {
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] = 0;
        stq<NV>(a.G + f * NV, g);
        return;
    }
    const int64_t c0 = a.fc0[f];
    double n[DIM], A[NV], B[NV];
#pragma unroll
    for (int d = 0; d < DIM; d++) n[d] = a.fn[f * DIM + d];
    ldq<NV>(a.Q + c0 * NV, A);
    if (kind == mst::KIND_INTERIOR || kind == mst::KIND_PERIODIC) {
        const int64_t c1 = a.fc1[f];
        ldq<NV>(a.Q + c1 * NV, B);
        if constexpr (SECOND) {
            double xf[DIM], dx[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                xf[d] = a.fxf[f * DIM + d];
                dx[d] = xf[d] - a.cxc[c0 * DIM + d];
            }
            recon<DIM>(A, a.lg + c0 * NV * DIM, dx);
            if (kind == mst::KIND_PERIODIC) {
#pragma unroll
                for (int d = 0; d < DIM; d++) xf[d] = a.pxf[(f - a.pstart) * DIM + d];
            }
#pragma unroll
            for (int d = 0; d < DIM; d++) dx[d] = xf[d] - a.cxc[c1 * DIM + d];
            recon<DIM>(B, a.lg + c1 * NV * DIM, dx);
        }
    } else {
        double Rc[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) Rc[k] = A[k];
        if constexpr (SECOND) {
            double dx[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) dx[d] = a.fxf[f * DIM + d] - a.cxc[c0 * DIM + d];
            recon<DIM>(Rc, a.lg + c0 * NV * DIM, dx);
        }
        if (kind == mst::KIND_WALL) {  // RhoSolver.cpp:142-179 / 285-319
#pragma unroll
            for (int k = 0; k < NV; k++) A[k] = B[k] = Rc[k];
            reflect<DIM>(B, A, n);
        } else if (kind == mst::KIND_SYMMETRY) {  // :184-211 / 322-350
#pragma unroll
            for (int k = 0; k < NV; k++) B[k] = Rc[k];
            reflect<DIM>(B, A, n);
        } else if (kind == mst::KIND_INLET) {  // :121-137 / 265-281
#pragma unroll
            for (int k = 0; k < NV; k++) {
                A[k] = Rc[k];
                B[k] = a.P.qin[k];
            }
        } else {  // outlet :216-226 / 353-363
#pragma unroll
            for (int k = 0; k < NV; k++) B[k] = Rc[k];
        }
    }
    riemann_face<DIM, SCHEME>(A, B, n, mask, a.P, g);
    stq<NV>(a.G + f * NV, g);
}

// ---------------------------------------------------------------------------------- gradient + limiter
struct GradArgs {
    int64_t nc;
    const int32_t *__restrict__ ell;
    const int32_t *__restrict__ fc0;
    const int32_t *__restrict__ fc1;
    const double *__restrict__ fn;
    const double *__restrict__ fxf;
    const double *__restrict__ feta;
    const double *__restrict__ Q;
    const double *__restrict__ cxc;
    const double *__restrict__ vol;
    double *__restrict__ lg;
    double *__restrict__ graw;
    double *__restrict__ phi;
};

// RhoSolver::updateGradFlux (RhoSolver.cpp:430-454) fused with the (new) per-cell limiter
template <int DIM, int MAXF, int LIM>
__global__ __launch_bounds__(BLOCK) void k_grad(const GradArgs a) {
    constexpr int NV = DIM + 2;
    const int64_t c = block_id() * BLOCK + threadIdx.x;
    if (c >= a.nc) return;
    double qc[NV], qmax[NV], qmin[NV], t[NV][DIM];
    ldq<NV>(a.Q + c * NV, qc);
#pragma unroll
    for (int k = 0; k < NV; k++) {
        qmax[k] = qmin[k] = qc[k];
#pragma unroll
        for (int d = 0; d < DIM; d++) t[k][d] = 0;
    }
    double xf[MAXF][DIM];
    int nfc = 0;
#pragma unroll
    for (int s = 0; s < MAXF; s++) {
        const int32_t e = a.ell[(int64_t)s * a.nc + c];
        if (e < 0) break;
        const int64_t f = e >> 1;
        const bool is1 = e & 1;
        const int64_t nb = is1 ? a.fc0[f] : a.fc1[f];
        double n[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double v = a.fn[f * DIM + d];
            n[d] = is1 ? -v : v;
            xf[s][d] = a.fxf[f * DIM + d];
        }
        double qf[NV];
        if (nb >= 0) {
            double qn[NV];
            ldq<NV>(a.Q + nb * NV, qn);
            const double eta = a.feta[f];
#pragma unroll
            for (int k = 0; k < NV; k++) {
                qf[k] = is1 ? eta * qn[k] + (1 - eta) * qc[k] : eta * qc[k] + (1 - eta) * qn[k];
                qmax[k] = fmax(qmax[k], qn[k]);
                qmin[k] = fmin(qmin[k], qn[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NV; k++) qf[k] = qc[k];
        }
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) t[k][d] += qf[k] * n[d];
        nfc = s + 1;
    }
    const double iV = frcp(a.vol[c]);
    double xc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) xc[d] = a.cxc[c * DIM + d];
    double ph[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        ph[k] = 1;
#pragma unroll
        for (int d = 0; d < DIM; d++) t[k][d] *= iV;
    }
    if constexpr (LIM != MST_LIM_NONE) {
        // Phi = min(1, min_f psi(r_f)) = min(1, psi(min_f r_f)) since psi is increasing: the smallest
        // ratio r = |a|/|delta| is tracked as a fraction (bn/bd) and psi is applied once per field
        double bn[NV], bd[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            bn[k] = 1.0;
            bd[k] = 0.0;  // "infinite" ratio: no face limits
        }
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            if (s >= nfc) break;
            double dx[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) dx[d] = xf[s][d] - xc[d];
#pragma unroll
            for (int k = 0; k < NV; k++) {
                double delta = t[k][0] * dx[0] + t[k][1] * dx[1];
                if constexpr (DIM == 3) delta += t[k][2] * dx[2];
                const double eps = 1e-12 * (fabs(qmax[k]) + fabs(qmin[k]));
                const double ext = delta > 0 ? qmax[k] : qmin[k];
                const double an = fabs(ext - qc[k]), ad = fabs(delta);
                if (fabs(delta) > eps && (bd[k] == 0.0 || an * bd[k] < bn[k] * ad)) {
                    bn[k] = an;
                    bd[k] = ad;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NV; k++) {
            if (bd[k] == 0.0) continue;
            if constexpr (LIM == MST_LIM_MINMOD)
                ph[k] = bn[k] < bd[k] ? bn[k] / bd[k] : 1.0;
            else
                ph[k] = fmin(1.0, 2.0 * bn[k] / (bd[k] + bn[k]));  // (r+|r|)/(1+|r|) at r = bn/bd >= 0
        }
    }
    double out[NV * DIM];
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
        for (int d = 0; d < DIM; d++) out[k * DIM + d] = ph[k] * t[k][d];
#pragma unroll
    for (int i = 0; i < NV * DIM; i++) a.lg[c * NV * DIM + i] = out[i];
    if (a.graw) {
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) a.graw[c * NV * DIM + k * DIM + d] = t[k][d];
    }
    if (a.phi) {
#pragma unroll
        for (int k = 0; k < NV; k++) a.phi[c * NV + k] = ph[k];
    }
}

// ---------------------------------------------------------------------------------- gather + update
struct UpdArgs {
    int64_t nc;
    const int32_t *__restrict__ ell;
    const double *__restrict__ G;
    const double *__restrict__ Qs;
    const double *__restrict__ Qn;
    const double *__restrict__ vol;
    const double *__restrict__ sh;  // half sums of |S_d| per cell (CFL)
    double *dt;                     // [0] dt of this stage; rewritten with the next step's dt (UPD_DT)
    double ca, cb;
    int stage0;
    double gamma, cfl;
    double *__restrict__ Qout;
    double *__restrict__ R;
    double *__restrict__ part_l2;   // per-block partial sums of (R/V)^2: [field][block]
    double *__restrict__ part_dt;   // per-block minimum CFL time step
    int *__restrict__ bad;          // set to 1 when an updated state is not physical
};

// mode bits: UPD_STAGE writes the stage update (else only R), UPD_L2 reduces |R/V|^2,
// UPD_DT reduces the CFL time step of the updated state (for the next step)
enum { UPD_RESIDUAL = 0, UPD_STAGE = 1, UPD_L2 = 2, UPD_DT = 4 };

// Block-level reduction of W values (fixed tree order): sums, or a min when MIN is set; thread 0
// of the block holds the result.  The per-block partials are combined by k_finish in block order,
// so every reduction is bit-reproducible (no float atomics, no cross-block fences).
template <int W, bool MIN>
__device__ __forceinline__ void block_partial(double (&v)[W], double *__restrict__ out, int64_t stride) {
    __shared__ double sh[W][BLOCK];
#pragma unroll
    for (int k = 0; k < W; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < W; k++) {
                const double y = sh[k][threadIdx.x + w];
                sh[k][threadIdx.x] = MIN ? fmin(sh[k][threadIdx.x], y) : sh[k][threadIdx.x] + y;
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < W) out[threadIdx.x * stride] = sh[threadIdx.x][0];
}

template <int DIM>
__device__ __forceinline__ double local_dt(const double *q, const double *__restrict__ sh, int64_t c, double V,
                                           double gamma, double cfl) {
    constexpr int NV = DIM + 2;
    double ke = 0;
#pragma unroll
    for (int k = 1; k <= DIM; k++) ke += q[k] * q[k];
    const double p = (q[NV - 1] - 0.5 * ke / q[0]) * (gamma - 1);
    const double snd = sqrt(gamma * p / q[0]);
    double lam = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) lam += (fabs(q[1 + d] / q[0]) + snd) * sh[c * DIM + d];
    const double v = cfl * V / lam;
    return v == v ? v : -1.0;  // a NaN state reports a negative time step
}

// RhoSolver::solve gather + update (RhoSolver.cpp:45-68), SSP-RK stage form (new)
template <int NV, int MAXF, int MODE>
__global__ __launch_bounds__(BLOCK) void k_update(const UpdArgs a) {
    constexpr int DIM = NV - 2;
    constexpr bool STAGE = MODE & UPD_STAGE, L2 = MODE & UPD_L2, DT = MODE & UPD_DT;
    const int64_t blk = block_id();
    const int64_t c = blk * BLOCK + threadIdx.x;
    const bool valid = c < a.nc;
    double r[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) r[k] = 0;
    double red[NV + 1];
#pragma unroll
    for (int k = 0; k < NV; k++) red[k] = 0;
    red[NV] = INFINITY;
    if (valid) {
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            const int32_t e = a.ell[(int64_t)s * a.nc + c];
            if (e < 0) break;
            double gf[NV];
            ldq<NV>(a.G + (int64_t)(e >> 1) * NV, gf);
            if (e & 1) {
#pragma unroll
                for (int k = 0; k < NV; k++) r[k] -= gf[k];
            } else {
#pragma unroll
                for (int k = 0; k < NV; k++) r[k] += gf[k];
            }
        }
        if constexpr (!STAGE) {
            stq<NV>(a.R + c * NV, r);
            return;
        } else {
            const double V = a.vol[c];
            const double dtv = a.dt[0] / V;
            double qs[NV], out[NV];
            ldq<NV>(a.Qs + c * NV, qs);
            if (a.stage0) {
#pragma unroll
                for (int k = 0; k < NV; k++) out[k] = qs[k] - dtv * r[k];
            } else {
                double qn[NV];
                ldq<NV>(a.Qn + c * NV, qn);
#pragma unroll
                for (int k = 0; k < NV; k++) out[k] = a.ca * qn[k] + a.cb * (qs[k] - dtv * r[k]);
            }
            stq<NV>(a.Qout + c * NV, out);
            double ke = 0;
#pragma unroll
            for (int k = 1; k < NV - 1; k++) ke += out[k] * out[k];
            const double p = (out[NV - 1] - 0.5 * ke / out[0]) * (a.gamma - 1);
            if (!(out[0] > 0) || !(p > 0)) a.bad[0] = 1;
            if constexpr (L2) {
                const double iV = frcp(V);
#pragma unroll
                for (int k = 0; k < NV; k++) {
                    const double x = r[k] * iV;
                    red[k] = x * x;
                }
            }
            if constexpr (DT) red[NV] = local_dt<DIM>(out, a.sh, c, V, a.gamma, a.cfl);
        }
    }
    if constexpr (STAGE && L2) {
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = red[k];
        block_partial<NV, false>(v, a.part_l2 + blk, gridDim.x);
    }
    if constexpr (STAGE && DT) {
        double v[1] = {red[NV]};
        block_partial<1, true>(v, a.part_dt + blk, 1);
    }
}

// ---------------------------------------------------------------------------------- reductions
// CFL: dt_c = cfl V / sum_d (|u_d| + a) Shalf_d (new; the reference's DT is fixed, time/Time.cpp:62)
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_cfl(int64_t nc, const double *__restrict__ Q,
                                               const double *__restrict__ vol, const double *__restrict__ sh,
                                               double gamma, double cfl, double *__restrict__ part_dt) {
    constexpr int NV = DIM + 2;
    const int64_t blk = block_id();
    const int64_t c = blk * BLOCK + threadIdx.x;
    double v[1] = {INFINITY};
    if (c < nc) {
        double q[NV];
        ldq<NV>(Q + c * NV, q);
        v[0] = local_dt<DIM>(q, sh, c, vol[c], gamma, cfl);
    }
    block_partial<1, true>(v, part_dt + blk, 1);
}

// Time::goNextTimeStep monitor (time/Time.cpp:69-76): max |new-old|/old, values that do not compare
// greater are skipped
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_maxrel(int64_t nc, const double *__restrict__ Qnew,
                                                  const double *__restrict__ Qold, double *__restrict__ part) {
    const int64_t blk = block_id();
    const int64_t c = blk * BLOCK + threadIdx.x;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0;
    if (c < nc) {
        double a[NV], b[NV];
        ldq<NV>(Qnew + c * NV, a);
        ldq<NV>(Qold + c * NV, b);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const double x = fabs(a[k] - b[k]) / b[k];
            if (v[k] < x) v[k] = x;
        }
    }
    __shared__ double s[NV][BLOCK];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < NV; k++)
                if (s[k][threadIdx.x] < s[k][threadIdx.x + w]) s[k][threadIdx.x] = s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < NV) part[threadIdx.x * (int64_t)gridDim.x + blk] = s[threadIdx.x][0];
}

// One-block finish of the per-block partials.  Slot k of the W slots is a sum (kind 0), a min
// (kind 1) or a max (kind 2); every thread walks the blocks with a fixed stride, then a fixed
// butterfly inside each wave and a fixed pass over the waves, so the result is bit-reproducible.
constexpr int FIN = 1024;
struct FinArgs {
    const double *src[12];
    int kind[12];
    double *dst[12];
    int post[12];  // 1: sqrt(v / count)
    int n;
    int64_t nblk;
    double count;
};

__device__ __forceinline__ double fin_op(int kind, double a, double b) {
    return kind == 0 ? a + b : (kind == 1 ? fmin(a, b) : (a < b ? b : a));
}

__global__ __launch_bounds__(FIN) void k_finish(const FinArgs a) {
    constexpr int MAXW = 12;
    __shared__ double s[MAXW][FIN / 64];
    double acc[MAXW];
#pragma unroll
    for (int k = 0; k < MAXW; k++) acc[k] = (k < a.n && a.kind[k] == 1) ? INFINITY : 0.0;
    for (int64_t i = threadIdx.x; i < a.nblk; i += FIN) {
#pragma unroll
        for (int k = 0; k < MAXW; k++)
            if (k < a.n) acc[k] = fin_op(a.kind[k], acc[k], a.src[k][i]);
    }
#pragma unroll
    for (int k = 0; k < MAXW; k++) {
        if (k < a.n) {
            for (int m = 32; m > 0; m >>= 1) acc[k] = fin_op(a.kind[k], acc[k], __shfl_xor(acc[k], m, 64));
            if ((threadIdx.x & 63) == 0) s[k][threadIdx.x >> 6] = acc[k];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < a.n) {
        const int k = threadIdx.x;
        double v = s[k][0];
        for (int w = 1; w < FIN / 64; w++) v = fin_op(a.kind[k], v, s[k][w]);
        a.dst[k][0] = a.post[k] ? sqrt(v / a.count) : v;
    }
}

// plain streaming copy: the reference point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes
__global__ __launch_bounds__(BLOCK) void k_copy(int64_t n2, const double2 *__restrict__ src,
                                                double2 *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n2) dst[i] = src[i];
}

// batched flux-scheme plugin: F = solverAll(dir) after set(QL, QR)
template <int DIM, int SCHEME>
__global__ __launch_bounds__(BLOCK) void k_riemann(int64_t n, const double *__restrict__ QL,
                                                   const double *__restrict__ QR, const int32_t *__restrict__ dir,
                                                   double *__restrict__ F, DevParams P) {
    constexpr int NV = DIM + 2;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double A[NV], B[NV], nrm[DIM], g[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        A[k] = QL[i * NV + k];
        B[k] = QR[i * NV + k];
    }
    const int d0 = dir[i];
#pragma unroll
    for (int d = 0; d < DIM; d++) nrm[d] = d == d0 ? 1.0 : 0.0;
    riemann_face<DIM, SCHEME>(A, B, nrm, (1 << DIM) - 1, P, g);
#pragma unroll
    for (int k = 0; k < NV; k++) F[i * NV + k] = g[k];
}

DevParams make_params(const mst_config &c) {
    DevParams p;
    p.gamma = c.gamma;
    p.gm1 = c.gamma - 1.0;
    p.eor = c.eor;
    p.delta = c.entropy_delta;
    p.dd = c.entropy_delta * c.entropy_delta;
    p.ihd = 0.5 / c.entropy_delta;
    p.astar_c = 2.0 * (c.gamma - 1.0) / (c.gamma + 1.0);
    p.compat = c.compat;
    for (int k = 0; k < 5; k++) p.qin[k] = c.q_inlet[k];
    return p;
}

int check_config(const mst_config *c, int dim) {
    if (!c) return mst::fail(MST_E_ARG, "null config");
    if (c->scheme != MST_SCHEME_ROE && c->scheme != MST_SCHEME_AUSM) return mst::fail(MST_E_ARG, "bad scheme");
    if (c->order != 1 && c->order != 2) return mst::fail(MST_E_ARG, "order must be 1 or 2");
    if (c->limiter < MST_LIM_NONE || c->limiter > MST_LIM_VANLEER) return mst::fail(MST_E_ARG, "bad limiter");
    if (c->rk < 1 || c->rk > 3) return mst::fail(MST_E_ARG, "rk must be 1, 2 or 3");
    if (!(c->gamma > 1) || !(c->entropy_delta > 0) || c->eor < 0 || c->cfl < 0)
        return mst::fail(MST_E_ARG, "bad gamma / entropy_delta / eor / cfl");
    if (c->q_inlet[0] <= 0 && dim > 0) return mst::fail(MST_E_ARG, "inlet density must be positive");
    return MST_OK;
}

}  // namespace

// ====================================================================================== context
struct mst_ctx {
    mst_prep prep;
    mst_config cfg;
    DevParams dp;
    hipStream_t stream = nullptr;
    int device = 0;
    int dim = 2, nv = 4, maxf = 3;
    int64_t nc = 0, nf = 0, nblk_c = 0, nblk_f = 0;
    int32_t *fc0 = nullptr, *fc1 = nullptr, *finfo = nullptr, *ell = nullptr;
    double *fn = nullptr, *fxf = nullptr, *pxf = nullptr, *feta = nullptr;
    double *cvol = nullptr, *cxc = nullptr, *cshalf = nullptr;
    double *Q[4] = {nullptr, nullptr, nullptr, nullptr};
    int iold = 0, inew = 1;
    double *G = nullptr, *lg = nullptr, *graw = nullptr, *phi = nullptr, *R = nullptr;
    double *scal = nullptr;  // [0] dt, [1..5] l2 norms, [6..10] maxrel
    double *part = nullptr;
    bool dt_valid = false;            // scal[0] holds the CFL time step of the current Old state
    int *bad = nullptr;
    bool timing = false;
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

template <typename K, typename... Args>
int launch(mst_ctx *x, int kclass, K kernel, int64_t nthreads, Args... args) {
    if (nthreads <= 0) return MST_OK;
    const int64_t grid = (nthreads + BLOCK - 1) / BLOCK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (x->timing && kclass >= 0) {
        e0 = take_event(x);
        e1 = take_event(x);
        if (e0) (void)hipEventRecord(e0, x->stream);
    }
    hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(BLOCK), 0, x->stream, args...);
    HIPCHK(hipGetLastError());
    if (e0 && e1) {
        (void)hipEventRecord(e1, x->stream);
        x->ev[kclass].push_back({e0, e1});
    }
    return MST_OK;
}

template <int DIM>
int run_grad_t(mst_ctx *x, const double *Qsrc, bool debug) {
    GradArgs a;
    a.nc = x->nc;
    a.ell = x->ell;
    a.fc0 = x->fc0;
    a.fc1 = x->fc1;
    a.fn = x->fn;
    a.fxf = x->fxf;
    a.feta = x->feta;
    a.Q = Qsrc;
    a.cxc = x->cxc;
    a.vol = x->cvol;
    a.lg = x->lg;
    a.graw = debug ? x->graw : nullptr;
    a.phi = debug ? x->phi : nullptr;
    const int lim = x->cfg.limiter;
#define MST_GRAD(MF)                                                                             \
    if (lim == MST_LIM_NONE) return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_NONE>, x->nc, a);     \
    if (lim == MST_LIM_MINMOD) return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_MINMOD>, x->nc, a); \
    return launch(x, MST_K_GRAD, k_grad<DIM, MF, MST_LIM_VANLEER>, x->nc, a);
    if (x->maxf <= 3) { MST_GRAD(3) }
    if (x->maxf <= 4) { MST_GRAD(4) }
    MST_GRAD(8)
#undef MST_GRAD
}

int run_grad(mst_ctx *x, const double *Qsrc, bool debug) {
    return x->dim == 2 ? run_grad_t<2>(x, Qsrc, debug) : run_grad_t<3>(x, Qsrc, debug);
}

template <int DIM>
int run_flux_t(mst_ctx *x, const double *Qsrc) {
    FluxArgs a;
    a.nf = x->nf;
    a.pstart = x->prep.P.pstart;
    a.fc0 = x->fc0;
    a.fc1 = x->fc1;
    a.finfo = x->finfo;
    a.fn = x->fn;
    a.fxf = x->fxf;
    a.pxf = x->pxf;
    a.Q = Qsrc;
    a.cxc = x->cxc;
    a.lg = x->lg;
    a.G = x->G;
    a.P = x->dp;
    const bool second = x->cfg.order >= 2;
    if (x->cfg.scheme == MST_SCHEME_ROE)
        return second ? launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_ROE, true>, x->nf, a)
                      : launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_ROE, false>, x->nf, a);
    return second ? launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_AUSM, true>, x->nf, a)
                  : launch(x, MST_K_FLUX, k_flux<DIM, MST_SCHEME_AUSM, false>, x->nf, a);
}

// residual evaluation up to the face fluxes: gradient + limiter (order 2) then the flux kernel
int run_face_flux(mst_ctx *x, const double *Qsrc) {
    if (x->cfg.order >= 2) {
        int rc = run_grad(x, Qsrc, false);
        if (rc) return rc;
    }
    return x->dim == 2 ? run_flux_t<2>(x, Qsrc) : run_flux_t<3>(x, Qsrc);
}

template <int NV>
int run_update_t(mst_ctx *x, int mode, const double *Qs, const double *Qn, double ca, double cb, int stage0,
                 double *Qout) {
    UpdArgs a;
    a.nc = x->nc;
    a.ell = x->ell;
    a.G = x->G;
    a.Qs = Qs;
    a.Qn = Qn;
    a.vol = x->cvol;
    a.sh = x->cshalf;
    a.dt = x->scal;
    a.ca = ca;
    a.cb = cb;
    a.stage0 = stage0;
    a.gamma = x->cfg.gamma;
    a.cfl = x->cfg.cfl;
    a.Qout = Qout;
    a.R = x->R;
    a.part_l2 = x->part;
    a.part_dt = x->part + x->nblk_c * x->nv;
    a.bad = x->bad;
#define MST_UPD(MF)                                                                                          \
    switch (mode) {                                                                                          \
        case UPD_RESIDUAL: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_RESIDUAL>, x->nc, a);         \
        case UPD_STAGE: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE>, x->nc, a);               \
        case UPD_STAGE | UPD_L2: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_L2>, x->nc, a); \
        case UPD_STAGE | UPD_DT: return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_DT>, x->nc, a); \
        default:                                                                                             \
            return launch(x, MST_K_UPDATE, k_update<NV, MF, UPD_STAGE | UPD_L2 | UPD_DT>, x->nc, a);         \
    }
    if (x->maxf <= 3) { MST_UPD(3) }
    if (x->maxf <= 4) { MST_UPD(4) }
    MST_UPD(8)
#undef MST_UPD
}

int run_update(mst_ctx *x, int mode, const double *Qs, const double *Qn, double ca, double cb, int stage0,
               double *Qout) {
    return x->nv == 4 ? run_update_t<4>(x, mode, Qs, Qn, ca, cb, stage0, Qout)
                      : run_update_t<5>(x, mode, Qs, Qn, ca, cb, stage0, Qout);
}

// one finish launch: L2 norms (nl2 fields) and/or the minimum time step, from the per-block partials
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
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN), 0, x->stream, f);
    HIPCHK(hipGetLastError());
    return MST_OK;
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

int upload_config(mst_ctx *x) {
    std::vector<int32_t> mask;
    mst::lr_masks(x->prep.P, x->cfg.compat, mask);
    std::vector<int32_t> info(x->nf);
    for (int64_t i = 0; i < x->nf; i++) info[i] = x->prep.P.fkind[i] | (mask[i] << 4);
    HIPCHK(hipMemcpy(x->finfo, info.data(), sizeof(int32_t) * x->nf, hipMemcpyHostToDevice));
    std::vector<double> eta = x->prep.P.feta;
    if ((x->cfg.compat & MST_COMPAT_GRAD_LASTINT) && x->prep.P.quirk_face >= 0) eta[x->prep.P.quirk_face] = 1.0;
    HIPCHK(hipMemcpy(x->feta, eta.data(), sizeof(double) * x->nf, hipMemcpyHostToDevice));
    x->dp = make_params(x->cfg);
    return MST_OK;
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

void release(mst_ctx *x) {
    if (!x) return;
    void *bufs[] = {x->fc0, x->fc1, x->finfo, x->ell, x->fn, x->fxf, x->pxf, x->feta, x->cvol, x->cxc,
                    x->cshalf, x->Q[0], x->Q[1], x->Q[2], x->Q[3], x->G, x->lg, x->graw, x->phi, x->R,
                    x->scal, x->part, x->bad};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
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
        return mst::fail(MST_E_STATE, "non-physical state (rho<=0, p<=0 or NaN) produced by the update");
    }
    return MST_OK;
}

}  // namespace

#define CTX_GUARD(x)                                           \
    if (!(x)) return mst::fail(MST_E_ARG, "null context");

extern "C" {

int mst_ctx_create(const mst_mesh *m, const mst_config *cfg, mst_ctx **out) {
    if (!m || !out) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    int rc = check_config(cfg, 2);
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
    TRY(dupload(&x->ell, P.ell));
    TRY(dupload(&x->fn, P.fn));
    TRY(dupload(&x->fxf, P.fxf));
    TRY(dupload(&x->pxf, P.pxf));
    TRY(dalloc(&x->feta, x->nf));
    TRY(dupload(&x->cvol, P.cvol));
    TRY(dupload(&x->cxc, P.cxc));
    TRY(dupload(&x->cshalf, P.cshalf));
    for (int i = 0; i < 4; i++) TRY(dalloc(&x->Q[i], x->nc * x->nv));
    TRY(dalloc(&x->G, x->nf * x->nv));
    TRY(dalloc(&x->lg, x->nc * x->nv * x->dim));
    TRY(dalloc(&x->scal, 16));
    TRY(dalloc(&x->part, std::max(x->nblk_c, x->nblk_f) * (x->nv + 1) + 16));
    TRY(dalloc(&x->bad, 1));
    if (hipMemset(x->bad, 0, sizeof(int)) != hipSuccess || hipMemset(x->scal, 0, 16 * sizeof(double)) != hipSuccess ||
        hipMemset(x->lg, 0, sizeof(double) * x->nc * x->nv * x->dim) != hipSuccess) {
        release(x);
        return mst::fail(MST_E_HIP, "hipMemset failed");
    }
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
    int rc = check_config(cfg, ctx->dim);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cfg = *cfg;
    ctx->dt_valid = false;
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
    return run_face_flux(ctx, ctx->Q[ctx->iold]);
}

int mst_residual_update(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    int rc = run_update(ctx, UPD_STAGE, ctx->Q[ctx->iold], ctx->Q[ctx->iold], 0.0, 1.0, 1, ctx->Q[ctx->inew]);
    if (rc) return rc;
    return check_bad(ctx);
}

int mst_solve(mst_ctx *ctx) {
    CTX_GUARD(ctx);
    int rc = run_face_flux(ctx, ctx->Q[ctx->iold]);
    if (rc) return rc;
    return mst_residual_update(ctx);
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
        // stage reduces |R/V|, the last one the CFL time step of the new state (next step's dt)
        int spare[2], w = 0;
        for (int i = 0; i < 4; i++)
            if (i != ctx->iold && i != ctx->inew) spare[w++] = i;
        const double *src = Qn;
        for (int s = 0; s < ns; s++) {
            double *dst = (s == ns - 1) ? ctx->Q[ctx->inew] : ctx->Q[spare[s & 1]];
            rc = run_face_flux(ctx, src);
            if (rc) return rc;
            int mode = UPD_STAGE | (s == 0 ? UPD_L2 : 0) | ((cfl && s == ns - 1) ? UPD_DT : 0);
            rc = run_update(ctx, mode, src, Qn, A[s], B[s], s == 0 ? 1 : 0, dst);
            if (rc) return rc;
            src = dst;
        }
        rc = run_finish(ctx, ctx->nv, cfl ? 1 : 0);
        if (rc) return rc;
        ctx->dt_valid = cfl;
        std::swap(ctx->iold, ctx->inew);
    }
    return check_bad(ctx);
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
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN), 0, ctx->stream, f);
    HIPCHK(hipGetLastError());
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, ctx->scal + 6, sizeof(double) * ctx->nv, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_get_residual(mst_ctx *ctx, double *R, int64_t n_cells) {
    CTX_GUARD(ctx);
    if (!R || n_cells != ctx->nc) return mst::fail(MST_E_ARG, "size does not match the mesh");
    if (!ctx->R) {
        int rc = dalloc(&ctx->R, ctx->nc * ctx->nv);
        if (rc) return rc;
    }
    int rc = run_face_flux(ctx, ctx->Q[ctx->iold]);
    if (rc) return rc;
    rc = run_update(ctx, UPD_RESIDUAL, nullptr, nullptr, 0, 0, 0, nullptr);
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
    for (int64_t f = 0; f < ctx->nf; f++)
        for (int k = 0; k < nv; k++) G[f * nv + k] = buf[(int64_t)new_of_old[f] * nv + k];
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
    int rc = run_grad(ctx, ctx->Q[ctx->iold], true);
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
    int rc = check_config(cfg, dim);
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
    DevParams P = make_params(*cfg);
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
    ctx->timing = on != 0;
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

// Compulsory HBM traffic of one launch: every array the kernel must touch, once.
int mst_algorithmic_bytes(const mst_ctx *ctx, int kernel, double *bytes) {
    CTX_GUARD(ctx);
    if (!bytes) return mst::fail(MST_E_ARG, "null argument");
    const double nc = (double)ctx->nc, nf = (double)ctx->nf, D = ctx->dim, NV = ctx->nv;
    const double ell = 4.0 * ctx->maxf * nc;
    const bool second = ctx->cfg.order >= 2;
    switch (kernel) {
        case MST_K_FLUX:  // c0,c1,info + normal + G out per face; Q per cell (+ centres, limited gradient)
            *bytes = nf * (12 + 8 * D + 8 * NV) + nc * 8 * NV;
            if (second) *bytes += nf * 8 * D + nc * (8 * D + 8 * NV * D) + (double)ctx->prep.P.pcount * 8 * D;
            break;
        case MST_K_GRAD:  // per face c0,c1, normal, centre, eta; per cell ell, Q, centre, V, limited gradient out
            *bytes = nf * (8 + 16 * D + 8) + ell + nc * (8 * NV + 8 * D + 8 + 8 * NV * D);
            break;
        case MST_K_UPDATE: {  // ell, G per face, Q_stage, V, Q_out; Q^n on stages after the first
            const double stages = ctx->cfg.rk;
            *bytes = ell + nf * 8 * NV + nc * (8 * NV + 8 + 8 * NV) + nc * 8 * NV * (stages - 1) / stages;
            break;
        }
        case MST_K_DT:
            *bytes = nc * (8 * NV + 8 + 8 * D);
            break;
        default:
            return mst::fail(MST_E_ARG, "bad kernel id");
    }
    return MST_OK;
}

}  // extern "C"
