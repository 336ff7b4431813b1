// riemann.hpp -- dimension-split Roe and AUSM+ face fluxes (device), real type T.
//
// Both return F[d][k], the flux of the 1D Riemann problem along axis d (rhoSolver/RhoSolver.cpp:
// 104-112) with the left state chosen per axis by bit d of `mask` (bit set: A, the c0/interior side,
// is the left state).  The caller projects G = sum_d n_d F_d with FP64 normals.
#pragma once
#include "device_common.hpp"

namespace mstk {

// Roe (SolverRoe.cpp:3-16,70-112).  The Roe average is computed once per face (it is symmetric in
// L/R); per axis d the dissipation R|Lambda|R^-1 (Q_R - Q_L) uses the analytic wave strengths of the
// reference's eigenvector matrix (same columns as SolverRoe.cpp:86,92; exact inverse including the
// (H - q^2/2) denominator the explicit 4x4 inverse produces) and changes sign when the ghost/c1 side
// is the Riemann left state.
template <int DIM, typename T>
__device__ __forceinline__ void roe_face(const T *A, const T *B, int mask, const Params<T> &P, T (&F)[DIM][DIM + 2]) {
    constexpr int NV = DIM + 2;
    const T h = T(0.5);
    const T iA = frcp(A[0]), iB = frcp(B[0]);
    T mA2 = 0, mB2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        mA2 += A[1 + k] * A[1 + k];
        mB2 += B[1 + k] * B[1 + k];
    }
    const T pA = (A[NV - 1] - h * mA2 * iA) * P.gm1;
    const T pB = (B[NV - 1] - h * mB2 * iB) * P.gm1;
    const T HA = (A[NV - 1] + pA) * iA, HB = (B[NV - 1] + pB) * iB;
    const T D = fsqrt(fabs(B[0] * iA));
    const T i1D = frcp(T(1) + D);
    T u[DIM], q2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        u[k] = (A[1 + k] * iA + D * (B[1 + k] * iB)) * i1D;
        q2 += u[k] * u[k];
    }
    const T Ht = (HA + D * HB) * i1D;
    const T h2 = h * q2;
    const T a = fsqrt(fabs(P.gm1 * (Ht - h2)));
    const T ia = frcp(a), iden = fdiv(T(1), Ht - h2);
    const T dr = B[0] - A[0], dE = B[NV - 1] - A[NV - 1];
    T beta[DIM], udm = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        const T dm = B[1 + k] - A[1 + k];
        udm += u[k] * dm;
        beta[k] = dm - u[k] * dr;
    }
    const T s = (dE - udm + h2 * dr) * iden;  // alpha(-) + alpha(+)
    const T a0 = dr - s;                      // entropy wave strength
    const T ieA = frcp(A[0] + P.eor), ieB = frcp(B[0] + P.eor);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const T bd = beta[d] * ia;
        const T ap = h * (s + bd), am = h * (s - bd);
        const T lm = hfix(fabs(u[d] - a), P), l0 = hfix(fabs(u[d]), P), lp = hfix(fabs(u[d] + a), P);
        const T wm = lm * am, wp = lp * ap, w0 = l0 * a0;
        const T sr = wm + w0 + wp;
        T Dv[NV];
        Dv[0] = sr;
        T esh = 0;
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            if (k == d) {
                Dv[1 + k] = u[k] * sr + a * (wp - wm);
            } else {
                const T t = l0 * beta[k];
                Dv[1 + k] = u[k] * sr + t;
                esh += u[k] * t;
            }
        }
        Dv[NV - 1] = Ht * (wm + wp) + u[d] * a * (wp - wm) + h2 * w0 + esh;
        const T sg = ((mask >> d) & 1) ? h : -h;
        const T mAd = A[1 + d], mBd = B[1 + d];
        F[d][0] = h * (mAd + mBd) - sg * Dv[0];
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            T fa = A[1 + k] * mAd * ieA, fb = B[1 + k] * mBd * ieB;
            if (k == d) {
                fa += pA;
                fb += pB;
            }
            F[d][1 + k] = h * (fa + fb) - sg * Dv[1 + k];
        }
        F[d][NV - 1] = h * (HA * mAd + HB * mBd) - sg * Dv[NV - 1];
    }
}

// AUSM+ with the reference's quirks (SolverAusm.cpp:3-26,53-63,105-143)
template <int DIM, typename T>
__device__ __forceinline__ void ausm_face(const T *A, const T *B, int mask, const Params<T> &P, T (&F)[DIM][DIM + 2]) {
    constexpr int NV = DIM + 2;
    const T h = T(0.5), q4 = T(0.25), e8 = T(0.125), c316 = T(0.1875), one = T(1), two = T(2);
    const T iA = frcp(A[0]), iB = frcp(B[0]);
    T mA2 = 0, mB2 = 0;
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        mA2 += A[1 + k] * A[1 + k];
        mB2 += B[1 + k] * B[1 + k];
    }
    const T pA = (A[NV - 1] - h * mA2 * iA) * P.gm1;
    const T pB = (B[NV - 1] - h * mB2 * iB) * P.gm1;
    const T HA = (A[NV - 1] + pA) * iA, HB = (B[NV - 1] + pB) * iB;
    const T asA = fsqrt(HA * P.astar_c), asB = fsqrt(HB * P.astar_c);
    const T UA = fsqrt(mA2 * iA * iA), UB = fsqrt(mB2 * iB * iB);
    const T tA = asA * asA * frcp(fmax(asA, UA)), tB = asB * asB * frcp(fmax(asB, UB));
    const T aF = fmin(tA, tB);
    const T iaF = frcp(aF);
    T aA = aF, aB = aF;
    if (P.compat & MST_COMPAT_AUSM_LOCAL_A) {
        aA = fsqrt(P.gamma * pA * iA);
        aB = fsqrt(P.gamma * pB * iB);
    }
    const bool unscaled = (P.compat & MST_COMPAT_AUSM_P_UNSCALED) != 0;
    const bool onesided = (P.compat & MST_COMPAT_AUSM_ONESIDED) != 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const bool left_is_A = (mask >> d) & 1;
        const T *L = left_is_A ? A : B, *R = left_is_A ? B : A;
        const T iL = left_is_A ? iA : iB, iR = left_is_A ? iB : iA;
        const T pL = left_is_A ? pA : pB, pR = left_is_A ? pB : pA;
        const T aL = left_is_A ? aA : aB, aR = left_is_A ? aB : aA;
        const T mL = L[1 + d] * iL * iaF, mR = R[1 + d] * iR * iaF;
        T Mp, Mm, Pp, Pm;
        const T qL = mL * mL - one, qR = mR * mR - one;
        // SolverAusm.cpp:116,122 test M <= 1 (compat); AUSM+ tests |M| <= 1
        const bool subL = onesided ? mL <= one : fabs(mL) <= one;
        const bool subR = onesided ? mR <= one : fabs(mR) <= one;
        if (subL) {
            Mp = q4 * (mL + one) * (mL + one) + e8 * qL * qL;
            const T poly = q4 * (mL + one) * (mL + one) * (two - mL);
            const T corr = c316 * mL * qL * qL;
            Pp = unscaled ? pL * poly + corr : pL * (poly + corr);
        } else {
            Mp = h * (mL + fabs(mL));
            Pp = fdiv(pL * h * (mL + fabs(mL)), mL);
        }
        if (subR) {
            Mm = -q4 * (mR - one) * (mR - one) - e8 * qR * qR;
            const T poly = q4 * (mR - one) * (mR - one) * (two + mR);
            const T corr = c316 * mR * qR * qR;
            Pm = unscaled ? pR * poly - corr : pR * (poly - corr);
        } else {
            Mm = h * (mR - fabs(mR));
            Pm = fdiv(pR * h * (mR - fabs(mR)), mR);
        }
        const T Mf = Mm + Mp, pf = Pm + Pp, aMf = fabs(Mf);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const T psL = (k == NV - 1) ? L[k] + pL : L[k];
            const T psR = (k == NV - 1) ? R[k] + pR : R[k];
            F[d][k] = h * (Mf * (aL * psL + aR * psR) - aMf * (aR * psR - aL * psL));
        }
        F[d][1 + d] += pf;
    }
}

// G_k = sum_d n_d F_d,k in FP64 (exact freestream cancellation over a closed cell in both forms)
template <int DIM, int SCHEME, typename T>
__device__ __forceinline__ void riemann_face(const T *A, const T *B, const double *n, int mask, const Params<T> &P,
                                             double *g) {
    constexpr int NV = DIM + 2;
    T F[DIM][NV];
    if constexpr (SCHEME == MST_SCHEME_ROE)
        roe_face<DIM, T>(A, B, mask, P, F);
    else
        ausm_face<DIM, T>(A, B, mask, P, F);
#pragma unroll
    for (int k = 0; k < NV; k++) g[k] = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++)
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] += n[d] * (double)F[d][k];
}

}  // namespace mstk
