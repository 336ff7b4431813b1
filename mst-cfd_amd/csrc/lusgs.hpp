// Steady driver: matrix-free LU-SGS over a cell colouring.
//
// The factorised operator (D+L) D^-1 (D+U) -- a forward sweep with the lower part, a division by D, a
// backward sweep with the upper part -- is the one SparseSolverNUM::solveILUSGS applies to its
// assembled scalar systems (lusolver/SparseSolverNUM.cpp:144-196).  Here the blocks are never
// assembled; they come from the Rusanov split of the face-flux Jacobian (Jameson & Yoon, AIAA J.
// 25(7), 1987):
//   r_f  = max over the face's cells of |u.S_f| + a |S_f|        (boundary face: its one cell)
//   D_i  = (1/CFL + omega/2) sum_f r_f                            (local step CFL V_i / sum_f r_f)
//   (i,j) block times dQ_j = 1/2 [F_n(Q_j + dQ_j) - F_n(Q_j)] - 1/2 omega r_f dQ_j,
//        n = the outward normal of cell i at f, F_n = sum_d n_d F_d (physical Euler flux)
// L / U hold the neighbours of lower / higher colour; boundary faces enter D only.
//
// Sequential Gauss-Seidel order is replaced by colour order: no two cells of one colour are
// neighbours, so one launch per colour updates all of its cells at once, and the result is the same
// as sweeping the cells one by one in colour order (the order inside a colour does not matter).
//   k_lu_radius  one lane per face: r_f
//   k_lu_sweep   one lane per cell of one colour; forward: dQ* = (-R - sum_L ...) / D;
//                backward: dQ = dQ* - sum_U ... / D, Q_new = Q + dQ (and the admissibility check)
// Both sweeps are memory-bound gathers over the neighbours of a cell, like the residual's gather.
#pragma once
#include "device_common.hpp"

namespace mstk {

struct LuArgs {
    int64_t nc;
    const int32_t *__restrict__ ell;     // slot-major: (face << 1) | is_c1, -1 past the last face
    const int32_t *__restrict__ fc0;
    const int32_t *__restrict__ fc1;     // neighbour (interior / periodic partner cell), -1 boundary
    const double *__restrict__ fn;       // outward normal of c0 (area-weighted)
    const int32_t *__restrict__ colour;  // per cell
    const double *__restrict__ rf;       // per face spectral radius
    const double *__restrict__ Q;        // state at the start of the iteration
    const double *__restrict__ R;        // its residual
    double *__restrict__ dQ;
    double *__restrict__ Qout;
    int *__restrict__ bad;
    double gamma, dscale, omega;         // dscale = 1/CFL + omega/2
};

template <int DIM>
__device__ __forceinline__ double lu_pressure(const double *q, double gamma) {
    constexpr int NV = DIM + 2;
    double ke = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) ke += q[1 + d] * q[1 + d];
    return (q[NV - 1] - 0.5 * ke / q[0]) * (gamma - 1);
}

template <int DIM>
__device__ __forceinline__ double lu_radius(const double *q, const double *n, double gamma) {
    double un = 0, s2 = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        un += q[1 + d] * n[d];
        s2 += n[d] * n[d];
    }
    const double p = lu_pressure<DIM>(q, gamma);
    return fabs(un / q[0]) + sqrt(gamma * p / q[0]) * sqrt(s2);
}

template <int DIM>
__device__ __forceinline__ void lu_flux(const double *q, const double *n, double gamma, double *F) {
    constexpr int NV = DIM + 2;
    double un = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) un += q[1 + d] * n[d];
    un /= q[0];
    const double p = lu_pressure<DIM>(q, gamma);
    F[0] = q[0] * un;
#pragma unroll
    for (int d = 0; d < DIM; d++) F[1 + d] = q[1 + d] * un + p * n[d];
    F[NV - 1] = (q[NV - 1] + p) * un;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_lu_radius(const LuArgs a, int64_t nf, double *__restrict__ rf) {
    constexpr int NV = DIM + 2;
    const int64_t f = block_id() * BLOCK + threadIdx.x;
    if (f >= nf) return;
    double n[DIM], q[NV];
#pragma unroll
    for (int d = 0; d < DIM; d++) n[d] = a.fn[f * DIM + d];
    ldq<NV>(a.Q + (int64_t)a.fc0[f] * NV, q);
    double r = lu_radius<DIM>(q, n, a.gamma);
    const int32_t c1 = a.fc1[f];
    if (c1 >= 0) {
        ldq<NV>(a.Q + (int64_t)c1 * NV, q);
        r = fmax(r, lu_radius<DIM>(q, n, a.gamma));
    }
    rf[f] = r;
}

// one colour's cells (cells[0..n)); FWD: forward sweep into dQ, else backward sweep, dQ and Qout
template <int DIM, bool FWD, int MAXF>
__global__ __launch_bounds__(BLOCK) void k_lu_sweep(const LuArgs a, const int32_t *__restrict__ cells, int64_t n,
                                                    int col) {
    constexpr int NV = DIM + 2;
    const int64_t t = block_id() * BLOCK + threadIdx.x;
    if (t >= n) return;
    const int64_t c = cells[t];
    double s[NV];
    if constexpr (FWD) {
        ldq<NV>(a.R + c * NV, s);
#pragma unroll
        for (int k = 0; k < NV; k++) s[k] = -s[k];
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) s[k] = 0;
    }
    double rsum = 0;
#pragma unroll
    for (int slot = 0; slot < MAXF; slot++) {
        const int32_t e = a.ell[(int64_t)slot * a.nc + c];
        if (e < 0) break;
        const int64_t f = e >> 1;
        const bool c1side = e & 1;
        const double r = a.rf[f];
        rsum += r;
        const int32_t o = c1side ? a.fc0[f] : a.fc1[f];
        if (o < 0 || o == c) continue;
        const int32_t oc = a.colour[o];
        if (FWD ? oc < col : oc > col) {
            double nrm[DIM], q[NV], dq[NV], qp[NV], F1[NV], F0[NV];
#pragma unroll
            for (int d = 0; d < DIM; d++) nrm[d] = c1side ? -a.fn[f * DIM + d] : a.fn[f * DIM + d];
            ldq<NV>(a.Q + (int64_t)o * NV, q);
            ldq<NV>(a.dQ + (int64_t)o * NV, dq);
#pragma unroll
            for (int k = 0; k < NV; k++) qp[k] = q[k] + dq[k];
            lu_flux<DIM>(qp, nrm, a.gamma, F1);
            lu_flux<DIM>(q, nrm, a.gamma, F0);
#pragma unroll
            for (int k = 0; k < NV; k++) s[k] -= 0.5 * ((F1[k] - F0[k]) - a.omega * r * dq[k]);
        }
    }
    const double D = a.dscale * rsum;
    double out[NV];
    if constexpr (FWD) {
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] = s[k] / D;
        stq<NV>(a.dQ + c * NV, out);
    } else {
        double dq[NV], q[NV];
        ldq<NV>(a.dQ + c * NV, dq);
        ldq<NV>(a.Q + c * NV, q);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            dq[k] += s[k] / D;
            out[k] = q[k] + dq[k];
        }
        stq<NV>(a.dQ + c * NV, dq);
        stq<NV>(a.Qout + c * NV, out);
        const double p = lu_pressure<DIM>(out, a.gamma);
        if (!(out[0] > 0) || !(p > 0)) a.bad[0] = 1;
    }
}

}  // namespace mstk
