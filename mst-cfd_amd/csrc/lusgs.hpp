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
//   k_lu_sweep   one lane per cell of one colour; forward: r_f of the cell's faces, D, and
//                dQ* = (-R - sum_L ...) / D; backward: dQ = dQ* - sum_U ... / D, Q_new = Q + dQ
//                (and the admissibility check)
// Both sweeps are memory-bound gathers over the neighbours of a cell, like the residual's gather.
#pragma once
#include "kernels.hpp"

namespace mstk {

// Cells are addressed by their position t in the colour-sorted list (colour k: positions
// [start_k, start_{k+1})); per (slot, position), slot-major, the sweeps read precomputed records so a
// neighbour costs one round of dependent loads (its state and update) instead of walking
// cell -> face -> neighbour -> colour.
struct LuArgs {
    int64_t nc;
    const int32_t *__restrict__ cells;   // position -> cell
    const int32_t *__restrict__ nb;      // (neighbour << 2) | 1 (lower colour) | 2 (higher colour);
                                         // -1 boundary face, -2 past the last face
    const double *__restrict__ n;        // outward normal of the cell at the slot (area-weighted)
    double *__restrict__ r;              // spectral radius of the slot's face (per iteration)
    double *__restrict__ D;              // diagonal (per iteration)
    const double *__restrict__ Q;        // state at the start of the iteration (cell order)
    const double *__restrict__ R;        // its residual (cell order)
    double *__restrict__ dQ;             // update (cell order)
    double *__restrict__ Qout;
    int *__restrict__ bad;
    double gamma, dscale, omega;         // dscale = 1/CFL + omega/2
    // first-order residual computed by the forward sweep itself (RSCHEME >= 0)
    const int32_t *__restrict__ face;    // (face << 1) | is_c1 per (slot, position)
    const int32_t *__restrict__ finfo;   // per face: kind | (L/R mask << 4)
    Params<double> P;
    // batches (pcase set): the case of each position, per case gamma and active flag (inactive cases
    // are skipped; the bad flag is per case); the residual is gathered from the face fluxes G
    const int32_t *__restrict__ pcase;
    const double *__restrict__ gam;
    const uint8_t *__restrict__ active;
    const double *__restrict__ G;
};

// residual sources of the forward sweep
constexpr int LU_R_ARRAY = -1;   // read R
constexpr int LU_R_GATHER = -2;  // sum +-G_f of the cell's faces in slot order (the gather kernels' order)

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

// One colour's positions [t0, t0 + n).
// FWD: the spectral radius of each of the cell's faces (max over the face's two cells: the same value
// from either side, exactly, since it depends on the normal only through |u.n| and |n|) and D are
// computed and stored for the backward sweep, then dQ* = (-R - sum_L ...) / D.  The last colour has
// no higher neighbours, so its backward sweep is the identity: LAST also writes Qout.
// Backward (!FWD): dQ = dQ* - sum_U ... / D, Qout = Q + dQ.
// RSCHEME = LU_R_GATHER (forward, cells of <= 4 faces): R is gathered from the face fluxes G.
// RSCHEME >= 0 (forward, order 1, cells of <= 4 faces): R of the cell is not read but computed here,
// from the states the sweep loads anyway: the same face solves (face_solve, the c0 outward normal,
// the c0 side first) summed in slot order as the gather kernels sum them, so R is the same bit for bit.
// One cell's sweep (position t, cell c).  qown (forward only): the cell's state, already in registers,
// in place of a.Q[c]; qnew (backward or LAST): receives the new state the sweep also stores in Qout.
template <int DIM, bool FWD, bool LAST, int MAXF, int RSCHEME>
__device__ __forceinline__ void lu_cell(const LuArgs &a, const int64_t t, const int64_t c, const double *qown,
                                        double *qnew) {
    static_assert(RSCHEME == LU_R_ARRAY || (FWD && MAXF <= 4), "computed residual: forward sweeps of small cells");
    constexpr int NV = DIM + 2;
    double s[NV], qc[NV];
    if constexpr (FWD) {
        if constexpr (RSCHEME == LU_R_ARRAY) {
            ldq<NV>(a.R + c * NV, s);
#pragma unroll
            for (int k = 0; k < NV; k++) s[k] = -s[k];
        }
        if (qown) {
#pragma unroll
            for (int k = 0; k < NV; k++) qc[k] = qown[k];
        } else {
            ldq<NV>(a.Q + c * NV, qc);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) s[k] = 0;
    }
    double dqc[NV], D = 0;  // backward: the cell's own update, state and diagonal, read up front
    if constexpr (!FWD) {
        ldq<NV>(a.dQ + c * NV, dqc);
        ldq<NV>(a.Q + c * NV, qc);
        D = a.D[t];
    }
    constexpr int32_t SIDE = FWD ? 1 : 2;
    double rsum = 0;
    if constexpr (MAXF <= 4) {
        // every memory read of the cell's slots is issued before the first is used: two rounds of
        // dependent loads per cell (records, then neighbour states and updates) instead of two per
        // slot.  The arithmetic below runs in the same order as the slot loop further down.
        int32_t code[MAXF], fcode[MAXF], info[MAXF];
        double nrm[MAXF][DIM], q[MAXF][NV], dq[MAXF][NV], rr[MAXF];
#pragma unroll
        for (int k = 0; k < MAXF; k++) code[k] = a.nb[(int64_t)k * a.nc + t];
        if constexpr (RSCHEME != LU_R_ARRAY) {
#pragma unroll
            for (int k = 0; k < MAXF; k++) {
                fcode[k] = a.face[(int64_t)k * a.nc + t];
                if constexpr (RSCHEME >= 0) info[k] = fcode[k] >= 0 ? a.finfo[fcode[k] >> 1] : 0;
            }
        }
        double gsl[MAXF][NV];  // gathered face fluxes (RSCHEME == LU_R_GATHER)
        if constexpr (RSCHEME == LU_R_GATHER) {
#pragma unroll
            for (int k = 0; k < MAXF; k++)
                if (fcode[k] >= 0) ldq<NV>(a.G + (int64_t)(fcode[k] >> 1) * NV, gsl[k]);
        }
#pragma unroll
        for (int k = 0; k < MAXF; k++) {
            const int64_t i = (int64_t)k * a.nc + t;
            const bool nbr = code[k] >= 0, use = nbr && (code[k] & SIDE);
            const int64_t o = code[k] >> 2;
            if (FWD ? code[k] != -2 : use) {
#pragma unroll
                for (int d = 0; d < DIM; d++) nrm[k][d] = a.n[i * DIM + d];
            }
            if (FWD ? nbr : use) ldq<NV>(a.Q + o * NV, q[k]);
            if (use) ldq<NV>(a.dQ + o * NV, dq[k]);
            if (!FWD && use) rr[k] = a.r[i];
        }
        if constexpr (RSCHEME == LU_R_GATHER) {
            double r[NV], cmp[NV];
#pragma unroll
            for (int j = 0; j < NV; j++) r[j] = cmp[j] = 0;
#pragma unroll
            for (int k = 0; k < MAXF; k++) {
                if (code[k] == -2) break;
                accumulate<NV, false>(r, cmp, gsl[k], fcode[k] & 1);
            }
#pragma unroll
            for (int j = 0; j < NV; j++) s[j] = -r[j];
        }
        if constexpr (RSCHEME >= 0) {
            double r[NV], cmp[NV];
#pragma unroll
            for (int j = 0; j < NV; j++) r[j] = cmp[j] = 0;
#pragma unroll
            for (int k = 0; k < MAXF; k++) {
                if (code[k] == -2) break;
                const bool c1side = fcode[k] & 1;
                double n64[DIM], X[NV], Y[NV], g[NV];
#pragma unroll
                for (int d = 0; d < DIM; d++) n64[d] = c1side ? -nrm[k][d] : nrm[k][d];
#pragma unroll
                for (int j = 0; j < NV; j++) {
                    X[j] = (code[k] >= 0 && c1side) ? q[k][j] : qc[j];
                    Y[j] = code[k] >= 0 ? (c1side ? qc[j] : q[k][j]) : qc[j];
                }
                face_solve<DIM, RSCHEME, double>(a.P, RSCHEME == MST_SCHEME_ROE, info[k] & 15, info[k] >> 4, n64, X, Y, g);
                accumulate<NV, false>(r, cmp, g, c1side);
            }
#pragma unroll
            for (int j = 0; j < NV; j++) s[j] = -r[j];
        }
#pragma unroll
        for (int k = 0; k < MAXF; k++) {
            if (code[k] == -2) break;  // the slots after it are padding too
            const bool nbr = code[k] >= 0, use = nbr && (code[k] & SIDE);
            double r;
            if constexpr (FWD) {
                r = lu_radius<DIM>(qc, nrm[k], a.gamma);
                if (nbr) r = fmax(r, lu_radius<DIM>(q[k], nrm[k], a.gamma));
                a.r[(int64_t)k * a.nc + t] = r;
                rsum += r;
            } else {
                r = rr[k];
            }
            if (!use) continue;
            double qp[NV], F1[NV], F0[NV];
#pragma unroll
            for (int j = 0; j < NV; j++) qp[j] = q[k][j] + dq[k][j];
            lu_flux<DIM>(qp, nrm[k], a.gamma, F1);
            lu_flux<DIM>(q[k], nrm[k], a.gamma, F0);
#pragma unroll
            for (int j = 0; j < NV; j++) s[j] -= 0.5 * ((F1[j] - F0[j]) - a.omega * r * dq[k][j]);
        }
    } else {
#pragma unroll
        for (int slot = 0; slot < MAXF; slot++) {
            const int64_t i = (int64_t)slot * a.nc + t;
            const int32_t code = a.nb[i];
            if (code == -2) break;
            if (!FWD && (code < 0 || !(code & SIDE))) continue;
            double nrm[DIM], q[NV];
#pragma unroll
            for (int d = 0; d < DIM; d++) nrm[d] = a.n[i * DIM + d];
            const int64_t o = code >> 2;
            if (code >= 0) ldq<NV>(a.Q + o * NV, q);
            double r;
            if constexpr (FWD) {
                r = lu_radius<DIM>(qc, nrm, a.gamma);
                if (code >= 0) r = fmax(r, lu_radius<DIM>(q, nrm, a.gamma));
                a.r[i] = r;
                rsum += r;
                if (code < 0 || !(code & SIDE)) continue;
            } else {
                r = a.r[i];
            }
            double dq[NV], qp[NV], F1[NV], F0[NV];
            ldq<NV>(a.dQ + o * NV, dq);
#pragma unroll
            for (int k = 0; k < NV; k++) qp[k] = q[k] + dq[k];
            lu_flux<DIM>(qp, nrm, a.gamma, F1);
            lu_flux<DIM>(q, nrm, a.gamma, F0);
#pragma unroll
            for (int k = 0; k < NV; k++) s[k] -= 0.5 * ((F1[k] - F0[k]) - a.omega * r * dq[k]);
        }
    }
    if constexpr (FWD) {
        D = a.dscale * rsum;
        a.D[t] = D;
    }
    double dq[NV], out[NV];
    if constexpr (FWD) {
#pragma unroll
        for (int k = 0; k < NV; k++) dq[k] = s[k] / D;
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) dq[k] = dqc[k] + s[k] / D;
    }
    stq<NV>(a.dQ + c * NV, dq);
    if constexpr (!FWD || LAST) {
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] = qc[k] + dq[k];
        stq<NV>(a.Qout + c * NV, out);
        const double p = lu_pressure<DIM>(out, a.gamma);
        if (!(out[0] > 0) || !(p > 0)) a.bad[0] = 1;
        if (qnew)
#pragma unroll
            for (int k = 0; k < NV; k++) qnew[k] = out[k];
    }
}

template <int DIM, bool FWD, bool LAST, int MAXF, int RSCHEME = LU_R_ARRAY>
__global__ __launch_bounds__(BLOCK) void k_lu_sweep(LuArgs a, int64_t t0, int64_t n) {
    const int64_t l = block_id() * BLOCK + threadIdx.x;
    if (l >= n) return;
    const int64_t t = t0 + l;
    if (a.pcase) {  // batch: this position's case
        const int32_t cs = a.pcase[t];
        if (!a.active[cs]) return;
        a.gamma = a.gam[cs];
        a.bad += cs;
    }
    lu_cell<DIM, FWD, LAST, MAXF, RSCHEME>(a, t, a.cells[t], nullptr, nullptr);
}

// Colour 0's backward sweep of one iteration fused with its forward sweep of the next (single
// context, first order, the residual computed in the forward sweeps, two colours or more).  Colour 0
// has no lower neighbours, so its forward sweep reads only the cell's new state (in registers here)
// and its neighbours' new states, which the iteration's other colours have finished; no other cell of
// the launch reads what either part writes.  The same code as the two separate launches: the same bits.
template <int DIM, int MAXF, int RSCHEME>
__global__ __launch_bounds__(BLOCK) void k_lu_sweep_bf(LuArgs a, int64_t t0, int64_t n) {
    constexpr int NV = DIM + 2;
    const int64_t l = block_id() * BLOCK + threadIdx.x;
    if (l >= n) return;
    const int64_t t = t0 + l, c = a.cells[t];
    double q1[NV];
    lu_cell<DIM, false, false, MAXF, LU_R_ARRAY>(a, t, c, nullptr, q1);
    LuArgs b = a;
    b.Q = a.Qout;  // the next iteration starts from the new state
    lu_cell<DIM, true, false, MAXF, RSCHEME>(b, t, c, q1, nullptr);
}

}  // namespace mstk
