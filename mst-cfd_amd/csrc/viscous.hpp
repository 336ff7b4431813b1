// Viscous flux: the intent of RhoSolver::updateViscid (rhoSolver/RhoSolver.cpp:371-429), which the
// reference cannot run as written and never adds to the residual (FLAGVISCID 0, include/CONST.h:14).
//   Q_f      the gradient kernel's face values (eta0 interpolation of the face's two cells; a
//            boundary face its cell; the last-interior-face quirk has eta0 = 1)         (:373-378)
//   W_f      = (u_1..u_dim, T) of Q_f, T = e / cv; on a wall face u = 0 (no-slip, adiabatic: the
//            cell's T), the wall the viscous branch of the convective boundary stands for (:160-163)
//   grad W_c = (1/V) sum_f W_f n_out                                                    (:380-391)
//   grad W_f = eta0 grad W_c0 + (1 - eta0) grad W_c1; boundary: grad W_c0                (:392-399)
//   tau_id   = mu (du_i/dx_d + du_d/dx_i) + lam (div u) delta_id                         (:401-413)
//   Fv[d]    = (0, tau_1d .. tau_dim,d, sum_i u_i tau_id + kappa dT/dx_d)                 (:415-419)
// and the face flux becomes G_f - sum_d n_d Fv[d] (the residual sums n_out . (F_conv - Fv)).
//   k_visc_grad  one lane per cell (the gradient kernel's slot-major incidence layout): grad W_c
//   k_visc_flux  one lane per face: grad W_f, W_f, tau, and G_f -= n . Fv, after the flux kernel
// Both are FP64 gathers over the state (memory-bound, like the inviscid gradient and flux kernels).
#pragma once
#include "device_common.hpp"

namespace mstk {

struct ViscArgs {
    int64_t nc, nf;
    const int32_t *__restrict__ inc_code;  // (neighbour << 1) | is_c1, -1 boundary, -2 empty
    const int32_t *__restrict__ ell;       // the same slots' (face << 1) | is_c1
    const int32_t *__restrict__ finfo;     // kind | (L/R mask << 4) per face
    const double *__restrict__ inc_w;      // eta0 of the face
    const double *__restrict__ inc_n;      // the cell's outward normal
    const double *__restrict__ vol;
    const int32_t *__restrict__ fc0;
    const int32_t *__restrict__ fc1;       // neighbour (interior / periodic partner cell), -1 boundary
    const double *__restrict__ feta;       // eta0 per face (quirk face: 1)
    const double *__restrict__ fn;         // c0 outward normal
    const double *__restrict__ Q;
    double *__restrict__ gw;               // [nc][dim + 1][dim]
    double *__restrict__ G;                // face fluxes, updated in place
    double mu, lam, kappa, cv;
};

template <int DIM>
__device__ __forceinline__ void visc_w(const double *q, double cv, double (&w)[DIM + 1]) {
    double ke = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        w[d] = q[1 + d] / q[0];
        ke += q[1 + d] * q[1 + d];
    }
    w[DIM] = (q[DIM + 1] - 0.5 * ke / q[0]) / q[0] / cv;
}

template <int DIM, int MAXF>
__global__ __launch_bounds__(BLOCK) void k_visc_grad(const ViscArgs a) {
    constexpr int NV = DIM + 2, NW = DIM + 1;
    const int64_t c = block_id() * BLOCK + threadIdx.x;
    if (c >= a.nc) return;
    double qc[NV], t[NW][DIM];
    ldq<NV>(a.Q + c * NV, qc);
#pragma unroll
    for (int k = 0; k < NW; k++)
#pragma unroll
        for (int d = 0; d < DIM; d++) t[k][d] = 0;
#pragma unroll
    for (int s = 0; s < MAXF; s++) {
        const int64_t i = (int64_t)s * a.nc + c;
        const int32_t code = a.inc_code[i];
        if (code == -2) break;
        double qf[NV];
        if (code >= 0) {
            const double eta = a.inc_w[i];
            double qn[NV];
            ldq<NV>(a.Q + (int64_t)(code >> 1) * NV, qn);
#pragma unroll
            for (int k = 0; k < NV; k++) qf[k] = (code & 1) ? eta * qn[k] + (1 - eta) * qc[k] : eta * qc[k] + (1 - eta) * qn[k];
        } else {
#pragma unroll
            for (int k = 0; k < NV; k++) qf[k] = qc[k];
        }
        double w[NW];
        visc_w<DIM>(qf, a.cv, w);
        if (code == -1 && (a.finfo[a.ell[i] >> 1] & 15) == mst::KIND_WALL) {
#pragma unroll
            for (int d = 0; d < DIM; d++) w[d] = 0;
        }
#pragma unroll
        for (int k = 0; k < NW; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) t[k][d] += w[k] * a.inc_n[i * DIM + d];
    }
    const double V = a.vol[c];
#pragma unroll
    for (int k = 0; k < NW; k++)
#pragma unroll
        for (int d = 0; d < DIM; d++) a.gw[(c * NW + k) * DIM + d] = t[k][d] / V;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_visc_flux(const ViscArgs a) {
    constexpr int NV = DIM + 2, NW = DIM + 1;
    const int64_t f = block_id() * BLOCK + threadIdx.x;
    if (f >= a.nf) return;
    const int64_t c0 = a.fc0[f], c1 = a.fc1[f];
    double q0[NV], qf[NV], g[NW][DIM];
    ldq<NV>(a.Q + c0 * NV, q0);
    if (c1 >= 0) {
        const double eta = a.feta[f];
        double q1[NV];
        ldq<NV>(a.Q + c1 * NV, q1);
#pragma unroll
        for (int k = 0; k < NV; k++) qf[k] = eta * q0[k] + (1 - eta) * q1[k];
#pragma unroll
        for (int k = 0; k < NW; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++)
                g[k][d] = eta * a.gw[(c0 * NW + k) * DIM + d] + (1 - eta) * a.gw[(c1 * NW + k) * DIM + d];
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) qf[k] = q0[k];
#pragma unroll
        for (int k = 0; k < NW; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) g[k][d] = a.gw[(c0 * NW + k) * DIM + d];
    }
    double w[NW];
    visc_w<DIM>(qf, a.cv, w);
    if (c1 < 0 && (a.finfo[f] & 15) == mst::KIND_WALL) {
#pragma unroll
        for (int d = 0; d < DIM; d++) w[d] = 0;
    }
    double div = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) div += g[d][d];
    double gf[NV];
    ldq<NV>(a.G + f * NV, gf);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double nd = a.fn[f * DIM + d];
        double en = 0;
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            double tau = a.mu * (g[i][d] + g[d][i]);
            if (i == d) tau += a.lam * div;
            gf[1 + i] -= nd * tau;
            en += w[i] * tau;
        }
        gf[NV - 1] -= nd * (en + a.kappa * g[DIM][d]);
    }
    stq<NV>(a.G + f * NV, gf);
}

}  // namespace mstk
