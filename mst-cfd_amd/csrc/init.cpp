// init.cpp -- initial conditions on a mesh's cell centres (host, file order).
//
// MST_INIT_REFERENCE restates Time::initialization (time/Time.cpp:13-36): every cell gets the given
// conserved state, and cells whose centre has x > 0.5 get rho * 0.125 and rho*E * 0.1 (momentum
// unchanged).  The other kinds are new: a uniform state, a general two-state (Riemann) split, Shu's
// isentropic vortex, a Gaussian density/pressure pulse on a uniform stream, and the freestream of a
// ramp case, all from primitive variables (rho, u, p) with p = (gamma - 1)(rho E - rho |u|^2 / 2).
#include <cmath>
#include <vector>

#include "mesh.hpp"
#include "mstcfd.h"

namespace {

void put(double *Q, int dim, double gamma, double rho, const double *u, double p) {
    double ke = 0;
    Q[0] = rho;
    for (int d = 0; d < dim; d++) {
        Q[1 + d] = rho * u[d];
        ke += u[d] * u[d];
    }
    Q[dim + 1] = p / (gamma - 1) + 0.5 * rho * ke;
}

int need(int n_params, int n) {
    if (n_params < n) return mst::fail(MST_E_ARG, "init_state: " + std::to_string(n) + " parameters needed");
    return MST_OK;
}

}  // namespace

extern "C" int mst_mesh_init_state(const mst_mesh *m, int kind, const double *params, int n_params, double *Q,
                                   int64_t n_cells) {
    if (!m || !Q || (n_params > 0 && !params)) return mst::fail(MST_E_ARG, "null argument");
    if (n_cells != m->n_cells) return mst::fail(MST_E_ARG, "state size does not match the mesh");
    const int dim = m->dim, nv = dim + 2;
    std::vector<double> cc;
    int rc;
    try {
        rc = mst::cell_centres(*m, cc);
    } catch (const std::bad_alloc &) {
        return mst::fail(MST_E_ARG, "out of host memory");
    }
    if (rc) return rc;
    const double *P = params;
    double u[3] = {0, 0, 0};
    switch (kind) {
        case MST_INIT_REFERENCE: {  // params: Q0 (nv conserved values)
            if ((rc = need(n_params, nv))) return rc;
            for (int64_t c = 0; c < n_cells; c++) {
                for (int k = 0; k < nv; k++) Q[c * nv + k] = P[k];
                if (cc[c * dim] > 0.5) {
                    Q[c * nv] *= 0.125;
                    Q[c * nv + nv - 1] *= 0.1;
                }
            }
            return MST_OK;
        }
        case MST_INIT_UNIFORM: {  // gamma, rho, u_1..u_dim, p
            if ((rc = need(n_params, 3 + dim))) return rc;
            for (int64_t c = 0; c < n_cells; c++) put(Q + c * nv, dim, P[0], P[1], P + 2, P[2 + dim]);
            return MST_OK;
        }
        case MST_INIT_RIEMANN: {  // gamma, axis, x0, rhoL, uL(dim), pL, rhoR, uR(dim), pR
            if ((rc = need(n_params, 3 + 2 * (dim + 2)))) return rc;
            const int axis = (int)P[1];
            if (axis < 0 || axis >= dim) return mst::fail(MST_E_ARG, "init_state: axis out of range");
            const double *L = P + 3, *R = P + 3 + dim + 2;
            for (int64_t c = 0; c < n_cells; c++) {
                const double *S = cc[c * dim + axis] <= P[2] ? L : R;
                put(Q + c * nv, dim, P[0], S[0], S + 1, S[1 + dim]);
            }
            return MST_OK;
        }
        case MST_INIT_VORTEX: {  // gamma, mach, beta, xc, yc: rho_inf = p_inf = 1, u_inf = mach sqrt(gamma)
            if (dim != 2) return mst::fail(MST_E_ARG, "init_state: the vortex is two-dimensional");
            if ((rc = need(n_params, 5))) return rc;
            const double g = P[0], uinf = P[1] * std::sqrt(g), beta = P[2];
            for (int64_t c = 0; c < n_cells; c++) {
                const double x = cc[c * 2] - P[3], y = cc[c * 2 + 1] - P[4], r2 = x * x + y * y;
                const double e = std::exp(0.5 * (1 - r2));
                const double T = 1 - (g - 1) * beta * beta / (8 * g * M_PI * M_PI) * std::exp(1 - r2);
                const double rho = std::pow(T, 1 / (g - 1));
                u[0] = uinf - beta / (2 * M_PI) * y * e;
                u[1] = beta / (2 * M_PI) * x * e;
                put(Q + c * nv, 2, g, rho, u, rho * T);
            }
            return MST_OK;
        }
        case MST_INIT_PULSE: {  // gamma, amp, sigma, centre(dim), mach (stream along x)
            if ((rc = need(n_params, 4 + dim))) return rc;
            const double g = P[0], amp = P[1], s2 = 2 * P[2] * P[2];
            for (int64_t c = 0; c < n_cells; c++) {
                double r2 = 0;
                for (int d = 0; d < dim; d++) {
                    const double dx = cc[c * dim + d] - P[3 + d];
                    r2 += dx * dx;
                }
                const double b = amp * std::exp(-r2 / s2);
                u[0] = P[3 + dim];
                put(Q + c * nv, dim, g, 1 + b, u, (1 + b) / g);
            }
            return MST_OK;
        }
        case MST_INIT_FREESTREAM: {  // gamma, mach, flow angle in degrees (x-y plane): rho = 1, p = 1/gamma
            if ((rc = need(n_params, 3))) return rc;
            const double a = P[2] * M_PI / 180.0;
            u[0] = P[1] * std::cos(a);
            u[1] = P[1] * std::sin(a);
            for (int64_t c = 0; c < n_cells; c++) put(Q + c * nv, dim, P[0], 1.0, u, 1.0 / P[0]);
            return MST_OK;
        }
        default:
            return mst::fail(MST_E_ARG, "init_state: unknown kind");
    }
}
