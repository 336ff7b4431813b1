/*
 * mst_oracle.h -- CPU restatement of MST-CFD's density-based explicit residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is compared against
 * (and the timed CPU baseline).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (mst-cfd_amd/) never links, loads or calls it.
 *
 * It is plain C99 (+ optional OpenMP for the baseline timing), written from the reference's
 * algorithm (file:line citations on every function in mst_oracle.c), in the reference's own
 * data order: faces in file order, each cell's faces in the order the reader attaches them,
 * per-direction Riemann solves with set()+solverAll(), an explicit matrix inverse for Roe's |A|.
 *
 * Layouts (all row-major C arrays, nv = dim + 2 fields):
 *   Q     [n_cells][nv]           conserved state (rho, m_1..m_dim, rhoE)
 *   F     [n_faces][dim][nv]      face flux, column d = flux of the Riemann problem along axis d
 *   grad  [n_cells][nv][dim]      Green-Gauss gradient of the conserved state
 *   phi   [n_cells][nv]           limiter factor in [0,1]
 *   R     [n_cells][nv]           residual  R_i = sum_f sum_d n_out,d(i,f) F_f[d]
 */
#ifndef MST_ORACLE_H
#define MST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fluent boundary types used by the reference reader/solver (rhoSolver/RhoSolver.cpp:97-230),
 * plus PERIODIC (12, Fluent's periodic type), which the reference does not have. */
enum { ORC_INTERIOR = 2, ORC_WALL = 3, ORC_OUTLET = 5, ORC_SYMMETRY = 7, ORC_INLET = 10, ORC_PERIODIC = 12 };
enum { ORC_ROE = 0, ORC_AUSM = 1 };
enum { ORC_LIM_NONE = 0, ORC_LIM_MINMOD = 1, ORC_LIM_VANLEER = 2 };

/* compat bits: reference behaviours that can be switched */
#define ORC_COMPAT_BRANCH2_LR      1  /* literal L/R flags of MshBlock.cpp:294-303 on "branch 2" faces */
#define ORC_COMPAT_GRAD_LASTINT    2  /* RhoSolver.cpp:437-440 boundary loop starting at N_int-1 */
#define ORC_COMPAT_AUSM_LOCAL_A    4  /* SolverAusm.cpp:112-113: flux scaled by local a_L, a_R */
#define ORC_COMPAT_AUSM_P_UNSCALED 8  /* SolverAusm.cpp:130,140: 3/16 term of P+- not multiplied by p */
#define ORC_COMPAT_AUSM_ONESIDED   16 /* SolverAusm.cpp:116,122,129,135: subsonic branch taken for M <= 1
                                         (the code's abs(M <= 1)) instead of |M| <= 1 */

/* Same field layout as mst_config in include/mstcfd.h. */
typedef struct {
    int32_t scheme;    /* ORC_ROE / ORC_AUSM */
    int32_t order;     /* 1 or 2 */
    int32_t limiter;   /* ORC_LIM_* (order 2 only) */
    int32_t rk;        /* 1 forward Euler, 2 SSP-RK2, 3 SSP-RK3 */
    int32_t compat;    /* ORC_COMPAT_* bits */
    int32_t precision; /* ignored: the oracle is the FP64 reference */
    double gamma;          /* 1.4 */
    double eor;            /* 1e-10: rho + EOR in Roe's momentum fluxes */
    double entropy_delta;  /* 0.125: Harten fix */
    double cfl;
    double q_inlet[5];     /* inlet ghost state */
    double mu;             /* NEW viscous terms: dynamic viscosity (0: inviscid), include/CONST.h:46 */
    double lam;            /* bulk coefficient lambda (CONST.h:47: -0.666667 mu) */
    double kappa;          /* thermal conductivity (CONST.h:48 TEMPK) */
    double cv;             /* T = e / cv (CONST.h:39) */
} orc_config;

typedef struct orc_mesh orc_mesh;

orc_mesh *orc_mesh_create(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int64_t n_faces,
                          int fnn, const int32_t *fnodes, const int32_t *c0, const int32_t *c1,
                          const int32_t *ftype, const int32_t *partner, int64_t n_int_file);
void orc_mesh_destroy(orc_mesh *m);
int orc_set_threads(int n);

/* geometry in file order; any output pointer may be NULL */
int orc_mesh_geometry(const orc_mesh *m, double *fcen, double *S, double *eta, int32_t *dac,
                      double *ccen, double *vol);
/* per-face, per-direction "c0 is the Riemann left state" flags (n_faces*dim bytes) */
int orc_mesh_lrflags(const orc_mesh *m, int compat, uint8_t *flags);

/* thermodynamics and directional Riemann fluxes (single state pair) */
double orc_pressure(const double *Q, int dim, double gamma);
double orc_enthalpy(const double *Q, int dim, double gamma);
void orc_roe_flux(const double *QL, const double *QR, int d, int dim, const orc_config *cfg, double *F);
void orc_ausm_flux(const double *QL, const double *QR, int d, int dim, const orc_config *cfg, double *F);
void orc_roe_eigvecs(const double *QL, const double *QR, int d, int dim, const orc_config *cfg,
                     double *Rm, double *Rinv, double *lam);
int orc_invert(int n, const double *A, double *Ainv);

/* hot path */
int orc_gradient(const orc_mesh *m, const orc_config *cfg, const double *Q, double *grad);
int orc_limiter(const orc_mesh *m, const orc_config *cfg, const double *Q, const double *grad, double *phi);
int orc_face_flux(const orc_mesh *m, const orc_config *cfg, const double *Q, double *F);
int orc_residual(const orc_mesh *m, const orc_config *cfg, const double *Q, double *R, double *Sabs);
int orc_euler(const orc_mesh *m, const orc_config *cfg, const double *Qold, double dt, double *Qnew);
double orc_cfl_dt(const orc_mesh *m, const orc_config *cfg, const double *Q);
int orc_rk_step(const orc_mesh *m, const orc_config *cfg, double *Q, double dt, double *res_l2);
int orc_monitor_maxrel(int64_t n_cells, int nv, const double *Qnew, const double *Qold, double *out);

/* NEW viscous flux (the intent of RhoSolver::updateViscid, rhoSolver/RhoSolver.cpp:371-429):
 * Fv[n_faces][dim][nv]; orc_residual subtracts it from the convective flux when mu or kappa > 0.
 * gradW: optional [n_cells][dim+1][dim] Green-Gauss gradients of (u_1..u_dim, T). */
int orc_viscous_flux(const orc_mesh *m, const orc_config *cfg, const double *Q, double *Fv, double *gradW);

/* NEW steady driver: matrix-free LU-SGS over a cell colouring (colour[c] >= 0 per file-order cell,
 * no two neighbours alike).  sweep: dQ from a given residual R; step: R(Q), res_l2 (as orc_rk_step),
 * the sweeps and Q += dQ. */
int orc_lusgs_sweep(const orc_mesh *m, const orc_config *cfg, const double *Q, const double *R,
                    const int32_t *colour, double cfl, double omega, double *dQ);
int orc_lusgs_step(const orc_mesh *m, const orc_config *cfg, double *Q, const int32_t *colour, double cfl,
                   double omega, double *res_l2);

#ifdef __cplusplus
}
#endif
#endif
