/*
 * mst_oracle.c -- CPU restatement of MST-CFD's density-based explicit residual path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); see mst_oracle.h.
 *
 * Every function cites the reference file:line it restates (paths relative to MST-CFD/).
 * Pieces the reference does not have (limiter, RK stages, CFL time step, periodic faces,
 * 3D generalisation, L2 residual) are marked NEW and follow the textbook definitions named there.
 *
 * Parity status (what pins each function; the tests in tests/test_oracle_*.py hold the checks):
 *   thermo            closed forms (p, H for hand-computed states)
 *   geometry          closed forms (Heron == 1/2|cross|, sum n_out = 0, sum V = domain area)
 *   Roe               consistency F(Q,Q)=F(Q), eigenvectors of the textbook Jacobian, mirror
 *                     symmetry, supersonic upwinding, Harten-fix continuity
 *   AUSM              split-polynomial identities, F(Q,Q) closed form exposing the a/a~ quirk,
 *                     supersonic upwinding
 *   wall/symmetry     ghost from a rotation into the frame of a slanted wall; vanishing mass and
 *                     energy fluxes on axis-aligned walls; uniform flow along a slanted wall is steady
 *   residual/update   freestream preservation and global conservation (2D and 3D); the column sums
 *                     of a y-uniform strip (and the slab sums of an x-only 3D box) equal a textbook
 *                     1D Roe / AUSM+ finite-volume difference; 3D residual symmetric under axis
 *                     permutations
 *   limiter (NEW)     textbook psi(r) values, psi=0 for r<=0, bounds of reconstructed values
 *   RK/CFL (NEW)      stage 1 == Euler bit-for-bit; SSP combination identities
 *   physics           Sod exact Riemann solution (L1 convergence), isentropic vortex
 * Planted mistakes (ORC_PLANT, below) check that these pins catch plausible slips.
 */
#include "mst_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNV 5
#define MAXD 3

/* Planted mistakes: tests/test_oracle_planted.py compiles this file with -DORC_PLANT=k into a
 * temporary library and checks that the pins of tests/test_oracle_*.py fail against it.  The
 * default build (ORC_PLANT 0) contains none of them. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

struct orc_mesh {
    int dim, nv, fnn;
    int64_t nn, nc, nf, n_int;
    double *xyz;
    int32_t *fnodes, *c0, *c1, *ftype, *partner;
    /* geometry, file order */
    double *fcen, *S, *eta;
    int32_t *dac;
    double *ccen, *vol, *shalf;
    /* cell -> faces, in the order the reader attaches them (ascending face id) */
    int64_t *cptr;
    int32_t *cface;
};

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ---------------------------------------------------------------- thermodynamics
 * work/FUNCTION.cpp:3-15.  The reference's kinetic energy uses components 1 and 2 only
 * (DIM==2).  For dim==3 the third momentum is added (NEW: no reference semantics in 3D). */
static double momsq(const double *Q, int dim) {
    double s = Q[1] * Q[1] + Q[2] * Q[2];
#if ORC_PLANT == 4 /* planted: the reference's 2D-only kinetic energy kept in 3D */
    (void)dim;
#else
    if (dim == 3) s += Q[3] * Q[3];
#endif
    return s;
}
double orc_pressure(const double *Q, int dim, double gamma) {
    return (Q[dim + 1] - 0.5 * momsq(Q, dim) / Q[0]) * (gamma - 1); /* FUNCTION.cpp:12-15 */
}
double orc_enthalpy(const double *Q, int dim, double gamma) {
    double p = (Q[dim + 1] - 0.5 * momsq(Q, dim) / Q[0]) * (gamma - 1); /* FUNCTION.cpp:3-7 */
    return (Q[dim + 1] + p) / Q[0];
}

/* ---------------------------------------------------------------- dense inverse
 * Stand-in for Eigen's Matrix::inverse() at rhoSolver/SolverRoe.cpp:103 (Gauss-Jordan with
 * partial pivoting; agrees with any exact inverse to rounding). */
int orc_invert(int n, const double *A, double *Ainv) {
    double a[MAXNV][2 * MAXNV];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++) a[i][j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return -1;
        if (piv != c)
            for (int j = 0; j < 2 * n; j++) {
                double t = a[c][j];
                a[c][j] = a[piv][j];
                a[piv][j] = t;
            }
        double inv = 1.0 / a[c][c];
        for (int j = 0; j < 2 * n; j++) a[c][j] *= inv;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            double f = a[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Ainv[i * n + j] = a[i][j + n];
    return 0;
}

/* ---------------------------------------------------------------- Roe
 * solverRoe::set (rhoSolver/SolverRoe.cpp:3-16), solverAll (:70-112), entropyRefine (:114-123). */
static double entropy_fix(double x, double delta) {
    if (x > delta) return x;                   /* SolverRoe.cpp:116-118 */
    return (x * x + delta * delta) / 2 / delta; /* SolverRoe.cpp:120-121 */
}

typedef struct {
    double u[MAXD], ht, a;
} roe_avg;

static void roe_set(const double *L, const double *R, int dim, double gamma, roe_avg *ra) {
    double D = sqrt(fabs(R[0] / L[0])); /* SolverRoe.cpp:7 */
    double usq = 0;
    for (int k = 0; k < dim; k++) {     /* SolverRoe.cpp:10-12 */
        ra->u[k] = (L[k + 1] / L[0] + D * R[k + 1] / R[0]) / (1. + D);
        usq += ra->u[k] * ra->u[k];
    }
    ra->ht = (orc_enthalpy(L, dim, gamma) + D * orc_enthalpy(R, dim, gamma)) / (1. + D); /* :13 */
    double nrm = sqrt(usq);                                                               /* roeU.norm() */
    ra->a = sqrt(fabs((gamma - 1.) * (ra->ht - 0.5 * (nrm * nrm))));                      /* :14 */
}

/* Right eigenvector matrix, row-major nv x nv.  Columns: acoustic(-), entropy, one shear wave per
 * transverse axis k != d (ascending k), acoustic(+).  For dim==2 this is exactly the comma
 * initialiser at SolverRoe.cpp:86 (x) and :92 (y); the dim==3 layout is NEW. */
static void roe_R(const roe_avg *ra, int d, int dim, double *Rm) {
    int nv = dim + 2;
    double q2h = 0;
    for (int k = 0; k < dim; k++) q2h += ra->u[k] * ra->u[k];
    q2h = 0.5 * q2h;
    for (int i = 0; i < nv * nv; i++) Rm[i] = 0;
    /* column 0: u_d - a */
    Rm[0 * nv + 0] = 1;
    for (int k = 0; k < dim; k++) Rm[(1 + k) * nv + 0] = (k == d) ? ra->u[k] - ra->a : ra->u[k];
    Rm[(nv - 1) * nv + 0] = ra->ht - ra->u[d] * ra->a;
    /* column 1: entropy */
    Rm[0 * nv + 1] = 1;
    for (int k = 0; k < dim; k++) Rm[(1 + k) * nv + 1] = ra->u[k];
    Rm[(nv - 1) * nv + 1] = q2h;
    /* shear columns */
    int col = 2;
    for (int k = 0; k < dim; k++) {
        if (k == d) continue;
        Rm[(1 + k) * nv + col] = 1;
#if ORC_PLANT == 5 /* planted: shear wave's energy entry from the wrong velocity component */
        Rm[(nv - 1) * nv + col] = ra->u[d];
#else
        Rm[(nv - 1) * nv + col] = ra->u[k];
#endif
        col++;
    }
    /* last column: u_d + a */
    Rm[0 * nv + nv - 1] = 1;
    for (int k = 0; k < dim; k++) Rm[(1 + k) * nv + nv - 1] = (k == d) ? ra->u[k] + ra->a : ra->u[k];
    Rm[(nv - 1) * nv + nv - 1] = ra->ht + ra->u[d] * ra->a;
}

/* physical flux along axis d with the reference's rho+EOR in the momentum terms
 * (SolverRoe.cpp:87-88 x, :93-94 y; p and H use rho without EOR). */
static void roe_phys_flux(const double *Q, int d, int dim, double gamma, double eor, double *F) {
    double p = orc_pressure(Q, dim, gamma);
    double ht = orc_enthalpy(Q, dim, gamma);
    F[0] = Q[d + 1];
    for (int k = 0; k < dim; k++) {
        F[1 + k] = Q[1 + k] * Q[1 + d] / (Q[0] + eor);
        if (k == d) F[1 + k] += p;
    }
    F[dim + 1] = ht * Q[d + 1];
}

void orc_roe_eigvecs(const double *QL, const double *QR, int d, int dim, const orc_config *cfg, double *Rm,
                     double *Rinv, double *lam) {
    roe_avg ra;
    int nv = dim + 2;
    roe_set(QL, QR, dim, cfg->gamma, &ra);
    roe_R(&ra, d, dim, Rm);
    if (Rinv) orc_invert(nv, Rm, Rinv);
    if (lam) {
        lam[0] = entropy_fix(fabs(ra.u[d] - ra.a), cfg->entropy_delta); /* SolverRoe.cpp:75 */
        for (int k = 0; k < dim; k++) lam[1 + k] = entropy_fix(fabs(ra.u[d]), cfg->entropy_delta); /* :76-78 */
        lam[nv - 1] = entropy_fix(fabs(ra.u[d] + ra.a), cfg->entropy_delta);                       /* :79 */
    }
}

void orc_roe_flux(const double *QL, const double *QR, int d, int dim, const orc_config *cfg, double *F) {
    int nv = dim + 2;
    double Rm[MAXNV * MAXNV], Ri[MAXNV * MAXNV], lam[MAXNV], absA[MAXNV * MAXNV];
    orc_roe_eigvecs(QL, QR, d, dim, cfg, Rm, Ri, lam);
    /* |A| = R |Lambda| R^-1  (SolverRoe.cpp:103-104) */
    for (int i = 0; i < nv; i++)
        for (int j = 0; j < nv; j++) {
            double s = 0;
            for (int k = 0; k < nv; k++) s += (Rm[i * nv + k] * lam[k]) * Ri[k * nv + j];
            absA[i * nv + j] = s;
        }
    double FL[MAXNV], FR[MAXNV], dU[MAXNV];
    roe_phys_flux(QL, d, dim, cfg->gamma, cfg->eor, FL);
    roe_phys_flux(QR, d, dim, cfg->gamma, cfg->eor, FR);
    for (int k = 0; k < nv; k++) dU[k] = QR[k] - QL[k]; /* SolverRoe.cpp:106 */
    for (int i = 0; i < nv; i++) {                      /* SolverRoe.cpp:109 */
        double s = 0;
        for (int j = 0; j < nv; j++) s += 0.5 * absA[i * nv + j] * dU[j];
        F[i] = 0.5 * (FL[i] + FR[i]) - s;
    }
}

/* ---------------------------------------------------------------- AUSM+
 * SolverAusm::set (rhoSolver/SolverAusm.cpp:3-26), solverAll (:53-63),
 * calculateMachAndPFace_AusmPlus (:105-143), with the reference's quirks behind compat bits. */
void orc_ausm_flux(const double *QL, const double *QR, int d, int dim, const orc_config *cfg, double *F) {
    const double g = cfg->gamma;
    int nv = dim + 2;
    /* set() */
    double aLs = sqrt(2 * orc_enthalpy(QL, dim, g) * (g - 1) / (g + 1)); /* :6 */
    double aRs = sqrt(2 * orc_enthalpy(QR, dim, g) * (g - 1) / (g + 1)); /* :7 */
    double UL = sqrt(momsq(QL, dim) / QL[0] / QL[0]);                    /* :10 (3D :15) */
    double UR = sqrt(momsq(QR, dim) / QR[0] / QR[0]);                    /* :11 */
    double aL = aLs * aLs / (aLs > UL ? aLs : UL);                       /* :20 std::max */
    double aR = aRs * aRs / (aRs > UR ? aRs : UR);                       /* :21 */
    double aF = aL < aR ? aL : aR;                                       /* :22 std::min */
    double pL = orc_pressure(QL, dim, g);                                /* :24 */
    double pR = orc_pressure(QR, dim, g);                                /* :25 */
    /* solverAll(d) */
    double mL = QL[d + 1] / QL[0] / aF; /* :54 */
    double mR = QR[d + 1] / QR[0] / aF; /* :55 */
    /* calculateMachAndPFace_AusmPlus() */
    double PsiL[MAXNV], PsiR[MAXNV];
    for (int k = 0; k < nv; k++) {
        PsiL[k] = QL[k];
        PsiR[k] = QR[k];
    }
    PsiL[nv - 1] += orc_pressure(QL, dim, g); /* :108 */
    PsiR[nv - 1] += orc_pressure(QR, dim, g); /* :110 */
    if (cfg->compat & ORC_COMPAT_AUSM_LOCAL_A) {
        aL = sqrt(g * orc_pressure(QL, dim, g) / QL[0]); /* :112 */
        aR = sqrt(g * orc_pressure(QR, dim, g) / QR[0]); /* :113 */
    } else {
        aL = aF;
        aR = aF;
    }
    double Mp, Mm, Pp, Pm;
    /* the reference writes abs(machL <= 1), i.e. the one-sided test machL <= 1 (:116,122,129,135);
     * without the compat bit the AUSM+ test |M| <= 1 */
    const int onesided = (cfg->compat & ORC_COMPAT_AUSM_ONESIDED) != 0;
    const int subL = onesided ? mL <= 1 : fabs(mL) <= 1;
    const int subR = onesided ? mR <= 1 : fabs(mR) <= 1;
    if (subL)
        Mp = 0.25 * (mL + 1) * (mL + 1) + 0.125 * (mL * mL - 1) * (mL * mL - 1); /* :117 */
    else
        Mp = 0.5 * (mL + fabs(mL)); /* :120 */
    if (subR)
        Mm = -0.25 * (mR - 1) * (mR - 1) - 0.125 * (mR * mR - 1) * (mR * mR - 1); /* :123 */
    else
        Mm = 0.5 * (mR - fabs(mR)); /* :126 */
    if (subL) {
        if (cfg->compat & ORC_COMPAT_AUSM_P_UNSCALED) /* :130 literal precedence */
            Pp = pL * 0.25 * (mL + 1) * (mL + 1) * (2 - mL) + 0.1875 * mL * (mL * mL - 1) * (mL * mL - 1);
        else
            Pp = pL * (0.25 * (mL + 1) * (mL + 1) * (2 - mL) + 0.1875 * mL * (mL * mL - 1) * (mL * mL - 1));
    } else
        Pp = pL * 0.5 * (mL + fabs(mL)) / mL; /* :133 */
    if (subR) {
        if (cfg->compat & ORC_COMPAT_AUSM_P_UNSCALED) /* :136 */
            Pm = pR * 0.25 * (mR - 1) * (mR - 1) * (2 + mR) - 0.1875 * mR * (mR * mR - 1) * (mR * mR - 1);
        else
            Pm = pR * (0.25 * (mR - 1) * (mR - 1) * (2 + mR) - 0.1875 * mR * (mR * mR - 1) * (mR * mR - 1));
    } else
        Pm = pR * 0.5 * (mR - fabs(mR)) / mR; /* :139 */
    double Mf = Mm + Mp; /* :141 */
    double pf = Pm + Pp; /* :142 */
    for (int k = 0; k < nv; k++) /* :58 */
        F[k] = 0.5 * (Mf * (aL * PsiL[k] + aR * PsiR[k]) - fabs(Mf) * (aR * PsiR[k] - aL * PsiL[k]));
    F[d + 1] += pf; /* :61 */
}

static void riemann(const orc_config *cfg, const double *L, const double *R, int d, int dim, double *F) {
    if (cfg->scheme == ORC_ROE)
        orc_roe_flux(L, R, d, dim, cfg, F);
    else
        orc_ausm_flux(L, R, d, dim, cfg, F);
}

/* ---------------------------------------------------------------- mesh + geometry
 * mesh/MshBlock.cpp:228-264 (face/cell attachment), :272-280 (centres, volumes),
 * :281-318 (orientation, per-cell outward normals), :319-321 (eta0);
 * mesh/Face.cpp:8-28 (centre, area vector), :62-69 (eta0); mesh/Cell.cpp:6-49 (centre, area). */
static double dot(const double *a, const double *b, int dim) {
    double s = a[0] * b[0] + a[1] * b[1];
    if (dim == 3) s += a[2] * b[2];
    return s;
}
static double norm(const double *a, int dim) { return sqrt(dot(a, a, dim)); }

orc_mesh *orc_mesh_create(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int64_t n_faces, int fnn,
                          const int32_t *fnodes, const int32_t *c0, const int32_t *c1, const int32_t *ftype,
                          const int32_t *partner, int64_t n_int_file) {
    if (dim != 2 && dim != 3) return NULL;
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->dim = dim;
    m->nv = dim + 2;
    m->fnn = fnn;
    m->nn = n_nodes;
    m->nc = n_cells;
    m->nf = n_faces;
    m->n_int = n_int_file;
    m->xyz = (double *)malloc(sizeof(double) * n_nodes * dim);
    memcpy(m->xyz, xyz, sizeof(double) * n_nodes * dim);
    m->fnodes = (int32_t *)malloc(sizeof(int32_t) * n_faces * fnn);
    memcpy(m->fnodes, fnodes, sizeof(int32_t) * n_faces * fnn);
    m->c0 = (int32_t *)malloc(sizeof(int32_t) * n_faces);
    m->c1 = (int32_t *)malloc(sizeof(int32_t) * n_faces);
    m->ftype = (int32_t *)malloc(sizeof(int32_t) * n_faces);
    m->partner = (int32_t *)malloc(sizeof(int32_t) * n_faces);
    memcpy(m->c0, c0, sizeof(int32_t) * n_faces);
    memcpy(m->c1, c1, sizeof(int32_t) * n_faces);
    memcpy(m->ftype, ftype, sizeof(int32_t) * n_faces);
    for (int64_t f = 0; f < n_faces; f++) m->partner[f] = partner ? partner[f] : -1;

    m->fcen = (double *)malloc(sizeof(double) * n_faces * dim);
    m->S = (double *)malloc(sizeof(double) * n_faces * dim);
    m->eta = (double *)malloc(sizeof(double) * n_faces);
    m->dac = (int32_t *)malloc(sizeof(int32_t) * n_faces);
    m->ccen = (double *)calloc((size_t)n_cells * dim, sizeof(double));
    m->vol = (double *)calloc((size_t)n_cells, sizeof(double));
    m->shalf = (double *)calloc((size_t)n_cells * dim, sizeof(double));

    /* cell -> face lists in attachment order (MshBlock.cpp:238-255: c0 then c1, face by face) */
    m->cptr = (int64_t *)calloc((size_t)n_cells + 1, sizeof(int64_t));
    for (int64_t f = 0; f < n_faces; f++) {
        m->cptr[c0[f] + 1]++;
        if (c1[f] >= 0) m->cptr[c1[f] + 1]++;
    }
    for (int64_t c = 0; c < n_cells; c++) m->cptr[c + 1] += m->cptr[c];
    m->cface = (int32_t *)malloc(sizeof(int32_t) * (m->cptr[n_cells] > 0 ? m->cptr[n_cells] : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (n_cells > 0 ? n_cells : 1));
    for (int64_t c = 0; c < n_cells; c++) fill[c] = m->cptr[c];
    for (int64_t f = 0; f < n_faces; f++) {
        m->cface[fill[c0[f]]++] = (int32_t)f;
        if (c1[f] >= 0) m->cface[fill[c1[f]]++] = (int32_t)f;
    }
    free(fill);

    /* Face::setCenter (Face.cpp:8-28) */
    for (int64_t f = 0; f < n_faces; f++) {
        double sum[MAXD] = {0, 0, 0};
        for (int j = 0; j < fnn; j++) {
            const double *p = xyz + (int64_t)fnodes[f * fnn + j] * dim;
            for (int k = 0; k < dim; k++) sum[k] += p[k];
        }
        for (int k = 0; k < dim; k++) m->fcen[f * dim + k] = sum[k] / (double)fnn;
        const double *p0 = xyz + (int64_t)fnodes[f * fnn + 0] * dim;
        const double *p1 = xyz + (int64_t)fnodes[f * fnn + 1] * dim;
        if (dim == 2) { /* Face.cpp:16-18 */
            double fx = p0[0] - p1[0], fy = p0[1] - p1[1];
            m->S[f * 2 + 0] = -fy;
            m->S[f * 2 + 1] = fx;
        } else { /* Face.cpp:23-27 (triangle faces) */
            const double *p2 = xyz + (int64_t)fnodes[f * fnn + 2] * dim;
            double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
#if ORC_PLANT == 6 /* planted: a transposed operand in the cross product's second component */
            double cr[3] = {a[1] * b[2] - a[2] * b[1], a[0] * b[2] - a[2] * b[0], a[0] * b[1] - a[1] * b[0]};
#else
            double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
#endif
            double s = fnn == 3 ? 0.5 : 1.0; /* Face.cpp:26 (tri) vs :33 (quad) */
            for (int k = 0; k < 3; k++) m->S[f * 3 + k] = s * cr[k];
        }
    }
    /* Cell::setCenter (Cell.cpp:6-14): mean of face centres in list order */
    for (int64_t c = 0; c < n_cells; c++) {
        double sum[MAXD] = {0, 0, 0};
        int64_t nfc = m->cptr[c + 1] - m->cptr[c];
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++)
            for (int k = 0; k < dim; k++) sum[k] += m->fcen[(int64_t)m->cface[j] * dim + k];
        for (int k = 0; k < dim; k++) m->ccen[c * dim + k] = sum[k] / (double)nfc;
    }
    /* orientation: MshBlock.cpp:282-283 -- directAndCells = -1 when S points into c0 */
    for (int64_t f = 0; f < n_faces; f++) {
        double dx[MAXD];
        for (int k = 0; k < dim; k++) dx[k] = m->fcen[f * dim + k] - m->ccen[(int64_t)c0[f] * dim + k];
        m->dac[f] = dot(m->S + f * dim, dx, dim) < 0 ? -1 : 1;
    }
    /* Cell::setVolume (Cell.cpp:15-49) */
    for (int64_t c = 0; c < n_cells; c++) {
        int64_t b = m->cptr[c];
        int64_t nfc = m->cptr[c + 1] - b;
        double V = 0;
        if (dim == 2 && nfc == 3) { /* Heron, Cell.cpp:18-27 */
            double l[3];
            for (int i = 0; i < 3; i++) l[i] = norm(m->S + (int64_t)m->cface[b + i] * 2, 2);
            double s = 0.5 * (l[0] + l[1] + l[2]);
            V = sqrt(s * (s - l[0]) * (s - l[1]) * (s - l[2]));
        } else if (dim == 2 && nfc == 4) { /* "middle line" heuristic, Cell.cpp:28-49 */
            const double *fc[4];
            double L[4];
            for (int i = 0; i < 4; i++) {
                fc[i] = m->fcen + (int64_t)m->cface[b + i] * 2;
                L[i] = norm(m->S + (int64_t)m->cface[b + i] * 2, 2);
            }
            double v1[2] = {fc[0][0] - fc[1][0], fc[0][1] - fc[1][1]};
            double v2[2] = {fc[2][0] - fc[3][0], fc[2][1] - fc[3][1]};
            int i0, i1, i2, i3;
            double ml;
            if (fabs(v1[0] * v2[1] - v2[0] * v1[1]) < 1e-7) { /* EOR2, CONST.h:43 */
                ml = 2 * norm(v1, 2);
                i0 = 0; i1 = 1; i2 = 2; i3 = 3;
            } else {
                double v3[2] = {fc[0][0] - fc[2][0], fc[0][1] - fc[2][1]};
                ml = 2 * norm(v3, 2);
                i0 = 0; i1 = 2; i2 = 1; i3 = 3;
            }
            double s = 0.5 * (L[i0] + L[i1] + ml);
            V = sqrt(s * (s - L[i0]) * (s - L[i1]) * (s - ml));
            double s2 = 0.5 * (L[i2] + L[i3] + ml);
            V += sqrt(s2 * (s2 - L[i2]) * (s2 - L[i3]) * (s2 - ml));
        } else { /* NEW: divergence theorem V = (1/dim) sum_f (x_f - x_c) . n_out (exact for simplices) */
            for (int64_t j = b; j < b + nfc; j++) {
                int32_t f = m->cface[j];
                double sg = (c == c0[f]) ? (double)m->dac[f] : -(double)m->dac[f];
                double dx[MAXD];
                for (int k = 0; k < dim; k++) dx[k] = m->fcen[(int64_t)f * dim + k] - m->ccen[c * dim + k];
                V += sg * dot(m->S + (int64_t)f * dim, dx, dim);
            }
            V /= dim;
        }
        m->vol[c] = V;
        /* NEW: half sums of |S_d| for the CFL estimate */
        for (int64_t j = b; j < b + nfc; j++)
            for (int k = 0; k < dim; k++) m->shalf[c * dim + k] += fabs(m->S[(int64_t)m->cface[j] * dim + k]);
        for (int k = 0; k < dim; k++) m->shalf[c * dim + k] *= 0.5;
    }
    /* Face::updateEta0 (Face.cpp:62-69); periodic faces (NEW) use the partner cell seen from the partner face */
    for (int64_t f = 0; f < n_faces; f++) {
        const double *xf = m->fcen + f * dim;
        double d0[MAXD], d1[MAXD];
        if (c1[f] >= 0) {
            for (int k = 0; k < dim; k++) {
                d0[k] = m->ccen[(int64_t)c0[f] * dim + k] - xf[k];
                d1[k] = m->ccen[(int64_t)c1[f] * dim + k] - xf[k];
            }
            m->eta[f] = norm(d1, dim) / (norm(d0, dim) + norm(d1, dim));
        } else if (ftype[f] == ORC_PERIODIC && m->partner[f] >= 0) {
            int32_t g = m->partner[f];
            for (int k = 0; k < dim; k++) {
                d0[k] = m->ccen[(int64_t)c0[f] * dim + k] - xf[k];
                d1[k] = m->ccen[(int64_t)c0[g] * dim + k] - m->fcen[(int64_t)g * dim + k];
            }
            m->eta[f] = norm(d1, dim) / (norm(d0, dim) + norm(d1, dim));
        } else {
            m->eta[f] = 1;
        }
    }
    return m;
}

void orc_mesh_destroy(orc_mesh *m) {
    if (!m) return;
    free(m->xyz); free(m->fnodes); free(m->c0); free(m->c1); free(m->ftype); free(m->partner);
    free(m->fcen); free(m->S); free(m->eta); free(m->dac); free(m->ccen); free(m->vol); free(m->shalf);
    free(m->cptr); free(m->cface);
    free(m);
}

int orc_mesh_geometry(const orc_mesh *m, double *fcen, double *S, double *eta, int32_t *dac, double *ccen,
                      double *vol) {
    int dim = m->dim;
    if (fcen) memcpy(fcen, m->fcen, sizeof(double) * m->nf * dim);
    if (S) memcpy(S, m->S, sizeof(double) * m->nf * dim);
    if (eta) memcpy(eta, m->eta, sizeof(double) * m->nf);
    if (dac) memcpy(dac, m->dac, sizeof(int32_t) * m->nf);
    if (ccen) memcpy(ccen, m->ccen, sizeof(double) * m->nc * dim);
    if (vol) memcpy(vol, m->vol, sizeof(double) * m->nc);
    return 0;
}

/* L/R flags, MshBlock.cpp:281-305.  flag true => (L,R) = (c0, c1 or ghost)  (RhoSolver.cpp:105-111).
 * Physical (default): flag_d = (n_out(c0)_d >= 0) -- the first-branch rule.  With
 * ORC_COMPAT_BRANCH2_LR the literal second branch (:294-303, inverted) is used for faces whose
 * area vector already points out of c0.  In 3D the reference never sets branch-2 flags
 * (MshBlock.cpp:635-647), so the physical rule is used. */
static int lr_flag(const orc_mesh *m, int compat, int64_t f, int d) {
    double s = (double)m->dac[f] * m->S[f * m->dim + d];
#if ORC_PLANT == 8 /* planted: the downwind side as Riemann "left" */
    int phys = (s < 0);
#else
    int phys = (s >= 0);
#endif
    if ((compat & ORC_COMPAT_BRANCH2_LR) && m->dim == 2 && m->dac[f] == 1) return !phys;
    return phys;
}

int orc_mesh_lrflags(const orc_mesh *m, int compat, uint8_t *flags) {
    for (int64_t f = 0; f < m->nf; f++)
        for (int d = 0; d < m->dim; d++) flags[f * m->dim + d] = (uint8_t)lr_flag(m, compat, f, d);
    return 0;
}

/* ---------------------------------------------------------------- gradient
 * RhoSolver::updateGradFlux, GRAD_INTERVAL==1 branch (rhoSolver/RhoSolver.cpp:430-454):
 * face values eta*Q_c0 + (1-eta)*Q_c1 (:433-436), boundary faces take Q_c0 (:437-440, starting
 * at N_int-1 under ORC_COMPAT_GRAD_LASTINT), then grad_c = (sum_f Q_f (x) n_out) / V (:441-452).
 * Periodic faces (NEW) take the eta-weighted value with the partner cell. */
static void face_values(const orc_mesh *m, const orc_config *cfg, const double *Q, double *Qf) {
    int nv = m->nv;
    int64_t nf = m->nf;
    int64_t nint = m->n_int;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; f++) {
        const double e = m->eta[f];
        if (m->c1[f] >= 0 && f < nint) {
            const double *q0 = Q + (int64_t)m->c0[f] * nv, *q1 = Q + (int64_t)m->c1[f] * nv;
            for (int k = 0; k < nv; k++) Qf[f * nv + k] = e * q0[k] + (1 - e) * q1[k];
        } else if (m->ftype[f] == ORC_PERIODIC && m->partner[f] >= 0) {
            const double *q0 = Q + (int64_t)m->c0[f] * nv, *q1 = Q + (int64_t)m->c0[m->partner[f]] * nv;
            for (int k = 0; k < nv; k++) Qf[f * nv + k] = e * q0[k] + (1 - e) * q1[k];
        } else {
            const double *q0 = Q + (int64_t)m->c0[f] * nv;
            for (int k = 0; k < nv; k++) Qf[f * nv + k] = q0[k];
        }
    }
    if ((cfg->compat & ORC_COMPAT_GRAD_LASTINT) && nint >= 1 && nint <= nf) {
        int64_t f = nint - 1; /* RhoSolver.cpp:438 loop start */
        if (m->ftype[f] != ORC_PERIODIC) {
            const double *q0 = Q + (int64_t)m->c0[f] * nv;
            for (int k = 0; k < nv; k++) Qf[f * nv + k] = q0[k];
        }
    }
}

static inline double nout(const orc_mesh *m, int64_t c, int32_t f, int d) {
#if ORC_PLANT == 7 /* planted: the c1 side keeps c0's outward normal */
    (void)c;
    return (double)m->dac[f] * m->S[(int64_t)f * m->dim + d];
#else
    return (c == m->c0[f]) ? (double)m->dac[f] * m->S[(int64_t)f * m->dim + d]
                           : (double)(-m->dac[f]) * m->S[(int64_t)f * m->dim + d];
#endif
}

int orc_gradient(const orc_mesh *m, const orc_config *cfg, const double *Q, double *grad) {
    int nv = m->nv, dim = m->dim;
    double *Qf = (double *)malloc(sizeof(double) * m->nf * nv);
    if (!Qf) return -1;
    face_values(m, cfg, Q, Qf);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->nc; c++) {
        double t[MAXNV * MAXD];
        for (int i = 0; i < nv * dim; i++) t[i] = 0;
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
            int32_t f = m->cface[j];
            for (int d = 0; d < dim; d++) {
                double n = nout(m, c, f, d);
                for (int k = 0; k < nv; k++) t[k * dim + d] += Qf[(int64_t)f * nv + k] * n;
            }
        }
        for (int i = 0; i < nv * dim; i++) grad[c * nv * dim + i] = t[i] / m->vol[c];
    }
    free(Qf);
    return 0;
}

/* ---------------------------------------------------------------- limiter (NEW)
 * Not in the reference (no limiter anywhere).  Per cell and field, the factor
 *   Phi = min(1, min_f psi(r_f)),  r_f = (Q_ext - Q_c) / delta_f,  delta_f = grad_c . (x_f - x_c),
 * with Q_ext the max (delta>0) or min (delta<0) over the cell and its face neighbours;
 * faces with |delta| <= 1e-12 (|Qmax|+|Qmin|) do not limit.
 *   minmod:   psi(r) = max(0, min(1, r))          (Barth-Jespersen form)
 *   van Leer: psi(r) = (r + |r|) / (1 + |r|)
 * With ORC_LIM_NONE, Phi = 1 and the reconstruction is the reference's unlimited one. */
static double psi(int lim, double r) {
    if (lim == ORC_LIM_MINMOD) return r < 0 ? 0 : (r < 1 ? r : 1);
    return (r + fabs(r)) / (1 + fabs(r));
}

int orc_limiter(const orc_mesh *m, const orc_config *cfg, const double *Q, const double *grad, double *phi) {
    int nv = m->nv, dim = m->dim;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->nc; c++) {
        double qmax[MAXNV], qmin[MAXNV], ph[MAXNV];
        const double *qc = Q + c * nv;
        for (int k = 0; k < nv; k++) {
            qmax[k] = qmin[k] = qc[k];
            ph[k] = 1;
        }
        if (cfg->limiter == ORC_LIM_NONE || cfg->order < 2) {
            for (int k = 0; k < nv; k++) phi[c * nv + k] = 1;
            continue;
        }
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
            int32_t f = m->cface[j];
            int64_t nb = -1;
            if (m->c1[f] >= 0)
                nb = (c == m->c0[f]) ? m->c1[f] : m->c0[f];
            else if (m->ftype[f] == ORC_PERIODIC && m->partner[f] >= 0)
                nb = m->c0[m->partner[f]];
            if (nb < 0) continue;
            for (int k = 0; k < nv; k++) {
                double q = Q[nb * nv + k];
                if (q > qmax[k]) qmax[k] = q;
                if (q < qmin[k]) qmin[k] = q;
            }
        }
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
            int32_t f = m->cface[j];
            double dx[MAXD];
            for (int d = 0; d < dim; d++) dx[d] = m->fcen[(int64_t)f * dim + d] - m->ccen[c * dim + d];
            for (int k = 0; k < nv; k++) {
                const double *g = grad + c * nv * dim + k * dim;
                double delta = g[0] * dx[0] + g[1] * dx[1];
                if (dim == 3) delta += g[2] * dx[2];
                double eps = 1e-12 * (fabs(qmax[k]) + fabs(qmin[k]));
                double p;
                if (delta > eps)
                    p = psi(cfg->limiter, (qmax[k] - qc[k]) / delta);
                else if (delta < -eps)
                    p = psi(cfg->limiter, (qmin[k] - qc[k]) / delta);
                else
                    p = 1;
                if (p < ph[k]) ph[k] = p;
            }
        }
        for (int k = 0; k < nv; k++) phi[c * nv + k] = ph[k];
    }
    return 0;
}

/* reconstructed state of cell c at face f: Q_c + Phi_c o (grad_c (x_f - x_c))  (RhoSolver.cpp:250,254) */
static void recon(const orc_mesh *m, const double *Q, const double *grad, const double *phi, int64_t c, int64_t f,
                  double *out) {
    int nv = m->nv, dim = m->dim;
    double dx[MAXD];
    for (int d = 0; d < dim; d++) dx[d] = m->fcen[f * dim + d] - m->ccen[c * dim + d];
    for (int k = 0; k < nv; k++) {
        const double *g = grad + c * nv * dim + k * dim;
        double delta = g[0] * dx[0] + g[1] * dx[1];
        if (dim == 3) delta += g[2] * dx[2];
        out[k] = Q[c * nv + k] + phi[c * nv + k] * delta;
    }
}

/* wall / symmetry ghost momentum: u - 2 n^ (n^ . u), n^ = S/|S| (RhoSolver.cpp:148-159, 187-198) */
static void reflect(const orc_mesh *m, int64_t f, const double *src, double *dst) {
    int dim = m->dim;
    const double *S = m->S + f * dim;
    double nrm = norm(S, dim);
    double nh[MAXD];
#if ORC_PLANT == 2 /* planted: the raw area vector instead of the unit normal */
    (void)nrm;
    for (int d = 0; d < dim; d++) nh[d] = S[d];
#else
    for (int d = 0; d < dim; d++) nh[d] = S[d] / nrm;
#endif
    double s = nh[0] * src[1] + nh[1] * src[2];
    if (dim == 3) s += nh[2] * src[3];
#if ORC_PLANT == 1 /* planted: the factor 2 dropped (normal momentum zeroed, not negated) */
    for (int d = 0; d < dim; d++) dst[1 + d] = src[1 + d] - nh[d] * s;
#elif ORC_PLANT == 3 /* planted: the normal component added instead of subtracted */
    for (int d = 0; d < dim; d++) dst[1 + d] = src[1 + d] + 2 * nh[d] * s;
#else
    for (int d = 0; d < dim; d++) dst[1 + d] = src[1 + d] - 2 * nh[d] * s;
#endif
}

/* ---------------------------------------------------------------- face flux
 * RhoSolver::updateFaceFlux (rhoSolver/RhoSolver.cpp:90-233) and updateFaceFlux2ndOrder (:234-369),
 * face by face with the zone type switch:
 *   2 interior  (:100-116 / :244-260)       10 inlet  (:121-137 / :265-281)
 *   3 wall      (:142-179 / :285-319)        7 symmetry (:184-211 / :322-350)
 *   5 outlet    (:216-226 / :353-363)       12 periodic (NEW: ghost = partner cell)
 * Second-order outlet: the reference writes through unallocated pointers (:360-361); the intent
 * (L = Q_c0, R = reconstructed Q_c0, no flag) is implemented.  Unknown types keep a zero flux. */
static void face_flux_core(const orc_mesh *m, const orc_config *cfg, const double *Q, const double *grad,
                           const double *phi, double *F) {
    int nv = m->nv, dim = m->dim;
    int second = (cfg->order >= 2);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < m->nf; f++) {
        double *Ff = F + f * dim * nv;
        int64_t c = m->c0[f];
        double in[MAXNV], gh[MAXNV];
        int t = m->ftype[f];
        int use_flags = 1;
        const double *qc = Q + c * nv;
        if (m->c1[f] >= 0) { /* interior */
            if (second) {
                recon(m, Q, grad, phi, c, f, in);
                recon(m, Q, grad, phi, m->c1[f], f, gh);
            } else {
                memcpy(in, qc, sizeof(double) * nv);
                memcpy(gh, Q + (int64_t)m->c1[f] * nv, sizeof(double) * nv);
            }
        } else if (t == ORC_INLET) {
            if (second) recon(m, Q, grad, phi, c, f, in);
            else memcpy(in, qc, sizeof(double) * nv);
            /* inletQ (RhoSolver.cpp:123): rho, rho u, rho v, E from the configuration */
            for (int k = 0; k < nv; k++) gh[k] = cfg->q_inlet[k];
        } else if (t == ORC_WALL) {
            if (second) recon(m, Q, grad, phi, c, f, in);
            else memcpy(in, qc, sizeof(double) * nv);
            memcpy(gh, in, sizeof(double) * nv);
#if ORC_PLANT == 10 /* planted: the inviscid mirror ghost kept on a viscous wall */
            reflect(m, f, in, gh);
#else
            if (cfg->mu > 0 || cfg->kappa > 0) /* FLAGVISCID wall: momentum negated (:160-163, :300-303) */
                for (int d = 0; d < dim; d++) gh[1 + d] = -in[1 + d];
            else
                reflect(m, f, in, gh);
#endif
        } else if (t == ORC_SYMMETRY) {
            memcpy(in, qc, sizeof(double) * nv); /* :325 interior side not reconstructed */
            if (second) recon(m, Q, grad, phi, c, f, gh);
            else memcpy(gh, qc, sizeof(double) * nv);
            reflect(m, f, qc, gh); /* momentum from the unreconstructed state (:328-335) */
        } else if (t == ORC_OUTLET) {
            use_flags = 0;
            memcpy(in, qc, sizeof(double) * nv);
            if (second) recon(m, Q, grad, phi, c, f, gh);
            else memcpy(gh, qc, sizeof(double) * nv);
        } else if (t == ORC_PERIODIC && m->partner[f] >= 0) {
            int64_t g = m->partner[f];
            int64_t cb = m->c0[g];
            if (second) {
                recon(m, Q, grad, phi, c, f, in);
                recon(m, Q, grad, phi, cb, g, gh);
            } else {
                memcpy(in, qc, sizeof(double) * nv);
                memcpy(gh, Q + cb * nv, sizeof(double) * nv);
            }
        } else {
            for (int i = 0; i < dim * nv; i++) Ff[i] = 0;
            continue;
        }
        for (int d = 0; d < dim; d++) {
            if (!use_flags) {
                /* outlet: 1st order set(out,in) (:219-224); 2nd order set(Q_c0, out) (:358-361) */
                if (second) riemann(cfg, in, gh, d, dim, Ff + d * nv);
                else riemann(cfg, gh, in, d, dim, Ff + d * nv);
            } else if (lr_flag(m, cfg->compat, f, d))
                riemann(cfg, in, gh, d, dim, Ff + d * nv);
            else
                riemann(cfg, gh, in, d, dim, Ff + d * nv);
        }
    }
}

int orc_face_flux(const orc_mesh *m, const orc_config *cfg, const double *Q, double *F) {
    int nv = m->nv, dim = m->dim;
    if (cfg->order >= 2) {
        double *grad = (double *)malloc(sizeof(double) * m->nc * nv * dim);
        double *phi = (double *)malloc(sizeof(double) * m->nc * nv);
        orc_gradient(m, cfg, Q, grad);
        orc_limiter(m, cfg, Q, grad, phi);
        face_flux_core(m, cfg, Q, grad, phi, F);
        free(grad);
        free(phi);
    } else {
        face_flux_core(m, cfg, Q, NULL, NULL, F);
    }
    return 0;
}

/* ---------------------------------------------------------------- residual gather
 * RhoSolver::solve (rhoSolver/RhoSolver.cpp:45-56): R_i = sum over the cell's faces (list order)
 * and directions d of n_out,d(i,f) * F_f[:,d].  Sabs (parity scale) = sum |n_d| |F_f[k,d]|. */
static void gather(const orc_mesh *m, const double *F, double *R, double *Sabs) {
    int nv = m->nv, dim = m->dim;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->nc; c++) {
        double r[MAXNV] = {0, 0, 0, 0, 0}, sa[MAXNV] = {0, 0, 0, 0, 0};
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
            int32_t f = m->cface[j];
            for (int d = 0; d < dim; d++) {
                double n = nout(m, c, f, d);
                const double *Fd = F + ((int64_t)f * dim + d) * nv;
                for (int k = 0; k < nv; k++) {
                    r[k] += n * Fd[k];
                    sa[k] += fabs(n) * fabs(Fd[k]);
                }
            }
        }
        for (int k = 0; k < nv; k++) {
            R[c * nv + k] = r[k];
            if (Sabs) Sabs[c * nv + k] = sa[k];
        }
    }
}

/* ---------------------------------------------------------------- viscous flux (NEW)
 * RhoSolver::updateViscid (rhoSolver/RhoSolver.cpp:371-429) states the algorithm but cannot run as
 * written (cell gradients read the face array by cell index, :383-388; velocity components read the
 * wrong state slots, :386; the stress indexes a third direction, :404; boundary faces read the face
 * gradient array by cell id, :398; the heat flux takes one matrix element, :418), and the residual
 * never uses it (FLAGVISCID 0, include/CONST.h:14).  Restated as intended:
 *   Q_f      as the gradient's face values (eta0 interpolation; boundary: the cell; :373-378)
 *   W_f      = (u_1..u_dim, T) of Q_f, T = e / cv with e the internal energy per mass; on a wall face
 *              u = 0 and T the cell's (no-slip, adiabatic): the wall of the reference's viscous
 *              branch, whose convective ghost negates the momentum (:160-163, :300-303)
 *   grad W_c = (1/V) sum_f W_f n_out                                              (:380-391)
 *   grad W_f = eta0 grad W_c0 + (1 - eta0) grad W_c1; boundary: grad W_c0          (:392-399)
 *   tau_id   = mu (du_i/dx_d + du_d/dx_i) + lam (div u) delta_id                   (:401-413)
 *   Fv[d]    = (0, tau_1d .. tau_dim,d, sum_i u_i tau_id + kappa dT/dx_d)           (:415-419)
 * and R_i = sum_f n_out . (F_conv - Fv). */
static void prim_w(const double *q, int dim, double cv, double *w) {
    double ke = 0;
    for (int d = 0; d < dim; d++) {
        w[d] = q[1 + d] / q[0];
        ke += q[1 + d] * q[1 + d];
    }
    w[dim] = (q[dim + 1] - 0.5 * ke / q[0]) / q[0] / cv;
}

int orc_viscous_flux(const orc_mesh *m, const orc_config *cfg, const double *Q, double *Fv, double *gradW) {
    const int nv = m->nv, dim = m->dim, nw = dim + 1;
    double *Qf = (double *)malloc(sizeof(double) * m->nf * nv);
    double *gw = (double *)calloc((size_t)m->nc * nw * dim, sizeof(double));
    if (!Qf || !gw) {
        free(Qf);
        free(gw);
        return -1;
    }
    face_values(m, cfg, Q, Qf);
    for (int64_t c = 0; c < m->nc; c++) {
        double *g = gw + c * nw * dim;
        for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
            const int32_t f = m->cface[j];
            double w[MAXD + 1];
            prim_w(Qf + (int64_t)f * nv, dim, cfg->cv, w);
#if ORC_PLANT != 9 /* planted (9): a slip wall in the viscous gradient */
            if (m->c1[f] < 0 && m->ftype[f] == ORC_WALL)
                for (int d = 0; d < dim; d++) w[d] = 0;
#endif
            for (int k = 0; k < nw; k++)
                for (int d = 0; d < dim; d++) g[k * dim + d] += w[k] * nout(m, c, f, d);
        }
        for (int k = 0; k < nw * dim; k++) g[k] /= m->vol[c];
    }
    const int quirk = (cfg->compat & ORC_COMPAT_GRAD_LASTINT) && m->n_int >= 1 && m->n_int <= m->nf;
    for (int64_t f = 0; f < m->nf; f++) {
        const double *g0 = gw + (int64_t)m->c0[f] * nw * dim;
        int64_t o = -1;
        if (m->c1[f] >= 0 && f < m->n_int && !(quirk && f == m->n_int - 1))
            o = m->c1[f];
        else if (m->ftype[f] == ORC_PERIODIC && m->partner[f] >= 0)
            o = m->c0[m->partner[f]];
        double gf[(MAXD + 1) * MAXD];
        const double e = m->eta[f];
        for (int k = 0; k < nw * dim; k++) gf[k] = o >= 0 ? e * g0[k] + (1 - e) * gw[o * nw * dim + k] : g0[k];
        double w[MAXD + 1];
        prim_w(Qf + f * nv, dim, cfg->cv, w);
        if (m->c1[f] < 0 && m->ftype[f] == ORC_WALL)
            for (int d = 0; d < dim; d++) w[d] = 0;
        double div = 0;
        for (int d = 0; d < dim; d++) div += gf[d * dim + d];
        for (int d = 0; d < dim; d++) {
            double *F = Fv + (f * dim + d) * nv;
            F[0] = 0;
            double en = 0;
            for (int i = 0; i < dim; i++) {
                double tau = cfg->mu * (gf[i * dim + d] + gf[d * dim + i]);
                if (i == d) tau += cfg->lam * div;
                F[1 + i] = tau;
                en += w[i] * tau;
            }
            F[nv - 1] = en + cfg->kappa * gf[dim * dim + d];
        }
    }
    if (gradW) memcpy(gradW, gw, sizeof(double) * m->nc * nw * dim);
    free(Qf);
    free(gw);
    return 0;
}

int orc_residual(const orc_mesh *m, const orc_config *cfg, const double *Q, double *R, double *Sabs) {
    const size_t n = (size_t)m->nf * m->dim * m->nv;
    double *F = (double *)malloc(sizeof(double) * n);
    if (!F) return -1;
    orc_face_flux(m, cfg, Q, F);
    if (cfg->mu > 0 || cfg->kappa > 0) {
        double *Fv = (double *)malloc(sizeof(double) * n);
        if (!Fv || orc_viscous_flux(m, cfg, Q, Fv, NULL)) {
            free(Fv);
            free(F);
            return -1;
        }
        for (size_t i = 0; i < n; i++) F[i] -= Fv[i];
        free(Fv);
    }
    gather(m, F, R, Sabs);
    free(F);
    return 0;
}

/* RhoSolver::solve forward-Euler update (rhoSolver/RhoSolver.cpp:64): Q_new = Q_old - DT/V * R */
int orc_euler(const orc_mesh *m, const orc_config *cfg, const double *Qold, double dt, double *Qnew) {
    int nv = m->nv;
    double *R = (double *)malloc(sizeof(double) * m->nc * nv);
    orc_residual(m, cfg, Qold, R, NULL);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->nc; c++)
        for (int k = 0; k < nv; k++) Qnew[c * nv + k] = Qold[c * nv + k] - dt / m->vol[c] * R[c * nv + k];
    free(R);
    return 0;
}

/* NEW: CFL time step dt = CFL * min_c V_c / sum_d (|u_d| + a) Shalf_c,d, Shalf_c,d = 1/2 sum_f |S_f,d|
 * (the reference uses a fixed DT = 1/STEP_TIME, time/Time.cpp:62). */
double orc_cfl_dt(const orc_mesh *m, const orc_config *cfg, const double *Q) {
    int nv = m->nv, dim = m->dim;
    double best = INFINITY;
    for (int64_t c = 0; c < m->nc; c++) {
        const double *q = Q + c * nv;
        double p = orc_pressure(q, dim, cfg->gamma);
        double a = sqrt(cfg->gamma * p / q[0]);
        double lam = 0;
        for (int d = 0; d < dim; d++) lam += (fabs(q[1 + d] / q[0]) + a) * m->shalf[c * dim + d];
        double dt = cfg->cfl * m->vol[c] / lam;
        if (dt < best) best = dt;
    }
    return best;
}

/* NEW: SSP Runge-Kutta (Shu-Osher form).  Stage s: Q(s+1) = a_s Q^n + b_s (Q(s) - dt/V R(Q(s))),
 * stage 0 is exactly the reference's Euler update.  res_l2 (optional, nv values) receives the L2
 * norm over cells of R(Q^n)/V per field. */
int orc_rk_step(const orc_mesh *m, const orc_config *cfg, double *Q, double dt, double *res_l2) {
    static const double A2[2] = {0, 0.5}, B2[2] = {1, 0.5};
    static const double A3[3] = {0, 0.75, 1.0 / 3.0}, B3[3] = {1, 0.25, 2.0 / 3.0};
    int nv = m->nv;
    int ns = cfg->rk == 3 ? 3 : (cfg->rk == 2 ? 2 : 1);
    const double *A = cfg->rk == 3 ? A3 : A2, *B = cfg->rk == 3 ? B3 : B2;
    size_t n = (size_t)m->nc * nv;
    double *Qn = (double *)malloc(sizeof(double) * n);
    double *R = (double *)malloc(sizeof(double) * n);
    memcpy(Qn, Q, sizeof(double) * n);
    for (int s = 0; s < ns; s++) {
        orc_residual(m, cfg, Q, R, NULL);
        if (s == 0 && res_l2) {
            for (int k = 0; k < nv; k++) res_l2[k] = 0;
            for (int64_t c = 0; c < m->nc; c++)
                for (int k = 0; k < nv; k++) {
                    double r = R[c * nv + k] / m->vol[c];
                    res_l2[k] += r * r;
                }
            for (int k = 0; k < nv; k++) res_l2[k] = sqrt(res_l2[k] / (double)m->nc);
        }
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < m->nc; c++)
            for (int k = 0; k < nv; k++) {
                double e = Q[c * nv + k] - dt / m->vol[c] * R[c * nv + k];
                Q[c * nv + k] = (s == 0) ? e : A[s] * Qn[c * nv + k] + B[s] * e;
            }
    }
    free(Qn);
    free(R);
    return 0;
}

/* Time::goNextTimeStep monitor (time/Time.cpp:69-76): r_k = max_i |new-old|/old, skipping values
 * that do not compare greater (NaN from 0/0, negatives from a signed denominator). */
int orc_monitor_maxrel(int64_t n_cells, int nv, const double *Qnew, const double *Qold, double *out) {
    for (int k = 0; k < nv; k++) out[k] = 0;
    for (int64_t i = 0; i < n_cells; i++)
        for (int k = 0; k < nv; k++) {
            double v = fabs(Qnew[i * nv + k] - Qold[i * nv + k]) / Qold[i * nv + k];
            if (out[k] < v) out[k] = v;
        }
    return 0;
}

/* ---------------------------------------------------------------- LU-SGS steady driver (NEW)
 * Matrix-free LU-SGS pseudo-time iteration.  The factorised operator (D+L) D^-1 (D+U) -- a forward
 * sweep with the lower part, a division by D, a backward sweep with the upper part -- is the one
 * SparseSolverNUM::solveILUSGS applies to its assembled scalar systems
 * (lusolver/SparseSolverNUM.cpp:144-196).  Here the blocks are never assembled: they come from the
 * Rusanov split of the face-flux Jacobian (Jameson & Yoon, AIAA J. 25(7), 929-935, 1987):
 *   r_f  = max over the face's cells of |u.S_f| + a |S_f|     (a boundary face: its one cell)
 *   D_i  = (1/CFL + omega/2) sum_f r_f                         (local step dt_i = CFL V_i / sum_f r_f)
 *   (i,j) block times dQ_j = 1/2 [F_n(Q_j + dQ_j) - F_n(Q_j)] - 1/2 omega r_f dQ_j,
 *        n = the outward normal of cell i at f, F_n = sum_d n_d F_d (physical Euler flux)
 * L / U hold the neighbours of lower / higher colour.  Boundary faces enter D only.  Cells are swept
 * colour by colour, ascending then descending; no two cells of one colour are neighbours, so the
 * order inside a colour does not change the result. */
static double lu_radius(const double *q, const double *n, int dim, double gamma) {
    double un = 0, s2 = 0;
    for (int d = 0; d < dim; d++) {
        un += q[1 + d] * n[d];
        s2 += n[d] * n[d];
    }
    const double p = orc_pressure(q, dim, gamma);
    return fabs(un / q[0]) + sqrt(gamma * p / q[0]) * sqrt(s2);
}

static void lu_flux(const double *q, const double *n, int dim, double gamma, double *F) {
    const int nv = dim + 2;
    double un = 0;
    for (int d = 0; d < dim; d++) un += q[1 + d] * n[d];
    un /= q[0];
    const double p = orc_pressure(q, dim, gamma);
    F[0] = q[0] * un;
    for (int d = 0; d < dim; d++) F[1 + d] = q[1 + d] * un + p * n[d];
    F[nv - 1] = (q[nv - 1] + p) * un;
}

/* the cell across face f from cell c (periodic: the partner face's cell), -1 on a boundary */
static int64_t lu_neighbour(const orc_mesh *m, int64_t c, int32_t f) {
    if (m->c1[f] >= 0) return c == m->c0[f] ? m->c1[f] : m->c0[f];
    if (m->ftype[f] == ORC_PERIODIC && m->partner[f] >= 0) return m->c0[m->partner[f]];
    return -1;
}

/* s -= 1/2 [F_n(q + dq) - F_n(q)] - 1/2 omega r dq */
static void lu_offdiag(const double *q, const double *dq, const double *n, double r, int dim, double gamma,
                       double omega, double *s) {
    const int nv = dim + 2;
    double qp[MAXNV], F1[MAXNV], F0[MAXNV];
    for (int k = 0; k < nv; k++) qp[k] = q[k] + dq[k];
    lu_flux(qp, n, dim, gamma, F1);
    lu_flux(q, n, dim, gamma, F0);
    for (int k = 0; k < nv; k++) s[k] -= 0.5 * ((F1[k] - F0[k]) - omega * r * dq[k]);
}

int orc_lusgs_sweep(const orc_mesh *m, const orc_config *cfg, const double *Q, const double *R,
                    const int32_t *colour, double cfl, double omega, double *dQ) {
    const int nv = m->nv, dim = m->dim;
    const double g = cfg->gamma;
    if (!(cfl > 0) || !(omega > 0)) return -1;
    int ncol = 0;
    for (int64_t c = 0; c < m->nc; c++) {
        if (colour[c] < 0) return -1;
        if (colour[c] + 1 > ncol) ncol = colour[c] + 1;
    }
    double *rf = (double *)malloc(sizeof(double) * (m->nf > 0 ? m->nf : 1));
    if (!rf) return -1;
    for (int64_t f = 0; f < m->nf; f++) {
        const double *n = m->S + f * dim;
        double r = lu_radius(Q + (int64_t)m->c0[f] * nv, n, dim, g);
        const int64_t o = lu_neighbour(m, m->c0[f], (int32_t)f);
        if (o >= 0) {
            const double r1 = lu_radius(Q + o * nv, n, dim, g);
            if (r1 > r) r = r1;
        }
        rf[f] = r;
    }
    const double dscale = 1.0 / cfl + 0.5 * omega;
    for (int pass = 0; pass < 2; pass++) {
        for (int step = 0; step < ncol; step++) {
            const int col = pass == 0 ? step : ncol - 1 - step;
            for (int64_t c = 0; c < m->nc; c++) {
                if (colour[c] != col) continue;
                double s[MAXNV], nrm[MAXD], rsum = 0;
                for (int k = 0; k < nv; k++) s[k] = pass == 0 ? -R[c * nv + k] : 0.0;
                for (int64_t j = m->cptr[c]; j < m->cptr[c + 1]; j++) {
                    const int32_t f = m->cface[j];
                    rsum += rf[f];
                    const int64_t o = lu_neighbour(m, c, f);
                    if (o < 0 || o == c) continue;
                    if (pass == 0 ? colour[o] < col : colour[o] > col) {
                        for (int d = 0; d < dim; d++) nrm[d] = nout(m, c, f, d);
                        lu_offdiag(Q + o * nv, dQ + o * nv, nrm, rf[f], dim, g, omega, s);
                    }
                }
                const double D = dscale * rsum;
                for (int k = 0; k < nv; k++) {
                    if (pass == 0)
                        dQ[c * nv + k] = s[k] / D;
                    else
                        dQ[c * nv + k] += s[k] / D;
                }
            }
        }
    }
    free(rf);
    return 0;
}

int orc_lusgs_step(const orc_mesh *m, const orc_config *cfg, double *Q, const int32_t *colour, double cfl,
                   double omega, double *res_l2) {
    const int nv = m->nv;
    const size_t n = (size_t)m->nc * nv;
    double *R = (double *)malloc(sizeof(double) * (n ? n : 1));
    double *dQ = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!R || !dQ) {
        free(R);
        free(dQ);
        return -1;
    }
    orc_residual(m, cfg, Q, R, NULL);
    if (res_l2) {
        for (int k = 0; k < nv; k++) res_l2[k] = 0;
        for (int64_t c = 0; c < m->nc; c++)
            for (int k = 0; k < nv; k++) {
                const double r = R[c * nv + k] / m->vol[c];
                res_l2[k] += r * r;
            }
        for (int k = 0; k < nv; k++) res_l2[k] = sqrt(res_l2[k] / (double)m->nc);
    }
    int rc = orc_lusgs_sweep(m, cfg, Q, R, colour, cfl, omega, dQ);
    if (rc == 0)
        for (size_t i = 0; i < n; i++) Q[i] += dQ[i];
    free(R);
    free(dQ);
    return rc;
}
