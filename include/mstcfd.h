/*
 * mstcfd.h -- C ABI of the MI355X (gfx950) density-based explicit residual path of MST-CFD.
 *
 * The reference's solver "plugin" is duck-typed C++ (rhoSolver/RhoSolver.h:14-24): a class built
 * from (MshBlock*, fstream*, AllData*) with setDT / solve / getOldValue / getNewValue /
 * updateNewToOld, selected in Time::goNextTimeStep (time/Time.cpp:57-61).  This header is the
 * C boundary that a GPU RhoSolver binds; every entry point names the reference interface it
 * replaces.  Plain pointers and sizes only.  Conventions:
 *   - return 0 on success, a negative MST_E_* code on error; mst_last_error() gives the text;
 *   - host arrays are owned by the caller; state arrays use the reference's AllData layout
 *     (cell-major, nv = dim+2 doubles per cell: rho, rho*u_1..rho*u_dim, rho*E) in FILE cell order;
 *   - a context owns its device buffers and one HIP stream; calls are stream-ordered, and the
 *     mst_get_* calls synchronise; one context per host thread (no internal locking).
 */
#ifndef MSTCFD_H
#define MSTCFD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MST_OK 0
#define MST_E_ARG (-1)      /* invalid argument / size mismatch */
#define MST_E_IO (-2)       /* file cannot be opened or parsed */
#define MST_E_HIP (-3)      /* HIP runtime error (no device, launch failure, out of memory) */
#define MST_E_MESH (-4)     /* mesh inconsistent (bad cell ids, open cell, unsupported element) */
#define MST_E_STATE (-5)    /* non-physical state detected (rho<=0, p<=0 or NaN) */
#define MST_E_UNSUPPORTED (-6)

/* Fluent boundary types as the reference's solver switches on them (rhoSolver/RhoSolver.cpp:97-230);
 * PERIODIC (12) is new: the ghost is the partner face's cell. */
#define MST_BC_INTERIOR 2
#define MST_BC_WALL 3
#define MST_BC_OUTLET 5
#define MST_BC_SYMMETRY 7
#define MST_BC_INLET 10
#define MST_BC_PERIODIC 12

#define MST_SCHEME_ROE 0   /* RHOSOLVER solverRoe   (rhoSolver/SolverRoe.h:9) */
#define MST_SCHEME_AUSM 1  /* RHOSOLVER SolverAusm  (rhoSolver/SolverAusm.h:9) */
#define MST_LIM_NONE 0     /* the reference's unlimited ACCURACY 2 reconstruction */
#define MST_LIM_MINMOD 1
#define MST_LIM_VANLEER 2

#define MST_PREC_FP64 0
#define MST_PREC_MIXED 1

/* reference behaviours behind switches (defaults: MST_COMPAT_REFERENCE) */
#define MST_COMPAT_BRANCH2_LR 1      /* literal inverted L/R of mesh/MshBlock.cpp:294-303 */
#define MST_COMPAT_GRAD_LASTINT 2    /* rhoSolver/RhoSolver.cpp:437-440 loop start N_int-1 */
#define MST_COMPAT_AUSM_LOCAL_A 4    /* rhoSolver/SolverAusm.cpp:112-113 */
#define MST_COMPAT_AUSM_P_UNSCALED 8 /* rhoSolver/SolverAusm.cpp:130,136 operator precedence */
#define MST_COMPAT_AUSM_ONESIDED 16  /* rhoSolver/SolverAusm.cpp:116,122,129,135 abs(M <= 1): subsonic
                                        polynomials for every M <= 1 (clear: |M| <= 1, AUSM+) */
#define MST_COMPAT_REFERENCE                                                                      \
    (MST_COMPAT_GRAD_LASTINT | MST_COMPAT_AUSM_LOCAL_A | MST_COMPAT_AUSM_P_UNSCALED | MST_COMPAT_AUSM_ONESIDED)

/* Runtime replacement of the compile-time macros of include/CONST.h:3-83
 * (RHOSOLVER :10, ACCURACY :6, GAMMA :41, EOR :42, inlet state :78-83). */
typedef struct {
    int32_t scheme;        /* MST_SCHEME_* */
    int32_t order;         /* 1 (ACCURACY 1) or 2 (ACCURACY 2) */
    int32_t limiter;       /* MST_LIM_* (order 2) */
    int32_t rk;            /* 1 forward Euler (time/Time.cpp:62-65), 2 SSP-RK2, 3 SSP-RK3 */
    int32_t compat;        /* MST_COMPAT_* */
    int32_t precision;     /* MST_PREC_FP64 (plain) or MST_PREC_MIXED (FP32 flux/limiter, FP64 gather+update) */
    double gamma;          /* 1.4 */
    double eor;            /* 1e-10 */
    double entropy_delta;  /* 0.125 (rhoSolver/SolverRoe.cpp:115) */
    double cfl;            /* > 0: mst_step computes dt = cfl * min V / sum (|u_d|+a) S_d; 0: use mst_set_dt */
    double q_inlet[5];     /* inlet ghost state (dim+2 used) */
    /* NEW viscous terms (the intent of RhoSolver::updateViscid, rhoSolver/RhoSolver.cpp:371-429; the
     * reference runs inviscid, FLAGVISCID 0): mu = kappa = 0 is inviscid.  Viscous runs use the face
     * strategy and FP64 or mixed single contexts (not batches). */
    double mu;             /* dynamic viscosity (include/CONST.h:46: 1.7894e-5) */
    double lam;            /* bulk coefficient (CONST.h:47: -0.666667 mu) */
    double kappa;          /* thermal conductivity (CONST.h:48 TEMPK: 0.0242) */
    double cv;             /* T = e / cv (CONST.h:39: 715.8); > 0 when viscous */
} mst_config;

typedef struct mst_mesh mst_mesh; /* host mesh in file order (MshBlock, mesh/MshBlock.h:19-76) */
typedef struct mst_ctx mst_ctx;   /* device solver context (RhoSolver + AllData device mirror) */

/* ---------------------------------------------------------------- errors */
/* last error text of the calling thread; returns the full length */
size_t mst_last_error(char *buf, size_t n);

/* ---------------------------------------------------------------- host mesh (integer work, bit-exact) */
/* MshBlock::readMsh (mesh/MshBlock.cpp:5-18, 2D :75-432, 3D :436-789): Fluent ASCII .msh */
int mst_mesh_read_msh(const char *path, mst_mesh **out);
/* inverse of the reader (Fluent ASCII, lowercase hex, the reference's zone layout) */
int mst_mesh_write_msh(const mst_mesh *m, const char *path);
/* mesh from raw arrays: faces in file order; c1[f] = -1 on boundary faces; ftype per face;
 * partner[f] = periodic partner face or -1 (may be NULL); fnn nodes per face */
int mst_mesh_from_arrays(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int64_t n_faces, int fnn,
                         const int32_t *fnodes, const int32_t *c0, const int32_t *c1, const int32_t *ftype,
                         const int32_t *partner, mst_mesh **out);
/* out[0..6] = dim, fnn, n_nodes, n_cells, n_faces, n_int (end of the interior zone, MshBlock.cpp:223-226),
 * n_zones */
int mst_mesh_counts(const mst_mesh *m, int64_t out[7]);
int mst_mesh_arrays(const mst_mesh *m, double *xyz, int32_t *fnodes, int32_t *c0, int32_t *c1, int32_t *ftype,
                    int32_t *partner);
/* zone table (FacesInf, mesh/FacesInf.h): n_zones x {zone id, first face, last face + 1, type} */
int mst_mesh_zones(const mst_mesh *m, int32_t *out);
void mst_mesh_destroy(mst_mesh *m);

/* Generators for the synthetic workloads (no reference counterpart; the reference only reads files).
 * 2D: nx x ny quads on [x0,x1]x[y0,y1], each split into 2 triangles, interior nodes jittered by
 * +-jitter*h (boundary nodes tangentially), seeded mt19937_64.  bc[4] = type of the x0, x1, y0, y1
 * sides (MST_BC_*; PERIODIC must be given on both sides of a pair).  Every face's area vector points
 * into its c0 cell. */
int mst_mesh_gen_tri(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
                     const int32_t bc[4], mst_mesh **out);
/* 3D: nx x ny x nz hexes on the box lo..hi, each split into 6 tets (Kuhn split); bc[6] = x0,x1,y0,y1,z0,z1 */
/* NEW (configuration C5): compression ramp, the gen_tri lattice mapped onto x in [x0, x1] (x0 < 0 < x1)
 * with a wall corner at x = 0 (bottom flat for x < 0, rising at theta_deg for x > 0) and a straight
 * top at y = height; bc = {left, right, bottom, top} */
int mst_mesh_gen_ramp(int nx, int ny, double x0, double x1, double height, double theta_deg, double jitter,
                      uint64_t seed, const int32_t bc[4], mst_mesh **out);
int mst_mesh_gen_tet(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
                     const int32_t bc[6], mst_mesh **out);

/* ---------------------------------------------------------------- initial conditions (file order)
 * Time::initialization (time/Time.cpp:13-36) and, new, the synthetic workloads' fields.  Q: n_cells*nv. */
#define MST_INIT_REFERENCE 0  /* params: Q0 (nv conserved); x_c > 0.5 gets rho*0.125, rho*E*0.1 (Time.cpp:26-36) */
#define MST_INIT_UNIFORM 1    /* params: gamma, rho, u_1..u_dim, p */
#define MST_INIT_RIEMANN 2    /* params: gamma, axis, x0, rhoL, uL(dim), pL, rhoR, uR(dim), pR (left: x_c[axis] <= x0) */
#define MST_INIT_VORTEX 3     /* 2D, params: gamma, mach, beta, xc, yc (Shu vortex on rho_inf = p_inf = 1) */
#define MST_INIT_PULSE 4      /* params: gamma, amp, sigma, centre(dim), mach: rho = 1+b, p = (1+b)/gamma, u = mach e_x */
#define MST_INIT_FREESTREAM 5 /* params: gamma, mach, angle (deg): rho = 1, p = 1/gamma, |u| = mach */
int mst_mesh_init_state(const mst_mesh *m, int kind, const double *params, int n_params, double *Q, int64_t n_cells);
/* cell centres (mean of face centres, mesh/Cell.cpp:6-14), n_cells*dim, file order */
int mst_mesh_cell_centres(const mst_mesh *m, double *ccen, int64_t n_cells);

/* ---------------------------------------------------------------- host preprocessing (no device needed)
 * The one-time part of the reference's mesh set-up (mesh/MshBlock.cpp:272-334, mesh/Face.cpp:8-69,
 * mesh/Cell.cpp:6-61) plus the device layout: geometry in FP64, Hilbert-curve (or reverse
 * Cuthill-McKee) cell order, faces sorted by kind then lowest cell, greedy face colouring, slot-major
 * cell->face table. */
typedef struct mst_prep mst_prep;
/* NEW: cell order of every later preprocessing (process-wide, default HILBERT).  HILBERT sorts the
 * cells along a Hilbert curve through their centres, so each run of consecutive cells is a compact
 * patch (most faces of a 256-cell block join two cells of the block); RCM minimises the bandwidth of
 * the cell graph.  Results agree to rounding (the gather order of a cell's faces is unchanged). */
#define MST_ORDER_HILBERT 0
#define MST_ORDER_RCM 1
int mst_set_cell_order(int order);
int mst_get_cell_order(int *order);
int mst_prep_create(const mst_mesh *m, mst_prep **out);
void mst_prep_destroy(mst_prep *p);
/* out[0..7] = dim, nv, n_cells, n_faces, max faces per cell, colours, interior faces, periodic faces */
int mst_prep_counts(const mst_prep *p, int64_t out[8]);
/* new index of each file-order cell / face, and the colour of each file-order face */
int mst_prep_permutation(const mst_prep *p, int32_t *cell_new_of_old, int32_t *face_new_of_old,
                         int32_t *face_colour);
/* geometry in file order, as the reference stores it on Face/Cell; any pointer may be NULL */
int mst_prep_geometry(const mst_prep *p, double *fcen, double *S, double *eta, int32_t *dac, double *ccen,
                      double *vol);
/* device-order integer tables: fc0, fc1 (n_faces), kind (n_faces), slot-major cell->face table
 * (maxf*n_cells, (face<<1)|is_c1, -1 padding); any pointer may be NULL */
int mst_prep_tables(const mst_prep *p, int32_t *fc0, int32_t *fc1, int32_t *kind, int32_t *cell_faces);
/* NEW: the cell colouring of the LU-SGS sweeps (no two cells sharing a face alike; two colours when
 * the cell graph is bipartite, else greedy in device order + Kempe-chain reduction of the top colour),
 * file order (colour[n_cells]), and the number of colours */
int mst_prep_cell_colours(const mst_prep *p, int32_t *colour, int *n_colours);

/* ---------------------------------------------------------------- on-disk formats (file order)
 * Tecplot ASCII result file: Work::writedataRhoBasedMshNodePlt (work/Work.cpp:204-304) -- cell
 * states carried to faces (interior eta0-weighted, boundary the cell's value, walls with zero
 * velocity when wall_noslip, as FLAG_WALLSMOOTH = 0), then to nodes (1/area-weighted over the node's
 * faces), written as X Y [Z] rho u v [w] T p Ma per node (T = e/cv, include/CONST.h:39) plus the
 * cells' nodes.  Fixed against the reference: zone type from the mesh (triangles, quadrilaterals,
 * tetrahedra; :231-235 uses the FELNUM macro), each node's own face weights (:292-293 indexes faces
 * by node number), values on symmetry and periodic faces, a well-formed TITLE and a 3D w column. */
int mst_write_tecplot(const mst_mesh *m, const double *Q, int64_t n_cells, double gamma, double cv, int wall_noslip,
                      const char *title, const char *path);
/* NEW: binary mesh cache (MshBlock::readMsh's result in file order, zones included), so a large
 * Fluent mesh is parsed once: written through a temporary file; read back bit for bit, checksummed
 * and validated like the reader's output (MST_E_IO for a damaged or foreign file). */
int mst_mesh_write_cache(const mst_mesh *m, const char *path);
int mst_mesh_read_cache(const char *path, mst_mesh **out);
/* Binary checkpoint replacing the text .cellPUVT dump (Work::writedataCellContinue /
 * readPbasedCellPlt, work/Work.cpp:432-514): the state (n_cells x nv doubles, file order) with a step
 * counter and a time, a mesh fingerprint and a state checksum.  Reading restores the state bit for
 * bit; a file for another mesh is MST_E_ARG, a damaged or truncated one MST_E_IO.  The write goes
 * through a temporary file renamed into place. */
int mst_checkpoint_write(const char *path, const mst_mesh *m, const double *Q, int64_t n_cells, int64_t step,
                         double time);
int mst_checkpoint_read(const char *path, const mst_mesh *m, double *Q, int64_t n_cells, int64_t *step, double *time);

/* ---------------------------------------------------------------- solver context
 * RhoSolver::RhoSolver (rhoSolver/RhoSolver.cpp:3-26) + AllData::createAllData (data/AllData.cpp:3-28):
 * preprocesses the mesh (as mst_prep_create), uploads the SoA arrays and allocates the state
 * buffers on the current HIP device.  Fails with MST_E_HIP when no device is present. */
int mst_ctx_create(const mst_mesh *m, const mst_config *cfg, mst_ctx **out);
int mst_ctx_destroy(mst_ctx *ctx);
/* replace the configuration (same mesh); rebuilds the per-face L/R masks */
int mst_ctx_set_config(mst_ctx *ctx, const mst_config *cfg);
/* the context's preprocessing (owned by the context) */
const mst_prep *mst_ctx_prep(const mst_ctx *ctx);

/* ---------------------------------------------------------------- state (AllData p1Old/p1NewCellQs) */
/* write p1OldCellQs (data/AllData.h getP1OldCellQs); Q: n_cells*nv doubles, file order */
int mst_set_state(mst_ctx *ctx, const double *Q, int64_t n_cells);
/* RhoSolver::getOldValue (rhoSolver/RhoSolver.cpp:504-506) */
int mst_get_state(mst_ctx *ctx, double *Q, int64_t n_cells);
/* RhoSolver::getNewValue (rhoSolver/RhoSolver.cpp:507-509) */
int mst_get_new_state(mst_ctx *ctx, double *Q, int64_t n_cells);

/* ---------------------------------------------------------------- hot path */
/* RhoSolver::setDT (rhoSolver/RhoSolver.cpp:33-35) */
int mst_set_dt(mst_ctx *ctx, double dt);
/* RhoSolver::updateFaceFlux / updateFaceFlux2ndOrder (+ updateGradFlux) (rhoSolver/RhoSolver.cpp:90-369,
 * 430-454) on the Old state: per face G_f = sum_d n_out,d(c0) F_f[:,d] (device-resident) */
int mst_face_flux(mst_ctx *ctx);
/* the gather + forward-Euler update of RhoSolver::solve (rhoSolver/RhoSolver.cpp:45-68): New = Old - DT/V R */
int mst_residual_update(mst_ctx *ctx);
/* RhoSolver::solve (rhoSolver/RhoSolver.cpp:37-89) = mst_face_flux + mst_residual_update */
int mst_solve(mst_ctx *ctx);
/* RhoSolver::updateNewToOld (rhoSolver/RhoSolver.cpp:513-517); a pointer swap on the device */
int mst_update_new_to_old(mst_ctx *ctx);
/* Time::goNextTimeStep (time/Time.cpp:54-80) generalised: nsteps steps of cfg->rk stages, dt from cfl
 * (device-side reduction) when cfg->cfl > 0 else the last mst_set_dt value.  Old holds the result. */
int mst_step(mst_ctx *ctx, int nsteps);
/* NEW: kernel strategy of mst_step and mst_solve.  FACE: a face kernel writes every face flux once,
 * a cell kernel gathers them and updates (the reference's updateFaceFlux + solve split).  CELL: one
 * fused cell kernel solves each cell's own faces (every interior face twice, bitwise the same flux
 * on both sides) and updates, with no face-flux array in memory.  PATCH: one kernel per block of 256
 * consecutive cells (a compact patch in the Hilbert cell order) solves the patch's faces once into
 * LDS, then sums and updates per cell; available for cells of at most 4 faces (MST_E_UNSUPPORTED
 * otherwise).  FUSED (order 2; order 1 runs PATCH): the patch kernel also computes the limited
 * gradients of its cells and of its halo and keeps the face states in LDS, one kernel per stage.
 * All give the same result bit for bit.  AUTO (default): the first stage after creation
 * or a configuration change times one stage of each available strategy on the context's own mesh
 * and keeps the fastest; mst_get_strategy then reports it.  mst_face_flux / mst_residual_update keep
 * the face split whatever the strategy. */
#define MST_STRATEGY_FACE 0
#define MST_STRATEGY_CELL 1
#define MST_STRATEGY_AUTO 2
#define MST_STRATEGY_PATCH 3
#define MST_STRATEGY_FUSED 4
int mst_set_strategy(mst_ctx *ctx, int strategy);
int mst_get_strategy(const mst_ctx *ctx, int *strategy);
/* NEW: steady driver.  Steps of mst_step (or LU-SGS iterations after
 * mst_set_steady_driver(MST_STEADY_LUSGS)) until max_f L2(R/V)_f / ref_f <= tol (checked every
 * check_every steps) or max_steps.  ref[nv]: reference norms; when ref[0] <= 0 on entry the first
 * step's norms are taken and written back.  steps_done / drop: steps taken and the final ratio. */
int mst_run_to_steady(mst_ctx *ctx, double tol, int max_steps, int check_every, double *ref, int *steps_done,
                      double *drop);
/* NEW: implicit steady driver.  The reference's LU-SGS (SparseSolverNUM::solveILUSGS,
 * lusolver/SparseSolverNUM.cpp:144-196) factorises assembled scalar systems of its pressure-based
 * path; here the same (D+L) D^-1 (D+U) factorisation is applied matrix-free to the density-based
 * residual (Jameson-Yoon scalar diagonal, Rusanov-split face Jacobians), over a greedy colouring of
 * the cell graph: one launch per colour and sweep direction, deterministic.
 * Per iteration: R(Q) with the configured scheme/order/limiter (its L2(R/V) norms as mst_step's),
 * r_f = max over the face's cells of |u.S_f| + a|S_f|, D_i = (1/cfl + omega/2) sum_f r_f,
 * forward sweep (-R - sum_L) / D, backward sweep, Q += dQ.  cfl is the local pseudo-time CFL number
 * (large values approach Newton-like LU-SGS); omega >= 1 keeps D dominant.
 * mst_set_steady_driver: MST_STEADY_EXPLICIT (mst_step, the default) or MST_STEADY_LUSGS for
 * mst_run_to_steady. */
#define MST_STEADY_EXPLICIT 0
#define MST_STEADY_LUSGS 1
int mst_set_steady_driver(mst_ctx *ctx, int driver, double cfl, double omega);
/* n_iter LU-SGS iterations (Old -> New, then swap, like mst_step); cfl/omega from
 * mst_set_steady_driver (defaults 1e4 and 1.0) */
int mst_lusgs_step(mst_ctx *ctx, int n_iter);
/* the cell colouring the sweeps use (as mst_prep_cell_colours), file order, and the number of colours */
int mst_lusgs_colours(mst_ctx *ctx, int32_t *colour, int64_t n_cells, int *n_colours);
/* NEW: dt = cfl * min_c V_c / sum_d (|u_d| + a) Shalf_c,d on the Old state */
int mst_compute_dt(mst_ctx *ctx, double cfl, double *dt);
/* NEW: L2 norm over cells of R(Q_old)/V per field (nv values) from the first stage of the last step of
 * the last mst_step call (or the last iteration of mst_lusgs_step): reduced once per call */
int mst_residual_norms(mst_ctx *ctx, double *out);
/* Time::goNextTimeStep monitor (time/Time.cpp:69-76): max_i |New-Old|/Old per field */
int mst_monitor_maxrel(mst_ctx *ctx, double *out);
/* NEW, no counterpart in the reference: the same monitor over caller-owned DEVICE arrays (cell-major,
 * nv doubles per cell; 16-byte aligned when nv == 4), writing its nv maxima to the device array out,
 * stream-ordered on stream (a hipStream_t; NULL: the default stream), without synchronising.  For
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops): it calls no
 * allocator; its scratch is a fixed 80 KiB array per device (2048 blocks x 5 partials), held until the
 * process ends.  n_cells up to MST_NOALLOC_MAX_CELLS.  Calls that overlap on one device race on that
 * scratch. */
#define MST_NOALLOC_MAX_CELLS ((int64_t)1 << 40)
int mst_monitor_maxrel_noalloc(const double *qnew, const double *qold, int64_t n_cells, int nv, double *out,
                               void *stream);

/* ---------------------------------------------------------------- inspection (tests) */
/* residual R_i = sum_f sum_d n_out,d F_f[:,d] (rhoSolver/RhoSolver.cpp:51-55) of the Old state, file order */
int mst_get_residual(mst_ctx *ctx, double *R, int64_t n_cells);
/* per-face G_f of the last mst_face_flux, file order, n_faces*nv */
int mst_get_face_flux(mst_ctx *ctx, double *G, int64_t n_faces);
/* gradient (n_cells*nv*dim) and limiter factor (n_cells*nv) of the Old state, file order */
int mst_get_gradient(mst_ctx *ctx, double *grad, double *phi, int64_t n_cells);

/* flux-scheme plugin, batched on the device: RHOSOLVER::set(L,R) + solverAll(dir[i])
 * (rhoSolver/SolverRoe.cpp:3-16,70-112; rhoSolver/SolverAusm.cpp:3-26,53-63,105-143).
 * QL, QR: n*(dim+2); dir: n; F: n*(dim+2); host arrays. */
int mst_riemann_flux(const mst_config *cfg, int dim, int64_t n, const double *QL, const double *QR,
                     const int32_t *dir, double *F);

/* ---------------------------------------------------------------- batches (NEW, no reference counterpart:
 * the reference runs one mesh per process, work/Work.cpp:53-89)
 * Many independent cases advanced together.  A step of every active case is one pass of persistent
 * kernels over a work list of (case, tile) items -- gradient + limiter, fluxes (Roe or AUSM per
 * case), gather + stage update -- and one reduction block per case; the whole RK step is replayed
 * from a hipGraph.  Cases share dim, order, limiter, rk and precision; scheme, compat, gamma, eor,
 * entropy_delta, q_inlet and cfl (> 0: each case advances with its own CFL time step) are per case.
 * Periodic faces are not supported.  States use each case's file cell order.  Results do not depend
 * on the tiling and equal the single-case context's. */
typedef struct mst_batch mst_batch;
#define MST_BATCH_BUCKETS 4 /* cases binned by cell count: < 2048, < 8192, < 32768, >= 32768 */
int mst_batch_create(int n_cases, const mst_mesh *const *meshes, const mst_config *cfgs, mst_batch **out);
int mst_batch_destroy(mst_batch *b);
/* out[6]: cases, total cells, total faces, cell work items, face work items, dim */
int mst_batch_counts(const mst_batch *b, int64_t *out);
int mst_batch_case_counts(const mst_batch *b, int case_id, int64_t *n_cells, int64_t *n_faces);
int mst_batch_set_state(mst_batch *b, int case_id, const double *Q, int64_t n_cells);
int mst_batch_get_state(mst_batch *b, int case_id, double *Q, int64_t n_cells);
/* nsteps RK steps of every active case */
int mst_batch_step(mst_batch *b, int nsteps);
/* per case (n_cases*nv): L2 over cells of R/V from the first stage of the last step */
int mst_batch_residual_norms(mst_batch *b, double *out);
/* per case: the CFL time step the next step will take */
int mst_batch_dt(mst_batch *b, double *out);
/* active[n_cases]: inactive cases keep their state and are left out of the work list */
int mst_batch_set_active(mst_batch *b, const uint8_t *active);
/* NEW: the LU-SGS steady driver over a whole batch (as mst_set_steady_driver / mst_lusgs_step for one
 * context): every active case's colour-k cells in one launch per colour and sweep direction; each
 * case's result equals the same case in its own context bit for bit.  FP64 batches of cells with at
 * most 4 faces.  mst_batch_run_to_steady then runs LU-SGS iterations. */
int mst_batch_set_steady_driver(mst_batch *b, int driver, double cfl, double omega);
int mst_batch_lusgs_step(mst_batch *b, int n_iter);
/* steady driver for every active case: a case that reaches max_f L2(R/V)_f / ref_f <= tol is frozen;
 * ref [n_cases*nv] in/out (a case with ref <= 0 takes its first step's norms), steps [n_cases] and
 * drop [n_cases] out.  The activity flags are restored at the end. */
int mst_batch_run_to_steady(mst_batch *b, double tol, int max_steps, int check_every, double *ref, int32_t *steps,
                            double *drop);
/* launch shape: block size (64, 128, 256) and cell tile per case-size bucket (faces: scaled by nf/nc) */
int mst_batch_set_tuning(mst_batch *b, int block, const int32_t *tiles);
int mst_batch_get_tuning(const mst_batch *b, int *block, int32_t *tiles);
/* kernel strategy of the batched step: MST_STRATEGY_FACE, MST_STRATEGY_CELL (default),
 * MST_STRATEGY_PATCH (one work item per 256-cell patch, its fluxes in LDS) or MST_STRATEGY_FUSED (the
 * patch item also computes its cells' and halo's gradients, order 2); block and tiles do not apply
 * to the patch strategies, which return MST_E_UNSUPPORTED for cells of more than 4 faces.  The
 * autotuner searches all four; the results are the same bit for bit. */
int mst_batch_set_strategy(mst_batch *b, int strategy);
int mst_batch_get_strategy(const mst_batch *b, int *strategy);
/* 1: replay the captured step graph (default); 0: launch the kernels one by one */
int mst_batch_set_graph(mst_batch *b, int on);
/* measured search over strategy, block size and per-bucket tiles (steps timed per trial); the best shape is
 * applied and stored in the JSON file cache_path under a key describing the batch; a cached entry
 * is applied without measuring unless force.  The state is restored afterwards. */
int mst_batch_autotune(mst_batch *b, const char *cache_path, int force, int steps, double *ms_per_step);

/* ---------------------------------------------------------------- learned warm start (NEW, no reference
 * counterpart).  A fully connected network n_in -> width (tanh) x hidden_layers -> n_out (linear),
 * FP32, trained and evaluated by the library's own kernels.  Parameters: per layer W [out x in]
 * row-major, then b [out].  Features are normalised (f - in_shift) * in_scale, outputs de-normalised
 * y * out_scale + out_shift; the loss is the mean squared error of the de-normalised outputs. */
typedef struct mst_mlp mst_mlp;
int mst_mlp_create(int n_in, int width, int hidden_layers, int n_out, uint64_t seed, mst_mlp **out);
int mst_mlp_destroy(mst_mlp *m);
int mst_mlp_num_params(const mst_mlp *m, int64_t *n);
int mst_mlp_get_params(mst_mlp *m, float *p, int64_t n);
int mst_mlp_set_params(mst_mlp *m, const float *p, int64_t n);
int mst_mlp_set_scaling(mst_mlp *m, const float *in_shift, const float *in_scale, const float *out_shift,
                        const float *out_scale);
/* training set, host arrays X [n x n_in], T [n x n_out] (de-normalised targets), copied to the device */
int mst_mlp_set_data(mst_mlp *m, const float *X, const float *T, int64_t n);
/* zero the Adam moments and the step counter */
int mst_mlp_reset_optimizer(mst_mlp *m);
/* steps of Adam (beta 0.9 / 0.999, eps 1e-8) on batches of `batch` samples drawn by a counter-based
 * hash of (step, row) (the whole set in order when batch >= n); use_graph: replay one captured step
 * (dropped whenever the training set or a buffer it uses is replaced; launched directly instead
 * while a rocprofiler-sdk tool is loaded in the process).  loss: mean squared error of the last batch */
int mst_mlp_train(mst_mlp *m, int steps, int batch, double lr, int use_graph, double *loss);
/* de-normalised outputs Y [n x n_out] of X [n x n_in] (host arrays) */
int mst_mlp_forward(mst_mlp *m, const float *X, int64_t n, float *Y);
/* per-dispatch timing of the network's kernels (start/stop HIP events carried by each launch; training
 * then launches directly instead of replaying its graph): count, summed milliseconds and summed
 * floating-point operations (GEMMs: 2 M N K) per kernel class */
#define MST_MLP_K_GEMM_FWD 0   /* forward GEMMs (bias + tanh epilogues) */
#define MST_MLP_K_GEMM_WGRAD 1 /* weight-gradient GEMMs (batch split) */
#define MST_MLP_K_GEMM_DGRAD 2 /* activation-gradient GEMMs (tanh' epilogue) */
#define MST_MLP_K_OTHER 3      /* batch gather, loss gradient, split sums, bias gradients, Adam */
#define MST_MLP_K_COUNT 4
int mst_mlp_timing_enable(mst_mlp *m, int on);
int mst_mlp_timing_get(mst_mlp *m, int kclass, int64_t *count, double *total_ms, double *total_flop);
int mst_mlp_timing_reset(mst_mlp *m);
/* gradient of the mean squared error over (X, T) w.r.t. the parameters, and the loss (host arrays) */
int mst_mlp_gradient(mst_mlp *m, const float *X, const float *T, int64_t n, float *grad, double *loss);
/* Old state := decoded network prediction at the cell centres.  Features: (x_c (dim), cond[0..n_cond)),
 * n_in = dim + n_cond, n_out = nv.  Outputs y decode as rho = exp(y_0), u_d = cond[0] * y_{1+d},
 * p = exp(y_{nv-1}) / gamma (freestream rho = 1, p = 1/gamma, |u| = cond[0] = Mach), so density and
 * pressure are positive; a NaN prediction returns MST_E_STATE (the caller keeps the freestream start). */
int mst_ctx_warm_start(mst_ctx *ctx, mst_mlp *m, const double *cond, int n_cond);

/* ---------------------------------------------------------------- timing (HIP events on the ctx stream) */
#define MST_K_GRAD 0
#define MST_K_FLUX 1
#define MST_K_UPDATE 2
#define MST_K_DT 3
#define MST_K_CELL 4 /* fused flux + gather + update (MST_STRATEGY_CELL) */
#define MST_K_PATCH 5 /* patch fluxes in LDS + gather + update (MST_STRATEGY_PATCH) */
#define MST_K_FUSED 6 /* gradients + face states + fluxes in LDS + gather + update (MST_STRATEGY_FUSED) */
#define MST_K_LUSGS 7 /* LU-SGS colour sweeps (mst_lusgs_step); algorithmic bytes: per launch, averaged */
#define MST_K_VISC 8 /* viscous gradient + face kernels (mu or kappa > 0); algorithmic bytes: per launch, averaged */
#define MST_K_COUNT 9
/* on: 0 off, 1 events on every launch, n > 1 on one launch in n of each kernel class (counted from
   this call; a timed dispatch costs a few microseconds of throughput) */
int mst_timing_enable(mst_ctx *ctx, int on);
/* per kernel class: launches timed and total milliseconds since the last reset */
int mst_timing_get(mst_ctx *ctx, int kernel, int64_t *count, double *total_ms);
int mst_timing_reset(mst_ctx *ctx);
int mst_sync(mst_ctx *ctx);
/* measurement helper (no reference counterpart): one launch of a plain device copy of nbytes
 * (read nbytes + write nbytes), the calibration point for hardware byte counters; ms = its time */
int mst_copy_probe(int64_t nbytes, double *ms);
/* algorithmic (compulsory) HBM bytes of one launch of each kernel class for this mesh/config */
int mst_algorithmic_bytes(const mst_ctx *ctx, int kernel, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
