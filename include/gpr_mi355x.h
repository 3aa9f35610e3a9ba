/* C ABI of libgpr_mi355x.so: the ADER-WENO time step of the GPR system on one
 * AMD Instinct MI355X (gfx950).
 *
 * Drop-in replacement for the pybind11 module GPRpy built from
 * code/cpp/bindings/bindings.cpp (haranjackson/phd).  Each entry point names
 * the reference function it replaces.  Conventions:
 *   - The caller owns every array; results are written in place, in the
 *     reference's C-order layouts:
 *       u   [nx(,ny)][V]               cells
 *       ub  [(nx+2N)(,ny+2N)][V]       ghost-extended by N per side
 *       wh  [(nx+2)(,ny+2)][N^ndim][V]
 *       qh  [(nx+2)(,ny+2)][N^(ndim+1)][V]
 *       mask[(nx+2)(,ny+2)] uint8
 *   - Functions without _dev take host pointers, copy to the device, run on
 *     the context's stream and synchronise before returning.  *_dev functions
 *     take device pointers and only enqueue work (the caller synchronises).
 *   - Return 0 on success, a negative GPR_E_* code otherwise; the message is
 *     in gpr_last_error() (per thread).
 *   - N (Gauss-Legendre nodes per dimension, 2 or 3), ndim (1 or 2) and
 *     thermal (V = 14 or 17) are fixed per context; the reference fixes them
 *     at compile time (code/cpp/src/options.h:5-35).
 */
#ifndef GPR_MI355X_H
#define GPR_MI355X_H

#include <stdint.h>

#include "gpr_par.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GPR_OK 0
#define GPR_E_ARG (-1)          /* bad argument / shape */
#define GPR_E_UNSUPPORTED (-2)  /* option outside this path (e.g. nmat > 1) */
#define GPR_E_HIP (-3)          /* HIP runtime error */
#define GPR_E_EOS (-4)          /* EOS code not recognised (mg.cpp:27) */
#define GPR_E_PREDICTOR (-5)    /* a cell's predictor converged in neither phase */
#define GPR_E_NONFINITE (-6)    /* state became non-finite (iterator.cpp:161-165) */

typedef struct gpr_ctx gpr_ctx;

/* Solver settings beyond the reference's compile-time constants. */
typedef struct gpr_solver_opts {
  int32_t dg_it;      /* Picard cap, reference DG_IT = 50 (options.h:31) */
  double dg_tol;      /* Picard tolerance, reference DG_TOL = 6e-6 (options.h:32) */
  int32_t stiff_it;   /* cap of the preconditioned fallback iteration */
  double stiff_tol;   /* its stopping tolerance (same rule as dg_tol) */
  int32_t bailout;    /* 0: Picard runs to dg_it as in the reference; k > 0:
                         leave it once its stopping statistic grew k times in a row */
  int32_t mode;       /* 0: Picard, then the fallback for cells it leaves, in
                            the same wavefront;
                         1: the preconditioned iteration for every cell;
                         2: as 0 with the same values, the device loop running
                            the fallback as a second launch over a list of the
                            cells Picard leaves */
  double contorted_tol; /* reset threshold of iterator(), its contorted_tol argument */
  int32_t precision;  /* device loop (gpr_dev_steps): 0 FP64 throughout;
                         1 FP32 bulk: WENO and Picard in FP32, the stiff
                         fallback, FV accumulation, post-step and CFL in FP64.
                         gpr_predictor without STIFF: 1 rounds wh to float and
                         runs the same two predictor launches */
  int32_t split;      /* device loop: 0 the ADER step; 1 the split-WENO step
                         (steppers.cpp:50-66, iterator's SPLIT), FP64 only */
  int32_t half_step;  /* with split: the midstepper between WENO and FV (HALF_STEP) */
  int32_t bail_from;  /* 0: off; j > 0: a cell that has run j Picard iterations
                         without converging leaves Picard the first time its
                         stopping statistic grows (in addition to bailout) */
  int32_t flux;       /* device loop: the FV Riemann solver, 0 Rusanov, 1 Roe,
                         2 Osher (fluxes.h:7-9); iterator() sets it from FLUX */
} gpr_solver_opts;

const char *gpr_last_error(void);
int gpr_version(void);

/* Context: device, stream, constants and work buffers. */
int gpr_ctx_create(gpr_ctx **ctx, int N, int ndim, int thermal);
int gpr_ctx_destroy(gpr_ctx *ctx);
int gpr_ctx_set_opts(gpr_ctx *ctx, const gpr_solver_opts *opts);
int gpr_ctx_get_opts(const gpr_ctx *ctx, gpr_solver_opts *opts);
/* hipStream_t of the context (as void*), for callers that share device work */
void *gpr_ctx_stream(gpr_ctx *ctx);

/* options.N(), options.NV(), ... (bindings.cpp:83-88) */
int gpr_options(const gpr_ctx *ctx, int32_t *N, int32_t *V, int32_t *ndim, int32_t *thermal);

/* Basis constants (etc/globals.cpp:16-50; bindings.cpp:217-221 and
 * solvers.weno.coefficient_matrices / oscillation_indicator, :236-237).
 * Arrays of N, N, 2N and N*N (row-major) doubles. */
int gpr_constants(int N, double *nodes, double *wghts, double *endvals, double *dervals,
                  double *mLinv, double *mRinv, double *mCLinv, double *sig, double *dg_end,
                  double *dg_der, double *dg_mat);

/* --- pointwise physics over `count` states (system.*, bindings.cpp:197-206) */
/* flux (equations.cpp:11-48): ret[count][V] = F_d(Q) (written, not accumulated) */
int gpr_flux(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP);
/* source (equations.cpp:50-67) */
int gpr_source(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, const gpr_par *MP);
/* Bdot (equations.cpp:149-167): ret = B_d(Q) x */
int gpr_bdot(gpr_ctx *ctx, double *ret, const double *Q, const double *x, int64_t count, int d,
             const gpr_par *MP);
/* max_abs_eigs (eig.cpp:66-77): ret[count] */
int gpr_max_abs_eigs(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP);
/* thermo_acoustic_tensor (eig.cpp:55-64): ret[count][n][n], n = 3 (4 if thermal) */
int gpr_thermo_acoustic_tensor(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, int d,
                               const gpr_par *MP);
/* solvers.fv.D_RUS (fluxes.cpp:66-75) and solvers.fv.Bint (fluxes.cpp:9-20) */
int gpr_d_rus(gpr_ctx *ctx, double *ret, const double *qL, const double *qR, int64_t count, int d,
              const gpr_par *MP);
int gpr_bint(gpr_ctx *ctx, double *ret, const double *qL, const double *qR, int64_t count, int d,
             const gpr_par *MP);
/* solvers.fv.D_ROE (fluxes.cpp:45-64) and solvers.fv.D_OSH (fluxes.cpp:22-43),
   bindings.cpp:246-247: ret[count][V]; GPR_E_NONFINITE if a QR iteration did
   not converge (that result is NaN) */
int gpr_d_roe(gpr_ctx *ctx, double *ret, const double *qL, const double *qR, int64_t count, int d,
              const gpr_par *MP);
int gpr_d_osh(gpr_ctx *ctx, double *ret, const double *qL, const double *qR, int64_t count, int d,
              const gpr_par *MP);
/* system.system_matrix (equations.cpp:169-176), system.dFdP and system.dPdQ
   (jacobians.cpp:10-152), bindings.cpp:201,208-209: ret[count][V][V] row-major */
int gpr_system_matrix(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP);
int gpr_dfdp(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP);
int gpr_dpdq(gpr_ctx *ctx, double *ret, const double *Q, int64_t count, const gpr_par *MP);

/* --- grid (boundaries / extend_mask, grid.cpp:147-230, bindings.cpp:66-67) */
int gpr_boundaries(gpr_ctx *ctx, const double *u, double *ub, const int32_t *nX, const int32_t *bcs);
int gpr_extend_mask(gpr_ctx *ctx, const uint8_t *mask, uint8_t *maskb, const int32_t *nX);

/* --- the three stages (bindings.cpp:235-253; steppers.cpp:10-22) */
/* solvers.weno.weno_launcher(wh, ub, nX, MP) (weno.h:6) */
int gpr_weno_launcher(gpr_ctx *ctx, double *wh, const double *ub, const int32_t *nX, const gpr_par *MP);
/* solvers.dg.predictor(qh, wh, dt, dX, STIFF, STIFF_IG, MP, mask) (dg.h:7-8);
 * ncell = mask length (ghost-extended cells), as the reference takes it from
 * the mask.  iters (optional, [ncell]) receives the Picard count, 1000 + the
 * fallback count, or -1 for a failed cell.  STIFF selects mode 1.  With the
 * solver option precision = 1 (and STIFF = 0), the FP32-bulk predictor. */
int gpr_predictor(gpr_ctx *ctx, double *qh, const double *wh, double dt, const double *dX,
                  int64_t ncell, int stiff, int stiff_ig, const gpr_par *MP, const uint8_t *mask,
                  int32_t *iters);
/* solvers.fv.fv_launcher(u, rec, nX, dt, dX, SOURCES, TIME, FLUX, MP, mask)
 * (fv.h:17-18), FLUX = 0 Rusanov, 1 Roe, 2 Osher (fluxes.h:7-9).  TIME = 1:
 * rec is the predictor's qh ([cells+halo][N time][N^d space][V]); TIME = 0: one time level per cell
 * ([cells+halo][N^d][V], the split step's wh), weighted by dt (fv.cpp:46-54). */
int gpr_fv_launcher(gpr_ctx *ctx, double *u, const double *qh, const int32_t *nX, double dt,
                    const double *dX, int sources, int time, int flux, const gpr_par *MP,
                    const uint8_t *mask);
/* solvers.fv.centers1 / centers2 (fv.cpp:46-54, 99-109): the cell integrals
 * of fv_launcher alone, u += dt W (S - B_x dq/dx - B_y dq/dy) */
int gpr_fv_centers(gpr_ctx *ctx, double *u, const double *rec, const int32_t *nX, double dt,
                   const double *dX, int sources, int time, const gpr_par *MP, const uint8_t *mask);
/* solvers.fv.interfs1 / interfs2 (fv.cpp:151-160, 270-279): the face terms of
 * fv_launcher alone, with the Riemann solver FLUX */
int gpr_fv_interfs(gpr_ctx *ctx, double *u, const double *rec, const int32_t *nX, double dt,
                   const double *dX, int time, int flux, const gpr_par *MP, const uint8_t *mask);
/* solvers.timestep (iterator.cpp:54-83) for one material: *dt_out */
int gpr_timestep(gpr_ctx *ctx, double *dt_out, const double *u, const int32_t *nX, const double *dX,
                 double CFL, double t, double tf, int64_t count, const gpr_par *MP);
/* solvers.ader_stepper(u, ub, nX, dt, dX, STIFF, FLUX, MP, mask) (steppers.h:7-8) */
int gpr_ader_stepper(gpr_ctx *ctx, double *u, const double *ub, const int32_t *nX, double dt,
                     const double *dX, int stiff, int flux, const gpr_par *MP, const uint8_t *mask);
/* gpr_ader_stepper on device buffers without allocating.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * d_u (in/out, [cells][V]), d_ub ([ghost-extended cells][V]) and d_mask
 * ([cells + 1 ghost layer], uint8) are device pointers; nX, dX and MP are
 * host pointers.  The step's work buffers (WENO x-pass, wh, qh, face sums,
 * failure counter) live in a file-scope device array of
 * GPR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a grid
 * needing more (2D, N = 3, V = 14: about 4.9 KB per cell, so up to about
 * 13,000 cells) fails with GPR_E_ARG before any launch.  The call is
 * asynchronous on the context's stream; the number of cells whose predictor
 * converged in neither phase is written to the device int *d_nfail.
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (contexts on different streams) race and must be serialised by the
 * caller.  Values equal gpr_ader_stepper's for the same inputs.  FLUX must
 * be 0: the Roe and Osher solvers keep per-lane matrices in private memory,
 * which the HIP runtime allocates when such a kernel first runs. */
#define GPR_NOALLOC_SCRATCH_BYTES (64u << 20)
int gpr_ader_stepper_noalloc(gpr_ctx *ctx, double *d_u, const double *d_ub, const int32_t *nX,
                             double dt, const double *dX, int stiff, int flux, const gpr_par *MP,
                             const uint8_t *d_mask, int32_t *d_nfail);
/* solvers.split_stepper(u, ub, nX, dt, dX, HALF_STEP, FLUX, MP, mask)
 * (steppers.h:13-14, steppers.cpp:50-66): half-step source ODE on u and on ub
 * (both written in place), WENO of ub, the midstepper if HALF_STEP, the FV
 * update with TIME = SOURCES = 0, and the closing half-step ODE on u.
 * solvers.split_stepper_para is the same step on a private copy of ub. */
int gpr_split_stepper(gpr_ctx *ctx, double *u, double *ub, const int32_t *nX, double dt,
                      const double *dX, int half_step, int flux, const gpr_par *MP,
                      const uint8_t *mask);
/* solvers.split.ode_launcher(u, dt, MP) (solvers/split/ode.h:6, ode.cpp:5-11):
 * the closed-form source ODE (relaxation/analytic.cpp:154-162) over dt in each
 * of the ncell states of u.  Reactive materials (REACTION > -1) are refused. */
int gpr_ode_launcher(gpr_ctx *ctx, double *u, int64_t ncell, double dt, const gpr_par *MP);
/* solvers.split.analyticSolver_distortion (part 0, analytic.cpp:72-129) and
 * analyticSolver_thermal (part 1, analytic.cpp:131-152) of count states in place */
int gpr_analytic_solver(gpr_ctx *ctx, double *Q, int64_t count, double dt, const gpr_par *MP, int part);
/* solvers.split.midstepper(wh, dt, dX, MP, mask) (homogeneous.h:7,
 * homogeneous.cpp:5-67): wh ([ncell][N^d][V]) advanced by dt/2 under the
 * homogeneous system in each masked cell. */
int gpr_midstepper(gpr_ctx *ctx, double *wh, double dt, const double *dX, int64_t ncell,
                   const gpr_par *MP, const uint8_t *mask);
/* solvers.renorm_distortion / contorted / reset_distortion (utils.cpp) */
int gpr_renorm_distortion(gpr_ctx *ctx, double *u, int64_t ncell, const gpr_par *MP);
int gpr_contorted(gpr_ctx *ctx, int32_t *result, const double *u, int64_t ncell, double tol);
int gpr_reset_distortion(gpr_ctx *ctx, double *u, int64_t ncell, const gpr_par *MP);

/* solvers.iterator (iterator.cpp:85-170) for one material (LSET = 0),
 * FLUX = 0, 1 or 2; SPLIT = 1 takes the split-WENO step (with HALF_STEP) instead of
 * the ADER step.  u is advanced in place to tf.  snapshots
 * ([100][ncell*V], may be null) receives u at each 1% of tf as the reference's
 * returned list; *nsnap the number filled.  On a non-finite state the loop
 * stops as at iterator.cpp:161-165: the last two snapshots hold the previous
 * and the non-finite state, u holds the non-finite state, GPR_E_NONFINITE. */
int gpr_iterator(gpr_ctx *ctx, double *u, double tf, const int32_t *nX, const double *dX, double CFL,
                 const int32_t *bcs, int split, int half_step, int stiff, int flux, const gpr_par *MPs,
                 int nmat, double contorted_tol, double *snapshots, int32_t *nsnap, int64_t *steps);

/* --- device memory for callers without another allocator ------------------
 * (no reference counterpart: the reference keeps everything in host memory).
 * gpr_dev_memcpy is synchronous on the context's stream; kind 0 copies host
 * to device, 1 device to host, 2 device to device. */
int gpr_dev_malloc(gpr_ctx *ctx, void **d_ptr, int64_t bytes);
int gpr_dev_free(gpr_ctx *ctx, void *d_ptr);
int gpr_dev_memcpy(gpr_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);

/* --- device-resident time loop (device pointers) ---------------------------
 * Prepare a run on a device-resident u (d_u, [cells][V] doubles): builds the
 * mask, sizes the work buffers and computes the first dt on the device. */
int gpr_dev_begin(gpr_ctx *ctx, double *d_u, const int32_t *nX, const double *dX, double CFL,
                  const int32_t *bcs, double t0, double tf, int64_t count0, const gpr_par *MP);
/* enqueue `nsteps` full steps (boundaries, WENO, predictor, FV + post, dt) */
int gpr_dev_steps(gpr_ctx *ctx, int64_t nsteps);
/* read back t, count, the last dt and the status flags (synchronises) */
int gpr_dev_status(gpr_ctx *ctx, double *t, int64_t *count, double *dt, int32_t *pred_fail,
                   int32_t *nonfinite);
/* Per-phase timing of gpr_dev_steps with HIP events on the context's stream.
 * enable = 1 clears the totals and starts recording; 0 stops. */
int gpr_dev_set_timing(gpr_ctx *ctx, int enable);
/* Totals since timing was enabled (synchronises): ms[5] = boundaries, WENO,
 * predictor, FV (faces and cells, or faces and the cell update), post-step +
 * reset + dt;
 * *steps = steps timed; iters[3] = Picard iterations, fallback iterations and
 * fallback cells summed over the predicted cells; cells = predicted cells
 * summed over the steps. */
int gpr_dev_timing(gpr_ctx *ctx, double *ms, int64_t *steps, int64_t *iters, int64_t *cells);

/* --- batches of independent 1D problems ------------------------------------
 * No single reference counterpart: each instance is what iterator() does for
 * one 1D problem (iterator.cpp:85-170, one material, SPLIT = 0, FLUX = 0).
 * The context must be 1D.  Instance k has nx[k] cells of width dx[k],
 * boundary codes bcs[2k], bcs[2k+1], material MPs[k] and end time tf[k]; the
 * state of all instances is one array, the instances' [nx[k]][V] blocks in
 * order.  Every stage of a step is one launch over all instances, each with
 * its own dt (CFL, the 0.2 ramp of the first 6 steps, the clamp to its tf);
 * an instance stops at its tf or when its state becomes non-finite. */
int gpr_batch_create(gpr_ctx *ctx, int count, const int32_t *nx, const double *dx, const int32_t *bcs,
                     const gpr_par *MPs, const double *tf, double CFL);
int gpr_batch_destroy(gpr_ctx *ctx);
/* host state in (t = 0, step count 0, first dt on the device) / out */
int gpr_batch_set_state(gpr_ctx *ctx, const double *u);
int gpr_batch_get_state(gpr_ctx *ctx, double *u);
/* block size of the element-wise kernels (a multiple of 64, <= 1024) */
int gpr_batch_set_block(gpr_ctx *ctx, int block);
/* enqueue nsteps steps on the context's stream */
int gpr_batch_steps(gpr_ctx *ctx, int64_t nsteps);
/* capture steps_per_graph steps in a hipGraph (kept until the next change of
 * steps_per_graph or block) and replay it until every instance has stopped or
 * max_graphs replays ran; *graphs_run replays, *running instances left */
int gpr_batch_run(gpr_ctx *ctx, int steps_per_graph, int64_t max_graphs, int64_t *graphs_run, int32_t *running);
/* per-instance t, step count, next dt, stopped and non-finite flags (arrays of
 * count; any may be null) and the predictor's failed-cell count */
int gpr_batch_status(gpr_ctx *ctx, double *t, int64_t *count, double *dt, int32_t *done, int32_t *nonfinite,
                     int32_t *pred_fail);

/* --- adjoint of the device loop (2D, N = 3, V = 14 or 17; Newtonian fluids
 * and elastic or power-law solids) ---------------------------------------------
 * No reference counterpart.  After gpr_dev_begin: gpr_dev_steps_dt advances
 * with prescribed time steps (dts[k] for step k) instead of the CFL rule;
 * gpr_adj_forward advances nsteps steps as gpr_dev_steps does, keeping a
 * checkpoint of u every K steps and the time steps taken (dts_out, may be
 * null); gpr_adj_backward takes lam_T = d loss / d u at the end (host,
 * [cells][V]) and returns d loss / d u at the start (lam_0) and the
 * derivative with respect to the material's relaxation parameter in *grad_mu:
 * mu for a Newtonian fluid, sigmaY for a power-law solid (0 for an elastic
 * solid), with the time steps held at the forward run's values; *unconverged counts
 * the cells whose transposed predictor iteration did not converge.
 * gpr_adj_forward fails with GPR_E_UNSUPPORTED if the global distortion reset
 * fires during the sweep (it is not differentiated), and with
 * GPR_E_PREDICTOR / GPR_E_NONFINITE on a failed step.
 * gpr_adj_release frees the checkpoints. */
int gpr_dev_steps_dt(gpr_ctx *ctx, const double *dts, int64_t nsteps);
int gpr_adj_forward(gpr_ctx *ctx, int64_t nsteps, int K, double *dts_out);
int gpr_adj_backward(gpr_ctx *ctx, const double *lam_T, double *lam_0, double *grad_mu, int32_t *unconverged);
int gpr_adj_release(gpr_ctx *ctx);
/* The transposed predictor solve of later gpr_adj_backward calls: with
 * start_preconditioned = 0 each cell runs the transposed Picard iteration
 * and switches to the transposed preconditioned form if that does not
 * converge; with 1 every cell uses the preconditioned form from the start
 * (both have the same solution; the default is 1, which takes about a third
 * of the iterations).  tol: stopping tolerance on the change of the iterate,
 * relative to 1 + |iterate| (default 1e-13).  dual_local = 1 takes the local
 * derivatives of the node physics by forward-mode dual numbers, one seed per
 * input, instead of the hand-written reverse mode (default 0; several times
 * slower).  The two agree to round-off: within 1e-12 relative on fluids, 2e-11
 * on a power-law solid, where the reverse mode takes the wave speed's
 * derivative at the computed eigenvalue and the dual numbers carry it through
 * the eigenvalue iteration. */
int gpr_adj_set_solver(gpr_ctx *ctx, int start_preconditioned, double tol, int dual_local);
/* d loss / d kappa (the heat conductivity) of the last gpr_adj_backward; 0 for
 * a context without heat conduction */
int gpr_adj_grad_kappa(gpr_ctx *ctx, double *out);
/* out[3]: of the last gpr_adj_backward, transposed predictor iterations summed
 * over cells and steps, cell-steps, and cell-steps that used the
 * preconditioned form */
int gpr_adj_stats(gpr_ctx *ctx, int64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* GPR_MI355X_H */
