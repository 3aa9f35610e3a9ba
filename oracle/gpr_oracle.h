/* CPU restatement of the reference's ADER-WENO step for the GPR system.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker.
 * The product (phd_amd/, libgpr_mi355x.so) never links, loads or calls it.
 *
 * Plain C11 + OpenMP.  Runtime N (nodes per dimension, 2 or 3), ndim (1 or 2)
 * and THERMAL (V = 14 or 17).  Data layouts are the reference's:
 *   u   [cells][V]                      C-order over (x, y)
 *   ub  [(nx+2N)(ny+2N)][V]             ghost-extended by N per side
 *   wh  [(nx+2)(ny+2)][N^ndim][V]       node index a*N+b (a = x-node)
 *   qh  [(nx+2)(ny+2)][N^(ndim+1)][V]   node index (t*N+a)*N+b
 *   mask[(nx+2)(ny+2)] uint8
 * Each function cites the reference file:line it restates.
 */
#ifndef GPR_ORACLE_H
#define GPR_ORACLE_H

#include <stdint.h>

#include "../include/gpr_par.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
  int N;       /* Gauss-Legendre nodes per dimension (2 or 3) */
  int thermal; /* 0: V = 14, 1: V = 17 */
  int ndim;    /* 1 or 2 */
  int dg_it;   /* Picard iteration cap (reference DG_IT = 50) */
  double dg_tol; /* Picard tolerance (reference DG_TOL = 6e-6) */
  int stiff_it;  /* cap on the stiff fallback iterations */
  double stiff_tol; /* stopping tolerance of the stiff fallback */
  int bailout;   /* 0: Picard runs to dg_it as in the reference; k > 0: leave
                    Picard once its stopping statistic grew k times in a row */
  int fallback;  /* 0: preconditioned Picard (what the device code runs);
                    1: Newton on the local source with lagged transport (an
                       independent solver of the same equations, for tests) */
  int flux;      /* Riemann solver of the FV interfaces: 0 Rusanov, 1 Roe, 2 Osher */
  int bail_from; /* 0: off; j > 0: after j Picard iterations without convergence,
                    leave Picard the first time its stopping statistic grows */
} orc_cfg;

/* per-call predictor statistics */
typedef struct orc_stats {
  long cells;          /* masked cells solved */
  long picard_iters;   /* total Picard iterations over all cells */
  long hist[64];       /* histogram of Picard iterations (index = count, 63 = failed) */
  long fallback_cells; /* cells solved by the stiff fallback */
  long fallback_iters; /* total fallback iterations */
  long failed_cells;   /* cells neither path converged */
} orc_stats;

void orc_default_cfg(orc_cfg *c, int N, int thermal, int ndim);

/* constants */
void orc_constants(int N, double *nodes, double *wghts, double *endvals,
                   double *dervals, double *mLinv, double *mRinv,
                   double *mCLinv, double *sig, double *dg_end, double *dg_der,
                   double *dg_mat);

/* pointwise physics, V = 14 + 3*thermal */
double orc_pressure(const double *Q, const gpr_par *p, int thermal);
void orc_flux(double *ret, const double *Q, int d, const gpr_par *p, int thermal);
void orc_source(double *ret, const double *Q, const gpr_par *p, int thermal);
void orc_bdot(double *ret, const double *Q, const double *x, int d, int thermal);
int orc_thermo_acoustic_tensor(double *ret, const double *Q, int d,
                               const gpr_par *p, int thermal);
double orc_max_abs_eigs(const double *Q, int d, const gpr_par *p, int thermal);
void orc_d_rus(double *ret, const double *qL, const double *qR, int d,
               const gpr_par *p, const orc_cfg *c);
void orc_bint(double *ret, const double *qL, const double *qR, int d,
              const gpr_par *p, const orc_cfg *c);
/* system matrix dF_d/dQ + B_d, V x V row-major */
void orc_system_matrix(double *ret, const double *Q, int d, const gpr_par *p, int thermal);
/* Roe and Osher fluxes; 0, or -1 if an eigendecomposition did not converge */
int orc_d_roe(double *ret, const double *qL, const double *qR, int d,
              const gpr_par *p, const orc_cfg *c);
int orc_d_osh(double *ret, const double *qL, const double *qR, int d,
              const gpr_par *p, const orc_cfg *c);
/* general real eigenvalues of an n x n row-major matrix (n <= 4) */
int orc_eigenvalues(const double *a, int n, double *wr, double *wi);

/* grid */
void orc_boundaries(const double *u, double *ub, const int32_t *nX,
                    const int32_t *bcs, const orc_cfg *c);
void orc_extend_mask(const uint8_t *mask, uint8_t *maskb, const int32_t *nX,
                     const orc_cfg *c);

/* stages */
void orc_weno_launcher(double *wh, const double *ub, const int32_t *nX,
                       const orc_cfg *c);
int orc_predictor(double *qh, const double *wh, double dt, const double *dX,
                  const int32_t *nX, const gpr_par *p, const uint8_t *mask,
                  const orc_cfg *c, orc_stats *st, int32_t *iters);
/* DG residual of one cell (the reference's obj1/obj2, dg1.cpp:54-66,
 * dg2.cpp:73-90): out = rhs(q) - (DG_MAT (x) W (x) W) q */
void orc_dg_residual(double *out, const double *q, const double *w, double dt,
                     const double *dX, const gpr_par *p, const orc_cfg *c);
void orc_fv_launcher(double *u, const double *qh, const int32_t *nX,
                     double dt, const double *dX, const gpr_par *p,
                     const uint8_t *mask, const orc_cfg *c);
double orc_timestep(const double *u, const int32_t *nX, const double *dX,
                    double CFL, double t, double tf, long count,
                    const gpr_par *p, const orc_cfg *c);
void orc_renorm_distortion(double *u, long ncell, const gpr_par *p, int V);
int orc_contorted(const double *u, long ncell, double tol, int V);
void orc_reset_distortion(double *u, long ncell, const gpr_par *p, int V);

/* One full step (iterator.cpp:110-165 body): dt, BCs, WENO, predictor, FV,
 * renorm, contorted -> reset.  Returns the status of the predictor (0 ok,
 * nonzero: some cell failed both solvers) or -1 on a non-finite u. */
int orc_step(double *u, const int32_t *nX, const double *dX, double CFL,
             const int32_t *bcs, double *t, long *count, double tf,
             double contorted_tol, const gpr_par *p, const orc_cfg *c,
             double *dt_out, orc_stats *st);

/* centers (parts & 1) and interfaces (parts & 2) of fv_launcher */
void orc_fv_parts(double *u, const double *qh, const int32_t *nX, double dt,
                  const double *dX, const gpr_par *p, const uint8_t *mask,
                  int sources, int time, int parts, const orc_cfg *c);
/* fv_launcher with its SOURCES and TIME flags (TIME = 0: one time level per cell) */
void orc_fv_launcher_ex(double *u, const double *qh, const int32_t *nX, double dt,
                        const double *dX, const gpr_par *p, const uint8_t *mask,
                        int sources, int time, const orc_cfg *c);

/* split-WENO step (steppers.cpp:50-94, solvers/split, relaxation/analytic.cpp) */
void orc_ode_launcher(double *u, long ncell, double dt, const gpr_par *p, const orc_cfg *c);
void orc_analytic_solver(double *Q, long count, double dt, const gpr_par *p, int part,
                         const orc_cfg *c);
void orc_midstepper(double *wh, long ncell, double dt, const double *dX, const gpr_par *p,
                    const uint8_t *mask, const orc_cfg *c);
/* One split step of the time loop; half_step runs the midstepper.  Returns 0,
 * or -1 on a non-finite u. */
int orc_split_step(double *u, const int32_t *nX, const double *dX, double CFL,
                   const int32_t *bcs, double *t, long *count, double tf,
                   double contorted_tol, int half_step, const gpr_par *p, const orc_cfg *c,
                   double *dt_out);

#ifdef __cplusplus
}
#endif

#endif
