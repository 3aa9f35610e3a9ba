/* CPU restatement of the reference's ADER-WENO step for the GPR system.
 *
 * TEST INFRASTRUCTURE ONLY (see gpr_oracle.h): the checker for the HIP
 * implementation, never part of the product.
 *
 * Follows, function by function, code/cpp/src of haranjackson/phd:
 *   constants   solvers/poly/basis.cpp:7-72, solvers/weno/weno_matrices.cpp:9-42,
 *               solvers/dg/dg_matrices.cpp:30-53, etc/globals.cpp:8-50
 *   physics     system/equations.cpp:11-167, system/eig.cpp:8-77,
 *               system/variables/{state,sources}.cpp, system/energy/{mg,eos,derivatives}.cpp,
 *               system/waves/{shear,speeds}.cpp, system/functions/matrices.cpp
 *   grid        etc/grid.cpp:12-230
 *   WENO        solvers/weno/weno.cpp:4-157
 *   predictor   solvers/dg/dg1.cpp, dg2.cpp (Picard, stopping rule dg2.cpp:273-298)
 *   FV          solvers/fv/fv.cpp:11-297, fluxes.cpp:9-20,66-75
 *   step        solvers/iterator.cpp:54-83,110-165, solvers/utils.cpp:6-85
 *
 * One deliberate difference, in the predictor: a cell whose Picard iteration
 * does not meet the tolerance within the iteration cap, or ends non-finite,
 * is re-solved by a locally implicit iteration (transport lagged, stiff
 * source implicit by Newton per spatial node) for the same DG equations.  The
 * reference leaves such a cell unassigned (dg2.cpp:273-298) or, on NaN, calls
 * a Newton-Krylov solver (newton_krylov.cpp:143-182); both solve the same
 * equations, so the converged predictor is the same up to the tolerance.
 */
#include "gpr_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 3
#define MAXV 17

static const double LAMS = 1.0, LAMC = 1e5, WENO_EPS = 1e-14;

void orc_default_cfg(orc_cfg *c, int N, int thermal, int ndim) {
  c->N = N;
  c->thermal = thermal;
  c->ndim = ndim;
  c->dg_it = 50;
  c->dg_tol = 6e-6;
  c->stiff_it = 200;
  c->stiff_tol = 6e-6;
  c->bailout = 0;
  c->fallback = 0;
  c->flux = 0;
  c->bail_from = 0;
}

/* ------------------------------------------------------------------ constants
 * Polynomials as ascending coefficient arrays of length <= 8. */
typedef struct {
  int n;
  double c[8];
} poly;

static double pev(const poly *p, double x) {
  double r = 0;
  for (int i = p->n - 1; i >= 0; i--)
    r = r * x + p->c[i];
  return r;
}
static poly pdiff(const poly *p) {
  poly r = {0, {0}};
  r.n = p->n > 1 ? p->n - 1 : 1;
  for (int i = 1; i < p->n; i++)
    r.c[i - 1] = i * p->c[i];
  return r;
}
static poly pint(const poly *p) {
  poly r = {p->n + 1, {0}};
  for (int i = 0; i < p->n; i++)
    r.c[i + 1] = p->c[i] / (i + 1);
  return r;
}
static poly pmul(const poly *a, const poly *b) {
  poly r = {a->n + b->n - 1, {0}};
  for (int i = 0; i < a->n; i++)
    for (int j = 0; j < b->n; j++)
      r.c[i + j] += a->c[i] * b->c[j];
  return r;
}

/* Gauss-Legendre nodes/weights on [0,1] (scaled_nodes/weights, basis.cpp:7-20) */
static void gl01(int N, double *x, double *w) {
  if (N == 2) {
    double a = 1.0 / sqrt(3.0);
    x[0] = (1 - a) / 2;
    x[1] = (1 + a) / 2;
    w[0] = w[1] = 0.5;
  } else if (N == 3) {
    double a = sqrt(0.6);
    x[0] = (1 - a) / 2;
    x[1] = 0.5;
    x[2] = (1 + a) / 2;
    w[0] = w[2] = 5.0 / 18.0;
    w[1] = 4.0 / 9.0;
  } else {
    x[0] = 0.5;
    w[0] = 1.0;
  }
}

/* Lagrange basis at the nodes (lagrange, basis.cpp:22-53) */
static void basis_polys(int N, const double *x, poly *psi) {
  for (int i = 0; i < N; i++) {
    poly p = {1, {1.0}};
    for (int j = 0; j < N; j++) {
      if (j == i)
        continue;
      poly f = {2, {-x[j] / (x[i] - x[j]), 1.0 / (x[i] - x[j])}};
      p = pmul(&p, &f);
    }
    psi[i] = p;
  }
}

static int inv_small(const double *a, double *ainv, int n) {
  double m[4][8];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++)
      m[i][j] = j < n ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(m[r][c]) > fabs(m[piv][c]))
        piv = r;
    if (m[piv][c] == 0.0)
      return -1;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) {
        double t = m[c][j];
        m[c][j] = m[piv][j];
        m[piv][j] = t;
      }
    double d = m[c][c];
    for (int j = 0; j < 2 * n; j++)
      m[c][j] /= d;
    for (int r = 0; r < n; r++) {
      if (r == c)
        continue;
      double f = m[r][c];
      for (int j = 0; j < 2 * n; j++)
        m[r][j] -= f * m[c][j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      ainv[i * n + j] = m[i][n + j];
  return 0;
}

typedef struct {
  int N;
  double x[MAXN], w[MAXN];
  double endv[2][MAXN];
  double der[MAXN][MAXN]; /* DERVALS(i,j) = psi_j'(x_i) */
  double mLinv[MAXN][MAXN], mRinv[MAXN][MAXN], mCLinv[MAXN][MAXN];
  double sig[MAXN][MAXN];
  double dgder[MAXN][MAXN]; /* DG_DER */
  double dgmat[MAXN][MAXN];
  double dgmatinv[MAXN][MAXN];
} consts_t;

static void build_consts(int N, consts_t *K) {
  memset(K, 0, sizeof(*K));
  K->N = N;
  gl01(N, K->x, K->w);
  poly psi[MAXN];
  basis_polys(N, K->x, psi);
  int FN2 = (N - 1) / 2, CN2 = N / 2; /* floor((N-1)/2), ceil((N-1)/2) */
  double mL[9], mR[9], mCL[9];
  for (int j = 0; j < N; j++) {
    poly dj = pdiff(&psi[j]);
    K->endv[0][j] = pev(&psi[j], 0.0);
    K->endv[1][j] = pev(&psi[j], 1.0);
    for (int i = 0; i < N; i++)
      K->der[i][j] = pev(&dj, K->x[i]);
  }
  /* coefficient_matrices, weno_matrices.cpp:9-29 */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      poly P = pint(&psi[j]);
      mL[i * N + j] = pev(&P, i - N + 2) - pev(&P, i - N + 1);
      mR[i * N + j] = pev(&P, i + 1) - pev(&P, i);
      mCL[i * N + j] = pev(&P, i - CN2 + 1) - pev(&P, i - CN2);
    }
  (void)FN2;
  double t[9];
  inv_small(mL, t, N);
  for (int i = 0; i < N * N; i++)
    K->mLinv[i / N][i % N] = t[i];
  inv_small(mR, t, N);
  for (int i = 0; i < N * N; i++)
    K->mRinv[i / N][i % N] = t[i];
  inv_small(mCL, t, N);
  for (int i = 0; i < N * N; i++)
    K->mCLinv[i / N][i % N] = t[i];
  /* oscillation_indicator, weno_matrices.cpp:31-42 (zero-initialised) */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      double s = 0;
      poly pi = psi[i], pj = psi[j];
      for (int a = 1; a < N; a++) {
        pi = pdiff(&pi);
        pj = pdiff(&pj);
        poly pr = pmul(&pi, &pj);
        poly P = pint(&pr);
        s += pev(&P, 1.0) - pev(&P, 0.0);
      }
      K->sig[i][j] = s;
    }
  /* end_value_products, derivative_products, DG_MAT (dg_matrices.cpp:30-53,
   * globals.cpp:40-42) */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      if (i == j)
        K->dgder[i][j] = (K->endv[1][i] * K->endv[1][i] -
                          K->endv[0][i] * K->endv[0][i]) / 2;
      else
        K->dgder[i][j] = K->w[i] * K->der[i][j];
    }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      K->dgmat[i][j] = K->endv[1][i] * K->endv[1][j] - K->dgder[j][i];
  double dm[9];
  for (int i = 0; i < N * N; i++)
    dm[i] = K->dgmat[i / N][i % N];
  inv_small(dm, t, N);
  for (int i = 0; i < N * N; i++)
    K->dgmatinv[i / N][i % N] = t[i];
}

static consts_t k2_, k3_;

__attribute__((constructor)) static void init_consts(void) {
  build_consts(2, &k2_);
  build_consts(3, &k3_);
}

static const consts_t *get_consts(int N) { return N == 2 ? &k2_ : &k3_; }

void orc_constants(int N, double *nodes, double *wghts, double *endvals,
                   double *dervals, double *mLinv, double *mRinv,
                   double *mCLinv, double *sig, double *dg_end, double *dg_der,
                   double *dg_mat) {
  const consts_t *K = get_consts(N);
  for (int i = 0; i < N; i++) {
    nodes[i] = K->x[i];
    wghts[i] = K->w[i];
    endvals[i] = K->endv[0][i];
    endvals[N + i] = K->endv[1][i];
    for (int j = 0; j < N; j++) {
      dervals[i * N + j] = K->der[i][j];
      mLinv[i * N + j] = K->mLinv[i][j];
      mRinv[i * N + j] = K->mRinv[i][j];
      mCLinv[i * N + j] = K->mCLinv[i][j];
      sig[i * N + j] = K->sig[i][j];
      dg_end[i * N + j] = K->endv[1][i] * K->endv[1][j];
      dg_der[i * N + j] = K->dgder[i][j];
      dg_mat[i * N + j] = K->dgmat[i][j];
    }
  }
}

/* ------------------------------------------------------------------- EOS
 * system/energy/mg.cpp:6-303 */
static double Gamma_MG(double r, const gpr_par *p) {
  switch (p->EOS) {
  case GPR_EOS_STIFFENED_GAS:
    return p->gamma - 1;
  case GPR_EOS_SHOCK_MG:
    return p->Gamma0 * p->rho0 / r;
  case GPR_EOS_GODUNOV_ROMENSKI:
    return p->gamma;
  case GPR_EOS_JWL:
  case GPR_EOS_COCHRAN_CHAN:
    return p->Gamma0;
  }
  return NAN;
}
static double dGamma_MG(double r, const gpr_par *p) {
  if (p->EOS == GPR_EOS_SHOCK_MG)
    return -p->Gamma0 * p->rho0 / (r * r);
  return 0.0;
}
static double p_ref(double r, const gpr_par *p) {
  double r0 = p->rho0;
  switch (p->EOS) {
  case GPR_EOS_STIFFENED_GAS:
    return -p->pINF;
  case GPR_EOS_SHOCK_MG:
    if (r > r0) {
      double t = 1 / r0 - p->s * (1 / r0 - 1 / r);
      return p->c02 * (1 / r0 - 1 / r) / (t * t);
    }
    return p->c02 * (r - r0);
  case GPR_EOS_GODUNOV_ROMENSKI: {
    double t = pow(r / r0, p->alpha);
    return p->c02 * r / p->alpha * (t - 1) * t;
  }
  case GPR_EOS_JWL: {
    double v = r0 / r;
    return p->A * exp(-p->R1 * v) + p->B * exp(-p->R2 * v);
  }
  case GPR_EOS_COCHRAN_CHAN: {
    double v = r0 / r;
    return p->A * pow(v, -p->R1) - p->B * pow(v, -p->R2);
  }
  }
  return NAN;
}
static double e_ref(double r, const gpr_par *p) {
  double r0 = p->rho0;
  switch (p->EOS) {
  case GPR_EOS_STIFFENED_GAS:
    return p->pINF / r;
  case GPR_EOS_SHOCK_MG:
    if (r > r0)
      return 0.5 * p_ref(r, p) * (1 / r0 - 1 / r);
    return 0.0;
  case GPR_EOS_GODUNOV_ROMENSKI: {
    double t = pow(r / r0, p->alpha);
    return p->c02 / (2 * p->alpha * p->alpha) * (t - 1) * (t - 1);
  }
  case GPR_EOS_JWL: {
    double v = r0 / r;
    return p->A / (r0 * p->R1) * exp(-p->R1 * v) +
           p->B / (r0 * p->R2) * exp(-p->R2 * v);
  }
  case GPR_EOS_COCHRAN_CHAN: {
    double v = r0 / r;
    return -p->A / (r0 * (1 - p->R1)) * (pow(v, 1 - p->R1) - 1) +
           p->B / (r0 * (1 - p->R2)) * (pow(v, 1 - p->R2) - 1);
  }
  }
  return NAN;
}
static double dp_ref(double r, const gpr_par *p) {
  double r0 = p->rho0;
  switch (p->EOS) {
  case GPR_EOS_STIFFENED_GAS:
    return 0.0;
  case GPR_EOS_SHOCK_MG:
    if (r > r0) {
      double t = p->s * (r - r0) - r;
      return p->c02 * r0 * r0 * (p->s * (r0 - r) - r) / (t * t * t);
    }
    return p->c02;
  case GPR_EOS_GODUNOV_ROMENSKI: {
    double t = pow(r / r0, p->alpha);
    return p->c02 / p->alpha * t * ((1 + p->alpha) * (t - 1) + p->alpha * t);
  }
  case GPR_EOS_JWL: {
    double v = r0 / r;
    return v / r * (p->A * p->R1 * exp(-p->R1 * v) + p->B * p->R2 * exp(-p->R2 * v));
  }
  case GPR_EOS_COCHRAN_CHAN: {
    double v = r0 / r;
    return p->A * p->R1 / r * pow(v, -p->R1) - p->B * p->R2 / r * pow(v, -p->R2);
  }
  }
  return NAN;
}
static double de_ref(double r, const gpr_par *p) {
  if (p->EOS == GPR_EOS_SHOCK_MG) {
    double r0 = p->rho0;
    if (r > r0) {
      double t = p->s * (r - r0) - r;
      return -(r - r0) * r0 * p->c02 / (t * t * t);
    }
    return 0.0;
  }
  return p_ref(r, p) / (r * r);
}
static double phi_int(double r, const gpr_par *p) {
  double r0 = p->rho0;
  switch (p->EOS) {
  case GPR_EOS_STIFFENED_GAS:
    return pow(r / r0, p->gamma - 1);
  case GPR_EOS_SHOCK_MG:
    return exp(p->Gamma0 * (1 - r0 / r));
  case GPR_EOS_GODUNOV_ROMENSKI:
    return pow(r / r0, p->gamma);
  case GPR_EOS_JWL:
  case GPR_EOS_COCHRAN_CHAN:
    return pow(r / r0, p->Gamma0);
  }
  return NAN;
}
static double temperature_mg(double r, double pr_, const gpr_par *p) {
  double G = Gamma_MG(r, p);
  return phi_int(r, p) * p->Tref + (pr_ - p_ref(r, p)) / (r * G * p->cv);
}
static double dedrho(double r, double pp, const gpr_par *p) {
  double G = Gamma_MG(r, p), dG = dGamma_MG(r, p), pr = p_ref(r, p),
         dpr = dp_ref(r, p), der = de_ref(r, p);
  return der - (dpr * r * G + (G + r * dG) * (pp - pr)) / ((r * G) * (r * G));
}
static double dedp(double r, const gpr_par *p) { return 1 / (r * Gamma_MG(r, p)); }
static double dTdrho(double r, double pp, const gpr_par *p) {
  double G = Gamma_MG(r, p), dG = dGamma_MG(r, p), pr = p_ref(r, p),
         dpr = dp_ref(r, p);
  return -(dpr * r * G + (G + r * dG) * (pp - pr)) / ((r * G) * (r * G)) / p->cv;
}
static double dTdp(double r, const gpr_par *p) { return dedp(r, p) / p->cv; }

/* waves/shear.cpp:6-41 */
static double c_s2(double r, const gpr_par *p) { return p->b02 * pow(r / p->rho0, p->beta); }
static double dc_s2drho(double r, const gpr_par *p) { return p->beta / r * c_s2(r, p); }
static double C_0(double r, const gpr_par *p) {
  return (1 - p->beta / Gamma_MG(r, p)) * c_s2(r, p);
}
static double dC_0drho(double r, const gpr_par *p) {
  double cs2 = c_s2(r, p), dcs2 = dc_s2drho(r, p), G = Gamma_MG(r, p),
         dG = dGamma_MG(r, p);
  return (1 - p->beta / G) * dcs2 + p->beta / (G * G) * dG * cs2;
}

/* functions/matrices.cpp:8-41 ; A row-major 3x3 */
static void AdevG(const double *A, double *out) {
  double G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = A[0 * 3 + i] * A[0 * 3 + j] + A[1 * 3 + i] * A[1 * 3 + j] +
                     A[2 * 3 + i] * A[2 * 3 + j];
  double x = (G[0] + G[4] + G[8]) / 3;
  G[0] -= x;
  G[4] -= x;
  G[8] -= x;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = A[i * 3 + 0] * G[0 * 3 + j] + A[i * 3 + 1] * G[1 * 3 + j] +
                       A[i * 3 + 2] * G[2 * 3 + j];
}
static double devGsq(const double *A) {
  double G00 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  double G11 = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  double G22 = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  double G01 = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  double G02 = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  double G12 = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  double t = (G00 + G11 + G22) / 3;
  return (G00 - t) * (G00 - t) + (G11 - t) * (G11 - t) + (G22 - t) * (G22 - t) +
         2 * (G01 * G01 + G02 * G02 + G12 * G12);
}
static double det3(const double *A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}
static double sigma_norm(const double *s) {
  double a = s[0] - s[4], b = s[4] - s[8], c = s[8] - s[0];
  double t1 = a * a + b * b + c * c;
  double t2 = s[1] * s[1] + s[5] * s[5] + s[6] * s[6];
  return sqrt(0.5 * t1 + 3 * t2);
}

/* energy/eos.cpp:21-25,72-94 */
static double E_2A(double r, const double *A, const gpr_par *p) {
  return C_0(r, p) / 4 * devGsq(A);
}
static double internal_energy(const double *Q, const gpr_par *p, int thermal) {
  double r = Q[0];
  double E = Q[1] / r;
  double v0 = Q[2] / r, v1 = Q[3] / r, v2 = Q[4] / r;
  double E1 = E - (v0 * v0 + v1 * v1 + v2 * v2) / 2;
  E1 -= E_2A(r, Q + 5, p);
  if (thermal) {
    double J0 = Q[14] / r, J1 = Q[15] / r, J2 = Q[16] / r;
    E1 -= p->calpha2 * (J0 * J0 + J1 * J1 + J2 * J2) / 2;
  }
  return E1;
}
/* variables/state.cpp:11-21 */
double orc_pressure(const double *Q, const gpr_par *p, int thermal) {
  double r = Q[0];
  double e = internal_energy(Q, p, thermal);
  return (e - e_ref(r, p)) * r * Gamma_MG(r, p) + p_ref(r, p);
}
/* dEdA_s = cs2 * AdevG(A) (derivatives.cpp:169-175) */
static void dEdA_s(const double *Q, const gpr_par *p, double *out) {
  double cs2 = c_s2(Q[0], p);
  AdevG(Q + 5, out);
  for (int i = 0; i < 9; i++)
    out[i] *= cs2;
}
/* sigma (state.cpp:23-37): -rho E_A^T A */
static void sigma_full(const double *Q, const gpr_par *p, double *s) {
  double EA[9];
  const double *A = Q + 5;
  dEdA_s(Q, p, EA);
  double r = Q[0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      s[i * 3 + j] = -r * (EA[0 * 3 + i] * A[0 * 3 + j] + EA[1 * 3 + i] * A[1 * 3 + j] +
                           EA[2 * 3 + i] * A[2 * 3 + j]);
}
static void sigma_col(const double *Q, const gpr_par *p, int d, double *s) {
  double EA[9];
  const double *A = Q + 5;
  dEdA_s(Q, p, EA);
  double r = Q[0];
  for (int i = 0; i < 3; i++)
    s[i] = -r * (EA[0 * 3 + i] * A[0 * 3 + d] + EA[1 * 3 + i] * A[1 * 3 + d] +
                 EA[2 * 3 + i] * A[2 * 3 + d]);
}

/* equations.cpp:11-48 (accumulates into ret) */
void orc_flux(double *ret, const double *Q, int d, const gpr_par *p, int thermal) {
  double r = Q[0], rE = Q[1];
  const double *rv = Q + 2;
  double pp = orc_pressure(Q, p, thermal);
  double vd = rv[d] / r;
  ret[0] += rv[d];
  ret[1] += vd * (rE + pp);
  ret[2] += vd * rv[0];
  ret[3] += vd * rv[1];
  ret[4] += vd * rv[2];
  ret[2 + d] += pp;
  {
    const double *A = Q + 5;
    double sd[3];
    sigma_col(Q, p, d, sd);
    ret[1] -= (sd[0] * rv[0] + sd[1] * rv[1] + sd[2] * rv[2]) / r;
    ret[2] -= sd[0];
    ret[3] -= sd[1];
    ret[4] -= sd[2];
    for (int k = 0; k < 3; k++) {
      double Av = (A[k * 3 + 0] * rv[0] + A[k * 3 + 1] * rv[1] + A[k * 3 + 2] * rv[2]) / r;
      ret[5 + d + 3 * k] += Av;
    }
  }
  if (thermal) {
    const double *rJ = Q + 14;
    double T = temperature_mg(r, pp, p);
    ret[1] += p->calpha2 * T * rJ[d] / r;
    ret[14] += vd * rJ[0];
    ret[15] += vd * rJ[1];
    ret[16] += vd * rJ[2];
    ret[14 + d] += T;
  }
}

/* variables/sources.cpp:13-91 */
static double theta1inv(const double *Q, const gpr_par *p) {
  double r = Q[0];
  double A53 = pow(det3(Q + 5), 5.0 / 3.0);
  double cs2 = c_s2(r, p);
  double n = p->n;
  if (p->SOLID) {
    if (p->POWER_LAW) {
      double s[9];
      sigma_full(Q, p, s);
      double sn = sigma_norm(s);
      sn = fmin(sn, 1e8);
      return 3 * A53 / (cs2 * p->tau0) * pow(sn / p->sigmaY, n);
    }
    return 0.0;
  } else if (p->BINGHAM) {
    double s[9], f = 0;
    sigma_full(Q, p, s);
    for (int i = 0; i < 9; i++)
      f += s[i] * s[i];
    double sn = sqrt(f) / sqrt(2.0);
    if (sn > p->sigmaY)
      return A53 * p->rho0 / (2 * p->mu) * (1. - p->sigmaY / sn);
    return 0.0;
  } else {
    double c = A53 * p->rho0 / (2 * pow(p->mu, 1 / n));
    if (p->POWER_LAW) {
      double s[9], f = 0;
      sigma_full(Q, p, s);
      for (int i = 0; i < 9; i++)
        f += s[i] * s[i];
      double sn = sqrt(f) / sqrt(2.0);
      return c * pow(sn, (1 - n) / n);
    }
    return c;
  }
}

/* equations.cpp:50-67 */
void orc_source(double *ret, const double *Q, const gpr_par *p, int thermal) {
  int V = thermal ? 17 : 14;
  for (int i = 0; i < V; i++)
    ret[i] = 0.0;
  ret[2] = p->dp[0];
  ret[3] = p->dp[1];
  ret[4] = p->dp[2];
  double EA[9];
  dEdA_s(Q, p, EA);
  double th = theta1inv(Q, p);
  for (int i = 0; i < 9; i++)
    ret[5 + i] = -EA[i] * th;
  if (thermal) {
    double r = Q[0];
    double T = temperature_mg(r, orc_pressure(Q, p, thermal), p);
    double th2 = T / (p->kappa * r);
    for (int i = 0; i < 3; i++)
      ret[14 + i] = -r * (p->calpha2 * Q[14 + i] / r) * th2;
  }
}

/* equations.cpp:189-267 */
void orc_bdot(double *ret, const double *Q, const double *x, int d, int thermal) {
  int V = thermal ? 17 : 14;
  double r = Q[0];
  double v0 = Q[2] / r, v1 = Q[3] / r, v2 = Q[4] / r;
  for (int i = 0; i < V; i++)
    ret[i] = 0.0;
  for (int k = 0; k < 3; k++) {
    const double *xk = x + 5 + 3 * k;
    double *rk = ret + 5 + 3 * k;
    if (d == 0) {
      rk[0] = -v1 * xk[1] - v2 * xk[2];
      rk[1] = v0 * xk[1];
      rk[2] = v0 * xk[2];
    } else if (d == 1) {
      rk[0] = v1 * xk[0];
      rk[1] = -v0 * xk[0] - v2 * xk[2];
      rk[2] = v1 * xk[2];
    } else {
      rk[0] = v2 * xk[0];
      rk[1] = v2 * xk[1];
      rk[2] = -v0 * xk[0] - v1 * xk[1];
    }
  }
}

/* eig.cpp:8-64 */
int orc_thermo_acoustic_tensor(double *ret, const double *Q, int d,
                               const gpr_par *p, int thermal) {
  double r = Q[0];
  double pp = orc_pressure(Q, p, thermal);
  const double *A = Q + 5;
  double cs2 = c_s2(r, p), dcs2 = dc_s2drho(r, p);
  double sd[3];
  sigma_col(Q, p, d, sd);
  double dsdr[3];
  for (int i = 0; i < 3; i++)
    dsdr[i] = (1 / r + dcs2 / cs2) * sd[i];
  /* dsigmadA (state.cpp:56-69) */
  double G[9], AG[9], M[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  AdevG(A, AG);
  for (int i = 0; i < 3; i++)
    AG[i * 3 + d] *= 2.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      AG[i * 3 + j] += 1.0 / 3.0 * A[i * 3 + d] * G[d * 3 + j];
  for (int i = 0; i < 9; i++)
    AG[i] += G[d * 3 + d] * A[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      M[i * 3 + j] = -r * cs2 * AG[j * 3 + i];
  int n = thermal ? 4 : 3;
  double X1[4][5], X2[5][4];
  memset(X1, 0, sizeof X1);
  memset(X2, 0, sizeof X2);
  X1[0][1] = 1 / r;
  for (int i = 0; i < 3; i++) {
    X1[i][0] = -1 / r * dsdr[i];
    for (int j = 0; j < 3; j++)
      X1[i][2 + j] = -1 / r * M[i * 3 + j];
  }
  /* c_0 (speeds.cpp:8-18) */
  double dEdr = dedrho(r, pp, p) + dC_0drho(r, p) / 4 * devGsq(A);
  double c0 = sqrt((pp / (r * r) - dEdr) / dedp(r, p));
  X2[0][0] = r;
  X2[1][d] = r * c0 * c0;
  for (int j = 0; j < 3; j++)
    X2[1][j] += sd[j] - r * dsdr[j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      X2[2 + i][j] = A[i * 3 + j];
  if (thermal) {
    X1[3][0] = dTdrho(r, pp, p) / r;
    X1[3][1] = dTdp(r, p) / r;
    double T = temperature_mg(r, pp, p);
    double ch = sqrt(p->calpha2 * T / p->cv) / r;
    X2[1][3] = r * ch * ch / dTdp(r, p);
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double s = 0;
      for (int k = 0; k < 5; k++)
        s += X1[i][k] * X2[k][j];
      ret[i * n + j] = s;
    }
  return n;
}

/* Eigenvalues of a general real n x n matrix (n <= 4): complex shifted QR
 * with Wilkinson shifts and deflation; backward stable like the real Schur
 * iteration behind Eigen's EigenSolver used at eig.cpp:72. */
int orc_eigenvalues(const double *a, int n, double *wr, double *wi) {
  double complex H[4][4];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      H[i][j] = a[i * n + j];
  int m = n;
  int its = 0;
  while (m > 0) {
    if (m == 1) {
      wr[0] = creal(H[0][0]);
      wi[0] = cimag(H[0][0]);
      m = 0;
      break;
    }
    double s = cabs(H[m - 1][m - 1]) + cabs(H[m - 2][m - 2]);
    double fro = 0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++)
        fro += cabs(H[i][j]);
    if (s == 0)
      s = fro;
    /* deflate when the whole last row left of the diagonal is negligible */
    double off = 0;
    for (int j = 0; j < m - 1; j++)
      off += cabs(H[m - 1][j]);
    if (off <= 1e-17 * s || off == 0.0) {
      wr[m - 1] = creal(H[m - 1][m - 1]);
      wi[m - 1] = cimag(H[m - 1][m - 1]);
      m--;
      its = 0;
      continue;
    }
    if (++its > 200)
      return -1;
    /* Wilkinson shift from the trailing 2x2 */
    double complex a11 = H[m - 2][m - 2], a12 = H[m - 2][m - 1],
                   a21 = H[m - 1][m - 2], a22 = H[m - 1][m - 1];
    double complex tr = a11 + a22, det = a11 * a22 - a12 * a21;
    double complex disc = csqrt(tr * tr / 4 - det);
    double complex l1 = tr / 2 + disc, l2 = tr / 2 - disc;
    double complex mu = cabs(l1 - a22) < cabs(l2 - a22) ? l1 : l2;
    if (its % 11 == 10)
      mu += 0.7 * s * I; /* exceptional shift */
    for (int i = 0; i < m; i++)
      H[i][i] -= mu;
    /* QR by Givens rotations on the leading m x m block: H = Q R */
    double complex cs[16], sn[16];
    int nr = 0;
    for (int j = 0; j < m - 1; j++)
      for (int i = j + 1; i < m; i++) {
        double complex x = H[j][j], y = H[i][j];
        double r = sqrt(creal(x * conj(x)) + creal(y * conj(y)));
        double complex c, sv;
        if (r == 0) {
          c = 1;
          sv = 0;
        } else {
          c = x / r;
          sv = y / r;
        }
        /* rows j, i <- [conj(c) conj(s); -s c] [row j; row i] */
        for (int k = 0; k < m; k++) {
          double complex hj = H[j][k], hi = H[i][k];
          H[j][k] = conj(c) * hj + conj(sv) * hi;
          H[i][k] = -sv * hj + c * hi;
        }
        cs[nr] = c;
        sn[nr] = sv;
        nr++;
      }
    /* H = R Q : apply the adjoint rotations from the right in order */
    nr = 0;
    for (int j = 0; j < m - 1; j++)
      for (int i = j + 1; i < m; i++) {
        double complex c = cs[nr], sv = sn[nr];
        nr++;
        for (int k = 0; k < m; k++) {
          double complex hj = H[k][j], hi = H[k][i];
          H[k][j] = hj * c + hi * sv;
          H[k][i] = -hj * conj(sv) + hi * conj(c);
        }
      }
    for (int i = 0; i < m; i++)
      H[i][i] += mu;
  }
  return 0;
}

/* eig.cpp:66-77 */
double orc_max_abs_eigs(const double *Q, int d, const gpr_par *p, int thermal) {
  double X[16], wr[4], wi[4];
  int n = orc_thermo_acoustic_tensor(X, Q, d, p, thermal);
  if (orc_eigenvalues(X, n, wr, wi) != 0)
    return NAN;
  double mx = 0;
  for (int i = 0; i < n; i++) {
    double a = hypot(wr[i], wi[i]);
    if (a > mx || isnan(a))
      mx = a;
  }
  double lam = sqrt(mx);
  double vd = Q[2 + d] / Q[0];
  return vd > 0 ? vd + lam : lam - vd;
}

/* fluxes.cpp:66-75 */
void orc_d_rus(double *ret, const double *qL, const double *qR, int d,
               const gpr_par *p, const orc_cfg *c) {
  int th = c->thermal, V = th ? 17 : 14;
  double m1 = orc_max_abs_eigs(qL, d, p, th), m2 = orc_max_abs_eigs(qR, d, p, th);
  double m = (m1 < m2) ? m2 : m1; /* std::max: a NaN on the right is dropped */
  for (int i = 0; i < V; i++)
    ret[i] = m * (qL[i] - qR[i]);
  orc_flux(ret, qL, d, p, th);
  orc_flux(ret, qR, d, p, th);
}

/* fluxes.cpp:9-20 */
void orc_bint(double *ret, const double *qL, const double *qR, int d,
              const gpr_par *p, const orc_cfg *c) {
  (void)p;
  const consts_t *K = get_consts(c->N);
  int th = c->thermal, V = th ? 17 : 14;
  double dq[MAXV], q[MAXV], tmp[MAXV];
  for (int i = 0; i < V; i++) {
    dq[i] = qR[i] - qL[i];
    ret[i] = 0.0;
  }
  for (int s = 0; s < c->N; s++) {
    for (int i = 0; i < V; i++)
      q[i] = qL[i] + K->x[s] * dq[i];
    orc_bdot(tmp, q, dq, d, th);
    for (int i = 0; i < V; i++)
      ret[i] += K->w[s] * tmp[i];
  }
}

/* ---------------------------------------------------------- Roe and Osher
 * The system matrix M_d = dF_d/dP . dP/dQ + B_d (equations.cpp:169-176,
 * jacobians.cpp:10-152) with P = (rho, p, v, A, J) the primitive variables,
 * and the two Riemann solvers that take |M| from its eigendecomposition
 * (fluxes.cpp:22-64). */

/* dE/drho at constant p and A, one material (derivatives.cpp:11-45) */
static double dEdrho_Q(const double *Q, double pp, const gpr_par *p) {
  return dedrho(Q[0], pp, p) + dC_0drho(Q[0], p) / 4 * devGsq(Q + 5);
}

/* dsigma_ij / dA_mn at constant rho (state.cpp:71-84); A, G = A^T A and
 * A dev(G) row-major */
static double dsig_dA(double r, double cs2, const double *A, const double *G,
                      const double *AdG, int i, int j, int m, int n) {
  double t = A[m * 3 + i] * G[j * 3 + n] + A[m * 3 + j] * G[i * 3 + n] -
             2. / 3. * G[i * 3 + j] * A[m * 3 + n];
  if (i == n)
    t += AdG[m * 3 + j];
  if (j == n)
    t += AdG[m * 3 + i];
  return -r * cs2 * t;
}

/* dF_d/dP (jacobians.cpp:10-106), V x V row-major */
static void jac_dFdP(const double *Q, int d, const gpr_par *p, int th, double *M) {
  int V = th ? 17 : 14;
  for (int i = 0; i < V * V; i++)
    M[i] = 0.0;
#define MM(i, j) M[(i) * V + (j)]
  double r = Q[0], E = Q[1] / r, pp = orc_pressure(Q, p, th);
  double v[3] = {Q[2] / r, Q[3] / r, Q[4] / r};
  const double *A = Q + 5;
  double Er = dEdrho_Q(Q, pp, p), Ep = dedp(r, p);
  double vd = v[d], rvd = r * vd;
  MM(0, 0) = vd;
  MM(0, 2 + d) = r;
  MM(1, 0) = (E + r * Er) * vd;
  MM(1, 1) = (r * Ep + 1) * vd;
  for (int k = 0; k < 3; k++)
    MM(1, 2 + k) = r * vd * v[k];
  MM(1, 2 + d) += r * E + pp;
  for (int k = 0; k < 3; k++) {
    MM(2 + k, 0) = vd * v[k];
    MM(2 + k, 2 + k) = rvd;
  }
  for (int k = 0; k < 3; k++)
    MM(2 + k, 2 + d) += r * v[k];
  MM(2 + d, 1) = 1.0;

  /* viscous part: the stress column sigma_{.d}, its rho derivative at
   * constant A, and dE/dA = C_0 A dev(G) */
  double sd[3], sr[3], G[9], AdG[9], psi[9];
  double cs2 = c_s2(r, p), C0 = C_0(r, p);
  sigma_col(Q, p, d, sd);
  for (int k = 0; k < 3; k++)
    sr[k] = (1 / r + dc_s2drho(r, p) / cs2) * sd[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  AdevG(A, AdG);
  for (int k = 0; k < 9; k++)
    psi[k] = C0 * AdG[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double t = 0.;
      for (int k = 0; k < 3; k++) {
        double ds = dsig_dA(r, cs2, A, G, AdG, d, k, i, j);
        MM(2 + k, 5 + i * 3 + j) = -ds;
        t -= v[k] * ds;
      }
      MM(1, 5 + i * 3 + j) = t;
    }
  MM(1, 0) -= sr[0] * v[0] + sr[1] * v[1] + sr[2] * v[2];
  for (int k = 0; k < 3; k++)
    MM(1, 2 + k) -= sd[k];
  for (int k = 0; k < 9; k++)
    MM(1, 5 + k) += r * vd * psi[k];
  for (int k = 0; k < 3; k++)
    MM(2 + k, 0) -= sr[k];
  for (int row = 0; row < 3; row++)
    for (int k = 0; k < 3; k++) {
      MM(5 + 3 * row + d, 2 + k) = A[row * 3 + k];
      MM(5 + 3 * row + d, 5 + 3 * row + k) = v[k];
    }

  if (th) {
    double T = temperature_mg(r, pp, p), Tr = dTdrho(r, pp, p), Tp = dTdp(r, p);
    double J[3] = {Q[14] / r, Q[15] / r, Q[16] / r};
    double H[3] = {p->calpha2 * Q[14] / r, p->calpha2 * Q[15] / r, p->calpha2 * Q[16] / r};
    MM(1, 0) += Tr * H[d];
    MM(1, 1) += Tp * H[d];
    for (int k = 0; k < 3; k++)
      MM(1, 14 + k) = rvd * H[k];
    MM(1, 14 + d) += p->calpha2 * T;
    for (int k = 0; k < 3; k++)
      MM(14 + k, 0) = vd * J[k];
    MM(14 + d, 0) += Tr;
    MM(14 + d, 1) = Tp;
    for (int k = 0; k < 3; k++) {
      MM(14 + k, 2 + d) = r * J[k];
      MM(14 + k, 14 + k) = rvd;
    }
  }
#undef MM
}

/* dP/dQ (jacobians.cpp:108-152), V x V row-major */
static void jac_dPdQ(const double *Q, const gpr_par *p, int th, double *M) {
  int V = th ? 17 : 14;
  for (int i = 0; i < V * V; i++)
    M[i] = i % (V + 1) == 0 ? 1.0 : 0.0;
#define MM(i, j) M[(i) * V + (j)]
  double r = Q[0], E = Q[1] / r, pp = orc_pressure(Q, p, th);
  double v[3] = {Q[2] / r, Q[3] / r, Q[4] / r};
  double Er = dEdrho_Q(Q, pp, p), Ep = dedp(r, p);
  double Gi = 1 / (r * Ep);
  double t = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - (E + r * Er);
  double J[3] = {0, 0, 0};
  if (th) {
    for (int k = 0; k < 3; k++)
      J[k] = Q[14 + k] / r;
    t += p->calpha2 * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
  }
  MM(1, 0) = Gi * t;
  MM(1, 1) = Gi;
  for (int k = 0; k < 3; k++) {
    MM(1, 2 + k) = -Gi * v[k];
    MM(2 + k, 0) = -v[k] / r;
    MM(2 + k, 2 + k) = 1 / r;
  }
  double AdG[9];
  AdevG(Q + 5, AdG);
  double C0 = C_0(r, p);
  for (int k = 0; k < 9; k++)
    MM(1, 5 + k) = -Gi * r * (C0 * AdG[k]);
  if (th)
    for (int k = 0; k < 3; k++) {
      MM(1, 14 + k) = -Gi * (p->calpha2 * Q[14 + k] / r);
      MM(14 + k, 0) = -J[k] / r;
      MM(14 + k, 14 + k) = 1 / r;
    }
#undef MM
}

/* equations.cpp:169-176 with block() (equations.cpp:69-87) */
void orc_system_matrix(double *ret, const double *Q, int d, const gpr_par *p, int thermal) {
  int V = thermal ? 17 : 14;
  double F[MAXV * MAXV], P[MAXV * MAXV];
  jac_dFdP(Q, d, p, thermal, F);
  jac_dPdQ(Q, p, thermal, P);
  for (int i = 0; i < V; i++)
    for (int j = 0; j < V; j++) {
      double s = 0;
      for (int k = 0; k < V; k++)
        s += F[i * V + k] * P[k * V + j];
      ret[i * V + j] = s;
    }
  double r = Q[0], vd = Q[2 + d] / r;
  for (int i = 5; i < 14; i++)
    ret[i * V + i] += vd;
  for (int row = 0; row < 3; row++)
    for (int k = 0; k < 3; k++)
      ret[(5 + 3 * row + d) * V + 5 + 3 * row + k] -= Q[2 + k] / r;
}

/* out = Re(R |Lambda| R^-1 x) for a real n x n row-major J = R Lambda R^-1,
 * the product D_ROE and D_OSH form from Eigen's EigenSolver and a pivoted QR
 * solve (fluxes.cpp:34-37,56-59).  Here: Householder reduction to Hessenberg
 * form, complex single-shift QR to a Schur form J = Z T Z^H (Wilkinson
 * shifts, an exceptional shift every 10th iteration), the eigenvectors of T
 * by back substitution (U unit upper triangular; a pivot below eps |T| is
 * raised to it, as in EISPACK's hqr2), R = Z U, and R b = x by Gaussian
 * elimination with partial pivoting.  Returns 0, or -1 if QR does not
 * converge in 60 n iterations. */
static int eig_abs_apply(const double *J, int n, const double *x, double *out) {
  double complex H[MAXV][MAXV], Z[MAXV][MAXV], U[MAXV][MAXV], R[MAXV][MAXV], b[MAXV];
  double Hr[MAXV][MAXV], Q[MAXV][MAXV];
  const double eps = 2.220446049250313e-16;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      Hr[i][j] = J[i * n + j];
      Q[i][j] = i == j;
    }
  /* Householder reduction to upper Hessenberg form, Q accumulated */
  for (int k = 0; k < n - 2; k++) {
    double v[MAXV], nrm = 0, tail = 0;
    for (int i = k + 1; i < n; i++) {
      v[i] = Hr[i][k];
      nrm += v[i] * v[i];
      if (i > k + 1)
        tail += v[i] * v[i];
    }
    if (tail == 0.0)
      continue;
    nrm = sqrt(nrm);
    double alpha = v[k + 1] >= 0 ? -nrm : nrm;
    v[k + 1] -= alpha;
    double vv = 0;
    for (int i = k + 1; i < n; i++)
      vv += v[i] * v[i];
    double beta = 2.0 / vv;
    for (int j = 0; j < n; j++) { /* H <- P H */
      double s = 0;
      for (int i = k + 1; i < n; i++)
        s += v[i] * Hr[i][j];
      s *= beta;
      for (int i = k + 1; i < n; i++)
        Hr[i][j] -= s * v[i];
    }
    for (int i = 0; i < n; i++) { /* H <- H P, Q <- Q P */
      double s = 0, t = 0;
      for (int j = k + 1; j < n; j++) {
        s += Hr[i][j] * v[j];
        t += Q[i][j] * v[j];
      }
      s *= beta;
      t *= beta;
      for (int j = k + 1; j < n; j++) {
        Hr[i][j] -= s * v[j];
        Q[i][j] -= t * v[j];
      }
    }
    Hr[k + 1][k] = alpha;
    for (int i = k + 2; i < n; i++)
      Hr[i][k] = 0.0;
  }
  double norm = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      H[i][j] = Hr[i][j];
      Z[i][j] = Q[i][j];
      norm += fabs(Hr[i][j]);
    }
  if (norm == 0.0) {
    for (int i = 0; i < n; i++)
      out[i] = 0.0;
    return 0;
  }
  /* complex QR on the active window [l, hi], bulge chased by rotations */
  int hi = n - 1, its = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; l--) {
      double s = cabs(H[l - 1][l - 1]) + cabs(H[l][l]);
      if (s == 0.0)
        s = norm;
      if (cabs(H[l][l - 1]) <= eps * s) {
        H[l][l - 1] = 0.0;
        break;
      }
    }
    if (l == hi) {
      hi--;
      its = 0;
      continue;
    }
    if (++total > 60 * n)
      return -1;
    its++;
    double complex mu;
    if (its % 10 == 0) {
      mu = H[hi][hi] + 0.75 * fabs(creal(H[hi][hi - 1]));
    } else {
      double complex a = H[hi - 1][hi - 1], bb = H[hi - 1][hi], c = H[hi][hi - 1],
                     dd = H[hi][hi];
      double complex tr2 = (a + dd) / 2, disc = csqrt((a - dd) * (a - dd) / 4 + bb * c);
      double complex l1 = tr2 + disc, l2 = tr2 - disc;
      mu = cabs(l1 - dd) <= cabs(l2 - dd) ? l1 : l2;
    }
    for (int k = l; k < hi; k++) {
      double complex xx = k == l ? H[l][l] - mu : H[k][k - 1];
      double complex yy = k == l ? H[l + 1][l] : H[k + 1][k - 1];
      double rr = sqrt(creal(xx * conj(xx)) + creal(yy * conj(yy)));
      if (rr == 0.0)
        continue;
      double complex c = xx / rr, s = yy / rr;
      if (k > l) {
        H[k][k - 1] = rr;
        H[k + 1][k - 1] = 0.0;
      }
      for (int j = k; j < n; j++) { /* rows k, k+1 <- G rows */
        double complex h0 = H[k][j], h1 = H[k + 1][j];
        H[k][j] = conj(c) * h0 + conj(s) * h1;
        H[k + 1][j] = -s * h0 + c * h1;
      }
      int top = k + 2 < hi ? k + 2 : hi;
      for (int i = 0; i <= top; i++) { /* columns k, k+1 <- columns G^H */
        double complex h0 = H[i][k], h1 = H[i][k + 1];
        H[i][k] = h0 * c + h1 * s;
        H[i][k + 1] = -h0 * conj(s) + h1 * conj(c);
      }
      for (int i = 0; i < n; i++) {
        double complex z0 = Z[i][k], z1 = Z[i][k + 1];
        Z[i][k] = z0 * c + z1 * s;
        Z[i][k + 1] = -z0 * conj(s) + z1 * conj(c);
      }
    }
  }
  /* eigenvectors of the triangular T = H */
  double tn = 0;
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++)
      tn += cabs(H[i][j]);
  double small = eps * tn;
  for (int k = 0; k < n; k++) {
    for (int j = 0; j < n; j++)
      U[j][k] = 0.0;
    U[k][k] = 1.0;
    for (int j = k - 1; j >= 0; j--) {
      double complex s = 0;
      for (int m = j + 1; m <= k; m++)
        s += H[j][m] * U[m][k];
      double complex den = H[j][j] - H[k][k];
      if (cabs(den) < small)
        den = small;
      U[j][k] = -s / den;
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double complex s = 0;
      for (int m = 0; m <= j; m++)
        s += Z[i][m] * U[m][j];
      R[i][j] = s;
    }
  /* R b = x, partial pivoting (R is overwritten) */
  double complex A[MAXV][MAXV];
  for (int i = 0; i < n; i++) {
    b[i] = x[i];
    for (int j = 0; j < n; j++)
      A[i][j] = R[i][j];
  }
  for (int k = 0; k < n; k++) {
    int piv = k;
    for (int i = k + 1; i < n; i++)
      if (cabs(A[i][k]) > cabs(A[piv][k]))
        piv = i;
    if (piv != k) {
      for (int j = 0; j < n; j++) {
        double complex t = A[k][j];
        A[k][j] = A[piv][j];
        A[piv][j] = t;
      }
      double complex t = b[k];
      b[k] = b[piv];
      b[piv] = t;
    }
    for (int i = k + 1; i < n; i++) {
      double complex f = A[i][k] / A[k][k];
      for (int j = k; j < n; j++)
        A[i][j] -= f * A[k][j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double complex s = b[i];
    for (int j = i + 1; j < n; j++)
      s -= A[i][j] * b[j];
    b[i] = s / A[i][i];
  }
  for (int i = 0; i < n; i++)
    b[i] *= cabs(H[i][i]);
  for (int i = 0; i < n; i++) {
    double complex s = 0;
    for (int j = 0; j < n; j++)
      s += R[i][j] * b[j];
    out[i] = creal(s);
  }
  return 0;
}

/* fluxes.cpp:45-64: |M(qbar)| (qL - qR) + F(qL) + F(qR), with
 * M(qbar) = sum_i W_i M(qR + x_i (qL - qR)) */
int orc_d_roe(double *ret, const double *qL, const double *qR, int d, const gpr_par *p,
              const orc_cfg *c) {
  const consts_t *K = get_consts(c->N);
  int th = c->thermal, V = th ? 17 : 14;
  double dq[MAXV], q[MAXV], M[MAXV * MAXV], Mb[MAXV * MAXV];
  for (int i = 0; i < V; i++)
    dq[i] = qL[i] - qR[i];
  for (int i = 0; i < V * V; i++)
    Mb[i] = 0.0;
  for (int s = 0; s < c->N; s++) {
    for (int i = 0; i < V; i++)
      q[i] = qR[i] + K->x[s] * dq[i];
    orc_system_matrix(M, q, d, p, th);
    for (int i = 0; i < V * V; i++)
      Mb[i] += K->w[s] * M[i];
  }
  int rc = eig_abs_apply(Mb, V, dq, ret);
  orc_flux(ret, qL, d, p, th);
  orc_flux(ret, qR, d, p, th);
  return rc;
}

/* fluxes.cpp:22-43: sum_i W_i |M(qR + x_i (qL - qR))| (qL - qR) + F(qL) + F(qR) */
int orc_d_osh(double *ret, const double *qL, const double *qR, int d, const gpr_par *p,
              const orc_cfg *c) {
  const consts_t *K = get_consts(c->N);
  int th = c->thermal, V = th ? 17 : 14;
  double dq[MAXV], q[MAXV], M[MAXV * MAXV], t[MAXV];
  int rc = 0;
  for (int i = 0; i < V; i++) {
    dq[i] = qL[i] - qR[i];
    ret[i] = 0.0;
  }
  for (int s = 0; s < c->N; s++) {
    for (int i = 0; i < V; i++)
      q[i] = qR[i] + K->x[s] * dq[i];
    orc_system_matrix(M, q, d, p, th);
    rc |= eig_abs_apply(M, V, dq, t);
    for (int i = 0; i < V; i++)
      ret[i] += K->w[s] * t[i];
  }
  orc_flux(ret, qL, d, p, th);
  orc_flux(ret, qR, d, p, th);
  return rc;
}

/* the Riemann solver the configuration selects (fv.cpp:129-139,202-212) */
static void riemann_flux(double *ret, const double *qL, const double *qR, int d,
                         const gpr_par *p, const orc_cfg *c) {
  if (c->flux == 1)
    orc_d_roe(ret, qL, qR, d, p, c);
  else if (c->flux == 2)
    orc_d_osh(ret, qL, qR, d, p, c);
  else
    orc_d_rus(ret, qL, qR, d, p, c);
}

/* ------------------------------------------------------------------ grid */
static void bc_copy_row(double *dst, const double *src, int V) {
  memcpy(dst, src, sizeof(double) * V);
}
/* boundaries1 (grid.cpp:12-125) on a strided line: rows of u at u + i*su,
 * rows of ub at ub + i*sb, with types (left, right) and direction d: 0
 * transmissive, 1 periodic, 2 slip, 3 stick, 4 moving (grid.cpp:44-54,
 * 100-110), 5 symmetric; DESTRESS = false (options.h:12). */
static void bc_line(const double *u, long su, double *ub, long sb, int n, int N,
                    const int32_t *bt, int d, int V) {
  for (int side = 0; side < 2; side++) {
    for (int i = 0; i < N; i++) {
      int dst = side == 0 ? i : i + n + N;
      int src;
      switch (bt[side]) {
      case 0:
        src = side == 0 ? 0 : n - 1;
        break;
      case 1:
        src = side == 0 ? n - N + i : i;
        break;
      default:
        src = side == 0 ? N - 1 - i : n - 1 - i;
        break;
      }
      double *row = ub + dst * sb;
      bc_copy_row(row, u + src * su, V);
      if (bt[side] == 2) {
        row[2 + d] *= -1.;
      } else if (bt[side] == 3) {
        row[2] *= -1.;
        row[3] *= -1.;
        row[4] *= -1.;
      } else if (bt[side] == 4) {
        const double s = u[src * su + 2];
        row[2] *= -1.;
        row[3] *= -1.;
        row[4] *= -1.;
        row[2] += 2. * row[0];
        row[1] += (row[2] * row[2] - s * s) / (2. * row[0]);
      } else if (bt[side] == 5) {
        row[2 + d] *= -1.;
        for (int k = 0; k < 3; k++)
          row[5 + d * 3 + k] *= -1.;
        for (int k = 0; k < 3; k++)
          row[5 + k * 3 + d] *= -1.;
      }
    }
  }
}

void orc_boundaries(const double *u, double *ub, const int32_t *nX,
                    const int32_t *bcs, const orc_cfg *c) {
  int N = c->N, V = c->thermal ? 17 : 14;
  int nx = nX[0];
  if (c->ndim == 1) {
    memcpy(ub + N * V, u, sizeof(double) * nx * V);
    bc_line(u, V, ub, V, nx, N, bcs, 0, V);
    return;
  }
  int ny = nX[1];
  long sy = (long)(ny + 2 * N) * V;
  for (int i = 0; i < nx; i++)
    memcpy(ub + ((long)(i + N) * (ny + 2 * N) + N) * V, u + (long)i * ny * V,
           sizeof(double) * ny * V);
  for (int j = 0; j < ny; j++)
    bc_line(u + (long)j * V, (long)ny * V, ub + (long)(j + N) * V, sy, nx, N, bcs, 0, V);
  for (int i = 0; i < nx + 2 * N; i++)
    bc_line(ub + ((long)i * (ny + 2 * N) + N) * V, V, ub + (long)i * (ny + 2 * N) * V,
            V, ny, N, bcs + 2, 1, V);
}

void orc_extend_mask(const uint8_t *mask, uint8_t *maskb, const int32_t *nX,
                     const orc_cfg *c) {
  int nx = nX[0];
  if (c->ndim == 1) {
    for (int i = 0; i < nx; i++)
      maskb[i + 1] = mask[i] || (i > 0 && mask[i - 1]) || (i < nx - 1 && mask[i + 1]);
    maskb[0] = maskb[1];
    maskb[nx + 1] = maskb[nx];
    return;
  }
  /* grid.cpp:195-226, with the border assignments hoisted out of the cell
   * loop (they do not depend on it) */
  int ny = nX[1];
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      int ind = i * ny + j;
      maskb[(i + 1) * (ny + 2) + (j + 1)] =
          mask[ind] || (i > 0 && mask[ind - ny]) || (i < nx - 1 && mask[ind + ny]) ||
          (j > 0 && mask[ind - 1]) || (j < ny - 1 && mask[ind + 1]);
    }
  for (int i = 0; i < nx; i++) {
    maskb[(i + 1) * (ny + 2)] = mask[i * ny];
    maskb[(i + 1) * (ny + 2) + (ny + 1)] = mask[i * ny + (ny - 1)];
  }
  for (int j = 0; j < ny; j++) {
    maskb[j + 1] = mask[j];
    maskb[(nx + 1) * (ny + 2) + (j + 1)] = mask[(nx - 1) * ny + j];
  }
  maskb[0] = 0;
  maskb[ny + 1] = 0;
  maskb[(nx + 1) * (ny + 2)] = 0;
  maskb[(nx + 1) * (ny + 2) + (ny + 1)] = 0;
}

/* ------------------------------------------------------------------ WENO */
/* coeffs (weno.cpp:17-73) on one line: data rows at data + r*sd (2N-1 rows),
 * output rows at out + a*so (N rows).
 *
 * Evaluated in difference form about the central value c = data[N-1]: every
 * stencil matrix maps constants to themselves (each row of M^-1 sums to 1,
 * since the basis sums to 1) and the oscillation indicator annihilates them,
 * so  w_s = c + M_s^-1 (d - c),  w_s' SIG w_s = (w_s - c)' SIG (w_s - c)  and
 * sum w_s o_s / sum o_s = c + sum (w_s - c) o_s / sum o_s  exactly.  The
 * direct form loses the indicator to cancellation when |c| >> |d - c| (its
 * result then moves by ~1e-8 relative under a change of evaluation order);
 * this form keeps it to a few ulps. */
static void weno_coeffs(double *out, long so, const double *data, long sd, int ncell,
                        int gap, int ind, int V, const consts_t *K) {
  int N = K->N;
  int FN2 = (N - 1) / 2;
  for (int v = 0; v < V; v++) {
    double c = data[(N - 1) * sd + v];
    double dd[2 * MAXN - 1];
    for (int r = 0; r < 2 * N - 1; r++)
      dd[r] = data[r * sd + v] - c;
    double wL[MAXN], wR[MAXN], wC[MAXN], oL, oR, oC = 0;
    for (int a = 0; a < N; a++) {
      double sl = 0, sr = 0, sc = 0;
      for (int k = 0; k < N; k++) {
        sl += K->mLinv[a][k] * dd[k];
        sr += K->mRinv[a][k] * dd[N - 1 + k];
        sc += K->mCLinv[a][k] * dd[FN2 + k];
      }
      wL[a] = sl;
      wR[a] = sr;
      wC[a] = sc;
    }
    double *ws[3] = {wL, wR, wC};
    double os[3];
    for (int s = 0; s < (N > 2 ? 3 : 2); s++) {
      double q = 0;
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          q += ws[s][i] * K->sig[i][j] * ws[s][j];
      double t = q + WENO_EPS;
      double t2 = t * t, t4 = t2 * t2, den = t4 * t4;
      os[s] = (s == 2 ? LAMC : LAMS) / den;
    }
    oL = os[0];
    oR = os[1];
    if (N > 2)
      oC = os[2];
    double num[MAXN] = {0, 0, 0}, den = 0;
    if (ind >= gap) {
      for (int a = 0; a < N; a++)
        num[a] += wL[a] * oL;
      den += oL;
    }
    if (ind <= ncell - gap - 1) {
      for (int a = 0; a < N; a++)
        num[a] += wR[a] * oR;
      den += oR;
    }
    if (N > 2 && ind + FN2 >= gap) {
      for (int a = 0; a < N; a++)
        num[a] += wC[a] * oC;
      den += oC;
    }
    for (int a = 0; a < N; a++)
      out[a * so + v] = c + num[a] / den;
  }
}

void orc_weno_launcher(double *wh, const double *ub, const int32_t *nX,
                       const orc_cfg *c) {
  const consts_t *K = get_consts(c->N);
  int N = c->N, V = c->thermal ? 17 : 14;
  if (c->ndim == 1) {
    int NX = nX[0] + 2;
    for (int i = 0; i < NX; i++)
      weno_coeffs(wh + (long)i * N * V, V, ub + (long)i * V, V, NX, 0, i, V, K);
    return;
  }
  int NX = nX[0] + 2, NY = nX[1] + 2;
  int ncol = NY + 2 * (N - 1); /* = ny + 2N */
  double *ux = malloc(sizeof(double) * (size_t)NX * ncol * N * V);
  /* x-pass (weno.cpp:93-104) */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ncol; j++) {
    int gap = (j < N || j >= ncol - N) ? N : 0;
    for (int i = 0; i < NX; i++)
      weno_coeffs(ux + ((long)i * ncol + j) * N * V, V, ub + ((long)i * ncol + j) * V,
                  (long)ncol * V, NX, gap, i, V, K);
  }
  /* y-pass (weno.cpp:109-121) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NX; i++) {
    int gap = (i == 0 || i == NX - 1) ? 1 : 0;
    for (int ii = 0; ii < N; ii++)
      for (int j = 0; j < NY; j++)
        weno_coeffs(wh + (((long)i * NY + j) * N + ii) * N * V, V,
                    ux + (((long)i * ncol + j) * N + ii) * V, (long)N * V, NY, gap, j,
                    V, K);
  }
  free(ux);
}

/* ------------------------------------------------------------- predictor */
typedef struct {
  const consts_t *K;
  const gpr_par *p;
  int N, V, ndim, th, nn; /* nn = nodes per cell */
  double dt, dx, dy;
} cellctx;

/* rhs1 / rhs2 (dg1.cpp:8-35, dg2.cpp:8-71).  With src = 0 the stiff source
 * term is left out (used by the fallback iteration). */
static void dg_rhs(const cellctx *C, const double *q, const double *Ww, double *ret,
                   int with_src) {
  const consts_t *K = C->K;
  int N = C->N, V = C->V, th = C->th;
  const gpr_par *p = C->p;
  double F[27 * MAXV], G[27 * MAXV], dqx[MAXV], dqy[MAXV], S[MAXV], b1[MAXV], b2[MAXV];
  if (C->ndim == 1) {
    for (int t = 0; t < N; t++)
      for (int i = 0; i < N; i++) {
        int ind = t * N + i;
        const double *qn = q + ind * V;
        for (int v = 0; v < V; v++) {
          double s = 0;
          for (int j = 0; j < N; j++)
            s += K->der[i][j] * q[(t * N + j) * V + v];
          dqx[v] = s / C->dx;
        }
        if (with_src)
          orc_source(S, qn, p, th);
        else
          memset(S, 0, sizeof(double) * V);
        orc_bdot(b1, qn, dqx, 0, th);
        for (int v = 0; v < V; v++)
          ret[ind * V + v] = (S[v] - b1[v]) * (K->w[t] * K->w[i]);
        memset(F + ind * V, 0, sizeof(double) * V);
        orc_flux(F + ind * V, qn, 0, p, th);
      }
    for (int t = 0; t < N; t++)
      for (int i = 0; i < N; i++)
        for (int v = 0; v < V; v++) {
          double s = 0;
          for (int j = 0; j < N; j++)
            s += K->dgder[i][j] * F[(t * N + j) * V + v];
          ret[(t * N + i) * V + v] -= K->w[t] * s / C->dx;
        }
  } else {
    for (int t = 0; t < N; t++)
      for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++) {
          int ind = (t * N + a) * N + b;
          const double *qn = q + ind * V;
          for (int v = 0; v < V; v++) {
            double sx = 0, sy = 0;
            for (int j = 0; j < N; j++) {
              sx += K->der[a][j] * q[((t * N + j) * N + b) * V + v];
              sy += K->der[b][j] * q[((t * N + a) * N + j) * V + v];
            }
            dqx[v] = sx / C->dx;
            dqy[v] = sy / C->dy;
          }
          if (with_src)
            orc_source(S, qn, p, th);
          else
            memset(S, 0, sizeof(double) * V);
          orc_bdot(b1, qn, dqx, 0, th);
          orc_bdot(b2, qn, dqy, 1, th);
          double wt = K->w[t] * K->w[a] * K->w[b];
          for (int v = 0; v < V; v++)
            ret[ind * V + v] = (S[v] - (b1[v] + b2[v])) * wt;
          memset(F + ind * V, 0, sizeof(double) * V);
          memset(G + ind * V, 0, sizeof(double) * V);
          orc_flux(F + ind * V, qn, 0, p, th);
          orc_flux(G + ind * V, qn, 1, p, th);
        }
    for (int t = 0; t < N; t++)
      for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++)
          for (int v = 0; v < V; v++) {
            double sx = 0, sy = 0;
            for (int j = 0; j < N; j++) {
              sx += K->dgder[a][j] * K->w[b] * F[((t * N + j) * N + b) * V + v];
              sy += K->dgder[b][j] * G[((t * N + a) * N + j) * V + v];
            }
            ret[((t * N + a) * N + b) * V + v] -=
                K->w[t] * sx / C->dx + K->w[t] * K->w[a] * sy / C->dy;
          }
  }
  int tot = C->nn * V;
  for (int i = 0; i < tot; i++)
    ret[i] = ret[i] * C->dt + Ww[i];
}

/* initial_condition1/2 (dg1.cpp:68-73, dg2.cpp:92-100) */
static void dg_Ww(const cellctx *C, const double *w, double *Ww) {
  const consts_t *K = C->K;
  int N = C->N, V = C->V, ns = C->nn / N;
  for (int t = 0; t < N; t++)
    for (int s = 0; s < ns; s++) {
      double ws = C->ndim == 1 ? K->w[s] : K->w[s / N] * K->w[s % N];
      for (int v = 0; v < V; v++)
        Ww[(t * ns + s) * V + v] = K->endv[0][t] * ws * w[s * V + v];
    }
}

/* solve (DG_MAT (x) W (x) W) q = rhs (the DG_U1/DG_U2 solves) */
static void dg_solve(const cellctx *C, const double *rhs, double *q) {
  const consts_t *K = C->K;
  int N = C->N, V = C->V, ns = C->nn / N;
  for (int t = 0; t < N; t++)
    for (int s = 0; s < ns; s++) {
      double ws = C->ndim == 1 ? K->w[s] : K->w[s / N] * K->w[s % N];
      for (int v = 0; v < V; v++) {
        double acc = 0;
        for (int k = 0; k < N; k++)
          acc += K->dgmatinv[t][k] * rhs[(k * ns + s) * V + v];
        q[(t * ns + s) * V + v] = acc / ws;
      }
    }
}

void orc_dg_residual(double *out, const double *q, const double *w, double dt,
                     const double *dX, const gpr_par *p, const orc_cfg *c) {
  cellctx C = {get_consts(c->N), p, c->N, c->thermal ? 17 : 14, c->ndim, c->thermal,
               c->ndim == 1 ? c->N * c->N : c->N * c->N * c->N, dt, dX[0],
               c->ndim == 2 ? dX[1] : 1.0};
  double Ww[27 * MAXV];
  dg_Ww(&C, w, Ww);
  dg_rhs(&C, q, Ww, out, 1);
  const consts_t *K = C.K;
  int N = C.N, V = C.V, ns = C.nn / N;
  for (int t = 0; t < N; t++)
    for (int s = 0; s < ns; s++) {
      double ws = C.ndim == 1 ? K->w[s] : K->w[s / N] * K->w[s % N];
      for (int v = 0; v < V; v++) {
        double acc = 0;
        for (int k = 0; k < N; k++)
          acc += K->dgmat[t][k] * ws * q[(k * ns + s) * V + v];
        out[(t * ns + s) * V + v] -= acc;
      }
    }
}

/* dense LU solve with partial pivoting, n <= 64, a is n x n row-major */
static int lu_solve(double *a, double *b, int n) {
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(a[r * n + c]) > fabs(a[piv * n + c]))
        piv = r;
    if (a[piv * n + c] == 0.0)
      return -1;
    if (piv != c) {
      for (int j = 0; j < n; j++) {
        double t = a[c * n + j];
        a[c * n + j] = a[piv * n + j];
        a[piv * n + j] = t;
      }
      double t = b[c];
      b[c] = b[piv];
      b[piv] = t;
    }
    for (int r = c + 1; r < n; r++) {
      double f = a[r * n + c] / a[c * n + c];
      if (f == 0.0)
        continue;
      for (int j = c; j < n; j++)
        a[r * n + j] -= f * a[c * n + j];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double s = b[r];
    for (int j = r + 1; j < n; j++)
      s -= a[r * n + j] * b[j];
    b[r] = s / a[r * n + r];
  }
  return 0;
}

/* Stiff fallback for one cell: solve the same DG equations with the transport
 * lagged and the source implicit.  At each spatial node s, with r_t the lagged
 * right-hand side divided by the spatial weights,
 *     sum_k DG_MAT(t,k) X_k - dt W_t S(X_t) = r_t,   t = 0..N-1,
 * is solved by Newton (finite-difference source Jacobian).  Iterates until
 * max |q_new - q| <= tol (1 + |q|). */
static int stiff_fallback(const cellctx *C, const double *w, const double *Ww,
                          double *q, int maxit, double tol, int *iters) {
  const consts_t *K = C->K;
  int N = C->N, V = C->V, ns = C->nn / N, th = C->th;
  int NV = N * V;
  double R[27 * MAXV], qn[27 * MAXV];
  for (int t = 0; t < N; t++) /* standard initial guess (initial_guess.cpp:9-13) */
    memcpy(q + t * ns * V, w, sizeof(double) * ns * V);
  for (int it = 1; it <= maxit; it++) {
    dg_rhs(C, q, Ww, R, 0);
    memcpy(qn, q, sizeof(double) * C->nn * V);
    for (int s = 0; s < ns; s++) {
      double ws = C->ndim == 1 ? K->w[s] : K->w[s / N] * K->w[s % N];
      double r[3 * MAXV], X[3 * MAXV];
      for (int t = 0; t < N; t++)
        for (int v = 0; v < V; v++) {
          r[t * V + v] = R[(t * ns + s) * V + v] / ws;
          X[t * V + v] = qn[(t * ns + s) * V + v];
        }
      for (int nit = 0; nit < 30; nit++) {
        double F[3 * MAXV], Jm[51 * 51], S0[MAXV], Sp[MAXV], Sm[MAXV];
        memset(Jm, 0, sizeof(double) * NV * NV);
        for (int t = 0; t < N; t++) {
          orc_source(S0, X + t * V, C->p, th);
          for (int v = 0; v < V; v++) {
            double acc = 0;
            for (int k = 0; k < N; k++)
              acc += K->dgmat[t][k] * X[k * V + v];
            F[t * V + v] = acc - C->dt * K->w[t] * S0[v] - r[t * V + v];
            for (int k = 0; k < N; k++)
              Jm[(t * V + v) * NV + k * V + v] += K->dgmat[t][k];
          }
          for (int j = 0; j < V; j++) {
            double xs[MAXV];
            memcpy(xs, X + t * V, sizeof(double) * V);
            double h = 1e-7 * (1.0 + fabs(xs[j]));
            xs[j] += h;
            orc_source(Sp, xs, C->p, th);
            xs[j] -= 2 * h;
            orc_source(Sm, xs, C->p, th);
            for (int v = 0; v < V; v++)
              Jm[(t * V + v) * NV + t * V + j] -= C->dt * K->w[t] * (Sp[v] - Sm[v]) / (2 * h);
          }
        }
        double nrm = 0;
        for (int i = 0; i < NV; i++)
          nrm = fmax(nrm, fabs(F[i]));
        if (lu_solve(Jm, F, NV) != 0)
          return -1;
        double dmax = 0;
        for (int i = 0; i < NV; i++) {
          X[i] -= F[i];
          dmax = fmax(dmax, fabs(F[i]) / (1.0 + fabs(X[i])));
        }
        if (!(dmax == dmax))
          return -1;
        if (dmax < 1e-15 || nrm == 0.0)
          break;
      }
      for (int t = 0; t < N; t++)
        for (int v = 0; v < V; v++)
          qn[(t * ns + s) * V + v] = X[t * V + v];
    }
    int more = 0, bad = 0;
    for (int i = 0; i < C->nn * V; i++) {
      double d = fabs(qn[i] - q[i]);
      if (d > tol * (1 + fabs(q[i])))
        more = 1;
      if (qn[i] != qn[i])
        bad = 1;
    }
    memcpy(q, qn, sizeof(double) * C->nn * V);
    if (bad)
      return -1;
    if (!more) {
      *iters = it;
      return 0;
    }
  }
  *iters = maxit;
  return -1;
}

/* Orthogonal polar factor of a 3x3 matrix with det > 0 (Newton iteration
 * X <- (X + X^-T) / 2). */
static void polar_rotation(const double *A, double *R) {
  double X[9];
  memcpy(X, A, sizeof X);
  for (int it = 0; it < 40; it++) {
    double d = det3(X);
    double inv[9];
    inv[0] = (X[4] * X[8] - X[5] * X[7]) / d;
    inv[1] = (X[2] * X[7] - X[1] * X[8]) / d;
    inv[2] = (X[1] * X[5] - X[2] * X[4]) / d;
    inv[3] = (X[5] * X[6] - X[3] * X[8]) / d;
    inv[4] = (X[0] * X[8] - X[2] * X[6]) / d;
    inv[5] = (X[2] * X[3] - X[0] * X[5]) / d;
    inv[6] = (X[3] * X[7] - X[4] * X[6]) / d;
    inv[7] = (X[1] * X[6] - X[0] * X[7]) / d;
    inv[8] = (X[0] * X[4] - X[1] * X[3]) / d;
    double ch = 0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double nx = 0.5 * (X[i * 3 + j] + inv[j * 3 + i]);
        ch = fmax(ch, fabs(nx - X[i * 3 + j]));
        X[i * 3 + j] = nx;
      }
    if (ch < 1e-15)
      break;
  }
  memcpy(R, X, sizeof X);
}

/* Stiff-source preconditioner of one spatial node, frozen at the WENO state
 * ws of that node.  At a relaxed state A = s R (R a rotation) the distortion
 * source has Jacobian dS_A/dA = -kappa Pi, with Pi(X) = R dev(sym(R^T X)) an
 * orthogonal projector and kappa = 2 theta1inv cs^2 s^2; the thermal source
 * has dS_J/d(rho J) ~ -cα² T / (κ ρ).  Writes
 *   Ainv = (DG_MAT + dt kappa W)^-1,  Jinv = (DG_MAT + dt rJ W)^-1,  R. */
static void stiff_precond(const cellctx *C, const double *ws, double *Ainv, double *Jinv,
                          double *R) {
  const consts_t *K = C->K;
  int N = C->N;
  double r = ws[0];
  double dA = det3(ws + 5);
  double s13 = cbrt(dA);
  double kap = 2 * theta1inv(ws, C->p) * c_s2(r, C->p) * s13 * s13;
  if (!(kap >= 0) || !isfinite(kap))
    kap = 0;
  double rJ = 0;
  if (C->th) {
    double T = temperature_mg(r, orc_pressure(ws, C->p, 1), C->p);
    rJ = C->p->calpha2 * T / (C->p->kappa * r);
    if (!(rJ >= 0) || !isfinite(rJ))
      rJ = 0;
  }
  double m[9];
  for (int t = 0; t < N; t++)
    for (int k = 0; k < N; k++)
      m[t * N + k] = K->dgmat[t][k] + (t == k ? C->dt * kap * K->w[t] : 0.0);
  inv_small(m, Ainv, N);
  for (int t = 0; t < N; t++)
    for (int k = 0; k < N; k++)
      m[t * N + k] = K->dgmat[t][k] + (t == k ? C->dt * rJ * K->w[t] : 0.0);
  inv_small(m, Jinv, N);
  if (dA > 0)
    polar_rotation(ws + 5, R);
  else
    for (int i = 0; i < 9; i++)
      R[i] = (i % 4 == 0) ? 1.0 : 0.0;
}

/* Pi(X) = R dev(sym(R^T X)) */
static void proj_dev_sym(const double *R, const double *X, double *out) {
  double Y[9], S[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      Y[i * 3 + j] = R[0 * 3 + i] * X[0 * 3 + j] + R[1 * 3 + i] * X[1 * 3 + j] +
                     R[2 * 3 + i] * X[2 * 3 + j];
  double tr = (Y[0] + Y[4] + Y[8]) / 3;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      S[i * 3 + j] = 0.5 * (Y[i * 3 + j] + Y[j * 3 + i]) - (i == j ? tr : 0.0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = R[i * 3 + 0] * S[0 * 3 + j] + R[i * 3 + 1] * S[1 * 3 + j] +
                       R[i * 3 + 2] * S[2 * 3 + j];
}

/* One preconditioned Picard update q -> qn for the DG equations
 *   (DG_MAT (x) W_s) q = rhs(q):
 *   g = DG_MAT q - rhs(q)/W_s  (per spatial node s, per time node),
 *   qn = q - P^-1 g,  P^-1 = DG_MAT^-1 off the stiff directions and the
 *   stiff-source inverses of stiff_precond on them.
 * With kappa = rJ = 0 this is exactly the plain Picard update. */
static void prec_picard_update(const cellctx *C, const double *q, const double *Ww,
                               const double *w, double *qn) {
  const consts_t *K = C->K;
  int N = C->N, V = C->V, ns = C->nn / N;
  double rhs[27 * MAXV];
  dg_rhs(C, q, Ww, rhs, 1);
  for (int s = 0; s < ns; s++) {
    double ws = C->ndim == 1 ? K->w[s] : K->w[s / N] * K->w[s % N];
    double Ainv[9], Jinv[9], R[9];
    stiff_precond(C, w + s * V, Ainv, Jinv, R);
    double g[MAXN][MAXV], P[MAXN][9];
    for (int t = 0; t < N; t++) {
      for (int v = 0; v < V; v++) {
        double acc = 0;
        for (int k = 0; k < N; k++)
          acc += K->dgmat[t][k] * q[(k * ns + s) * V + v];
        g[t][v] = acc - rhs[(t * ns + s) * V + v] / ws;
      }
      proj_dev_sym(R, g[t] + 5, P[t]);
    }
    for (int t = 0; t < N; t++)
      for (int v = 0; v < V; v++) {
        double d = 0;
        if (v >= 5 && v < 14) {
          for (int k = 0; k < N; k++)
            d += K->dgmatinv[t][k] * (g[k][v] - P[k][v - 5]) + Ainv[t * N + k] * P[k][v - 5];
        } else if (v >= 14) {
          for (int k = 0; k < N; k++)
            d += Jinv[t * N + k] * g[k][v];
        } else {
          for (int k = 0; k < N; k++)
            d += K->dgmatinv[t][k] * g[k][v];
        }
        qn[(t * ns + s) * V + v] = q[(t * ns + s) * V + v] - d;
      }
  }
}

/* Preconditioned Picard from the standard initial guess; stopping rule of
 * dg2.cpp:280 with tolerance tol. */
static int prec_picard(const cellctx *C, const double *w, const double *Ww, double *q,
                       int maxit, double tol, int *iters) {
  int N = C->N, V = C->V, ns = C->nn / N;
  double qn[27 * MAXV];
  for (int t = 0; t < N; t++)
    memcpy(q + t * ns * V, w, sizeof(double) * ns * V);
  for (int it = 1; it <= maxit; it++) {
    prec_picard_update(C, q, Ww, w, qn);
    int more = 0, nan = 0;
    for (int i = 0; i < C->nn * V; i++) {
      if (fabs(qn[i] - q[i]) > tol * (1 + fabs(q[i])))
        more = 1;
      if (qn[i] != qn[i])
        nan = 1;
    }
    memcpy(q, qn, sizeof(double) * C->nn * V);
    if (!more) {
      *iters = it;
      return nan ? -1 : 0;
    }
  }
  *iters = maxit;
  return -1;
}

/* predictor1 / predictor2 (dg1.cpp:58-114, dg2.cpp:102-171) */
int orc_predictor(double *qh, const double *wh, double dt, const double *dX,
                  const int32_t *nX, const gpr_par *p, const uint8_t *mask,
                  const orc_cfg *c, orc_stats *st, int32_t *iters) {
  int N = c->N, V = c->thermal ? 17 : 14;
  long ncell = c->ndim == 1 ? nX[0] + 2 : (long)(nX[0] + 2) * (nX[1] + 2);
  int ns = c->ndim == 1 ? N : N * N;
  int nn = ns * N;
  cellctx C0 = {get_consts(N), p, N, V, c->ndim, c->thermal, nn, dt, dX[0],
                c->ndim == 2 ? dX[1] : 1.0};
  long s_cells = 0, s_iters = 0, s_fb = 0, s_fbit = 0, s_fail = 0;
  long hist[64] = {0};
  int status = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : s_cells, s_iters, s_fb, s_fbit, s_fail) reduction(| : status)
  for (long ind = 0; ind < ncell; ind++) {
    if (!mask[ind])
      continue;
    cellctx C = C0;
    const double *w = wh + ind * ns * V;
    double *qi = qh + ind * nn * V;
    double Ww[27 * MAXV], q0[27 * MAXV], q1[27 * MAXV], rhs[27 * MAXV];
    dg_Ww(&C, w, Ww);
    for (int t = 0; t < N; t++)
      memcpy(q0 + t * ns * V, w, sizeof(double) * ns * V);
    int done = 0, count, grow = 0;
    double prev = INFINITY;
    for (count = 0; count < c->dg_it; count++) {
      dg_rhs(&C, q0, Ww, rhs, 1);
      dg_solve(&C, rhs, q1);
      int more = 0, nan = 0;
      double ratio = 0;
      for (int i = 0; i < nn * V; i++) {
        double dd = fabs(q1[i] - q0[i]), lim = c->dg_tol * (1 + fabs(q0[i]));
        if (dd > lim)
          more = 1;
        if (q1[i] != q1[i])
          nan = 1;
        ratio = fmax(ratio, dd / lim);
      }
      if (more) {
        memcpy(q0, q1, sizeof(double) * nn * V);
        /* optional early exit: the stopping statistic grew bailout times in a row */
        grow = ratio > prev ? grow + 1 : 0;
        prev = ratio;
        if (c->bailout > 0 && grow >= c->bailout)
          break;
        if (c->bail_from > 0 && count + 1 > c->bail_from && grow > 0)
          break;
        continue;
      } else if (nan) {
        break;
      } else {
        memcpy(qi, q1, sizeof(double) * nn * V);
        done = 1;
        break;
      }
    }
    s_cells++;
    s_iters += done ? count + 1 : count;
    int hidx = done ? (count + 1 < 63 ? count + 1 : 62) : 63;
#pragma omp atomic
    hist[hidx]++;
    int it_used = done ? count + 1 : -1;
    if (!done) {
      int fit = 0;
      int rc = c->fallback == 1
                   ? stiff_fallback(&C, w, Ww, qi, c->stiff_it, c->stiff_tol, &fit)
                   : prec_picard(&C, w, Ww, qi, c->stiff_it, c->stiff_tol, &fit);
      s_fb++;
      s_fbit += fit;
      if (rc != 0) {
        s_fail++;
        status |= 1;
      }
      it_used = 1000 + fit;
    }
    if (iters)
      iters[ind] = it_used;
  }
  if (st) {
    st->cells += s_cells;
    st->picard_iters += s_iters;
    st->fallback_cells += s_fb;
    st->fallback_iters += s_fbit;
    st->failed_cells += s_fail;
    for (int i = 0; i < 64; i++)
      st->hist[i] += hist[i];
  }
  return status;
}

/* ------------------------------------------------------------------- FV */
void orc_fv_launcher(double *u, const double *qh, const int32_t *nX, double dt,
                     const double *dX, const gpr_par *p, const uint8_t *mask,
                     const orc_cfg *c) {
  orc_fv_launcher_ex(u, qh, nX, dt, dX, p, mask, 1, 1, c);
}

/* fv.cpp:11-297 with fv_launcher's SOURCES and TIME flags: TIME = 0 reads one
 * time level per cell (nt = 1) with weight dt (fv.cpp:46-54, 99-109, 151-160,
 * 270-279) */
void orc_fv_launcher_ex(double *u, const double *qh, const int32_t *nX, double dt,
                        const double *dX, const gpr_par *p, const uint8_t *mask,
                        int sources, int time, const orc_cfg *c) {
  orc_fv_parts(u, qh, nX, dt, dX, p, mask, sources, time, 3, c);
}

/* centers1/2 (parts & 1) and interfs1/2 (parts & 2), fv.cpp:11-279 */
void orc_fv_parts(double *u, const double *qh, const int32_t *nX, double dt,
                  const double *dX, const gpr_par *p, const uint8_t *mask,
                  int sources, int time, int parts, const orc_cfg *c) {
  const consts_t *K = get_consts(c->N);
  int N = c->N, V = c->thermal ? 17 : 14, th = c->thermal;
  int nt = time ? N : 1;
  double S[MAXV], b1[MAXV], b2[MAXV], f[MAXV], bb[MAXV];
  if (c->ndim == 1) {
    int nx = nX[0];
    double dx = dX[0];
    /* centers1 (fv.cpp:11-54) */
    for (int t = 0; t < ((parts & 1) ? nt : 0); t++) {
      double wt = time ? dt * K->w[t] : dt;
      for (int i = 0; i < nx; i++) {
        if (!mask[i + 1])
          continue;
        const double *qc = qh + ((long)(i + 1) * nt + t) * N * V;
        for (int s = 0; s < N; s++) {
          double dq[MAXV];
          for (int v = 0; v < V; v++) {
            double acc = 0;
            for (int j = 0; j < N; j++)
              acc += K->der[s][j] * qc[j * V + v];
            dq[v] = acc;
          }
          if (sources)
            orc_source(S, qc + s * V, p, th);
          else
            memset(S, 0, sizeof S);
          orc_bdot(b1, qc + s * V, dq, 0, th);
          for (int v = 0; v < V; v++)
            u[(long)i * V + v] += wt * K->w[s] * (S[v] - b1[v] / dx);
        }
      }
    }
    /* interfs1 (fv.cpp:111-160) */
    for (int t = 0; t < ((parts & 2) ? nt : 0); t++) {
      double k = (time ? dt * K->w[t] : dt) / (2. * dx);
      for (int i = 0; i < nx + 1; i++) {
        if (!(mask[i] && mask[i + 1]))
          continue;
        const double *ql = qh + ((long)i * nt + t) * N * V;
        const double *qr = qh + ((long)(i + 1) * nt + t) * N * V;
        double L[MAXV], R[MAXV];
        for (int v = 0; v < V; v++) {
          double a = 0, b = 0;
          for (int j = 0; j < N; j++) {
            a += K->endv[1][j] * ql[j * V + v];
            b += K->endv[0][j] * qr[j * V + v];
          }
          L[v] = a;
          R[v] = b;
        }
        riemann_flux(f, L, R, 0, p, c);
        orc_bint(bb, L, R, 0, p, c);
        if (i > 0)
          for (int v = 0; v < V; v++)
            u[(long)(i - 1) * V + v] -= k * (bb[v] + f[v]);
        if (i < nx)
          for (int v = 0; v < V; v++)
            u[(long)i * V + v] -= k * (bb[v] - f[v]);
      }
    }
    return;
  }
  int nx = nX[0], ny = nX[1];
  double dx = dX[0], dy = dX[1];
  int NNV = N * N * V;
  /* centers2 (fv.cpp:56-109) */
  for (int t = 0; t < ((parts & 1) ? nt : 0); t++) {
    double wt = time ? dt * K->w[t] : dt;
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        long e = (long)(i + 1) * (ny + 2) + (j + 1);
        if (!mask[e])
          continue;
        const double *qc = qh + (e * nt + t) * NNV;
        for (int a = 0; a < N; a++)
          for (int b = 0; b < N; b++) {
            double dqx[MAXV], dqy[MAXV];
            for (int v = 0; v < V; v++) {
              double sx = 0, sy = 0;
              for (int k = 0; k < N; k++) {
                sx += K->der[a][k] * qc[(k * N + b) * V + v];
                sy += K->der[b][k] * qc[(a * N + k) * V + v];
              }
              dqx[v] = sx;
              dqy[v] = sy;
            }
            const double *qs = qc + (a * N + b) * V;
            if (sources)
              orc_source(S, qs, p, th);
            else
              memset(S, 0, sizeof S);
            orc_bdot(b1, qs, dqx, 0, th);
            orc_bdot(b2, qs, dqy, 1, th);
            for (int v = 0; v < V; v++) {
              double sv = S[v] - b1[v] / dx;
              sv -= b2[v] / dy;
              u[((long)i * ny + j) * V + v] += wt * K->w[a] * K->w[b] * sv;
            }
          }
      }
  }
  /* interfs2 (fv.cpp:162-279) */
  for (int t = 0; t < ((parts & 2) ? nt : 0); t++) {
    double wts = time ? dt * K->w[t] : dt;
    for (int i = 0; i < nx + 1; i++)
      for (int j = 0; j < ny + 1; j++) {
        long e0 = (long)i * (ny + 2) + j;
        if (!mask[e0])
          continue;
        const double *q0c = qh + (e0 * nt + t) * NNV;
        for (int d = 0; d < 2; d++) {
          long e1 = d == 0 ? e0 + (ny + 2) : e0 + 1;
          if (!mask[e1] || (d == 0 ? j <= 0 : i <= 0))
            continue;
          const double *q1c = qh + (e1 * nt + t) * NNV;
          double u0[MAXV] = {0}, u1[MAXV] = {0};
          double kk = wts / (2. * (d == 0 ? dx : dy));
          for (int s = 0; s < N; s++) {
            double L[MAXV], R[MAXV];
            for (int v = 0; v < V; v++) {
              double a = 0, b = 0;
              for (int k = 0; k < N; k++) {
                if (d == 0) {
                  a += K->endv[1][k] * q0c[(k * N + s) * V + v];
                  b += K->endv[0][k] * q1c[(k * N + s) * V + v];
                } else {
                  a += K->endv[1][k] * q0c[(s * N + k) * V + v];
                  b += K->endv[0][k] * q1c[(s * N + k) * V + v];
                }
              }
              L[v] = a;
              R[v] = b;
            }
            riemann_flux(f, L, R, d, p, c);
            orc_bint(bb, L, R, d, p, c);
            double xw = kk * K->w[s];
            int lo = d == 0 ? i > 0 : j > 0;
            int hi = d == 0 ? i < nx : j < ny;
            for (int v = 0; v < V; v++) {
              if (lo)
                u0[v] += xw * (bb[v] + f[v]);
              if (hi)
                u1[v] += xw * (bb[v] - f[v]);
            }
          }
          long uind0 = ((long)(i - 1) * ny + (j - 1)) * V;
          long uind1 = d == 0 ? ((long)i * ny + (j - 1)) * V : ((long)(i - 1) * ny + j) * V;
          if (d == 0 ? i > 0 : j > 0)
            for (int v = 0; v < V; v++)
              u[uind0 + v] -= u0[v];
          if (d == 0 ? i < nx : j < ny)
            for (int v = 0; v < V; v++)
              u[uind1 + v] -= u1[v];
        }
      }
  }
}

/* ------------------------------------------------------------- time step */
double orc_timestep(const double *u, const int32_t *nX, const double *dX,
                    double CFL, double t, double tf, long count,
                    const gpr_par *p, const orc_cfg *c) {
  int V = c->thermal ? 17 : 14;
  long ncell = c->ndim == 1 ? nX[0] : (long)nX[0] * nX[1];
  double MAX = 0.;
#pragma omp parallel for reduction(max : MAX)
  for (long ind = 0; ind < ncell; ind++) {
    double tmp = 0.;
    for (int d = 0; d < c->ndim; d++)
      tmp += orc_max_abs_eigs(u + ind * V, d, p, c->thermal) / dX[d];
    if (tmp > MAX)
      MAX = tmp;
  }
  double dt = CFL / MAX;
  if (count <= 5)
    dt *= 0.2;
  if (t + dt > tf)
    return tf - t;
  return dt;
}

/* utils.cpp:6-23 */
void orc_renorm_distortion(double *u, long ncell, const gpr_par *p, int V) {
  for (long i = 0; i < ncell; i++) {
    double r = u[i * V];
    if (r > 0) {
      double c = cbrt(r / (p->rho0 * det3(u + i * V + 5)));
      for (int k = 0; k < 9; k++)
        u[i * V + 5 + k] *= c;
    }
  }
}

/* utils.cpp:65-85: the singular values enter only through sum s_i^2 = |A|_F^2
 * and s0 s1 s2 = |det A| */
int orc_contorted(const double *u, long ncell, double tol, int V) {
  for (long i = 0; i < ncell; i++) {
    const double *A = u + i * V + 5;
    double f = 0;
    for (int k = 0; k < 9; k++)
      f += A[k] * A[k];
    double d13 = cbrt(fabs(det3(A)));
    double m0 = f / (d13 * d13) / 3.;
    if (m0 - 1 > tol)
      return 1;
  }
  return 0;
}

/* utils.cpp:44-63 */
void orc_reset_distortion(double *u, long ncell, const gpr_par *p, int V) {
  for (long i = 0; i < ncell; i++) {
    double r = u[i * V];
    double c = cbrt(r / p->rho0);
    u[i * V + 1] -= r * E_2A(r, u + i * V + 5, p);
    for (int k = 0; k < 9; k++)
      u[i * V + 5 + k] = 0.0;
    u[i * V + 5] = c;
    u[i * V + 9] = c;
    u[i * V + 13] = c;
  }
}

int orc_step(double *u, const int32_t *nX, const double *dX, double CFL,
             const int32_t *bcs, double *t, long *count, double tf,
             double contorted_tol, const gpr_par *p, const orc_cfg *c,
             double *dt_out, orc_stats *st) {
  int N = c->N, V = c->thermal ? 17 : 14;
  long ncell = c->ndim == 1 ? nX[0] : (long)nX[0] * nX[1];
  long next = c->ndim == 1 ? nX[0] + 2 : (long)(nX[0] + 2) * (nX[1] + 2);
  long nb = c->ndim == 1 ? nX[0] + 2 * N : (long)(nX[0] + 2 * N) * (nX[1] + 2 * N);
  int ns = c->ndim == 1 ? N : N * N;
  double dt = orc_timestep(u, nX, dX, CFL, *t, tf, *count, p, c);
  double *ub = malloc(sizeof(double) * nb * V);
  double *wh = malloc(sizeof(double) * next * ns * V);
  double *qh = calloc((size_t)next * ns * N * V, sizeof(double));
  uint8_t *m = malloc(ncell), *mb = malloc(next);
  memset(m, 1, ncell);
  orc_extend_mask(m, mb, nX, c);
  orc_boundaries(u, ub, nX, bcs, c);
  orc_weno_launcher(wh, ub, nX, c);
  int rc = orc_predictor(qh, wh, dt, dX, nX, p, mb, c, st, NULL);
  orc_fv_launcher(u, qh, nX, dt, dX, p, mb, c);
  *t += dt;
  *count += 1;
  orc_renorm_distortion(u, ncell, p, V);
  if (orc_contorted(u, ncell, contorted_tol, V))
    orc_reset_distortion(u, ncell, p, V);
  free(ub);
  free(wh);
  free(qh);
  free(m);
  free(mb);
  if (dt_out)
    *dt_out = dt;
  for (long i = 0; i < ncell * V; i++)
    if (u[i] != u[i])
      return -1;
  return rc;
}

/* ========================================================= split-WENO step
 * solvers/steppers.cpp:50-94, solvers/split/{ode,homogeneous}.cpp,
 * system/relaxation/analytic.cpp. */

/* 3x3 SVD A = U diag(s) V^T: cyclic Jacobi on the symmetric G = A^T A gives V
 * and s^2, then U = A V / s.  s descending by selection of the largest
 * remaining value, ties kept in column order (JacobiSVD's ordering). */
static void orc_svd3(const double *A, double *U, double *s, double *Vm) {
  double G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  for (int k = 0; k < 9; k++)
    Vm[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = G[1] * G[1] + G[2] * G[2] + G[5] * G[5];
    double dia = G[0] * G[0] + G[4] * G[4] + G[8] * G[8];
    if (off <= 1e-34 * dia)
      break;
    for (int pr = 0; pr < 3; pr++) {
      int p = pr == 2 ? 1 : 0, q = pr == 0 ? 1 : 2;
      double apq = G[p * 3 + q];
      if (apq == 0.0)
        continue;
      double th = (G[q * 3 + q] - G[p * 3 + p]) / (2 * apq);
      double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
      double c = 1 / sqrt(t * t + 1), sn = t * c;
      for (int k = 0; k < 3; k++) { /* G P */
        double a = G[k * 3 + p], b = G[k * 3 + q];
        G[k * 3 + p] = c * a - sn * b;
        G[k * 3 + q] = sn * a + c * b;
      }
      for (int k = 0; k < 3; k++) { /* P^T G */
        double a = G[p * 3 + k], b = G[q * 3 + k];
        G[p * 3 + k] = c * a - sn * b;
        G[q * 3 + k] = sn * a + c * b;
      }
      for (int k = 0; k < 3; k++) {
        double a = Vm[k * 3 + p], b = Vm[k * 3 + q];
        Vm[k * 3 + p] = c * a - sn * b;
        Vm[k * 3 + q] = sn * a + c * b;
      }
    }
  }
  for (int j = 0; j < 3; j++) {
    double sj = sqrt(G[j * 3 + j] > 0 ? G[j * 3 + j] : 0);
    s[j] = sj;
    for (int i = 0; i < 3; i++) {
      double x = A[i * 3 + 0] * Vm[0 * 3 + j] + A[i * 3 + 1] * Vm[1 * 3 + j] + A[i * 3 + 2] * Vm[2 * 3 + j];
      U[i * 3 + j] = sj > 0 ? x / sj : (i == j ? 1.0 : 0.0);
    }
  }
  for (int i = 0; i < 2; i++) {
    int pos = i;
    for (int j = i + 1; j < 3; j++)
      if (s[j] > s[pos])
        pos = j;
    if (pos != i) {
      double tmp = s[i];
      s[i] = s[pos];
      s[pos] = tmp;
      for (int k = 0; k < 3; k++) {
        tmp = U[k * 3 + i];
        U[k * 3 + i] = U[k * 3 + pos];
        U[k * 3 + pos] = tmp;
        tmp = Vm[k * 3 + i];
        Vm[k * 3 + i] = Vm[k * 3 + pos];
        Vm[k * 3 + pos] = tmp;
      }
    }
  }
}

static double posv(double x) { return x > 0. ? x : 0.; } /* std::max(0., x) */
static const double ORC_PI = 3.14159265358979323846;

/* analytic.cpp:17-70 */
static double nondim_time(double r, double detA3, double m0, double u0, double dt,
                          const gpr_par *p) {
  if (p->POWER_LAW) {
    double n = p->n;
    double cs2 = c_s2(r, p);
    double a = 9. * m0 - u0 - 9.;
    double b = 6. * m0 - u0 - 6.;
    double c, lam, t1, t2;
    if (p->SOLID) {
      c = (108. * a - 324. * b + 108. * a * a - 396. * a * b + 297. * b * b -
           24. * (a * a * b - 2. * a * b * b + b * b * b) - 4. * pow(a - b, 4.));
      if (c <= 0.)
        return 0.;
      lam = c / (18. * a - 36. * b + 9. * a * a - 132. / 5. * a * b + 33. / 2. * b * b -
                 8. / 7. * a * a * b + 2. * a * b * b - 8. / 9. * b * b * b - pow(a, 4.) / 6. +
                 16. / 27. * a * a * a * b - 4. / 5. * a * a * b * b + 16. / 33. * a * b * b * b -
                 pow(b, 4.) / 9.);
      t1 = n * lam / p->tau0 * pow(detA3, 4. * n + 7.);
      t2 = pow(sqrt(c) * r * cs2 / (6. * p->sigmaY), n);
      return 2. / (n * lam) * log(t1 * t2 * dt + 1.);
    }
    double k = (1 - n) / n;
    c = (108. * a - 324. * b + 180. * a * a - 612. * a * b + 459. * b * b -
         24. * (a * a * b - 2. * a * b * b + b * b * b) - 4. * pow(a - b, 4.));
    if (c <= 0.)
      return 0.;
    lam = c / (18. * a - 36. * b + 15. * a * a - 204. / 5. * a * b + 51. / 2. * b * b -
               8. / 7. * a * a * b + 2. * a * b * b - 8. / 9. * b * b * b - pow(a, 4.) / 6. +
               16. / 27. * a * a * a * b - 4. / 5. * a * a * b * b + 16. / 33. * a * b * b * b -
               pow(b, 4.) / 9.);
    double tau1 = 6 * pow(p->mu, 1. / n) / (p->rho0 * p->b02);
    t1 = k * lam / tau1 * pow(detA3, 4. * k + 7.);
    t2 = pow(sqrt(c) * r * cs2 / (6. * sqrt(3.)), k);
    return 2. / (k * lam) * log(t1 * t2 * dt + 1.);
  }
  return p->rho0 * p->b02 * pow(detA3, 7.) / (3. * p->mu) * dt;
}

/* analytic.cpp:72-129 */
static void analytic_distortion(double *Q, double dt, const gpr_par *p) {
  double *A = Q + 5;
  double U[9], s[3], Vm[9];
  orc_svd3(A, U, s, Vm);
  double detA = s[0] * s[1] * s[2];
  double r = detA * p->rho0;
  double d13 = cbrt(detA);
  double d23 = d13 * d13;
  double s0 = s[0] * s[0] / d23, s1 = s[1] * s[1] / d23, s2 = s[2] * s[2] / d23;
  double m0 = (s0 + s1 + s2) / 3.;
  double u0 = ((s0 - s1) * (s0 - s1) + (s1 - s2) * (s1 - s2) + (s2 - s0) * (s2 - s0)) / 3.;
  if (!(u0 >= 1e-12))
    return;
  double tau = nondim_time(r, d13, m0, u0, dt, p);
  double e6 = exp(-6. * tau), e9 = exp(-9. * tau);
  double a = 3. * m0 - u0 / 3. - 3.;
  double b = 2. * m0 - u0 / 3. - 2.;
  double m = 1. + a * e6 - b * e9;
  double u = posv(6. * a * e6 - 9. * b * e9);
  double D = -2. * m * m * m + m * u + 2.;
  double arg1 = posv(6. * u * u * u - 81. * D * D);
  double x1;
  if (fabs(D) < 1e-12)
    x1 = sqrt(6. * u) / 3. * cos(ORC_PI / 6.) + m;
  else if (arg1 == 0.) {
    double t = cbrt(54. * D);
    x1 = t / 6. + u / t + m;
  } else {
    double th = atan(sqrt(arg1) / (9. * D));
    x1 = sqrt(6. * u) / 3. * cos(th / 3.) + m;
  }
  double t2 = 3. * m - x1;
  double arg2 = posv(x1 * t2 * t2 - 4.);
  double x2 = 0.5 * (sqrt(arg2 / x1) + t2);
  double x3 = 1. / (x1 * x2);
  double x[3] = {x1, x2, x3}, tmp; /* ascending */
  if (x[1] < x[0]) { tmp = x[0]; x[0] = x[1]; x[1] = tmp; }
  if (x[2] < x[1]) { tmp = x[1]; x[1] = x[2]; x[2] = tmp; }
  if (x[1] < x[0]) { tmp = x[0]; x[0] = x[1]; x[1] = tmp; }
  double g[3];
  for (int i = 0; i < 3; i++)
    g[i] = d13 * sqrt(x[2 - i]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      A[i * 3 + j] = U[i * 3 + 0] * (g[0] * Vm[j * 3 + 0]) + U[i * 3 + 1] * (g[1] * Vm[j * 3 + 1]) +
                     U[i * 3 + 2] * (g[2] * Vm[j * 3 + 2]);
}

/* analytic.cpp:131-152 */
static void analytic_thermal(double *Q, double dt, const gpr_par *p) {
  double r = Q[0];
  double E = Q[1] / r;
  double v0 = Q[2] / r, v1 = Q[3] / r, v2 = Q[4] / r;
  double c1 = E - E_2A(r, Q + 5, p) - (v0 * v0 + v1 * v1 + v2 * v2) / 2;
  double c2 = p->calpha2 / 2.;
  double k = 2 * p->calpha2 / (p->kappa * r * p->cv);
  c1 *= k;
  c2 *= k;
  double ea = exp(-c1 * dt / 2);
  double J2 = Q[14] * Q[14] + Q[15] * Q[15] + Q[16] * Q[16];
  double den = 1 - c2 / c1 * (1 - ea * ea) * J2 / (r * r);
  for (int i = 14; i < 17; i++)
    Q[i] *= ea / sqrt(den);
}

void orc_analytic_solver(double *Q, long count, double dt, const gpr_par *p, int part,
                         const orc_cfg *c) {
  int V = c->thermal ? 17 : 14;
  for (long i = 0; i < count; i++) {
    if (part == 0)
      analytic_distortion(Q + i * V, dt, p);
    else if (c->thermal)
      analytic_thermal(Q + i * V, dt, p);
  }
}

/* ode.cpp:5-11 with ode_stepper_analytic (analytic.cpp:154-162); the
 * numerical (reaction) part does nothing for REACTION = -1 */
void orc_ode_launcher(double *u, long ncell, double dt, const gpr_par *p, const orc_cfg *c) {
  int V = c->thermal ? 17 : 14;
  for (long i = 0; i < ncell; i++) {
    double *Q = u + i * V;
    for (int k = 0; k < 3; k++)
      Q[2 + k] += p->dp[k] * dt;
    analytic_distortion(Q, dt, p);
    if (c->thermal)
      analytic_thermal(Q, dt, p);
  }
}

/* homogeneous.cpp:5-67 */
void orc_midstepper(double *wh, long ncell, double dt, const double *dX, const gpr_par *p,
                    const uint8_t *mask, const orc_cfg *c) {
  const consts_t *K = get_consts(c->N);
  int N = c->N, V = c->thermal ? 17 : 14, th = c->thermal;
  int ns = c->ndim == 1 ? N : N * N;
  double F[MAXN * MAXN][MAXV], G[MAXN * MAXN][MAXV];
  double dwx[MAXN * MAXN][MAXV], dwy[MAXN * MAXN][MAXV], dFx[MAXN * MAXN][MAXV], dGy[MAXN * MAXN][MAXV];
  for (long ind = 0; ind < ncell; ind++) {
    if (!mask[ind])
      continue;
    double *w = wh + ind * ns * V;
    for (int sN = 0; sN < ns; sN++) {
      memset(F[sN], 0, sizeof F[sN]);
      memset(G[sN], 0, sizeof G[sN]);
      orc_flux(F[sN], w + sN * V, 0, p, th);
      if (c->ndim == 2)
        orc_flux(G[sN], w + sN * V, 1, p, th);
    }
    for (int a = 0; a < (c->ndim == 1 ? N : N); a++)
      for (int b = 0; b < (c->ndim == 1 ? 1 : N); b++) {
        int sN = c->ndim == 1 ? a : a * N + b;
        for (int v = 0; v < V; v++) {
          double sx = 0, sfx = 0, sy = 0, sgy = 0;
          for (int k = 0; k < N; k++) {
            int nx_ = c->ndim == 1 ? k : k * N + b;
            sx += K->der[a][k] * w[nx_ * V + v];
            sfx += K->der[a][k] * F[nx_][v];
            if (c->ndim == 2) {
              int ny_ = a * N + k;
              sy += K->der[b][k] * w[ny_ * V + v];
              sgy += K->der[b][k] * G[ny_][v];
            }
          }
          dwx[sN][v] = sx;
          dFx[sN][v] = sfx;
          dwy[sN][v] = sy;
          dGy[sN][v] = sgy;
        }
      }
    for (int sN = 0; sN < ns; sN++) {
      double bx[MAXV], by[MAXV];
      orc_bdot(bx, w + sN * V, dwx[sN], 0, th);
      if (c->ndim == 2)
        orc_bdot(by, w + sN * V, dwy[sN], 1, th);
      for (int v = 0; v < V; v++) {
        double inc = (dFx[sN][v] + bx[v]) / dX[0];
        if (c->ndim == 2)
          inc += (dGy[sN][v] + by[v]) / dX[1];
        w[sN * V + v] -= dt / 2 * inc;
      }
    }
  }
}

/* One split step inside the time loop (iterator.cpp:110-165 with SPLIT,
 * steppers.cpp:50-66) */
int orc_split_step(double *u, const int32_t *nX, const double *dX, double CFL,
                   const int32_t *bcs, double *t, long *count, double tf,
                   double contorted_tol, int half_step, const gpr_par *p, const orc_cfg *c,
                   double *dt_out) {
  int N = c->N, V = c->thermal ? 17 : 14;
  long ncell = c->ndim == 1 ? nX[0] : (long)nX[0] * nX[1];
  long next = c->ndim == 1 ? nX[0] + 2 : (long)(nX[0] + 2) * (nX[1] + 2);
  long nb = c->ndim == 1 ? nX[0] + 2 * N : (long)(nX[0] + 2 * N) * (nX[1] + 2 * N);
  int ns = c->ndim == 1 ? N : N * N;
  double dt = orc_timestep(u, nX, dX, CFL, *t, tf, *count, p, c);
  double *ub = malloc(sizeof(double) * nb * V);
  double *wh = malloc(sizeof(double) * next * ns * V);
  uint8_t *m = malloc(ncell), *mb = malloc(next);
  memset(m, 1, ncell);
  orc_extend_mask(m, mb, nX, c);
  orc_boundaries(u, ub, nX, bcs, c);
  orc_ode_launcher(u, ncell, dt / 2, p, c);
  orc_ode_launcher(ub, nb, dt / 2, p, c);
  orc_weno_launcher(wh, ub, nX, c);
  if (half_step)
    orc_midstepper(wh, next, dt, dX, p, mb, c);
  orc_fv_launcher_ex(u, wh, nX, dt, dX, p, mb, 0, 0, c);
  orc_ode_launcher(u, ncell, dt / 2, p, c);
  *t += dt;
  *count += 1;
  orc_renorm_distortion(u, ncell, p, V);
  if (orc_contorted(u, ncell, contorted_tol, V))
    orc_reset_distortion(u, ncell, p, V);
  free(ub);
  free(wh);
  free(m);
  free(mb);
  if (dt_out)
    *dt_out = dt;
  for (long i = 0; i < ncell * V; i++)
    if (u[i] != u[i])
      return -1;
  return 0;
}
