// Roe and Osher Riemann solvers (FLUX = 1, 2; solvers/fv/fluxes.cpp:22-64)
// on the device, and the system matrix they decompose
// (system/equations.cpp:169-176, system/jacobians.cpp:10-152).
//
// Both take |M| = R |Lambda| R^-1 of a V x V system matrix (14 or 17 rows):
// Roe of the path average of M over the N Gauss nodes of the straight path
// qR + x (qL - qR), Osher the Gauss sum of |M| at those nodes.  The
// decomposition restates the steps the reference takes through Eigen's
// EigenSolver and a pivoted solve: Householder reduction to Hessenberg
// form, complex single-shift QR to a triangular Schur form (Wilkinson
// shifts, an exceptional shift every 10th iteration), eigenvectors of the
// triangle by back substitution, R = Z U, and R b = dq by Gaussian
// elimination with partial pivoting.  One lane solves one problem.  Its V x V
// matrices live in a device workspace, not in private memory: element k of
// a lane's array is at base[k * stride + lane] (LaneMem), so the lanes of a
// wavefront touch consecutive words, and a kernel that uses it runs a
// bounded grid over its work so the workspace is (grid lanes) x
// EIG_WS_DOUBLES(V) doubles.  Private memory of this size per lane would
// make the HIP runtime reserve it for every wavefront slot of the device.
#pragma once

#include "gpr_device.hpp"

namespace gpr {

// A lane's strided view of a workspace array: element k at p[k * stride].
struct LaneMem {
  double *p;
  long stride;
  __device__ __forceinline__ double &operator[](long k) const { return p[k * stride]; }
  __device__ __forceinline__ LaneMem at(long k) const { return {p + k * stride, stride}; }
};
// workspace doubles per lane: M, Mbar, dF/dP, dP/dQ, then H and Z (real and
// imaginary parts)
constexpr long EIG_WS_DOUBLES(int V) { return 8L * V * V; }

// ------------------------------------------------- EOS derivatives (mg.cpp, shear.cpp)
__device__ __forceinline__ double ef_dedrho(double r, double pp, const gpr_par &p) {
  const double G = Gamma_MG<double>(r, p), dG = dGamma_MG<double>(r, p), pr = p_ref<double>(r, p),
               dpr = dp_ref<double>(r, p), der = de_ref<double>(r, p);
  return der - (dpr * r * G + (G + r * dG) * (pp - pr)) / ((r * G) * (r * G));
}
__device__ __forceinline__ double ef_dedp(double r, const gpr_par &p) { return 1 / (r * Gamma_MG<double>(r, p)); }
__device__ __forceinline__ double ef_dTdrho(double r, double pp, const gpr_par &p) {
  const double G = Gamma_MG<double>(r, p), dG = dGamma_MG<double>(r, p), pr = p_ref<double>(r, p),
               dpr = dp_ref<double>(r, p);
  return -(dpr * r * G + (G + r * dG) * (pp - pr)) / ((r * G) * (r * G)) / p.cv;
}
__device__ __forceinline__ double ef_cs2(double r, const gpr_par &p) { return p.b02 * pow(r / p.rho0, p.beta); }
__device__ __forceinline__ double ef_C0(double r, const gpr_par &p) {
  return (1 - p.beta / Gamma_MG<double>(r, p)) * ef_cs2(r, p);
}
__device__ __forceinline__ double ef_dC0drho(double r, const gpr_par &p) {
  const double cs2 = ef_cs2(r, p), dcs2 = p.beta / r * cs2, G = Gamma_MG<double>(r, p),
               dG = dGamma_MG<double>(r, p);
  return (1 - p.beta / G) * dcs2 + p.beta / (G * G) * dG * cs2;
}
// A dev(A^T A), A row-major (functions/matrices.cpp:8-41)
__device__ __forceinline__ void ef_AdevG(const double *A, double *out) {
  double G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = A[0 * 3 + i] * A[0 * 3 + j] + A[1 * 3 + i] * A[1 * 3 + j] + A[2 * 3 + i] * A[2 * 3 + j];
  const double x = (G[0] + G[4] + G[8]) / 3;
  G[0] -= x;
  G[4] -= x;
  G[8] -= x;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = A[i * 3 + 0] * G[0 * 3 + j] + A[i * 3 + 1] * G[1 * 3 + j] + A[i * 3 + 2] * G[2 * 3 + j];
}
// dsigma_ij / dA_mn at constant rho (state.cpp:71-84); G = A^T A, AdG = A dev(G)
__device__ __forceinline__ double ef_dsig_dA(double r, double cs2, const double *A, const double *G,
                                             const double *AdG, int i, int j, int m, int n) {
  double t = A[m * 3 + i] * G[j * 3 + n] + A[m * 3 + j] * G[i * 3 + n] - 2. / 3. * G[i * 3 + j] * A[m * 3 + n];
  if (i == n) t += AdG[m * 3 + j];
  if (j == n) t += AdG[m * 3 + i];
  return -r * cs2 * t;
}

// dF_d/dP with P = (rho, p, v, A, J) (jacobians.cpp:10-106), V x V row-major
template <bool TH>
__device__ __noinline__ void ef_dFdP(const double *Q, int d, const Mat &m, LaneMem M) {
  constexpr int V = Dims<TH>::V;
  const gpr_par &p = m.p;
  for (int i = 0; i < V * V; i++) M[i] = 0.0;
#define MM(i, j) M[(i) * V + (j)]
  NodeState<double, TH> s;
  node_state<double, TH>(Q, m, s);
  const double r = Q[0], E = Q[1] / r, pp = s.p;
  const double v[3] = {Q[2] / r, Q[3] / r, Q[4] / r};
  const double *A = Q + 5;
  const double Er = ef_dedrho(r, pp, p) + ef_dC0drho(r, p) / 4 * s.devGsq, Ep = ef_dedp(r, p);
  const double vd = v[d], rvd = r * vd;
  MM(0, 0) = vd;
  MM(0, 2 + d) = r;
  MM(1, 0) = (E + r * Er) * vd;
  MM(1, 1) = (r * Ep + 1) * vd;
  for (int k = 0; k < 3; k++) MM(1, 2 + k) = r * vd * v[k];
  MM(1, 2 + d) += r * E + pp;
  for (int k = 0; k < 3; k++) {
    MM(2 + k, 0) = vd * v[k];
    MM(2 + k, 2 + k) = rvd;
  }
  for (int k = 0; k < 3; k++) MM(2 + k, 2 + d) += r * v[k];
  MM(2 + d, 1) = 1.0;
  // viscous part: the stress column sigma_{.d}, its rho derivative at
  // constant A, and dE/dA = C_0 A dev(G)
  double sd[3], sr[3], G[9], AdG[9], psi[9];
  const double cs2 = ef_cs2(r, p), C0 = ef_C0(r, p);
  ef_AdevG(A, AdG);
  // sigma_col as the reference forms it: -rho (cs2 AdG)^T A, column d
  for (int i = 0; i < 3; i++)
    sd[i] = -r * ((cs2 * AdG[0 * 3 + i]) * A[0 * 3 + d] + (cs2 * AdG[1 * 3 + i]) * A[1 * 3 + d] +
                  (cs2 * AdG[2 * 3 + i]) * A[2 * 3 + d]);
  const double dcs2 = p.beta / r * cs2;
  for (int k = 0; k < 3; k++) sr[k] = (1 / r + dcs2 / cs2) * sd[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  for (int k = 0; k < 9; k++) psi[k] = C0 * AdG[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double t = 0.;
      for (int k = 0; k < 3; k++) {
        const double ds = ef_dsig_dA(r, cs2, A, G, AdG, d, k, i, j);
        MM(2 + k, 5 + i * 3 + j) = -ds;
        t -= v[k] * ds;
      }
      MM(1, 5 + i * 3 + j) = t;
    }
  MM(1, 0) -= sr[0] * v[0] + sr[1] * v[1] + sr[2] * v[2];
  for (int k = 0; k < 3; k++) MM(1, 2 + k) -= sd[k];
  for (int k = 0; k < 9; k++) MM(1, 5 + k) += r * vd * psi[k];
  for (int k = 0; k < 3; k++) MM(2 + k, 0) -= sr[k];
  for (int row = 0; row < 3; row++)
    for (int k = 0; k < 3; k++) {
      MM(5 + 3 * row + d, 2 + k) = A[row * 3 + k];
      MM(5 + 3 * row + d, 5 + 3 * row + k) = v[k];
    }
  if (TH) {
    const double G0 = Gamma_MG<double>(r, p), pr = p_ref<double>(r, p);
    double T = (pp - pr) / (r * G0 * p.cv);
    if (p.Tref != 0.0) T += phi_int<double>(r, p) * p.Tref;
    const double Tr = ef_dTdrho(r, pp, p), Tp = ef_dedp(r, p) / p.cv;
    const double J[3] = {Q[14] / r, Q[15] / r, Q[16] / r};
    const double H[3] = {p.calpha2 * Q[14] / r, p.calpha2 * Q[15] / r, p.calpha2 * Q[16] / r};
    MM(1, 0) += Tr * H[d];
    MM(1, 1) += Tp * H[d];
    for (int k = 0; k < 3; k++) MM(1, 14 + k) = rvd * H[k];
    MM(1, 14 + d) += p.calpha2 * T;
    for (int k = 0; k < 3; k++) MM(14 + k, 0) = vd * J[k];
    MM(14 + d, 0) += Tr;
    MM(14 + d, 1) = Tp;
    for (int k = 0; k < 3; k++) {
      MM(14 + k, 2 + d) = r * J[k];
      MM(14 + k, 14 + k) = rvd;
    }
  }
#undef MM
}

// dP/dQ (jacobians.cpp:108-152), V x V row-major
template <bool TH>
__device__ __noinline__ void ef_dPdQ(const double *Q, const Mat &m, LaneMem M) {
  constexpr int V = Dims<TH>::V;
  const gpr_par &p = m.p;
  for (int i = 0; i < V * V; i++) M[i] = i % (V + 1) == 0 ? 1.0 : 0.0;
#define MM(i, j) M[(i) * V + (j)]
  NodeState<double, TH> s;
  node_state<double, TH>(Q, m, s);
  const double r = Q[0], E = Q[1] / r, pp = s.p;
  const double v[3] = {Q[2] / r, Q[3] / r, Q[4] / r};
  const double Er = ef_dedrho(r, pp, p) + ef_dC0drho(r, p) / 4 * s.devGsq, Ep = ef_dedp(r, p);
  const double Gi = 1 / (r * Ep);
  double t = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - (E + r * Er);
  double J[3] = {0, 0, 0};
  if (TH) {
    for (int k = 0; k < 3; k++) J[k] = Q[14 + k] / r;
    t += p.calpha2 * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
  }
  MM(1, 0) = Gi * t;
  MM(1, 1) = Gi;
  for (int k = 0; k < 3; k++) {
    MM(1, 2 + k) = -Gi * v[k];
    MM(2 + k, 0) = -v[k] / r;
    MM(2 + k, 2 + k) = 1 / r;
  }
  double AdG[9];
  ef_AdevG(Q + 5, AdG);
  const double C0 = ef_C0(r, p);
  for (int k = 0; k < 9; k++) MM(1, 5 + k) = -Gi * r * (C0 * AdG[k]);
  if (TH)
    for (int k = 0; k < 3; k++) {
      MM(1, 14 + k) = -Gi * (p.calpha2 * Q[14 + k] / r);
      MM(14 + k, 0) = -J[k] / r;
      MM(14 + k, 14 + k) = 1 / r;
    }
#undef MM
}

// M_d = dF_d/dP dP/dQ + B_d (equations.cpp:169-176, block() :69-87) into ret;
// F and P: workspace for dF/dP and dP/dQ
template <bool TH>
__device__ __noinline__ void system_matrix(const double *Q, int d, const Mat &m, LaneMem ret, LaneMem F, LaneMem P) {
  constexpr int V = Dims<TH>::V;
  ef_dFdP<TH>(Q, d, m, F);
  ef_dPdQ<TH>(Q, m, P);
  for (int i = 0; i < V; i++)
    for (int j = 0; j < V; j++) {
      double s = 0;
      for (int k = 0; k < V; k++) s += F[i * V + k] * P[k * V + j];
      ret[i * V + j] = s;
    }
  const double r = Q[0], vd = Q[2 + d] / r;
  for (int i = 5; i < 14; i++) ret[i * V + i] += vd;
  for (int row = 0; row < 3; row++)
    for (int k = 0; k < 3; k++) ret[(5 + 3 * row + d) * V + 5 + 3 * row + k] -= Q[2 + k] / r;
}

// ------------------------------------------------------------ complex helpers
__device__ __forceinline__ cplx cmk(double re, double im = 0.0) { return {re, im}; }
__device__ __forceinline__ cplx cneg(cplx a) { return {-a.re, -a.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double r) { return {a.re * r, a.im * r}; }
// a / b by Smith's scaling (what C99 complex division does for finite operands)
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, den = b.re + b.im * t;
    return {(a.re + a.im * t) / den, (a.im - a.re * t) / den};
  }
  const double t = b.re / b.im, den = b.re * t + b.im;
  return {(a.re * t + a.im) / den, (a.im * t - a.re) / den};
}

// out = Re(R |Lambda| R^-1 x) for a real n x n row-major J (n = V, in a
// workspace view).  w holds 6 n^2 doubles: A (the copy of R the elimination
// overwrites), H (Hessenberg, then triangular T with U, the unit upper
// eigenvector matrix, transposed into its strictly lower triangle) and Z
// (Schur vectors, then R = Z U in place), each as real and imaginary parts;
// A's parts first hold the real Hessenberg reduction's H and Q.  Returns 0,
// or -1 if QR does not converge in 60 n iterations (out is then NaN).
template <int n>
__device__ __noinline__ int eig_abs_apply(LaneMem J, const double *x, double *out, LaneMem w) {
  constexpr long nn = (long)n * n;
  constexpr long AB = 0, HB = 2 * nn, ZB = 4 * nn;  // re at B, im at B + nn
  auto ld = [&](long B, int i, int j) -> cplx { return {w[B + i * n + j], w[B + nn + i * n + j]}; };
  auto st = [&](long B, int i, int j, cplx v) {
    w[B + i * n + j] = v.re;
    w[B + nn + i * n + j] = v.im;
  };
  const double eps = 2.220446049250313e-16;
  // real Householder reduction in A's two halves: Hr = A.re, Qr = A.im
  auto Hr = [&](int i, int j) -> double & { return w[AB + i * n + j]; };
  auto Qr = [&](int i, int j) -> double & { return w[AB + nn + i * n + j]; };
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      Hr(i, j) = J[i * n + j];
      Qr(i, j) = i == j;
    }
  for (int k = 0; k < n - 2; k++) {
    double v[n], nrm = 0, tail = 0;
    for (int i = k + 1; i < n; i++) {
      v[i] = Hr(i, k);
      nrm += v[i] * v[i];
      if (i > k + 1) tail += v[i] * v[i];
    }
    if (tail == 0.0) continue;
    nrm = sqrt(nrm);
    const double alpha = v[k + 1] >= 0 ? -nrm : nrm;
    v[k + 1] -= alpha;
    double vv = 0;
    for (int i = k + 1; i < n; i++) vv += v[i] * v[i];
    const double beta = 2.0 / vv;
    for (int j = 0; j < n; j++) {  // H <- P H
      double s = 0;
      for (int i = k + 1; i < n; i++) s += v[i] * Hr(i, j);
      s *= beta;
      for (int i = k + 1; i < n; i++) Hr(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; i++) {  // H <- H P, Q <- Q P
      double s = 0, t = 0;
      for (int j = k + 1; j < n; j++) {
        s += Hr(i, j) * v[j];
        t += Qr(i, j) * v[j];
      }
      s *= beta;
      t *= beta;
      for (int j = k + 1; j < n; j++) {
        Hr(i, j) -= s * v[j];
        Qr(i, j) -= t * v[j];
      }
    }
    Hr(k + 1, k) = alpha;
    for (int i = k + 2; i < n; i++) Hr(i, k) = 0.0;
  }
  double norm = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double h = Hr(i, j);
      st(HB, i, j, cmk(h));
      st(ZB, i, j, cmk(Qr(i, j)));
      norm += fabs(h);
    }
  if (norm == 0.0) {
    for (int i = 0; i < n; i++) out[i] = 0.0;
    return 0;
  }
  // complex QR on the active window [l, hi], the bulge chased by rotations
  int hi = n - 1, its = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; l--) {
      double s = cabs_(ld(HB, l - 1, l - 1)) + cabs_(ld(HB, l, l));
      if (s == 0.0) s = norm;
      if (cabs_(ld(HB, l, l - 1)) <= eps * s) {
        st(HB, l, l - 1, cmk(0.0));
        break;
      }
    }
    if (l == hi) {
      hi--;
      its = 0;
      continue;
    }
    if (++total > 60 * n) {
      for (int i = 0; i < n; i++) out[i] = __builtin_nan("");
      return -1;
    }
    its++;
    cplx mu;
    if (its % 10 == 0) {
      mu = cadd(ld(HB, hi, hi), cmk(0.75 * fabs(ld(HB, hi, hi - 1).re)));
    } else {
      const cplx a = ld(HB, hi - 1, hi - 1), bb = ld(HB, hi - 1, hi), c = ld(HB, hi, hi - 1), dd = ld(HB, hi, hi);
      const cplx tr2 = cscale(cadd(a, dd), 0.5);
      const cplx amd = csub(a, dd);
      const cplx disc = csqrt_(cadd(cscale(cmul(amd, amd), 0.25), cmul(bb, c)));
      const cplx l1 = cadd(tr2, disc), l2 = csub(tr2, disc);
      mu = cabs_(csub(l1, dd)) <= cabs_(csub(l2, dd)) ? l1 : l2;
    }
    for (int k = l; k < hi; k++) {
      const cplx xx = k == l ? csub(ld(HB, l, l), mu) : ld(HB, k, k - 1);
      const cplx yy = k == l ? ld(HB, l + 1, l) : ld(HB, k + 1, k - 1);
      const double rr = sqrt(cabs2(xx) + cabs2(yy));
      if (rr == 0.0) continue;
      const cplx c = cdivr(xx, rr), sn = cdivr(yy, rr);
      if (k > l) {
        st(HB, k, k - 1, cmk(rr));
        st(HB, k + 1, k - 1, cmk(0.0));
      }
      for (int j = k; j < n; j++) {  // rows k, k+1 <- G rows
        const cplx h0 = ld(HB, k, j), h1 = ld(HB, k + 1, j);
        st(HB, k, j, cadd(cmul(cconj(c), h0), cmul(cconj(sn), h1)));
        st(HB, k + 1, j, cadd(cmul(cneg(sn), h0), cmul(c, h1)));
      }
      const int top = k + 2 < hi ? k + 2 : hi;
      for (int i = 0; i <= top; i++) {  // columns k, k+1 <- columns G^H
        const cplx h0 = ld(HB, i, k), h1 = ld(HB, i, k + 1);
        st(HB, i, k, cadd(cmul(h0, c), cmul(h1, sn)));
        st(HB, i, k + 1, cadd(cmul(cneg(h0), cconj(sn)), cmul(h1, cconj(c))));
      }
      for (int i = 0; i < n; i++) {
        const cplx z0 = ld(ZB, i, k), z1 = ld(ZB, i, k + 1);
        st(ZB, i, k, cadd(cmul(z0, c), cmul(z1, sn)));
        st(ZB, i, k + 1, cadd(cmul(cneg(z0), cconj(sn)), cmul(z1, cconj(c))));
      }
    }
  }
  // eigenvectors of the triangular T = H: U[j][k] (j < k) kept at H[k][j]
  double tn = 0;
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++) tn += cabs_(ld(HB, i, j));
  const double small = eps * tn;
  for (int k = 0; k < n; k++) {
    const cplx hkk = ld(HB, k, k);
    for (int j = k - 1; j >= 0; j--) {
      cplx s = cmk(0.0);
      for (int m = j + 1; m <= k; m++) s = cadd(s, cmul(ld(HB, j, m), m == k ? cmk(1.0) : ld(HB, k, m)));
      cplx den = csub(ld(HB, j, j), hkk);
      if (cabs_(den) < small) den = cmk(small);
      st(HB, k, j, cneg(cdiv(s, den)));
    }
  }
  // R = Z U in place, column by column from the last (column j needs Z's
  // columns m <= j only)
  for (int j = n - 1; j >= 0; j--)
    for (int i = 0; i < n; i++) {
      cplx s = cmk(0.0);
      for (int m = 0; m <= j; m++) s = cadd(s, cmul(ld(ZB, i, m), m == j ? cmk(1.0) : ld(HB, j, m)));
      st(ZB, i, j, s);
    }
  // R b = x with partial pivoting on the copy A of R
  cplx b[n];
  for (int i = 0; i < n; i++) {
    b[i] = cmk(x[i]);
    for (int j = 0; j < n; j++) st(AB, i, j, ld(ZB, i, j));
  }
  for (int k = 0; k < n; k++) {
    int piv = k;
    double best = cabs_(ld(AB, k, k));
    for (int i = k + 1; i < n; i++) {
      const double a = cabs_(ld(AB, i, k));
      if (a > best) {
        best = a;
        piv = i;
      }
    }
    if (piv != k) {
      for (int j = 0; j < n; j++) {
        const cplx t = ld(AB, k, j);
        st(AB, k, j, ld(AB, piv, j));
        st(AB, piv, j, t);
      }
      const cplx t = b[k];
      b[k] = b[piv];
      b[piv] = t;
    }
    const cplx akk = ld(AB, k, k);
    for (int i = k + 1; i < n; i++) {
      const cplx f = cdiv(ld(AB, i, k), akk);
      for (int j = k; j < n; j++) st(AB, i, j, csub(ld(AB, i, j), cmul(f, ld(AB, k, j))));
      b[i] = csub(b[i], cmul(f, b[k]));
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    cplx s = b[i];
    for (int j = i + 1; j < n; j++) s = csub(s, cmul(ld(AB, i, j), b[j]));
    b[i] = cdiv(s, ld(AB, i, i));
  }
  for (int i = 0; i < n; i++) b[i] = cscale(b[i], cabs_(ld(HB, i, i)));
  for (int i = 0; i < n; i++) {
    cplx s = cmk(0.0);
    for (int j = 0; j < n; j++) s = cadd(s, cmul(ld(ZB, i, j), b[j]));
    out[i] = s.re;
  }
  return 0;
}

// D_ROE (fluxes.cpp:45-64): |M(qbar)| (qL - qR) + F(qL) + F(qR), M(qbar) the
// Gauss average of M over the straight path; D_OSH (fluxes.cpp:22-43): the
// Gauss sum of |M| (qL - qR) on the path + F(qL) + F(qR).  which: 1 Roe,
// 2 Osher.  ws: EIG_WS_DOUBLES(V) doubles of workspace (M, Mbar, then the
// system matrix's dF/dP and dP/dQ, which eig_abs_apply's A reuses, H, Z).
template <int N, bool TH>
__device__ __noinline__ int eig_flux(const double *qL, const double *qR, int d, const Mat &m, const Basis &K,
                                     int which, double *ret, LaneMem ws) {
  constexpr int V = Dims<TH>::V;
  constexpr long nn = (long)V * V;
  const LaneMem M = ws, Mb = ws.at(nn), F = ws.at(2 * nn), P = ws.at(3 * nn), W = ws.at(2 * nn);
  double dq[V], q[V];
  int rc = 0;
  for (int i = 0; i < V; i++) dq[i] = qL[i] - qR[i];
  if (which == 1) {
    for (long i = 0; i < nn; i++) Mb[i] = 0.0;
    for (int s = 0; s < N; s++) {
      for (int i = 0; i < V; i++) q[i] = qR[i] + K.x[s] * dq[i];
      system_matrix<TH>(q, d, m, M, F, P);
      for (long i = 0; i < nn; i++) Mb[i] += K.w[s] * M[i];
    }
    rc = eig_abs_apply<V>(Mb, dq, ret, W);
  } else {
    double t[V];
    for (int i = 0; i < V; i++) ret[i] = 0.0;
    for (int s = 0; s < N; s++) {
      for (int i = 0; i < V; i++) q[i] = qR[i] + K.x[s] * dq[i];
      system_matrix<TH>(q, d, m, M, F, P);
      rc |= eig_abs_apply<V>(M, dq, t, W);
      for (int i = 0; i < V; i++) ret[i] += K.w[s] * t[i];
    }
  }
  double Fx[V];
  {
    NodeState<double, TH> s;
    node_state<double, TH>(qL, m, s);
    flux<double, TH>(qL, s, m, d, Fx);
    for (int i = 0; i < V; i++) ret[i] += Fx[i];
  }
  {
    NodeState<double, TH> s;
    node_state<double, TH>(qR, m, s);
    flux<double, TH>(qR, s, m, d, Fx);
    for (int i = 0; i < V; i++) ret[i] += Fx[i];
  }
  return rc;
}

}  // namespace gpr
