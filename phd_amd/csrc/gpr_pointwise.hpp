// Pointwise numerics shared by the kernels: the WENO reconstruction of one
// line of cell averages, the Rusanov flux and the path-conservative jump of
// one face, and the stiff fallback's preconditioner of one spatial node and
// its projection.  Templated on the number type R; with GPR_HOST_ONLY
// defined (gpr_device.hpp) they also compile for the host, where
// flopcount.cpp evaluates them with a counting number type.
#pragma once

#include <type_traits>

#include "gpr_device.hpp"

namespace gpr {

// ---------------------------------------------------------------------- WENO
// Reconstruction of one variable over one stencil line (coeffs, weno.cpp:17-73)
// in difference form about the central value c: w_s = c + M_s^-1 (d - c), whose
// oscillation indicator and weighted mean are those of the direct form exactly
// (the basis sums to one, so M_s^-1 maps constants to themselves and SIG
// annihilates them) but do not cancel when |c| >> |d - c|.
template <typename R, int N>
__device__ __forceinline__ void weno_point(const R *d, const Basis &K, int ncell, int gap, int ind, R *out) {
  constexpr int FN2 = (N - 1) / 2;
  R c = d[N - 1];
  R dd[2 * N - 1];
#pragma unroll
  for (int r = 0; r < 2 * N - 1; r++) dd[r] = d[r] - c;
  R ws[3][N];
#pragma unroll
  for (int a = 0; a < N; a++) {
    R sl = 0, sr = 0, sc = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      sl += R(K.mLinv[a][k]) * dd[k];
      sr += R(K.mRinv[a][k]) * dd[N - 1 + k];
      sc += R(K.mCLinv[a][k]) * dd[FN2 + k];
    }
    ws[0][a] = sl;
    ws[1][a] = sr;
    ws[2][a] = sc;
  }
  constexpr int NSTEN = N > 2 ? 3 : 2;
  bool useL = ind >= gap, useR = ind <= ncell - gap - 1, useC = N > 2 && ind + FN2 >= gap;
  R o[3], tt[3];
#pragma unroll
  for (int s = 0; s < NSTEN; s++) {
    R qv = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
      for (int j = 0; j < N; j++) qv += ws[s][i] * R(K.sig[i][j]) * ws[s][j];
    tt[s] = qv + R(1e-14);
  }
  if constexpr (!std::is_same<R, float>::value) {
#pragma unroll
    for (int s = 0; s < NSTEN; s++) {
      R t2 = tt[s] * tt[s], t4 = t2 * t2;
      o[s] = R(s == 2 ? 1e5 : 1.0) / (t4 * t4);
    }
  } else {
    // lambda_s / t_s^8 scaled by the smallest t^8 among the stencils in use:
    // the normalised weights are unchanged, and t^8 (down to 1e-112) would
    // underflow in single precision
    R tmin = R(3e38);
    if (useL) tmin = fmin(tmin, tt[0]);
    if (useR) tmin = fmin(tmin, tt[1]);
    if (useC) tmin = fmin(tmin, tt[NSTEN - 1]);
#pragma unroll
    for (int s = 0; s < NSTEN; s++) {
      R r = tmin / tt[s], r2 = r * r, r4 = r2 * r2;
      o[s] = R(s == 2 ? 1e5 : 1.0) * (r4 * r4);
    }
  }
  R num[N], den = 0;
#pragma unroll
  for (int a = 0; a < N; a++) num[a] = 0;
  if (useL) {
#pragma unroll
    for (int a = 0; a < N; a++) num[a] += ws[0][a] * o[0];
    den += o[0];
  }
  if (useR) {
#pragma unroll
    for (int a = 0; a < N; a++) num[a] += ws[1][a] * o[1];
    den += o[1];
  }
  if (useC) {
#pragma unroll
    for (int a = 0; a < N; a++) num[a] += ws[2][a] * o[2];
    den += o[2];
  }
#pragma unroll
  for (int a = 0; a < N; a++) out[a] = c + num[a] / den;
}

// ------------------------------------------------------------------------ FV
// Rusanov flux D_RUS (fluxes.cpp:66-75) and the path integral Bint
// (fluxes.cpp:9-20) between two states, accumulated as
//   lo += k (b + f),  hi += k (b - f).
// Each side's node state serves both its wave-speed bound and its flux.
template <bool TH, typename R = double, int MK = MK_ANY>
__device__ __forceinline__ void rusanov(const R *qL, const R *qR, int d, const Mat &m, R *f) {
  constexpr int V = Dims<TH>::V;
  R FL[V], FR[V], sL, sR;
  {
    NodeState<R, TH> st;
    node_state<R, TH, MK>(qL, m, st);
    sL = max_abs_eigs_of<TH, R, MK>(qL, st, m, d);
    flux<R, TH>(qL, st, m, d, FL);
  }
  {
    NodeState<R, TH> st;
    node_state<R, TH, MK>(qR, m, st);
    sR = max_abs_eigs_of<TH, R, MK>(qR, st, m, d);
    flux<R, TH>(qR, st, m, d, FR);
  }
  R smax = (sL < sR) ? sR : sL;  // std::max(max1, max2): a NaN on the right is dropped
#pragma unroll
  for (int v = 0; v < V; v++) f[v] = smax * (qL[v] - qR[v]);
#pragma unroll
  for (int v = 0; v < V; v++) f[v] += FL[v];
#pragma unroll
  for (int v = 0; v < V; v++) f[v] += FR[v];
}

template <int N, bool TH, typename R = double>
__device__ __forceinline__ void bint(const R *qL, const R *qR, int d, const Basis &K, R *b) {
  constexpr int V = Dims<TH>::V;
#pragma unroll
  for (int v = 0; v < V; v++) b[v] = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    R xi = K.x[i];
    R r = qL[0] + xi * (qR[0] - qL[0]);
    R vv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) vv[c] = (qL[2 + c] + xi * (qR[2 + c] - qL[2 + c])) / r;
    R dq[V];
#pragma unroll
    for (int v = 0; v < V; v++) dq[v] = qR[v] - qL[v];
    bdot_sub(vv, dq, d, R(-K.w[i]), b);
  }
}

template <int N, bool TH, typename R = double, int MK = MK_ANY>
__device__ __forceinline__ void riemann_accum(const R *qL, const R *qR, int d, const Mat &m,
                                              const Basis &K, double k, R *lo, R *hi) {
  constexpr int V = Dims<TH>::V;
  R f[V], b[V];
  rusanov<TH, R, MK>(qL, qR, d, m, f);
  bint<N, TH, R>(qL, qR, d, K, b);
#pragma unroll
  for (int v = 0; v < V; v++) {
    lo[v] += k * (b[v] + f[v]);
    hi[v] += k * (b[v] - f[v]);
  }
}

// Preconditioner of one spatial node, frozen at its WENO state ws:
// out[0..8] = (DG_MAT + dt kappa W)^-1, out[9..17] = polar rotation of A,
// out[18..26] = (DG_MAT + dt rJ W)^-1 (thermal only).  At a relaxed state
// A = s Rot the distortion source has Jacobian -kappa Pi with
// The FV cell integrand of one space-time node (centers1/2, fv.cpp:11-109):
// wgt (S(q) - B_x dq/dx - B_y dq/dy), wgt = dt W_t W_a W_b (TIME) or
// dt W_a W_b, with dq from the cell's nodes in myQ ([node][V], node =
// t NS + a N + b).  SRC = false drops S(q).  The FV cell kernel and the
// predictor's fused epilogue both evaluate it here, so they round alike.
template <int N, bool TH, int ND, bool TIME, bool SRC, int MK = MK_ANY>
__device__ __forceinline__ void cell_integrand(const double *q, const double *myQ, int t, int ia, int ib,
                                               const Mat &m, const Basis &K, double dt, double dx, double dy,
                                               double *S) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = ND == 1 ? N : N * N;
  NodeState<double, TH> st;
  node_state<double, TH, MK>(q, m, st);
  if (SRC) {
    source<double, TH, MK>(q, st, m, S);
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) S[v] = 0.0;
  }
  double dq[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    double acc = 0;
#pragma unroll
    for (int k = 0; k < N; k++) acc += K.der[ia][k] * myQ[(t * NS + (ND == 1 ? k : k * N + ib)) * V + v];
    dq[v] = acc / dx;
  }
  bdot_sub(st.v, dq, 0, 1.0, S);
  if (ND == 2) {
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += K.der[ib][k] * myQ[(t * NS + ia * N + k) * V + v];
      dq[v] = acc / dy;
    }
    bdot_sub(st.v, dq, 1, 1.0, S);
  }
  const double wt = TIME ? dt * K.w[t] : dt;
  const double wgt = ND == 2 ? wt * K.w[ia] * K.w[ib] : wt * K.w[ia];
#pragma unroll
  for (int v = 0; v < V; v++) S[v] = wgt * S[v];
}

// Face states of the FV corrector's fused form, per extended cell:
// [side][t][face node][V], sides 0 / 1 the x-faces at x = 0 / 1 of the
// cell, 2 / 3 the y-faces at y = 0 / 1; face node b on x-sides, a on
// y-sides.  Each value is sum_k psi_k(end) q(..., k, ...) along the normal,
// in the order k = 0..N-1 that the face kernel's own extrapolation uses.
template <int N, int ND, bool TIME = true> struct FaceStates {
  static constexpr int NSF = ND == 2 ? N : 1;        // face nodes
  static constexpr int NPS = (TIME ? N : 1) * NSF;  // nodes per side
  static constexpr int SIDES = 2 * ND;
};

// Pi(X) = Rot dev(sym(Rot^T X)), kappa = 2 theta1inv cs^2 det(A)^(2/3);
// the thermal source -(c_alpha^2 T / (kappa rho)) rho J has rate rJ.
template <int N, bool TH, int MK = MK_ANY, typename R = double>
__device__ __forceinline__ void stiff_precond(const R *ws, const Mat &m, R dt, const Basis &K, R *out) {
  NodeState<R, TH> s;
  node_state<R, TH, MK>(ws, m, s);
  R dA = det3(ws + 5);
  R s13 = cbrt(dA);
  R kap = R(2.0) * theta1inv<R, TH, MK>(ws, s, m) * s.cs2 * s13 * s13;
  if (!(kap >= 0.0) || !(kap < 1e300)) kap = 0.0;
  R rJ = 0.0;
  if (TH) {
    rJ = m.p.calpha2 * s.T / (m.p.kappa * s.r);
    if (!(rJ >= 0.0) || !(rJ < 1e300)) rJ = 0.0;
  }
#pragma unroll
  for (int which = 0; which < (TH ? 2 : 1); which++) {
    R rate = which == 0 ? kap : rJ;
    // Gauss-Jordan with partial pivoting.  Every index is a compile-time
    // constant (the pivot row is swapped in by predicated exchanges), so the
    // matrix stays in registers rather than in scratch memory.
    R M[N][2 * N];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
      for (int j = 0; j < 2 * N; j++)
        M[i][j] = j < N ? R(K.dgmat[i][j]) + (i == j ? dt * rate * R(K.w[i]) : R(0.0)) : R(j - N == i ? 1.0 : 0.0);
#pragma unroll
    for (int c = 0; c < N; c++) {
      int piv = c;
      R best = fabs(M[c][c]);
#pragma unroll
      for (int r = c + 1; r < N; r++)
        if (fabs(M[r][c]) > best) {
          best = fabs(M[r][c]);
          piv = r;
        }
#pragma unroll
      for (int r = c + 1; r < N; r++)
        if (piv == r)
#pragma unroll
          for (int j = 0; j < 2 * N; j++) {
            R tt = M[c][j];
            M[c][j] = M[r][j];
            M[r][j] = tt;
          }
      R dg = M[c][c];
#pragma unroll
      for (int j = 0; j < 2 * N; j++) M[c][j] /= dg;
#pragma unroll
      for (int r = 0; r < N; r++) {
        if (r == c) continue;
        R f = M[r][c];
#pragma unroll
        for (int j = 0; j < 2 * N; j++) M[r][j] -= f * M[c][j];
      }
    }
    R *dst = which == 0 ? out : out + 18;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
      for (int j = 0; j < N; j++) dst[i * N + j] = M[i][N + j];
  }
  R X[9];
  for (int i = 0; i < 9; i++) X[i] = ws[5 + i];
  if (dA > 0) {
    for (int it = 0; it < 40; it++) {  // X <- (X + X^-T) / 2
      R d = det3(X), inv[9];
      inv[0] = (X[4] * X[8] - X[5] * X[7]) / d;
      inv[1] = (X[2] * X[7] - X[1] * X[8]) / d;
      inv[2] = (X[1] * X[5] - X[2] * X[4]) / d;
      inv[3] = (X[5] * X[6] - X[3] * X[8]) / d;
      inv[4] = (X[0] * X[8] - X[2] * X[6]) / d;
      inv[5] = (X[2] * X[3] - X[0] * X[5]) / d;
      inv[6] = (X[3] * X[7] - X[4] * X[6]) / d;
      inv[7] = (X[1] * X[6] - X[0] * X[7]) / d;
      inv[8] = (X[0] * X[4] - X[1] * X[3]) / d;
      R ch = 0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          R nx = R(0.5) * (X[i * 3 + j] + inv[j * 3 + i]);
          ch = fmax(ch, fabs(nx - X[i * 3 + j]));
          X[i * 3 + j] = nx;
        }
      if (ch < 1e-15) break;
    }
  } else {
    for (int i = 0; i < 9; i++) X[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  for (int i = 0; i < 9; i++) out[9 + i] = X[i];
}

// Pi(X) = Rot dev(sym(Rot^T X))
template <typename R>
__device__ __forceinline__ void proj_dev_sym(const R *Rot, const R *X, R *out) {
  R Y[9], S[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Y[i * 3 + j] = Rot[0 * 3 + i] * X[0 * 3 + j] + Rot[1 * 3 + i] * X[1 * 3 + j] + Rot[2 * 3 + i] * X[2 * 3 + j];
  R tr = (Y[0] + Y[4] + Y[8]) / R(3.0);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) S[i * 3 + j] = R(0.5) * (Y[i * 3 + j] + Y[j * 3 + i]) - (i == j ? tr : R(0.0));
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = Rot[i * 3 + 0] * S[0 * 3 + j] + Rot[i * 3 + 1] * S[1 * 3 + j] + Rot[i * 3 + 2] * S[2 * 3 + j];
}

}  // namespace gpr
