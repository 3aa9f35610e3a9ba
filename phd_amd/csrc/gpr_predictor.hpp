// Local space-time DG predictor (solvers/dg/dg1.cpp, dg2.cpp) on gfx950.
//
// One lane per space-time node (t, a[, b]) of a cell: 27 lanes per cell in 2D
// with N = 3, two cells per 64-lane wavefront (CPW cells per wave in
// general).  Each lane evaluates the pointwise physics at its node; the
// derivative, flux and time contractions between nodes go through LDS.
//
// Per cell, iterate q -> (DG_MAT (x) W_s)^-1 rhs(q) (Picard, dg2.cpp:273-298)
// with the reference's stopping rule over the whole block
//   any |q1 - q0| > DG_TOL (1 + |q0|)  -> iterate again.
// A cell that does not converge within dg_it iterations, ends non-finite, or
// (bailout > 0) whose stopping statistic max |q1-q0| / (DG_TOL (1+|q0|)) grew
// `bailout` times in a row, restarts from the standard initial guess in the
// preconditioned iteration
//   q <- q - P^-1 (DG_MAT q - rhs(q) / W_s),
//   P^-1 = DG_MAT^-1 off the stiff directions, (DG_MAT + dt kappa W)^-1 on the
//   distortion directions Pi (and (DG_MAT + dt rJ W)^-1 on the thermal ones),
// which solves the same equations and reduces to Picard when kappa = rJ = 0.
#pragma once

#include "gpr_device.hpp"

// waves per SIMD the predictor is compiled for (register budget 512 / waves)
#ifndef GPR_PRED_WAVES
#define GPR_PRED_WAVES 1
#endif

namespace gpr {

constexpr int WAVE = 64;

struct StepState;

struct PredArgs {
  double *qh;              // [ncell][NN][V]
  const double *wh;        // [ncell][NS][V]
  const uint8_t *mask;     // [ncell]
  int32_t *iters;          // optional per-cell iteration count (fallback: 1000 + its)
  int *fail;               // device counter of cells that converged in neither phase
  unsigned long long *itc; // optional: [0] Picard, [1] fallback iterations, [2] cells, [3] fallback cells
  const double *dt_dev;    // if non-null, dt is read from here
  double dt, dx, dy;
  long ncell;
  int dg_it, stiff_it, bailout, mode;  // mode 0: Picard then fallback; 1: fallback only
  double dg_tol, stiff_tol;
  Mat m;
};

// One wavefront of 64 lanes holds CPW cells, one space-time node per lane.
template <int N, int ND> struct PredShape {
  static constexpr int NS = ND == 1 ? N : N * N;  // spatial nodes per cell
  static constexpr int NN = NS * N;                // space-time nodes per cell
  static constexpr int CPW = WAVE / NN;            // cells per wavefront
};

// Preconditioner of one spatial node, frozen at its WENO state ws:
// out[0..8] = (DG_MAT + dt kappa W)^-1, out[9..17] = polar rotation of A,
// out[18..26] = (DG_MAT + dt rJ W)^-1 (thermal only).  At a relaxed state
// A = s Rot the distortion source has Jacobian -kappa Pi with
// Pi(X) = Rot dev(sym(Rot^T X)), kappa = 2 theta1inv cs^2 det(A)^(2/3);
// the thermal source -(c_alpha^2 T / (kappa rho)) rho J has rate rJ.
template <int N, bool TH>
__device__ void stiff_precond(const double *ws, const PredArgs &a, double dt, double *out) {
  const Basis &K = basis_dev<N>();
  NodeState<double, TH> s;
  node_state<double, TH>(ws, a.m, s);
  double dA = det3(ws + 5);
  double s13 = cbrt(dA);
  double kap = 2.0 * theta1inv<double, TH>(ws, s, a.m) * s.cs2 * s13 * s13;
  if (!(kap >= 0.0) || !(kap < 1e300)) kap = 0.0;
  double rJ = 0.0;
  if (TH) {
    rJ = a.m.p.calpha2 * s.T / (a.m.p.kappa * s.r);
    if (!(rJ >= 0.0) || !(rJ < 1e300)) rJ = 0.0;
  }
  for (int which = 0; which < (TH ? 2 : 1); which++) {
    double rate = which == 0 ? kap : rJ;
    double M[N][2 * N];
    for (int i = 0; i < N; i++)
      for (int j = 0; j < 2 * N; j++)
        M[i][j] = j < N ? K.dgmat[i][j] + (i == j ? dt * rate * K.w[i] : 0.0) : (j - N == i ? 1.0 : 0.0);
    for (int c = 0; c < N; c++) {
      int piv = c;
      for (int r = c + 1; r < N; r++)
        if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
      if (piv != c)
        for (int j = 0; j < 2 * N; j++) {
          double tt = M[c][j];
          M[c][j] = M[piv][j];
          M[piv][j] = tt;
        }
      double dg = M[c][c];
      for (int j = 0; j < 2 * N; j++) M[c][j] /= dg;
      for (int r = 0; r < N; r++) {
        if (r == c) continue;
        double f = M[r][c];
        for (int j = 0; j < 2 * N; j++) M[r][j] -= f * M[c][j];
      }
    }
    double *dst = which == 0 ? out : out + 18;
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) dst[i * N + j] = M[i][N + j];
  }
  double X[9];
  for (int i = 0; i < 9; i++) X[i] = ws[5 + i];
  if (dA > 0) {
    for (int it = 0; it < 40; it++) {  // X <- (X + X^-T) / 2
      double d = det3(X), inv[9];
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
      if (ch < 1e-15) break;
    }
  } else {
    for (int i = 0; i < 9; i++) X[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  for (int i = 0; i < 9; i++) out[9 + i] = X[i];
}

// The predictor's workgroup is exactly one wavefront, so exchanging values
// between its lanes through LDS needs ordering within the wave, not a
// workgroup barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a / b to within a few ulps: hardware reciprocal refined by one Newton step
__device__ __forceinline__ double fast_div(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return a * r;
}

// Pi(X) = Rot dev(sym(Rot^T X))
__device__ __forceinline__ void proj_dev_sym(const double *Rot, const double *X, double *out) {
  double Y[9], S[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Y[i * 3 + j] = Rot[0 * 3 + i] * X[0 * 3 + j] + Rot[1 * 3 + i] * X[1 * 3 + j] + Rot[2 * 3 + i] * X[2 * 3 + j];
  double tr = (Y[0] + Y[4] + Y[8]) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) S[i * 3 + j] = 0.5 * (Y[i * 3 + j] + Y[j * 3 + i]) - (i == j ? tr : 0.0);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = Rot[i * 3 + 0] * S[0 * 3 + j] + Rot[i * 3 + 1] * S[1 * 3 + j] + Rot[i * 3 + 2] * S[2 * 3 + j];
}

template <int N, bool TH, int ND>
__global__ __launch_bounds__(64, GPR_PRED_WAVES) void k_predictor(PredArgs a) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = PredShape<N, ND>::NS;
  constexpr int NN = PredShape<N, ND>::NN;
  constexpr int CPW = PredShape<N, ND>::CPW;
  constexpr int PRE = TH ? 27 : 18;
  __shared__ double sQ[CPW * NN * V];   // iterate; in the fallback update also P g
  __shared__ double sF[CPW * NN * V];   // fluxes, then rhs or g
  __shared__ double sPre[CPW * NS * PRE];
  __shared__ double sRat[WAVE];

  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / NN;
  const int node = lane - slot * NN;
  const int t = node / NS;
  const int s = node - t * NS;
  const int ia = ND == 1 ? s : s / N;
  const int ib = ND == 1 ? 0 : s - (s / N) * N;
  const long cell = (long)blockIdx.x * CPW + slot;
  const bool lane_ok = slot < CPW;  // lanes past the last whole cell write no LDS
  const bool valid = lane_ok && cell < a.ncell;
  const bool active = valid && a.mask[cell];
  const int sl = lane_ok ? slot : 0;
  double *myQ = sQ + (sl * NN) * V;
  double *myF = sF + (sl * NN) * V;
  double *myPre = sPre + (sl * NS + s) * PRE;
  const double *wsrc = a.wh + (cell * NS + s) * V;

  const double dt = dev_or(a.dt_dev, a.dt);
  const double Ws = ND == 1 ? K.w[ia] : K.w[ia] * K.w[ib];
  const double iWs = 1.0 / Ws;
  const double Wt = K.w[t];
  const double cW = K.endv[0][t] * Ws;  // initial-condition weight psi_t(0) W_s
  const double inv_dx = 1.0 / a.dx, inv_dy = 1.0 / a.dy;
  // this lane's rows of the basis matrices
  double cdx[N], cdy[N], cgx[N], cgy[N], cmi[N], cmt[N];
#pragma unroll
  for (int j = 0; j < N; j++) {
    cdx[j] = K.der[ia][j];
    cdy[j] = K.der[ib][j];
    cgx[j] = K.dgder[ia][j];
    cgy[j] = K.dgder[ib][j];
    cmi[j] = K.dgmatinv[t][j];
    cmt[j] = K.dgmat[t][j];
  }

  double q0[V], w[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    w[v] = active ? wsrc[v] : (v == 0 || v == 5 || v == 9 || v == 13 ? 1.0 : 0.0);
    q0[v] = w[v];
  }

  // phase: 0 Picard, 1 preconditioned, 2 converged, 3 idle / given up
  int phase = active ? (a.mode == 1 ? 1 : 0) : 3;
  int it = 0, pit = 0, grow = 0;
  double prev = 1e308;
  const unsigned long long cellbits = (NN == 64 ? ~0ull : ((1ull << NN) - 1)) << (sl * NN);
  bool need_pre = phase == 1;

  while (__any(phase == 0 || phase == 1)) {
    if (need_pre && t == 0 && lane_ok) {
      stiff_precond<N, TH>(w, a, dt, myPre);
    }
    need_pre = false;
    // (1) share the iterate
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myQ[node * V + v] = q0[v];
    wave_sync();

    // (2) spatial derivatives, source, non-conservative products, fluxes
    double rhs[V], G[V];
    {
      NodeState<double, TH> st;
      node_state<double, TH>(q0, a.m, st);
      source<double, TH>(q0, st, a.m, rhs);
      // B_d(q) . dq/dx_d reads only the distortion rows of dq (equations.cpp:189-247)
      double dq[14];
#pragma unroll
      for (int v = 5; v < 14; v++) {
        double acc = 0;
#pragma unroll
        for (int j = 0; j < N; j++) acc += cdx[j] * myQ[(t * NS + (ND == 1 ? j : j * N + ib)) * V + v];
        dq[v] = acc * inv_dx;
      }
      bdot_sub(st.v, dq, 0, 1.0, rhs);
      if (ND == 2) {
#pragma unroll
        for (int v = 5; v < 14; v++) {
          double acc = 0;
#pragma unroll
          for (int j = 0; j < N; j++) acc += cdy[j] * myQ[(t * NS + ia * N + j) * V + v];
          dq[v] = acc * inv_dy;
        }
        bdot_sub(st.v, dq, 1, 1.0, rhs);
      }
      const double wts = Wt * Ws;
#pragma unroll
      for (int v = 0; v < V; v++) rhs[v] *= wts;
      double F[V];
      flux<double, TH>(q0, st, a.m, 0, F);
      const double wb = ND == 2 ? K.w[ib] : 1.0;
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myF[node * V + v] = F[v] * wb;
      if (ND == 2) flux<double, TH>(q0, st, a.m, 1, G);
    }
    wave_sync();
    // (3) flux contractions (rhs2, dg2.cpp:173-201) through one buffer
    {
      const double cx = Wt * inv_dx;
#pragma unroll
      for (int v = 0; v < V; v++) {
        double sx = 0;
#pragma unroll
        for (int j = 0; j < N; j++) sx += cgx[j] * myF[(t * NS + (ND == 1 ? j : j * N + ib)) * V + v];
        rhs[v] -= cx * sx;
      }
    }
    if (ND == 2) {
      wave_sync();
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myF[node * V + v] = G[v];
      wave_sync();
      const double cy = Wt * K.w[ia] * inv_dy;
#pragma unroll
      for (int v = 0; v < V; v++) {
        double sy = 0;
#pragma unroll
        for (int j = 0; j < N; j++) sy += cgy[j] * myF[(t * NS + ia * N + j) * V + v];
        rhs[v] -= cy * sy;
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++) rhs[v] = rhs[v] * dt + cW * w[v];
    // (4) time solve.  Picard: q1 = (DG_MAT (x) W_s)^-1 rhs.  Fallback:
    // g = DG_MAT q0 - rhs / W_s (q0 of the time column from sQ), P g in sQ.
    double P[9];
    if (phase == 1) {
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < N; k++) acc += cmt[k] * myQ[(k * NS + s) * V + v];
        rhs[v] = acc - rhs[v] * iWs;  // rhs now holds g
      }
      proj_dev_sym(myPre + 9, rhs + 5, P);
    }
    wave_sync();
    if (lane_ok) {
#pragma unroll
      for (int v = 0; v < V; v++) myF[node * V + v] = rhs[v];
      if (phase == 1)
#pragma unroll
        for (int k = 0; k < 9; k++) myQ[node * V + 5 + k] = P[k];
    }
    wave_sync();
    if (phase == 1) {
      const double *Ainv = myPre, *Jinv = myPre + 18;
#pragma unroll
      for (int v = 0; v < V; v++) {
        double d = 0;
        if (v >= 5 && v < 14) {
#pragma unroll
          for (int k = 0; k < N; k++) {
            double gk = myF[(k * NS + s) * V + v];
            double pk = myQ[(k * NS + s) * V + v];
            d += cmi[k] * (gk - pk) + Ainv[t * N + k] * pk;
          }
        } else if (v >= 14) {
#pragma unroll
          for (int k = 0; k < N; k++) d += Jinv[t * N + k] * myF[(k * NS + s) * V + v];
        } else {
#pragma unroll
          for (int k = 0; k < N; k++) d += cmi[k] * myF[(k * NS + s) * V + v];
        }
        rhs[v] = q0[v] - d;  // rhs now holds q1
      }
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < N; k++) acc += cmi[k] * myF[(k * NS + s) * V + v];
        rhs[v] = acc * iWs;  // rhs now holds q1
      }
    }
    // (5) stopping rule over the cell's whole block (dg2.cpp:278-297)
    const double tol = phase == 1 ? a.stiff_tol : a.dg_tol;
    bool more = false, nan = false;
    double ratio = 0;
#pragma unroll
    for (int v = 0; v < V; v++) {
      double dd = fabs(rhs[v] - q0[v]), lim = tol * (1.0 + fabs(q0[v]));
      more |= dd > lim;
      nan |= rhs[v] != rhs[v];
      if (a.bailout > 0) ratio = fmax(ratio, fast_div(dd, lim));
    }
    const unsigned long long bm = __ballot(more), bn = __ballot(nan);
    const bool cmore = (bm & cellbits) != 0, cnan = (bn & cellbits) != 0;
    if (a.bailout > 0) {
      sRat[lane] = ratio;
      wave_sync();
      double cr = 0;
      for (int k = 0; k < NN; k++) cr = fmax(cr, sRat[sl * NN + k]);
      ratio = cr;
    }
    if (phase == 0) {
      it++;
      bool to_fallback = false;
      if (cmore) {
#pragma unroll
        for (int v = 0; v < V; v++) q0[v] = rhs[v];
        grow = ratio > prev ? grow + 1 : 0;
        prev = ratio;
        if (it >= a.dg_it || (a.bailout > 0 && grow >= a.bailout)) to_fallback = true;
      } else if (cnan) {
        to_fallback = true;
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) a.qh[(cell * NN + node) * V + v] = rhs[v];
        if (node == 0) {
          if (a.iters) a.iters[cell] = it;
          if (a.itc) {
            atomicAdd(&a.itc[0], (unsigned long long)it);
            atomicAdd(&a.itc[2], 1ull);
          }
        }
        phase = 2;
      }
      if (to_fallback) {
        phase = 1;
        pit = 0;
        need_pre = true;
#pragma unroll
        for (int v = 0; v < V; v++) q0[v] = w[v];
      }
    } else if (phase == 1) {
      pit++;
#pragma unroll
      for (int v = 0; v < V; v++) q0[v] = rhs[v];
      if (!cmore || pit >= a.stiff_it) {
#pragma unroll
        for (int v = 0; v < V; v++) a.qh[(cell * NN + node) * V + v] = rhs[v];
        bool ok = !cmore && !cnan;
        if (node == 0) {
          if (a.iters) a.iters[cell] = ok ? 1000 + pit : -1;
          if (!ok) atomicAdd(a.fail, 1);
          if (a.itc) {
            atomicAdd(&a.itc[0], (unsigned long long)it);
            atomicAdd(&a.itc[1], (unsigned long long)pit);
            atomicAdd(&a.itc[2], 1ull);
            atomicAdd(&a.itc[3], 1ull);
          }
        }
        phase = ok ? 2 : 3;
      }
    }
    if (a.bailout > 0) wave_sync();  // sRat is rewritten next iteration
  }
}

}  // namespace gpr
