// The split-WENO step's pieces that the ADER step does not have: the source
// ODE over half a step solved in closed form, and the midstepper.
//
//   ode_analytic   system/relaxation/analytic.cpp:13-162 (body force, the
//                  distortion's relaxation along its singular values, the
//                  thermal impulse), applied per cell by ode_launcher
//                  (solvers/split/ode.cpp:5-11)
//   k_ode          ode_launcher over a block of cells, one thread per cell
//   k_midstep      midstepper (solvers/split/homogeneous.cpp:5-67): the WENO
//                  reconstruction advanced by dt/2 under the homogeneous
//                  system, one lane per spatial node
#pragma once

#include "gpr_predictor.hpp"

namespace gpr {

// Singular value decomposition A = U diag(s) V^T of a row-major 3x3 matrix by
// one-sided Jacobi (Hestenes): plane rotations applied on the right make the
// columns of A V mutually orthogonal; s are their norms and U the normalised
// columns.  s is ordered as a selection sort on the largest remaining value
// leaves it, ties kept in column order (JacobiSVD's ordering, used by
// analytic.cpp:88 through svd.singularValues()).
__device__ __forceinline__ void svd3(const double *A, double *U, double *s, double *Vm) {
  double W[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    W[k] = A[k];
    Vm[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 12; sweep++) {
    bool rotated = false;
#pragma unroll
    for (int pr = 0; pr < 3; pr++) {
      const int p = pr == 2 ? 1 : 0, q = pr == 0 ? 1 : 2;
      double al = 0, be = 0, ga = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        al += W[k * 3 + p] * W[k * 3 + p];
        be += W[k * 3 + q] * W[k * 3 + q];
        ga += W[k * 3 + p] * W[k * 3 + q];
      }
      if (fabs(ga) > 1e-15 * sqrt(al * be)) {
        rotated = true;
        const double ze = (be - al) / (2.0 * ga);
        const double t = (ze >= 0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const double wp = W[k * 3 + p], wq = W[k * 3 + q];
          W[k * 3 + p] = c * wp - sn * wq;
          W[k * 3 + q] = sn * wp + c * wq;
          const double vp = Vm[k * 3 + p], vq = Vm[k * 3 + q];
          Vm[k * 3 + p] = c * vp - sn * vq;
          Vm[k * 3 + q] = sn * vp + c * vq;
        }
      }
    }
    if (!rotated) break;
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double n = sqrt(W[j] * W[j] + W[3 + j] * W[3 + j] + W[6 + j] * W[6 + j]);
    s[j] = n;
#pragma unroll
    for (int k = 0; k < 3; k++) U[k * 3 + j] = n > 0 ? W[k * 3 + j] / n : (k == j ? 1.0 : 0.0);
  }
  // descending; the first of equal maxima stays first
#pragma unroll
  for (int i = 0; i < 2; i++) {
    int pos = i;
#pragma unroll
    for (int j = i + 1; j < 3; j++)
      if (s[j] > s[pos]) pos = j;
    if (pos != i) {
      double tmp = s[i];
      s[i] = s[pos];
      s[pos] = tmp;
#pragma unroll
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

// The relaxation time of the linearised distortion equations over dt
// (nondimensionalized_time, analytic.cpp:17-70): power-law solid, power-law
// fluid, otherwise the Newtonian rate.  d13 = det(A)^(1/3).
__device__ __forceinline__ double relax_time(double r, double d13, double m0, double u0, double dt,
                                             const gpr_par &p) {
  if (p.POWER_LAW) {
    const double n = p.n;
    const double cs2 = p.b02 * pow(r / p.rho0, p.beta);
    const double a = 9.0 * m0 - u0 - 9.0;
    const double b = 6.0 * m0 - u0 - 6.0;
    const double ab4 = pow(a - b, 4.0), a4 = pow(a, 4.0), b4 = pow(b, 4.0);
    const double a2b = a * a * b - 2.0 * a * b * b + b * b * b;
    const double tail = -8.0 / 7.0 * a * a * b + 2.0 * a * b * b - 8.0 / 9.0 * b * b * b - a4 / 6.0 +
                        16.0 / 27.0 * a * a * a * b - 4.0 / 5.0 * a * a * b * b + 16.0 / 33.0 * a * b * b * b -
                        b4 / 9.0;
    if (p.SOLID) {
      const double c = 108.0 * a - 324.0 * b + 108.0 * a * a - 396.0 * a * b + 297.0 * b * b - 24.0 * a2b -
                       4.0 * ab4;
      if (c <= 0.0) return 0.0;
      const double lam = c / (18.0 * a - 36.0 * b + 9.0 * a * a - 132.0 / 5.0 * a * b + 33.0 / 2.0 * b * b + tail);
      const double t1 = n * lam / p.tau0 * pow(d13, 4.0 * n + 7.0);
      const double t2 = pow(sqrt(c) * r * cs2 / (6.0 * p.sigmaY), n);
      return 2.0 / (n * lam) * log(t1 * t2 * dt + 1.0);
    }
    const double k = (1 - n) / n;
    const double c = 108.0 * a - 324.0 * b + 180.0 * a * a - 612.0 * a * b + 459.0 * b * b - 24.0 * a2b -
                     4.0 * ab4;
    if (c <= 0.0) return 0.0;
    const double lam = c / (18.0 * a - 36.0 * b + 15.0 * a * a - 204.0 / 5.0 * a * b + 51.0 / 2.0 * b * b + tail);
    const double tau1 = 6 * pow(p.mu, 1.0 / n) / (p.rho0 * p.b02);
    const double t1 = k * lam / tau1 * pow(d13, 4.0 * k + 7.0);
    const double t2 = pow(sqrt(c) * r * cs2 / (6.0 * sqrt(3.0)), k);
    return 2.0 / (k * lam) * log(t1 * t2 * dt + 1.0);
  }
  return p.rho0 * p.b02 * pow(d13, 7.0) / (3.0 * p.mu) * dt;
}

// analyticSolver_distortion (analytic.cpp:72-129): the normalised squares of
// A's singular values have mean m0 and spread u0; both relax in closed form
// over dt, the three new values are the roots of the cubic they fix (with
// product one), and A is rebuilt with them on its own singular vectors.
__device__ __forceinline__ void ode_distortion(double *A, double dt, const gpr_par &p) {
  double U[9], s[3], Vm[9];
  svd3(A, U, s, Vm);
  const double detA = s[0] * s[1] * s[2];
  const double r = detA * p.rho0;
  const double d13 = cbrt(detA);
  const double d23 = d13 * d13;
  const double s0 = s[0] * s[0] / d23, s1 = s[1] * s[1] / d23, s2 = s[2] * s[2] / d23;
  const double m0 = (s0 + s1 + s2) / 3.0;
  const double u0 = ((s0 - s1) * (s0 - s1) + (s1 - s2) * (s1 - s2) + (s2 - s0) * (s2 - s0)) / 3.0;
  if (!(u0 >= 1e-12)) return;
  const double tau = relax_time(r, d13, m0, u0, dt, p);
  const double e6 = exp(-6.0 * tau), e9 = exp(-9.0 * tau);
  const double a = 3.0 * m0 - u0 / 3.0 - 3.0;
  const double b = 2.0 * m0 - u0 / 3.0 - 2.0;
  const double m = 1.0 + a * e6 - b * e9;
  const double u = fmax(0.0, 6.0 * a * e6 - 9.0 * b * e9);
  const double D = -2.0 * m * m * m + m * u + 2.0;
  const double arg1 = fmax(0.0, 6.0 * u * u * u - 81.0 * D * D);
  double x1;
  if (fabs(D) < 1e-12) {
    x1 = sqrt(6.0 * u) / 3.0 * cos(M_PI / 6.0) + m;
  } else if (arg1 == 0.0) {
    const double t = cbrt(54.0 * D);
    x1 = t / 6.0 + u / t + m;
  } else {
    const double th = atan(sqrt(arg1) / (9.0 * D));
    x1 = sqrt(6.0 * u) / 3.0 * cos(th / 3.0) + m;
  }
  const double t2 = 3.0 * m - x1;
  const double arg2 = fmax(0.0, x1 * t2 * t2 - 4.0);
  const double x2 = 0.5 * (sqrt(arg2 / x1) + t2);
  const double x3 = 1.0 / (x1 * x2);
  // ascending order of the roots (a three-element sorting network); the
  // largest goes with the largest singular value
  double lo = x1, mid = x2, hi = x3, tmp;
  if (mid < lo) { tmp = lo; lo = mid; mid = tmp; }
  if (hi < mid) { tmp = mid; mid = hi; hi = tmp; }
  if (mid < lo) { tmp = lo; lo = mid; mid = tmp; }
  const double g[3] = {d13 * sqrt(hi), d13 * sqrt(mid), d13 * sqrt(lo)};
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      A[i * 3 + j] = U[i * 3 + 0] * (g[0] * Vm[j * 3 + 0]) + U[i * 3 + 1] * (g[1] * Vm[j * 3 + 1]) +
                     U[i * 3 + 2] * (g[2] * Vm[j * 3 + 2]);
}

// analyticSolver_thermal (analytic.cpp:131-152): the thermal impulse's ODE in
// closed form, with the distortion already relaxed.
__device__ __forceinline__ void ode_thermal(double *q, double dt, const gpr_par &p) {
  const double r = q[0];
  const double E = q[1] / r;
  const double v0 = q[2] / r, v1 = q[3] / r, v2 = q[4] / r;
  // |dev G|^2 from the rows of A, as devGsq (functions/matrices.cpp:17-27)
  const double *A = q + 5;
  double G[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G[i * 3 + j] = A[i * 3] * A[j * 3] + A[i * 3 + 1] * A[j * 3 + 1] + A[i * 3 + 2] * A[j * 3 + 2];
  const double tr = (G[0] + G[4] + G[8]) / 3;
  const double dev2 = (G[0] - tr) * (G[0] - tr) + (G[4] - tr) * (G[4] - tr) + (G[8] - tr) * (G[8] - tr) +
                      2 * (G[1] * G[1] + G[2] * G[2] + G[5] * G[5]);
  const double Gm = Gamma_MG(r, p);
  const double cs2 = p.b02 * pow(r / p.rho0, p.beta);
  const double E2A = (1 - p.beta / Gm) * cs2 / 4 * dev2;
  double c1 = E - E2A - (v0 * v0 + v1 * v1 + v2 * v2) / 2;
  double c2 = p.calpha2 / 2.0;
  const double k = 2 * p.calpha2 / (p.kappa * r * p.cv);
  c1 *= k;
  c2 *= k;
  const double ea = exp(-c1 * dt / 2);
  const double J2 = q[14] * q[14] + q[15] * q[15] + q[16] * q[16];
  const double den = 1 - c2 / c1 * (1 - ea * ea) * J2 / (r * r);
  const double f = ea / sqrt(den);
  q[14] *= f;
  q[15] *= f;
  q[16] *= f;
}

// ode_stepper_analytic (analytic.cpp:154-162); VISCOUS is on in this build.
template <bool TH> __device__ __forceinline__ void ode_analytic(double *q, double dt, const gpr_par &p) {
  q[2] += p.dp[0] * dt;
  q[3] += p.dp[1] * dt;
  q[4] += p.dp[2] * dt;
  ode_distortion(q + 5, dt, p);
  if (TH) ode_thermal(q, dt, p);
}

// ode_launcher over n cells of x, dt = *dt_dev * scale (or dt_host * scale).
// With u non-null, x is the ghost-extended grid ub ([nx+2N][ny+2N] cells, or
// [nx+2N] in 1D) and each interior result is also stored into u: the same
// values as the reference's separate calls on u and on ub
// (steppers.cpp:54-55), since ub's interior is a copy of u.
template <bool TH, int ND>
__global__ __launch_bounds__(256) void k_ode(double *x, long n, double *u, int nx, int ny, int N,
                                            const double *dt_dev, double dt_host, double scale, gpr_par p) {
  constexpr int V = Dims<TH>::V;
  const double dt = dev_or(dt_dev, dt_host) * scale;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = x[c * V + v];
    ode_analytic<TH>(q, dt, p);
#pragma unroll
    for (int v = 0; v < V; v++) x[c * V + v] = q[v];
    if (u) {
      const long nyb = ND == 2 ? ny + 2 * N : 1;
      const long i = c / nyb - N, j = ND == 2 ? c % nyb - N : 0;
      if (i >= 0 && i < nx && j >= 0 && j < (ND == 2 ? ny : 1)) {
        const long ci = ND == 2 ? i * ny + j : i;
#pragma unroll
        for (int v = 0; v < V; v++) u[ci * V + v] = q[v];
      }
    }
  }
}

// midstepper (homogeneous.cpp:5-67): for each masked cell of wh,
//   w_s -= dt/2 ((dF/dx + B_x dw/dx)/dx + (dG/dy + B_y dw/dy)/dy)
// at its spatial nodes s, the derivatives by DERVALS along each axis.  One
// lane per node; the cell's w, F and G are exchanged through LDS.
template <int N, int ND> struct MidShape {
  static constexpr int NS = ND == 1 ? N : N * N;
  static constexpr int CPW = WAVE / NS;
};

template <int N, bool TH, int ND>
__global__ __launch_bounds__(64) void k_midstep(double *wh, long ncell, const uint8_t *mask, const double *dt_dev,
                                               double dt_host, double dx, double dy, Mat m) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = MidShape<N, ND>::NS;
  constexpr int CPW = MidShape<N, ND>::CPW;
  constexpr int NF = ND;  // flux directions
  __shared__ double sW[CPW * NS * V];
  __shared__ double sF[CPW * NF * NS * V];
  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / NS, node = lane - slot * NS;
  const bool lane_ok = slot < CPW;
  const long c = (long)blockIdx.x * CPW + slot;
  const bool active = lane_ok && c < ncell && mask[c];
  const int ia = ND == 1 ? node : node / N, ib = ND == 1 ? 0 : node - (node / N) * N;
  const double dt = dev_or(dt_dev, dt_host);
  const int sl = lane_ok ? slot : 0;
  double *myW = sW + sl * NS * V;
  double *myF = sF + sl * NF * NS * V;
  double w[V];
#pragma unroll
  for (int v = 0; v < V; v++) w[v] = active ? wh[(c * NS + node) * V + v] : (v == 0 ? 1.0 : 0.0);
  {
    NodeState<double, TH> st;
    node_state<double, TH>(w, m, st);
    double F[V];
    flux<double, TH>(w, st, m, 0, F);
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myF[node * V + v] = F[v];
    if (ND == 2) {
      flux<double, TH>(w, st, m, 1, F);
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myF[(NS + node) * V + v] = F[v];
    }
  }
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myW[node * V + v] = w[v];
  __syncthreads();
  if (!active) return;
  const double vel[3] = {w[2] / w[0], w[3] / w[0], w[4] / w[0]};
  double out[V];
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = w[v];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    // derivative along axis d at this node: x runs over the first index
    double dw[V], dF[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double aw = 0, af = 0;
#pragma unroll
      for (int k = 0; k < N; k++) {
        const int nk = ND == 1 ? k : (d == 0 ? k * N + ib : ia * N + k);
        const double dk = K.der[d == 0 ? ia : ib][k];
        aw += dk * myW[nk * V + v];
        af += dk * myF[(d * NS + nk) * V + v];
      }
      dw[v] = aw;
      dF[v] = af;
    }
    double Bw[V];
    bdot<double, V>(vel, dw, d, Bw);
    const double h = d == 0 ? dx : dy;
#pragma unroll
    for (int v = 0; v < V; v++) dF[v] = (dF[v] + Bw[v]) / h;
    if (ND == 1) {
#pragma unroll
      for (int v = 0; v < V; v++) out[v] -= dt / 2 * dF[v];
    } else if (d == 0) {
#pragma unroll
      for (int v = 0; v < V; v++) out[v] = dF[v];  // x term, combined below
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) out[v] = w[v] - dt / 2 * (out[v] + dF[v]);
    }
  }
#pragma unroll
  for (int v = 0; v < V; v++) wh[(c * NS + node) * V + v] = out[v];
}

}  // namespace gpr
