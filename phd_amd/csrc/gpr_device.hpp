// Device-side GPR physics and basis constants for the ADER-WENO kernels.
//
// Pointwise functions follow code/cpp/src/system of haranjackson/phd:
//   fluxes        equations.cpp:11-48       sources   equations.cpp:50-67
//   B(Q).x        equations.cpp:189-267     pressure  variables/state.cpp:11-21
//   stress        variables/state.cpp:23-84 theta1inv variables/sources.cpp:13-81
//   EOS           energy/mg.cpp, energy/eos.cpp, energy/derivatives.cpp
//   wave speeds   waves/shear.cpp, waves/speeds.cpp, eig.cpp:8-77
// Each node's flux, source and stress share one evaluation of A^T A, dev G,
// A dev G and the pressure instead of recomputing them per call as the
// reference does.  State layout per cell: [ρ, ρE, ρv(3), A row-major (9), ρJ(3)].
#pragma once

#include <type_traits>

#ifdef GPR_HOST_ONLY
// host compile of the pointwise physics (flopcount.cpp): the templated
// functions below are instantiated with a counting number type
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
#define __constant__
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "../../include/gpr_par.h"

namespace gpr {

constexpr int MAXN = 3;

// Basis, WENO and DG constants for one N, built on the host (gpr_host.cpp)
// and passed by value to every kernel.
struct Basis {
  int N;
  int pad_;
  double x[MAXN], w[MAXN];
  double endv[2][MAXN];      // psi_j(0), psi_j(1)
  double der[MAXN][MAXN];    // psi_j'(x_i)                      (DERVALS)
  double dgder[MAXN][MAXN];  // <psi_i, psi_j'>                  (DG_DER)
  double dgmat[MAXN][MAXN];  // DG_END - DG_DER^T                (DG_MAT)
  double dgmatinv[MAXN][MAXN];
  double mLinv[MAXN][MAXN], mRinv[MAXN][MAXN], mCLinv[MAXN][MAXN];
  double sig[MAXN][MAXN];    // WENO oscillation indicator
};

// The basis tables for N = 2 and N = 3 in the constant segment, written at
// context creation.  Kernels index them per lane; the same per-lane index
// into a table held in a by-value kernel argument would make the compiler
// copy the whole argument to the stack.  Each translation unit that launches
// kernels has its own copy (internal linkage), written by that unit.
static __constant__ Basis c_basis[2];
template <int N> __device__ __forceinline__ const Basis &basis_dev() { return c_basis[N - 2]; }

// *p if p is non-null, else v.  The load is atomic so that the compiler cannot
// merge the two sources into a load through a selected pointer, which would
// take the address of the by-value kernel argument holding v and copy the
// whole argument to the stack.
#ifndef GPR_HOST_ONLY
__device__ __forceinline__ double dev_or(const double *p, double v) {
  return p ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : v;
}
#endif

// Material parameters plus the few derived scalars the kernels use often.
struct Mat {
  gpr_par p;
  double mu_fac;   // rho0 / (2 mu^(1/n)) for the Newtonian relaxation time
  int beta_zero;   // b02 (rho/rho0)^beta == b02
  int pad_;
};

template <bool TH> struct Dims {
  static constexpr int V = TH ? 17 : 14;
};

// Material kinds a kernel can be compiled for.  MK_ANY reads the EOS and the
// relaxation model from the parameters at run time; the others fix them at
// compile time, so that a kernel carries no code (and keeps no hoisted
// constants in registers) for the branches its material never takes.
enum { MK_ANY = 0, MK_GAS_NEWTONIAN = 1, MK_GR_PLASTIC = 2 };
template <int MK> struct MatKind {
  // -1: read from the parameters
  static constexpr int eos = MK == MK_GAS_NEWTONIAN ? GPR_EOS_STIFFENED_GAS
                             : MK == MK_GR_PLASTIC  ? GPR_EOS_GODUNOV_ROMENSKI
                                                    : -1;
  static constexpr int solid = MK == MK_GAS_NEWTONIAN ? 0 : MK == MK_GR_PLASTIC ? 1 : -1;
  static constexpr int power_law = MK == MK_GAS_NEWTONIAN ? 0 : MK == MK_GR_PLASTIC ? 1 : -1;
  static constexpr int bingham = MK == MK_ANY ? -1 : 0;
  static constexpr int beta_zero = MK == MK_GAS_NEWTONIAN ? 1 : -1;
};
template <int MK> __host__ __device__ __forceinline__ int eos_code(const gpr_par &p) {
  return MatKind<MK>::eos >= 0 ? MatKind<MK>::eos : p.EOS;
}
template <int MK> __host__ __device__ __forceinline__ bool mk_flag(int fixed, int runtime) {
  return fixed >= 0 ? fixed != 0 : runtime != 0;
}
// The most specific material kind a parameter set fits (MK_ANY if none).
inline int material_kind(const Mat &m) {
  const gpr_par &p = m.p;
  if (p.BINGHAM) return MK_ANY;
  if (p.EOS == GPR_EOS_STIFFENED_GAS && !p.SOLID && !p.POWER_LAW && m.beta_zero) return MK_GAS_NEWTONIAN;
  if (p.EOS == GPR_EOS_GODUNOV_ROMENSKI && p.SOLID && p.POWER_LAW) return MK_GR_PLASTIC;
  return MK_ANY;
}

// ------------------------------------------------------------------ EOS
// The value of an exponent, which decides how dpow evaluates (the dual-number
// and counting types overload it beside their own definitions).
__device__ __forceinline__ double exponent_value(double b) { return b; }
__device__ __forceinline__ double exponent_value(float b) { return b; }

// a^b.  A whole exponent from 2 to 16, such as a Godunov-Romenski shear
// exponent beta = 3 or a power-law exponent n = 10, is evaluated by
// repeated squaring (a^10 = a^2 a^8: four products), which costs a fraction
// of pow's logarithm and exponential and differs from it by a few ulps.
// The exponents are material constants, so every lane takes the same branch.
template <typename R> __device__ __forceinline__ R dpow(R a, R b) {
  const double e = exponent_value(b);
  if (e >= 2.0 && e <= 16.0 && e == double(int(e))) {
    int k = int(e) - 1;  // out = a^(1 + k)
    R base = a, out = a;
    while (k) {
      if (k & 1) out = out * base;
      k >>= 1;
      if (k) base = base * base;
    }
    return out;
  }
  return pow(a, b);
}

template <typename R, int MK = MK_ANY> __device__ __forceinline__ R Gamma_MG(R r, const gpr_par &p) {
  switch (eos_code<MK>(p)) {
  case GPR_EOS_STIFFENED_GAS: return R(p.gamma - 1);
  case GPR_EOS_SHOCK_MG: return R(p.Gamma0 * p.rho0) / r;
  case GPR_EOS_GODUNOV_ROMENSKI: return R(p.gamma);
  default: return R(p.Gamma0);
  }
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R dGamma_MG(R r, const gpr_par &p) {
  if (eos_code<MK>(p) == GPR_EOS_SHOCK_MG) return -R(p.Gamma0 * p.rho0) / (r * r);
  return R(0);
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R p_ref(R r, const gpr_par &p) {
  R r0 = R(p.rho0);
  switch (eos_code<MK>(p)) {
  case GPR_EOS_STIFFENED_GAS: return -R(p.pINF);
  case GPR_EOS_SHOCK_MG:
    if (r > r0) {
      R t = R(1) / r0 - R(p.s) * (R(1) / r0 - R(1) / r);
      return R(p.c02) * (R(1) / r0 - R(1) / r) / (t * t);
    }
    return R(p.c02) * (r - r0);
  case GPR_EOS_GODUNOV_ROMENSKI: {
    R t = p.alpha == 1.0 ? r / r0 : dpow(r / r0, R(p.alpha));
    return R(p.c02) * r / R(p.alpha) * (t - R(1)) * t;
  }
  case GPR_EOS_JWL: {
    R v = r0 / r;
    return R(p.A) * exp(-R(p.R1) * v) + R(p.B) * exp(-R(p.R2) * v);
  }
  default: {
    R v = r0 / r;
    return R(p.A) * dpow(v, -R(p.R1)) - R(p.B) * dpow(v, -R(p.R2));
  }
  }
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R e_ref(R r, const gpr_par &p) {
  R r0 = R(p.rho0);
  switch (eos_code<MK>(p)) {
  case GPR_EOS_STIFFENED_GAS: return R(p.pINF) / r;
  case GPR_EOS_SHOCK_MG:
    if (r > r0) return R(0.5) * p_ref<R, MK>(r, p) * (R(1) / r0 - R(1) / r);
    return R(0);
  case GPR_EOS_GODUNOV_ROMENSKI: {
    R t = p.alpha == 1.0 ? r / r0 : dpow(r / r0, R(p.alpha));
    return R(p.c02) / R(2 * p.alpha * p.alpha) * (t - R(1)) * (t - R(1));
  }
  case GPR_EOS_JWL: {
    R v = r0 / r;
    return R(p.A) / (r0 * R(p.R1)) * exp(-R(p.R1) * v) + R(p.B) / (r0 * R(p.R2)) * exp(-R(p.R2) * v);
  }
  default: {
    R v = r0 / r;
    return -R(p.A) / (r0 * R(1 - p.R1)) * (dpow(v, R(1 - p.R1)) - R(1)) +
           R(p.B) / (r0 * R(1 - p.R2)) * (dpow(v, R(1 - p.R2)) - R(1));
  }
  }
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R dp_ref(R r, const gpr_par &p) {
  R r0 = R(p.rho0);
  switch (eos_code<MK>(p)) {
  case GPR_EOS_STIFFENED_GAS: return R(0);
  case GPR_EOS_SHOCK_MG:
    if (r > r0) {
      R t = R(p.s) * (r - r0) - r;
      return R(p.c02) * r0 * r0 * (R(p.s) * (r0 - r) - r) / (t * t * t);
    }
    return R(p.c02);
  case GPR_EOS_GODUNOV_ROMENSKI: {
    R t = p.alpha == 1.0 ? r / r0 : dpow(r / r0, R(p.alpha));
    return R(p.c02) / R(p.alpha) * t * (R(1 + p.alpha) * (t - R(1)) + R(p.alpha) * t);
  }
  case GPR_EOS_JWL: {
    R v = r0 / r;
    return v / r * (R(p.A * p.R1) * exp(-R(p.R1) * v) + R(p.B * p.R2) * exp(-R(p.R2) * v));
  }
  default: {
    R v = r0 / r;
    return R(p.A * p.R1) / r * dpow(v, -R(p.R1)) - R(p.B * p.R2) / r * dpow(v, -R(p.R2));
  }
  }
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R de_ref(R r, const gpr_par &p) {
  if (eos_code<MK>(p) == GPR_EOS_SHOCK_MG) {
    R r0 = R(p.rho0);
    if (r > r0) {
      R t = R(p.s) * (r - r0) - r;
      return -(r - r0) * r0 * R(p.c02) / (t * t * t);
    }
    return R(0);
  }
  return p_ref<R, MK>(r, p) / (r * r);
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R phi_int(R r, const gpr_par &p) {
  R r0 = R(p.rho0);
  switch (eos_code<MK>(p)) {
  case GPR_EOS_STIFFENED_GAS: return dpow(r / r0, R(p.gamma - 1));
  case GPR_EOS_SHOCK_MG: return exp(R(p.Gamma0) * (R(1) - r0 / r));
  case GPR_EOS_GODUNOV_ROMENSKI: return dpow(r / r0, R(p.gamma));
  default: return dpow(r / r0, R(p.Gamma0));
  }
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R temperature_mg(R r, R pp, const gpr_par &p) {
  R G = Gamma_MG<R, MK>(r, p);
  R T = (pp - p_ref<R, MK>(r, p)) / (r * G * R(p.cv));
  if (p.Tref != 0.0) T += phi_int<R, MK>(r, p) * R(p.Tref);
  return T;
}
template <typename R, int MK = MK_ANY> __device__ __forceinline__ R c_s2(R r, const Mat &m) {
  return mk_flag<MK>(MatKind<MK>::beta_zero, m.beta_zero) ? R(m.p.b02)
                                                          : R(m.p.b02) * dpow(r / R(m.p.rho0), R(m.p.beta));
}

template <typename R> __device__ __forceinline__ R det3(const R *A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// Everything a node needs from one state: pressure, stress, dE/dA.
template <typename R, bool TH> struct NodeState {
  R r, ir, v[3], p, T;
  R EA[9];   // dE/dA at constant entropy = cs^2 A dev(A^T A)
  R sig[9];  // viscous stress, symmetric: -rho EA^T A
  R cs2;
  R devGsq;
};

template <typename R, bool TH, int MK = MK_ANY>
__device__ __forceinline__ void node_state(const R *q, const Mat &m, NodeState<R, TH> &s) {
  const gpr_par &p = m.p;
  s.r = q[0];
  s.ir = R(1) / s.r;
  s.v[0] = q[2] * s.ir;
  s.v[1] = q[3] * s.ir;
  s.v[2] = q[4] * s.ir;
  const R *A = q + 5;
  R G[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++) {
      R g = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
      G[i * 3 + j] = g;
      G[j * 3 + i] = g;
    }
  R tr = (G[0] + G[4] + G[8]) / R(3);
  R d0 = G[0] - tr, d1 = G[4] - tr, d2 = G[8] - tr;
  s.devGsq = d0 * d0 + d1 * d1 + d2 * d2 + R(2) * (G[1] * G[1] + G[2] * G[2] + G[5] * G[5]);
  G[0] = d0;
  G[4] = d1;
  G[8] = d2;
  s.cs2 = c_s2<R, MK>(s.r, m);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      s.EA[i * 3 + j] = s.cs2 * (A[i * 3 + 0] * G[0 * 3 + j] + A[i * 3 + 1] * G[1 * 3 + j] +
                                 A[i * 3 + 2] * G[2 * 3 + j]);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++) {
      R x = -s.r * (s.EA[0 * 3 + i] * A[0 * 3 + j] + s.EA[1 * 3 + i] * A[1 * 3 + j] +
                    s.EA[2 * 3 + i] * A[2 * 3 + j]);
      s.sig[i * 3 + j] = x;
      s.sig[j * 3 + i] = x;
    }
  R Gm = Gamma_MG<R, MK>(s.r, p);
  R C0 = (R(1) - R(p.beta) / Gm) * s.cs2;
  R e = q[1] * s.ir - (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]) / R(2) -
        C0 / R(4) * s.devGsq;
  if (TH) {
    R J0 = q[14] * s.ir, J1 = q[15] * s.ir, J2 = q[16] * s.ir;
    e -= R(p.calpha2) * (J0 * J0 + J1 * J1 + J2 * J2) / R(2);
  }
  s.p = (e - e_ref<R, MK>(s.r, p)) * s.r * Gm + p_ref<R, MK>(s.r, p);
  s.T = TH ? temperature_mg<R, MK>(s.r, s.p, p) : R(0);
}

// Relaxation rate of the distortion (variables/sources.cpp:13-81).
template <typename R, bool TH, int MK = MK_ANY>
__device__ __forceinline__ R theta1inv(const R *q, const NodeState<R, TH> &s, const Mat &m) {
  const gpr_par &p = m.p;
  R dA = det3(q + 5);
  R c13 = cbrt(dA);
  R A53 = dA * c13 * c13;
  if (mk_flag<MK>(MatKind<MK>::solid, p.SOLID)) {
    if (mk_flag<MK>(MatKind<MK>::power_law, p.POWER_LAW)) {
      const R *S = s.sig;
      R a = S[0] - S[4], b = S[4] - S[8], c = S[8] - S[0];
      R sn = sqrt(R(0.5) * (a * a + b * b + c * c) + R(3) * (S[1] * S[1] + S[5] * S[5] + S[6] * S[6]));
      sn = fmin(sn, R(1e8));
      return R(3) * A53 / (s.cs2 * R(p.tau0)) * dpow(sn / R(p.sigmaY), R(p.n));
    }
    return R(0);
  }
  if (mk_flag<MK>(MatKind<MK>::bingham, p.BINGHAM)) {
    R f = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) f += s.sig[i] * s.sig[i];
    R sn = sqrt(f) / sqrt(R(2));
    if (sn > R(p.sigmaY)) return A53 * R(p.rho0) / R(2 * p.mu) * (R(1) - R(p.sigmaY) / sn);
    return R(0);
  }
  R c = A53 * R(m.mu_fac);
  if (mk_flag<MK>(MatKind<MK>::power_law, p.POWER_LAW)) {
    R f = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) f += s.sig[i] * s.sig[i];
    R sn = sqrt(f) / sqrt(R(2));
    return c * dpow(sn, (R(1) - R(p.n)) / R(p.n));
  }
  return c;
}

// Algebraic source (equations.cpp:50-67).
template <typename R, bool TH, int MK = MK_ANY>
__device__ __forceinline__ void source(const R *q, const NodeState<R, TH> &s, const Mat &m, R *S) {
  constexpr int V = Dims<TH>::V;
  S[0] = R(0);
  S[1] = R(0);
  S[2] = R(m.p.dp[0]);
  S[3] = R(m.p.dp[1]);
  S[4] = R(m.p.dp[2]);
  R th = theta1inv<R, TH, MK>(q, s, m);
#pragma unroll
  for (int i = 0; i < 9; i++) S[5 + i] = -s.EA[i] * th;
  if (TH) {
    R th2 = s.T / (R(m.p.kappa) * s.r);
#pragma unroll
    for (int i = 0; i < 3; i++) S[14 + i] = -s.r * (R(m.p.calpha2) * q[14 + i] * s.ir) * th2;
  }
  (void)V;
}

// Conservative flux in direction d (equations.cpp:11-48), written (not accumulated).
template <typename R, bool TH>
__device__ __forceinline__ void flux(const R *q, const NodeState<R, TH> &s, const Mat &m, int d, R *F) {
  R vd = s.v[d];
  R sd0 = s.sig[0 * 3 + d], sd1 = s.sig[1 * 3 + d], sd2 = s.sig[2 * 3 + d];
  F[0] = q[2 + d];
  F[1] = vd * (q[1] + s.p) - (sd0 * q[2] + sd1 * q[3] + sd2 * q[4]) * s.ir;
  F[2] = vd * q[2] - sd0;
  F[3] = vd * q[3] - sd1;
  F[4] = vd * q[4] - sd2;
  F[2 + d] += s.p;
#pragma unroll
  for (int k = 0; k < 9; k++) F[5 + k] = R(0);
#pragma unroll
  for (int k = 0; k < 3; k++)
    F[5 + d + 3 * k] = q[5 + 3 * k] * s.v[0] + q[6 + 3 * k] * s.v[1] + q[7 + 3 * k] * s.v[2];
  if (TH) {
    F[1] += R(m.p.calpha2) * s.T * q[14 + d] * s.ir;
    F[14] = vd * q[14];
    F[15] = vd * q[15];
    F[16] = vd * q[16];
    F[14 + d] += s.T;
  }
}

// The flux components that can be nonzero in direction d (equations.cpp:11-48):
// mass, energy, momentum, column d of the distortion (F[5 + d + 3k]) and the
// thermal rows.  The other six distortion components are zero, so the
// predictor neither stores nor contracts them.
template <bool TH> struct FluxPack {
  static constexpr int NF = TH ? 11 : 8;        // nonzero components
  static constexpr int STRIDE = TH ? 14 : 10;   // LDS doubles per node (16-byte aligned records)
};
__host__ __device__ constexpr int flux_var(int i, int d) { return i < 5 ? i : i < 8 ? 5 + d + 3 * (i - 5) : 6 + i; }

// out -= scale * B_d(q) . x  on the distortion rows (equations.cpp:189-267).
template <typename R>
__device__ __forceinline__ void bdot_sub(const R *v, const R *x, int d, R scale, R *out) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const R *xk = x + 5 + 3 * k;
    R *o = out + 5 + 3 * k;
    if (d == 0) {
      o[0] -= scale * (-v[1] * xk[1] - v[2] * xk[2]);
      o[1] -= scale * (v[0] * xk[1]);
      o[2] -= scale * (v[0] * xk[2]);
    } else {
      o[0] -= scale * (v[1] * xk[0]);
      o[1] -= scale * (-v[0] * xk[0] - v[2] * xk[2]);
      o[2] -= scale * (v[1] * xk[2]);
    }
  }
}
// B_d(q) . x written into out (all V entries)
template <typename R, int V>
__device__ __forceinline__ void bdot(const R *v, const R *x, int d, R *out) {
#pragma unroll
  for (int i = 0; i < V; i++) out[i] = R(0);
  bdot_sub(v, x, d, R(-1), out);
}

// ------------------------------------------------------------------ eigenvalues
// Which root max_abs_eig3 settles on: kind 0 the real root `root` (the value
// is |root|), kind 1 the complex pair beside the real root `root` (the value
// is sqrt(|c0 / root|)); c2, c1, c0 the characteristic polynomial's
// coefficients.
struct Eig3Root {
  int kind;
  double root, c2, c1, c0;
};

// Largest |eigenvalue| of a real 3x3 matrix: roots of the characteristic
// polynomial (trigonometric form when all are real, Cardano otherwise), the
// dominant real root refined by Newton steps on the polynomial.  info (R =
// double only) reports the root taken.
template <typename R> __device__ __forceinline__ R max_abs_eig3(const R *a, Eig3Root *info = nullptr) {
  R c2 = a[0] + a[4] + a[8];
  R c1 = a[0] * a[4] - a[1] * a[3] + a[0] * a[8] - a[2] * a[6] + a[4] * a[8] - a[5] * a[7];
  R c0 = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
              a[2] * (a[3] * a[7] - a[4] * a[6]);
  // lambda^3 - c2 lambda^2 + c1 lambda - c0 ; shift lambda = y + c2/3
  R m = c2 / 3.0;
  R P = c1 - c2 * c2 / 3.0;                                 // y^3 + P y + Qc
  R Qc = -2.0 * m * m * m + c1 * m - c0;
  R disc = (Qc * Qc) / 4.0 + (P * P * P) / 27.0;
  R best;
  if (disc <= 0.0 && c2 >= 0.0 && c1 >= 0.0 && c0 >= 0.0) {
    // three real roots, all >= 0 (signs of the coefficients): Newton from
    // above the largest root decreases monotonically onto it.  The start is
    // the Laguerre-Samuelson bound m + (2/3) sqrt(c2^2 - 3 c1), which every
    // root of a cubic with real roots lies below and which equals the largest
    // root when the other two coincide (an unsheared state's acoustic tensor),
    // raised by a margin for its rounding (the square root keeps only half
    // the digits of c2^2 - 3 c1 when the roots nearly coincide) and capped by
    // the trace: 2-4 steps where starting from the trace took 7-10.
    const R margin = std::is_same<R, float>::value ? R(1e-3) : R(1e-7);
    R x = fmin(c2, m + R(2.0 / 3.0) * sqrt(fmax(c2 * c2 - R(3) * c1, R(0))) + margin * c2);
    for (int it = 0; it < 60; it++) {
      R f = ((x - c2) * x + c1) * x - c0;
      R fp = (3 * x - 2 * c2) * x + c1;
      if (!(fp > 0.0)) break;
      R nx = x - f / fp;
      if (!(nx < x)) break;
      x = nx;
    }
    if constexpr (std::is_same<R, double>::value)
      if (info) *info = {0, x, c2, c1, c0};
    return fabs(x);
  }
  if (disc <= 0.0) {
    R rr = sqrt(fmax(-P / 3.0, 0.0));
    R arg = (rr > 0) ? (-Qc / 2.0) / (rr * rr * rr) : R(0.0);
    arg = fmin(fmax(arg, -1.0), 1.0);
    R th = acos(arg) / 3.0;
    R y0 = 2 * rr * cos(th), y1 = 2 * rr * cos(th - 2.0943951023931955),
           y2 = 2 * rr * cos(th + 2.0943951023931955);
    R l0 = y0 + m, l1 = y1 + m, l2 = y2 + m;
    best = l0;
    if (fabs(l1) > fabs(best)) best = l1;
    if (fabs(l2) > fabs(best)) best = l2;
    for (int it = 0; it < 3; it++) {
      R f = ((best - c2) * best + c1) * best - c0;
      R fp = (3 * best - 2 * c2) * best + c1;
      if (fp == 0.0) break;
      R nb = best - f / fp;
      if (!(fabs(nb - best) > 0)) { best = nb; break; }
      best = nb;
    }
    if constexpr (std::is_same<R, double>::value)
      if (info) *info = {0, best, c2, c1, c0};
    return fabs(best);
  }
  R sq = sqrt(disc);
  R u = cbrt(-Qc / 2.0 + sq), v = cbrt(-Qc / 2.0 - sq);
  R y = u + v;
  R lr = y + m;
  for (int it = 0; it < 3; it++) {
    R f = ((lr - c2) * lr + c1) * lr - c0;
    R fp = (3 * lr - 2 * c2) * lr + c1;
    if (fp == 0.0) break;
    lr = lr - f / fp;
  }
  // complex pair: product of all roots = c0
  R mod2 = (lr != 0.0) ? c0 / lr : R(0.0);
  R mc = sqrt(fabs(mod2));
  if constexpr (std::is_same<R, double>::value)
    if (info) *info = {fabs(lr) >= mc ? 0 : 1, lr, c2, c1, c0};
  return fmax(fabs(lr), mc);
}

// Largest |eigenvalue| of a real n x n matrix (n <= 4): complex QR iteration
// with Wilkinson shifts and deflation.
struct cplx { double re, im; };
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ double cabs_(cplx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cplx cdivr(cplx a, double r) { return {a.re / r, a.im / r}; }
__device__ __forceinline__ cplx csqrt_(cplx z) {
  double r = cabs_(z);
  double sre = sqrt(fmax(0.0, (r + z.re) / 2)), sim = sqrt(fmax(0.0, (r - z.re) / 2));
  return {sre, z.im < 0 ? -sim : sim};
}

// With lam_out, the eigenvalue of largest modulus itself (one of a complex pair).
__device__ double max_abs_eig4(const double *a, cplx *lam_out = nullptr) {
  cplx H[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) H[i][j] = {a[i * 4 + j], 0.0};
  double best = 0;
  cplx lbest = {0.0, 0.0};
  int m = 4, its = 0;
  while (m > 0) {
    if (m == 1) {
      if (cabs_(H[0][0]) > best) lbest = H[0][0];
      best = fmax(best, cabs_(H[0][0]));
      break;
    }
    double s = cabs_(H[m - 1][m - 1]) + cabs_(H[m - 2][m - 2]);
    double off = 0;
    for (int j = 0; j < m - 1; j++) off += cabs_(H[m - 1][j]);
    if (s == 0) {
      for (int i = 0; i < m; i++) for (int j = 0; j < m; j++) s += cabs_(H[i][j]);
    }
    if (off <= 1e-17 * s || off == 0.0) {
      if (cabs_(H[m - 1][m - 1]) > best) lbest = H[m - 1][m - 1];
      best = fmax(best, cabs_(H[m - 1][m - 1]));
      m--;
      its = 0;
      continue;
    }
    if (++its > 200) return __builtin_nan("");
    cplx a11 = H[m - 2][m - 2], a12 = H[m - 2][m - 1], a21 = H[m - 1][m - 2], a22 = H[m - 1][m - 1];
    cplx tr = cadd(a11, a22), det = csub(cmul(a11, a22), cmul(a12, a21));
    cplx h = cdivr(tr, 2.0);
    cplx disc = csqrt_(csub(cmul(h, h), det));
    cplx l1 = cadd(h, disc), l2 = csub(h, disc);
    cplx mu = cabs2(csub(l1, a22)) < cabs2(csub(l2, a22)) ? l1 : l2;
    if (its % 11 == 10) mu.im += 0.7 * s;
    for (int i = 0; i < m; i++) H[i][i] = csub(H[i][i], mu);
    cplx cs[6], sn[6];
    int nr = 0;
    for (int j = 0; j < m - 1; j++)
      for (int i = j + 1; i < m; i++) {
        cplx x = H[j][j], y = H[i][j];
        double r = sqrt(cabs2(x) + cabs2(y));
        cplx c = {1.0, 0.0}, sv = {0.0, 0.0};
        if (r != 0) { c = cdivr(x, r); sv = cdivr(y, r); }
        for (int k = 0; k < m; k++) {
          cplx hj = H[j][k], hi = H[i][k];
          H[j][k] = cadd(cmul(cconj(c), hj), cmul(cconj(sv), hi));
          H[i][k] = csub(cmul(c, hi), cmul(sv, hj));
        }
        cs[nr] = c; sn[nr] = sv; nr++;
      }
    nr = 0;
    for (int j = 0; j < m - 1; j++)
      for (int i = j + 1; i < m; i++) {
        cplx c = cs[nr], sv = sn[nr]; nr++;
        for (int k = 0; k < m; k++) {
          cplx hj = H[k][j], hi = H[k][i];
          H[k][j] = cadd(cmul(hj, c), cmul(hi, sv));
          H[k][i] = csub(cmul(hi, cconj(c)), cmul(hj, cconj(sv)));
        }
      }
    for (int i = 0; i < m; i++) H[i][i] = cadd(H[i][i], mu);
  }
  if (lam_out) *lam_out = lbest;
  return best;
}

// Thermo-acoustic tensor Xi = Xi1 Xi2 (eig.cpp:8-64) into X (n x n, n = 3 or 4),
// from the node state s of q (node_state<R, TH, MK>(q, mm, s)).
template <bool TH, typename R = double, int MK = MK_ANY>
__device__ __forceinline__ int thermo_acoustic_tensor_of(const R *q, const NodeState<R, TH> &s, const Mat &mm, int d,
                                                         R *X) {
  const gpr_par &p = mm.p;
  const R *A = q + 5;
  R r = s.r, cs2 = s.cs2;
  R dcs2 = mk_flag<MK>(MatKind<MK>::beta_zero, mm.beta_zero) ? R(0.0) : p.beta / r * cs2;
  R sd[3] = {s.sig[0 * 3 + d], s.sig[1 * 3 + d], s.sig[2 * 3 + d]};
  // one reciprocal of cs2 for the ten quotients below (within an ulp of
  // dividing each)
  const R ics2 = R(1) / cs2;
  R fac = 1.0 / r + dcs2 * ics2;
  R dsdr[3] = {fac * sd[0], fac * sd[1], fac * sd[2]};
  // dsigma/dA (state.cpp:56-69): M = -rho cs2 (AdevG with col d doubled + A_d G_d/3 + G_dd A)^T
  R G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  R T9[9];
  for (int i = 0; i < 9; i++) T9[i] = s.EA[i] * ics2;  // A dev G
  for (int i = 0; i < 3; i++) T9[i * 3 + d] *= 2.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T9[i * 3 + j] += 1.0 / 3.0 * A[i * 3 + d] * G[d * 3 + j] + G[d * 3 + d] * A[i * 3 + j];
  R M[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[i * 3 + j] = -r * cs2 * T9[j * 3 + i];
  // c0^2 (speeds.cpp:8-18)
  R Gm = Gamma_MG<R, MK>(r, p), dGm = dGamma_MG<R, MK>(r, p), pr = p_ref<R, MK>(r, p), dpr = dp_ref<R, MK>(r, p),
    der = de_ref<R, MK>(r, p);
  R dedr = der - (dpr * r * Gm + (Gm + r * dGm) * (s.p - pr)) / ((r * Gm) * (r * Gm));
  R dC0 = (1 - p.beta / Gm) * dcs2 + p.beta / (Gm * Gm) * dGm * cs2;
  R dEdr = dedr + dC0 / 4 * s.devGsq;
  R dEdp = 1.0 / (r * Gm);
  R c02 = (s.p / (r * r) - dEdr) / dEdp;
  constexpr int n = TH ? 4 : 3;
  R X1[4][5] = {}, X2[5][4] = {};
  X1[0][1] = 1 / r;
  for (int i = 0; i < 3; i++) {
    X1[i][0] = -1 / r * dsdr[i];
    for (int j = 0; j < 3; j++) X1[i][2 + j] = -1 / r * M[i * 3 + j];
  }
  X2[0][0] = r;
  X2[1][d] = r * c02;
  for (int j = 0; j < 3; j++) X2[1][j] += sd[j] - r * dsdr[j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) X2[2 + i][j] = A[i * 3 + j];
  if (TH) {
    R dTdp = dEdp / p.cv;
    X1[3][0] = -(dpr * r * Gm + (Gm + r * dGm) * (s.p - pr)) / ((r * Gm) * (r * Gm)) / p.cv / r;
    X1[3][1] = dTdp / r;
    R ch = sqrt(p.calpha2 * s.T / p.cv) / r;  // c_h (speeds.cpp:20-25): NaN for T < 0 as there
    X2[1][3] = r * ch * ch / dTdp;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      R acc = 0;
      for (int k = 0; k < 5; k++) acc += X1[i][k] * X2[k][j];
      X[i * n + j] = acc;
    }
  return n;
}

template <bool TH, typename R = double>
__device__ __forceinline__ int thermo_acoustic_tensor(const R *q, const Mat &mm, int d, R *X) {
  NodeState<R, TH> s;
  node_state<R, TH>(q, mm, s);
  return thermo_acoustic_tensor_of<TH, R>(q, s, mm, d, X);
}

// max |characteristic speed| in direction d (eig.cpp:66-77), from the node
// state s of q
template <bool TH, typename R = double, int MK = MK_ANY>
__device__ __forceinline__ R max_abs_eigs_of(const R *q, const NodeState<R, TH> &s, const Mat &mm, int d) {
  R X[16];
  thermo_acoustic_tensor_of<TH, R, MK>(q, s, mm, d, X);
  R lam2;
  if constexpr (TH)
    lam2 = max_abs_eig4(X);
  else
    lam2 = max_abs_eig3<R>(X);
  R lam = sqrt(lam2);
  R vd = q[2 + d] / q[0];
  return vd > 0 ? vd + lam : lam - vd;
}
template <bool TH, typename R = double>
__device__ __forceinline__ R max_abs_eigs(const R *q, const Mat &mm, int d) {
  NodeState<R, TH> s;
  node_state<R, TH>(q, mm, s);
  return max_abs_eigs_of<TH, R>(q, s, mm, d);
}

// The CFL speed of a cell (timestep, iterator.cpp:58-72): the sum over the
// ND directions of max |characteristic speed| / h, from one node state.
template <bool TH, int ND, typename R = double, int MK = MK_ANY>
__device__ __forceinline__ R cfl_speed(const R *q, const Mat &mm, R dx, R dy) {
  NodeState<R, TH> s;
  node_state<R, TH, MK>(q, mm, s);
  R sp = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) sp += max_abs_eigs_of<TH, R, MK>(q, s, mm, d) / (d == 0 ? dx : dy);
  return sp;
}

}  // namespace gpr
