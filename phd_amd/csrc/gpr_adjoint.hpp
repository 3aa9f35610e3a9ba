// Adjoint (reverse mode) of one plain FP64 ADER-WENO step in 2D, V = 14 or
// 17 (heat conduction), with the step's dt held fixed (the time steps of the
// forward run).  With lambda the gradient of a loss with respect to u_{n+1},
// the kernels below give its gradient with respect to u_n and accumulate its
// derivatives with respect to the material's relaxation parameter (mu of a
// Newtonian fluid, sigmaY of a power-law solid) and, with heat conduction,
// the conductivity kappa.  Order, the reverse of the step:
//   ka_renorm   u_{n+1} = renorm(u_fv)                    (utils.cpp:6-23)
//   ka_faces    u_fv -= face terms of qh                  (fv.cpp:162-268)
//   ka_cells    u_fv += cell terms of qh; gathers faces  (fv.cpp:56-109)
//   ka_predictor  qh solves the DG equations R(qh; wh, parameters) = 0
//               (dg2.cpp): (dR/dq)^T m = lambda_qh by a transposed
//               fixed-point iteration, then lambda_wh = -(dR/dwh)^T m
//   ka_weno_y, ka_weno_x   wh = WENO_y(WENO_x(ub))         (weno.cpp:83-124)
//   ka_boundaries          ub from u (grid.cpp:147-170), plus the identity
//                          path u_fv = u_n + ...
// Local derivatives of the node physics come from the hand-written reverse
// mode node_vjp, or (DUAL) from forward-mode dual numbers (gpr_dual.hpp) on
// the same templated physics the forward step runs; the wave speed's come
// from dual numbers (heat conduction: Jacobi's formula on the 4x4 tensor's
// eigenvalue).  The predictor is differentiated at its solution (implicit
// function theorem); the forward run that a gradient is checked against
// solves it tightly (dg_tol 1e-13).
#pragma once

#include "gpr_dual.hpp"
#include "gpr_kernels.hpp"

namespace gpr {

struct AdjArgs {
  const double *u_fv;   // [cells][V] state after FV, before renormalisation
  const double *lam1;   // [cells][V] d loss / d u_{n+1}
  double *lam_fv;       // [cells][V]
  double *lam_face_x;   // [(nx+1) ny][N][NSF][2][V]: adjoints of qL, qR at each x face node
  double *lam_face_y;   // [nx (ny+1)][N][NSF][2][V]
  const double *qh;     // [(nx+2)(ny+2)][NN][V]
  const double *wh;     // [(nx+2)(ny+2)][NS][V] WENO states (the stiff preconditioner's)
  double *lam_qh;       // [(nx+2)(ny+2)][NN][V]
  double *lam_wh;       // [(nx+2)(ny+2)][NS][V]
  const double *ux;     // WENO x-pass output [(nx+2)(ny+2N)][N][V]
  const double *ub;     // [(nx+2N)(ny+2N)][V]
  double *lam_ux;       // like ux
  double *lam_ub;       // like ub
  double *lam0;         // [cells][V] d loss / d u_n (output)
  const uint8_t *mask;  // [(nx+2)(ny+2)]
  double *gmu;          // d loss / d (the material parameter), accumulated
  double *gkap;         // d loss / d kappa (heat conduction), accumulated
  int *fail;            // cells whose transposed Picard iteration did not converge
  unsigned long long *itc;  // ITC_SLOTS x ITC_STRIDE words: [0] transposed predictor iterations, [1] cells,
                            // [2] cells in the preconditioned form
  double dt, dx, dy;
  double par_fac;       // d S_A / d parameter = par_fac S_A: -1/mu (Newtonian), -n/sigmaY (power-law solid)
  double kappa;         // heat conductivity: d S_J / d kappa = -S_J / kappa
  int nx, ny;
  int it_plain;  // transposed Picard iterations before the preconditioned form
  int start_pre; // 1: every cell starts in the preconditioned form
  int it_max;    // iterations of the preconditioned form
  double tol;
  int4 bc;
  Mat m;
};

// Reverse-mode derivative of one node's terms for the materials the adjoint
// covers (V = 14, no thermal terms, any EOS; theta1inv of a Newtonian fluid
// or of a solid, elastic or power-law plastic):
//   out[j] = d/dq_j [ ah . (S(q) - B_x(v) dqx - B_y(v) dqy)     (H)
//                   + aF . F_x(q)                               (FX)
//                   + aG . F_y(q) ]                             (FY)
// with the weights and dq held fixed.  It recomputes the forward quantities
// of node_state / source / flux (gpr_device.hpp) and runs them backwards;
// the EOS scalars of rho (c_s^2, Gamma, e_ref, p_ref) take their derivatives
// from one dual-number evaluation.  The stress and dev G are symmetric
// functions of q, so their full 3x3 formulas are differentiated (the forward
// code fills one triangle and mirrors it, which is the same function of q).
template <bool TH, bool H, bool FX, bool FY>
__device__ __forceinline__ void node_vjp(const double *q, const double *dqx, const double *dqy, const Mat &m,
                                         const double *ah, const double *aF, const double *aG, double *out) {
  constexpr int V = Dims<TH>::V;
  const gpr_par &p = m.p;
  // ---- forward
  const double r = q[0], ir = 1.0 / r;
  const double v[3] = {q[2] * ir, q[3] * ir, q[4] * ir};
  const double *A = q + 5;
  double G[9], D[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G[i * 3 + j] = A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j];
  const double tr3 = (G[0] + G[4] + G[8]) / 3.0;
#pragma unroll
  for (int k = 0; k < 9; k++) D[k] = G[k] - (k % 4 == 0 ? tr3 : 0.0);
  double devGsq = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) devGsq += D[k] * D[k];
  const Dual rd(r, 1.0);
  const Dual cs2d = c_s2(rd, m), Gmd = Gamma_MG(rd, p), erd = e_ref(rd, p), prd = p_ref(rd, p);
  const double cs2 = cs2d.v, Gm = Gmd.v;
  double AD[9], EA[9], sig[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      AD[i * 3 + j] = A[i * 3 + 0] * D[0 * 3 + j] + A[i * 3 + 1] * D[1 * 3 + j] + A[i * 3 + 2] * D[2 * 3 + j];
      EA[i * 3 + j] = cs2 * AD[i * 3 + j];
    }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      sig[i * 3 + j] = -r * (EA[0 * 3 + i] * A[0 * 3 + j] + EA[1 * 3 + i] * A[1 * 3 + j] + EA[2 * 3 + i] * A[2 * 3 + j]);
  const double C0 = (1.0 - p.beta / Gm) * cs2;
  double Jv[3] = {0.0, 0.0, 0.0};  // J = (rho J) / rho
  if constexpr (TH)
#pragma unroll
    for (int i = 0; i < 3; i++) Jv[i] = q[14 + i] * ir;
  const double e = q[1] * ir - 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - 0.25 * C0 * devGsq -
                   (TH ? 0.5 * p.calpha2 * (Jv[0] * Jv[0] + Jv[1] * Jv[1] + Jv[2] * Jv[2]) : 0.0);
  const double pr = (e - erd.v) * r * Gm + prd.v;
  // temperature T(rho, p) (mg.cpp) and its partial derivatives
  double Tt = 0, dTdr = 0, dTdp = 0, Tb = 0;
  if constexpr (TH) {
    const Dual Td = temperature_mg(rd, Dual(pr, 0.0), p);
    Tt = Td.v;
    dTdr = Td.d;
    dTdp = 1.0 / (r * Gm * p.cv);
  }
  // ---- reverse
  double qb[V], vb[3] = {0, 0, 0}, sigb[9], EAb[9], Ab[9];
#pragma unroll
  for (int k = 0; k < V; k++) qb[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    sigb[k] = 0.0;
    EAb[k] = 0.0;
    Ab[k] = 0.0;
  }
  double pb = 0, irb = 0, rb = 0, thb = 0, cs2b = 0;
  if constexpr (H) {
    // S_A = -EA theta1inv (sources.cpp:13-81):
    //   Newtonian fluid   theta1inv = det(A)^(5/3) mu_fac
    //   power-law solid   theta1inv = 3 det(A)^(5/3) / (cs2 tau0) (min(|sigma'|, 1e8) / sigmaY)^n
    //   elastic solid     theta1inv = 0
    const double dA = det3(A), c13 = cbrt(dA), A53 = dA * c13 * c13;
    double th = 0, th_A53 = 0, th_cs2 = 0, th_sn = 0, sn = 0;
    if (p.SOLID) {
      if (p.POWER_LAW) {
        const double a1 = sig[0] - sig[4], b1 = sig[4] - sig[8], c1 = sig[8] - sig[0];
        sn = sqrt(0.5 * (a1 * a1 + b1 * b1 + c1 * c1) + 3.0 * (sig[1] * sig[1] + sig[5] * sig[5] + sig[6] * sig[6]));
        const double base = 3.0 / (cs2 * p.tau0), pw = pow(fmin(sn, 1e8) / p.sigmaY, p.n);
        th = A53 * base * pw;
        th_A53 = base * pw;
        th_cs2 = -th / cs2;
        if (sn < 1e8) th_sn = sn > 0 ? p.n * th / sn : (p.n == 1.0 ? A53 * base / p.sigmaY : 0.0);
      }
    } else {
      th = A53 * m.mu_fac;
      th_A53 = m.mu_fac;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
      EAb[k] -= ah[5 + k] * th;
      thb -= ah[5 + k] * EA[k];
    }
    cs2b += thb * th_cs2;
    if (th_sn != 0.0 && sn > 0) {
      // |sigma'|^2 = ((s00 - s11)^2 + (s11 - s22)^2 + (s22 - s00)^2) / 2 + 3 (s01^2 + s12^2 + s20^2)
      const double g = thb * th_sn / (2.0 * sn);
      const double a1 = sig[0] - sig[4], b1 = sig[4] - sig[8], c1 = sig[8] - sig[0];
      sigb[0] += g * (a1 - c1);
      sigb[4] += g * (b1 - a1);
      sigb[8] += g * (c1 - b1);
      sigb[1] += g * 6.0 * sig[1];
      sigb[5] += g * 6.0 * sig[5];
      sigb[6] += g * 6.0 * sig[6];
    }
    const double f = thb * th_A53 * (5.0 / 3.0) * c13 * c13;  // d theta / d det(A)
    Ab[0] += f * (A[4] * A[8] - A[5] * A[7]);
    Ab[1] += f * (A[5] * A[6] - A[3] * A[8]);
    Ab[2] += f * (A[3] * A[7] - A[4] * A[6]);
    Ab[3] += f * (A[2] * A[7] - A[1] * A[8]);
    Ab[4] += f * (A[0] * A[8] - A[2] * A[6]);
    Ab[5] += f * (A[1] * A[6] - A[0] * A[7]);
    Ab[6] += f * (A[1] * A[5] - A[2] * A[4]);
    Ab[7] += f * (A[2] * A[3] - A[0] * A[5]);
    Ab[8] += f * (A[0] * A[4] - A[1] * A[3]);
    // -B_d(v) dq_d is linear in v: its v_l coefficient is -B_d(e_l) dq_d
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const double *x = d == 0 ? dqx : dqy;
      if (x == nullptr) continue;
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const double el[3] = {l == 0 ? 1.0 : 0.0, l == 1 ? 1.0 : 0.0, l == 2 ? 1.0 : 0.0};
        double o[14];
#pragma unroll
        for (int k = 0; k < 14; k++) o[k] = 0.0;
        bdot_sub(el, x, d, 1.0, o);
        double acc = 0;
#pragma unroll
        for (int k = 5; k < 14; k++) acc += ah[k] * o[k];
        vb[l] += acc;
      }
    }
  }
  if constexpr (H && TH) {
    // S_J = -rho (c_alpha^2 J) T / (kappa rho) = -(c_alpha^2 / kappa) (rho J) T / rho
    const double c = -p.calpha2 / p.kappa;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      qb[14 + i] += ah[14 + i] * c * Tt * ir;
      Tb += ah[14 + i] * c * q[14 + i] * ir;
      irb += ah[14 + i] * c * q[14 + i] * Tt;
    }
  }
  // fluxes (flux in gpr_device.hpp)
#pragma unroll
  for (int d = 0; d < 2; d++) {
    if ((d == 0 && !FX) || (d == 1 && !FY)) continue;
    const double *a = d == 0 ? aF : aG;
    const double vd = v[d];
    qb[2 + d] += a[0];
    // F1 = vd (q1 + p) - (sum_i sig_id q_{2+i}) / rho
    double sq = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) sq += sig[i * 3 + d] * q[2 + i];
    vb[d] += a[1] * (q[1] + pr);
    qb[1] += a[1] * vd;
    pb += a[1] * vd;
    irb -= a[1] * sq;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      sigb[i * 3 + d] -= a[1] * q[2 + i] * ir;
      qb[2 + i] -= a[1] * sig[i * 3 + d] * ir;
      // F_{2+i} = vd q_{2+i} - sig_id (+ p on i = d)
      vb[d] += a[2 + i] * q[2 + i];
      qb[2 + i] += a[2 + i] * vd;
      sigb[i * 3 + d] -= a[2 + i];
    }
    pb += a[2 + d];
    // F_{5+d+3k} = A row k . v
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double w = a[5 + d + 3 * k];
#pragma unroll
      for (int l = 0; l < 3; l++) {
        Ab[3 * k + l] += w * v[l];
        vb[l] += w * A[3 * k + l];
      }
    }
    if constexpr (TH) {
      // F1 += c_alpha^2 T (rho J_d) / rho; F_{14+i} = v_d (rho J_i); F_{14+d} += T
      Tb += a[1] * p.calpha2 * q[14 + d] * ir;
      qb[14 + d] += a[1] * p.calpha2 * Tt * ir;
      irb += a[1] * p.calpha2 * Tt * q[14 + d];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        vb[d] += a[14 + i] * q[14 + i];
        qb[14 + i] += a[14 + i] * vd;
      }
      Tb += a[14 + d];
    }
  }
  // T = T(rho, p)
  if constexpr (TH) {
    pb += Tb * dTdp;
    rb += Tb * dTdr;
  }
  // p = (e - e_ref) rho Gamma + p_ref
  const double eb = pb * r * Gm;
  double Gmb = pb * (e - erd.v) * r;
  rb += pb * ((e - erd.v) * Gm - erd.d * r * Gm + prd.d);
  // e = q1 / rho - |v|^2 / 2 - C0 devGsq / 4,  C0 = (1 - beta / Gamma) cs2
  qb[1] += eb * ir;
  irb += eb * q[1];
#pragma unroll
  for (int i = 0; i < 3; i++) vb[i] -= eb * v[i];
  if constexpr (TH)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      qb[14 + i] -= eb * p.calpha2 * Jv[i] * ir;
      irb -= eb * p.calpha2 * Jv[i] * q[14 + i];
    }
  const double C0b = -0.25 * eb * devGsq, devGsqb = -0.25 * eb * C0;
  Gmb += C0b * p.beta / (Gm * Gm) * cs2;
  cs2b += C0b * (1.0 - p.beta / Gm);
  // sig = -rho EA^T A
  double Db[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    rb += sigb[k] * sig[k] * ir;
    Db[k] = 2.0 * devGsqb * D[k];
  }
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double e1 = 0, a1 = 0;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        e1 += A[k * 3 + j] * sigb[i * 3 + j];
        a1 += EA[k * 3 + j] * sigb[j * 3 + i];
      }
      EAb[k * 3 + i] -= r * e1;  // d/dEA_ki of -rho sum_j EA_ki A_kj sigb_ij
      Ab[k * 3 + i] -= r * a1;   // d/dA_ki of -rho sum_j EA_kj A_ki sigb_ji
    }
  // EA = cs2 A D
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cs2b += EAb[i * 3 + j] * AD[i * 3 + j];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        Ab[i * 3 + k] += cs2 * EAb[i * 3 + j] * D[k * 3 + j];
        Db[k * 3 + j] += cs2 * A[i * 3 + k] * EAb[i * 3 + j];
      }
    }
  // D = G - tr(G)/3 I, G = A^T A
  const double tb = (Db[0] + Db[4] + Db[8]) / 3.0;
  double Gb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Gb[k] = Db[k] - (k % 4 == 0 ? tb : 0.0);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double acc = 0;
#pragma unroll
      for (int j = 0; j < 3; j++) acc += (Gb[i * 3 + j] + Gb[j * 3 + i]) * A[k * 3 + j];
      Ab[k * 3 + i] += acc;
    }
  rb += cs2b * cs2d.d + Gmb * Gmd.d;
  // v = (rho v) / rho, 1 / rho
#pragma unroll
  for (int i = 0; i < 3; i++) {
    qb[2 + i] += vb[i] * ir;
    irb += vb[i] * q[2 + i];
  }
  rb -= irb * ir * ir;
  qb[0] += rb;
#pragma unroll
  for (int k = 0; k < 9; k++) qb[5 + k] += Ab[k];
#pragma unroll
  for (int k = 0; k < V; k++) out[k] = qb[k];
}

// Renormalisation of A (k_post / utils.cpp:6-23) on a dual state
template <typename R> __device__ __forceinline__ void renorm_state(R *q, double rho0) {
  if (q[0] > 0) {
    R cf = cbrt(q[0] / (rho0 * det3(q + 5)));
#pragma unroll
    for (int k = 0; k < 9; k++) q[5 + k] *= cf;
  }
}

// one thread per cell: lam_fv = (d renorm / d u_fv)^T lam1
template <bool TH>
__global__ void ka_renorm(AdjArgs a, long ncell) {
  constexpr int V = Dims<TH>::V;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    const double *u = a.u_fv + c * V;
    const double *l = a.lam1 + c * V;
    double out[V];
#pragma unroll
    for (int v = 0; v < V; v++) out[v] = l[v];  // rho E and rho v pass through; A and rho below
    for (int j = 0; j < 14; j++) {
      if (j >= 1 && j < 5) continue;
      Dual q[V];
#pragma unroll
      for (int v = 0; v < V; v++) q[v] = Dual(u[v], v == j ? 1.0 : 0.0);
      renorm_state(q, a.m.p.rho0);
      double g = j == 0 ? l[0] : 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) g += l[5 + k] * q[5 + k].d;
      out[j] = g;
    }
#pragma unroll
    for (int v = 0; v < V; v++) a.lam_fv[c * V + v] = out[v];
  }
}

// (B_d(v))^T w on the distortion rows: out[j] = w . (B_d e_j), j = 5..13
template <int V = 14>
__device__ __forceinline__ void bdot_T(const double *vel, const double *w, int d, double *out) {
  for (int j = 0; j < V; j++) out[j] = 0.0;
#pragma unroll
  for (int j = 5; j < 14; j++) {
    double e[14], o[14];
#pragma unroll
    for (int v = 0; v < 14; v++) {
      e[v] = v == j ? 1.0 : 0.0;
      o[v] = 0.0;
    }
    bdot_sub(vel, e, d, -1.0, o);  // o = B_d e_j (distortion rows only)
    double s = 0;
#pragma unroll
    for (int v = 5; v < 14; v++) s += w[v] * o[v];
    out[j] = s;
  }
}

// adj(M) of a complex 4x4 matrix (row-major):
// adj(M)_rc = (-1)^(r+c) det(M without row c, column r)
__device__ __forceinline__ cplx cdet3(const cplx *a) {
  const cplx m0 = csub(cmul(a[4], a[8]), cmul(a[5], a[7]));
  const cplx m1 = csub(cmul(a[3], a[8]), cmul(a[5], a[6]));
  const cplx m2 = csub(cmul(a[3], a[7]), cmul(a[4], a[6]));
  return cadd(csub(cmul(a[0], m0), cmul(a[1], m1)), cmul(a[2], m2));
}
__device__ __forceinline__ void adjugate4(const cplx *M, cplx *adj) {
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      cplx sub[9];
      int k = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (i == c) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (j == r) continue;
          sub[k++] = M[i * 4 + j];
        }
      }
      const cplx d = cdet3(sub);
      adj[r * 4 + c] = ((r + c) % 2) ? cplx{-d.re, -d.im} : d;
    }
}

// Gradient of max_abs_eigs<true> (the heat-conducting 4x4 acoustic tensor X,
// whose eigenvalue of largest modulus max_abs_eig4 finds by QR iteration).
// That eigenvalue lam (real, or one of a complex pair: the y tensor has such
// pairs) is a simple root of det(lam I - X), so by Jacobi's formula
//   d lam = sum_ik adj(lam I - X)_ki dX_ik / tr adj(lam I - X)
// in complex arithmetic, and d|lam| = Re(conj(lam) d lam) / |lam|, with dX
// from dual numbers through thermo_acoustic_tensor.  The speed is
// sqrt(|lam|) plus |v_D| (eig.cpp:66-77).
template <int D>
__device__ __forceinline__ void speed_grad_th(const double *q, const Mat &m, double *out) {
  constexpr int V = 17;
  double X[16];
  thermo_acoustic_tensor<true, double>(q, m, D, X);
  cplx lam;
  const double alam = max_abs_eig4(X, &lam);
  cplx M[16], adj[16];
#pragma unroll
  for (int k = 0; k < 16; k++) M[k] = {(k % 5 == 0 ? lam.re : 0.0) - X[k], k % 5 == 0 ? lam.im : 0.0};
  adjugate4(M, adj);
  const cplx tr = cadd(cadd(adj[0], adj[5]), cadd(adj[10], adj[15]));
  const double tr2 = cabs2(tr);
  const double sq = sqrt(alam), vd = q[2 + D] / q[0];
  for (int j = 0; j < V; j++) {
    Dual x[V], Xd[16];
#pragma unroll
    for (int v = 0; v < V; v++) x[v] = Dual(q[v], v == j ? 1.0 : 0.0);
    thermo_acoustic_tensor<true, Dual>(x, m, D, Xd);
    cplx num = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int k = 0; k < 4; k++) {
        num.re += adj[k * 4 + i].re * Xd[i * 4 + k].d;
        num.im += adj[k * 4 + i].im * Xd[i * 4 + k].d;
      }
    const cplx dl = cdivr(cmul(num, cconj(tr)), tr2);           // d lam
    const double dabs = (lam.re * dl.re + lam.im * dl.im) / alam;  // d |lam|
    const double dv = (j == 2 + D ? 1.0 / q[0] : 0.0) - (j == 0 ? vd / q[0] : 0.0);
    out[j] = dabs / (2.0 * sq) + (vd > 0 ? dv : -dv);
  }
}

// d s / d q_j, j < 14, of the wave speed s = max_abs_eigs<false>(q, m, D)
// without differentiating the root iteration: at the root r of
// p(x) = x^3 - c2 x^2 + c1 x - c0 that max_abs_eig3 settles on,
// dr = (r^2 dc2 - r dc1 + dc0) / p'(r), the coefficients' derivatives taken
// from the thermo-acoustic tensor in dual numbers (one seed per input); a
// complex pair's modulus sqrt(|c0 / r|) follows from r's.
template <int D>
__device__ __forceinline__ void speed_grad(const double *q, const Mat &m, double *out) {
  constexpr int V = 14;
  double X[9];
  thermo_acoustic_tensor<false, double>(q, m, D, X);
  Eig3Root e;
  const double lam2 = max_abs_eig3<double>(X, &e);
  const double r = e.root, dp = (3 * r - 2 * e.c2) * r + e.c1;
  const double mod2 = e.kind ? e.c0 / r : 0.0;
  const double sq = sqrt(lam2), vd = q[2 + D] / q[0];
  for (int j = 0; j < V; j++) {
    Dual x[V], Xd[9];
#pragma unroll
    for (int v = 0; v < V; v++) x[v] = Dual(q[v], v == j ? 1.0 : 0.0);
    thermo_acoustic_tensor<false, Dual>(x, m, D, Xd);
    const Dual c2 = Xd[0] + Xd[4] + Xd[8];
    const Dual c1 = Xd[0] * Xd[4] - Xd[1] * Xd[3] + Xd[0] * Xd[8] - Xd[2] * Xd[6] + Xd[4] * Xd[8] - Xd[5] * Xd[7];
    const Dual c0 = Xd[0] * (Xd[4] * Xd[8] - Xd[5] * Xd[7]) - Xd[1] * (Xd[3] * Xd[8] - Xd[5] * Xd[6]) +
                    Xd[2] * (Xd[3] * Xd[7] - Xd[4] * Xd[6]);
    const double dr = (r * r * c2.d - r * c1.d + c0.d) / dp;
    double dlam2;
    if (e.kind == 0) {
      dlam2 = r >= 0 ? dr : -dr;
    } else {
      const double dmod2 = (c0.d - mod2 * dr) / r;
      dlam2 = (mod2 >= 0 ? dmod2 : -dmod2) / (2.0 * sqrt(fabs(mod2)));
    }
    const double dv = (j == 2 + D ? 1.0 / q[0] : 0.0) - (j == 0 ? vd / q[0] : 0.0);
    out[j] = dlam2 / (2.0 * sq) + (vd > 0 ? dv : -dv);
  }
}

// The transposed Riemann terms of one face node in direction D.  With
// f = smax (qL - qR) + F(qL) + F(qR), smax = max(s(qL), s(qR)) (rusanov) and
// b = Bint(qL, qR), out[j] = sum_v ls[v] db_v/dx_j + ld[v] df_v/dx_j for the
// 2V inputs x = (qL, qR).  A seed on one side moves only that side's flux and
// wave speed, and smax only through the side the maximum takes (a tie takes
// the left one, as rusanov does), so each side's flux is differentiated with
// its own V seeds and one wave speed with V seeds.
template <int N, int D, bool DUAL, bool TH>
__device__ __forceinline__ void riemann_T(const double *qL, const double *qR, const double *ls, const double *ld,
                                          const Mat &m, const Basis &K, double *out) {
  constexpr int V = Dims<TH>::V;
  const double sL = max_abs_eigs<TH, double>(qL, m, D), sR = max_abs_eigs<TH, double>(qR, m, D);
  const bool Lmax = !(sL < sR);
  const double smax = Lmax ? sL : sR;
  double dd = 0;  // ld . (qL - qR), the weight of d smax
#pragma unroll
  for (int v = 0; v < V; v++) dd += ld[v] * (qL[v] - qR[v]);
  for (int side = 0; side < 2; side++) {
    const double *qs = side == 0 ? qL : qR;
    const bool with_speed = side == 0 ? Lmax : !Lmax;
    double fl[V];  // ld . dF(qs)/dq_j
    if constexpr (!DUAL) node_vjp<TH, false, D == 0, D == 1>(qs, nullptr, nullptr, m, nullptr, ld, ld, fl);
    double sg[V];  // the wave speed's gradient (the dual-number form differentiates it directly)
    if constexpr (TH) {
      if (with_speed) speed_grad_th<D>(qs, m, sg);
    } else if constexpr (!DUAL) {
      if (with_speed) speed_grad<D>(qs, m, sg);
    }
    for (int j = 0; j < V; j++) {
      Dual x[V];
#pragma unroll
      for (int v = 0; v < V; v++) x[v] = Dual(qs[v], v == j ? 1.0 : 0.0);
      double g = side == 0 ? smax * ld[j] : -smax * ld[j];
      if constexpr (DUAL) {
        Dual F[V];
        NodeState<Dual, TH> st;
        node_state<Dual, TH>(x, m, st);
        flux<Dual, TH>(x, st, m, D, F);
#pragma unroll
        for (int v = 0; v < V; v++) g += ld[v] * F[v].d;
      } else {
        g += fl[j];
      }
      if (with_speed) {
        if constexpr (TH || !DUAL)
          g += dd * sg[j];
        else
          g += dd * max_abs_eigs<false, Dual>(x, m, D).d;
      }
      out[side * V + j] = g;
    }
  }
  if constexpr (DUAL) {
    for (int j = 0; j < 2 * V; j++) {
      Dual xL[V], xR[V], b[V];
#pragma unroll
      for (int v = 0; v < V; v++) {
        xL[v] = Dual(qL[v], j == v ? 1.0 : 0.0);
        xR[v] = Dual(qR[v], j == V + v ? 1.0 : 0.0);
      }
      bint<N, TH, Dual>(xL, xR, D, K, b);
      double g = 0;
#pragma unroll
      for (int v = 0; v < V; v++) g += ls[v] * b[v].d;
      out[j] += g;
    }
  } else {
    // b = sum_i w_i B_D(v_i) dq with dq = qR - qL and v_i = m_i / r_i on the
    // straight path (bint): linear in dq, and linear in v_i, whose momentum
    // and density are (1 - x_i) qL + x_i qR
    double dq[V];
#pragma unroll
    for (int v = 0; v < V; v++) dq[v] = qR[v] - qL[v];
    double cv[3];  // ls . B_D(e_c) dq
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double ec[3] = {c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, c == 2 ? 1.0 : 0.0};
      double o[V];
#pragma unroll
      for (int v = 0; v < V; v++) o[v] = 0.0;
      bdot_sub(ec, dq, D, -1.0, o);
      double acc = 0;
#pragma unroll
      for (int v = 5; v < V; v++) acc += ls[v] * o[v];
      cv[c] = acc;
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double xi = K.x[i], wi = K.w[i];
      const double r = qL[0] + xi * (qR[0] - qL[0]);
      double vv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) vv[c] = (qL[2 + c] + xi * (qR[2 + c] - qL[2 + c])) / r;
      double gdq[V];  // w_i B_D(v_i)^T ls
      bdot_T<V>(vv, ls, D, gdq);
      double rbar = 0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double mbar = wi * cv[c] / r;
        out[2 + c] += (1.0 - xi) * mbar;
        out[V + 2 + c] += xi * mbar;
        rbar -= mbar * vv[c];
      }
      out[0] += (1.0 - xi) * rbar;
      out[V] += xi * rbar;
#pragma unroll
      for (int v = 5; v < V; v++) {
        out[v] -= wi * gdq[v];
        out[V + v] += wi * gdq[v];
      }
    }
  }
}

// Faces: one lane per (face, time node, face node), as k_fv_faces.  For the
// face's lo = k (b + f) (taken from its left cell) and hi = k (b - f) (from its
// right cell), the adjoints of qL and qR.
template <int N, int ND, bool DUAL, bool TH>
__global__ __launch_bounds__(64) void ka_faces(AdjArgs a) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = ND == 1 ? N : N * N;
  constexpr int NN = NS * N;
  constexpr int NSF = ND == 2 ? N : 1;
  constexpr int NT = N * NSF;
  constexpr int FPW = WAVE / NT;
  const Basis &K = basis_dev<N>();
  const int nx = a.nx, ny = ND == 2 ? a.ny : 1;
  const long nfx = (long)(nx + 1) * ny;
  const long nfy = ND == 2 ? (long)nx * (ny + 1) : 0;
  const int lane = threadIdx.x;
  const int fl = lane / NT, r = lane - fl * NT;
  const int t = r / NSF, sf = r - t * NSF;
  const long id = (long)blockIdx.x * FPW + fl;
  if (fl >= FPW || id >= nfx + nfy) return;
  const int d = id < nfx ? 0 : 1;
  long e0, e1, cl = -1, cr = -1;  // extended cells, real cells on either side (-1: none)
  double *out;
  if (d == 0) {
    int i = id / ny, j = id % ny;
    e0 = ND == 2 ? (long)i * (ny + 2) + (j + 1) : i;
    e1 = ND == 2 ? (long)(i + 1) * (ny + 2) + (j + 1) : i + 1;
    if (i >= 1) cl = (long)(i - 1) * ny + j;
    if (i < nx) cr = (long)i * ny + j;
    out = a.lam_face_x + ((id * N + t) * NSF + sf) * 2 * V;
  } else {
    long k = id - nfx;
    int i = k / (ny + 1), j = k % (ny + 1);
    e0 = (long)(i + 1) * (ny + 2) + j;
    e1 = e0 + 1;
    if (j >= 1) cl = (long)i * ny + (j - 1);
    if (j < ny) cr = (long)i * ny + j;
    out = a.lam_face_y + ((k * N + t) * NSF + sf) * 2 * V;
  }
  if (!(a.mask[e0] && a.mask[e1])) {
    for (int v = 0; v < 2 * V; v++) out[v] = 0.0;
    return;
  }
  const double h = d == 0 ? a.dx : a.dy;
  const double kk = a.dt * K.w[t] / (2.0 * h) * (ND == 2 ? K.w[sf] : 1.0);
  double qL[V], qR[V];
  const double *pl = a.qh + (e0 * NN + t * NS) * V;
  const double *pr = a.qh + (e1 * NN + t * NS) * V;
#pragma unroll
  for (int v = 0; v < V; v++) {
    double l = 0, rr = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      int nl = ND == 1 ? k : (d == 0 ? k * N + sf : sf * N + k);
      l += K.endv[1][k] * pl[nl * V + v];
      rr += K.endv[0][k] * pr[nl * V + v];
    }
    qL[v] = l;
    qR[v] = rr;
  }
  // u_fv[cl] -= lo, u_fv[cr] -= hi
  double ls[V], ld[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    double lL = cl >= 0 ? a.lam_fv[cl * V + v] : 0.0, lR = cr >= 0 ? a.lam_fv[cr * V + v] : 0.0;
    ls[v] = -kk * (lL + lR);  // weight of db
    ld[v] = -kk * (lL - lR);  // weight of df
  }
  double g[2 * V];
  if (d == 0)
    riemann_T<N, 0, DUAL, TH>(qL, qR, ls, ld, a.m, K, g);
  else
    riemann_T<N, 1, DUAL, TH>(qL, qR, ls, ld, a.m, K, g);
  for (int j = 0; j < 2 * V; j++) out[j] = g[j];
}

// The cell terms of one node, h = S(q) - B_x(q) dqx - B_y(q) dqy (dq fixed),
// as k_fv_cells and the predictor's right-hand side form them.
template <typename R, int ND, bool TH>
__device__ __forceinline__ void node_h(const R *q, const double *dqx, const double *dqy, const Mat &m, R *h) {
  constexpr int V = Dims<TH>::V;
  NodeState<R, TH> st;
  node_state<R, TH>(q, m, st);
  source<R, TH>(q, st, m, h);
  R dq[V];
#pragma unroll
  for (int v = 0; v < V; v++) dq[v] = R(dqx[v]);
  bdot_sub(st.v, dq, 0, R(1.0), h);
  if (ND == 2) {
#pragma unroll
    for (int v = 0; v < V; v++) dq[v] = R(dqy[v]);
    bdot_sub(st.v, dq, 1, R(1.0), h);
  }
}


// Cells: one lane per space-time node of every extended cell, as k_predictor's
// layout.  lam_qh = (cell terms)^T lam_fv for real cells + face extrapolation
// adjoints; the parameter derivative of the cell terms into gmu.
template <int N, int ND, bool DUAL, bool TH>
__global__ __launch_bounds__(64) void ka_cells(AdjArgs a, long ncell_ext) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = PredShape<N, ND>::NS;
  constexpr int NN = PredShape<N, ND>::NN;
  constexpr int CPW = PredShape<N, ND>::CPW;
  constexpr int NSF = ND == 2 ? N : 1;
  __shared__ double sQ[CPW * NN * V];
  __shared__ double sG[CPW * NN * V];
  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / NN;
  const int node = lane - slot * NN;
  const int t = node / NS;
  const int s = node - t * NS;
  const int ia = ND == 1 ? s : s / N;
  const int ib = ND == 1 ? 0 : s - (s / N) * N;
  const bool lane_ok = slot < CPW;
  const long e = (long)blockIdx.x * CPW + slot;
  const bool valid = lane_ok && e < ncell_ext && a.mask[e];
  const int nx = a.nx, ny = ND == 2 ? a.ny : 1;
  const int I = ND == 2 ? (int)(e / (ny + 2)) : (int)e, J = ND == 2 ? (int)(e % (ny + 2)) : 0;
  const bool real = valid && I >= 1 && I <= nx && (ND == 1 || (J >= 1 && J <= ny));
  const long c = real ? (ND == 2 ? (long)(I - 1) * ny + (J - 1) : I - 1) : 0;
  const int sl = lane_ok ? slot : 0;
  double *myQ = sQ + sl * NN * V;
  double *myG = sG + sl * NN * V;
  double q[V];
#pragma unroll
  for (int v = 0; v < V; v++) q[v] = valid ? a.qh[(e * NN + node) * V + v] : (v == 0 || v == 5 || v == 9 || v == 13 ? 1.0 : 0.0);
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = q[v];
  __syncthreads();
  double dqx[V], dqy[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    double ax = 0, ay = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      ax += K.der[ia][k] * myQ[(t * NS + (ND == 1 ? k : k * N + ib)) * V + v];
      if (ND == 2) ay += K.der[ib][k] * myQ[(t * NS + ia * N + k) * V + v];
    }
    dqx[v] = ax / a.dx;
    dqy[v] = ND == 2 ? ay / a.dy : 0.0;
  }
  double g[V], wl[V];
#pragma unroll
  for (int v = 0; v < V; v++) g[v] = 0.0;
  double gx[V], gy[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    gx[v] = 0.0;
    gy[v] = 0.0;
  }
  if (real) {
    const double wgt = ND == 2 ? a.dt * K.w[t] * K.w[ia] * K.w[ib] : a.dt * K.w[t] * K.w[ia];
#pragma unroll
    for (int v = 0; v < V; v++) wl[v] = wgt * a.lam_fv[c * V + v];
    if constexpr (DUAL) {
      for (int j = 0; j < V; j++) {
        Dual x[V], h[V];
#pragma unroll
        for (int v = 0; v < V; v++) x[v] = Dual(q[v], v == j ? 1.0 : 0.0);
        node_h<Dual, ND, TH>(x, dqx, dqy, a.m, h);
        double s2 = 0;
#pragma unroll
        for (int v = 0; v < V; v++) s2 += wl[v] * h[v].d;
        g[j] = s2;
      }
    } else {
      node_vjp<TH, true, false, false>(q, dqx, ND == 2 ? dqy : nullptr, a.m, wl, nullptr, nullptr, g);
    }
    // the dq inputs: h depends on dqx through -B_x(v) dqx
    NodeState<double, TH> st;
    node_state<double, TH>(q, a.m, st);
    double tmp[V];
    bdot_T<V>(st.v, wl, 0, tmp);
#pragma unroll
    for (int v = 0; v < V; v++) gx[v] = -tmp[v] / a.dx;
    if (ND == 2) {
      bdot_T<V>(st.v, wl, 1, tmp);
#pragma unroll
      for (int v = 0; v < V; v++) gy[v] = -tmp[v] / a.dy;
    }
    // d h / d parameter = par_fac S_A (theta1^-1 ~ 1/mu, or ~ sigmaY^-n)
    double S[V];
    source<double, TH>(q, st, a.m, S);
    double gm = 0;
#pragma unroll
    for (int k = 5; k < 14; k++) gm += wl[k] * S[k] * a.par_fac;
    atomicAdd(a.gmu, gm);
    if constexpr (TH) {
      double gk = 0;
#pragma unroll
      for (int k = 14; k < 17; k++) gk -= wl[k] * S[k] / a.kappa;
      atomicAdd(a.gkap, gk);
    }
  }
  __syncthreads();
  // scatter the dq adjoints along the lines: q_{t,k,ib} gets der[ia][k] gx(node)
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myG[node * V + v] = gx[v];
  __syncthreads();
#pragma unroll
  for (int v = 0; v < V; v++) {
    double acc = 0;
#pragma unroll
    for (int k = 0; k < N; k++) acc += K.der[k][ia] * myG[(t * NS + (ND == 1 ? k : k * N + ib)) * V + v];
    g[v] += acc;
  }
  if (ND == 2) {
    __syncthreads();
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myG[node * V + v] = gy[v];
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += K.der[k][ib] * myG[(t * NS + ia * N + k) * V + v];
      g[v] += acc;
    }
  }
  // face extrapolations: this cell is qL of its x+ (y+) face and qR of its x- (y-) face
  if (valid) {
    if (ND == 1 || (J >= 1 && J <= ny)) {
      const int j = ND == 2 ? J - 1 : 0;
      if (I <= nx) {  // x face i = I: this cell on its left (qL)
        const double *p = a.lam_face_x + ((((long)I * ny + j) * N + t) * NSF + (ND == 2 ? ib : 0)) * 2 * V;
#pragma unroll
        for (int v = 0; v < V; v++) g[v] += K.endv[1][ia] * p[v];
      }
      if (I >= 1) {  // x face i = I - 1: this cell on its right (qR)
        const double *p = a.lam_face_x + ((((long)(I - 1) * ny + j) * N + t) * NSF + (ND == 2 ? ib : 0)) * 2 * V;
#pragma unroll
        for (int v = 0; v < V; v++) g[v] += K.endv[0][ia] * p[V + v];
      }
    }
    if (ND == 2 && I >= 1 && I <= nx) {
      const int i = I - 1;
      if (J <= ny) {  // y face j = J: lower side (qL)
        const double *p = a.lam_face_y + ((((long)i * (ny + 1) + J) * N + t) * NSF + ia) * 2 * V;
#pragma unroll
        for (int v = 0; v < V; v++) g[v] += K.endv[1][ib] * p[v];
      }
      if (J >= 1) {  // y face j = J - 1: upper side (qR)
        const double *p = a.lam_face_y + ((((long)i * (ny + 1) + J - 1) * N + t) * NSF + ia) * 2 * V;
#pragma unroll
        for (int v = 0; v < V; v++) g[v] += K.endv[0][ib] * p[V + v];
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++) a.lam_qh[(e * NN + node) * V + v] = g[v];
  }
}

// Predictor: one lane per space-time node, CPW cells per wavefront.  At the
// solution q of R(q) = DG_U2 q - rhs(q, w) = 0 solve (dR/dq)^T m = lam_qh by
//   m <- DG_U2^-T (lam_qh + (d rhs / d q)^T m),
// the transposed Picard iteration, which converges exactly where the forward
// one does.  A cell where it does not (NaN, five growing updates in a row, or
// it_plain iterations) restarts with the transposed form of the forward's
// preconditioned fallback (with start_pre, every cell starts in that form),
// on g = W_s^-1 (lam_qh - (dR/dq)^T m):
//   m <- m + P^-T g,  P^-1 = DG_MAT^-1 (x) (I - Pi_s) + A_s^-1 (x) Pi_s
// on the distortion rows (DG_MAT^-1 on the others), with Pi_s and
// A_s = DG_MAT + dt kappa_s W as stiff_precond builds them at the node's WENO
// state; Pi_s is an orthogonal projection, so Pi_s^T = Pi_s.  Both forms have
// the same fixed point.  Then lam_wh = (d rhs / d w)^T m =
// sum_t psi_t(0) W_s m(t, s), and gmu += (d rhs / d parameter)^T m.
template <int N, int ND, bool DUAL, bool TH>
__global__ __launch_bounds__(64) void ka_predictor(AdjArgs a, long ncell_ext) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = PredShape<N, ND>::NS;
  constexpr int NN = PredShape<N, ND>::NN;
  constexpr int CPW = PredShape<N, ND>::CPW;
  __shared__ double sQ[CPW * NN * V];  // q (for dq), then the rhs adjoint pieces
  __shared__ double sM[CPW * NN * V];  // m
  __shared__ double sA[CPW * NN * V];  // scratch: a_F, a_G, dq adjoints
  __shared__ double sM0[CPW * NN * V];  // the starting iterate, for a restart
  // each node's dq/dx, dq/dy (read by its own lane only): in LDS for V = 14,
  // where it takes register pressure off the iteration; in registers for
  // V = 17, where the LDS would no longer fit four wavefronts per CU
  __shared__ double sDQ[TH ? 1 : 2 * CPW * NN * V];
  double dq_reg[TH ? 2 * V : 1];
  __shared__ double sRat[WAVE];
  constexpr int PRE = TH ? 27 : 18;
  __shared__ double sPre[CPW * NS * PRE];
  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / NN;
  const int node = lane - slot * NN;
  const int t = node / NS;
  const int s = node - t * NS;
  const int ia = ND == 1 ? s : s / N;
  const int ib = ND == 1 ? 0 : s - (s / N) * N;
  const bool lane_ok = slot < CPW;
  const long e = (long)blockIdx.x * CPW + slot;
  const bool active = lane_ok && e < ncell_ext && a.mask[e];
  const int sl = lane_ok ? slot : 0;
  double *myQ = sQ + sl * NN * V;
  double *myM = sM + sl * NN * V;
  double *myA = sA + sl * NN * V;
  const double dt = a.dt;
  const double Ws = ND == 1 ? K.w[ia] : K.w[ia] * K.w[ib];
  const double Wt = K.w[t];
  double q[V], lq[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    q[v] = active ? a.qh[(e * NN + node) * V + v] : (v == 0 || v == 5 || v == 9 || v == 13 ? 1.0 : 0.0);
    lq[v] = active ? a.lam_qh[(e * NN + node) * V + v] : 0.0;
  }
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = q[v];
  __syncthreads();
  double *dqx = TH ? dq_reg : sDQ + (sl * NN + node) * V;
  double *dqy = TH ? dq_reg + V : sDQ + (CPW * NN + sl * NN + node) * V;
#pragma unroll
  for (int v = 0; v < V; v++) {
    double ax = 0, ay = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      ax += K.der[ia][k] * myQ[(t * NS + (ND == 1 ? k : k * N + ib)) * V + v];
      if (ND == 2) ay += K.der[ib][k] * myQ[(t * NS + ia * N + k) * V + v];
    }
    if (TH || lane_ok) {  // in LDS, lanes past the last whole cell read slot 0's values and write none
      dqx[v] = ax / a.dx;
      dqy[v] = ND == 2 ? ay / a.dy : 0.0;
    }
  }
  double vel[3];  // the node's velocity, for the transposed B products
#pragma unroll
  for (int k = 0; k < 3; k++) vel[k] = q[2 + k] / q[0];
  // m0 = DG_U2^-T lam_qh: m(t, s) = sum_k DG_MAT^-1[k][t] x(k, s) / W_s
  auto solveT = [&](const double *x, double *out) {
    __syncthreads();
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myA[node * V + v] = x[v];
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += K.dgmatinv[k][t] * myA[(k * NS + s) * V + v];
      out[v] = acc / Ws;
    }
  };
  // m0 is kept in LDS for a restart
  double mu_[V];
  solveT(lq, mu_);
  double *myM0 = sM0 + sl * NN * V;
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myM0[node * V + v] = mu_[v];
  const unsigned long long cellbits = lane_ok ? (((1ull << NN) - 1ull) << (sl * NN)) : 0ull;
  const double *myPre = sPre + (sl * NS + s) * PRE;
  int phase = active ? (a.start_pre ? 1 : 0) : 2;  // 0 transposed Picard, 1 preconditioned, 2 done
  int it = 0, itot = 0, grow = 0;
  double prev = 0.0;
  bool failed = false, have_pre = false, used_pre = phase == 1;
  while (__ballot(phase < 2) != 0ull) {
    if (!have_pre && __ballot(phase == 1) != 0ull) {
      if (active && t == 0) stiff_precond<N, TH>(a.wh + (e * NS + s) * V, a.m, dt, K, sPre + (sl * NS + s) * PRE);
      have_pre = true;
      __syncthreads();
    }
    // (d rhs / d q)^T m at this node
    __syncthreads();
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myM[node * V + v] = mu_[v];
    __syncthreads();
    double ah[V], aF[V], aG[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      ah[v] = dt * Wt * Ws * mu_[v];
      // rhs_n -= dt Wt / dx sum_j DG_DER[a_n][j] W_b F(q_{t,j,b}): node (t, ia, ib) is j = ia
      double sF = 0, sG2 = 0;
#pragma unroll
      for (int k = 0; k < N; k++) {
        const int n_idx = ND == 1 ? k : k * N + ib;  // nodes (t, k, ib)
        sF += K.dgder[k][ia] * myM[(t * NS + n_idx) * V + v];
        if (ND == 2) sG2 += K.dgder[k][ib] * myM[(t * NS + ia * N + k) * V + v];
      }
      aF[v] = -dt * Wt / a.dx * (ND == 2 ? K.w[ib] : 1.0) * sF;
      aG[v] = ND == 2 ? -dt * Wt * K.w[ia] / a.dy * sG2 : 0.0;
    }
    double g[V];
    if constexpr (!DUAL) node_vjp<TH, true, true, ND == 2>(q, dqx, ND == 2 ? dqy : nullptr, a.m, ah, aF, aG, g);
    for (int j = 0; j < (DUAL ? V : 0); j++) {
      Dual x[V], h[V], F[V], G[V];
#pragma unroll
      for (int v = 0; v < V; v++) x[v] = Dual(q[v], v == j ? 1.0 : 0.0);
      NodeState<Dual, TH> sd;
      node_state<Dual, TH>(x, a.m, sd);
      source<Dual, TH>(x, sd, a.m, h);
      Dual dq[V];
#pragma unroll
      for (int v = 0; v < V; v++) dq[v] = Dual(dqx[v]);
      bdot_sub(sd.v, dq, 0, Dual(1.0), h);
      if (ND == 2) {
#pragma unroll
        for (int v = 0; v < V; v++) dq[v] = Dual(dqy[v]);
        bdot_sub(sd.v, dq, 1, Dual(1.0), h);
      }
      flux<Dual, TH>(x, sd, a.m, 0, F);
      if (ND == 2) flux<Dual, TH>(x, sd, a.m, 1, G);
      double acc = 0;
#pragma unroll
      for (int v = 0; v < V; v++) {
        acc += ah[v] * h[v].d + aF[v] * F[v].d;
        if (ND == 2) acc += aG[v] * G[v].d;
      }
      g[j] = acc;
    }
    // the dq paths: h -= B_d(v) dq_d, dq_x at node n = sum_k der[a_n][k] q_{t,k,b} / dx
    double tx[V], ty[V];
    bdot_T<V>(vel, ah, 0, tx);
    if (ND == 2) bdot_T<V>(vel, ah, 1, ty);
    __syncthreads();
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myA[node * V + v] = -tx[v] / a.dx;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += K.der[k][ia] * myA[(t * NS + (ND == 1 ? k : k * N + ib)) * V + v];
      g[v] += acc;
    }
    if (ND == 2) {
      __syncthreads();
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myA[node * V + v] = -ty[v] / a.dy;
      __syncthreads();
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < N; k++) acc += K.der[k][ib] * myA[(t * NS + ia * N + k) * V + v];
        g[v] += acc;
      }
    }
    // z = (lam_qh + (d rhs / d q)^T m) / W_s, less DG_MAT^T m in the preconditioned form
    double z[V], P[9];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double mt = 0;
#pragma unroll
      for (int k = 0; k < N; k++) mt += K.dgmat[k][t] * myM[(k * NS + s) * V + v];
      z[v] = (lq[v] + g[v]) / Ws - (phase == 1 ? mt : 0.0);
    }
    if (phase == 1) proj_dev_sym(myPre + 9, z + 5, P);
    __syncthreads();
    if (lane_ok) {
#pragma unroll
      for (int v = 0; v < V; v++) myA[node * V + v] = z[v];
      if (phase == 1)
#pragma unroll
        for (int k = 0; k < 9; k++) myQ[node * V + 5 + k] = P[k];
    }
    __syncthreads();
    double mn[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += K.dgmatinv[k][t] * myA[(k * NS + s) * V + v];
      if (phase == 1 && v >= 5 && v < 14)
#pragma unroll
        for (int k = 0; k < N; k++) acc += (myPre[k * N + t] - K.dgmatinv[k][t]) * myQ[(k * NS + s) * V + v];
      if (TH && phase == 1 && v >= 14)  // J rows: (DG_MAT + dt rJ W)^-T
#pragma unroll
        for (int k = 0; k < N; k++) acc += (myPre[18 + k * N + t] - K.dgmatinv[k][t]) * myA[(k * NS + s) * V + v];
      mn[v] = phase == 1 ? mu_[v] + acc : acc;
    }
    // stopping rule over the cell's block, as the forward iteration's
    bool more = false, nan = false;
    double ratio = 0;
#pragma unroll
    for (int v = 0; v < V; v++) {
      const double dd = fabs(mn[v] - mu_[v]), lim = a.tol * (1.0 + fabs(mu_[v]));
      more |= dd > lim;
      nan |= mn[v] != mn[v];
      ratio = fmax(ratio, dd / lim);
    }
    const unsigned long long bm = __ballot(more), bn = __ballot(nan);
    const bool cmore = (bm & cellbits) != 0ull, cnan = (bn & cellbits) != 0ull;
    sRat[lane] = ratio;
    __syncthreads();
    double cr = 0;
    if (lane_ok)
      for (int k = 0; k < NN; k++) cr = fmax(cr, sRat[sl * NN + k]);
    __syncthreads();
    if (phase < 2) itot++;
    if (phase == 0) {
      it++;
      bool restart = cnan;
      if (!cnan) {
#pragma unroll
        for (int v = 0; v < V; v++) mu_[v] = mn[v];
        if (!cmore) {
          phase = 2;
        } else {
          grow = cr > prev ? grow + 1 : 0;
          prev = cr;
          restart = it >= a.it_plain || grow >= 5;
        }
      }
      if (restart) {
        phase = 1;
        used_pre = true;
        it = 0;
#pragma unroll
        for (int v = 0; v < V; v++) mu_[v] = myM0[node * V + v];
      }
    } else if (phase == 1) {
      it++;
#pragma unroll
      for (int v = 0; v < V; v++) mu_[v] = mn[v];
      if (!cmore && !cnan) {
        phase = 2;
      } else if (cnan || it >= a.it_max) {
        phase = 2;
        failed = true;
      }
    }
  }
  if (failed && node == 0) atomicAdd(a.fail, 1);
  if (active && node == 0) {
    unsigned long long *slot = a.itc + (blockIdx.x % ITC_SLOTS) * ITC_STRIDE;
    atomicAdd(&slot[0], (unsigned long long)itot);
    atomicAdd(&slot[1], 1ull);
    if (used_pre) atomicAdd(&slot[2], 1ull);
  }
  // lam_wh(s) = sum_t psi_t(0) W_s m(t, s); d rhs / d mu through S
  __syncthreads();
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myM[node * V + v] = mu_[v];
  __syncthreads();
  if (active) {
    if (t == 0) {
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < N; k++) acc += K.endv[0][k] * Ws * myM[(k * NS + s) * V + v];
        a.lam_wh[(e * NS + s) * V + v] = acc;
      }
    }
    NodeState<double, TH> st;
    node_state<double, TH>(q, a.m, st);
    double S[V];
    source<double, TH>(q, st, a.m, S);
    double gm = 0;
#pragma unroll
    for (int k = 5; k < 14; k++) gm += dt * Wt * Ws * mu_[k] * S[k] * a.par_fac;
    atomicAdd(a.gmu, gm);
    if constexpr (TH) {
      double gk = 0;
#pragma unroll
      for (int k = 14; k < 17; k++) gk -= dt * Wt * Ws * mu_[k] * S[k] / a.kappa;
      atomicAdd(a.gkap, gk);
    }
  }
}

// WENO y-pass transposed: lam_ux[i][j + r][a][v] += d wh[i][j][a][.][v] / d ux[i][j + r][a][v]
// . lam_wh.  One thread per line (i, j, a, v), as k_weno_y.
template <int N, int V>
__global__ void ka_weno_y(AdjArgs a) {
  const Basis &K = basis_dev<N>();
  const int nx = a.nx, ny = a.ny;
  const int NX = nx + 2, NY = ny + 2, ncol = ny + 2 * N;
  const long total = (long)NX * NY * N * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int v = id % V;
    long c = id / V;
    const int aa = c % N;
    c /= N;
    const int j = c % NY, i = c / NY;
    const int gap = (i == 0 || i == NX - 1) ? 1 : 0;
    double lw[N];
#pragma unroll
    for (int b = 0; b < N; b++) lw[b] = a.lam_wh[((((long)i * NY + j) * N + aa) * N + b) * V + v];
    for (int r = 0; r < 2 * N - 1; r++) {
      Dual d[2 * N - 1], o[N];
#pragma unroll
      for (int k = 0; k < 2 * N - 1; k++) d[k] = Dual(a.ux[(((long)i * ncol + j + k) * N + aa) * V + v], k == r ? 1.0 : 0.0);
      weno_point<Dual, N>(d, K, NY, gap, j, o);
      double g = 0;
#pragma unroll
      for (int b = 0; b < N; b++) g += lw[b] * o[b].d;
      atomicAdd(&a.lam_ux[(((long)i * ncol + j + r) * N + aa) * V + v], g);
    }
  }
}

// WENO x-pass transposed: lam_ub[i + r][j][v] += d ux[i][j][.][v] / d ub[i + r][j][v] . lam_ux
template <int N, int V>
__global__ void ka_weno_x(AdjArgs a) {
  const Basis &K = basis_dev<N>();
  const int nx = a.nx, ny = a.ny;
  const int NX = nx + 2, ncol = ny + 2 * N;
  const long total = (long)NX * ncol * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int v = id % V;
    const long c = id / V;
    const int j = c % ncol, i = c / ncol;
    const int gap = (j < N || j >= ncol - N) ? N : 0;
    double lw[N];
#pragma unroll
    for (int b = 0; b < N; b++) lw[b] = a.lam_ux[(c * N + b) * V + v];
    for (int r = 0; r < 2 * N - 1; r++) {
      Dual d[2 * N - 1], o[N];
#pragma unroll
      for (int k = 0; k < 2 * N - 1; k++) d[k] = Dual(a.ub[((long)(i + k) * ncol + j) * V + v], k == r ? 1.0 : 0.0);
      weno_point<Dual, N>(d, K, NX, gap, i, o);
      double g = 0;
#pragma unroll
      for (int b = 0; b < N; b++) g += lw[b] * o[b].d;
      atomicAdd(&a.lam_ub[((long)(i + r) * ncol + j) * V + v], g);
    }
  }
}

// Boundaries transposed (codes 0, 1, 2, 3, 5: linear copies with signs) and
// the identity path: lam0 = lam_fv + sum over the ghost-extended positions
// that copy each cell.  One thread per (ghost-extended cell, variable).
template <int N, int V>
__global__ void ka_boundaries(AdjArgs a) {
  const int nx = a.nx, ny = a.ny;
  const long nby = ny + 2 * N;
  const long total = (long)(nx + 2 * N) * nby * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int v = id % V;
    const long c = id / V;
    const int J = c % nby, I = c / nby;
    bool rx, ry;
    const int codex = I < N ? a.bc.x : a.bc.y;
    const int i = bc_src(I, nx, N, codex, rx);
    const int codey = J < N ? a.bc.z : a.bc.w;
    const int j = bc_src(J, ny, N, codey, ry);
    double sgn = (rx ? bc_sign(v, codex, 0) : 1.0) * (ry ? bc_sign(v, codey, 1) : 1.0);
    const double g = sgn * a.lam_ub[id];
    if (g != 0.0) atomicAdd(&a.lam0[((long)i * ny + j) * V + v], g);
  }
}

// The transposed faces, cells and predictor of one (DUAL, TH) pair are
// compiled in a translation unit of their own (gpr_adj_heavy.hip, built once
// per pair), which also holds its own copy of the basis table.
template <bool DUAL, bool TH>
hipError_t adj_heavy_launch(dim3 gf, dim3 gc, hipStream_t s, const AdjArgs &a, long ne);
template <bool DUAL, bool TH> hipError_t adj_heavy_upload(const Basis *hb);

}  // namespace gpr
