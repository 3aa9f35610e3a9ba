// Instrumented FP64 operation counts of the ADER-WENO step (the roofline
// numerator).  Host program: the kernels' templated pointwise physics
// (gpr_device.hpp, gpr_pointwise.hpp: node_state, source, flux, bdot_sub,
// weno_point, rusanov, bint, max_abs_eigs, stiff_precond, proj_dev_sym) is
// instantiated with a number type that counts every operation it performs,
// and the contractions between nodes are replayed as the kernels perform them
// per lane.  Counting convention: +, -, *, / one operation each (a fused
// multiply-add is a multiply and an add, so it counts 2), and sqrt, cbrt,
// pow, exp, cos, acos one each; fabs, fmin, fmax and comparisons count 0.
// Loop-invariant scalar set-up (per-lane weights, 1/dx) is not counted per
// iteration.
//
//   g++ -std=c++17 -O1 -DGPR_HOST_ONLY flopcount.cpp -o flopcount
//   ./flopcount            # prints one JSON object
//
// Workload materials: the 2D 512^2 double shear layer's ideal-gas Newtonian
// fluid (gamma 1.4, c_s 8, mu 1e-3) and the 1024^2 Godunov-Romenski plastic
// solid; P2 (N = 3), V = 14, 2D.
#include <cmath>
#include <cstdio>
#include <cstring>

namespace fc {

struct Counts {
  long add = 0, mul = 0, div = 0, special = 0;
  long total() const { return add + mul + div + special; }
};
Counts g;

// a double that counts the arithmetic performed on it
struct F {
  double v;
  F() : v(0) {}
  F(double x) : v(x) {}
};
inline F operator+(F a, F b) { g.add++; return F(a.v + b.v); }
inline F operator-(F a, F b) { g.add++; return F(a.v - b.v); }
inline F operator*(F a, F b) { g.mul++; return F(a.v * b.v); }
inline F operator/(F a, F b) { g.div++; return F(a.v / b.v); }
inline F operator-(F a) { return F(-a.v); }
inline F &operator+=(F &a, F b) { return a = a + b; }
inline F &operator-=(F &a, F b) { return a = a - b; }
inline F &operator*=(F &a, F b) { return a = a * b; }
inline F &operator/=(F &a, F b) { return a = a / b; }
inline bool operator<(F a, F b) { return a.v < b.v; }
inline bool operator>(F a, F b) { return a.v > b.v; }
inline bool operator<=(F a, F b) { return a.v <= b.v; }
inline bool operator>=(F a, F b) { return a.v >= b.v; }
inline bool operator==(F a, F b) { return a.v == b.v; }
inline bool operator!=(F a, F b) { return a.v != b.v; }
inline F fabs(F a) { return F(std::fabs(a.v)); }
inline F fmin(F a, F b) { return F(std::fmin(a.v, b.v)); }
inline F fmax(F a, F b) { return F(std::fmax(a.v, b.v)); }
inline F sqrt(F a) { g.special++; return F(std::sqrt(a.v)); }
inline F cbrt(F a) { g.special++; return F(std::cbrt(a.v)); }
inline F pow(F a, F b) { g.special++; return F(std::pow(a.v, b.v)); }
inline double exponent_value(F b) { return b.v; }
inline F exp(F a) { g.special++; return F(std::exp(a.v)); }
inline F cos(F a) { g.special++; return F(std::cos(a.v)); }
inline F acos(F a) { g.special++; return F(std::acos(a.v)); }

}  // namespace fc

using fc::F;

#include "gpr_pointwise.hpp"

using namespace gpr;

namespace {

constexpr int N = 3, V = 14, NS = 9, NN = 27;

// Gauss-Legendre basis on [0, 1] for N = 3; only the nodes and weights enter
// a count through branches-free formulas, so the matrices are placeholders
Basis basis3() {
  Basis K;
  std::memset(&K, 0, sizeof K);
  K.N = 3;
  const double a = 0.5 * std::sqrt(0.6);
  const double x[3] = {0.5 - a, 0.5, 0.5 + a}, w[3] = {5.0 / 18, 4.0 / 9, 5.0 / 18};
  for (int i = 0; i < 3; i++) {
    K.x[i] = x[i];
    K.w[i] = w[i];
    for (int j = 0; j < 3; j++) {
      const double e = i == j ? 1.0 : 0.1;
      K.der[i][j] = K.dgder[i][j] = K.dgmat[i][j] = K.dgmatinv[i][j] = e;
      K.mLinv[i][j] = K.mRinv[i][j] = K.mCLinv[i][j] = K.sig[i][j] = e;
    }
    K.endv[0][i] = K.endv[1][i] = 1.0 / 3;
  }
  return K;
}

Mat make_mat(bool solid) {
  Mat m;
  std::memset(&m, 0, sizeof m);
  gpr_par &p = m.p;
  p.REACTION = -1;
  if (!solid) {  // ideal gas, Newtonian (the 512^2 double shear layer)
    p.EOS = GPR_EOS_STIFFENED_GAS;
    p.rho0 = 1.0;
    p.cv = 1.0;
    p.gamma = 1.4;
    p.b02 = 64.0;
    p.mu = 1e-3;
    p.n = 1.0;
    m.beta_zero = 1;
  } else {  // Godunov-Romenski plastic solid (the 1024^2 pulse)
    p.EOS = GPR_EOS_GODUNOV_ROMENSKI;
    p.rho0 = 8.9;
    p.cv = 4e-4;
    p.gamma = 2.0;
    p.c02 = 3.909 * 3.909;
    p.alpha = 1.0;
    p.beta = 3.0;
    p.b02 = 2.1 * 2.1;
    p.sigmaY = 0.4;
    p.tau0 = 1.0;
    p.n = 10.0;
    p.SOLID = 1;
    p.POWER_LAW = 1;
    m.beta_zero = 0;
  }
  m.mu_fac = p.rho0 / (2.0 * std::pow(p.mu > 0 ? p.mu : 1.0, 1.0 / p.n));
  return m;
}

// a representative state: density rho, velocity, a sheared distortion
void state(const Mat &m, double rho, F *q) {
  const double v[3] = {0.3, -0.1, 0.0};
  const double A[9] = {1.0, 0.05, 0.0, -0.02, 1.0, 0.01, 0.0, 0.0, 1.0};
  const double s = std::cbrt(rho / m.p.rho0);
  q[0] = rho;
  for (int k = 0; k < 3; k++) q[2 + k] = rho * v[k];
  for (int k = 0; k < 9; k++) q[5 + k] = s * A[k];
  q[1] = rho * (2.5 + 0.5 * (v[0] * v[0] + v[1] * v[1])) + 0.01;
}

long counted(const fc::Counts &a) { return fc::g.total() - a.total(); }

// Contraction over N nodes as a lane performs it: acc = sum_j c_j x_j
F contract(const F *c, const F *x) {
  F acc = 0.0;
  for (int j = 0; j < N; j++) acc += c[j] * x[j];
  return acc;
}

template <int MK> struct Counter {
  Mat m;
  Basis K;
  F q0[V], w[V], nbr[N][V], c[N];
  Counter(bool solid) : m(make_mat(solid)), K(basis3()) {
    state(m, solid ? 8.9 : 1.0, q0);
    for (int v = 0; v < V; v++) w[v] = q0[v];
    for (int j = 0; j < N; j++) {
      c[j] = 0.1 * (j + 1);
      for (int v = 0; v < V; v++) nbr[j][v] = q0[v];
    }
  }

  // the part of an iteration both phases share: derivatives, physics,
  // flux contractions, initial term (k_predictor steps 1-3)
  long common(F *rhs) {
    fc::Counts s0 = fc::g;
    NodeState<F, false> st;
    node_state<F, false, MK>(q0, m, st);
    source<F, false, MK>(q0, st, m, rhs);
    F dq[14], inv_dx = 512.0;
    for (int d = 0; d < 2; d++) {
      for (int v = 5; v < 14; v++) {
        F col[N];
        for (int j = 0; j < N; j++) col[j] = nbr[j][v];
        dq[v] = contract(c, col) * inv_dx;
      }
      bdot_sub(st.v, dq, d, F(1.0), rhs);
    }
    const F wts = 0.2;
    for (int v = 0; v < V; v++) rhs[v] *= wts;
    F Fx[V], Fy[V];
    flux<F, false>(q0, st, m, 0, Fx);
    flux<F, false>(q0, st, m, 1, Fy);
    const F wb = 0.4, cx = 1.0, cy = 1.0;
    for (int i = 0; i < FluxPack<false>::NF; i++) (void)(Fx[flux_var(i, 0)] * wb);  // the weighted x fluxes stored
    for (int d = 0; d < 2; d++)
      for (int i = 0; i < FluxPack<false>::NF; i++) {
        F col[N];
        for (int j = 0; j < N; j++) col[j] = (d == 0 ? Fx : Fy)[flux_var(i, d)];
        rhs[flux_var(i, d)] -= (d == 0 ? cx : cy) * contract(c, col);
      }
    const F dt = 1e-4, cW = 0.3;
    for (int v = 0; v < V; v++) rhs[v] = rhs[v] * dt + cW * w[v];
    return counted(s0);
  }

  // the stopping test with the bailout's ratio (k_predictor step 5)
  long stop(const F *q1) {
    fc::Counts s0 = fc::g;
    const F tol = 6e-6;
    for (int v = 0; v < V; v++) {
      F dd = fabs(q1[v] - q0[v]), lim = tol * (F(1.0) + fabs(q0[v]));
      (void)(dd / lim);
    }
    return counted(s0);
  }

  long picard_iteration() {
    fc::Counts s0 = fc::g;
    F rhs[V];
    common(rhs);
    F q1[V];
    const F iWs = 2.0;
    for (int v = 0; v < V; v++) {
      F col[N];
      for (int k = 0; k < N; k++) col[k] = rhs[v];
      q1[v] = contract(c, col) * iWs;
    }
    stop(q1);
    return counted(s0);
  }

  long fallback_iteration(const F *pre) {
    fc::Counts s0 = fc::g;
    F rhs[V];
    common(rhs);
    const F iWs = 2.0;
    F gv[V];
    for (int v = 0; v < V; v++) {  // g = DG_MAT q0 (time column) - rhs / W_s
      F col[N];
      for (int k = 0; k < N; k++) col[k] = nbr[k][v];
      gv[v] = contract(c, col) - rhs[v] * iWs;
    }
    F P[9];
    proj_dev_sym(pre + 9, gv + 5, P);
    for (int k = 0; k < 9; k++) gv[5 + k] -= P[k];
    F q1[V];
    for (int v = 0; v < V; v++) {
      F d = 0.0;
      for (int k = 0; k < N; k++) {
        if (v >= 5 && v < 14)
          d += c[k] * gv[v] + pre[k] * P[v - 5];
        else
          d += c[k] * gv[v];
      }
      q1[v] = q0[v] - d;
    }
    stop(q1);
    return counted(s0);
  }

  long setup(F *pre, int &polar_its) {
    fc::Counts s0 = fc::g;
    stiff_precond<N, false, MK, F>(w, m, F(1e-4), K, pre);
    // polar iterations of this state: rerun the Newton iteration on doubles
    double X[9];
    for (int i = 0; i < 9; i++) X[i] = w[5 + i].v;
    polar_its = 0;
    for (int it = 0; it < 40; it++) {
      polar_its++;
      double d = X[0] * (X[4] * X[8] - X[5] * X[7]) - X[1] * (X[3] * X[8] - X[5] * X[6]) +
                 X[2] * (X[3] * X[7] - X[4] * X[6]);
      double inv[9] = {(X[4] * X[8] - X[5] * X[7]) / d, (X[2] * X[7] - X[1] * X[8]) / d,
                       (X[1] * X[5] - X[2] * X[4]) / d, (X[5] * X[6] - X[3] * X[8]) / d,
                       (X[0] * X[8] - X[2] * X[6]) / d, (X[2] * X[3] - X[0] * X[5]) / d,
                       (X[3] * X[7] - X[4] * X[6]) / d, (X[1] * X[6] - X[0] * X[7]) / d,
                       (X[0] * X[4] - X[1] * X[3]) / d};
      double ch = 0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double nx = 0.5 * (X[i * 3 + j] + inv[j * 3 + i]);
          ch = std::fmax(ch, std::fabs(nx - X[i * 3 + j]));
          X[i * 3 + j] = nx;
        }
      if (ch < 1e-15) break;
    }
    return counted(s0);
  }

  // WENO of one variable along one line (k_weno_x / k_weno_y, one call each)
  long weno_line() {
    fc::Counts s0 = fc::g;
    F d[2 * N - 1], out[N];
    for (int r = 0; r < 2 * N - 1; r++) d[r] = 1.0 + 0.01 * r * r;
    weno_point<F, N>(d, K, 100, 1, 50, out);
    return counted(s0);
  }

  // one face lane of the FV corrector (k_fv_faces): both sides extrapolated
  // to the face node from the N nodes along the normal, one Rusanov flux,
  // one path integral, weighted into the two one-sided sums
  long face_lane() {
    fc::Counts s0 = fc::g;
    F qL[V], qR[V], lo[V], hi[V];
    for (int v = 0; v < V; v++) {
      F col[N];
      for (int j = 0; j < N; j++) col[j] = nbr[j][v];
      qL[v] = contract(c, col);
      qR[v] = contract(c, col);
      lo[v] = hi[v] = 0.0;
    }
    qR[0] = qR[0] * F(1.01);
    riemann_accum<N, false, F>(qL, qR, 0, m, K, 0.1, lo, hi);
    return counted(s0);
  }

  // one node of the FV cell update (k_fv_cells): dt W (S - B_x dq/dx - B_y dq/dy)
  long cell_node() {
    fc::Counts s0 = fc::g;
    NodeState<F, false> st;
    node_state<F, false, MK>(q0, m, st);
    F S[V], dq[14];
    source<F, false, MK>(q0, st, m, S);
    for (int d = 0; d < 2; d++) {
      for (int v = 5; v < 14; v++) {
        F col[N];
        for (int j = 0; j < N; j++) col[j] = nbr[j][v];
        dq[v] = contract(c, col) * F(512.0);
      }
      bdot_sub(st.v, dq, d, F(1.0), S);
    }
    const F wgt = 0.01;
    F acc[V];
    for (int v = 0; v < V; v++) acc[v] = acc[v] + wgt * S[v];
    return counted(s0);
  }

  // the post-step of one cell (k_post): renormalisation, contortion measure,
  // the two wave speeds of the next CFL maximum
  long post_cell() {
    fc::Counts s0 = fc::g;
    F u[V];
    for (int v = 0; v < V; v++) u[v] = q0[v];
    F cf = cbrt(u[0] / (F(m.p.rho0) * det3(u + 5)));
    for (int k = 0; k < 9; k++) u[5 + k] *= cf;
    F f = 0.0;
    for (int k = 0; k < 9; k++) f += u[5 + k] * u[5 + k];
    F d13 = cbrt(fabs(det3(u + 5)));
    (void)(f / (d13 * d13) / F(3.0) - F(1.0));
    F s = cfl_speed<false, 2, F, MK>(u, m, F(1.0 / 512), F(1.0 / 512));
    (void)s;
    return counted(s0);
  }
};

template <int MK> void report(const char *name, bool solid, bool last) {
  Counter<MK> C(solid);
  F pre[27];
  int polar = 0;
  const long setup = C.setup(pre, polar);
  const long pic = C.picard_iteration();
  const long fb = C.fallback_iteration(pre);
  const long weno = C.weno_line(), face = C.face_lane(), cell = C.cell_node(), post = C.post_cell();
  // per cell-update of the interior (2D, N = 3): WENO V lines along x and
  // N V along y; FV 2 faces x N time x N face nodes lanes, NN cell nodes
  const long weno_cell = (long)V * (1 + N) * weno;
  const long fv_cell = 2L * N * N * face + (long)NN * cell;
  std::printf("  \"%s\": {\"picard_node\": %ld, \"fallback_node\": %ld, \"fallback_setup_spatial\": %ld, "
              "\"polar_iterations\": %d, \"weno_line\": %ld, \"weno_cell\": %ld, \"fv_face_lane\": %ld, "
              "\"fv_cell_node\": %ld, \"fv_cell\": %ld, \"post_cell\": %ld}%s\n",
              name, pic, fb, setup, polar, weno, weno_cell, face, cell, fv_cell, post, last ? "" : ",");
}

}  // namespace

int main() {
  std::printf("{\n");
  report<MK_GAS_NEWTONIAN>("gas_newtonian_2d_p2", false, false);
  report<MK_GR_PLASTIC>("gr_plastic_2d_p2", true, true);
  std::printf("}\n");
  return 0;
}
