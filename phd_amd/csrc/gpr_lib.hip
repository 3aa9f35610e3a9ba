// libgpr_mi355x.so: host side of the ADER-WENO GPR step on gfx950 and its C ABI
// (include/gpr_mi355x.h).  Kernels are in gpr_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpr_mi355x.h"
#include "gpr_adjoint.hpp"
#include "gpr_batch.hpp"
#include "gpr_units.hpp"

using namespace gpr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(GPR_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

// ------------------------------------------------------------ basis constants
// Lagrange basis on the Gauss-Legendre nodes of [0, 1] (basis.cpp:7-72), the
// WENO stencil matrices (weno_matrices.cpp:9-29), the oscillation indicator
// (weno_matrices.cpp:31-42) and the DG time matrices (dg_matrices.cpp:30-53).
// Polynomials are monomial coefficient arrays; every integral is of degree
// <= 2N - 1 and is taken exactly with the same N-point Gauss rule.
using Poly = std::array<double, 8>;

double peval(const Poly &p, double x) {
  double r = 0;
  for (int i = 7; i >= 0; i--) r = r * x + p[i];
  return r;
}
Poly pderiv(const Poly &p) {
  Poly r{};
  for (int i = 1; i < 8; i++) r[i - 1] = i * p[i];
  return r;
}

int invert(int n, const double *a, double *out) {
  double m[4][8];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) m[i][j] = j < n ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (m[piv][c] == 0.0) return -1;
    for (int j = 0; j < 2 * n; j++) std::swap(m[c][j], m[piv][j]);
    double d = m[c][c];
    for (int j = 0; j < 2 * n; j++) m[c][j] /= d;
    for (int r = 0; r < n; r++)
      if (r != c) {
        double f = m[r][c];
        for (int j = 0; j < 2 * n; j++) m[r][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) out[i * n + j] = m[i][n + j];
  return 0;
}

Basis make_basis(int N) {
  Basis K;
  std::memset(&K, 0, sizeof K);
  K.N = N;
  if (N == 2) {
    double a = 0.5 / std::sqrt(3.0);
    K.x[0] = 0.5 - a;
    K.x[1] = 0.5 + a;
    K.w[0] = K.w[1] = 0.5;
  } else {
    double a = 0.5 * std::sqrt(0.6);
    K.x[0] = 0.5 - a;
    K.x[1] = 0.5;
    K.x[2] = 0.5 + a;
    K.w[0] = K.w[2] = 5.0 / 18.0;
    K.w[1] = 4.0 / 9.0;
  }
  Poly psi[MAXN];
  for (int j = 0; j < N; j++) {
    Poly p{};
    p[0] = 1.0;
    for (int m = 0; m < N; m++) {
      if (m == j) continue;
      double den = K.x[j] - K.x[m];
      Poly q{};
      for (int i = 0; i < 7; i++) {
        q[i + 1] += p[i] / den;
        q[i] -= p[i] * K.x[m] / den;
      }
      p = q;
    }
    psi[j] = p;
  }
  auto integ = [&](const Poly &p, double lo) {  // integral over [lo, lo + 1]
    double s = 0;
    for (int q = 0; q < N; q++) s += K.w[q] * peval(p, lo + K.x[q]);
    return s;
  };
  const int CN2 = N / 2;  // ceil((N - 1) / 2)
  const int FN2 = (N - 1) / 2;
  double mL[9], mR[9], mCL[9];
  for (int j = 0; j < N; j++) {
    Poly d1 = pderiv(psi[j]);
    K.endv[0][j] = peval(psi[j], 0.0);
    K.endv[1][j] = peval(psi[j], 1.0);
    for (int i = 0; i < N; i++) {
      K.der[i][j] = peval(d1, K.x[i]);
      mL[i * N + j] = integ(psi[j], i - N + 1);
      mR[i * N + j] = integ(psi[j], i);
      mCL[i * N + j] = integ(psi[j], i - CN2);
    }
  }
  (void)FN2;
  double t[9];
  invert(N, mL, t);
  for (int i = 0; i < N * N; i++) K.mLinv[i / N][i % N] = t[i];
  invert(N, mR, t);
  for (int i = 0; i < N * N; i++) K.mRinv[i / N][i % N] = t[i];
  invert(N, mCL, t);
  for (int i = 0; i < N * N; i++) K.mCLinv[i / N][i % N] = t[i];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      Poly pi = psi[i], pj = psi[j];
      double s = 0;
      for (int a = 1; a < N; a++) {
        pi = pderiv(pi);
        pj = pderiv(pj);
        for (int q = 0; q < N; q++) s += K.w[q] * peval(pi, K.x[q]) * peval(pj, K.x[q]);
      }
      K.sig[i][j] = s;
    }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      K.dgder[i][j] = i == j ? (K.endv[1][i] * K.endv[1][i] - K.endv[0][i] * K.endv[0][i]) / 2
                             : K.w[i] * K.der[i][j];
  double dm[9];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      K.dgmat[i][j] = K.endv[1][i] * K.endv[1][j] - K.dgder[j][i];
      dm[i * N + j] = K.dgmat[i][j];
    }
  invert(N, dm, t);
  for (int i = 0; i < N * N; i++) K.dgmatinv[i / N][i % N] = t[i];
  return K;
}

const Basis &basis(int N) {
  static const Basis b2 = make_basis(2), b3 = make_basis(3);
  return N == 2 ? b2 : b3;
}

int make_mat(const gpr_par *p, Mat &m) {
  if (!p) return fail(GPR_E_ARG, "null material parameters");
  if (p->EOS < 0 || p->EOS > 4) return fail(GPR_E_EOS, "EOS not recognized");
  if (p->REACTION > -1) return fail(GPR_E_UNSUPPORTED, "reactive materials are not on this path");
  std::memset(&m, 0, sizeof m);
  m.p = *p;
  m.mu_fac = p->rho0 / (2.0 * std::pow(p->mu, 1.0 / p->n));
  m.beta_zero = p->beta == 0.0;
  return GPR_OK;
}

}  // namespace

// ------------------------------------------------- allocation-free scratch
// Device scratch of gpr_ader_stepper_noalloc: the WENO x-pass, wh, qh, the
// face sums and a failure counter of one step, carved from one file-scope
// array so that the call allocates nothing.  Each device holds its own copy
// (GPR_NOALLOC_SCRATCH_BYTES) from the library's load until the process
// ends.  The host may not use the array's name as an address: its device
// address is taken once per device with hipGetSymbolAddress (at context
// creation) and kept in g_scratch_dev[device].
static __device__ __align__(256) unsigned char scratch[GPR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[MAX_DEVICES] = {};

// the device address of `scratch` on `device` (the current device)
int scratch_setup(int device, void **out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(GPR_E_HIP, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GPR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return GPR_OK;
}

__global__ void k_copy_int(int32_t *dst, const int *src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// double *p[0, n) = 0 as a kernel (graph capture records a kernel node)
__global__ void k_zero_doubles(double *p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}
}  // namespace

// ---------------------------------------------------------------- context
struct gpr_ctx {
  int N, ndim, thermal, V;
  gpr_solver_opts opts;
  hipStream_t stream = nullptr;
  int device = 0;
  void *scratch = nullptr;  // this device's address of the no-alloc scratch array
  std::map<std::string, std::pair<void *, size_t>> bufs;
  StepState *d_st = nullptr;
  int *d_fail = nullptr;
  // device-resident run
  double *run_u = nullptr;
  int run_nx = 0, run_ny = 1;
  double run_dx = 1, run_dy = 1;
  int4 run_bc{0, 0, 0, 0};
  Mat run_mat;
  bool run_ready = false;
  // per-phase timing of gpr_dev_steps
  bool timing = false;
  std::vector<hipEvent_t> events;  // 6 per timed step
  size_t ev_used = 0;
  int64_t timed_steps = 0;
  unsigned long long *d_itc = nullptr;  // Picard, fallback iterations, cells
  struct BatchHost *batch = nullptr;    // gpr_batch_* (1D contexts)
  struct AdjHost *adj = nullptr;        // gpr_adj_* (2D, V = 14)
};

namespace {

void *buf(gpr_ctx *c, const char *name, size_t bytes, int &rc) {
  auto &e = c->bufs[name];
  if (e.second >= bytes && e.first) return e.first;
  if (e.first) (void)hipFree(e.first);
  e.first = nullptr;
  e.second = 0;
  void *p = nullptr;
  hipError_t err = hipMalloc(&p, std::max<size_t>(bytes, 256));
  if (err != hipSuccess) {
    rc = fail(GPR_E_HIP, std::string("hipMalloc(") + name + "): " + hipGetErrorString(err));
    return nullptr;
  }
  e.first = p;
  e.second = bytes;
  rc = GPR_OK;
  return p;
}

int grid_for(long total, int block) {
  long g = (total + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (1L << 20)) g = 1L << 20;
  return (int)g;
}

long ext_cells(const gpr_ctx *c, const int32_t *nX, int ext) {
  long r = nX[0] + 2 * ext;
  if (c->ndim == 2) r *= nX[1] + 2 * ext;
  return r;
}

int check_nX(const gpr_ctx *c, const int32_t *nX) {
  if (!nX) return fail(GPR_E_ARG, "null nX");
  for (int d = 0; d < c->ndim; d++)
    if (nX[d] < 1) return fail(GPR_E_ARG, "grid dimensions must be positive");
  if (c->ndim == 2 && (nX[0] < 1 || nX[1] < 1)) return fail(GPR_E_ARG, "bad grid");
  return GPR_OK;
}

// ---------------------------------------------------------- dispatch helpers
template <template <int, bool, int> class Op, typename... A> int dispatch(const gpr_ctx *c, A &&...args) {
  int N = c->N;
  bool th = c->thermal;
  int nd = c->ndim;
#define GPR_CASE(n, t, d) \
  if (N == n && th == t && nd == d) return Op<n, t, d>::run(args...);
  GPR_CASE(3, false, 2) GPR_CASE(3, true, 2) GPR_CASE(3, false, 1) GPR_CASE(3, true, 1)
  GPR_CASE(2, false, 2) GPR_CASE(2, true, 2) GPR_CASE(2, false, 1) GPR_CASE(2, true, 1)
#undef GPR_CASE
  return fail(GPR_E_UNSUPPORTED, "unsupported (N, thermal, ndim)");
}

// ------------------------------------------------------------- launches
template <int N, bool TH, int ND> struct BoundariesOp {
  static int run(gpr_ctx *c, const double *d_u, double *d_ub, int nx, int ny, int4 bc) {
    constexpr int V = Dims<TH>::V;
    long total = (long)(nx + 2 * N) * (ND == 2 ? ny + 2 * N : 1) * V;
    hipLaunchKernelGGL((k_boundaries<V, ND>), dim3(grid_for(total, 256)), dim3(256), 0, c->stream, d_u, d_ub, nx,
                       ny, N, bc);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

// WENO in the type R (double: the plain method; float: the FP32-bulk form),
// ub read in double, wh written in R.
// 2D: the fused x/y launch (k_weno_xy), unless the x-pass output is wanted:
// in ux_given, or (keep_ux) in the context's "ux" buffer, which the adjoint's
// reverse sweep reads; the values are the same.
template <typename R, int N, bool TH, int ND>
int weno_run(gpr_ctx *c, const double *d_ub, R *d_wh, int nx, int ny, R *ux_given = nullptr, bool keep_ux = false) {
  constexpr int V = Dims<TH>::V;
  const Basis &K = basis(N);
  if (ND == 1) {
    long total = (long)(nx + 2) * V;
    hipLaunchKernelGGL((k_weno_1d<R, N, V>), dim3(grid_for(total, 256)), dim3(256), 0, c->stream, d_ub, d_wh, nx,
                       K);
  } else if (!ux_given && !keep_ux) {
    const long items = (long)(nx + 2) * ((ny + 2 + WENO_JT - 1) / WENO_JT);
    hipLaunchKernelGGL((k_weno_xy<R, N, V>), dim3(8 * ((items + 7) / 8)), dim3(256), 0, c->stream, d_ub, d_wh, nx,
                       ny, K);
  } else {
    int rc;
    long nux = (long)(nx + 2) * (ny + 2 * N) * N * V;
    R *ux = ux_given ? ux_given : (R *)buf(c, sizeof(R) == 8 ? "ux" : "uxf", nux * sizeof(R), rc);
    if (!ux) return rc;
    long t1 = (long)(nx + 2) * (ny + 2 * N) * V;
    hipLaunchKernelGGL((k_weno_x<R, N, V>), dim3(grid_for(t1, 256)), dim3(256), 0, c->stream, d_ub, ux, nx, ny, K);
    long t2 = (long)(nx + 2) * (ny + 2) * N * V;
    hipLaunchKernelGGL((k_weno_y<R, N, V>), dim3(grid_for(t2, 256)), dim3(256), 0, c->stream, ux, d_wh, nx, ny, K);
  }
  HIP_TRY(hipGetLastError());
  return GPR_OK;
}

template <int N, bool TH, int ND> struct WenoOp {
  static int run(gpr_ctx *c, const double *d_ub, double *d_wh, int nx, int ny, double *ux = nullptr,
                 bool keep_ux = false) {
    return weno_run<double, N, TH, ND>(c, d_ub, d_wh, nx, ny, ux, keep_ux);
  }
};

template <int N, bool TH, int ND> struct WenoF32Op {
  static int run(gpr_ctx *c, const double *d_ub, float *d_wh, int nx, int ny) {
    return weno_run<float, N, TH, ND>(c, d_ub, d_wh, nx, ny);
  }
};

// The predictor and FV kernels are compiled in units of their own
// (gpr_pred.hip per N, gpr_fv.hip per (N, dimensions)); see gpr_units.hpp.
template <int N, bool TH, int ND> struct PredictorOp {
  static int run(gpr_ctx *c, PredArgs a) {
    HIP_TRY((pred_launch<N, TH, ND>(c->stream, a, PRED_ONE_LAUNCH)));
    return GPR_OK;
  }
};

template <int N, bool TH, int ND> struct PredictorF32Op {
  static int run(gpr_ctx *c, PredArgs a) {
    HIP_TRY((pred_launch<N, TH, ND>(c->stream, a, PRED_SPLIT_FP32)));
    return GPR_OK;
  }
};

template <int N, bool TH, int ND> struct PredictorSplitOp {
  static int run(gpr_ctx *c, PredArgs a) {
    HIP_TRY((pred_launch<N, TH, ND>(c->stream, a, PRED_SPLIT_FP64)));
    return GPR_OK;
  }
};

// the post-step on a.u (k_post; with a.ode the split step's closing ODE first)
template <int N, bool TH, int ND> struct PostOp {
  static int run(gpr_ctx *c, FvArgs a) {
    long nc = (long)a.nx * (ND == 2 ? a.ny : 1);
    if (a.ode)
      hipLaunchKernelGGL((k_post<TH, ND, true>), dim3(grid_for(nc, 256)), dim3(256), 0, c->stream, a.u, nc, a.dx,
                         a.dy, a.m, a.st);
    else
      hipLaunchKernelGGL((k_post<TH, ND>), dim3(grid_for(nc, 256)), dim3(256), 0, c->stream, a.u, nc, a.dx, a.dy,
                         a.m, a.st);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

// the FV corrector, then the post-step if a.post
template <int N, bool TH, int ND> struct FvOp {
  static int run(gpr_ctx *c, FvArgs a) {
    HIP_TRY((fv_launch<N, TH, ND>(c->stream, a)));
    if (a.post) return PostOp<N, TH, ND>::run(c, a);
    return GPR_OK;
  }
};

// ode_launcher (solvers/split/ode.cpp:5-11) over n cells of x; with u, x is
// the extended grid ub and interior results also go to u (k_ode).
template <int N, bool TH, int ND> struct OdeOp {
  static int run(gpr_ctx *c, double *x, long n, double *u, int nx, int ny, const double *dt_dev, double dt,
                 double scale, const gpr_par &p) {
    hipLaunchKernelGGL((k_ode<TH, ND>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, x, n, u, nx, ny, N, dt_dev,
                       dt, scale, p);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

// midstepper (solvers/split/homogeneous.cpp:5-67) over the ncell masked cells of wh
template <int N, bool TH, int ND> struct MidstepOp {
  static int run(gpr_ctx *c, double *wh, long ncell, const uint8_t *mask, const double *dt_dev, double dt,
                 double dx, double dy, const Mat &m) {
    constexpr int CPW = MidShape<N, ND>::CPW;
    hipLaunchKernelGGL((k_midstep<N, TH, ND>), dim3((ncell + CPW - 1) / CPW), dim3(WAVE), 0, c->stream, wh, ncell,
                       mask, dt_dev, dt, dx, dy, m);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

template <int N, bool TH, int ND> struct CflOp {
  static int run(gpr_ctx *c, const double *d_u, long ncell, double dx, double dy, const Mat &m,
                 unsigned long long *dst) {
    hipLaunchKernelGGL((k_cfl<TH, ND>), dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, d_u, ncell, dx, dy, m,
                       dst);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

template <int N, bool TH, int ND> struct ResetOp {
  static int run(gpr_ctx *c, double *d_u, long ncell, double dx, double dy, const Mat &m, StepState *st) {
    hipLaunchKernelGGL((k_reset<TH, ND>), dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, d_u, ncell, dx, dy,
                       m, st);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

// pointwise operations over a batch of states
enum PointOp {
  OP_FLUX, OP_SOURCE, OP_BDOT, OP_EIGS, OP_TAT, OP_DRUS, OP_BINT, OP_RENORM, OP_RESET, OP_CONTORTED,
  OP_ODE_DISTORTION, OP_ODE_THERMAL,  // dt in tol
  OP_SYSMAT, OP_DFDP, OP_DPDQ, OP_DROE, OP_DOSH
};

// the system matrix, its factors and the Roe / Osher fluxes, one state (or
// pair) per thread over a bounded grid; ws: a workspace slice per thread
template <int N, bool TH>
__global__ void k_pointwise_eig(int op, double *ret, const double *A, const double *B, long count, int d, Mat m,
                                Basis K, int *flag, double *ws) {
  constexpr int V = Dims<TH>::V;
  constexpr long nn = (long)V * V;
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x, nthr = (long)gridDim.x * blockDim.x;
  const LaneMem ews{ws + tid, nthr};
  for (long i = tid; i < count; i += nthr) {
    double q[V];
    for (int v = 0; v < V; v++) q[v] = A[i * V + v];
    switch (op) {
    case OP_SYSMAT: system_matrix<TH>(q, d, m, LaneMem{ret + i * nn, 1}, ews, ews.at(nn)); break;
    case OP_DFDP: ef_dFdP<TH>(q, d, m, LaneMem{ret + i * nn, 1}); break;
    case OP_DPDQ: ef_dPdQ<TH>(q, m, LaneMem{ret + i * nn, 1}); break;
    default: {  // OP_DROE, OP_DOSH
      double qR[V], r[V];
      for (int v = 0; v < V; v++) qR[v] = B[i * V + v];
      if (eig_flux<N, TH>(q, qR, d, m, K, op == OP_DROE ? 1 : 2, r, ews)) atomicOr(flag, 1);
      for (int v = 0; v < V; v++) ret[i * V + v] = r[v];
    }
    }
  }
}

template <int N, bool TH>
__global__ void k_pointwise(int op, double *ret, const double *A, const double *B, long count, int d, Mat m,
                            Basis K, double tol, int *flag) {
  constexpr int V = Dims<TH>::V;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    double q[V];
    for (int v = 0; v < V; v++) q[v] = A[i * V + v];
    switch (op) {
    case OP_FLUX: {
      NodeState<double, TH> s;
      node_state<double, TH>(q, m, s);
      double F[V];
      flux<double, TH>(q, s, m, d, F);
      for (int v = 0; v < V; v++) ret[i * V + v] = F[v];
    } break;
    case OP_SOURCE: {
      NodeState<double, TH> s;
      node_state<double, TH>(q, m, s);
      double S[V];
      source<double, TH>(q, s, m, S);
      for (int v = 0; v < V; v++) ret[i * V + v] = S[v];
    } break;
    case OP_BDOT: {
      double vv[3] = {q[2] / q[0], q[3] / q[0], q[4] / q[0]};
      double r[V];
      bdot<double, V>(vv, B + i * V, d, r);
      for (int v = 0; v < V; v++) ret[i * V + v] = r[v];
    } break;
    case OP_EIGS: ret[i] = max_abs_eigs<TH>(q, m, d); break;
    case OP_TAT: {
      double X[16];
      int n = thermo_acoustic_tensor<TH>(q, m, d, X);
      for (int k = 0; k < n * n; k++) ret[i * n * n + k] = X[k];
    } break;
    case OP_DRUS:
    case OP_BINT: {
      double qR[V], r[V];
      for (int v = 0; v < V; v++) qR[v] = B[i * V + v];
      if (op == OP_DRUS)
        rusanov<TH>(q, qR, d, m, r);
      else
        bint<N, TH>(q, qR, d, K, r);
      for (int v = 0; v < V; v++) ret[i * V + v] = r[v];
    } break;
    case OP_RENORM: {
      if (q[0] > 0) {
        double cf = cbrt(q[0] / (m.p.rho0 * det3(q + 5)));
        for (int k = 0; k < 9; k++) ret[i * V + 5 + k] = q[5 + k] * cf;
      }
    } break;
    case OP_RESET: {
      NodeState<double, TH> s;
      node_state<double, TH>(q, m, s);
      double r = q[0], cf = cbrt(r / m.p.rho0);
      double Gm = Gamma_MG(r, m.p);
      double C0 = (1 - m.p.beta / Gm) * s.cs2;
      ret[i * V + 1] = q[1] - r * (C0 / 4 * s.devGsq);
      for (int k = 0; k < 9; k++) ret[i * V + 5 + k] = 0.0;
      ret[i * V + 5] = cf;
      ret[i * V + 9] = cf;
      ret[i * V + 13] = cf;
    } break;
    case OP_CONTORTED: {
      double f = 0;
      for (int k = 0; k < 9; k++) f += q[5 + k] * q[5 + k];
      double d13 = cbrt(fabs(det3(q + 5)));
      if (f / (d13 * d13) / 3.0 - 1.0 > tol) atomicOr(flag, 1);
    } break;
    case OP_ODE_DISTORTION:
    case OP_ODE_THERMAL: {
      if (op == OP_ODE_DISTORTION)
        ode_distortion(q + 5, tol, m.p);
      else if (TH)
        ode_thermal(q, tol, m.p);
      for (int v = 0; v < V; v++) ret[i * V + v] = q[v];
    } break;
    }
  }
}

int pointwise(gpr_ctx *c, int op, double *ret, const double *A, const double *B, int64_t count, int d,
              const gpr_par *MP, size_t ret_per, double tol = 0, int32_t *flag_out = nullptr) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (count < 0 || !A) return fail(GPR_E_ARG, "bad arguments");
  if (d < 0 || d > 2) return fail(GPR_E_ARG, "direction out of range");
  Mat m;
  std::memset(&m, 0, sizeof m);
  if (MP) {
    int rc = make_mat(MP, m);
    if (rc) return rc;
  }
  if (count == 0) return GPR_OK;
  int rc;
  size_t V = c->V;
  double *dA = (double *)buf(c, "pA", count * V * 8, rc);
  if (!dA) return rc;
  double *dB = nullptr;
  if (B) {
    dB = (double *)buf(c, "pB", count * V * 8, rc);
    if (!dB) return rc;
    HIP_TRY(hipMemcpyAsync(dB, B, count * V * 8, hipMemcpyHostToDevice, c->stream));
  }
  double *dR = (double *)buf(c, "pR", count * std::max<size_t>(ret_per, 1) * 8, rc);
  if (!dR) return rc;
  HIP_TRY(hipMemcpyAsync(dA, A, count * V * 8, hipMemcpyHostToDevice, c->stream));
  if (ret_per) HIP_TRY(hipMemcpyAsync(dR, ret, count * ret_per * 8, hipMemcpyHostToDevice, c->stream));
  int *dflag = c->d_fail;
  HIP_TRY(hipMemsetAsync(dflag, 0, sizeof(int), c->stream));
  const Basis &K = basis(c->N);
  int g = grid_for(count, 128);
  if (op >= OP_SYSMAT) {  // a bounded grid and its workspace
    g = std::min(g, 64);
    double *dws = (double *)buf(c, "pws", (size_t)g * 128 * EIG_WS_DOUBLES(c->V) * 8, rc);
    if (!dws) return rc;
#define GPR_PWE(n, t)                                                                                      \
  if (c->N == n && (bool)c->thermal == t)                                                                \
    hipLaunchKernelGGL((k_pointwise_eig<n, t>), dim3(g), dim3(128), 0, c->stream, op, dR, dA, dB, (long)count, d, \
                       m, K, dflag, dws);
    GPR_PWE(3, false) GPR_PWE(3, true) GPR_PWE(2, false) GPR_PWE(2, true)
#undef GPR_PWE
  } else {
#define GPR_PW(n, t)                                                                                          \
  if (c->N == n && (bool)c->thermal == t)                                                                   \
    hipLaunchKernelGGL((k_pointwise<n, t>), dim3(g), dim3(128), 0, c->stream, op, dR, dA, dB, (long)count, d, m, \
                       K, tol, dflag);
    GPR_PW(3, false) GPR_PW(3, true) GPR_PW(2, false) GPR_PW(2, true)
#undef GPR_PW
  }
  HIP_TRY(hipGetLastError());
  if (ret_per) HIP_TRY(hipMemcpyAsync(ret, dR, count * ret_per * 8, hipMemcpyDeviceToHost, c->stream));
  if (flag_out) HIP_TRY(hipMemcpyAsync(flag_out, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

void host_extend_mask(const gpr_ctx *c, const uint8_t *mask, uint8_t *maskb, const int32_t *nX) {
  // extend_mask (grid.cpp:176-230); the 2D border assignments are loop-invariant
  int nx = nX[0];
  if (c->ndim == 1) {
    for (int i = 0; i < nx; i++)
      maskb[i + 1] = mask[i] || (i > 0 && mask[i - 1]) || (i < nx - 1 && mask[i + 1]);
    maskb[0] = maskb[1];
    maskb[nx + 1] = maskb[nx];
    return;
  }
  int ny = nX[1];
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      long ind = (long)i * ny + j;
      maskb[(long)(i + 1) * (ny + 2) + (j + 1)] = mask[ind] || (i > 0 && mask[ind - ny]) ||
                                                  (i < nx - 1 && mask[ind + ny]) || (j > 0 && mask[ind - 1]) ||
                                                  (j < ny - 1 && mask[ind + 1]);
    }
  for (int i = 0; i < nx; i++) {
    maskb[(long)(i + 1) * (ny + 2)] = mask[(long)i * ny];
    maskb[(long)(i + 1) * (ny + 2) + (ny + 1)] = mask[(long)i * ny + (ny - 1)];
  }
  for (int j = 0; j < ny; j++) {
    maskb[j + 1] = mask[j];
    maskb[(long)(nx + 1) * (ny + 2) + (j + 1)] = mask[(long)(nx - 1) * ny + j];
  }
  maskb[0] = 0;
  maskb[ny + 1] = 0;
  maskb[(long)(nx + 1) * (ny + 2)] = 0;
  maskb[(long)(nx + 1) * (ny + 2) + (ny + 1)] = 0;
}

int check_bcs(const gpr_ctx *c, const int32_t *bcs, int4 &bc) {
  if (!bcs) return fail(GPR_E_ARG, "null boundary codes");
  int v[4] = {0, 0, 0, 0};
  for (int i = 0; i < 2 * c->ndim; i++) {
    if (bcs[i] < 0 || bcs[i] > 5)
      return fail(GPR_E_ARG, "boundary code must be 0..5 (grid.cpp:5-10)");
    v[i] = bcs[i];
  }
  bc = make_int4(v[0], v[1], v[2], v[3]);
  return GPR_OK;
}

PredArgs pred_args(gpr_ctx *c, const Mat &m, double dt, const double *dX) {
  PredArgs a;
  std::memset(&a, 0, sizeof a);
  a.dt = dt;
  a.dx = dX[0];
  a.dy = c->ndim == 2 ? dX[1] : 1.0;
  a.dg_it = c->opts.dg_it;
  a.dg_tol = c->opts.dg_tol;
  a.stiff_it = c->opts.stiff_it;
  a.stiff_tol = c->opts.stiff_tol;
  a.bailout = c->opts.bailout;
  a.bail_from = c->opts.bail_from;
  a.mode = c->opts.mode;
  a.m = m;
  a.fail = c->d_fail;
  return a;
}

FvArgs fv_args(gpr_ctx *c, const Mat &m, double dt, const double *dX, const int32_t *nX) {
  FvArgs a;
  std::memset(&a, 0, sizeof a);
  a.dt = dt;
  a.dx = dX[0];
  a.dy = c->ndim == 2 ? dX[1] : 1.0;
  a.nx = nX[0];
  a.ny = c->ndim == 2 ? nX[1] : 1;
  a.time = 1;
  a.sources = 1;
  a.parts = 3;
  a.m = m;
  return a;
}

int fv_buffers(gpr_ctx *c, FvArgs &a) {
  int rc;
  long nfx = (long)(a.nx + 1) * a.ny, nfy = c->ndim == 2 ? (long)a.nx * (a.ny + 1) : 1;
  a.fx = (double *)buf(c, "fx", nfx * 2 * c->V * 8, rc);
  if (!a.fx) return rc;
  a.fy = (double *)buf(c, "fy", nfy * 2 * c->V * 8, rc);
  if (!a.fy) return rc;
  if (a.flux != 0) {  // the Roe / Osher face launch's workspace (one slice per lane of its grid)
    const long lanes_per_face = (long)(a.time ? c->N : 1) * (c->ndim == 2 ? c->N : 1);
    const long fpw = WAVE / lanes_per_face;
    const long items = ((c->ndim == 2 ? nfx + nfy : nfx) + fpw - 1) / fpw;
    a.eig_ws = (double *)buf(c, "eig_ws", eig_face_grid(items) * WAVE * EIG_WS_DOUBLES(c->V) * 8, rc);
    if (!a.eig_ws) return rc;
  }
  return GPR_OK;
}

__global__ void k_set_dt(StepState *st, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st->dt = dt;
}
// *dst = the dt the coming step takes
__global__ void k_record_dt(const StepState *st, double *dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = st->dt;
}
int dev_steps_impl(gpr_ctx *c, int64_t nsteps, const double *dts);

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char *gpr_last_error(void) { return g_err.c_str(); }
int gpr_version(void) { return 1; }

int gpr_ctx_create(gpr_ctx **ctx, int N, int ndim, int thermal) {
  if (!ctx) return fail(GPR_E_ARG, "null ctx pointer");
  *ctx = nullptr;
  if (N != 2 && N != 3) return fail(GPR_E_UNSUPPORTED, "N must be 2 or 3");
  if (ndim != 1 && ndim != 2) return fail(GPR_E_UNSUPPORTED, "ndim must be 1 or 2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(GPR_E_HIP, "no HIP device available");
  gpr_ctx *c = new gpr_ctx();
  c->N = N;
  c->ndim = ndim;
  c->thermal = thermal ? 1 : 0;
  c->V = thermal ? 17 : 14;
  c->opts.dg_it = 50;
  c->opts.dg_tol = 6e-6;
  c->opts.stiff_it = 200;
  c->opts.stiff_tol = 6e-6;
  c->opts.bailout = 0;
  c->opts.mode = 0;
  c->opts.precision = 0;
  c->opts.contorted_tol = 1.0;
  c->opts.split = 0;
  c->opts.half_step = 0;
  c->opts.bail_from = 0;
  c->opts.flux = 0;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_st, sizeof(StepState));
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_fail, sizeof(int) * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_itc, sizeof(unsigned long long) * ITC_SLOTS * ITC_STRIDE);
  if (e == hipSuccess) {
    // basis tables of the constant segment on this device (same values for
    // every context, so concurrent writers agree)
    const Basis hb[2] = {basis(2), basis(3)};
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_basis), hb, sizeof(hb));
    // the copies of the translation units that launch the adjoint's kernels
    if (e == hipSuccess) e = adj_heavy_upload<false, false>(hb);
    if (e == hipSuccess) e = adj_heavy_upload<false, true>(hb);
    if (e == hipSuccess) e = adj_heavy_upload<true, false>(hb);
    if (e == hipSuccess) e = adj_heavy_upload<true, true>(hb);
    if (e == hipSuccess) e = units_upload_basis(hb);
  }
  if (e != hipSuccess) {
    int rc = fail(GPR_E_HIP, std::string("context setup: ") + hipGetErrorString(e));
    gpr_ctx_destroy(c);
    return rc;
  }
  if (int rc = scratch_setup(c->device, &c->scratch)) {
    gpr_ctx_destroy(c);
    return rc;
  }
  *ctx = c;
  return GPR_OK;
}

int gpr_ctx_destroy(gpr_ctx *c) {
  if (!c) return GPR_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->bufs)
    if (kv.second.first) (void)hipFree(kv.second.first);
  if (c->d_st) (void)hipFree(c->d_st);
  if (c->d_fail) (void)hipFree(c->d_fail);
  if (c->d_itc) (void)hipFree(c->d_itc);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  if (c->batch) (void)gpr_batch_destroy(c);
  if (c->adj) (void)gpr_adj_release(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GPR_OK;
}

int gpr_ctx_set_opts(gpr_ctx *c, const gpr_solver_opts *o) {
  if (!c || !o) return fail(GPR_E_ARG, "null argument");
  if (o->dg_it < 1 || o->stiff_it < 1 || !(o->dg_tol > 0) || !(o->stiff_tol > 0) || o->bailout < 0 ||
      o->mode < 0 || o->mode > 2 || o->precision < 0 || o->precision > 1 || o->split < 0 || o->split > 1 ||
      o->half_step < 0 || o->half_step > 1 || o->bail_from < 0 || o->flux < 0 || o->flux > 2)
    return fail(GPR_E_ARG, "invalid solver options");
  c->opts = *o;
  return GPR_OK;
}

int gpr_ctx_get_opts(const gpr_ctx *c, gpr_solver_opts *o) {
  if (!c || !o) return fail(GPR_E_ARG, "null argument");
  *o = c->opts;
  return GPR_OK;
}

void *gpr_ctx_stream(gpr_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gpr_options(const gpr_ctx *c, int32_t *N, int32_t *V, int32_t *ndim, int32_t *thermal) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (N) *N = c->N;
  if (V) *V = c->V;
  if (ndim) *ndim = c->ndim;
  if (thermal) *thermal = c->thermal;
  return GPR_OK;
}

int gpr_constants(int N, double *nodes, double *wghts, double *endvals, double *dervals, double *mLinv,
                  double *mRinv, double *mCLinv, double *sig, double *dg_end, double *dg_der, double *dg_mat) {
  if (N != 2 && N != 3) return fail(GPR_E_UNSUPPORTED, "N must be 2 or 3");
  const Basis &K = basis(N);
  for (int i = 0; i < N; i++) {
    if (nodes) nodes[i] = K.x[i];
    if (wghts) wghts[i] = K.w[i];
    if (endvals) {
      endvals[i] = K.endv[0][i];
      endvals[N + i] = K.endv[1][i];
    }
    for (int j = 0; j < N; j++) {
      if (dervals) dervals[i * N + j] = K.der[i][j];
      if (mLinv) mLinv[i * N + j] = K.mLinv[i][j];
      if (mRinv) mRinv[i * N + j] = K.mRinv[i][j];
      if (mCLinv) mCLinv[i * N + j] = K.mCLinv[i][j];
      if (sig) sig[i * N + j] = K.sig[i][j];
      if (dg_end) dg_end[i * N + j] = K.endv[1][i] * K.endv[1][j];
      if (dg_der) dg_der[i * N + j] = K.dgder[i][j];
      if (dg_mat) dg_mat[i * N + j] = K.dgmat[i][j];
    }
  }
  return GPR_OK;
}

int gpr_flux(gpr_ctx *c, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_FLUX, ret, Q, nullptr, count, d, MP, c->V);
}
int gpr_source(gpr_ctx *c, double *ret, const double *Q, int64_t count, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_SOURCE, ret, Q, nullptr, count, 0, MP, c->V);
}
int gpr_bdot(gpr_ctx *c, double *ret, const double *Q, const double *x, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (!x) return fail(GPR_E_ARG, "null x");
  return pointwise(c, OP_BDOT, ret, Q, x, count, d, MP, c->V);
}
int gpr_max_abs_eigs(gpr_ctx *c, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP) {
  return pointwise(c, OP_EIGS, ret, Q, nullptr, count, d, MP, 1);
}
int gpr_thermo_acoustic_tensor(gpr_ctx *c, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  int n = c->thermal ? 4 : 3;
  return pointwise(c, OP_TAT, ret, Q, nullptr, count, d, MP, n * n);
}
int gpr_d_rus(gpr_ctx *c, double *ret, const double *qL, const double *qR, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (!qR) return fail(GPR_E_ARG, "null qR");
  return pointwise(c, OP_DRUS, ret, qL, qR, count, d, MP, c->V);
}
int gpr_bint(gpr_ctx *c, double *ret, const double *qL, const double *qR, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (!qR) return fail(GPR_E_ARG, "null qR");
  return pointwise(c, OP_BINT, ret, qL, qR, count, d, MP, c->V);
}
static int eig_pointwise(gpr_ctx *c, int op, double *ret, const double *qL, const double *qR, int64_t count, int d,
                         const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (!qR) return fail(GPR_E_ARG, "null qR");
  int32_t bad = 0;
  int rc = pointwise(c, op, ret, qL, qR, count, d, MP, c->V, 0.0, &bad);
  if (rc) return rc;
  if (bad) return fail(GPR_E_NONFINITE, "eigendecomposition: QR did not converge (result NaN)");
  return GPR_OK;
}
int gpr_d_roe(gpr_ctx *c, double *ret, const double *qL, const double *qR, int64_t count, int d, const gpr_par *MP) {
  return eig_pointwise(c, OP_DROE, ret, qL, qR, count, d, MP);
}
int gpr_d_osh(gpr_ctx *c, double *ret, const double *qL, const double *qR, int64_t count, int d, const gpr_par *MP) {
  return eig_pointwise(c, OP_DOSH, ret, qL, qR, count, d, MP);
}
int gpr_system_matrix(gpr_ctx *c, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_SYSMAT, ret, Q, nullptr, count, d, MP, (size_t)c->V * c->V);
}
int gpr_dfdp(gpr_ctx *c, double *ret, const double *Q, int64_t count, int d, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_DFDP, ret, Q, nullptr, count, d, MP, (size_t)c->V * c->V);
}
int gpr_dpdq(gpr_ctx *c, double *ret, const double *Q, int64_t count, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_DPDQ, ret, Q, nullptr, count, 0, MP, (size_t)c->V * c->V);
}
int gpr_renorm_distortion(gpr_ctx *c, double *u, int64_t ncell, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_RENORM, u, u, nullptr, ncell, 0, MP, c->V);
}
int gpr_reset_distortion(gpr_ctx *c, double *u, int64_t ncell, const gpr_par *MP) {
  if (!c) return fail(GPR_E_ARG, "null context");
  return pointwise(c, OP_RESET, u, u, nullptr, ncell, 0, MP, c->V);
}
int gpr_contorted(gpr_ctx *c, int32_t *result, const double *u, int64_t ncell, double tol) {
  if (!c || !result) return fail(GPR_E_ARG, "null argument");
  *result = 0;
  return pointwise(c, OP_CONTORTED, nullptr, u, nullptr, ncell, 0, nullptr, 0, tol, result);
}

int gpr_boundaries(gpr_ctx *c, const double *u, double *ub, const int32_t *nX, const int32_t *bcs) {
  if (!c || !u || !ub) return fail(GPR_E_ARG, "null argument");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  int4 bc;
  rc = check_bcs(c, bcs, bc);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0), nb = ext_cells(c, nX, c->N);
  double *du = (double *)buf(c, "u", nc * c->V * 8, rc);
  if (!du) return rc;
  double *dub = (double *)buf(c, "ub", nb * c->V * 8, rc);
  if (!dub) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, nc * c->V * 8, hipMemcpyHostToDevice, c->stream));
  rc = dispatch<BoundariesOp>(c, c, (const double *)du, dub, nX[0], c->ndim == 2 ? nX[1] : 1, bc);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ub, dub, nb * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_extend_mask(gpr_ctx *c, const uint8_t *mask, uint8_t *maskb, const int32_t *nX) {
  if (!c || !mask || !maskb) return fail(GPR_E_ARG, "null argument");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  host_extend_mask(c, mask, maskb, nX);
  return GPR_OK;
}

int gpr_weno_launcher(gpr_ctx *c, double *wh, const double *ub, const int32_t *nX, const gpr_par *MP) {
  if (!c || !wh || !ub) return fail(GPR_E_ARG, "null argument");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  (void)MP;  // PRIM_RECONSTRUCT = false (options.h:28): the material does not enter
  long nb = ext_cells(c, nX, c->N), ne = ext_cells(c, nX, 1);
  long nw = ne * (c->ndim == 2 ? c->N * c->N : c->N) * c->V;
  double *dub = (double *)buf(c, "ub", nb * c->V * 8, rc);
  if (!dub) return rc;
  double *dwh = (double *)buf(c, "wh", nw * 8, rc);
  if (!dwh) return rc;
  HIP_TRY(hipMemcpyAsync(dub, ub, nb * c->V * 8, hipMemcpyHostToDevice, c->stream));
  rc = dispatch<WenoOp>(c, c, (const double *)dub, dwh, nX[0], c->ndim == 2 ? nX[1] : 1);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(wh, dwh, nw * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_predictor(gpr_ctx *c, double *qh, const double *wh, double dt, const double *dX, int64_t ncell,
                  int stiff, int stiff_ig, const gpr_par *MP, const uint8_t *mask, int32_t *iters) {
  if (!c || !qh || !wh || !dX || !mask) return fail(GPR_E_ARG, "null argument");
  (void)stiff_ig;  // stiff_initial_guess is the standard guess (initial_guess.cpp:15-51)
  if (ncell < 0) return fail(GPR_E_ARG, "negative cell count");
  Mat m;
  int rc = make_mat(MP, m);
  if (rc) return rc;
  if (ncell == 0) return GPR_OK;
  long ne = ncell;
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  long nw = ne * NS * c->V, nq = ne * NS * c->N * c->V;
  double *dwh = (double *)buf(c, "wh", nw * 8, rc);
  if (!dwh) return rc;
  double *dqh = (double *)buf(c, "qh", nq * 8, rc);
  if (!dqh) return rc;
  uint8_t *dm = (uint8_t *)buf(c, "mask", ne, rc);
  if (!dm) return rc;
  int32_t *dit = nullptr;
  if (iters) {
    dit = (int32_t *)buf(c, "iters", ne * 4, rc);
    if (!dit) return rc;
  }
  HIP_TRY(hipMemcpyAsync(dwh, wh, nw * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dqh, qh, nq * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dm, mask, ne, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_fail, 0, sizeof(int), c->stream));
  PredArgs a = pred_args(c, m, dt, dX);
  if (stiff) a.mode = 1;
  a.qh = dqh;
  a.wh = dwh;
  a.mask = dm;
  a.iters = dit;
  a.ncell = ne;
  std::vector<float> whf;
  if (c->opts.precision == 1 && !stiff) {
    // the FP32-bulk form: wh rounded to float, Picard in FP32, the cells it
    // leaves solved again in FP64 by the fallback launch over a list
    whf.resize(nw);
    for (long i = 0; i < nw; i++) whf[i] = (float)wh[i];
    float *dwf = (float *)buf(c, "whf", nw * sizeof(float), rc);
    if (!dwf) return rc;
    int *dlist = (int *)buf(c, "plist", ne * sizeof(int), rc);
    if (!dlist) return rc;
    int *dlcnt = (int *)buf(c, "plcnt", sizeof(int), rc);
    if (!dlcnt) return rc;
    HIP_TRY(hipMemcpyAsync(dwf, whf.data(), nw * sizeof(float), hipMemcpyHostToDevice, c->stream));
    a.wh = dwf;
    a.list = dlist;
    a.list_count = dlcnt;
    rc = dispatch<PredictorF32Op>(c, c, a);
  } else {
    rc = dispatch<PredictorOp>(c, c, a);
  }
  if (rc) return rc;
  int nfail = 0;
  HIP_TRY(hipMemcpyAsync(qh, dqh, nq * 8, hipMemcpyDeviceToHost, c->stream));
  if (iters) HIP_TRY(hipMemcpyAsync(iters, dit, ne * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&nfail, c->d_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (nfail) return fail(GPR_E_PREDICTOR, std::to_string(nfail) + " cell(s): predictor converged in neither phase");
  return GPR_OK;
}

static int fv_host(gpr_ctx *c, double *u, const double *qh, const int32_t *nX, double dt, const double *dX,
                   int sources, int time, int flux, const gpr_par *MP, const uint8_t *mask, int parts);

int gpr_fv_launcher(gpr_ctx *c, double *u, const double *qh, const int32_t *nX, double dt, const double *dX,
                    int sources, int time, int flux, const gpr_par *MP, const uint8_t *mask) {
  return fv_host(c, u, qh, nX, dt, dX, sources, time, flux, MP, mask, 3);
}

int gpr_fv_centers(gpr_ctx *c, double *u, const double *rec, const int32_t *nX, double dt, const double *dX,
                   int sources, int time, const gpr_par *MP, const uint8_t *mask) {
  return fv_host(c, u, rec, nX, dt, dX, sources, time, 0, MP, mask, 1);
}

int gpr_fv_interfs(gpr_ctx *c, double *u, const double *rec, const int32_t *nX, double dt, const double *dX,
                   int time, int flux, const gpr_par *MP, const uint8_t *mask) {
  return fv_host(c, u, rec, nX, dt, dX, 1, time, flux, MP, mask, 2);
}

static int fv_host(gpr_ctx *c, double *u, const double *qh, const int32_t *nX, double dt, const double *dX,
                   int sources, int time, int flux, const gpr_par *MP, const uint8_t *mask, int parts) {
  if (!c || !u || !qh || !dX || !mask) return fail(GPR_E_ARG, "null argument");
  if (flux < 0 || flux > 2) return fail(GPR_E_ARG, "FLUX must be 0 (Rusanov), 1 (Roe) or 2 (Osher) (fluxes.h:7-9)");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0), ne = ext_cells(c, nX, 1);
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  long nq = ne * NS * (time ? c->N : 1) * c->V;
  double *du = (double *)buf(c, "u", nc * c->V * 8, rc);
  if (!du) return rc;
  double *dqh = (double *)buf(c, "qh", nq * 8, rc);
  if (!dqh) return rc;
  uint8_t *dm = (uint8_t *)buf(c, "mask", ne, rc);
  if (!dm) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, nc * c->V * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dqh, qh, nq * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dm, mask, ne, hipMemcpyHostToDevice, c->stream));
  FvArgs a = fv_args(c, m, dt, dX, nX);
  a.u = du;
  a.qh = dqh;
  a.mask = dm;
  a.time = time != 0;
  a.sources = sources != 0;
  a.parts = parts;
  a.flux = flux;
  rc = fv_buffers(c, a);
  if (rc) return rc;
  rc = dispatch<FvOp>(c, c, a);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(u, du, nc * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_timestep(gpr_ctx *c, double *dt_out, const double *u, const int32_t *nX, const double *dX, double CFL,
                 double t, double tf, int64_t count, const gpr_par *MP) {
  if (!c || !dt_out || !u || !dX) return fail(GPR_E_ARG, "null argument");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0);
  double *du = (double *)buf(c, "u", nc * c->V * 8, rc);
  if (!du) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, nc * c->V * 8, hipMemcpyHostToDevice, c->stream));
  StepState h;
  std::memset(&h, 0, sizeof h);
  h.t = t;
  h.tf = tf;
  h.cfl = CFL;
  h.count = count;
  HIP_TRY(hipMemcpyAsync(c->d_st, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
  rc = dispatch<CflOp>(c, c, (const double *)du, nc, dX[0], c->ndim == 2 ? dX[1] : 1.0, m, &c->d_st->maxspeed);
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, c->stream, c->d_st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&h, c->d_st, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *dt_out = h.dt;
  return GPR_OK;
}

int gpr_ader_stepper(gpr_ctx *c, double *u, const double *ub, const int32_t *nX, double dt, const double *dX,
                     int stiff, int flux, const gpr_par *MP, const uint8_t *mask) {
  if (!c || !u || !ub || !dX || !mask) return fail(GPR_E_ARG, "null argument");
  if (flux < 0 || flux > 2) return fail(GPR_E_ARG, "FLUX must be 0 (Rusanov), 1 (Roe) or 2 (Osher) (fluxes.h:7-9)");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0), ne = ext_cells(c, nX, 1), nb = ext_cells(c, nX, c->N);
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  long nw = ne * NS * c->V, nq = ne * NS * c->N * c->V;
  double *du = (double *)buf(c, "u", nc * c->V * 8, rc);
  if (!du) return rc;
  double *dub = (double *)buf(c, "ub", nb * c->V * 8, rc);
  if (!dub) return rc;
  double *dwh = (double *)buf(c, "wh", nw * 8, rc);
  if (!dwh) return rc;
  double *dqh = (double *)buf(c, "qh", nq * 8, rc);
  if (!dqh) return rc;
  uint8_t *dm = (uint8_t *)buf(c, "mask", ne, rc);
  if (!dm) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, nc * c->V * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dub, ub, nb * c->V * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dm, mask, ne, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(dqh, 0, nq * 8, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_fail, 0, sizeof(int), c->stream));
  int ny = c->ndim == 2 ? nX[1] : 1;
  rc = dispatch<WenoOp>(c, c, (const double *)dub, dwh, nX[0], ny);
  if (rc) return rc;
  PredArgs pa = pred_args(c, m, dt, dX);
  if (stiff) pa.mode = 1;
  pa.qh = dqh;
  pa.wh = dwh;
  pa.mask = dm;
  pa.ncell = ne;
  rc = dispatch<PredictorOp>(c, c, pa);
  if (rc) return rc;
  FvArgs fa = fv_args(c, m, dt, dX, nX);
  fa.flux = flux;
  fa.u = du;
  fa.qh = dqh;
  fa.mask = dm;
  rc = fv_buffers(c, fa);
  if (rc) return rc;
  rc = dispatch<FvOp>(c, c, fa);
  if (rc) return rc;
  int nfail = 0;
  HIP_TRY(hipMemcpyAsync(u, du, nc * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&nfail, c->d_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (nfail) return fail(GPR_E_PREDICTOR, std::to_string(nfail) + " cell(s): predictor converged in neither phase");
  return GPR_OK;
}

int gpr_ader_stepper_noalloc(gpr_ctx *c, double *d_u, const double *d_ub, const int32_t *nX, double dt,
                             const double *dX, int stiff, int flux, const gpr_par *MP, const uint8_t *d_mask,
                             int32_t *d_nfail) {
  if (!c || !d_u || !d_ub || !dX || !d_mask || !d_nfail) return fail(GPR_E_ARG, "null argument");
  if (flux != 0)
    return fail(GPR_E_UNSUPPORTED, "FLUX = 1, 2 need per-lane scratch that the HIP runtime allocates: use gpr_ader_stepper");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  if (!c->scratch) return fail(GPR_E_HIP, "the context has no scratch address");
  const int ny = c->ndim == 2 ? nX[1] : 1;
  const long ne = ext_cells(c, nX, 1);
  const int NS = c->ndim == 2 ? c->N * c->N : c->N;
  const long nw = ne * NS * c->V, nq = ne * NS * c->N * c->V;
  const long nux = c->ndim == 2 ? (long)(nX[0] + 2) * (ny + 2 * c->N) * c->N * c->V : 0;
  const long nfx = (long)(nX[0] + 1) * ny * 2 * c->V;
  const long nfy = c->ndim == 2 ? (long)nX[0] * (ny + 1) * 2 * c->V : 2 * c->V;
  // carve the scratch into 256-byte aligned pieces
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t o_fail = take(sizeof(int)), o_ux = take(nux * 8), o_wh = take(nw * 8), o_qh = take(nq * 8),
               o_fx = take(nfx * 8), o_fy = take(nfy * 8);
  if (off > (size_t)GPR_NOALLOC_SCRATCH_BYTES)
    return fail(GPR_E_ARG, "grid too large for the allocation-free stepper: it needs " + std::to_string(off) +
                               " bytes of scratch, the maximum is " + std::to_string(GPR_NOALLOC_SCRATCH_BYTES));
  unsigned char *base = (unsigned char *)c->scratch;
  int *dfail = (int *)(base + o_fail);
  double *dux = (double *)(base + o_ux), *dwh = (double *)(base + o_wh), *dqh = (double *)(base + o_qh);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, c->stream, dfail);
  hipLaunchKernelGGL(k_zero_doubles, dim3(grid_for(nq, 256)), dim3(256), 0, c->stream, dqh, nq);
  HIP_TRY(hipGetLastError());
  rc = dispatch<WenoOp>(c, c, d_ub, dwh, nX[0], ny, dux);
  if (rc) return rc;
  PredArgs pa = pred_args(c, m, dt, dX);
  if (stiff) pa.mode = 1;
  pa.qh = dqh;
  pa.wh = dwh;
  pa.mask = d_mask;
  pa.ncell = ne;
  pa.fail = dfail;
  rc = dispatch<PredictorOp>(c, c, pa);
  if (rc) return rc;
  FvArgs fa = fv_args(c, m, dt, dX, nX);
  fa.flux = flux;
  fa.u = d_u;
  fa.qh = dqh;
  fa.mask = d_mask;
  fa.fx = (double *)(base + o_fx);
  fa.fy = (double *)(base + o_fy);
  rc = dispatch<FvOp>(c, c, fa);
  if (rc) return rc;
  hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, c->stream, d_nfail, (const int *)dfail);
  HIP_TRY(hipGetLastError());
  return GPR_OK;
}

int gpr_split_stepper(gpr_ctx *c, double *u, double *ub, const int32_t *nX, double dt, const double *dX,
                      int half_step, int flux, const gpr_par *MP, const uint8_t *mask) {
  if (!c || !u || !ub || !dX || !mask) return fail(GPR_E_ARG, "null argument");
  if (flux < 0 || flux > 2) return fail(GPR_E_ARG, "FLUX must be 0 (Rusanov), 1 (Roe) or 2 (Osher) (fluxes.h:7-9)");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0), ne = ext_cells(c, nX, 1), nb = ext_cells(c, nX, c->N);
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  long nw = ne * NS * c->V;
  double *du = (double *)buf(c, "u", nc * c->V * 8, rc);
  if (!du) return rc;
  double *dub = (double *)buf(c, "ub", nb * c->V * 8, rc);
  if (!dub) return rc;
  double *dwh = (double *)buf(c, "wh", nw * 8, rc);
  if (!dwh) return rc;
  uint8_t *dm = (uint8_t *)buf(c, "mask", ne, rc);
  if (!dm) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, nc * c->V * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dub, ub, nb * c->V * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dm, mask, ne, hipMemcpyHostToDevice, c->stream));
  int ny = c->ndim == 2 ? nX[1] : 1;
  // u and ub are the caller's and need not agree: two ODE launches
  rc = dispatch<OdeOp>(c, c, du, nc, (double *)nullptr, nX[0], ny, (const double *)nullptr, dt, 0.5, m.p);
  if (!rc) rc = dispatch<OdeOp>(c, c, dub, nb, (double *)nullptr, nX[0], ny, (const double *)nullptr, dt, 0.5, m.p);
  if (!rc) rc = dispatch<WenoOp>(c, c, (const double *)dub, dwh, nX[0], ny);
  if (!rc && half_step)
    rc = dispatch<MidstepOp>(c, c, dwh, ne, (const uint8_t *)dm, (const double *)nullptr, dt, dX[0],
                             c->ndim == 2 ? dX[1] : 1.0, m);
  if (rc) return rc;
  FvArgs fa = fv_args(c, m, dt, dX, nX);
  fa.flux = flux;
  fa.u = du;
  fa.qh = dwh;
  fa.mask = dm;
  fa.time = 0;
  fa.sources = 0;
  rc = fv_buffers(c, fa);
  if (!rc) rc = dispatch<FvOp>(c, c, fa);
  if (!rc) rc = dispatch<OdeOp>(c, c, du, nc, (double *)nullptr, nX[0], ny, (const double *)nullptr, dt, 0.5, m.p);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(u, du, nc * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(ub, dub, nb * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_ode_launcher(gpr_ctx *c, double *u, int64_t ncell, double dt, const gpr_par *MP) {
  if (!c || !u) return fail(GPR_E_ARG, "null argument");
  if (ncell < 0) return fail(GPR_E_ARG, "negative cell count");
  Mat m;
  int rc = make_mat(MP, m);
  if (rc) return rc;
  if (ncell == 0) return GPR_OK;
  double *du = (double *)buf(c, "u", ncell * c->V * 8, rc);
  if (!du) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, ncell * c->V * 8, hipMemcpyHostToDevice, c->stream));
  rc = dispatch<OdeOp>(c, c, du, (long)ncell, (double *)nullptr, 1, 1, (const double *)nullptr, dt, 1.0, m.p);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(u, du, ncell * c->V * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_analytic_solver(gpr_ctx *c, double *Q, int64_t count, double dt, const gpr_par *MP, int part) {
  if (!c || !Q) return fail(GPR_E_ARG, "null argument");
  if (count < 0) return fail(GPR_E_ARG, "negative count");
  if (part != 0 && part != 1) return fail(GPR_E_ARG, "part must be 0 (distortion) or 1 (thermal)");
  if (part == 1 && !c->thermal) return fail(GPR_E_UNSUPPORTED, "the thermal solver needs THERMAL (V = 17)");
  Mat m;
  int rc = make_mat(MP, m);
  if (rc) return rc;
  return pointwise(c, part == 0 ? OP_ODE_DISTORTION : OP_ODE_THERMAL, Q, Q, nullptr, count, 0, MP, c->V, dt);
}

int gpr_midstepper(gpr_ctx *c, double *wh, double dt, const double *dX, int64_t ncell, const gpr_par *MP,
                   const uint8_t *mask) {
  if (!c || !wh || !dX || !mask) return fail(GPR_E_ARG, "null argument");
  if (ncell < 0) return fail(GPR_E_ARG, "negative cell count");
  Mat m;
  int rc = make_mat(MP, m);
  if (rc) return rc;
  if (ncell == 0) return GPR_OK;
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  long nw = ncell * NS * c->V;
  double *dwh = (double *)buf(c, "wh", nw * 8, rc);
  if (!dwh) return rc;
  uint8_t *dm = (uint8_t *)buf(c, "mask", ncell, rc);
  if (!dm) return rc;
  HIP_TRY(hipMemcpyAsync(dwh, wh, nw * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dm, mask, ncell, hipMemcpyHostToDevice, c->stream));
  rc = dispatch<MidstepOp>(c, c, dwh, (long)ncell, (const uint8_t *)dm, (const double *)nullptr, dt, dX[0],
                           c->ndim == 2 ? dX[1] : 1.0, m);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(wh, dwh, nw * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

// ------------------------------------------------------------ device memory
int gpr_dev_malloc(gpr_ctx *c, void **d, int64_t bytes) {
  if (!c || !d || bytes <= 0) return fail(GPR_E_ARG, "gpr_dev_malloc: bad argument");
  *d = nullptr;
  HIP_TRY(hipMalloc(d, (size_t)bytes));
  return GPR_OK;
}

int gpr_dev_free(gpr_ctx *c, void *d) {
  if (!c) return fail(GPR_E_ARG, "null ctx");
  if (d) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(d));
  }
  return GPR_OK;
}

int gpr_dev_memcpy(gpr_ctx *c, void *dst, const void *src, int64_t bytes, int kind) {
  if (!c || !dst || !src || bytes < 0 || kind < 0 || kind > 2) return fail(GPR_E_ARG, "gpr_dev_memcpy: bad argument");
  const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
  HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, k[kind], c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

// ------------------------------------------------------- device-resident loop
int gpr_dev_begin(gpr_ctx *c, double *d_u, const int32_t *nX, const double *dX, double CFL, const int32_t *bcs,
                  double t0, double tf, int64_t count0, const gpr_par *MP) {
  if (!c || !d_u || !dX) return fail(GPR_E_ARG, "null argument");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  int4 bc;
  rc = check_bcs(c, bcs, bc);
  if (rc) return rc;
  Mat m;
  rc = make_mat(MP, m);
  if (rc) return rc;
  if (c->opts.split && c->opts.precision != 0)
    return fail(GPR_E_UNSUPPORTED, "the split-WENO step runs in FP64 only");
  c->run_ready = false;
  c->run_u = d_u;
  c->run_nx = nX[0];
  c->run_ny = c->ndim == 2 ? nX[1] : 1;
  c->run_dx = dX[0];
  c->run_dy = c->ndim == 2 ? dX[1] : 1.0;
  c->run_bc = bc;
  c->run_mat = m;
  long nc = ext_cells(c, nX, 0), ne = ext_cells(c, nX, 1), nb = ext_cells(c, nX, c->N);
  int NS = c->ndim == 2 ? c->N * c->N : c->N;
  if (!buf(c, "ub", nb * c->V * 8, rc)) return rc;
  if (!buf(c, "wh", ne * NS * c->V * 8, rc)) return rc;
  if (!c->opts.split && !buf(c, "qh", ne * NS * c->N * c->V * 8, rc)) return rc;
  // the run's mask is set here once and read by every step: a buffer of its
  // own, so that host-pointer stage calls on this context cannot change it
  uint8_t *dm = (uint8_t *)buf(c, "run_mask", ne, rc);
  if (!dm) return rc;
  if (c->ndim == 2 && !buf(c, "ux", (long)(nX[0] + 2) * (nX[1] + 2 * c->N) * c->N * c->V * 8, rc)) return rc;
  FvArgs fa = fv_args(c, m, 0, dX, nX);
  rc = fv_buffers(c, fa);
  if (rc) return rc;
  std::vector<uint8_t> m1(nc, 1), mb(ne, 0);
  host_extend_mask(c, m1.data(), mb.data(), nX);
  HIP_TRY(hipMemcpyAsync(dm, mb.data(), ne, hipMemcpyHostToDevice, c->stream));
  StepState h;
  std::memset(&h, 0, sizeof h);
  h.t = t0;
  h.tf = tf;
  h.cfl = CFL;
  h.count = count0;
  h.contorted_tol = c->opts.contorted_tol;
  HIP_TRY(hipMemcpyAsync(c->d_st, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_fail, 0, sizeof(int), c->stream));
  rc = dispatch<CflOp>(c, c, (const double *)d_u, nc, c->run_dx, c->run_dy, m, &c->d_st->maxspeed);
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, c->stream, c->d_st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));  // mask vector leaves scope
  c->run_ready = true;
  return GPR_OK;
}

int gpr_dev_steps(gpr_ctx *c, int64_t nsteps) { return dev_steps_impl(c, nsteps, nullptr); }

int gpr_dev_steps_dt(gpr_ctx *c, const double *dts, int64_t nsteps) {
  if (!dts) return fail(GPR_E_ARG, "null time steps");
  return dev_steps_impl(c, nsteps, dts);
}

}  // extern "C"

namespace {
int dev_steps_impl(gpr_ctx *c, int64_t nsteps, const double *dts) {
  if (!c || !c->run_ready) return fail(GPR_E_ARG, "gpr_dev_begin has not been called");
  int rc;
  int nX[2] = {c->run_nx, c->run_ny};
  double dX[2] = {c->run_dx, c->run_dy};
  long nc = ext_cells(c, nX, 0), ne = ext_cells(c, nX, 1), nb = ext_cells(c, nX, c->N);
  const bool split = c->opts.split != 0;
  if (split && c->opts.precision != 0) return fail(GPR_E_UNSUPPORTED, "the split-WENO step runs in FP64 only");
  double *dub = (double *)c->bufs["ub"].first;
  double *dwh = (double *)c->bufs["wh"].first;
  double *dqh = split ? nullptr : (double *)c->bufs["qh"].first;
  uint8_t *dm = (uint8_t *)c->bufs["run_mask"].first;
  if (!split && !dqh) return fail(GPR_E_ARG, "gpr_dev_begin ran for the split step; call it again");
  for (int64_t k = 0; k < nsteps; k++) {
    hipEvent_t *ev = nullptr;
    if (c->timing) {
      while (c->events.size() < c->ev_used + 6) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        c->events.push_back(e);
      }
      ev = &c->events[c->ev_used];
      c->ev_used += 6;
      c->timed_steps++;
      HIP_TRY(hipEventRecord(ev[0], c->stream));
    }
    if (dts) hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->stream, c->d_st, dts[k]);
    rc = dispatch<BoundariesOp>(c, c, (const double *)c->run_u, dub, c->run_nx, c->run_ny, c->run_bc);
    if (rc) return rc;
    // split step (steppers.cpp:50-66): the half-step ODE on u and ub, one launch
    if (split) rc = dispatch<OdeOp>(c, c, dub, nb, c->run_u, c->run_nx, c->run_ny, (const double *)&c->d_st->dt, 0.0,
                                    0.5, c->run_mat.p);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[1], c->stream));
    const bool f32 = c->opts.precision == 1;
    float *dwhf = nullptr;
    int *dlist = (int *)buf(c, "plist", ne * sizeof(int), rc);
    if (!dlist) return rc;
    int *dlcnt = (int *)buf(c, "plcnt", sizeof(int), rc);
    if (!dlcnt) return rc;
    if (f32) {
      int NS = c->ndim == 2 ? c->N * c->N : c->N;
      dwhf = (float *)buf(c, "whf", ne * NS * c->V * sizeof(float), rc);
      if (!dwhf) return rc;
      rc = dispatch<WenoF32Op>(c, c, (const double *)dub, dwhf, c->run_nx, c->run_ny);
    } else {
      rc = dispatch<WenoOp>(c, c, (const double *)dub, dwh, c->run_nx, c->run_ny);
    }
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[2], c->stream));
    if (split) {
      if (c->opts.half_step)
        rc = dispatch<MidstepOp>(c, c, dwh, ne, (const uint8_t *)dm, (const double *)&c->d_st->dt, 0.0, dX[0], dX[1],
                                 c->run_mat);
      if (rc) return rc;
      if (ev) HIP_TRY(hipEventRecord(ev[3], c->stream));
      FvArgs fs = fv_args(c, c->run_mat, 0, dX, nX);
      fs.flux = c->opts.flux;
      fs.st = c->d_st;
      fs.u = c->run_u;
      fs.qh = dwh;
      fs.mask = dm;
      fs.post = 1;
      fs.time = 0;
      fs.sources = 0;
      fs.ode = 1;
      rc = fv_buffers(c, fs);
      if (rc) return rc;
      rc = dispatch<FvOp>(c, c, fs);
      if (rc) return rc;
      if (ev) HIP_TRY(hipEventRecord(ev[4], c->stream));
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_st);
      rc = dispatch<ResetOp>(c, c, c->run_u, nc, dX[0], dX[1], c->run_mat, c->d_st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, c->stream, c->d_st);
      HIP_TRY(hipGetLastError());
      if (ev) HIP_TRY(hipEventRecord(ev[5], c->stream));
      continue;
    }
    PredArgs pa = pred_args(c, c->run_mat, 0, dX);
    pa.dt_dev = &c->d_st->dt;
    pa.qh = dqh;
    pa.wh = f32 ? (const void *)dwhf : (const void *)dwh;
    pa.list = dlist;
    pa.list_count = dlcnt;
    pa.mask = dm;
    pa.ncell = ne;
    pa.itc = c->timing ? c->d_itc : nullptr;
    // the one-launch FP64 predictor hands the FV corrector face states and
    // cell integrals instead of qh (same values).  Measured per material
    // kind: on the 2D P2 ideal-gas Newtonian fluid the epilogue costs less
    // than the qh round trip it removes (FV 1.04 -> 0.71 ms, predictor
    // +0.07 ms at 512^2); on the Godunov-Romenski plastic solid the epilogue
    // costs more than k_fv_cells (1.85 against 1.31 ms at 1024^2), so that
    // kind and the others keep the qh path.
    const bool fused = !f32 && c->opts.mode == 0 && c->ndim == 2 && c->N == 3 && !c->thermal &&
                       c->opts.flux == 0 && material_kind(c->run_mat) == MK_GAS_NEWTONIAN;
    if (fused) {
      const long nps = (long)c->N * (c->ndim == 2 ? c->N : 1);
      pa.fs = (double *)buf(c, "fs", ne * 2 * c->ndim * nps * c->V * sizeof(double), rc);
      if (!pa.fs) return rc;
      pa.cint = (double *)buf(c, "cint", nc * c->V * sizeof(double), rc);
      if (!pa.cint) return rc;
      pa.fnx = c->run_nx;
      pa.fny = c->ndim == 2 ? c->run_ny : 1;
      pa.qh = nullptr;
    }
    if (f32)
      rc = dispatch<PredictorF32Op>(c, c, pa);
    else if (c->opts.mode == 2)
      rc = dispatch<PredictorSplitOp>(c, c, pa);
    else
      rc = dispatch<PredictorOp>(c, c, pa);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[3], c->stream));
    FvArgs fa = fv_args(c, c->run_mat, 0, dX, nX);
    fa.flux = c->opts.flux;
    fa.st = c->d_st;
    fa.u = c->run_u;
    fa.qh = dqh;
    fa.mask = dm;
    fa.post = 0;  // the post-step is launched (and timed) on its own below
    if (fused) {
      fa.fs = pa.fs;
      fa.cint = pa.cint;
    }
    rc = fv_buffers(c, fa);
    if (rc) return rc;
    rc = dispatch<FvOp>(c, c, fa);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[4], c->stream));
    rc = dispatch<PostOp>(c, c, fa);
    if (rc) return rc;
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_st);
    rc = dispatch<ResetOp>(c, c, c->run_u, nc, dX[0], dX[1], c->run_mat, c->d_st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, c->stream, c->d_st);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[5], c->stream));
  }
  return GPR_OK;
}
}  // namespace

extern "C" {

int gpr_dev_set_timing(gpr_ctx *c, int enable) {
  if (!c) return fail(GPR_E_ARG, "null context");
  if (enable) {
    HIP_TRY(hipMemsetAsync(c->d_itc, 0, sizeof(unsigned long long) * ITC_SLOTS * ITC_STRIDE, c->stream));
    c->ev_used = 0;
    c->timed_steps = 0;
  }
  c->timing = enable != 0;
  return GPR_OK;
}

int gpr_dev_timing(gpr_ctx *c, double *ms, int64_t *steps, int64_t *iters, int64_t *cells) {
  if (!c) return fail(GPR_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  double tot[5] = {0, 0, 0, 0, 0};
  for (size_t s = 0; s + 6 <= c->ev_used; s += 6)
    for (int p = 0; p < 5; p++) {
      float f = 0;
      HIP_TRY(hipEventElapsedTime(&f, c->events[s + p], c->events[s + p + 1]));
      tot[p] += f;
    }
  unsigned long long h[4] = {0, 0, 0, 0};
  {
    std::vector<unsigned long long> slots((size_t)ITC_SLOTS * ITC_STRIDE);
    HIP_TRY(hipMemcpy(slots.data(), c->d_itc, slots.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int s = 0; s < ITC_SLOTS; s++)
      for (int k = 0; k < 4; k++) h[k] += slots[(size_t)s * ITC_STRIDE + k];
  }
  if (ms)
    for (int p = 0; p < 5; p++) ms[p] = tot[p];
  if (steps) *steps = c->timed_steps;
  if (iters) {
    iters[0] = (int64_t)h[0];
    iters[1] = (int64_t)h[1];
    iters[2] = (int64_t)h[3];
  }
  if (cells) *cells = (int64_t)h[2];
  return GPR_OK;
}

int gpr_dev_status(gpr_ctx *c, double *t, int64_t *count, double *dt, int32_t *pred_fail, int32_t *nonfinite) {
  if (!c) return fail(GPR_E_ARG, "null context");
  StepState h;
  int nfail = 0;
  HIP_TRY(hipMemcpyAsync(&h, c->d_st, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&nfail, c->d_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (t) *t = h.t;
  if (count) *count = h.count;
  if (dt) *dt = h.dt;
  if (pred_fail) *pred_fail = nfail;
  if (nonfinite) *nonfinite = h.nonfinite;
  return GPR_OK;
}

int gpr_iterator(gpr_ctx *c, double *u, double tf, const int32_t *nX, const double *dX, double CFL,
                 const int32_t *bcs, int split, int half_step, int stiff, int flux, const gpr_par *MPs, int nmat,
                 double contorted_tol, double *snapshots, int32_t *nsnap, int64_t *steps) {
  if (!c || !u || !dX || !MPs) return fail(GPR_E_ARG, "null argument");
  if (flux < 0 || flux > 2) return fail(GPR_E_ARG, "FLUX must be 0 (Rusanov), 1 (Roe) or 2 (Osher) (fluxes.h:7-9)");
  if (nmat != 1) return fail(GPR_E_UNSUPPORTED, "one material (LSET = 0) is supported");
  int rc = check_nX(c, nX);
  if (rc) return rc;
  long nc = ext_cells(c, nX, 0);
  size_t bytes = nc * c->V * 8;
  double *du = (double *)buf(c, "it_u", bytes, rc);
  if (!du) return rc;
  double *dprev = (double *)buf(c, "it_prev", bytes, rc);
  if (!dprev) return rc;
  HIP_TRY(hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, c->stream));
  gpr_solver_opts saved = c->opts;
  c->opts.contorted_tol = contorted_tol;
  if (stiff) c->opts.mode = 1;
  c->opts.split = split != 0;
  c->opts.half_step = half_step != 0;
  c->opts.flux = flux;
  if (split) c->opts.precision = 0;
  rc = gpr_dev_begin(c, du, nX, dX, CFL, bcs, 0.0, tf, 0, MPs);
  if (rc) {
    c->opts = saved;
    return rc;
  }

  int push = 0;
  int64_t k = 0;
  int status = GPR_OK;
  double t = 0;
  // one step at a time: snapshots at each 1% of tf and the NaN exit need t
  while (t < tf) {
    HIP_TRY(hipMemcpyAsync(dprev, du, bytes, hipMemcpyDeviceToDevice, c->stream));
    rc = gpr_dev_steps(c, 1);
    if (rc) {
      c->opts = saved;
      return rc;
    }
    int64_t cnt;
    int32_t pf, nf;
    double dtl;
    rc = gpr_dev_status(c, &t, &cnt, &dtl, &pf, &nf);
    if (rc) {
      c->opts = saved;
      return rc;
    }
    k++;
    if (pf) status = fail(GPR_E_PREDICTOR, std::to_string(pf) + " cell(s): predictor converged in neither phase");
    if (nf) {
      if (snapshots && push + 2 <= 100) {
        HIP_TRY(hipMemcpyAsync(snapshots + (size_t)push * nc * c->V, dprev, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(snapshots + (size_t)(push + 1) * nc * c->V, du, bytes, hipMemcpyDeviceToHost,
                               c->stream));
        push += 2;
      }
      status = fail(GPR_E_NONFINITE, "state became non-finite");
      break;
    }
    if (snapshots && t >= double(push + 1) / 100.0 * tf && push < 99) {
      HIP_TRY(hipMemcpyAsync(snapshots + (size_t)push * nc * c->V, du, bytes, hipMemcpyDeviceToHost, c->stream));
      push++;
    }
  }
  if (snapshots && status != GPR_E_NONFINITE) {
    HIP_TRY(hipMemcpyAsync(snapshots + (size_t)99 * nc * c->V, du, bytes, hipMemcpyDeviceToHost, c->stream));
    push = std::max(push, 100);
  }
  HIP_TRY(hipMemcpyAsync(u, du, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->opts = saved;
  if (nsnap) *nsnap = push;
  if (steps) *steps = k;
  return status;
}

}  // extern "C"

// ===================================================================== batch
// Many independent 1D problems in one set of launches per stage; the step
// loop is captured in a hipGraph and replayed.
struct BatchHost {
  int count = 0;
  std::vector<BatchInst> inst;
  std::vector<long> u0s, b0s, e0s;
  long nu = 0, nb = 0, ne = 0, nf = 0;
  int nblk_p = 0, nblk_c = 0, nblk_f = 0;
  BatchArgs args{};
  std::vector<void *> allocs;
  int block = 256;  // block size of the element-wise kernels
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  bool ready = false;
  int view_count = 0;  // instances the launches cover
  long vnu = 0, vnb = 0, vne = 0;  // their cells, ghost-extended rows, extended cells
  int mk = MK_ANY;  // the material kind every instance has (MK_ANY if they differ)
};

namespace {

template <int N, bool TH> struct BatchStep {
  static int run(gpr_ctx *c, BatchHost *B) {
    constexpr int V = Dims<TH>::V;
    const BatchArgs &b = B->args;
    const int blk = B->block;
    hipLaunchKernelGGL((kb_weno<N, V>), dim3(grid_for(B->vne * V, blk)), dim3(blk), 0, c->stream, b);
    PredArgs pa = pred_args(c, B->inst[0].m, 0.0, &B->inst[0].dx);
    pa.qh = b.qh;
    pa.wh = b.wh;
    pa.ncell = B->ne;
    pa.binst = b.inst;
    pa.pblk = (const int *)B->allocs[3];
    pa.bst = b.st;
    // P2 batches of one material kind without heat conduction run the
    // predictor and the faces compiled for that kind (the same values)
    // the ideal-gas Newtonian kind also fuses the corrector into the
    // predictor's epilogue: face states in qh's space and cell integrals,
    // read by the faces and by the post-step's cell update (the same values)
    bool kinded = false;
    if constexpr (N == 3 && !TH) {
      if (B->mk == MK_GAS_NEWTONIAN) {
        pa.qh = nullptr;
        pa.fs = (double *)b.fs;
        pa.cint = (double *)b.cint;
        hipLaunchKernelGGL((k_predictor<N, TH, 1, double, double, true, KIND_BOTH, MK_GAS_NEWTONIAN>),
                           dim3(B->nblk_p), dim3(WAVE), 0, c->stream, pa);
        hipLaunchKernelGGL((kb_faces<N, TH, MK_GAS_NEWTONIAN, true>), dim3(B->nblk_f), dim3(WAVE), 0, c->stream,
                           b);
        kinded = true;
      }
    }
    if (!kinded) {
      hipLaunchKernelGGL((k_predictor<N, TH, 1, double, double, true>), dim3(B->nblk_p), dim3(WAVE), 0, c->stream,
                         pa);
      hipLaunchKernelGGL((kb_faces<N, TH>), dim3(B->nblk_f), dim3(WAVE), 0, c->stream, b);
      hipLaunchKernelGGL((kb_cells<N, TH>), dim3(B->nblk_c), dim3(WAVE), 0, c->stream, b);
    }
    if (kinded)  // the fused corrector's cell update, then the post-step
      hipLaunchKernelGGL((kb_post<TH, true>), dim3((B->vnu + blk - 1) / blk), dim3(blk), 0, c->stream, b, 0);
    else
      hipLaunchKernelGGL((kb_post<TH>), dim3((B->vnu + blk - 1) / blk), dim3(blk), 0, c->stream, b, 0);
    hipLaunchKernelGGL(kb_advance, dim3((B->count + 63) / 64), dim3(64), 0, c->stream, b);
    hipLaunchKernelGGL((kb_reset<TH>), dim3((B->vnu + blk - 1) / blk), dim3(blk), 0, c->stream, b);
    hipLaunchKernelGGL(kb_finalize, dim3((B->count + 63) / 64), dim3(64), 0, c->stream, b);
    HIP_TRY(hipGetLastError());
    return GPR_OK;
  }
};

int batch_step(gpr_ctx *c, BatchHost *B) {
  if (c->N == 3 && !c->thermal) return BatchStep<3, false>::run(c, B);
  if (c->N == 3 && c->thermal) return BatchStep<3, true>::run(c, B);
  if (c->N == 2 && !c->thermal) return BatchStep<2, false>::run(c, B);
  return BatchStep<2, true>::run(c, B);
}

int batch_first_dt(gpr_ctx *c, BatchHost *B) {
  const int blk = B->block;
  if (c->thermal)
    hipLaunchKernelGGL((kb_post<true>), dim3((B->vnu + blk - 1) / blk), dim3(blk), 0, c->stream, B->args, 1);
  else
    hipLaunchKernelGGL((kb_post<false>), dim3((B->vnu + blk - 1) / blk), dim3(blk), 0, c->stream, B->args, 1);
  hipLaunchKernelGGL(kb_finalize, dim3((B->count + 63) / 64), dim3(64), 0, c->stream, B->args);
  HIP_TRY(hipGetLastError());
  return GPR_OK;
}

void batch_drop_graph(BatchHost *B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  B->exec = nullptr;
  B->graph_steps = 0;
}

// The block -> instance maps of the predictor, face and cell launches, over
// the instances with on[k] set: each such instance gets its blocks (and its
// blk_p / blk_c / blk_f), the others none, so a launch covers running
// instances only.  A map that would be empty gets one block of instance 0,
// whose work the kernels skip once it is done.  Waits for the stream.
int batch_build_view(gpr_ctx *c, BatchHost *B, const std::vector<char> &on) {
  const int N = c->N, NN = N * N;
  const int CPW = WAVE / NN, FPW = WAVE / N;
  std::vector<int> pblk, cblk, fblk;
  int n_on = 0;
  for (int k = 0; k < B->count; k++) {
    if (!on[k]) continue;
    n_on++;
    BatchInst &I = B->inst[k];
    I.blk_p = (int)pblk.size();
    for (int b = 0; b < (I.nx + 2 + CPW - 1) / CPW; b++) pblk.push_back(k);
    I.blk_c = (int)cblk.size();
    for (int b = 0; b < (I.nx + CPW - 1) / CPW; b++) cblk.push_back(k);
    I.blk_f = (int)fblk.size();
    for (int b = 0; b < (I.nx + 1 + FPW - 1) / FPW; b++) fblk.push_back(k);
  }
  if (pblk.empty()) {
    B->inst[0].blk_p = B->inst[0].blk_c = B->inst[0].blk_f = 0;
    pblk.push_back(0);
    cblk.push_back(0);
    fblk.push_back(0);
  }
  // the element-wise launches' index space: the covered instances' cells,
  // rows and extended cells back to back (instance 0 alone if none)
  std::vector<int> vinst;
  for (int k = 0; k < B->count; k++)
    if (on[k]) vinst.push_back(k);
  if (vinst.empty()) vinst.push_back(0);
  const int nv = (int)vinst.size();
  std::vector<long> voff(3 * (B->count + 1), 0);
  long *vu = voff.data(), *vb = vu + (B->count + 1), *ve = vb + (B->count + 1);
  for (int j = 0; j < nv; j++) {
    const BatchInst &I = B->inst[vinst[j]];
    vu[j + 1] = vu[j] + I.nx;
    vb[j + 1] = vb[j] + I.nx + 2 * N;
    ve[j + 1] = ve[j] + I.nx + 2;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[11], vinst.data(), sizeof(int) * nv, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[12], voff.data(), sizeof(long) * voff.size(), hipMemcpyHostToDevice, c->stream));
  B->args.vinst = (const int *)B->allocs[11];
  B->args.vcount = nv;
  B->args.vu0s = (const long *)B->allocs[12];
  B->args.vb0s = B->args.vu0s + (B->count + 1);
  B->args.ve0s = B->args.vb0s + (B->count + 1);
  B->vnu = vu[nv];
  B->vnb = vb[nv];
  B->vne = ve[nv];
  HIP_TRY(hipMemcpyAsync(B->allocs[0], B->inst.data(), sizeof(BatchInst) * B->count, hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[3], pblk.data(), sizeof(int) * pblk.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[4], cblk.data(), sizeof(int) * cblk.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[5], fblk.data(), sizeof(int) * fblk.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  batch_drop_graph(B);  // launch sizes and the view count are in the captured launches
  B->nblk_p = (int)pblk.size();
  B->nblk_c = (int)cblk.size();
  B->nblk_f = (int)fblk.size();
  B->view_count = n_on;
  return GPR_OK;
}

}  // namespace

extern "C" {

int gpr_batch_destroy(gpr_ctx *c) {
  if (!c) return fail(GPR_E_ARG, "null context");
  BatchHost *B = c->batch;
  if (!B) return GPR_OK;
  (void)hipStreamSynchronize(c->stream);
  batch_drop_graph(B);
  for (void *p : B->allocs)
    if (p) (void)hipFree(p);
  delete B;
  c->batch = nullptr;
  return GPR_OK;
}

int gpr_batch_create(gpr_ctx *c, int count, const int32_t *nx, const double *dx, const int32_t *bcs,
                     const gpr_par *MPs, const double *tf, double CFL) {
  if (!c || count < 1 || !nx || !dx || !bcs || !MPs || !tf) return fail(GPR_E_ARG, "gpr_batch_create: bad argument");
  if (c->ndim != 1) return fail(GPR_E_UNSUPPORTED, "batches are of 1D problems: create the context with ndim = 1");
  if (c->batch) gpr_batch_destroy(c);
  const int N = c->N, V = c->V, NN = N * N;
  const int CPW = WAVE / NN, FPW = WAVE / N;
  BatchHost *B = new BatchHost();
  B->count = count;
  B->inst.resize(count);
  B->u0s.assign(count + 1, 0);
  B->b0s.assign(count + 1, 0);
  B->e0s.assign(count + 1, 0);
  std::vector<long> f0s(count + 1, 0);
  std::vector<int> pblk, cblk, fblk;
  for (int k = 0; k < count; k++) {
    if (nx[k] < N || !(dx[k] > 0) || !(tf[k] >= 0)) {
      delete B;
      return fail(GPR_E_ARG, "instance " + std::to_string(k) + ": need nx >= N, dx > 0, tf >= 0");
    }
    for (int s = 0; s < 2; s++)
      if (bcs[2 * k + s] < 0 || bcs[2 * k + s] > 5) {
        delete B;
        return fail(GPR_E_ARG, "instance " + std::to_string(k) + ": boundary code must be 0..5");
      }
    BatchInst &I = B->inst[k];
    std::memset(&I, 0, sizeof I);
    int rc = make_mat(&MPs[k], I.m);
    if (rc) {
      delete B;
      return fail(rc, "instance " + std::to_string(k) + ": " + g_err);
    }
    I.nx = nx[k];
    I.dx = dx[k];
    I.bc0 = bcs[2 * k];
    I.bc1 = bcs[2 * k + 1];
    I.u0 = B->u0s[k];
    I.b0 = B->b0s[k];
    I.e0 = B->e0s[k];
    I.f0 = f0s[k];
    B->u0s[k + 1] = I.u0 + nx[k];
    B->b0s[k + 1] = I.b0 + nx[k] + 2 * N;
    B->e0s[k + 1] = I.e0 + nx[k] + 2;
    f0s[k + 1] = I.f0 + nx[k] + 1;
    I.blk_p = (int)pblk.size();
    for (int b = 0; b < (nx[k] + 2 + CPW - 1) / CPW; b++) pblk.push_back(k);
    I.blk_c = (int)cblk.size();
    for (int b = 0; b < (nx[k] + CPW - 1) / CPW; b++) cblk.push_back(k);
    I.blk_f = (int)fblk.size();
    for (int b = 0; b < (nx[k] + 1 + FPW - 1) / FPW; b++) fblk.push_back(k);
  }
  B->nu = B->u0s[count];
  B->nb = B->b0s[count];
  B->ne = B->e0s[count];
  B->nf = f0s[count];
  B->mk = material_kind(B->inst[0].m);
  for (int k = 1; k < count; k++)
    if (material_kind(B->inst[k].m) != B->mk) B->mk = MK_ANY;
  B->nblk_p = (int)pblk.size();
  B->nblk_c = (int)cblk.size();
  B->nblk_f = (int)fblk.size();
  B->view_count = count;
  // allocations: 0 inst, 1 st, 2 offsets (3 x (count + 1)), 3 pblk, 4 cblk, 5 fblk, 6 u, 7 (unused: WENO
  // reads the ghost cells from u), 8 wh, 9 qh
  // (or the fused form's face states: 2N of its N^2 nodes per cell), 10 fx, 11 view instances, 12 view
  // offsets (3 x (count + 1)), 13 the fused form's cell integrals
  const bool fused = N == 3 && !c->thermal && B->mk == MK_GAS_NEWTONIAN;
  const size_t sizes[14] = {sizeof(BatchInst) * count,
                            sizeof(BatchState) * count,
                            sizeof(long) * 3 * (count + 1),
                            sizeof(int) * pblk.size(),
                            sizeof(int) * cblk.size(),
                            sizeof(int) * fblk.size(),
                            sizeof(double) * B->nu * V,
                            0,
                            sizeof(double) * B->ne * N * V,
                            sizeof(double) * B->ne * NN * V,
                            sizeof(double) * B->nf * 2 * V,
                            sizeof(int) * count,
                            sizeof(long) * 3 * (count + 1),
                            fused ? sizeof(double) * B->nu * V : 0};
  B->allocs.assign(14, nullptr);
  c->batch = B;
  for (int i = 0; i < 14; i++) HIP_TRY(hipMalloc(&B->allocs[i], std::max<size_t>(sizes[i], 256)));
  std::vector<long> offs;
  offs.insert(offs.end(), B->u0s.begin(), B->u0s.end());
  offs.insert(offs.end(), B->b0s.begin(), B->b0s.end());
  offs.insert(offs.end(), B->e0s.begin(), B->e0s.end());
  HIP_TRY(hipMemcpyAsync(B->allocs[0], B->inst.data(), sizes[0], hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[2], offs.data(), sizes[2], hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[3], pblk.data(), sizes[3], hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[4], cblk.data(), sizes[4], hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(B->allocs[5], fblk.data(), sizes[5], hipMemcpyHostToDevice, c->stream));
  std::vector<BatchState> st(count);
  for (int k = 0; k < count; k++) {
    std::memset(&st[k], 0, sizeof(BatchState));
    st[k].tf = tf[k];
    st[k].cfl = CFL;
    st[k].contorted_tol = c->opts.contorted_tol;
    st[k].done = 1;  // until gpr_batch_set_state
  }
  HIP_TRY(hipMemcpyAsync(B->allocs[1], st.data(), sizes[1], hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  BatchArgs &a = B->args;
  a.inst = (const BatchInst *)B->allocs[0];
  a.st = (BatchState *)B->allocs[1];
  a.count = count;
  a.u0s = (const long *)B->allocs[2];
  a.b0s = a.u0s + (count + 1);
  a.e0s = a.b0s + (count + 1);
  a.cblk = (const int *)B->allocs[4];
  a.fblk = (const int *)B->allocs[5];
  a.u = (double *)B->allocs[6];
  a.wh = (double *)B->allocs[8];
  a.qh = (double *)B->allocs[9];
  a.fx = (double *)B->allocs[10];
  a.fs = a.qh;
  a.cint = (const double *)B->allocs[13];
  return batch_build_view(c, B, std::vector<char>(count, 1));
}

int gpr_batch_set_state(gpr_ctx *c, const double *u) {
  if (!c || !c->batch || !u) return fail(GPR_E_ARG, "gpr_batch_set_state: no batch or null state");
  BatchHost *B = c->batch;
  HIP_TRY(hipMemcpyAsync(B->args.u, u, sizeof(double) * B->nu * c->V, hipMemcpyHostToDevice, c->stream));
  std::vector<BatchState> st(B->count);
  HIP_TRY(hipMemcpyAsync(st.data(), B->args.st, sizeof(BatchState) * B->count, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto &s : st) {
    s.t = 0.0;
    s.count = 0;
    s.dt = 0.0;
    s.maxspeed = s.maxspeed2 = 0ull;
    s.done = s.contorted = s.nonfinite = 0;
    s.contorted_tol = c->opts.contorted_tol;
  }
  HIP_TRY(hipMemcpyAsync(B->args.st, st.data(), sizeof(BatchState) * B->count, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_fail, 0, sizeof(int), c->stream));
  if (B->view_count != B->count) {
    int rv = batch_build_view(c, B, std::vector<char>(B->count, 1));
    if (rv) return rv;
  }
  int rc = batch_first_dt(c, B);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  B->ready = true;
  return GPR_OK;
}

int gpr_batch_get_state(gpr_ctx *c, double *u) {
  if (!c || !c->batch || !u) return fail(GPR_E_ARG, "gpr_batch_get_state: no batch or null buffer");
  BatchHost *B = c->batch;
  HIP_TRY(hipMemcpyAsync(u, B->args.u, sizeof(double) * B->nu * c->V, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GPR_OK;
}

int gpr_batch_set_block(gpr_ctx *c, int block) {
  if (!c || !c->batch) return fail(GPR_E_ARG, "no batch");
  if (block < 64 || block > 1024 || block % 64) return fail(GPR_E_ARG, "block must be a multiple of 64 in [64, 1024]");
  if (block != c->batch->block) batch_drop_graph(c->batch);
  c->batch->block = block;
  return GPR_OK;
}

int gpr_batch_steps(gpr_ctx *c, int64_t nsteps) {
  if (!c || !c->batch || !c->batch->ready) return fail(GPR_E_ARG, "gpr_batch_set_state has not been called");
  if (c->opts.flux != 0) return fail(GPR_E_UNSUPPORTED, "batches take the Rusanov flux (FLUX = 0)");
  for (int64_t k = 0; k < nsteps; k++) {
    int rc = batch_step(c, c->batch);
    if (rc) return rc;
  }
  return GPR_OK;
}

int gpr_batch_status(gpr_ctx *c, double *t, int64_t *count, double *dt, int32_t *done, int32_t *nonfinite,
                     int32_t *pred_fail) {
  if (!c || !c->batch) return fail(GPR_E_ARG, "no batch");
  BatchHost *B = c->batch;
  std::vector<BatchState> st(B->count);
  int pf = 0;
  HIP_TRY(hipMemcpyAsync(st.data(), B->args.st, sizeof(BatchState) * B->count, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&pf, c->d_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < B->count; k++) {
    if (t) t[k] = st[k].t;
    if (count) count[k] = st[k].count;
    if (dt) dt[k] = st[k].dt;
    if (done) done[k] = st[k].done;
    if (nonfinite) nonfinite[k] = st[k].nonfinite;
  }
  if (pred_fail) *pred_fail = pf;
  return GPR_OK;
}

int gpr_batch_run(gpr_ctx *c, int steps_per_graph, int64_t max_graphs, int64_t *graphs_run, int32_t *running) {
  if (!c || !c->batch || !c->batch->ready) return fail(GPR_E_ARG, "gpr_batch_set_state has not been called");
  if (steps_per_graph < 1 || max_graphs < 1) return fail(GPR_E_ARG, "steps_per_graph and max_graphs must be >= 1");
  if (c->opts.flux != 0) return fail(GPR_E_UNSUPPORTED, "batches take the Rusanov flux (FLUX = 0)");
  BatchHost *B = c->batch;
  if (B->exec && B->graph_steps != steps_per_graph) batch_drop_graph(B);
  std::vector<BatchState> st(B->count);
  int64_t n = 0;
  int run = B->count;
  while (n < max_graphs && run > 0) {
    if (!B->exec) {
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = GPR_OK;
      for (int k = 0; k < steps_per_graph && rc == GPR_OK; k++) rc = batch_step(c, B);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(GPR_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(GPR_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      B->graph_steps = steps_per_graph;
    }
    HIP_TRY(hipGraphLaunch(B->exec, c->stream));
    n++;
    HIP_TRY(hipMemcpyAsync(st.data(), B->args.st, sizeof(BatchState) * B->count, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    run = 0;
    for (auto &s : st) run += s.done ? 0 : 1;
    // Once half of the covered instances are done, the block-mapped launches
    // are rebuilt over the running ones (and the graph captured again).
    if (run > 0 && run <= B->view_count / 2) {
      std::vector<char> on(B->count);
      for (int k = 0; k < B->count; k++) on[k] = st[k].done ? 0 : 1;
      int rv = batch_build_view(c, B, on);
      if (rv) return rv;
    }
  }
  if (graphs_run) *graphs_run = n;
  if (running) *running = run;
  return GPR_OK;
}

}  // extern "C"

// ==================================================================== adjoint
// Forward run with checkpoints and its dt schedule, then the reverse sweep:
// each segment of K steps is recomputed from its checkpoint, then its steps
// are differentiated in reverse (gpr_adjoint.hpp).
struct AdjHost {
  int64_t nsteps = 0;
  int K = 0;
  std::vector<double> dts;
  double *chk = nullptr;  // [nchk][cells][V]
  double *seg = nullptr;  // [K][cells][V]
  double *qhs = nullptr;  // [K][predictor cells][NN][V], the segment's predictor results (or none)
  size_t qh_bytes = 0;
  std::vector<std::pair<void *, size_t>> bufs;
  int start_pre = 1;      // transposed predictor solve: 1 = preconditioned form from the start
  int dual_local = 0;     // 1: local derivatives by dual numbers instead of node_vjp
  double tol = 1e-13;     // its stopping tolerance
  int64_t stats[3] = {0, 0, 0};  // the last reverse sweep's iterations, cells, preconditioned cells
  double grad_kappa = 0.0;        // the last reverse sweep's d loss / d kappa
};

namespace {

template <int V>
__global__ void k_renorm_only(double *u, long ncell, double rho0) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    double q[14];
#pragma unroll
    for (int v = 0; v < 14; v++) q[v] = u[c * V + v];
    if (q[0] > 0) {  // as k_post
      double cf = cbrt(q[0] / (rho0 * det3(q + 5)));
#pragma unroll
      for (int k = 0; k < 9; k++) q[5 + k] *= cf;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) u[c * V + 5 + k] = q[5 + k];
  }
}

void *abuf(AdjHost *A, int i, size_t bytes, int &rc) {
  if ((int)A->bufs.size() <= i) A->bufs.resize(i + 1, {nullptr, 0});
  auto &e = A->bufs[i];
  if (e.first && e.second >= bytes) return e.first;
  if (e.first) (void)hipFree(e.first);
  e = {nullptr, 0};
  if (hipMalloc(&e.first, std::max<size_t>(bytes, 256)) != hipSuccess) {
    rc = fail(GPR_E_HIP, "adjoint buffer allocation");
    return nullptr;
  }
  e.second = bytes;
  rc = GPR_OK;
  return e.first;
}

// the stages of one step from u_n with a given dt: ub, ux, wh, qh and u_fv
// (u_n plus the FV update, before renormalisation).  With qh_in, the
// predictor is not run: qh_in holds its result for this step and the FV
// update reads it.
int adj_stages(gpr_ctx *c, const double *u_n, double *u_fv, double dt, const double *qh_in = nullptr) {
  int rc;
  int nX[2] = {c->run_nx, c->run_ny};
  double dX[2] = {c->run_dx, c->run_dy};
  long ne = ext_cells(c, nX, 1), nc = ext_cells(c, nX, 0);
  double *dub = (double *)c->bufs["ub"].first;
  double *dwh = (double *)c->bufs["wh"].first;
  double *dqh = (double *)c->bufs["qh"].first;
  uint8_t *dm = (uint8_t *)c->bufs["run_mask"].first;
  if (!dub || !dwh || !dm || (!qh_in && !dqh))
    return fail(GPR_E_ARG, "the device run has no ADER stage buffers (gpr_dev_begin in the ADER FP64 form is needed)");
  rc = dispatch<BoundariesOp>(c, c, u_n, dub, c->run_nx, c->run_ny, c->run_bc);
  if (rc) return rc;
  // the reverse sweep reads the x-pass output (ux)
  rc = dispatch<WenoOp>(c, c, (const double *)dub, dwh, c->run_nx, c->run_ny, nullptr, true);
  if (rc) return rc;
  if (!qh_in) {
    PredArgs pa = pred_args(c, c->run_mat, dt, dX);
    pa.qh = dqh;
    pa.wh = dwh;
    pa.mask = dm;
    pa.ncell = ne;
    rc = dispatch<PredictorOp>(c, c, pa);
    if (rc) return rc;
  }
  HIP_TRY(hipMemcpyAsync(u_fv, u_n, sizeof(double) * nc * c->V, hipMemcpyDeviceToDevice, c->stream));
  FvArgs fa = fv_args(c, c->run_mat, dt, dX, nX);
  fa.u = u_fv;
  fa.qh = qh_in ? qh_in : dqh;
  fa.mask = dm;
  fa.post = 0;
  rc = fv_buffers(c, fa);
  if (rc) return rc;
  return dispatch<FvOp>(c, c, fa);
}

template <int N, bool TH>
int adj_step_back(gpr_ctx *c, AdjHost *A, const double *u_n, double dt, double *lam, double *d_gmu,
                  double *d_gkap, const double *qh_saved) {
  constexpr int V = Dims<TH>::V;
  int rc;
  const int nx = c->run_nx, ny = c->run_ny;
  long nc = (long)nx * ny, ne = (long)(nx + 2) * (ny + 2);
  long nb = (long)(nx + 2 * N) * (ny + 2 * N), nux = (long)(nx + 2) * (ny + 2 * N);
  long nfx = (long)(nx + 1) * ny, nfy = (long)nx * (ny + 1);
  double *u_fv = (double *)abuf(A, 0, sizeof(double) * nc * V, rc);
  if (!u_fv) return rc;
  rc = adj_stages(c, u_n, u_fv, dt, qh_saved);
  if (rc) return rc;
  AdjArgs a;
  std::memset(&a, 0, sizeof a);
  a.u_fv = u_fv;
  a.lam1 = lam;
  if (!(a.lam_fv = (double *)abuf(A, 1, sizeof(double) * nc * V, rc))) return rc;
  if (!(a.lam_face_x = (double *)abuf(A, 2, sizeof(double) * nfx * N * N * 2 * V, rc))) return rc;
  if (!(a.lam_face_y = (double *)abuf(A, 3, sizeof(double) * nfy * N * N * 2 * V, rc))) return rc;
  if (!(a.lam_qh = (double *)abuf(A, 4, sizeof(double) * ne * N * N * N * V, rc))) return rc;
  if (!(a.lam_wh = (double *)abuf(A, 5, sizeof(double) * ne * N * N * V, rc))) return rc;
  if (!(a.lam_ux = (double *)abuf(A, 6, sizeof(double) * nux * N * V, rc))) return rc;
  if (!(a.lam_ub = (double *)abuf(A, 7, sizeof(double) * nb * V, rc))) return rc;
  if (!(a.lam0 = (double *)abuf(A, 8, sizeof(double) * nc * V, rc))) return rc;
  a.qh = qh_saved ? qh_saved : (const double *)c->bufs["qh"].first;
  a.ux = (const double *)c->bufs["ux"].first;
  a.ub = (const double *)c->bufs["ub"].first;
  a.mask = (const uint8_t *)c->bufs["run_mask"].first;
  a.gmu = d_gmu;
  a.gkap = d_gkap;
  a.kappa = c->run_mat.p.kappa;
  a.fail = c->d_fail;
  a.dt = dt;
  a.dx = c->run_dx;
  a.dy = c->run_dy;
  {
    const gpr_par &P = c->run_mat.p;
    a.par_fac = P.SOLID ? (P.POWER_LAW ? -P.n / P.sigmaY : 0.0) : -1.0 / P.mu;
  }
  a.nx = nx;
  a.ny = ny;
  a.wh = (const double *)c->bufs["wh"].first;
  a.it_plain = c->opts.dg_it;
  a.start_pre = A->start_pre;
  a.it_max = 400;
  a.tol = A->tol;
  if (!(a.itc = (unsigned long long *)abuf(A, 12, ITC_SLOTS * ITC_STRIDE * sizeof(unsigned long long), rc)))
    return rc;
  a.bc = c->run_bc;
  a.m = c->run_mat;
  HIP_TRY(hipMemsetAsync(a.lam_wh, 0, sizeof(double) * ne * N * N * V, c->stream));
  HIP_TRY(hipMemsetAsync(a.lam_ux, 0, sizeof(double) * nux * N * V, c->stream));
  HIP_TRY(hipMemsetAsync(a.lam_ub, 0, sizeof(double) * nb * V, c->stream));
  hipLaunchKernelGGL((ka_renorm<TH>), dim3(grid_for(nc, 256)), dim3(256), 0, c->stream, a, nc);
  constexpr int FPW = FaceShape<N, TH, 2>::FPW;
  constexpr int CPW = PredShape<N, 2>::CPW;
  const dim3 gf((nfx + nfy + FPW - 1) / FPW), gc((ne + CPW - 1) / CPW);
  static_assert(N == 3, "the adjoint's transposed faces, cells and predictor are built for N = 3");
  if (A->dual_local)
    HIP_TRY((adj_heavy_launch<true, TH>(gf, gc, c->stream, a, ne)));
  else
    HIP_TRY((adj_heavy_launch<false, TH>(gf, gc, c->stream, a, ne)));
  hipLaunchKernelGGL((ka_weno_y<N, V>), dim3(grid_for(ne * N * V, 256)), dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL((ka_weno_x<N, V>), dim3(grid_for(nux * V, 256)), dim3(256), 0, c->stream, a);
  HIP_TRY(hipMemcpyAsync(a.lam0, a.lam_fv, sizeof(double) * nc * V, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL((ka_boundaries<N, V>), dim3(grid_for(nb * V, 256)), dim3(256), 0, c->stream, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(lam, a.lam0, sizeof(double) * nc * V, hipMemcpyDeviceToDevice, c->stream));
  return GPR_OK;
}

}  // namespace

extern "C" {

int gpr_adj_release(gpr_ctx *c) {
  if (!c) return fail(GPR_E_ARG, "null context");
  AdjHost *A = c->adj;
  if (!A) return GPR_OK;
  (void)hipStreamSynchronize(c->stream);
  if (A->chk) (void)hipFree(A->chk);
  if (A->seg) (void)hipFree(A->seg);
  if (A->qhs) (void)hipFree(A->qhs);
  for (auto &b : A->bufs)
    if (b.first) (void)hipFree(b.first);
  delete A;
  c->adj = nullptr;
  return GPR_OK;
}

int gpr_adj_forward(gpr_ctx *c, int64_t nsteps, int K, double *dts_out) {
  if (!c || !c->run_ready) return fail(GPR_E_ARG, "gpr_dev_begin has not been called");
  if (c->ndim != 2 || c->N != 3) return fail(GPR_E_UNSUPPORTED, "the adjoint is built for 2D, N = 3");
  if (nsteps < 1 || K < 1) return fail(GPR_E_ARG, "nsteps and K must be >= 1");
  // the reverse sweep transposes the plain FP64 ADER step; a forward sweep
  // in the split or FP32-bulk form would be a different map
  if (c->opts.split != 0) return fail(GPR_E_UNSUPPORTED, "the adjoint covers the ADER step, not the split step");
  if (c->opts.precision != 0) return fail(GPR_E_UNSUPPORTED, "the adjoint covers the FP64 step, not the FP32-bulk form");
  if (c->opts.flux != 0) return fail(GPR_E_UNSUPPORTED, "the adjoint transposes the Rusanov flux only");
  if (c->run_mat.p.BINGHAM || (c->run_mat.p.POWER_LAW && !c->run_mat.p.SOLID))
    return fail(GPR_E_UNSUPPORTED,
                "the adjoint covers Newtonian fluids and elastic or power-law solids, not Bingham or power-law fluids");
  for (int k = 0; k < 4; k++) {
    int code = k == 0 ? c->run_bc.x : k == 1 ? c->run_bc.y : k == 2 ? c->run_bc.z : c->run_bc.w;
    if (code == 4) return fail(GPR_E_UNSUPPORTED, "the adjoint has no moving-wall boundary");
  }
  if (c->adj) gpr_adj_release(c);
  AdjHost *A = new AdjHost();
  c->adj = A;
  A->nsteps = nsteps;
  A->K = K;
  A->dts.resize(nsteps);
  const long nc = (long)c->run_nx * c->run_ny;
  const size_t sz = sizeof(double) * nc * c->V;
  const int64_t nchk = (nsteps + K - 1) / K;
  HIP_TRY(hipMalloc((void **)&A->chk, sz * nchk));
  HIP_TRY(hipMalloc((void **)&A->seg, sz * K));
  StepState h0;
  HIP_TRY(hipMemcpyAsync(&h0, c->d_st, sizeof h0, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the steps' dt are recorded on the device as they are taken (no host
  // synchronisation per step)
  int rc = GPR_OK;
  double *d_dts = (double *)abuf(A, 13, sizeof(double) * nsteps, rc);
  if (!d_dts) return rc;
  for (int64_t k = 0; k < nsteps; k++) {
    if (k % K == 0)
      HIP_TRY(hipMemcpyAsync(A->chk + (k / K) * nc * c->V, c->run_u, sz, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(1), 0, c->stream, c->d_st, d_dts + k);
    rc = dev_steps_impl(c, 1, nullptr);
    if (rc) return rc;
  }
  HIP_TRY(hipMemcpyAsync(A->dts.data(), d_dts, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
  StepState h1;
  HIP_TRY(hipMemcpyAsync(&h1, c->d_st, sizeof h1, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (dts_out) std::memcpy(dts_out, A->dts.data(), sizeof(double) * nsteps);
  // the reverse sweep differentiates renormalisation, not the global reset,
  // and a failed or non-finite step has no derivative
  if (h1.resets != h0.resets)
    return fail(GPR_E_UNSUPPORTED, "a global distortion reset happened in the forward sweep; the adjoint does not differentiate it");
  if (h1.pred_fail != h0.pred_fail) return fail(GPR_E_PREDICTOR, "a predictor cell failed in the forward sweep");
  if (h1.nonfinite) return fail(GPR_E_NONFINITE, "the forward sweep became non-finite");
  return GPR_OK;
}

int gpr_adj_backward(gpr_ctx *c, const double *lam_T, double *lam_0, double *grad_mu, int32_t *unconverged) {
  if (!c || !c->adj) return fail(GPR_E_ARG, "gpr_adj_forward has not been called");
  AdjHost *A = c->adj;
  int rc;
  const long nc = (long)c->run_nx * c->run_ny;
  const size_t sz = sizeof(double) * nc * c->V;
  double *lam = (double *)abuf(A, 9, sz, rc);
  if (!lam) return rc;
  double *d_gmu = (double *)abuf(A, 10, sizeof(double), rc);
  if (!d_gmu) return rc;
  double *d_gkap = (double *)abuf(A, 14, sizeof(double), rc);
  if (!d_gkap) return rc;
  HIP_TRY(hipMemsetAsync(d_gkap, 0, sizeof(double), c->stream));
  double *work = (double *)abuf(A, 11, sz, rc);
  if (!work) return rc;
  HIP_TRY(hipMemcpyAsync(lam, lam_T, sz, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(d_gmu, 0, sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(c->d_fail, 0, sizeof(int), c->stream));
  unsigned long long *d_itc =
      (unsigned long long *)abuf(A, 12, ITC_SLOTS * ITC_STRIDE * sizeof(unsigned long long), rc);
  if (!d_itc) return rc;
  HIP_TRY(hipMemsetAsync(d_itc, 0, ITC_SLOTS * ITC_STRIDE * sizeof(unsigned long long), c->stream));
  const int K = A->K;
  const int64_t nchk = (A->nsteps + K - 1) / K;
  // Keep each recomputed step's predictor result for its transposed step
  // (K of them; 5 GB at 256^2 with K = 25), so the predictor runs once per
  // step in the reverse sweep.  Without the memory, it runs twice.
  const size_t qhb = sizeof(double) * (size_t)(c->run_nx + 2) * (c->run_ny + 2) * 27 * c->V;  // N = 3, 2D
  if (!A->qhs || A->qh_bytes != qhb) {
    if (A->qhs) (void)hipFree(A->qhs);
    A->qhs = nullptr;
    A->qh_bytes = qhb;
    if (hipMalloc((void **)&A->qhs, qhb * K) != hipSuccess) {
      (void)hipGetLastError();
      A->qhs = nullptr;
    }
  }
  for (int64_t sgi = nchk - 1; sgi >= 0; sgi--) {
    const int64_t n0 = sgi * K, n1 = std::min<int64_t>(n0 + K, A->nsteps);
    // recompute u_n0 .. u_{n1-1} from the checkpoint
    HIP_TRY(hipMemcpyAsync(work, A->chk + sgi * nc * c->V, sz, hipMemcpyDeviceToDevice, c->stream));
    for (int64_t n = n0; n < n1; n++) {
      HIP_TRY(hipMemcpyAsync(A->seg + (n - n0) * nc * c->V, work, sz, hipMemcpyDeviceToDevice, c->stream));
      if (n + 1 < n1 || A->qhs) {
        double *u_fv = (double *)abuf(A, 0, sz, rc);
        if (!u_fv) return rc;
        rc = adj_stages(c, work, u_fv, A->dts[n]);
        if (rc) return rc;
        if (A->qhs)
          HIP_TRY(hipMemcpyAsync((char *)A->qhs + (n - n0) * qhb, c->bufs["qh"].first, qhb, hipMemcpyDeviceToDevice,
                                 c->stream));
        if (n + 1 < n1) {
          if (c->thermal)
            hipLaunchKernelGGL(k_renorm_only<17>, dim3(grid_for(nc, 256)), dim3(256), 0, c->stream, u_fv, nc,
                               c->run_mat.p.rho0);
          else
            hipLaunchKernelGGL(k_renorm_only<14>, dim3(grid_for(nc, 256)), dim3(256), 0, c->stream, u_fv, nc,
                             c->run_mat.p.rho0);
          HIP_TRY(hipMemcpyAsync(work, u_fv, sz, hipMemcpyDeviceToDevice, c->stream));
        }
      }
    }
    for (int64_t n = n1 - 1; n >= n0; n--) {
      const double *qsave = A->qhs ? (const double *)((char *)A->qhs + (n - n0) * qhb) : nullptr;
      rc = c->thermal ? adj_step_back<3, true>(c, A, A->seg + (n - n0) * nc * c->V, A->dts[n], lam, d_gmu, d_gkap, qsave)
                      : adj_step_back<3, false>(c, A, A->seg + (n - n0) * nc * c->V, A->dts[n], lam, d_gmu, d_gkap,
                                                qsave);
      if (rc) return rc;
    }
  }
  int nf = 0;
  HIP_TRY(hipMemcpyAsync(lam_0, lam, sz, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(grad_mu, d_gmu, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&nf, c->d_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  std::vector<unsigned long long> itc_slots((size_t)ITC_SLOTS * ITC_STRIDE);
  HIP_TRY(hipMemcpyAsync(itc_slots.data(), d_itc, itc_slots.size() * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&A->grad_kappa, d_gkap, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; k++) {
    unsigned long long tot = 0;
    for (int s = 0; s < ITC_SLOTS; s++) tot += itc_slots[(size_t)s * ITC_STRIDE + k];
    A->stats[k] = (int64_t)tot;
  }
  if (unconverged) *unconverged = nf;
  return GPR_OK;
}

int gpr_adj_set_solver(gpr_ctx *c, int start_preconditioned, double tol, int dual_local) {
  if (!c || !c->adj) return fail(GPR_E_ARG, "gpr_adj_forward has not been called");
  if (start_preconditioned < 0 || start_preconditioned > 1 || !(tol > 0) || !(tol < 1e-3) || dual_local < 0 ||
      dual_local > 1)
    return fail(GPR_E_ARG, "start_preconditioned and dual_local must be 0 or 1, and 0 < tol < 1e-3");
  c->adj->start_pre = start_preconditioned;
  c->adj->tol = tol;
  c->adj->dual_local = dual_local;
  return GPR_OK;
}

int gpr_adj_grad_kappa(gpr_ctx *c, double *out) {
  if (!c || !c->adj || !out) return fail(GPR_E_ARG, "gpr_adj_forward has not been called");
  *out = c->adj->grad_kappa;
  return GPR_OK;
}

int gpr_adj_stats(gpr_ctx *c, int64_t *out) {
  if (!c || !c->adj || !out) return fail(GPR_E_ARG, "gpr_adj_forward has not been called");
  for (int k = 0; k < 3; k++) out[k] = c->adj->stats[k];
  return GPR_OK;
}

}  // extern "C"
