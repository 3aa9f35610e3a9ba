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

#include "gpr_pointwise.hpp"

// waves per SIMD the predictor is compiled for (register budget 512 / waves)
#ifndef GPR_PRED_WAVES
#define GPR_PRED_WAVES 2
#endif

namespace gpr {

constexpr int WAVE = 64;

// Iteration counters (PredArgs::itc and the adjoint's) are summed into
// ITC_SLOTS slots of ITC_STRIDE words (one 128-byte line each), slot =
// block % ITC_SLOTS, so that atomics from different workgroups rarely meet
// on one address; the host adds the slots.
constexpr int ITC_SLOTS = 256, ITC_STRIDE = 16;

struct StepState;

// A batch of independent 1D problems in one set of launches.  Instance k's
// extended cells are [e0, e0 + nx + 2) of the concatenated wh / qh, its cells
// [u0, u0 + nx) of u, its ghost-extended rows [b0, b0 + nx + 2N) of ub and its
// faces [f0, f0 + nx + 1).  Kernels that give one wavefront a group of cells
// or faces never mix instances: the instance's groups start at block blk_p
// (predictor), blk_c (cell update) and blk_f (faces).
struct BatchInst {
  long e0, u0, b0, f0;
  int nx, blk_p, blk_c, blk_f;
  int bc0, bc1;  // boundary codes, left and right
  double dx;
  Mat m;
};
// per-instance time-loop state
struct BatchState {
  double t, tf, cfl, dt, contorted_tol;
  long long count;
  unsigned long long maxspeed;   // bit pattern of the CFL maximum (non-negative double)
  unsigned long long maxspeed2;  // the same after a distortion reset
  int done, contorted, nonfinite, pad_;
};

struct PredArgs {
  double *qh;              // [ncell][NN][V]
  const void *wh;          // [ncell][NS][V], double or float (the kernel's WIn)
  int *list;               // split launches: cells Picard leaves, appended ...
  int *list_count;         // ... at list[atomicAdd(list_count, 1)]
  const uint8_t *mask;     // [ncell]
  int32_t *iters;          // optional per-cell iteration count (fallback: 1000 + its)
  int *fail;               // device counter of cells that converged in neither phase
  unsigned long long *itc; // optional, ITC_SLOTS x ITC_STRIDE words: [0] Picard, [1] fallback iterations,
                           // [2] cells, [3] fallback cells
  const double *dt_dev;    // if non-null, dt is read from here
  double dt, dx, dy;
  long ncell;
  int dg_it, stiff_it, bailout, mode;  // mode 1: fallback only; 0 or 2: Picard then fallback
  int bail_from;           // > 0: after this many Picard iterations, one growth leaves Picard
  // fused FV corrector (the one-launch FP64 kernels of the P2 ideal-gas
  // Newtonian kind, 2D and batched, the only ones compiled with it): if fs
  // is set, the launch writes each active cell's face states (FaceStates
  // layout) and each real cell's FV cell integral (sources, TIME) to cint
  // ([fnx][fny][V]; batched: the instances' cells concatenated); qh may then
  // be null
  double *fs;
  double *cint;
  int fnx, fny;
  double dg_tol, stiff_tol;
  Mat m;
  // batched launches only
  const BatchInst *binst;
  const int *pblk;         // predictor block -> instance
  const BatchState *bst;
};

// One wavefront of 64 lanes holds CPW cells, one space-time node per lane.
// A cell of more than 16 nodes (N = 3 in 2D: 27) takes a 32-lane half of
// the wavefront, so that reductions over its lanes run within the half
// (cross-lane DPP moves); smaller cells are packed densely.
template <int N, int ND> struct PredShape {
  static constexpr int NS = ND == 1 ? N : N * N;  // spatial nodes per cell
  static constexpr int NN = NS * N;                // space-time nodes per cell
  static constexpr int LPC = NN > 16 ? 32 : NN;    // lanes per cell slot
  static constexpr int CPW = WAVE / LPC;           // cells per wavefront
};


// max over each 32-lane half of the wavefront of a non-negative double (NaN
// counts as 0, as fmax ignores it), for every lane of the half.  Doubles >= 0
// order as their bit patterns, so the maximum is taken on (hi, lo) words
// moved by DPP: prefix maxima within each 16-lane row (row_shr 1, 2, 4, 8),
// then row 0's (row 2's) maximum into row 1 (row 3) by row_bcast:15, which
// leaves the half's maximum in lanes 31 and 63.  Every lane must be active.
__device__ __forceinline__ double half_wave_max(double x) {
  unsigned long long b = x == x ? (unsigned long long)__double_as_longlong(x) : 0ull;
  unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
#define GPR_DPP_MAX_STEP(ctrl, rmask)                                                \
  {                                                                                 \
    const unsigned l2 = __builtin_amdgcn_update_dpp(0u, lo, ctrl, rmask, 0xf, false); \
    const unsigned h2 = __builtin_amdgcn_update_dpp(0u, hi, ctrl, rmask, 0xf, false); \
    const bool gt = h2 > hi || (h2 == hi && l2 > lo);                                \
    lo = gt ? l2 : lo;                                                              \
    hi = gt ? h2 : hi;                                                              \
  }
  GPR_DPP_MAX_STEP(0x111, 0xf)  // row_shr:1
  GPR_DPP_MAX_STEP(0x112, 0xf)  // row_shr:2
  GPR_DPP_MAX_STEP(0x114, 0xf)  // row_shr:4
  GPR_DPP_MAX_STEP(0x118, 0xf)  // row_shr:8
  GPR_DPP_MAX_STEP(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
#undef GPR_DPP_MAX_STEP
  const bool upper = (threadIdx.x & 32) != 0;
  const unsigned l = upper ? __builtin_amdgcn_readlane(lo, 63) : __builtin_amdgcn_readlane(lo, 31);
  const unsigned h = upper ? __builtin_amdgcn_readlane(hi, 63) : __builtin_amdgcn_readlane(hi, 31);
  return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
}

// The predictor's workgroup is exactly one wavefront, so exchanging values
// between its lanes through LDS needs ordering within the wave, not a
// workgroup barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scheduling fence every GPR_PRED_CHUNK variables of an unrolled contraction:
// without it the scheduler issues all of a contraction's LDS reads before its
// arithmetic, which needs more registers than two wavefronts per SIMD have.
#ifndef GPR_PRED_CHUNK
#define GPR_PRED_CHUNK 4
#endif
#define GPR_CHUNK_FENCE(v) \
  if ((v) % GPR_PRED_CHUNK == GPR_PRED_CHUNK - 1) __builtin_amdgcn_sched_barrier(0)

// Materialise x here: an empty asm that reads and writes x keeps the
// arithmetic producing x ahead of this point.  Without it, LLVM sinks a
// contraction's arithmetic below the next conditional LDS-write block, to
// where its result is next used, while the LDS reads stay behind the wave
// fence above: every value read stays live across the blocks in between.
template <typename T> __device__ __forceinline__ void pin(T &x) { asm volatile("" : "+v"(x)); }

// a / b to within a few ulps: hardware reciprocal refined by one Newton step
__device__ __forceinline__ double fast_div(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return a * r;
}
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// What one launch does (KIND):
//   KIND_BOTH      Picard, and the preconditioned fallback in the same
//                  wavefront for the cells Picard leaves (R = double);
//   KIND_PICARD    Picard only; the cells it leaves are appended to a.list;
//   KIND_FALLBACK  the fallback only, from the standard initial guess, for
//                  the cells of a.list, CPW of them per wavefront (R = double).
// A cell's result does not depend on which launch or wavefront solves it:
// KIND_PICARD + KIND_FALLBACK gives KIND_BOTH's values.
enum { KIND_BOTH = 0, KIND_PICARD = 1, KIND_FALLBACK = 2 };

// R: the type the iteration computes in (double, or float for the FP32-bulk
// form, KIND_PICARD only).  WIn: the element type of wh.
// BATCH: the launch covers a batch of 1D instances (a.binst, a.pblk, a.bst)
// with per-instance material, dx and dt; every extended cell is predicted.
// MK: the material kind the physics is compiled for (MK_ANY: any material).
template <int N, bool TH, int ND, typename R = double, typename WIn = double, bool BATCH = false,
          int KIND = KIND_BOTH, int MK = MK_ANY>
__global__ __launch_bounds__(64, KIND == KIND_PICARD && sizeof(R) == sizeof(float) ? 2 : GPR_PRED_WAVES) void k_predictor(PredArgs a) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = PredShape<N, ND>::NS;
  constexpr int NN = PredShape<N, ND>::NN;
  constexpr int CPW = PredShape<N, ND>::CPW;
  constexpr int LPC = PredShape<N, ND>::LPC;
  constexpr int NF = FluxPack<TH>::NF, FS = FluxPack<TH>::STRIDE;
  constexpr int PRE = TH ? 27 : 18;
  constexpr bool FB = KIND != KIND_PICARD;  // this instantiation has the fallback
  static_assert(!FB || sizeof(R) == sizeof(double), "the fallback computes in double");
  // LDS per wavefront (17.5 KB for N = 3, 2D, V = 14, so that two
  // wavefronts per SIMD fit in 160 KB).  sQ holds the iterate (the
  // derivative contractions and the fallback's time column read it), then in
  // the fallback P g on the distortion rows; sF holds the x fluxes, then the
  // y fluxes, then the time solve's right-hand side.
  __shared__ R sQ[CPW * NN * V];
  __shared__ R sF[CPW * NN * V];
  __shared__ double sPre[FB ? CPW * NS * PRE : 1];
  __shared__ R sW[CPW * NS * V];       // the WENO state of each spatial node
  __shared__ R sB[4 * N * N];          // der, dgder, dgmatinv, dgmat
  __shared__ R sRat[LPC == 32 ? 1 : WAVE];

  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / LPC;
  const bool lane_ok = slot < CPW && lane - slot * LPC < NN;  // lanes outside a whole cell write no LDS
  const int node = lane_ok ? lane - slot * LPC : 0;
  const int t = node / NS;
  const int s = node - t * NS;
  const int ia = ND == 1 ? s : s / N;
  const int ib = ND == 1 ? 0 : s - (s / N) * N;
  long cell;
  bool valid;
  double dt, dxv;
  [[maybe_unused]] long real_cell = -1;  // batched: the cell's index in u if it is a real cell
  // the material: a reference straight into the kernel argument unless
  // batched (a pointer chosen on a branch would make the compiler copy the
  // whole argument to the stack)
  const Mat &M = [&]() -> const Mat & {
    if constexpr (BATCH)
      return a.binst[a.pblk[blockIdx.x]].m;
    else
      return a.m;
  }();
  if constexpr (BATCH) {
    const int k = a.pblk[blockIdx.x];
    const BatchInst &I = a.binst[k];
    const long lc = (long)(blockIdx.x - I.blk_p) * CPW + slot;
    cell = I.e0 + lc;
    valid = lane_ok && lc < I.nx + 2 && !a.bst[k].done;
    dt = a.bst[k].dt;
    dxv = I.dx;
    if (lc >= 1 && lc <= I.nx) real_cell = I.u0 + lc - 1;
  } else if constexpr (KIND == KIND_FALLBACK) {
    const long g0 = (long)blockIdx.x * CPW;
    const long cnt = *a.list_count;
    if (g0 >= cnt) return;  // uniform over the block
    valid = lane_ok && g0 + slot < cnt;
    cell = valid ? a.list[g0 + slot] : 0;
    dt = dev_or(a.dt_dev, a.dt);
    dxv = a.dx;
  } else {
    cell = (long)blockIdx.x * CPW + slot;
    valid = lane_ok && cell < a.ncell;
    dt = dev_or(a.dt_dev, a.dt);
    dxv = a.dx;
  }
  const bool active = valid && (BATCH || KIND == KIND_FALLBACK || a.mask[cell]);
  const int sl = lane_ok ? slot : 0;
  R *myQ = sQ + (sl * NN) * V;
  R *myF = sF + (sl * NN) * V;
  R *myW = sW + (sl * NS) * V;
  double *myPre = sPre + (FB ? (sl * NS + s) * PRE : 0);

  // basis tables and the WENO state into LDS
  if (lane < 4 * N * N) {
    const int tb = lane / (N * N), e = lane - tb * N * N;
    const double *src = tb == 0 ? &K.der[0][0] : tb == 1 ? &K.dgder[0][0] : tb == 2 ? &K.dgmatinv[0][0] : &K.dgmat[0][0];
    sB[lane] = R(src[(e / N) * MAXN + e % N]);
  }
  R q0[V];
  {
    const WIn *wsrc = (const WIn *)a.wh + (cell * NS + s) * V;
#pragma unroll
    for (int v = 0; v < V; v++) q0[v] = active ? R(wsrc[v]) : R(v == 0 || v == 5 || v == 9 || v == 13 ? 1 : 0);
    if (lane_ok && t == 0)
#pragma unroll
      for (int v = 0; v < V; v++) myW[s * V + v] = q0[v];
  }
  const R *Bder = sB, *Bdgder = sB + N * N, *Bmi = sB + 2 * N * N, *Bmt = sB + 3 * N * N;

  const R dtR = R(dt);
  const R Ws = R(ND == 1 ? K.w[ia] : K.w[ia] * K.w[ib]);
  const R iWs = R(1.0 / (ND == 1 ? K.w[ia] : K.w[ia] * K.w[ib]));
  const R Wt = R(K.w[t]);
  const R cW = R(K.endv[0][t] * (ND == 1 ? K.w[ia] : K.w[ia] * K.w[ib]));  // psi_t(0) W_s
  const R inv_dx = R(1.0 / dxv), inv_dy = R(1.0 / a.dy);

  // phase: 0 Picard, 1 preconditioned, 2 converged, 3 idle / given up
  int phase = active ? ((KIND == KIND_FALLBACK || (FB && a.mode == 1)) ? 1 : 0) : 3;
  int it = 0, pit = 0, grow = 0;
  R prev = R(1e30);
  bool fb_used = phase == 1;  // the cell ran (or runs) the fallback
  const unsigned long long cellbits = (NN == 64 ? ~0ull : ((1ull << NN) - 1)) << (sl * LPC);
  bool need_pre = phase == 1;
  wave_sync();  // sB, sW

  while (__any(phase == 0 || phase == 1)) {
    if constexpr (FB) {
      if (need_pre && t == 0 && lane_ok) {
        R w[V];
#pragma unroll
        for (int v = 0; v < V; v++) w[v] = myW[s * V + v];
        stiff_precond<N, TH, MK>(w, M, dt, K, myPre);
      }
    }
    need_pre = false;
    // (1) share the iterate
    if (lane_ok)
#pragma unroll
      for (int v = 0; v < V; v++) myQ[node * V + v] = q0[v];
    wave_sync();

    // (2) spatial derivatives, source, non-conservative products, fluxes
    R rhs[V], G[ND == 2 ? V : 1];
    {
      NodeState<R, TH> st;
      node_state<R, TH, MK>(q0, M, st);
      source<R, TH, MK>(q0, st, M, rhs);
      // B_d(q) . dq/dx_d reads only the distortion rows of dq (equations.cpp:189-247)
      R dq[14];
      {
        R c0 = Bder[ia * N + 0], c1 = Bder[ia * N + 1], c2 = N > 2 ? Bder[ia * N + N - 1] : R(0);
#pragma unroll
        for (int v = 5; v < 14; v++) {
          R acc = 0;
#pragma unroll
          for (int j = 0; j < N; j++)
            acc += (j == 0 ? c0 : j == 1 ? c1 : c2) * myQ[(t * NS + (ND == 1 ? j : j * N + ib)) * V + v];
          dq[v] = acc * inv_dx;
          GPR_CHUNK_FENCE(v);
        }
      }
      bdot_sub(st.v, dq, 0, R(1), rhs);
#pragma unroll
      for (int v = 5; v < 14; v++) pin(rhs[v]);
      if (ND == 2) {
        R c0 = Bder[ib * N + 0], c1 = Bder[ib * N + 1], c2 = N > 2 ? Bder[ib * N + N - 1] : R(0);
#pragma unroll
        for (int v = 5; v < 14; v++) {
          R acc = 0;
#pragma unroll
          for (int j = 0; j < N; j++) acc += (j == 0 ? c0 : j == 1 ? c1 : c2) * myQ[(t * NS + ia * N + j) * V + v];
          dq[v] = acc * inv_dy;
          GPR_CHUNK_FENCE(v);
        }
        bdot_sub(st.v, dq, 1, R(1), rhs);
      }
      const R wts = Wt * Ws;
#pragma unroll
      for (int v = 0; v < V; v++) {
        rhs[v] *= wts;
        pin(rhs[v]);
      }
      R F[V];
      flux<R, TH>(q0, st, M, 0, F);
      const R wb = R(ND == 2 ? K.w[ib] : 1.0);
      if (lane_ok)
#pragma unroll
        for (int i = 0; i < NF; i++) myF[node * FS + i] = F[flux_var(i, 0)] * wb;
      if constexpr (ND == 2) {
        flux<R, TH>(q0, st, M, 1, G);
#pragma unroll
        for (int i = 0; i < NF; i++) pin(G[flux_var(i, 1)]);
      }
    }
    wave_sync();
    // (3) flux contractions (rhs2, dg2.cpp:173-201): x, then y through the
    // same buffer
    {
      const R cx = Wt * inv_dx;
      R c0 = Bdgder[ia * N + 0], c1 = Bdgder[ia * N + 1], c2 = N > 2 ? Bdgder[ia * N + N - 1] : R(0);
#pragma unroll
      for (int i = 0; i < NF; i++) {
        R sx = 0;
#pragma unroll
        for (int j = 0; j < N; j++)
          sx += (j == 0 ? c0 : j == 1 ? c1 : c2) * myF[(t * NS + (ND == 1 ? j : j * N + ib)) * FS + i];
        rhs[flux_var(i, 0)] -= cx * sx;
        pin(rhs[flux_var(i, 0)]);
        GPR_CHUNK_FENCE(i);
      }
    }
    if constexpr (ND == 2) {
      wave_sync();
      if (lane_ok)
#pragma unroll
        for (int i = 0; i < NF; i++) myF[node * FS + i] = G[flux_var(i, 1)];
      wave_sync();
      const R cy = Wt * R(K.w[ia]) * inv_dy;
      R c0 = Bdgder[ib * N + 0], c1 = Bdgder[ib * N + 1], c2 = N > 2 ? Bdgder[ib * N + N - 1] : R(0);
#pragma unroll
      for (int i = 0; i < NF; i++) {
        R sy = 0;
#pragma unroll
        for (int j = 0; j < N; j++) sy += (j == 0 ? c0 : j == 1 ? c1 : c2) * myF[(t * NS + ia * N + j) * FS + i];
        rhs[flux_var(i, 1)] -= cy * sy;
        pin(rhs[flux_var(i, 1)]);
        GPR_CHUNK_FENCE(i);
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
      rhs[v] = rhs[v] * dtR + cW * myW[s * V + v];
      pin(rhs[v]);
    }
    // (4) time solve.  Picard: q1 = (DG_MAT (x) W_s)^-1 rhs.  Fallback:
    // g = DG_MAT q0 - rhs / W_s (q0 of the time column from sQ); sF gets
    // g - P g on the distortion rows and g elsewhere, and P g replaces the
    // distortion rows of q0 in sQ.
    [[maybe_unused]] R P[9];
    if constexpr (FB) {
      if (phase == 1) {
        {
          R m0 = Bmt[t * N + 0], m1 = Bmt[t * N + 1], m2 = N > 2 ? Bmt[t * N + N - 1] : R(0);
#pragma unroll
          for (int v = 0; v < V; v++) {
            R acc = 0;
#pragma unroll
            for (int k = 0; k < N; k++) acc += (k == 0 ? m0 : k == 1 ? m1 : m2) * myQ[(k * NS + s) * V + v];
            rhs[v] = acc - rhs[v] * iWs;  // rhs now holds g
            pin(rhs[v]);
            GPR_CHUNK_FENCE(v);
          }
        }
        proj_dev_sym(myPre + 9, rhs + 5, P);
#pragma unroll
        for (int k = 0; k < 9; k++) rhs[5 + k] -= P[k];
      }
    }
    wave_sync();  // every lane has read its time column of q0
    if (lane_ok) {
#pragma unroll
      for (int v = 0; v < V; v++) myF[node * V + v] = rhs[v];
      if constexpr (FB) {
        if (phase == 1)
#pragma unroll
          for (int k = 0; k < 9; k++) myQ[node * V + 5 + k] = P[k];
      }
    }
    wave_sync();
    {
      R m0 = Bmi[t * N + 0], m1 = Bmi[t * N + 1], m2 = N > 2 ? Bmi[t * N + N - 1] : R(0);
      if (FB && phase == 1) {
        if constexpr (FB) {
          const double *Ainv = myPre, *Jinv = myPre + 18;
#pragma unroll
          for (int v = 0; v < V; v++) {
            R d = 0;
            if (v >= 5 && v < 14) {
#pragma unroll
              for (int k = 0; k < N; k++) {
                R hk = myF[(k * NS + s) * V + v];
                R pk = myQ[(k * NS + s) * V + v];
                d += (k == 0 ? m0 : k == 1 ? m1 : m2) * hk + Ainv[t * N + k] * pk;
              }
            } else if (v >= 14) {
#pragma unroll
              for (int k = 0; k < N; k++) d += Jinv[t * N + k] * myF[(k * NS + s) * V + v];
            } else {
#pragma unroll
              for (int k = 0; k < N; k++) d += (k == 0 ? m0 : k == 1 ? m1 : m2) * myF[(k * NS + s) * V + v];
            }
            rhs[v] = q0[v] - d;  // rhs now holds q1
            pin(rhs[v]);
            GPR_CHUNK_FENCE(v);
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) {
          R acc = 0;
#pragma unroll
          for (int k = 0; k < N; k++) acc += (k == 0 ? m0 : k == 1 ? m1 : m2) * myF[(k * NS + s) * V + v];
          rhs[v] = acc * iWs;  // rhs now holds q1
          pin(rhs[v]);
          GPR_CHUNK_FENCE(v);
        }
      }
    }
    // (5) stopping rule over the cell's whole block (dg2.cpp:278-297)
    const R tol = R(phase == 1 ? a.stiff_tol : a.dg_tol);
    bool more = false, nan = false;
    R ratio = 0;
    // the stopping statistic feeds the growth tests only: every iteration for
    // bailout, and from iteration bail_from on (its value there is the first
    // one compared) for bail_from
    const bool want_ratio = __any(a.bailout > 0 || (a.bail_from > 0 && phase == 0 && it + 1 >= a.bail_from));
#pragma unroll
    for (int v = 0; v < V; v++) {
      R dd = fabs(rhs[v] - q0[v]), lim = tol * (R(1) + fabs(q0[v]));
      more |= dd > lim;
      nan |= rhs[v] != rhs[v];
      if (want_ratio) ratio = fmax(ratio, fast_div(dd, lim));
    }
    const unsigned long long bm = __ballot(more), bn = __ballot(nan);
    const bool cmore = (bm & cellbits) != 0, cnan = (bn & cellbits) != 0;
    if (want_ratio) {  // the cell's largest ratio, in each of its lanes (wave-uniform branch)
      if constexpr (LPC == 32) {
        ratio = R(half_wave_max(lane_ok ? double(ratio) : 0.0));
      } else {
        sRat[lane] = ratio;
        wave_sync();
        R cr = 0;
        for (int k = 0; k < NN; k++) cr = fmax(cr, sRat[sl * NN + k]);
        ratio = cr;
      }
    }
    // the next iterate (reset below on a restart); a cell already solved keeps
    // its result, which the fused epilogue reads
    if (phase == 0 || phase == 1)
#pragma unroll
      for (int v = 0; v < V; v++) q0[v] = rhs[v];
    [[maybe_unused]] bool push = false;  // KIND_PICARD: append this cell to a.list
    if (phase == 0) {
      it++;
      bool to_fallback = false;
      if (cmore) {
        grow = ratio > prev ? grow + 1 : 0;
        prev = ratio;
        if (it >= a.dg_it || (a.bailout > 0 && grow >= a.bailout) || (a.bail_from > 0 && it > a.bail_from && grow > 0))
          to_fallback = true;
      } else if (cnan) {
        to_fallback = true;
      } else {
        if (a.qh)
#pragma unroll
          for (int v = 0; v < V; v++) a.qh[(cell * NN + node) * V + v] = rhs[v];
        if (node == 0) {
          if (a.iters) a.iters[cell] = it;
        }
        phase = 2;
      }
      if (to_fallback) {
        if constexpr (FB) {
          phase = 1;
          pit = 0;
          need_pre = true;
          fb_used = true;
#pragma unroll
          for (int v = 0; v < V; v++) q0[v] = myW[s * V + v];
        } else {
          // solved again by the KIND_FALLBACK launch over a.list
          if (node == 0) push = true;
          phase = 3;
        }
      }
    } else if (FB && phase == 1) {
      pit++;
      if (!cmore || pit >= a.stiff_it) {
        if (a.qh)
#pragma unroll
          for (int v = 0; v < V; v++) a.qh[(cell * NN + node) * V + v] = rhs[v];
        bool ok = !cmore && !cnan;
        if (node == 0) {
          if (a.iters) a.iters[cell] = ok ? 1000 + pit : -1;
          if (!ok) atomicAdd(a.fail, 1);
        }
        phase = ok ? 2 : 3;
      }
    }
    if constexpr (!FB) {
      // one atomic per wavefront for the cells it appends to a.list
      const unsigned long long pm = __ballot(push);
      if (pm) {
        int base = 0;
        if (lane == 0) base = atomicAdd(a.list_count, (int)__popcll(pm));
        base = __shfl(base, 0);
        if (push) a.list[base + __popcll(pm & ((1ull << lane) - 1))] = (int)cell;
      }
    }
    wave_sync();  // sQ, sF and sRat are rewritten next iteration
  }
  // Fused FV corrector: the final iterate of every node (q0 holds it) gives
  // the cell's face states and its cell integral (what k_fv_faces / kb_faces
  // extrapolate and k_fv_cells / kb_cells integrate from qh), so qh need not
  // go through HBM.  Batched: cell indexes the instances' concatenated
  // extended cells and real_cell the concatenated cells.
  if constexpr (KIND == KIND_BOTH && sizeof(R) == sizeof(double) && MK == MK_GAS_NEWTONIAN) {
    if (a.fs) {  // uniform
      using FS = FaceStates<N, ND>;
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myQ[node * V + v] = q0[v];
      wave_sync();
      if (active) {
        // the cell's SIDES x NPS face-node states (side, time node tt, face
        // node fn), dealt over its lanes: the extrapolation to side e of the
        // N nodes along the side's normal
        for (int job = node; job < FS::SIDES * FS::NPS; job += NN) {
          const int side = job / FS::NPS, rem = job - side * FS::NPS;
          const int tt = rem / FS::NSF, fn = rem - tt * FS::NSF;
          const int e = side & 1;
          double *dst = a.fs + ((cell * FS::SIDES + side) * FS::NPS + rem) * V;
#pragma unroll
          for (int v = 0; v < V; v++) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < N; k++) {
              const int nl = ND == 1 ? k : (side < 2 ? k * N + fn : fn * N + k);
              acc += K.endv[e][k] * (double)myQ[(tt * NS + nl) * V + v];
            }
            dst[v] = acc;
          }
        }
      }
      // the cell integral of the real cells (extended cells minus the ghost ring)
      const long ey = ND == 2 ? a.fny + 2 : 1;
      const long ci = ND == 2 ? cell / ey - 1 : cell - 1, cj = ND == 2 ? cell % ey - 1 : 0;
      const bool real = BATCH ? valid && real_cell >= 0
                              : valid && ci >= 0 && ci < a.fnx && cj >= 0 && cj < (ND == 2 ? a.fny : 1);
      double S[V];
      cell_integrand<N, TH, ND, true, true, MK>((const double *)q0, (const double *)myQ, t, ia, ib, M, K, dt, dxv,
                                                a.dy, S);
      wave_sync();  // every lane has read sQ
      if (lane_ok)
#pragma unroll
        for (int v = 0; v < V; v++) myF[node * V + v] = active ? S[v] : 0.0;
      wave_sync();
      if (real) {
        const long c = BATCH ? real_cell : ND == 2 ? ci * a.fny + cj : ci;
        for (int v = node; v < V; v += NN) {
          double du = 0;
          for (int k = 0; k < NN; k++) du += (double)myF[k * V + v];
          a.cint[c * V + v] = du;
        }
      }
    }
  }
  if (a.itc) {  // every lane is active here
    // node 0 of each solved cell: its Picard iterations (before the fallback,
    // if it ran), fallback iterations, and whether it was solved / fell back
    unsigned long long n_pic = 0, n_fb = 0, n_cells = 0, n_fbcells = 0;
    if (active && lane_ok && node == 0) {
      n_pic = (unsigned long long)it;
      if (fb_used) {
        n_fb = (unsigned long long)pit;
        n_cells = 1;
        n_fbcells = 1;
      } else if (phase == 2) {
        n_cells = 1;
      }
    }
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      n_pic += __shfl_xor(n_pic, o);
      n_fb += __shfl_xor(n_fb, o);
      n_cells += __shfl_xor(n_cells, o);
      n_fbcells += __shfl_xor(n_fbcells, o);
    }
    if (lane == 0) {
      unsigned long long *slot = a.itc + (blockIdx.x % ITC_SLOTS) * ITC_STRIDE;
      if (n_pic) atomicAdd(&slot[0], n_pic);
      if (n_fb) atomicAdd(&slot[1], n_fb);
      if (n_cells) atomicAdd(&slot[2], n_cells);
      if (n_fbcells) atomicAdd(&slot[3], n_fbcells);
    }
  }
}

}  // namespace gpr
