// A batch of independent 1D GPR problems (uneven sizes, own material, own dt,
// own end time) advanced together: each stage of the step is one launch over
// every instance, and the step loop is captured in a hipGraph by the host.
//
// Layout: the instances' arrays are concatenated (BatchInst offsets).  Kernels
// that give a wavefront a group of cells (predictor, cell update) or faces
// never mix instances, so a wavefront's material, dx and dt are uniform and
// come from the instance's BatchInst / BatchState; the element-wise kernels
// find their instance by binary search over the offsets.
#pragma once

#include "gpr_kernels.hpp"

namespace gpr {

struct BatchArgs {
  const BatchInst *inst;
  BatchState *st;
  int count;
  const long *u0s, *b0s, *e0s;  // instance offsets (count + 1 entries, last = total)
  const int *cblk, *fblk;       // cell-update and face block -> instance
  double *u;                    // [sum nx][V]
  double *wh;                   // [sum (nx + 2)][N][V]
  double *qh;                   // [sum (nx + 2)][N^2][V]
  double *fx;                   // [sum (nx + 1)][2][V]
  // fused corrector (the predictor's epilogue): face states per extended
  // cell (FaceStates<N, 1> layout, in qh's space) and cell integrals [sum nx][V]
  const double *fs;
  const double *cint;
  // the view: the instances the element-wise launches cover (all, or the
  // running ones), vinst[j] for j < vcount, with their cumulative cell,
  // ghost-extended row and extended cell counts (vcount + 1 entries each)
  const int *vinst;
  int vcount;
  const long *vu0s, *vb0s, *ve0s;
};

// index k with starts[k] <= x < starts[k + 1]
__device__ __forceinline__ int find_inst(const long *starts, int count, long x) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// weno_launcher in 1D (weno.cpp:145-147): one thread per (extended cell,
// variable).  The stencil reads the ghost-extended row I of the instance
// straight from u through boundaries1's map (grid.cpp:12-145: the source
// cell, and the boundary code's reflection), so no ghost-extended copy of u
// is written.
template <int N, int V>
__global__ void kb_weno(BatchArgs b) {
  const Basis &K = basis_dev<N>();
  const long total = b.ve0s[b.vcount] * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int v = id % V;
    const long ve = id / V;
    const int j = find_inst(b.ve0s, b.vcount, ve);
    const int k = b.vinst[j];
    const BatchInst &I = b.inst[k];
    if (b.st[k].done) continue;
    const int i = (int)(ve - b.ve0s[j]);
    const long e = I.e0 + i;
    double d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) {
      const int row = i + r;  // row of the ghost-extended instance
      bool refl;
      const int code = row < N ? I.bc0 : I.bc1;
      const double *src = b.u + (I.u0 + bc_src(row, I.nx, N, code, refl)) * V;
      d[r] = refl ? bc_value(v, code, 0, [&](int m) { return src[m]; }) : src[v];
    }
    weno_point<double, N>(d, K, I.nx + 2, 0, i, o);
#pragma unroll
    for (int a = 0; a < N; a++) b.wh[(e * N + a) * V + v] = o[a];
  }
}

// interfs1 (fv.cpp:111-160): one lane per (face, time node), as k_fv_faces
// MK: the material kind Rusanov's physics is compiled for (every instance
// of the batch of that kind, or MK_ANY); the values do not depend on it.
// FS: the face states come from b.fs (the predictor's fused epilogue)
// instead of being extrapolated from qh here; the values are the same.
template <int N, bool TH, int MK = MK_ANY, bool FS = false>
__global__ __launch_bounds__(64) void kb_faces(BatchArgs b) {
  constexpr int V = Dims<TH>::V;
  constexpr int NT = N, FPW = WAVE / NT, NN = N * N;
  __shared__ double sR[FPW * NT * 2 * V];
  const Basis &K = basis_dev<N>();
  const int k = b.fblk[blockIdx.x];
  const BatchInst &I = b.inst[k];
  if (b.st[k].done) return;  // uniform over the block
  const int lane = threadIdx.x;
  const int fl = lane / NT, t = lane - fl * NT;
  const long lf0 = (long)(blockIdx.x - I.blk_f) * FPW;
  const long lf = lf0 + fl;
  const bool ok = fl < FPW && lf < I.nx + 1;
  const double dt = b.st[k].dt;
  double lo[V], hi[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    lo[v] = 0.0;
    hi[v] = 0.0;
  }
  if (ok) {
    const long e0 = I.e0 + lf, e1 = e0 + 1;
    const double kk = dt * K.w[t] / (2.0 * I.dx);
    double qL[V], qR[V];
    if constexpr (FS) {
      using F = FaceStates<N, 1>;  // e0's right side (1), e1's left side (0)
      const double *pl = b.fs + ((e0 * F::SIDES + 1) * F::NPS + t) * V;
      const double *pr = b.fs + ((e1 * F::SIDES + 0) * F::NPS + t) * V;
#pragma unroll
      for (int v = 0; v < V; v++) {
        qL[v] = pl[v];
        qR[v] = pr[v];
      }
    } else {
      const double *pl = b.qh + (e0 * NN + t * N) * V;
      const double *pr = b.qh + (e1 * NN + t * N) * V;
#pragma unroll
      for (int v = 0; v < V; v++) {
        double l = 0, rr = 0;
#pragma unroll
        for (int a = 0; a < N; a++) {
          l += K.endv[1][a] * pl[a * V + v];
          rr += K.endv[0][a] * pr[a * V + v];
        }
        qL[v] = l;
        qR[v] = rr;
      }
    }
    riemann_accum<N, TH, double, MK>(qL, qR, 0, I.m, K, kk, lo, hi);
  }
  if (fl < FPW) {
    double *my = sR + (fl * NT + t) * 2 * V;
#pragma unroll
    for (int v = 0; v < V; v++) {
      my[v] = lo[v];
      my[V + v] = hi[v];
    }
  }
  __syncthreads();
  for (int o = lane; o < FPW * 2 * V; o += WAVE) {
    const int f = o / (2 * V), c = o - f * 2 * V;
    if (lf0 + f >= I.nx + 1) continue;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < NT; q++) acc += sR[(f * NT + q) * 2 * V + c];
    b.fx[(I.f0 + lf0 + f) * 2 * V + c] = acc;
  }
}

// centers1 (fv.cpp:11-54) and the face terms: one lane per space-time node,
// as k_fv_cells in 1D
template <int N, bool TH>
__global__ __launch_bounds__(64) void kb_cells(BatchArgs b) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = N, NN = N * N, CPW = WAVE / NN;
  __shared__ double sQ[CPW * NN * V];
  const Basis &K = basis_dev<N>();
  const int k = b.cblk[blockIdx.x];
  const BatchInst &I = b.inst[k];
  if (b.st[k].done) return;  // uniform over the block
  const int lane = threadIdx.x;
  const int slot = lane / NN;
  const int node = lane - slot * NN;
  const int t = node / NS, ia = node - t * NS;
  const bool lane_ok = slot < CPW;
  const long lc = (long)(blockIdx.x - I.blk_c) * CPW + slot;
  const bool valid = lane_ok && lc < I.nx;
  const long e = I.e0 + lc + 1;
  const double dt = b.st[k].dt;
  const int sl = lane_ok ? slot : 0;
  double *myQ = sQ + sl * NN * V;
  double q[V];
#pragma unroll
  for (int v = 0; v < V; v++) q[v] = valid ? b.qh[(e * NN + node) * V + v] : (v == 0 ? 1.0 : 0.0);
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = q[v];
  __syncthreads();
  double S[V];
  {
    NodeState<double, TH> st;
    node_state<double, TH>(q, I.m, st);
    source<double, TH>(q, st, I.m, S);
    double dq[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = 0;
#pragma unroll
      for (int j = 0; j < N; j++) acc += K.der[ia][j] * myQ[(t * NS + j) * V + v];
      dq[v] = acc / I.dx;
    }
    bdot_sub(st.v, dq, 0, 1.0, S);
    const double wgt = dt * K.w[t] * K.w[ia];
#pragma unroll
    for (int v = 0; v < V; v++) S[v] = valid ? wgt * S[v] : 0.0;
  }
  __syncthreads();
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = S[v];
  __syncthreads();
  for (int v = node; valid && v < V; v += NN) {
    double du = 0;
    for (int j = 0; j < NN; j++) du += myQ[j * V + v];
    const double *fl = b.fx + (I.f0 + lc) * 2 * V;      // left face: hi side
    const double *fr = b.fx + (I.f0 + lc + 1) * 2 * V;  // right face: lo side
    du -= fl[V + v] + fr[v];
    b.u[(I.u0 + lc) * V + v] += du;
  }
}

// CFL maximum of one value per lane into its instance: one atomic per
// wavefront when the wavefront's cells share an instance.  sp is never NaN
// (callers keep a running maximum that NaN cannot enter, as std::max).
__device__ __forceinline__ void batch_max(BatchState *st, int k, bool work, double sp, bool second) {
  const int k0 = __shfl(k, 0);
  if (__all(k == k0)) {
    const double mx = wave_max(sp);
    if ((threadIdx.x & 63) == 0) atomic_max_pos(second ? &st[k0].maxspeed2 : &st[k0].maxspeed, mx);
  } else if (work) {
    atomic_max_pos(second ? &st[k].maxspeed2 : &st[k].maxspeed, sp);
  }
}

// Post-step work, one thread per cell, as k_post: renormalise A, the
// instance's contortion and non-finite flags, and its CFL maximum for the
// next dt (cfl_only: the maximum alone, for the first step).
// UPD: first the fused corrector's cell update, u += cint - the two face
// terms (the operations with which kb_cells ends), on the state the
// post-step then works on; the whole updated state is written back.
template <bool TH, bool UPD = false>
__global__ void kb_post(BatchArgs b, int cfl_only) {
  constexpr int V = Dims<TH>::V;
  const long total = b.vu0s[b.vcount];
  const long vc = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool in = vc < total;
  const int j = find_inst(b.vu0s, b.vcount, in ? vc : total - 1);
  const int k = b.vinst[j];
  const BatchInst &I = b.inst[k];
  const long c = I.u0 + (in ? vc : total - 1) - b.vu0s[j];
  BatchState &S = b.st[k];
  const bool work = in && !S.done;
  double sp = 0.0;
  if (work) {
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = b.u[c * V + v];
    if constexpr (UPD) {
      const long lc = c - I.u0;
      const double *fl = b.fx + (I.f0 + lc) * 2 * V;      // left face: hi side
      const double *fr = b.fx + (I.f0 + lc + 1) * 2 * V;  // right face: lo side
#pragma unroll
      for (int v = 0; v < V; v++) {
        double du = b.cint[c * V + v];
        du -= fl[V + v] + fr[v];
        q[v] += du;
      }
    }
    if (!cfl_only) {
      if (q[0] > 0) {
        double cf = cbrt(q[0] / (I.m.p.rho0 * det3(q + 5)));
#pragma unroll
        for (int j = 0; j < 9; j++) q[5 + j] *= cf;
      }
      double f = 0;
#pragma unroll
      for (int j = 0; j < 9; j++) f += q[5 + j] * q[5 + j];
      double d13 = cbrt(fabs(det3(q + 5)));
      if (f / (d13 * d13) / 3.0 - 1.0 > S.contorted_tol) atomicOr(&S.contorted, 1);
      bool nonfinite = false;
#pragma unroll
      for (int v = 0; v < V; v++)
        if (!(q[v] - q[v] == 0.0)) nonfinite = true;
      if (nonfinite) atomicOr(&S.nonfinite, 1);
      if constexpr (UPD) {
#pragma unroll
        for (int v = 0; v < 5; v++) b.u[c * V + v] = q[v];
#pragma unroll
        for (int v = 14; v < V; v++) b.u[c * V + v] = q[v];
      }
#pragma unroll
      for (int j = 0; j < 9; j++) b.u[c * V + 5 + j] = q[5 + j];
    }
    const double s = max_abs_eigs<TH>(q, I.m, 0) / I.dx;
    if (s > sp) sp = s;
  }
  batch_max(b.st, k, work, sp, false);
}

// reset_distortion of every cell of a contorted instance (utils.cpp:44-63,
// iterator.cpp:150-151), with the CFL maximum recomputed into maxspeed2
template <bool TH>
__global__ void kb_reset(BatchArgs b) {
  constexpr int V = Dims<TH>::V;
  const long total = b.vu0s[b.vcount];
  const long vc = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool in = vc < total;
  const int j = find_inst(b.vu0s, b.vcount, in ? vc : total - 1);
  const int k = b.vinst[j];
  const BatchInst &I = b.inst[k];
  const long c = I.u0 + (in ? vc : total - 1) - b.vu0s[j];
  BatchState &S = b.st[k];
  const bool work = in && !S.done && S.contorted;
  double sp = 0.0;
  if (work) {
    const Mat &m = I.m;
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = b.u[c * V + v];
    double r = q[0];
    double cf = cbrt(r / m.p.rho0);
    NodeState<double, TH> s;
    node_state<double, TH>(q, m, s);
    double Gm = Gamma_MG(r, m.p);
    double C0 = (1 - m.p.beta / Gm) * s.cs2;
    q[1] -= r * (C0 / 4 * s.devGsq);
#pragma unroll
    for (int j = 0; j < 9; j++) q[5 + j] = 0.0;
    q[5] = cf;
    q[9] = cf;
    q[13] = cf;
#pragma unroll
    for (int v = 0; v < V; v++) b.u[c * V + v] = q[v];
    const double x = max_abs_eigs<TH>(q, m, 0) / I.dx;
    if (x > sp) sp = x;
  }
  batch_max(b.st, k, work, sp, true);
}

// t += dt, count += 1 for the instances still running (iterator.cpp:142-143)
__global__ void kb_advance(BatchArgs b) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.count) return;
  BatchState &S = b.st[k];
  if (S.done) return;
  S.t += S.dt;
  S.count += 1;
}

// the next dt of each running instance (timestep, iterator.cpp:54-83); an
// instance that reached its end time, or whose state became non-finite
// (iterator.cpp:161-165), stops
__global__ void kb_finalize(BatchArgs b) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.count) return;
  BatchState &S = b.st[k];
  if (S.done) return;
  unsigned long long bits = S.contorted ? S.maxspeed2 : S.maxspeed;
  double MAX = __longlong_as_double((long long)bits);
  double dt = S.cfl / MAX;
  if (S.count <= 5) dt *= 0.2;
  if (S.t + dt > S.tf) dt = S.tf - S.t;
  S.dt = dt;
  S.maxspeed = 0ull;
  S.maxspeed2 = 0ull;
  S.contorted = 0;
  if (S.t >= S.tf || S.nonfinite) {
    S.done = 1;
    S.dt = 0.0;
  }
}

}  // namespace gpr
