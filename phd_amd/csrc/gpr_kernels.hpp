// HIP kernels of one ADER-WENO step for the GPR system on gfx950.
//
// Pipeline (one step, all on one stream, dt kept on the device):
//   k_boundaries  u -> ub (ghost layers, etc/grid.cpp:147-170)
//   k_weno_*      ub -> wh (solvers/weno/weno.cpp:75-157)
//   k_predictor   wh -> qh (solvers/dg/dg1.cpp, dg2.cpp): one space-time node
//                 per lane, the cell's nodes in one wavefront, contractions
//                 through LDS, Picard with the reference stopping rule and a
//                 stiff-source preconditioned fallback
//   k_fv_faces    qh -> face Rusanov + path-conservative jumps (fv.cpp:162-268)
//   k_fv_cells    cell sources/B terms + face gather -> u (fv.cpp:56-109),
//                 fused with the post-step distortion renormalisation, the
//                 contortion flag and the next step's CFL maximum
//   k_reset / k_finalize_dt   (iterator.cpp:54-83,145-151, utils.cpp:44-85)
#pragma once

#include <type_traits>

#include "gpr_eigflux.hpp"
#include "gpr_pointwise.hpp"
#include "gpr_relax.hpp"

namespace gpr {


// Device-resident time-loop state.
struct StepState {
  double t, tf, cfl, dt;
  long long count;
  unsigned long long maxspeed;  // bit pattern of a non-negative double (max over cells)
  unsigned long long maxspeed2; // same, recomputed after a global distortion reset
  int contorted;
  int nonfinite;
  int pred_fail;
  int done;
  double contorted_tol;
  int pushcount;  // snapshots taken
  int resets;     // global distortion resets so far
};

__device__ __forceinline__ void atomic_max_pos(unsigned long long *addr, double v) {
  if (!(v >= 0.0)) return;  // NaN (and negatives) never enter the maximum, as in std::max
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
  return v;
}

// ---------------------------------------------------------------- boundaries
// One thread per (ghost-extended cell, variable).  Boundary codes 0
// transmissive, 1 periodic, 2 slip, 3 stick, 4 moving, 5 symmetric
// (grid.cpp:5-10); DESTRESS is false in the reference (options.h:12).
__device__ __forceinline__ int bc_src(int I, int n, int N, int code, bool &refl) {
  refl = false;
  if (I >= N && I < n + N) return I - N;
  bool left = I < N;
  int i = left ? I : I - n - N;  // 0..N-1
  switch (code) {
  case 0: return left ? 0 : n - 1;
  case 1: return left ? n - N + i : i;
  default: refl = true; return left ? N - 1 - i : n - 1 - i;
  }
}
__device__ __forceinline__ double bc_sign(int v, int code, int d) {
  if (code == 2) return v == 2 + d ? -1.0 : 1.0;
  if (code == 3) return (v >= 2 && v <= 4) ? -1.0 : 1.0;
  if (code == 5) {
    if (v == 2 + d) return -1.0;
    if (v >= 5 && v < 14) {
      int k = v - 5, r = k / 3, c = k % 3;
      return ((r == d) != (c == d)) ? -1.0 : 1.0;  // row d and column d flipped: the (d,d) entry twice
    }
  }
  return 1.0;
}

// Moving wall (code 4, grid.cpp:44-54,100-110): the stick reflection, then
// rho v_x += 2 rho (a wall sliding at unit speed along x, whatever the wall's
// direction) and rho E += ((rho v_x)'^2 - (rho v_x)^2) / (2 rho), with
// (rho v_x) the source cell's value.  g(k) gives the source cell's variable k.
template <typename G> __device__ __forceinline__ double bc_value(int v, int code, int d, const G &g) {
  if (code == 4) {
    if (v == 1 || v == 2) {
      const double r = g(0), s = g(2);
      const double m = -s + 2.0 * r;
      return v == 2 ? m : g(1) + (m * m - s * s) / (2.0 * r);
    }
    return (v == 3 || v == 4) ? -g(v) : g(v);
  }
  return bc_sign(v, code, d) * g(v);
}

template <int V, int ND>
__global__ void k_boundaries(const double *__restrict__ u, double *__restrict__ ub, int nx, int ny, int N,
                             int4 bc) {
  long nby = ND == 2 ? ny + 2 * N : 1;
  long total = (long)(nx + 2 * N) * nby * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    int v = id % V;
    long c = id / V;
    int J = ND == 2 ? c % nby : 0;
    int I = ND == 2 ? c / nby : c;
    bool rx, ry = false;
    int codex = I < N ? bc.x : bc.y;
    int i = bc_src(I, nx, N, codex, rx);
    int j = 0, codey = 0;
    if (ND == 2) {
      codey = J < N ? bc.z : bc.w;
      j = bc_src(J, ny, N, codey, ry);
    }
    // the x pass (rows of u) and then the y pass over the x-extended rows, as
    // boundaries2 (grid.cpp:147-170) composes them
    const double *row = u + (ND == 2 ? (long)i * ny + j : (long)i) * V;
    auto xval = [&](int k) { return rx ? bc_value(k, codex, 0, [&](int m) { return row[m]; }) : row[k]; };
    ub[id] = ry ? bc_value(v, codey, 1, xval) : xval(v);
  }
}

// 1D: wh[(nx+2)][N][V] from ub[(nx+2N)][V] (weno_launcher ndim 1, weno.cpp:145-147)
template <typename R, int N, int V>
__global__ void k_weno_1d(const double *__restrict__ ub, R *__restrict__ wh, int nx, Basis K) {
  int NX = nx + 2;
  long total = (long)NX * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    int v = id % V, i = id / V;
    R d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) d[r] = R(ub[(long)(i + r) * V + v]);
    weno_point<R, N>(d, K, NX, 0, i, o);
#pragma unroll
    for (int a = 0; a < N; a++) wh[((long)i * N + a) * V + v] = o[a];
  }
}

// 2D x-pass: ux[(nx+2)][(ny+2N)][N][V] (weno.cpp:93-104, NO_CORNERS gap N on
// the first and last N columns of the ghost-extended grid)
template <typename R, int N, int V>
__global__ void k_weno_x(const double *__restrict__ ub, R *__restrict__ ux, int nx, int ny, Basis K) {
  int NX = nx + 2, ncol = ny + 2 * N;
  long total = (long)NX * ncol * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    int v = id % V;
    long c = id / V;
    int j = c % ncol, i = c / ncol;
    int gap = (j < N || j >= ncol - N) ? N : 0;
    R d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) d[r] = R(ub[((long)(i + r) * ncol + j) * V + v]);
    weno_point<R, N>(d, K, NX, gap, i, o);
#pragma unroll
    for (int a = 0; a < N; a++) ux[((c)*N + a) * V + v] = o[a];
  }
}

// 2D y-pass: wh[(nx+2)(ny+2)][N][N][V] (weno.cpp:109-121, gap 1 on the first
// and last ghost rows)
template <typename R, int N, int V>
__global__ void k_weno_y(const R *__restrict__ ux, R *__restrict__ wh, int nx, int ny, Basis K) {
  int NX = nx + 2, NY = ny + 2, ncol = ny + 2 * N;
  long total = (long)NX * NY * N * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    int v = id % V;
    long c = id / V;
    int a = c % N;
    c /= N;
    int j = c % NY, i = c / NY;
    int gap = (i == 0 || i == NX - 1) ? 1 : 0;
    R d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) d[r] = ux[(((long)i * ncol + j + r) * N + a) * V + v];
    weno_point<R, N>(d, K, NY, gap, j, o);
#pragma unroll
    for (int b = 0; b < N; b++) wh[((((long)i * NY + j) * N + a) * N + b) * V + v] = o[b];
  }
}

// 2D x- and y-passes in one launch: block (row i, tile of WENO_JT output
// columns) reconstructs the x-pass of the columns its y-pass stencils reach
// into LDS (the values k_weno_x writes to ux), then the y-pass from LDS, so
// the x-pass output never goes through HBM.  Rows are dealt to the XCDs in
// contiguous bands (block b on XCD b % 8 takes item (b % 8) M + b / 8), so
// that the five rows of ub a row's x-pass reads are shared in one L2.
constexpr int WENO_JT = 64;
template <typename R, int N, int V>
__global__ __launch_bounds__(256) void k_weno_xy(const double *__restrict__ ub, R *__restrict__ wh, int nx, int ny,
                                                 Basis K) {
  constexpr int NC = WENO_JT + 2 * N - 2;  // x-pass columns a tile's y-pass reads
  __shared__ R sx[NC * N * V];
  const int NX = nx + 2, NY = ny + 2, ncol = ny + 2 * N;
  const int ntile = (NY + WENO_JT - 1) / WENO_JT;
  const long nitems = (long)NX * ntile;
  const long M = (nitems + 7) / 8;
  const long item = (blockIdx.x % 8) * M + blockIdx.x / 8;
  if (item >= nitems) return;  // uniform over the block
  const int i = item / ntile, j0 = (item - (long)i * ntile) * WENO_JT;
  // x-pass of columns j0 .. j0 + NC - 1 of row i (weno.cpp:93-104)
  for (int k = threadIdx.x; k < NC * V; k += blockDim.x) {
    const int jj = k / V, v = k - jj * V, j = j0 + jj;
    if (j >= ncol) continue;
    const int gap = (j < N || j >= ncol - N) ? N : 0;
    R d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) d[r] = R(ub[((long)(i + r) * ncol + j) * V + v]);
    weno_point<R, N>(d, K, NX, gap, i, o);
#pragma unroll
    for (int a = 0; a < N; a++) sx[(jj * N + a) * V + v] = o[a];
  }
  __syncthreads();
  // y-pass of output columns j0 .. j0 + WENO_JT - 1 (weno.cpp:109-121)
  const int gap = (i == 0 || i == NX - 1) ? 1 : 0;
  for (int k = threadIdx.x; k < WENO_JT * N * V; k += blockDim.x) {
    const int v = k % V, a = (k / V) % N, jj = k / (V * N), j = j0 + jj;
    if (j >= NY) continue;
    R d[2 * N - 1], o[N];
#pragma unroll
    for (int r = 0; r < 2 * N - 1; r++) d[r] = sx[((jj + r) * N + a) * V + v];
    weno_point<R, N>(d, K, NY, gap, j, o);
#pragma unroll
    for (int b = 0; b < N; b++) wh[((((long)i * NY + j) * N + a) * N + b) * V + v] = o[b];
  }
}

struct FvArgs {
  double *u;             // [cells][V]
  const double *qh;      // [(nx+2)(ny+2)][NN][V]
  const uint8_t *mask;   // [(nx+2)(ny+2)]
  double *fx;            // x faces [(nx+1)][ny][2][V]
  double *fy;            // y faces [nx][(ny+1)][2][V]
  StepState *st;         // if non-null: dt from st->dt, post-step reductions into st
  double dt, dx, dy;
  int nx, ny;
  int post;              // fuse renormalisation, contortion flag and CFL maximum
  int time;              // 1: qh holds N time nodes (ADER); 0: one, weight dt (split step's wh)
  int sources;           // 1: the cell integral includes S(q)
  int ode;               // with post: the split step's closing half-step ODE first
  int parts;             // 1: cell integrals (centers1/2), 2: face terms (interfs1/2), 3: both
  int flux;              // Riemann solver (fluxes.h:7-9): 0 Rusanov, 1 Roe, 2 Osher
  double *eig_ws;        // Roe / Osher: eig_face_grid(items) x WAVE lanes x EIG_WS_DOUBLES(V)
  // fused form (ADER step, TIME and SOURCES): the predictor wrote each active
  // extended cell's face states (FaceStates layout) and each real cell's
  // cell integral; qh is then unused
  const double *fs;
  const double *cint;    // [nx][ny][V]
  Mat m;
};

// Faces: interfs1/interfs2 (fv.cpp:111-160, 162-268).  One lane per (face,
// time node t, face node s): it extrapolates both cells' states to the face
// node, evaluates one Rusanov flux and one path integral, and weights them;
// the NT = NTT * NSF lanes of a face are then summed through LDS.  Face d = 0
// joins extended cells (i, j+1) and (i+1, j+1) for i in [0, nx], j in
// [0, ny); face d = 1 joins (i+1, j) and (i+1, j+1) for i in [0, nx),
// j in [0, ny].  TIME = false is fv_launcher's TIME = false: the input holds
// one time level per cell (the split step's WENO reconstruction) and the
// time weight is dt (fv.cpp:156-159, 273-279).
template <int N, bool TH, int ND, bool TIME = true> struct FaceShape {
  static constexpr int NTT = TIME ? N : 1;     // time nodes
  static constexpr int NSF = ND == 2 ? N : 1;  // face nodes
  static constexpr int NT = NTT * NSF;         // lanes per face
  static constexpr int FPW = WAVE / NT;        // faces per wavefront
};

// Roe / Osher face launches run a bounded grid (each block loops over its
// share of the face groups) so that their workspace is bounded too.
constexpr long EIG_FACE_BLOCKS = 2048;
inline long eig_face_grid(long items) { return items < EIG_FACE_BLOCKS ? items : EIG_FACE_BLOCKS; }

// FS: the face states come from a.fs (the predictor's fused epilogue)
// instead of being extrapolated from qh here; the values are the same.
// EIG: the Riemann solver is a.flux's Roe (1) or Osher (2) instead of Rusanov.
// MK: the material kind Rusanov's physics is compiled for (MK_ANY: any
// material); the values do not depend on it.
template <int N, bool TH, int ND, bool TIME = true, bool FS = false, bool EIG = false, int MK = MK_ANY>
__global__ __launch_bounds__(64) void k_fv_faces(FvArgs a) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = ND == 1 ? N : N * N;
  constexpr int NN = NS * FaceShape<N, TH, ND, TIME>::NTT;
  constexpr int NSF = FaceShape<N, TH, ND, TIME>::NSF;
  constexpr int NT = FaceShape<N, TH, ND, TIME>::NT;
  constexpr int FPW = FaceShape<N, TH, ND, TIME>::FPW;
  __shared__ double sR[FPW * NT * 2 * V];
  const Basis &K = basis_dev<N>();
  const int nx = a.nx, ny = ND == 2 ? a.ny : 1;
  const long nfx = (long)(nx + 1) * ny;
  const long nfy = ND == 2 ? (long)nx * (ny + 1) : 0;
  const double dt = dev_or(a.st ? &a.st->dt : nullptr, a.dt);
  const int lane = threadIdx.x;
  const int fl = lane / NT, r = lane - fl * NT;
  const int t = r / NSF, sf = r - t * NSF;
  // Work order.  Faces are numbered row by row: row i's x-faces (between
  // cell rows i and i+1, ny of them), then its y-faces (ny + 1), for
  // i = 0..nx-1, then the last row of x-faces; every cell's qh is read by
  // its four faces within about two rows of work.  Blocks are dispatched to
  // the XCDs round-robin (block b on XCD b % 8), so block b takes item
  // (b % 8) * M + b / 8: each XCD works through one contiguous band of rows,
  // and the re-reads of a cell's qh hit that XCD's L2.
  const long nitems = (nfx + nfy + FPW - 1) / FPW;
  const long M = (nitems + 7) / 8;
  const long b = blockIdx.x;
  [[maybe_unused]] const LaneMem ews{EIG ? a.eig_ws + (b * WAVE + lane) : nullptr, (long)gridDim.x * WAVE};
  __shared__ long sOut[FPW];  // each face's output slot: fid in fx, or nfx + fid in fy
  // one group of FPW faces (item: its index in row order)
  auto group = [&](long item) {
    const bool item_ok = item < nitems;
    const long id0 = item * FPW;
    const long id = id0 + fl;  // the face in row order
    const bool ok = item_ok && fl < FPW && id < nfx + nfy;
    int d = 0;
    long e0 = 0, e1 = 0, fid = 0;  // fid: the face's slot in fx (d = 0) or fy (d = 1)
    if (ok) {
      if (ND == 1) {
        fid = id;
        e0 = id;
        e1 = id + 1;
      } else {
        const long per = 2L * ny + 1;  // faces of one row
        const long i = id / per, k = id - i * per;
        if (i == nx || k < ny) {  // x-face (i, j)
          const long j = k;
          fid = i * ny + j;
          e0 = i * (ny + 2) + (j + 1);
          e1 = (i + 1) * (ny + 2) + (j + 1);
        } else {  // y-face (i, j) of row i
          d = 1;
          const long j = k - ny;
          fid = i * (ny + 1) + j;
          e0 = (i + 1) * (ny + 2) + j;
          e1 = e0 + 1;
        }
      }
    }
    double lo[V], hi[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      lo[v] = 0.0;
      hi[v] = 0.0;
    }
    if (ok && a.mask[e0] && a.mask[e1]) {
      const double h = d == 0 ? a.dx : a.dy;
      const double kt = (TIME ? dt * K.w[t] : dt) / (2.0 * h);
      const double kk = ND == 2 ? kt * K.w[sf] : kt;
      double qL[V], qR[V];
      if constexpr (FS) {
        using F = FaceStates<N, ND, TIME>;
        const int sl = 2 * d;  // e0's high side is sl + 1, e1's low side sl
        const double *pl = a.fs + ((e0 * F::SIDES + sl + 1) * F::NPS + t * F::NSF + sf) * V;
        const double *pr = a.fs + ((e1 * F::SIDES + sl) * F::NPS + t * F::NSF + sf) * V;
#pragma unroll
        for (int v = 0; v < V; v++) {
          qL[v] = pl[v];
          qR[v] = pr[v];
        }
      } else {
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
      }
      if constexpr (EIG) {
        // D_ROE / D_OSH (fv.cpp:129-139,202-212), then Bint, weighted as Rusanov's
        double f[V], b[V];
        eig_flux<N, TH>(qL, qR, d, a.m, K, a.flux, f, ews);
        bint<N, TH>(qL, qR, d, K, b);
#pragma unroll
        for (int v = 0; v < V; v++) {
          lo[v] += kk * (b[v] + f[v]);
          hi[v] += kk * (b[v] - f[v]);
        }
      } else {
        // the direction as a compile-time constant in each branch, so that the
        // physics indexes its small arrays statically (no stack frame)
        if (d == 0)
          riemann_accum<N, TH, double, MK>(qL, qR, 0, a.m, K, kk, lo, hi);
        else
          riemann_accum<N, TH, double, MK>(qL, qR, 1, a.m, K, kk, lo, hi);
      }
    }
    if (fl < FPW) {
      double *my = sR + (fl * NT + r) * 2 * V;
#pragma unroll
      for (int v = 0; v < V; v++) {
        my[v] = lo[v];
        my[V + v] = hi[v];
      }
      if (r == 0) sOut[fl] = ok ? (d == 0 ? fid : nfx + fid) : -1;
    }
    __syncthreads();
    // one lane per (face, output): sum the face's NT lanes in a fixed order
    for (int o = lane; o < FPW * 2 * V; o += WAVE) {
      const int f = o / (2 * V), k = o - f * 2 * V;
      const long slot = sOut[f];
      if (slot < 0) continue;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < NT; q++) acc += sR[(f * NT + q) * 2 * V + k];
      double *out = slot < nfx ? a.fx + slot * 2 * V : a.fy + (slot - nfx) * 2 * V;
      out[k] = acc;
    }
  };
  if constexpr (EIG) {
    // block b takes groups b, b + grid, ... (a bounded grid: its workspace)
    for (long item = b; item < nitems; item += gridDim.x) {
      group(item);
      __syncthreads();  // sR and sOut are rewritten for the next group
    }
  } else {
    group((b % 8) * M + b / 8);  // the XCD-banded group of this block
  }
}

// Cell centres (fv.cpp:11-109) with the face terms: one lane per space-time
// node of a real cell, as in the predictor.  Each lane forms
// dt W_t W_a W_b (S(q) - B_x dq/dx - B_y dq/dy) at its node; the cell's
// lanes are summed through LDS and the face jumps subtracted.  TIME = false:
// one time level, weight dt; SRC = false: no S(q) (fv_launcher's SOURCES).
template <int N, int ND, bool TIME> struct CellShape {
  static constexpr int NS = ND == 1 ? N : N * N;  // spatial nodes per cell
  static constexpr int NN = NS * (TIME ? N : 1);  // nodes per cell
  static constexpr int CPW = WAVE / NN;           // cells per wavefront
};

template <int N, bool TH, int ND, bool TIME = true, bool SRC = true>
__global__ __launch_bounds__(64) void k_fv_cells(FvArgs a) {
  constexpr int V = Dims<TH>::V;
  constexpr int NS = CellShape<N, ND, TIME>::NS;
  constexpr int NN = CellShape<N, ND, TIME>::NN;
  constexpr int CPW = CellShape<N, ND, TIME>::CPW;
  __shared__ double sQ[CPW * NN * V];
  const Basis &K = basis_dev<N>();
  const int lane = threadIdx.x;
  const int slot = lane / NN;
  const int node = lane - slot * NN;
  const int t = node / NS;
  const int s = node - t * NS;
  const int ia = ND == 1 ? s : s / N;
  const int ib = ND == 1 ? 0 : s - (s / N) * N;
  const int nx = a.nx, ny = ND == 2 ? a.ny : 1;
  const long ncell = (long)nx * ny;
  const long c = (long)blockIdx.x * CPW + slot;
  const bool lane_ok = slot < CPW;
  const bool valid = lane_ok && c < ncell;
  const int ci = valid ? (ND == 2 ? c / ny : c) : 0, cj = valid ? (ND == 2 ? c % ny : 0) : 0;
  const long e = ND == 2 ? (long)(ci + 1) * (ny + 2) + (cj + 1) : ci + 1;
  const bool active = valid && a.mask[e];
  const double dt = dev_or(a.st ? &a.st->dt : nullptr, a.dt);
  const int sl = lane_ok ? slot : 0;
  double *myQ = sQ + sl * NN * V;

  double q[V];
#pragma unroll
  for (int v = 0; v < V; v++) q[v] = active ? a.qh[(e * NN + node) * V + v] : (v == 0 ? 1.0 : 0.0);
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = q[v];
  __syncthreads();
  double S[V];
  {
    cell_integrand<N, TH, ND, TIME, SRC>(q, myQ, t, ia, ib, a.m, K, dt, a.dx, a.dy, S);
    const bool centers = a.parts & 1;
#pragma unroll
    for (int v = 0; v < V; v++) S[v] = active && centers ? S[v] : 0.0;
  }
  __syncthreads();
  if (lane_ok)
#pragma unroll
    for (int v = 0; v < V; v++) myQ[node * V + v] = S[v];
  __syncthreads();
  // one lane per variable (a cell has fewer nodes than variables in 1D)
  for (int v = node; valid && v < V; v += NN) {
    double du = 0;
    for (int k = 0; k < NN; k++) du += myQ[k * V + v];
    if (!(a.parts & 2)) {
      a.u[c * V + v] += du;
      continue;
    }
    const double *fl = a.fx + ((long)ci * ny + cj) * 2 * V;        // left face: hi side
    const double *fr = a.fx + ((long)(ci + 1) * ny + cj) * 2 * V;  // right face: lo side
    du -= fl[V + v] + fr[v];
    if (ND == 2) {
      const double *fb = a.fy + ((long)ci * (ny + 1) + cj) * 2 * V;
      const double *ft = a.fy + ((long)ci * (ny + 1) + cj + 1) * 2 * V;
      du -= fb[V + v] + ft[v];
    }
    a.u[c * V + v] += du;
  }
}

// Fused form's cell update: u += cint - (face terms), one thread per (cell,
// variable), in the order k_fv_cells adds them.
template <bool TH, int ND>
__global__ __launch_bounds__(256) void k_fv_update(FvArgs a) {
  constexpr int V = Dims<TH>::V;
  const int nx = a.nx, ny = ND == 2 ? a.ny : 1;
  const long total = (long)nx * ny * V;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const long c = id / V;
    const int v = (int)(id - c * V);
    const long ci = ND == 2 ? c / ny : c, cj = ND == 2 ? c - ci * ny : 0;
    double du = a.cint[id];
    const double *fl = a.fx + (ci * ny + cj) * 2 * V;        // left face: hi side
    const double *fr = a.fx + ((ci + 1) * ny + cj) * 2 * V;  // right face: lo side
    du -= fl[V + v] + fr[v];
    if (ND == 2) {
      const double *fb = a.fy + (ci * (ny + 1) + cj) * 2 * V;
      const double *ft = a.fy + (ci * (ny + 1) + cj + 1) * 2 * V;
      du -= fb[V + v] + ft[v];
    }
    a.u[id] += du;
  }
}

// Post-step work on the updated u, one thread per cell: renorm_distortion
// (utils.cpp:6-23), the contortion flag (utils.cpp:65-85), a non-finite flag
// (iterator.cpp:161) and the CFL maximum for the next dt (iterator.cpp:58-72).
// ODE: the split step's closing half-step source ODE first
// (steppers.cpp:65), which the iterator's post-step then follows.
template <bool TH, int ND, bool ODE = false>
__global__ __launch_bounds__(256) void k_post(double *u, long ncell, double dx, double dy, Mat m, StepState *st) {
  constexpr int V = Dims<TH>::V;
  double local_max = 0.0;
  int contorted = 0, nonfinite = 0;
  const double hdt = ODE ? st->dt / 2 : 0.0;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = u[c * V + v];
    if (ODE) ode_analytic<TH>(q, hdt, m.p);
    if (q[0] > 0) {
      double cf = cbrt(q[0] / (m.p.rho0 * det3(q + 5)));
#pragma unroll
      for (int k = 0; k < 9; k++) q[5 + k] *= cf;
    }
    double f = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) f += q[5 + k] * q[5 + k];
    double d13 = cbrt(fabs(det3(q + 5)));
    if (f / (d13 * d13) / 3.0 - 1.0 > st->contorted_tol) contorted = 1;
#pragma unroll
    for (int v = 0; v < V; v++)
      if (!(q[v] - q[v] == 0.0)) nonfinite = 1;
    const double sp = cfl_speed<TH, ND>(q, m, dx, dy);
    if (sp > local_max) local_max = sp;
    if (ODE) {
#pragma unroll
      for (int v = 0; v < V; v++) u[c * V + v] = q[v];
    } else {
#pragma unroll
      for (int k = 0; k < 9; k++) u[c * V + 5 + k] = q[5 + k];
    }
  }
  double mx = wave_max(local_max);
  int cf = __any(contorted), nf = __any(nonfinite);
  if ((threadIdx.x & 63) == 0) {
    atomic_max_pos(&st->maxspeed, mx);
    if (cf) atomicOr(&st->contorted, 1);
    if (nf) atomicOr(&st->nonfinite, 1);
  }
}

// CFL maximum of u (timestep, iterator.cpp:58-72) into *dst.
template <bool TH, int ND>
__global__ __launch_bounds__(256) void k_cfl(const double *u, long ncell, double dx, double dy, Mat m,
                                            unsigned long long *dst) {
  constexpr int V = Dims<TH>::V;
  double local_max = 0.0;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = u[c * V + v];
    const double sp = cfl_speed<TH, ND>(q, m, dx, dy);
    if (sp > local_max) local_max = sp;
  }
  double mx = wave_max(local_max);
  if ((threadIdx.x & 63) == 0) atomic_max_pos(dst, mx);
}

// reset_distortion on every cell when the contortion flag is set (utils.cpp:44-63,
// iterator.cpp:150-151); the CFL maximum is then recomputed into maxspeed2.
template <bool TH, int ND>
__global__ __launch_bounds__(256) void k_reset(double *u, long ncell, double dx, double dy, Mat m, StepState *st) {
  constexpr int V = Dims<TH>::V;
  if (!st->contorted) return;
  double local_max = 0.0;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    double q[V];
#pragma unroll
    for (int v = 0; v < V; v++) q[v] = u[c * V + v];
    double r = q[0];
    double cf = cbrt(r / m.p.rho0);
    NodeState<double, TH> s;
    node_state<double, TH>(q, m, s);
    double Gm = Gamma_MG(r, m.p);
    double C0 = (1 - m.p.beta / Gm) * s.cs2;
    q[1] -= r * (C0 / 4 * s.devGsq);
#pragma unroll
    for (int k = 0; k < 9; k++) q[5 + k] = 0.0;
    q[5] = cf;
    q[9] = cf;
    q[13] = cf;
#pragma unroll
    for (int v = 0; v < V; v++) u[c * V + v] = q[v];
    const double sp = cfl_speed<TH, ND>(q, m, dx, dy);
    if (sp > local_max) local_max = sp;
  }
  double mx = wave_max(local_max);
  if ((threadIdx.x & 63) == 0) atomic_max_pos(&st->maxspeed2, mx);
}

// dt for the coming step from the CFL maximum (timestep, iterator.cpp:74-82).
// first: the maximum was computed by k_cfl into maxspeed; otherwise by the
// previous step's post work (maxspeed, or maxspeed2 after a reset).
static __global__ void k_finalize_dt(StepState *st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long bits = st->contorted ? st->maxspeed2 : st->maxspeed;
  double MAX = __longlong_as_double((long long)bits);
  if (st->contorted) st->resets += 1;
  double dt = st->cfl / MAX;
  if (st->count <= 5) dt *= 0.2;
  if (st->t + dt > st->tf) dt = st->tf - st->t;
  st->dt = dt;
  st->maxspeed = 0ull;
  st->maxspeed2 = 0ull;
  st->contorted = 0;
  st->done = st->t >= st->tf ? 1 : 0;
}

// *p = 0 as a kernel (a node that graph capture records reliably)
static __global__ void k_zero_int(int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// t += dt, count += 1 (iterator.cpp:142-143)
static __global__ void k_advance(StepState *st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->t += st->dt;
  st->count += 1;
}

}  // namespace gpr
