// Launches of the FV corrector's face and cell kernels (gpr_kernels.hpp) for
// N = GPR_UNIT_N in GPR_UNIT_ND dimensions; phd_amd/build.py compiles this
// unit once per pair.
#include <hip/hip_runtime.h>

#include "gpr_units.hpp"

#ifndef GPR_UNIT_N
#error "GPR_UNIT_N (2 or 3) must be defined"
#endif
#ifndef GPR_UNIT_ND
#error "GPR_UNIT_ND (1 or 2) must be defined"
#endif

namespace gpr {

// fv_launcher's faces and cells for one (TIME, SOURCES) pair, Rusanov's
// physics compiled for material kind MK (the cell kernel measured no faster
// compiled per kind, and stays general)
template <int N, bool TH, int ND, bool TIME, bool SRC, int MK = MK_ANY> void fv_pair(hipStream_t s, const FvArgs &a) {
  const long nf = (long)(a.nx + 1) * (ND == 2 ? a.ny : 1) + (ND == 2 ? (long)a.nx * (a.ny + 1) : 0);
  constexpr int FPW = FaceShape<N, TH, ND, TIME>::FPW;
  // k_fv_faces maps block b to work item (b % 8) * M + b / 8 with
  // M = ceil(items / 8): the grid is 8 M blocks, those past the items idle
  const long items = (nf + FPW - 1) / FPW;
  if ((a.parts & 2) && a.flux != 0)
    hipLaunchKernelGGL((k_fv_faces<N, TH, ND, TIME, false, true>), dim3(eig_face_grid(items)), dim3(WAVE), 0, s,
                       a);
  else if (a.parts & 2)
    hipLaunchKernelGGL((k_fv_faces<N, TH, ND, TIME, false, false, MK>), dim3(8 * ((items + 7) / 8)), dim3(WAVE), 0,
                       s, a);
  const long nc = (long)a.nx * (ND == 2 ? a.ny : 1);
  constexpr int CPW = CellShape<N, ND, TIME>::CPW;
  hipLaunchKernelGGL((k_fv_cells<N, TH, ND, TIME, SRC>), dim3((nc + CPW - 1) / CPW), dim3(WAVE), 0, s, a);
}

// the fused form: faces from the predictor's face states, then the cell update
template <int N, bool TH, int ND> void fv_fused(hipStream_t s, const FvArgs &a) {
  const long nf = (long)(a.nx + 1) * (ND == 2 ? a.ny : 1) + (ND == 2 ? (long)a.nx * (a.ny + 1) : 0);
  constexpr int FPW = FaceShape<N, TH, ND, true>::FPW;
  const long items = (nf + FPW - 1) / FPW;
  // the predictor writes face states for the 2D ideal-gas Newtonian kind only
  hipLaunchKernelGGL((k_fv_faces<N, TH, ND, true, true, false, MK_GAS_NEWTONIAN>), dim3(8 * ((items + 7) / 8)),
                     dim3(WAVE), 0, s, a);
  const long tot = (long)a.nx * (ND == 2 ? a.ny : 1) * Dims<TH>::V;
  hipLaunchKernelGGL((k_fv_update<TH, ND>), dim3((tot + 255) / 256), dim3(256), 0, s, a);
}

template <int N, bool TH, int ND> hipError_t fv_launch(hipStream_t s, const FvArgs &a) {
  if (a.fs) {
    if (!(N == 3 && ND == 2 && !TH) || material_kind(a.m) != MK_GAS_NEWTONIAN) return hipErrorInvalidValue;
    if constexpr (N == 3 && ND == 2 && !TH) fv_fused<N, TH, ND>(s, a);
  } else if (a.time && a.sources) {
    // the ADER step's corrector: the face kernel compiled for the material
    // kind on the 2D P2 single-material workloads without heat conduction
    int mk = MK_ANY;
    if constexpr (N == 3 && ND == 2 && !TH)
      if (a.flux == 0) mk = material_kind(a.m);
    if constexpr (N == 3 && ND == 2 && !TH) {
      if (mk == MK_GAS_NEWTONIAN)
        fv_pair<N, TH, ND, true, true, MK_GAS_NEWTONIAN>(s, a);
      else if (mk == MK_GR_PLASTIC)
        fv_pair<N, TH, ND, true, true, MK_GR_PLASTIC>(s, a);
      else
        fv_pair<N, TH, ND, true, true>(s, a);
    } else {
      fv_pair<N, TH, ND, true, true>(s, a);
    }
  } else if (!a.time && !a.sources)
    fv_pair<N, TH, ND, false, false>(s, a);
  else if (a.time)
    fv_pair<N, TH, ND, true, false>(s, a);
  else
    fv_pair<N, TH, ND, false, true>(s, a);
  return hipGetLastError();
}

template <int N, int ND> hipError_t fv_upload(const Basis *hb) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_basis), hb, 2 * sizeof(Basis));
}

template hipError_t fv_launch<GPR_UNIT_N, false, GPR_UNIT_ND>(hipStream_t, const FvArgs &);
template hipError_t fv_launch<GPR_UNIT_N, true, GPR_UNIT_ND>(hipStream_t, const FvArgs &);
template hipError_t fv_upload<GPR_UNIT_N, GPR_UNIT_ND>(const Basis *);

}  // namespace gpr
