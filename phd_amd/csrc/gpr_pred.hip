// Launches of the space-time predictor (gpr_predictor.hpp) for N =
// GPR_UNIT_N; phd_amd/build.py compiles this unit once per N.
#include <hip/hip_runtime.h>

#include "gpr_units.hpp"

#ifndef GPR_UNIT_N
#error "GPR_UNIT_N (2 or 3) must be defined"
#endif

namespace gpr {

// Picard in R over a.mask (wh elements WIn), appending the cells it leaves
// to a.list; then the FP64 preconditioned fallback over the list, CPW listed
// cells per wavefront.  R = double gives the one-launch form's values;
// R = float is the FP32-bulk form.
template <int N, bool TH, int ND, typename R, typename WIn, int MK = MK_ANY>
void pred_split(hipStream_t s, const PredArgs &a) {
  constexpr int CPW = PredShape<N, ND>::CPW;
  const long blocks = (a.ncell + CPW - 1) / CPW;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, a.list_count);
  hipLaunchKernelGGL((k_predictor<N, TH, ND, R, WIn, false, KIND_PICARD, MK>), dim3(blocks), dim3(WAVE), 0, s, a);
  hipLaunchKernelGGL((k_predictor<N, TH, ND, double, WIn, false, KIND_FALLBACK, MK>), dim3(blocks), dim3(WAVE), 0,
                     s, a);
}

template <int N, bool TH, int ND, int MK> void pred_launch_mk(hipStream_t s, const PredArgs &a, int how) {
  constexpr int CPW = PredShape<N, ND>::CPW;
  if (how == PRED_SPLIT_FP32)
    pred_split<N, TH, ND, float, float, MK>(s, a);
  else if (how == PRED_SPLIT_FP64)
    pred_split<N, TH, ND, double, double>(s, a);
  else
    hipLaunchKernelGGL((k_predictor<N, TH, ND, double, double, false, KIND_BOTH, MK>),
                       dim3((a.ncell + CPW - 1) / CPW), dim3(WAVE), 0, s, a);
}

// The 2D P2 single-material workloads (no heat conduction) get kernels
// compiled for their material kind: the one-launch predictor and the FP32
// pair.  Every other case, and the two-launch FP64 form, uses MK_ANY; the
// values do not depend on the choice.
template <int N, bool TH, int ND> hipError_t pred_launch(hipStream_t s, const PredArgs &a, int how) {
  // the fused FV epilogue exists in the one-launch kernel of one material kind only
  if (a.fs && !(N == 3 && ND == 2 && !TH && how == PRED_ONE_LAUNCH && material_kind(a.m) == MK_GAS_NEWTONIAN))
    return hipErrorInvalidValue;
  if constexpr (N == 3 && ND == 2 && !TH) {
    switch (material_kind(a.m)) {
    case MK_GAS_NEWTONIAN: pred_launch_mk<N, TH, ND, MK_GAS_NEWTONIAN>(s, a, how); break;
    case MK_GR_PLASTIC: pred_launch_mk<N, TH, ND, MK_GR_PLASTIC>(s, a, how); break;
    default: pred_launch_mk<N, TH, ND, MK_ANY>(s, a, how);
    }
  } else {
    pred_launch_mk<N, TH, ND, MK_ANY>(s, a, how);
  }
  return hipGetLastError();
}

template <int N> hipError_t pred_upload(const Basis *hb) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_basis), hb, 2 * sizeof(Basis));
}

constexpr int UN = GPR_UNIT_N;
template hipError_t pred_launch<UN, false, 1>(hipStream_t, const PredArgs &, int);
template hipError_t pred_launch<UN, true, 1>(hipStream_t, const PredArgs &, int);
template hipError_t pred_launch<UN, false, 2>(hipStream_t, const PredArgs &, int);
template hipError_t pred_launch<UN, true, 2>(hipStream_t, const PredArgs &, int);
template hipError_t pred_upload<UN>(const Basis *);

}  // namespace gpr
