// The adjoint's transposed Riemann faces, cell terms and predictor solve
// (gpr_adjoint.hpp) for one (dual-number local derivatives, heat conduction)
// pair, GPR_ADJ_DUAL and GPR_ADJ_TH.  phd_amd/build.py compiles this unit
// once per pair so that the four, the slowest kernels of the library to
// compile, build in parallel.
#include <hip/hip_runtime.h>

#include "gpr_adjoint.hpp"

#ifndef GPR_ADJ_DUAL
#error "GPR_ADJ_DUAL (0 or 1) must be defined"
#endif
#ifndef GPR_ADJ_TH
#error "GPR_ADJ_TH (0 or 1) must be defined"
#endif

namespace gpr {

template <bool DUAL, bool TH>
hipError_t adj_heavy_launch(dim3 gf, dim3 gc, hipStream_t s, const AdjArgs &a, long ne) {
  hipLaunchKernelGGL((ka_faces<3, 2, DUAL, TH>), gf, dim3(WAVE), 0, s, a);
  hipLaunchKernelGGL((ka_cells<3, 2, DUAL, TH>), gc, dim3(WAVE), 0, s, a, ne);
  hipLaunchKernelGGL((ka_predictor<3, 2, DUAL, TH>), gc, dim3(WAVE), 0, s, a, ne);
  return hipGetLastError();
}

template <bool DUAL, bool TH> hipError_t adj_heavy_upload(const Basis *hb) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_basis), hb, 2 * sizeof(Basis));
}

template hipError_t adj_heavy_launch<GPR_ADJ_DUAL != 0, GPR_ADJ_TH != 0>(dim3, dim3, hipStream_t, const AdjArgs &,
                                                                        long);
template hipError_t adj_heavy_upload<GPR_ADJ_DUAL != 0, GPR_ADJ_TH != 0>(const Basis *);

}  // namespace gpr
