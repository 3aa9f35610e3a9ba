// Launchers of the kernels compiled in translation units of their own, so
// that the library's units build in parallel (phd_amd/build.py):
//   gpr_pred.hip, once per N          the space-time predictor (K2)
//   gpr_fv.hip, once per (N, dims)    the FV faces and cell update (K3)
// Each unit holds its own copy of the constant basis table (c_basis), which
// units_upload_basis writes.
#pragma once

#include <hip/hip_runtime.h>

#include "gpr_kernels.hpp"

namespace gpr {

enum {
  PRED_ONE_LAUNCH = 0,  // Picard and the fallback in one launch (KIND_BOTH)
  PRED_SPLIT_FP64 = 1,  // Picard, then the fallback over the listed cells
  PRED_SPLIT_FP32 = 2,  // the same with Picard in FP32 on FP32 wh
};

template <int N, bool TH, int ND> hipError_t pred_launch(hipStream_t s, const PredArgs &a, int how);
template <int N> hipError_t pred_upload(const Basis *hb);

// fv_launcher's face terms (a.parts & 2) and cell update for a.time / a.sources
template <int N, bool TH, int ND> hipError_t fv_launch(hipStream_t s, const FvArgs &a);
template <int N, int ND> hipError_t fv_upload(const Basis *hb);

inline hipError_t units_upload_basis(const Basis *hb) {
  hipError_t e = pred_upload<2>(hb);
  if (e == hipSuccess) e = pred_upload<3>(hb);
  if (e == hipSuccess) e = fv_upload<2, 1>(hb);
  if (e == hipSuccess) e = fv_upload<2, 2>(hb);
  if (e == hipSuccess) e = fv_upload<3, 1>(hb);
  if (e == hipSuccess) e = fv_upload<3, 2>(hb);
  return e;
}

}  // namespace gpr
