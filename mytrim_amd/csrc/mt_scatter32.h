// The scattering step of the single-precision cascade paths (trim.C:172-272
// evaluated in FP32): on MT_FP32_SURROGATE the learned map answers inside its
// trained domain for eps <= 10 (above 10 the Rutherford closed form is cheaper
// than the network), everywhere else the FP32 solve runs.  Shared by the
// cascade kernel and the per-function batch kernel, so a batch check tests
// the code the cascades run.
#pragma once
#include "mt_mlp.h"
#include "mt_physics.h"

namespace mt {

// iters: Newton iterations of the solve, 0 for the Rutherford branch, -1 where
// the surrogate answered
template <int ARITH, int POT>
__device__ __forceinline__ void scatter_single(const MlpDev& mlp, float eps, float b, float leps, float& s2,
                                               float& ct, float& st, int& iters) {
  float s2f = -1.0f;
  if constexpr (ARITH == 2) {
    if (eps <= 10.0f) s2f = mlp_s2(mlp, eps, b);
  }
  if (s2f >= 0.0f) {
    s2 = s2f;
    ct = 1.0f - 2.0f * s2f;
    st = 2.0f * m_sqrt(fmaxf(0.0f, s2f * (1.0f - s2f)));
    iters = -1;
  } else {
    scatter_f32<POT>(eps, b, leps, s2, ct, st, iters);
  }
}

}  // namespace mt
