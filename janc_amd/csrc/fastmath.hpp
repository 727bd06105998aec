// Reciprocals for the hot loops.  IEEE division compiles to a ~10-instruction
// scale/rcp/Newton/fixup sequence on gfx950; where the divisor is a
// well-scaled positive number (densities, WENO weight sums, cp - R, dx) the
// hardware reciprocal refined by two Newton steps (FP64, <= ~1 ulp) or taken
// as is (FP32, 1 ulp) gives the same result to rounding at a third of the
// cost.  Host builds (used to check these headers against the oracle) divide.
#pragma once
#include <math.h>

#include "janc_types.hpp"

namespace jh {

__host__ __device__ __forceinline__ double rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / b;
#endif
}

__host__ __device__ __forceinline__ float rcp(float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(b);
#else
  return 1.0f / b;
#endif
}

}  // namespace jh
