// Directed FP64 -> FP32 rounding for device code, shared by the traversal kernels and the device builder.
// Results equal the host's (float)x corrected by std::nextafter toward -inf / +inf.
#pragma once
#include <hip/hip_runtime.h>

namespace brt {

// Adjacent float toward -inf / +inf (finite or infinite input, never NaN here).
__device__ __forceinline__ float float_below(float f) {
  if (f == 0.0f) return __int_as_float((int)0x80000001);
  const int b = __float_as_int(f);
  return __int_as_float(f > 0.0f ? b - 1 : b + 1);
}
__device__ __forceinline__ float float_above(float f) {
  if (f == 0.0f) return __int_as_float(1);
  const int b = __float_as_int(f);
  return __int_as_float(f > 0.0f ? b + 1 : b - 1);
}
// FP64 -> FP32 rounded toward -inf / +inf.
__device__ __forceinline__ float f32_down(double x) {
  const float f = (float)x;
  return ((double)f > x) ? float_below(f) : f;
}
__device__ __forceinline__ float f32_up(double x) {
  const float f = (float)x;
  return ((double)f < x) ? float_above(f) : f;
}

}  // namespace brt
