// Bit-exact device arithmetic of blazeRT's ray setup and node test, in the scene type R (double or float).
//
// Shared by libblazert_amd's traversal kernels and by user device code that traverses a brt_bvh tree
// (include/blazert_amd/user_geometry.h). Every function evaluates the reference's expressions in the reference's
// order with no fused multiply-add (each body disables contraction, so the results do not depend on the
// including translation unit's -ffp-contract setting).
//
// HIP device code; include after <hip/hip_runtime.h>.
#pragma once
#include <float.h>
#include <stdint.h>

#include <limits>

namespace brt {
namespace exact {

// Constants of the scene type T (blazert::Scene<T>): double for the drop-in path, float for Scene<float>.
template <class R>
struct Real;
template <>
struct Real<double> {
  static constexpr double eps = DBL_EPSILON;
  static constexpr double max = DBL_MAX;
};
template <>
struct Real<float> {
  static constexpr float eps = FLT_EPSILON;
  static constexpr float max = FLT_MAX;
};

// std::min / std::max argument order (the NaN behaviour the reference's slab test inherits)
template <class R>
__device__ __forceinline__ R std_min(R a, R b) { return (b < a) ? b : a; }
template <class R>
__device__ __forceinline__ R std_max(R a, R b) { return (a < b) ? b : a; }

// blazert/datatypes.h:83-85
__device__ __forceinline__ double product_sign(double x, double y) {
  const unsigned long long ux = (unsigned long long)__double_as_longlong(x);
  const unsigned long long uy = (unsigned long long)__double_as_longlong(y);
  return __longlong_as_double((long long)(ux ^ (uy & 0x8000000000000000ull)));
}
__device__ __forceinline__ float product_sign(float x, float y) {
  return __uint_as_float(__float_as_uint(x) ^ (__float_as_uint(y) & 0x80000000u));
}

// correctly rounded in both precisions
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }

template <class R>
struct RayT {
  R o[3];
  R d[3];    // normalize(raw direction)
  R inv[3];  // 1 / raw direction
  uint32_t sign;  // bit i: raw direction[i] < 0; bit 3: no slab term can be NaN (see setup_ray)
  R tmin, tmax;
};
constexpr uint32_t RAY_NAN_FREE = 8u;

// blazert/ray.h:56-65 with normalize(v) = v * (len != 0 ? 1/len : 0)
template <class R>
__device__ __forceinline__ void setup_ray(const R org[3], const R raw[3], R tmin, R tmax, RayT<R> &r) {
#pragma clang fp contract(off)
  r.o[0] = org[0];
  r.o[1] = org[1];
  r.o[2] = org[2];
  const R len = dsqrt(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2]);
  const R il = (len != R(0)) ? R(1) / len : R(0);
  r.sign = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r.d[k] = raw[k] * il;
    r.inv[k] = R(1) / raw[k];
    r.sign |= (raw[k] < R(0) ? 1u : 0u) << k;
  }
  r.tmin = tmin;
  r.tmax = tmax;
  // With a finite origin, finite non-zero direction components (so 1/d is finite and non-zero), a non-NaN tmin and
  // node boxes free of NaN (the caller clears the flag otherwise, also for boxes with min > max), (b - o) * inv is
  // never NaN: the slab chains reduce to plain max/min.
  bool ok = r.tmin == r.tmin;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ok = ok && fabs(r.o[k]) <= Real<R>::max && raw[k] != R(0) && fabs(raw[k]) <= Real<R>::max &&
         fabs(r.inv[k]) <= Real<R>::max;
  }
  if (ok) r.sign |= RAY_NAN_FREE;
}

// max / min of two values known not to be NaN. fmax / fmin in IEEE mode first quiet every operand the compiler
// cannot prove canonical, which it cannot for values computed in an earlier basic block: one extra v_max per
// operand. On non-NaN operands the bare instructions return what fmax / fmin return.
__device__ __forceinline__ double max_nn(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double min_nn(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float max_nn(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float min_nn(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Exact slab test of one child box (blazert/bvh/intersect.h:32-51 in the scene type), split into its hit-distance
// independent part. Returns false when the box fails for every hit distance; otherwise key is the value the current
// hit distance h must not be below (pass iff !(h < key)); key = -inf when a NaN far-plane term removes h from the
// reference's min chain. b = {min x, y, z, max x, y, z}.
//
// RAY_NAN_FREE requires, besides the ray's own conditions, a tree whose boxes all have min <= max and no NaN (the
// caller clears the bit otherwise). Then the reference's near-plane distance (b_near - o) * inv, with the plane
// chosen by the sign of the direction, is the smaller of the two planes' distances and the far one the larger:
// rounding is monotone, so for inv > 0 (b_min - o) * inv <= (b_max - o) * inv, and the order flips for inv < 0.
// min / max of the two computed distances thus pick exactly the reference's values (a zero's sign aside, which no
// comparison sees), without selecting the planes by sign first.
template <class R>
__device__ __forceinline__ bool slab_exact(const R *b, const RayT<R> &r, R &key) {
#pragma clang fp contract(off)
  const R L1 = R(1) + R(4) * Real<R>::eps;  // far-plane widening of blazert/bvh/intersect.h:40
  // both planes' distances; the reference's near / far terms are these values, chosen by the direction's sign
  R t0[3], t1[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    t0[i] = (b[i] - r.o[i]) * r.inv[i];
    t1[i] = (b[3 + i] - r.o[i]) * r.inv[i];
  }
  if (r.sign & RAY_NAN_FREE) {
    R A = max_nn(r.tmin, min_nn(t0[0], t1[0]));
    R M = max_nn(t0[0], t1[0]) * L1;  // the reference's min chain starts from +inf
#pragma unroll
    for (int i = 1; i < 3; i++) {
      A = max_nn(A, min_nn(t0[i], t1[i]));
      M = min_nn(M, max_nn(t0[i], t1[i]) * L1);
    }
    key = A;
    return A <= M;
  }
  R acc = r.tmin;
  R m[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool s = (r.sign >> i) & 1u;
    const R mn = s ? t1[i] : t0[i];
    m[i] = (s ? t0[i] : t1[i]) * L1;
    acc = std_max(mn, acc);
  }
  const bool n0 = m[0] != m[0], n1 = m[1] != m[1], n2 = m[2] != m[2];
  R M;
  bool hdep = false;
  if (n2) {
    return false;
  } else if (n1) {
    M = m[2];
  } else if (n0) {
    M = std_min(m[2], m[1]);
  } else {
    M = fmin(fmin(m[0], m[1]), m[2]);
    hdep = true;
  }
  if (!(acc <= M)) return false;
  key = hdep ? acc : -std::numeric_limits<R>::infinity();
  return true;
}

}  // namespace exact
}  // namespace brt
