// HIP kernels for gfx950: closest-hit traversal of one geometry-type BVH per launch.
//
// FP64 path (the drop-in): bit-identical to blazeRT's traverse() (blazert/bvh/accel.h:79-120). The reference
// pops a node, slab-tests it against [tmin, current hit] and pushes far then near. Here an internal record holds
// both child boxes: the near child is tested at the parent (the reference tests it at the very next pop, with the
// same hit distance), and the far child goes on the stack with the h-independent half of its slab test already
// applied and its entry distance as the key, so the pop-time check `!(h < key)` reproduces the reference's
// decision exactly, NaN propagation of std::max/std::min included (blazert/bvh/intersect.h:32-51).
//
// FP32 path (split precision): the same DFS with slab tests on outward-rounded FP32 child boxes and a per-ray
// absolute error bound, so it visits a superset of the FP64 path's nodes; primitive tests stay FP64.
//
// All arithmetic is compiled with -ffp-contract=off so that products and sums round exactly as the reference's.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "../../include/blazert_amd/device/exact.h"
#include "engine.h"
#include "f32_round.h"

namespace brt {

namespace {

constexpr uint32_t PREC_FP64 = 0;
constexpr uint32_t FLAG_INOUT = 1u;
constexpr uint32_t FLAG_ANY_HIT = 2u;
// Internal: an FP64 launch that traces only the rays a split-precision launch of the same geometry type deferred
// (see deferred_to_exact), and none at all when that launch deferred nothing.
constexpr uint32_t FLAG_DEFERRED_ONLY = 1u << 31;
// Exact paths: primitives of a leaf tested per traversal step, per primitive type (triangles, spheres, planes,
// cylinders). They are tested in leaf order, each against the hit distance the previous one left, so the count
// changes speed only.
#ifndef BRT_EXACT_GROUP_TRI
#define BRT_EXACT_GROUP_TRI 2  // 1.31M-triangle incoherent set: 2-3% faster than 1, bit-identical hits
#endif
#ifndef BRT_EXACT_GROUP_SPH
#define BRT_EXACT_GROUP_SPH 1  // mixed scene: 2 and 4 per step 4% and 15% slower (the FP64 sphere test is long)
#endif
#ifndef BRT_EXACT_GROUP_PLA
#define BRT_EXACT_GROUP_PLA 1  // mixed scene: 2 per step 13% slower
#endif
#ifndef BRT_EXACT_GROUP_CYL
#define BRT_EXACT_GROUP_CYL 1  // mixed scene: 2 per step 18% slower
#endif
template <uint32_t KIND>
constexpr int exact_group() {
  return KIND == TRIANGLES ? BRT_EXACT_GROUP_TRI
         : KIND == SPHERES ? BRT_EXACT_GROUP_SPH
         : KIND == PLANES  ? BRT_EXACT_GROUP_PLA
                           : BRT_EXACT_GROUP_CYL;
}

// Deepest stack a traversal can build: the binary FP64 DFS pushes at most one entry per level (trees are limited
// to depth 63 at upload); the 4-wide FP32 DFS at most three per Node4 record, and a record spans at least one
// binary level on every path, so at most 3 * 63 = 189.
template <bool SPLIT>
constexpr int stack_cap() { return SPLIT ? 192 : 64; }

// The reference's ray setup and exact slab test (Real, RayT, setup_ray, slab_exact, product_sign, dsqrt) are shared
// with user-collection traversal: include/blazert_amd/device/exact.h.
using namespace exact;
using Ray = RayT<double>;

// Rays the split-precision path hands to the exact FP64 path: those whose reference slab test can produce a NaN
// term or culls boxes its line crosses, i.e. every ray without RAY_NAN_FREE (a zero, non-finite or reciprocal-
// overflowing direction component, a non-finite origin, a NaN tmin). For a raw component -0.0 the reference takes
// the near plane from `direction < 0` (false) but multiplies by 1/-0.0 = -inf, so it culls every box the ray's line
// lies in; an origin on a box plane of a zero component gives 0 * inf = NaN. The FP32 path's |1/d| > 1e30 axes
// never cull, so on such rays it would reach primitives the reference cannot. Every other ray has NaN-free slab
// terms in the reference, and the FP32 path's conservative boxes visit a superset of its nodes.
template <class R>
__device__ __forceinline__ bool deferred_to_exact(const RayT<R> &r) {
  return !(r.sign & RAY_NAN_FREE);
}

// A pointer into device memory viewed in the global address space. Pointers loaded from memory (the batch kernel's
// job table) are generic, and accesses through generic pointers are flat accesses; through this view they are global.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Rays are read once and hits written once per launch: streamed with non-temporal hints, so that they do not evict
// the tree's lines from the caches (a launch of 4M rays streams ~335 MB, more than the last-level cache).
template <class T>
__device__ __forceinline__ T ld_stream(const T *p, uint64_t i) {
  return __builtin_nontemporal_load(gmem(p) + i);
}
template <class T, class V>
__device__ __forceinline__ void st_stream(T *p, uint64_t i, V v) {
  __builtin_nontemporal_store((T)v, gmem(p) + i);
}

template <class R>
__device__ __forceinline__ void make_ray(const RaysViewT<R> &V, uint64_t i, RayT<R> &r) {
  const R org[3] = {ld_stream(V.ox, i), ld_stream(V.oy, i), ld_stream(V.oz, i)};
  const R raw[3] = {ld_stream(V.dx, i), ld_stream(V.dy, i), ld_stream(V.dz, i)};
  setup_ray(org, raw, V.tmin ? ld_stream(V.tmin, i) : V.tmin0, V.tmax ? ld_stream(V.tmax, i) : V.tmax0, r);
}

// Conservative FP32 slab test: per axis lower/upper bounds of the FP64 entry/exit distances.
struct Ray32 {
  float inv[3];
  float offn[3];  // rounded-down (-o*inv - E)
  float offf[3];  // rounded-up   (-o*inv + E)
  float tmin;     // rounded down
  float om;       // max |o_i| rounded up; +inf when above FILTER_MAX_EXTENT (no FP32 triangle pre-test)
  uint32_t sign;
};

// E_i = 5 * 2^-24 * |inv_i| * (S + |o_i|) bounds |fma(b32, inv32, off32) - (b - o) inv| for every |b| <= S,
// including the FP64 reference's own rounding and its 1 + 4 eps far-plane factor.
__device__ __forceinline__ void make_ray32(const Ray &r, double S, Ray32 &q) {
  q.sign = r.sign;
  q.tmin = f32_down(r.tmin);
  const double om = fmax(fabs(r.o[0]), fmax(fabs(r.o[1]), fabs(r.o[2])));
  q.om = om <= FILTER_MAX_EXTENT ? f32_up(om) : __builtin_inff();
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double inv = r.inv[i];
    const double ainv = fabs(inv);
    if (!(ainv <= 1e30)) {  // zero / tiny direction component: this axis never culls
      q.inv[i] = 0.0f;
      q.offn[i] = -__builtin_inff();
      q.offf[i] = __builtin_inff();
      continue;
    }
    const double off = -(r.o[i] * inv);
    const double E = 5.0 * 5.9604644775390625e-08 * ainv * (S + fabs(r.o[i]));
    q.inv[i] = (float)inv;
    q.offn[i] = f32_down(off - E);
    q.offf[i] = f32_up(off + E);
  }
}

__device__ __forceinline__ bool slab32(const float *b, const Ray32 &q, float &key) {
  float A = q.tmin;
  float M = __builtin_inff();
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool s = (q.sign >> i) & 1u;
    const float nearp = s ? b[3 + i] : b[i];
    const float farp = s ? b[i] : b[3 + i];
    A = fmaxf(A, __builtin_fmaf(nearp, q.inv[i], q.offn[i]));
    M = fminf(M, __builtin_fmaf(farp, q.inv[i], q.offf[i]));
  }
  key = A;
  return A <= M;
}

// Conservative FP32 pre-test of one triangle for the split path: returns false only when the FP64 test isect_tri
// (blazert/primitives/trimesh.h:153-181) rejects the triangle for the ray at the current hit distance h <= h32.
//
// With u = 2^-24, Pm = vmax + om >= |v0_i| + |o_i|, sa = sum |e1_i|, sb = sum |e2_i|, the values computed here from
// v0, e1, e2, n and o rounded to nearest and the unit direction d rounded to nearest differ from the exact ones (and
// from the FP64 test's, whose own rounding is ~2^-29 smaller) by at most
//   u, v:  13.1 u Pm sb and 13.1 u Pm sa     (c = v0 - o: 2.01 u Pm per component; r = d x c: 5.1 u sqrt2 Pm)
//   det:    5.1 u sa sb                      (n rounded once from the FP64 test's own n)
//   t:      6.1 u sa sb Pm
//   w = |det| - u - v: the sum of those plus 2 u (|det| + |u| + |v|) <= 0.2 Ed + 0.18 (Eu + Ev)
// The bounds used (Eu = kub Pm, Ev = kua Pm, Ed = ed, Et = 1.2 ed Pm, Ew = 1.3 ed + 1.2 (kua + kub) Pm) keep a
// margin for the rounding of the bounds themselves; ed and Et hold an absolute 1e-29, Eu, Ev and Ew the reference's
// tolerance, for results in the FP32 denormal range. The reference accepts u, v, w >= -4 DBL_EPSILON (tol), t >= |det| tmin and |det| h > t with the signs
// taken from det; where the sign of det is not certain (|det| <= Ed) the triangle passes. Scenes and ray origins
// above FILTER_MAX_EXTENT are not pre-tested, so no value here overflows; every rejection is a comparison that is
// false for NaN operands.
__device__ __forceinline__ bool tri_filter(const TriFilt &T, const float D[3], const float O[3], float om,
                                           float tmin32, float h32) {
  const float u = 5.9604644775390625e-08f;
  const float tol = 9e-16f;  // 4 DBL_EPSILON rounded up
  // fused multiply-adds round once where a product and a sum round twice, so the bounds below hold for them too
  auto dot3 = [](float a0, float a1, float a2, float b0, float b1, float b2) {
    return __builtin_fmaf(a0, b0, __builtin_fmaf(a1, b1, a2 * b2));
  };
  const float det = dot3(T.n[0], T.n[1], T.n[2], D[0], D[1], D[2]);
  const float ad = fabsf(det);
  if (!(ad > T.ed)) return true;  // the sign of det is not certain
  const float c0 = T.v0[0] - O[0], c1 = T.v0[1] - O[1], c2 = T.v0[2] - O[2];
  const float rx = __builtin_fmaf(D[1], c2, -(D[2] * c1));
  const float ry = __builtin_fmaf(D[2], c0, -(D[0] * c2));
  const float rz = __builtin_fmaf(D[0], c1, -(D[1] * c0));
  const float Pm = T.vmax + om;
  const float *a = T.e1, *b = T.e2;
  const float us = product_sign(dot3(rx, ry, rz, b[0], b[1], b[2]), det);
  const float vs = product_sign(dot3(rx, ry, rz, a[0], a[1], a[2]), det);
  if (us < -__builtin_fmaf(T.kub, Pm, tol) || vs < -__builtin_fmaf(T.kua, Pm, tol)) return false;
  const float Ew = __builtin_fmaf(1.2f * (T.kua + T.kub), Pm, __builtin_fmaf(1.3f, T.ed, tol));
  if ((ad - us) - vs < -Ew) return false;
  const float ts = product_sign(dot3(T.n[0], T.n[1], T.n[2], c0, c1, c2), det);
  const float Et = __builtin_fmaf(1.2f * T.ed, Pm, 1e-29f);
  // t >= |det| tmin: a lower bound of the right side (tmin32 <= tmin; no bound for negative tmin)
  if (tmin32 >= 0.0f) {
    const float L = tmin32 > 0.0f ? ((ad - T.ed) * tmin32) * (1.0f - 4.0f * u) : 0.0f;
    if (ts + Et < L) return false;
  }
  // |det| h > t: an upper bound of the left side (h32 >= h)
  if (ts - Et >= ((ad + T.ed) * h32) * (1.0f + 4.0f * u)) return false;
  return true;
}

// Conservative FP32 pre-test of one sphere for the split path: returns false only when the FP64 test isect_sph
// (blazert/primitives/sphere.h:364-406) reports no hit for the ray at the current hit distance h <= h32.
//
// The FP64 test computes, from v = o - c and the unit direction d, A = d.d, B = 2 v.d, C = v.v - r r and
// D = B B - 4 A C, and accepts a root t = (-B -+ sqrt(D)) / 2A with tmin < t < h when D >= 0. With exact arithmetic
// on its inputs D = 4 (A r^2 - |v x d|^2); its roundings move D by at most 1e-13 (P^2 + r^2), where
// P >= |c_i| + |o_i| (P = (max |c32_i| + om) (1 + 4u), u = 2^-24). So
//   (a) D < 0 whenever |v x d| > r + 1.61e-7 (P + r), and
//   (b) every root lies within r + 1.61e-7 (P + r) of -v.d.
// Here v, v x d (one FMA per component) and v.d are computed in FP32 from the rounded center, origin and direction:
// each component of v is within 2.01 u P of the exact one, v x d within 12 u P (as a vector) and v.d within 10 u P.
// The record's radius is |r| rounded up. The test rejects when |v x d| in FP32 exceeds the radius plus both error
// terms (compared as squares, each rounding covered by a factor), or when the interval of (b), widened by the error
// of -v.d and the rounding of its ends, lies below tmin32 <= tmin or at or above h32 >= h (s.h carries the FP64 unit
// direction's norm, within 4 DBL_EPSILON of 1: covered by the 1 + 4u factor). A radius term below 1e-18, whose
// square would fall in the FP32 denormal range, skips test (a). Scenes and origins above FILTER_MAX_EXTENT are not
// pre-tested, and every rejection is a comparison that is false for NaN operands.
__device__ __forceinline__ bool sph_filter(const SphFilt &S, const float D[3], const float O[3], float om,
                                           float tmin32, float h32) {
  const float u = 5.9604644775390625e-08f;
  const float cm = fmaxf(fabsf(S.c[0]), fmaxf(fabsf(S.c[1]), fabsf(S.c[2])));
  const float P = (cm + om) * (1.0f + 4.0f * u);
  const float v0 = O[0] - S.c[0], v1 = O[1] - S.c[1], v2 = O[2] - S.c[2];
  const float wx = __builtin_fmaf(v1, D[2], -(v2 * D[1]));
  const float wy = __builtin_fmaf(v2, D[0], -(v0 * D[2]));
  const float wz = __builtin_fmaf(v0, D[1], -(v1 * D[0]));
  const float dist2 = __builtin_fmaf(wx, wx, __builtin_fmaf(wy, wy, wz * wz));
  const float e64 = __builtin_fmaf(1.61e-7f, P + S.r, 1e-30f);  // the FP64 test's own rounding, as a distance
  const float R = (S.r + __builtin_fmaf(12.0f * u, P, e64)) * (1.0f + 4.0f * u);
  if (R >= 1e-18f && dist2 * (1.0f - 8.0f * u) > (R * R) * (1.0f + 4.0f * u)) return false;
  const float tc = -__builtin_fmaf(v0, D[0], __builtin_fmaf(v1, D[1], v2 * D[2]));
  float rho = S.r + __builtin_fmaf(10.0f * u, P, e64);
  rho = __builtin_fmaf(4.0f * u, fabsf(tc) + rho, rho);
  if (tc + rho < tmin32) return false;
  if (tc - rho >= h32 * (1.0f + 4.0f * u)) return false;
  return true;
}

// ------------------------------------------------------------------------------------------------ intersectors
// Closest-hit state of one traversal (prepare_traversal: hit = tmax, uv = 0, prim invalid).
template <class R>
struct HitT {
  R h;
  R u, v;
  uint32_t leaf_idx;  // leaf-order index of the current hit; the user primitive id for records that carry it
};

// FP64 triangle records carry their user primitive id (TriRec::prim).
template <class Rec>
__device__ constexpr bool id_in_record() {
  return std::is_same<Rec, TriRec>::value;
}
using Hit = HitT<double>;

// Every intersector runs in the scene type of its record (Rec::real): the same operation sequence for double and
// float, with the reference's constants of that type.

// blazert/primitives/trimesh.h:153-181
template <class Rec>
__device__ __forceinline__ bool isect_tri(const Rec &T, const RayT<typename Rec::real> &r,
                                          HitT<typename Rec::real> &s, uint32_t idx) {
  using R = typename Rec::real;
  const R tol = R(4) * Real<R>::eps;
  const R c0 = T.v0[0] - r.o[0], c1 = T.v0[1] - r.o[1], c2 = T.v0[2] - r.o[2];
  const R rx = r.d[1] * c2 - r.d[2] * c1;
  const R ry = r.d[2] * c0 - r.d[0] * c2;
  const R rz = r.d[0] * c1 - r.d[1] * c0;
  const R *e1 = T.e1, *e2 = T.e2;
  R nx, ny, nz;  // cross(e1, e2)
  if constexpr (Rec::has_normal) {
    nx = T.n[0];
    ny = T.n[1];
    nz = T.n[2];
  } else {
    R n[3];
    tri_normal(e1, e2, n);
    nx = n[0];
    ny = n[1];
    nz = n[2];
  }
  const R det = nx * r.d[0] + ny * r.d[1] + nz * r.d[2];
  const R abs_det = fabs(det);
  const R u = product_sign(rx * e2[0] + ry * e2[1] + rz * e2[2], det);
  const R v = product_sign(rx * e1[0] + ry * e1[1] + rz * e1[2], det);
  const R w = abs_det - u - v;
  if (u >= -tol && v >= -tol && w >= -tol) {
    const R t = product_sign(nx * c0 + ny * c1 + nz * c2, det);
    if (t >= abs_det * r.tmin && abs_det * s.h > t) {
      const R inv_det = R(1) / abs_det;
      s.h = t * inv_det;
      s.u = u * inv_det;
      s.v = v * inv_det;
      if constexpr (id_in_record<Rec>()) s.leaf_idx = (uint32_t)T.prim;
      else s.leaf_idx = idx;
      return true;
    }
  }
  return false;
}

template <class R>
__device__ __forceinline__ R norm_scaled(R t, const R *d) {
  const R x = t * d[0], y = t * d[1], z = t * d[2];
  return dsqrt(x * x + y * y + z * z);
}

// blazert/primitives/sphere.h:364-406
template <class Rec>
__device__ __forceinline__ bool isect_sph(const Rec &S, const RayT<typename Rec::real> &r,
                                          HitT<typename Rec::real> &s, uint32_t idx) {
  using R = typename Rec::real;
  const R v0 = r.o[0] - S.c[0], v1 = r.o[1] - S.c[1], v2 = r.o[2] - S.c[2];
  const R A = r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2];
  const R B = R(2) * (v0 * r.d[0] + v1 * r.d[1] + v2 * r.d[2]);
  const R C = (v0 * v0 + v1 * v1 + v2 * v2) - S.r * S.r;
  const R D = B * B - R(4) * A * C;
  if (D < R(0)) return false;
  const R Q = dsqrt(D);
  const R rcpA = R(1) / A;
  const R t0 = R(0.5) * rcpA * (-B - Q);
  const R t1 = R(0.5) * rcpA * (-B + Q);
  if ((t0 > r.tmin) && (t0 < s.h)) {
    s.h = norm_scaled(t0, r.d);
    s.leaf_idx = idx;
    return true;
  }
  if ((t1 > r.tmin) && (t1 < s.h)) {
    s.h = norm_scaled(t1, r.d);
    s.leaf_idx = idx;
    return true;
  }
  return false;
}

// trans(R) * v and R * v for row-major R
template <class R>
__device__ __forceinline__ void matTvec(const R *M, const R *v, R *o) {
  o[0] = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
  o[1] = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
  o[2] = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
}
template <class R>
__device__ __forceinline__ void matvec(const R *M, const R *v, R *o) {
  o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
  o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
  o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
}

// Plane: local frame, t = -o_z / d_z, interior / edge / corner branches (blazert/primitives/plane.h:180-270).
// Returns the branch (0 = miss, 1 interior, 2..5 edges, 6..9 corners) for the given [tnear, tfar].
template <class Rec>
__device__ __forceinline__ int plane_branch(const Rec &P, const RayT<typename Rec::real> &r,
                                            typename Rec::real tnear, typename Rec::real tfar,
                                            typename Rec::real *org, typename Rec::real *dir,
                                            typename Rec::real &t1, bool check_t) {
  using R = typename Rec::real;
  const R oc[3] = {r.o[0] - P.c[0], r.o[1] - P.c[1], r.o[2] - P.c[2]};
  matTvec(P.R, oc, org);
  matTvec(P.R, r.d, dir);
  t1 = -org[2] / dir[2];
  const R X = org[0] + t1 * dir[0];
  const R Y = org[1] + t1 * dir[1];
  const R x_min = -P.dx / 2, x_max = P.dx / 2, y_min = -P.dy / 2, y_max = P.dy / 2;
  const bool tok = !check_t || ((t1 > tnear) && (t1 < tfar));
  if (!tok) return 0;
  if ((X > x_min) && (X < x_max) && (Y > y_min) && (Y < y_max)) return 1;
  if ((X == x_min) && (X < x_max) && (Y > y_min) && (Y < y_max)) return 2;
  if ((X > x_min) && (X == x_max) && (Y > y_min) && (Y < y_max)) return 3;
  if ((X > x_min) && (X < x_max) && (Y == y_min) && (Y < y_max)) return 4;
  if ((X > x_min) && (X < x_max) && (Y > y_min) && (Y == y_max)) return 5;
  if ((X == x_min) && (X < x_max) && (Y == y_min) && (Y < y_max)) return 6;
  if ((X == x_min) && (X < x_max) && (Y > y_min) && (Y == y_max)) return 7;
  if ((X > x_min) && (X == x_max) && (Y == y_min) && (Y < y_max)) return 8;
  if ((X > x_min) && (X == x_max) && (Y > y_min) && (Y == y_max)) return 9;
  return 0;
}

template <class Rec>
__device__ __forceinline__ bool isect_pla(const Rec &P, const RayT<typename Rec::real> &r,
                                          HitT<typename Rec::real> &s, uint32_t idx) {
  using R = typename Rec::real;
  R org[3], dir[3], t1;
  const int br = plane_branch(P, r, r.tmin, s.h, org, dir, t1, true);
  if (!br) return false;
  s.h = norm_scaled(t1, dir);
  s.leaf_idx = idx;
  return true;
}

// Plane normal of the accepted branch (before the final normalize of post_traversal).
template <class Rec>
__device__ void plane_normal(const Rec &P, const RayT<typename Rec::real> &r, typename Rec::real *nrm) {
  using R = typename Rec::real;
  R org[3], dir[3], t1;
  const int br = plane_branch(P, r, R(0), R(0), org, dir, t1, false);
  const R sgn = org[2] / fabs(org[2]);
  if (br == 1) {
    const R ez[3] = {R(0), R(0), R(1)};
    R rz[3];
    matvec(P.R, ez, rz);
    nrm[0] = sgn * rz[0];
    nrm[1] = sgn * rz[1];
    nrm[2] = sgn * rz[2];
    return;
  }
  const R lx[10] = {0, 0, -1, 1, 0, 0, -1, -1, 1, 1};
  const R ly[10] = {0, 0, 0, 0, -1, 1, -1, 1, -1, 1};
  const R l[3] = {lx[br], ly[br], sgn};
  matvec(P.R, l, nrm);
}

// Cylinder (blazert/primitives/cylinder.h:181-495). No hit-distance checks: the last accepted cylinder wins.
// Returns true on a hit with the parametric t and the (pre-normalisation) normal.
template <class R>
struct CylOut {
  R t;
  R n[3];
};

template <class R>
__device__ __forceinline__ void cyl_ng(const R *M, const R *nn, R *out) {
  R rn[3];
  matvec(M, nn, rn);
  const R len = dsqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  const R is = R(1) / len;
  out[0] = rn[0] * is;
  out[1] = rn[1] * is;
  out[2] = rn[2] * is;
}

template <class Rec>
__device__ bool cyl_eval(const Rec &Cy, const RayT<typename Rec::real> &r, typename Rec::real *dir,
                         CylOut<typename Rec::real> &out, bool want_normal) {
  using R = typename Rec::real;
  const R oc[3] = {r.o[0] - Cy.c[0], r.o[1] - Cy.c[1], r.o[2] - Cy.c[2]};
  R org[3];
  matTvec(Cy.R, oc, org);
  matTvec(Cy.R, r.d, dir);
  const R x0 = org[0], y0 = org[1], z0 = org[2];
  const R j = dir[0], k = dir[1], l = dir[2];
  const R a = Cy.a, b = Cy.b, h = Cy.h;
  const R zt = h / R(2), zb = -h / R(2);

  // shell quadratic shared by every region
  auto shell = [&](bool ordered, bool &hit) {
    hit = false;
    const R A = a * a * k * k + b * b * j * j;
    const R B = a * a * k * y0 + b * b * j * x0;
    const R C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
    if (C < 0) return true;  // reference returns false immediately
    // R(-1) / A is -(R(1) / A) exactly (round to nearest is symmetric under negation): one division for both roots
    const R inv_a = R(1) / A;
    const R s0 = inv_a * (-B + a * b * dsqrt(C));
    const R s1 = -inv_a * (B + a * b * dsqrt(C));
    const R i0z = org[2] + dir[2] * s0;
    const R i1z = org[2] + dir[2] * s1;
    if ((s1 > 0) && (!ordered || s1 < s0) && (i1z <= zt) && (i1z >= zb)) {
      out.t = s1;
      if (want_normal) {
        const R nn[3] = {2 * (org[0] + dir[0] * s1) / (a * a), 2 * (org[1] + dir[1] * s1) / (b * b), 0};
        cyl_ng(Cy.R, nn, out.n);
      }
      hit = true;
      return true;
    }
    if ((s0 > 0) && (!ordered || s0 < s1) && (i0z <= zt) && (i0z >= zb)) {
      out.t = s0;
      if (want_normal) {
        const R nn[3] = {2 * (org[0] + dir[0] * s0) / (a * a), 2 * (org[1] + dir[1] * s0) / (b * b), 0};
        cyl_ng(Cy.R, nn, out.n);
      }
      hit = true;
      return true;
    }
    return false;  // fall through
  };
  auto cap = [&](R zc, R sgn) {
    out.t = (zc - z0) / l;
    if (want_normal) {
      const R nz[3] = {R(0), R(0), sgn};
      cyl_ng(Cy.R, nz, out.n);
    }
  };

  if (z0 > zt || z0 < zb) {  // areas 1 and 2
    const bool top = z0 > zt;
    if (l == 0) return false;
    const R t0 = top ? (zt - z0) / l : (zb - z0) / l;
    if (t0 < 0) return false;
    const R ix = org[0] + t0 * dir[0], iy = org[1] + t0 * dir[1];
    if (b * b * ix * ix + a * a * iy * iy <= a * a * b * b) {
      cap(top ? zt : zb, top ? R(1) : R(-1));
      return true;
    }
    bool hit;
    if (shell(true, hit)) return hit;
  } else if ((z0 > zb) && (z0 < zt) && (x0 * x0 * b * b + y0 * y0 * a * a > a * a * b * b)) {  // area 3
    bool hit;
    if (shell(true, hit)) return hit;
  }
  if ((z0 >= zb) && (z0 <= zt) && (x0 * x0 * b * b + y0 * y0 * a * a <= a * a * b * b)) {  // area 4
    if (l == 0) {
      if ((z0 == zb) || (z0 == zt)) return false;
      bool hit;
      if (shell(false, hit)) return hit;
    } else if ((j == R(0)) && (k == R(0))) {
      if (b * b * x0 * x0 + a * a * y0 * y0 == a * a * b * b) return false;
      if (l < 0) {
        cap(zb, R(-1));
        return true;
      }
      if (l > 0) {
        cap(zt, R(1));
        return true;
      }
    } else {
      const R A = a * a * k * k + b * b * j * j;
      const R B = a * a * k * y0 + b * b * j * x0;
      const R C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
      if (C < 0) return false;
      const R inv_a = R(1) / A;
      const R s0 = inv_a * (-B + a * b * dsqrt(C));
      const R s1 = -inv_a * (B + a * b * dsqrt(C));
      const R i0z = org[2] + dir[2] * s0;
      const R i1z = org[2] + dir[2] * s1;
      if ((s1 > 0) && (i1z <= zt) && (i1z >= zb)) {
        out.t = s1;
        if (want_normal) {
          const R nn[3] = {2 * (org[0] + dir[0] * s1) / (a * a), 2 * (org[1] + dir[1] * s1) / (b * b), 0};
          cyl_ng(Cy.R, nn, out.n);
        }
        return true;
      }
      if ((s0 > 0) && (i0z <= zt) && (i0z >= zb)) {
        out.t = s0;
        if (want_normal) {  // reference stores the un-rotated normal here (cylinder.h:462-463)
          out.n[0] = 2 * (org[0] + dir[0] * s0) / (a * a);
          out.n[1] = 2 * (org[1] + dir[1] * s0) / (b * b);
          out.n[2] = 0;
        }
        return true;
      }
      if (l < 0) {
        cap(zb, R(-1));
        return true;
      }
      if (l > 0) {
        cap(zt, R(1));
        return true;
      }
    }
  }
  return false;
}

template <class Rec>
__device__ __forceinline__ bool isect_cyl(const Rec &Cy, const RayT<typename Rec::real> &r,
                                          HitT<typename Rec::real> &s, uint32_t idx) {
  using R = typename Rec::real;
  R dir[3];
  CylOut<R> o;
  if (!cyl_eval(Cy, r, dir, o, false)) return false;
  s.h = norm_scaled(o.t, dir);
  s.leaf_idx = idx;
  return true;
}

template <class Rec>
__device__ __forceinline__ bool isect(const Rec &rec, const RayT<typename Rec::real> &r, HitT<typename Rec::real> &s,
                                      uint32_t idx) {
  if constexpr (Rec::kind == TRIANGLES) return isect_tri(rec, r, s, idx);
  else if constexpr (Rec::kind == SPHERES) return isect_sph(rec, r, s, idx);
  else if constexpr (Rec::kind == PLANES) return isect_pla(rec, r, s, idx);
  else return isect_cyl(rec, r, s, idx);
}

template <class R>
__device__ __forceinline__ void normalize3(R *v) {
  const R len = dsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const R il = (len != R(0)) ? R(1) / len : R(0);
  v[0] *= il;
  v[1] *= il;
  v[2] *= il;
}

// post_traversal normals (trimesh.h:133-143, sphere.h:336-343, plane.h:153-159, cylinder.h:154-160)
template <class Rec>
__device__ void post_normal(const BvhView &bvh, const RayT<typename Rec::real> &r, const HitT<typename Rec::real> &s,
                            uint32_t pid, typename Rec::real *n) {
  using R = typename Rec::real;
  const Rec *recs = reinterpret_cast<const Rec *>(bvh.prims);
  if constexpr (Rec::kind == TRIANGLES) {
    const uint32_t *f = bvh.faces + 3ull * pid;
    const R *vn;
    if constexpr (sizeof(R) == 8) vn = bvh.vnormals;
    else vn = bvh.vnormals32;
    const R *n0 = vn + 3ull * f[0];
    const R *n1 = vn + 3ull * f[1];
    const R *n2 = vn + 3ull * f[2];
    const R w = (R(1) - s.u) - s.v;
#pragma unroll
    for (int k = 0; k < 3; k++) n[k] = (w * n0[k] + s.u * n1[k]) + s.v * n2[k];
  } else if constexpr (Rec::kind == SPHERES) {
    const Rec S = recs[s.leaf_idx];
#pragma unroll
    for (int k = 0; k < 3; k++) n[k] = (r.o[k] + s.h * r.d[k]) - S.c[k];
  } else if constexpr (Rec::kind == PLANES) {
    plane_normal(recs[s.leaf_idx], r, n);
  } else {
    R dir[3];
    CylOut<R> o;
    cyl_eval(recs[s.leaf_idx], r, dir, o, true);
    n[0] = o.n[0];
    n[1] = o.n[1];
    n[2] = o.n[2];
  }
  normalize3(n);
}

// ------------------------------------------------------------------------------------------------ traversal
// Per-lane traversal state. The DFS of the reference is advanced one unit at a time (one internal node, or one
// whole leaf), so a persistent wave can retire finished rays and pull new ones between steps.
struct Trav {
  uint32_t cur, cnt;  // current node reference and leaf count
  uint32_t pos;       // next primitive of the current leaf (with the triangle pre-test: first of the current group)
  uint32_t cand;      // split-precision triangles: group members that passed the FP32 pre-test, still to test
  bool active;        // traversal still running
  float h32;          // split precision: hit distance rounded up to FP32
};

// Traversal stack: the top `ring` entries of each lane live in LDS (slot-major, lane-minor, so a wave's accesses
// spread over all banks); older entries spill to private memory only when a ray goes deeper than the ring, and come
// back when the ring drains. Invariant: entries [lo, sp) are in LDS, [0, lo) in the private spill area.
// The ring depth is a launch parameter (a power of two, 4..16): deeper rings spill less but take more LDS per block.
// Defaults: 8 for FP64 keys; 16 for the FP32 path over triangles, whose 4-wide steps push up to three entries at a
// time (1.31M-triangle incoherent set); 8 for the FP32 path over the other types, where the fourth wave per SIMD
// that a shallower ring leaves room for pays more than the spills it causes (mixed scene, 8M rays: spheres 3.87 ->
// 3.35 ms, cylinders 4.97 -> 4.32, planes 1.66 -> 1.47 with 8 instead of 16).
template <class Rec, bool SPLIT>
constexpr uint32_t ring_default() { return (SPLIT && Rec::kind == TRIANGLES) ? 16u : 8u; }
// Default refill threshold (idle lanes of a wave before it takes new rays). Exact path: 32 for triangles (1.31M-
// triangle incoherent set: 16 is 3% slower, 40 5% slower), 16 for the other types, whose longer primitive tests
// favour earlier refills (mixed scene: spheres 3% and cylinders 5% faster than at 32). 4-wide split path: 16 for
// every type (icosphere: 32 is 7% slower).
template <class Rec, bool SPLIT>
constexpr uint32_t refill_default() { return (SPLIT || Rec::kind != TRIANGLES) ? 16u : 32u; }
// Default block size of a single-scene launch: 256, except 128 for FP64 spheres on the exact path (mixed scene, 8M
// rays: 4.30-4.33 ms against 4.44-4.51 with 256; the other types and the split path do not change between 64 and
// 256).
template <class Rec, bool SPLIT>
constexpr uint32_t block_default() { return (!SPLIT && std::is_same<Rec, SphRec>::value) ? 128u : 256u; }

// Stack keys: FP32 entry distances on the split path, otherwise the scene type's.
template <bool SPLIT, class R>
using StackKey = typename std::conditional<SPLIT, float, R>::type;

// LDS pointers of the ring stack carry the LDS address space explicitly: as generic pointers the compiler had to
// resolve their address space at run time in places (flat accesses), which costs time on the coherent sets.
#define BRT_LDS __attribute__((address_space(3)))

template <bool SPLIT, class R = double>
struct SpillArea {  // private (scratch) memory for entries pushed deeper than the ring
  using Key = StackKey<SPLIT, R>;
  Key key[stack_cap<SPLIT>()];
  uint32_t ref[stack_cap<SPLIT>()], cnt[stack_cap<SPLIT>()];
};

// The counters and pointers are kept apart from the spill arrays so that they stay in registers.
template <bool SPLIT, class R = double>
struct RingStack {
  using Key = StackKey<SPLIT, R>;
  BRT_LDS Key *lkey;  // this lane's slot 0 in LDS; slot j is at [64 * j]
  BRT_LDS uint32_t *lref, *lcnt;
  SpillArea<SPLIT, R> *spill;
  int sp, lo;
  int ring;  // LDS slots per lane (power of two)

  __device__ __forceinline__ void reset() {
    sp = 0;
    lo = 0;
  }
  __device__ __forceinline__ void push(uint32_t r, uint32_t c, Key k) {
    if (sp - lo == ring) {
      const int j = (lo & (ring - 1)) * 64;
      spill->key[lo] = lkey[j];
      spill->ref[lo] = lref[j];
      spill->cnt[lo] = lcnt[j];
      lo++;
    }
    const int j = (sp & (ring - 1)) * 64;
    lkey[j] = k;
    lref[j] = r;
    lcnt[j] = c;
    sp++;
  }
  // requires sp > 0
  __device__ __forceinline__ void pop(uint32_t &r, uint32_t &c, Key &k) {
    sp--;
    if (sp < lo) {
      lo = sp;
      r = spill->ref[sp];
      c = spill->cnt[sp];
      k = spill->key[sp];
    } else {
      const int j = (sp & (ring - 1)) * 64;
      r = lref[j];
      c = lcnt[j];
      k = lkey[j];
    }
  }
};

// LDS bytes of the ring stacks of one block
template <bool SPLIT, class R = double>
constexpr size_t ring_lds_bytes(uint32_t block, uint32_t ring) {
  return (size_t)(block / 64) * ring * 64 * (sizeof(StackKey<SPLIT, R>) + 8);
}

template <bool SPLIT, class R>
__device__ __forceinline__ void ring_bind(RingStack<SPLIT, R> &st, unsigned char *smem, SpillArea<SPLIT, R> *spill,
                                          uint32_t ring) {
  using Key = StackKey<SPLIT, R>;
  const uint32_t waves = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  BRT_LDS unsigned char *lds = (BRT_LDS unsigned char *)smem;
  const size_t slots = (size_t)waves * ring * 64;
  BRT_LDS Key *keys = (BRT_LDS Key *)lds;
  BRT_LDS uint32_t *refs = (BRT_LDS uint32_t *)(lds + slots * sizeof(Key));
  BRT_LDS uint32_t *cnts = (BRT_LDS uint32_t *)(lds + slots * (sizeof(Key) + 4));
  st.lkey = keys + (size_t)wave * ring * 64 + lane;
  st.lref = refs + (size_t)wave * ring * 64 + lane;
  st.lcnt = cnts + (size_t)wave * ring * 64 + lane;
  st.ring = (int)ring;
  st.spill = spill;
  st.reset();
}

// Ascending sort of four (key, ref, cnt) triples: compare-exchange network (0,1)(2,3)(0,2)(1,3)(1,2), every index
// a compile-time constant so the arrays stay in registers.
template <int I, int J>
__device__ __forceinline__ void cmpx(float (&k)[4], uint32_t (&r)[4], uint32_t (&c)[4]) {
  const bool sw = k[J] < k[I];
  const float ki = k[I], kj = k[J];
  const uint32_t ri = r[I], rj = r[J], ci = c[I], cj = c[J];
  k[I] = sw ? kj : ki;
  k[J] = sw ? ki : kj;
  r[I] = sw ? rj : ri;
  r[J] = sw ? ri : rj;
  c[I] = sw ? cj : ci;
  c[J] = sw ? ci : cj;
}
__device__ __forceinline__ void sort4(float (&k)[4], uint32_t (&r)[4], uint32_t (&c)[4]) {
  cmpx<0, 1>(k, r, c);
  cmpx<2, 3>(k, r, c);
  cmpx<0, 2>(k, r, c);
  cmpx<1, 3>(k, r, c);
  cmpx<1, 2>(k, r, c);
}

// Root test (the reference pushes node 0 and slab-tests it on the first pop with the ray's tmax).
template <bool SPLIT, class R>
__device__ __forceinline__ void trav_begin(const BvhView &bvh, const RayT<R> &r, const Ray32 &q, const HitT<R> &s,
                                           Trav &t) {
  t.cur = SPLIT ? bvh.root_ref4 : bvh.root_ref;
  t.cnt = bvh.root_cnt;
  t.pos = 0;
  t.cand = 0;
  if constexpr (SPLIT) {
    float k32;
    const float rb[6] = {bvh.root_min32[0], bvh.root_min32[1], bvh.root_min32[2],
                         bvh.root_max32[0], bvh.root_max32[1], bvh.root_max32[2]};
    t.h32 = f32_up(s.h);
    t.active = slab32(rb, q, k32) && !(t.h32 < k32);
  } else {
    // the descriptor holds the root box in FP64; a Scene<float> box converts back exactly
    R key;
    const R rb[6] = {R(bvh.root_min[0]), R(bvh.root_min[1]), R(bvh.root_min[2]),
                     R(bvh.root_max[0]), R(bvh.root_max[1]), R(bvh.root_max[2])};
    t.h32 = 0.0f;
    t.active = slab_exact(rb, r, key) && !(s.h < key);
  }
}

// One unit of the DFS: an internal node (test both children, push far, descend near) or a leaf's next primitives
// (exact_group per type, in order, blazert/bvh/intersect.h:66-78), followed by pops until a node passes or the stack
// is empty. On the split path with triangle or sphere pre-test records, a leaf step is either a pre-test of the
// next FILTER_GROUP triangles or SPH_FILTER_GROUP spheres in FP32 (tri_filter, sph_filter) or the FP64 test of the
// next one that passed it, in leaf order.
template <class Rec, bool SPLIT>
__device__ __forceinline__ void trav_step(const BvhView &bvh, const RayT<typename Rec::real> &r, const Ray32 &q,
                                          bool any_hit, HitT<typename Rec::real> &s, Trav &t,
                                          RingStack<SPLIT, typename Rec::real> &st) {
  using R = typename Rec::real;
  using Node = typename std::conditional<SPLIT, Node4,
                                         typename std::conditional<sizeof(R) == 8, Node64, Node32>::type>::type;
  // One fetch per lane, issued together: the internal record for lanes at an inner node, the primitive record (or
  // the group of pre-test records) for lanes at a leaf. Written as separate branches, each would wait for its own
  // loads, and a wave holding several kinds of lanes would pay several memory round trips per step. The chunks the
  // records have in common are loaded from a per-lane address in one set of instructions; the rest of the longer
  // ones follow in short branches (from the same lines, already on their way), and the compute branches below run
  // without waiting in between.
  // split precision, triangles, spheres and cylinders: FP32 pre-test records, FG of them per step
  constexpr bool FILTER = SPLIT && (Rec::kind == TRIANGLES || Rec::kind == SPHERES || Rec::kind == CYLINDERS);
  using Filt = typename std::conditional<Rec::kind == TRIANGLES, TriFilt, SphFilt>::type;
  constexpr uint32_t FG = Rec::kind == TRIANGLES ? FILTER_GROUP : SPH_FILTER_GROUP;
  constexpr int NODE_Q = SPLIT ? 8 : (sizeof(R) == 8 ? 7 : 4);  // 16-byte chunks of the record used (Node64: 112 B)
  constexpr int REC_Q = (int)(sizeof(Rec) / 16);
  constexpr int FILT_Q = FILTER ? (int)(FG * sizeof(Filt) / 16) : 0;
  // exact paths: primitives tested per leaf step
  constexpr int PG = SPLIT ? 1 : exact_group<Rec::kind>();
  constexpr int QMAX0 = NODE_Q < PG * REC_Q ? PG * REC_Q : NODE_Q;
  constexpr int QMAX = QMAX0 < FILT_Q ? FILT_Q : QMAX0;
  static_assert(sizeof(Rec) % 16 == 0 && sizeof(Node) >= 16 * NODE_Q, "record chunking");
  const bool at_leaf = (t.cur & LEAF_BIT) != 0;
  const Filt *frecs = nullptr;
  if constexpr (FILTER) {
    if constexpr (Rec::kind == TRIANGLES) frecs = bvh.tri_filter;
    else frecs = bvh.sph_filter;
  }
  const bool filt = FILTER && frecs != nullptr;  // the same for every lane of a launch
  const bool filter_step = filt && at_leaf && t.cand == 0u && t.pos < t.cnt;
  const bool fetch_rec = at_leaf && t.pos < t.cnt && !filter_step;
  const uint32_t p = (t.cur & ~LEAF_BIT) + t.pos + (filt ? (uint32_t)__builtin_ctz(t.cand | 0x80000000u) : 0u);
  const Node *nodes;
  if constexpr (SPLIT) nodes = reinterpret_cast<const Node *>(bvh.nodes4);
  else if constexpr (sizeof(R) == 8) nodes = reinterpret_cast<const Node *>(bvh.nodes64);
  else nodes = reinterpret_cast<const Node *>(bvh.nodes32);
  // The loads go through a pointer in the global address space: the batch kernel's tree pointers come from its job
  // table in memory, and as generic pointers every access through them would be a flat access.
  using u32x4 = unsigned int __attribute__((ext_vector_type(4)));
  using GlobalChunk = const __attribute__((address_space(1))) u32x4;
  GlobalChunk *src = !at_leaf ? (GlobalChunk *)(nodes + t.cur)
                     : filter_step ? (GlobalChunk *)(frecs + ((t.cur & ~LEAF_BIT) + t.pos))
                                   : (GlobalChunk *)(reinterpret_cast<const Rec *>(bvh.prims) + p);
  u32x4 chunk[QMAX];
  if constexpr (FILTER) {
    // three record kinds: each chunk is loaded by the lanes whose record has it
    const int nq = !at_leaf ? NODE_Q : (filter_step ? FILT_Q : (fetch_rec ? REC_Q : 0));
#pragma unroll
    for (int k = 0; k < QMAX; k++)
      if (k < nq) chunk[k] = src[k];
  } else {
    constexpr int COMMON = NODE_Q < REC_Q ? NODE_Q : REC_Q;
    if (!at_leaf || fetch_rec) {
#pragma unroll
      for (int k = 0; k < COMMON; k++) chunk[k] = src[k];
    }
    if (!at_leaf) {
#pragma unroll
      for (int k = COMMON; k < NODE_Q; k++) chunk[k] = src[k];
    } else if (fetch_rec) {
#pragma unroll
      for (int k = COMMON; k < REC_Q; k++) chunk[k] = src[k];
#pragma unroll
      for (int g = 1; g < PG; g++)
        if (t.pos + g < t.cnt) {
#pragma unroll
          for (int k = g * REC_Q; k < (g + 1) * REC_Q; k++) chunk[k] = src[k];
        }
    }
  }
  // an empty asm taking every chunk as an operand: the loads cannot be sunk into the compute branches below, so all
  // of them are issued first and waited for once
#pragma unroll
  for (int k = 0; k < QMAX; k++) asm volatile("" : "+v"(chunk[k]));
  if (!at_leaf) {
    bool near_ok, far_ok;
    uint32_t nref, ncnt, fref, fcnt;
    StackKey<SPLIT, R> fkey;
    Node nd;
    __builtin_memcpy(&nd, chunk, 16 * NODE_Q);
    if constexpr (SPLIT) {
      // up to four children per record, visited nearest entry first (the FP32 path needs no particular order:
      // it only has to test every primitive the FP64 path tests)
      float k[4];
      uint32_t rf[4], cn[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        float A = q.tmin, M = t.h32;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const bool sg = (q.sign >> a) & 1u;
          const float lo = nd.lo[a][c], hi = nd.hi[a][c];
          A = fmaxf(A, __builtin_fmaf(sg ? hi : lo, q.inv[a], q.offn[a]));
          M = fminf(M, __builtin_fmaf(sg ? lo : hi, q.inv[a], q.offf[a]));
        }
        k[c] = (nd.cnt[c] != EMPTY_SLOT && A <= M) ? A : __builtin_inff();
        rf[c] = nd.ref[c];
        cn[c] = nd.cnt[c];
      }
      sort4(k, rf, cn);
      const float INF = __builtin_inff();
      if (k[3] < INF) st.push(rf[3], cn[3], k[3]);
      if (k[2] < INF) st.push(rf[2], cn[2], k[2]);
      if (k[1] < INF) st.push(rf[1], cn[1], k[1]);
      if (k[0] < INF) {
        t.cur = rf[0];
        t.cnt = cn[0];
        t.pos = 0;
        return;
      }
      near_ok = far_ok = false;
      nref = ncnt = fref = fcnt = 0;
      fkey = 0.0f;
    } else {
      const uint32_t axis = nd.cnt[1] >> AXIS_SHIFT;
      const uint32_t nc = (r.sign >> axis) & 1u;
      R k0 = R(0), k1 = R(0);
      const bool p0 = slab_exact(nd.b, r, k0);
      const bool p1 = slab_exact(nd.b + 6, r, k1);
      near_ok = (nc ? p1 : p0) && !(s.h < (nc ? k1 : k0));
      far_ok = nc ? p0 : p1;
      fkey = nc ? k0 : k1;
      const uint32_t r0 = nd.ref[0], r1 = nd.ref[1], c0 = nd.cnt[0] & CNT_MASK, c1 = nd.cnt[1] & CNT_MASK;
      nref = nc ? r1 : r0;
      ncnt = nc ? c1 : c0;
      fref = nc ? r0 : r1;
      fcnt = nc ? c0 : c1;
    }
    if (far_ok) st.push(fref, fcnt, fkey);
    if (near_ok) {
      t.cur = nref;
      t.cnt = ncnt;
      t.pos = 0;
      return;
    }
  } else if (filter_step) {
    if constexpr (FILTER) {
      // pre-test the group; members past the leaf's end are padding or the next leaf's primitives
      Filt g[FG];
      __builtin_memcpy(g, chunk, sizeof g);
      const float D[3] = {(float)r.d[0], (float)r.d[1], (float)r.d[2]};
      const float O[3] = {(float)r.o[0], (float)r.o[1], (float)r.o[2]};
      const uint32_t left = t.cnt - t.pos;
      uint32_t mask = 0;
#pragma unroll
      for (uint32_t j = 0; j < FG; j++) {
        bool pass;
        if constexpr (Rec::kind == TRIANGLES) pass = tri_filter(g[j], D, O, q.om, q.tmin, t.h32);
        else if constexpr (Rec::kind == SPHERES) pass = sph_filter(g[j], D, O, q.om, q.tmin, t.h32);
        else pass = sph_filter(g[j], D, O, q.om, 0.0f, __builtin_inff());  // cylinders: any hit at t >= 0 counts
        if (j < left && pass) mask |= 1u << j;
      }
      t.cand = mask;
      if (mask) return;
      t.pos += FG;
      if (t.pos < t.cnt) return;
    }
  } else {
    // a few primitives per step: a leaf's tests (blazert/bvh/intersect.h:66-78, in order, each against the hit
    // distance the previous one left) are spread over consecutive steps, so lanes at leaves and lanes at inner
    // nodes run branches of similar length instead of inner-node lanes idling through a whole leaf
    if (fetch_rec) {
      Rec rec;
      __builtin_memcpy(&rec, chunk, sizeof(Rec));
      if (isect(rec, r, s, p)) {
        if (SPLIT) t.h32 = f32_up(s.h);
        if (any_hit) {
          t.active = false;
          return;
        }
      }
#pragma unroll
      for (int g = 1; g < PG; g++) {  // the leaf's next primitives, each against the hit distance just left
        if (t.pos + 1 >= t.cnt) break;
        __builtin_memcpy(&rec, chunk + g * REC_Q, sizeof(Rec));
        t.pos++;
        if (isect(rec, r, s, p + g) && any_hit) {
          t.active = false;
          return;
        }
      }
      if (filt) {
        t.cand &= t.cand - 1u;
        if (t.cand) return;
        t.pos += FG;
        if (t.pos < t.cnt) return;
      } else if (++t.pos < t.cnt) {
        return;
      }
    }
  }
  // pop the next far child whose entry distance does not exceed the current hit distance
  while (st.sp > 0) {
    uint32_t ref, cnt;
    StackKey<SPLIT, R> key;
    st.pop(ref, cnt, key);
    bool ok;
    if constexpr (SPLIT) ok = !(t.h32 < key);
    else ok = !(s.h < key);
    if (ok) {
      t.cur = ref;
      t.cnt = cnt;
      t.pos = 0;
      return;
    }
  }
  t.active = false;
}

// Merge into the hit record (blazert/scene.h:169-209): triangles replace unconditionally, later types only if
// strictly closer; the first type of the scene also writes the default RayHit for misses.
template <class Rec>
__device__ __forceinline__ void finish_ray(const BvhView &bvh, const HitsViewT<typename Rec::real> &hits, uint64_t i,
                                           const RayT<typename Rec::real> &r, const HitT<typename Rec::real> &s,
                                           uint32_t flags, uint32_t first, uint32_t is_tri) {
  using R = typename Rec::real;
  const bool hit = s.h < r.tmax;
  bool write = false;
  if (hit) {
    if (is_tri) write = true;
    else write = s.h < ((first && !(flags & FLAG_INOUT)) ? Real<R>::max : gmem(hits.t)[i]);
  }
  if (write) {
    const uint32_t pid = id_in_record<Rec>() ? s.leaf_idx : gmem(bvh.prim_ids)[s.leaf_idx];
    if (hits.merged) st_stream(hits.merged, i, 1u);
    st_stream(hits.t, i, s.h);
    if (hits.u) st_stream(hits.u, i, s.u);
    if (hits.v) st_stream(hits.v, i, s.v);
    st_stream(hits.prim_id, i, pid);
    if (hits.geom_id) st_stream(hits.geom_id, i, bvh.geom_id);
    if (hits.nx) {
      R nrm[3];
      post_normal<Rec>(bvh, r, s, pid, nrm);
      st_stream(hits.nx, i, nrm[0]);
      st_stream(hits.ny, i, nrm[1]);
      st_stream(hits.nz, i, nrm[2]);
    }
  } else if (first && !(flags & FLAG_INOUT)) {
    st_stream(hits.t, i, Real<R>::max);
    if (hits.u) st_stream(hits.u, i, R(0));
    if (hits.v) st_stream(hits.v, i, R(0));
    st_stream(hits.prim_id, i, INVALID_ID);
    if (hits.geom_id) st_stream(hits.geom_id, i, INVALID_ID);
    if (hits.nx) {
      st_stream(hits.nx, i, R(0));
      st_stream(hits.ny, i, R(0));
      st_stream(hits.nz, i, R(0));
    }
  }
}

template <bool SPLIT, class R>
__device__ __forceinline__ void start_ray(const BvhView &bvh, const RaysViewT<R> &rays, uint64_t i, double S,
                                          RayT<R> &r, Ray32 &q, HitT<R> &s) {
  make_ray(rays, i, r);
  if (bvh.irregular_boxes) r.sign &= ~RAY_NAN_FREE;
  if constexpr (SPLIT) make_ray32(r, S, q);
  s.h = r.tmax;  // prepare_traversal
  s.u = R(0);
  s.v = R(0);
  s.leaf_idx = INVALID_ID;
}

// One ray per thread (grid-stride).
template <class Rec, bool SPLIT>
__global__ void __launch_bounds__(256) k_trace(BvhView bvh, RaysViewT<typename Rec::real> rays,
                                               HitsViewT<typename Rec::real> hits, uint64_t n, uint32_t flags,
                                               uint32_t first, uint32_t is_tri, double S, const uint32_t *order,
                                               uint32_t ring, unsigned long long *deferred) {
  using R = typename Rec::real;
  extern __shared__ unsigned char smem[];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const bool any_hit = (flags & FLAG_ANY_HIT) != 0;
  const bool deferred_only = !SPLIT && (flags & FLAG_DEFERRED_ONLY);
  if (deferred_only && __atomic_load_n(deferred, __ATOMIC_RELAXED) == 0ull) return;
  SpillArea<SPLIT, R> spill;
  RingStack<SPLIT, R> st;
  ring_bind<SPLIT, R>(st, smem, &spill, ring);
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const uint64_t i = order ? (uint64_t)order[k] : k;
    RayT<R> r;
    Ray32 q;
    HitT<R> s;
    start_ray<SPLIT>(bvh, rays, i, S, r, q, s);
    if (SPLIT && deferred_to_exact(r)) {
      __atomic_store_n(deferred, 1ull, __ATOMIC_RELAXED);
      continue;
    }
    if (deferred_only && !deferred_to_exact(r)) continue;
    Trav t;
    trav_begin<SPLIT>(bvh, r, q, s, t);
    st.reset();
    while (t.active) trav_step<Rec, SPLIT>(bvh, r, q, any_hit, s, t, st);
    finish_ray<Rec>(bvh, hits, i, r, s, flags, first, is_tri);
  }
}

// Persistent waves: every lane owns one ray at a time; when at least `refill` lanes of a wave are idle, they start
// the next indices of the wave's pool. Skewed per-ray work (many rays miss the root box, a few visit hundreds of
// nodes) then no longer idles the wave until its slowest ray ends. The pool is a chunk of `chunk` consecutive indices
// (POOL_CHUNK unless the launch asks for 32 or 128) claimed with one atomic add. Claims on one address serialise
// across the GPU (about 14 ns each), which bounded short launches and image-order rays, so they are spread over
// CLAIM_COUNTERS counters on separate 128-byte lines: counter k hands out chunks k, k + CLAIM_COUNTERS,
// k + 2 CLAIM_COUNTERS, ... A wave starts at the counter of its block (block index mod CLAIM_COUNTERS: blocks are
// dealt to the eight XCDs in turn, so each XCD's waves start at a counter of their own) and moves on to the next
// counter once one has passed n. When all have, the pool is exhausted; every lane leaves the loop once its pool is
// exhausted and its own ray is done, so the grid always drains.
constexpr uint32_t POOL_CHUNK = 64;

// Wave-uniform pool of claimed, not yet started ray indices [next, end).
struct WavePool {
  uint64_t next = 0, end = 0;
  bool last = false;  // every counter has passed n
};

// Called by every lane of a wave with the same pool state when the wave refills. An idle lane handed an index gets
// it in `idx` and `got`; `exhausted` is set once nothing is left to hand out. `counter` is the first of the
// launch's CLAIM_COUNTERS counters (COUNTER_STRIDE slots apart).
__device__ __forceinline__ void pool_refill(WavePool &P, unsigned long long *counter, uint64_t n, uint32_t chunk,
                                            unsigned long long idle, uint32_t nidle, uint32_t lane, bool &got,
                                            uint64_t &idx, bool &exhausted) {
  got = false;
  if (P.next == P.end && !P.last) {
    unsigned long long base = n;
    const uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
    if (lane == leader) {
      const uint32_t home = blockIdx.x % (uint32_t)CLAIM_COUNTERS;
      for (uint32_t a = 0; a < (uint32_t)CLAIM_COUNTERS; a++) {
        const uint32_t k = (home + a) % (uint32_t)CLAIM_COUNTERS;
        const unsigned long long c = atomicAdd(counter + k * COUNTER_STRIDE, 1ull);
        const unsigned long long start = (c * CLAIM_COUNTERS + k) * chunk;
        if (start < n) {
          base = start;
          break;
        }
      }
    }
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)base, (int)leader);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(base >> 32), (int)leader);
    base = ((unsigned long long)hi << 32) | lo;
    if (base >= n) {
      P.last = true;
      P.next = P.end = n;
    } else {
      P.next = base;
      P.end = base + chunk < n ? base + chunk : n;
    }
  }
  const uint64_t avail = P.end - P.next;
  const uint32_t take = avail < nidle ? (uint32_t)avail : nidle;
  const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
  if (((idle >> lane) & 1ull) && rank < take) {
    idx = P.next + rank;
    got = true;
  }
  P.next += take;
  if (P.last && P.next == P.end) exhausted = true;
}
// Occupancy target of the persistent kernels (waves per SIMD the register allocation must allow).
#ifndef BRT_WAVES_PER_EU
#define BRT_WAVES_PER_EU 4
#endif

// The ray and hit column pointers (17 of them) are used only where a lane starts or finishes a ray. Held in
// scalar registers for the whole kernel they overflowed the scalar register file, and the compiler reloaded them
// from spill lanes as whole 16-register tuples at every start and finish. The persistent triangle kernels instead
// keep one copy per block in LDS and re-read it (volatile, so it is not hoisted back into registers) where it is
// used: incoherent triangle launches 1-2% and coherent ones 3-5% faster. The cylinder launch was 2% slower that way
// and the sphere and plane launches no faster, so they keep the kernel arguments.
template <class T>
__device__ __forceinline__ void lds_store(uint64_t *dst, const T &x) {
  static_assert(sizeof(T) % 8 == 0, "8-byte words");
  __builtin_memcpy(dst, &x, sizeof(T));
}
template <class T>
__device__ __forceinline__ T lds_fresh(const uint64_t *src) {
  using lds_u64 = const volatile __attribute__((address_space(3))) uint64_t;
  lds_u64 *v = (lds_u64 *)src;
  uint64_t w[sizeof(T) / 8];
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 8); k++) {
    const uint64_t x = v[k];  // the same in every lane: moved to scalar registers for the address arithmetic
    w[k] = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  }
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

template <class Rec, bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BRT_WAVES_PER_EU)))
k_trace_persistent(BvhView bvh, RaysViewT<typename Rec::real> rays, HitsViewT<typename Rec::real> hits, uint64_t n,
                   uint32_t flags, uint32_t first, uint32_t is_tri, double S, unsigned long long *counter,
                   uint32_t refill, const uint32_t *order, uint32_t ring, unsigned long long *deferred,
                   uint32_t chunk) {
  using R = typename Rec::real;
  extern __shared__ unsigned char smem[];
  const bool any_hit = (flags & FLAG_ANY_HIT) != 0;
  const bool deferred_only = !SPLIT && (flags & FLAG_DEFERRED_ONLY);
  if (deferred_only && __atomic_load_n(deferred, __ATOMIC_RELAXED) == 0ull) return;
  constexpr bool LDS_VIEWS = Rec::kind == TRIANGLES;
  constexpr int RW = (int)(sizeof(RaysViewT<R>) / 8), HW = (int)(sizeof(HitsViewT<R>) / 8);
  __shared__ uint64_t views[RW + HW];
  if constexpr (LDS_VIEWS) {
    if (threadIdx.x == 0) {
      lds_store(views, rays);
      lds_store(views + RW, hits);
    }
    __syncthreads();
  }
  auto rays_view = [&]() -> RaysViewT<R> {
    if constexpr (LDS_VIEWS) return lds_fresh<RaysViewT<R>>(views);
    else return rays;
  };
  auto hits_view = [&]() -> HitsViewT<R> {
    if constexpr (LDS_VIEWS) return lds_fresh<HitsViewT<R>>(views + RW);
    else return hits;
  };
  const uint32_t lane = __lane_id();
  SpillArea<SPLIT, R> spill;
  RingStack<SPLIT, R> st;
  ring_bind<SPLIT, R>(st, smem, &spill, ring);
  RayT<R> r;
  Ray32 q;
  HitT<R> s;
  Trav t;
  t.active = false;
  bool has_ray = false;
  bool exhausted = false;  // wave-uniform
  uint64_t idx = 0;
  WavePool pool;
  for (;;) {
    if (!exhausted) {
      const unsigned long long idle = __ballot(!has_ray);
      const uint32_t nidle = (uint32_t)__popcll(idle);
      if (nidle >= refill || nidle == 64u) {
        bool got;
        uint64_t cand = 0;
        pool_refill(pool, counter, n, chunk, idle, nidle, lane, got, cand, exhausted);
        if (got) {
          idx = order ? (uint64_t)order[cand] : cand;
          start_ray<SPLIT>(bvh, rays_view(), idx, S, r, q, s);
          if (SPLIT && deferred_to_exact(r)) {
            __atomic_store_n(deferred, 1ull, __ATOMIC_RELAXED);
          } else if (!deferred_only || deferred_to_exact(r)) {
            has_ray = true;
            trav_begin<SPLIT>(bvh, r, q, s, t);
            st.reset();
          }
        }
      }
    }
    // lanes may drop the rays they were handed (deferred or skipped), so an idle wave leaves only once the pool is
    // exhausted
    if (!__ballot(has_ray) && exhausted) break;
    if (has_ray) {
      if (t.active) trav_step<Rec, SPLIT>(bvh, r, q, any_hit, s, t, st);
      if (!t.active) {
        finish_ray<Rec>(bvh, hits_view(), idx, r, s, flags, first, is_tri);
        has_ray = false;
      }
    }
  }
}

// Many scenes in one launch: the claimed index runs over the concatenated rays of all jobs (one job = one scene's
// BVH of this geometry type + its own rays and hits). A lane finds its job by binary search over the prefix offsets
// and reads that job's BVH view from the table (a few cached loads per step), so waves freely mix scenes and the
// uneven batch drains with no per-scene tail.
constexpr uint32_t BATCH_LDS_JOBS = 32;

__device__ __forceinline__ uint32_t find_job(const BatchJob *jobs, uint32_t njobs, uint64_t g) {
  uint32_t lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (jobs[mid].offset <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class Rec, bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BRT_WAVES_PER_EU)))
k_trace_batch(const BatchJob *__restrict__ jobs, uint32_t njobs, uint64_t total, uint32_t flags,
              unsigned long long *counter, uint32_t refill, uint32_t ring, unsigned long long *deferred,
              uint32_t chunk) {
  extern __shared__ unsigned char smem[];
  const bool any_hit = (flags & FLAG_ANY_HIT) != 0;
  const bool deferred_only = !SPLIT && (flags & FLAG_DEFERRED_ONLY);
  if (deferred_only && __atomic_load_n(deferred, __ATOMIC_RELAXED) == 0ull) return;
  // Finishing a ray needs its job's hit columns and flags: for batches of up to BATCH_LDS_JOBS jobs every block
  // keeps them in LDS, so that a finish does not wait on a load from the job table.
  constexpr int JW = (int)(sizeof(HitsView) / 8) + 1;  // hit column pointers, then first | is_tri << 32
  __shared__ uint64_t job_fin[BATCH_LDS_JOBS * JW];
  const bool fin_lds = njobs <= BATCH_LDS_JOBS;
  if (fin_lds) {
    for (uint32_t k = threadIdx.x; k < njobs; k += blockDim.x) {
      lds_store(job_fin + k * JW, jobs[k].hits);
      job_fin[k * JW + JW - 1] = (uint64_t)jobs[k].first | ((uint64_t)jobs[k].is_tri << 32);
    }
    __syncthreads();
  }
  const uint32_t lane = __lane_id();
  SpillArea<SPLIT> spill;
  RingStack<SPLIT> st;
  ring_bind<SPLIT>(st, smem, &spill, ring);
  Ray r;
  Ray32 q;
  Hit s;
  Trav t;
  t.active = false;
  bool has_ray = false;
  bool exhausted = false;  // wave-uniform
  uint64_t idx = 0;
  uint32_t j = 0;
  BvhView bv{};  // this lane's job's tree, kept across steps (the step reads only its node and record pointers)
  WavePool pool;
  for (;;) {
    if (!exhausted) {
      const unsigned long long idle = __ballot(!has_ray);
      const uint32_t nidle = (uint32_t)__popcll(idle);
      if (nidle >= refill || nidle == 64u) {
        bool got;
        uint64_t g = 0;
        pool_refill(pool, counter, total, chunk, idle, nidle, lane, got, g, exhausted);
        // A claim is a run of consecutive indices and usually lies in one job: find the jobs of its first and last
        // index once per wave (wave-uniform, so scalar loads), and search per lane only when they differ.
        const unsigned long long gm = __ballot(got);
        if (gm) {
          const int lo = __ffsll((long long)gm) - 1, hi = 63 - __clzll((long long)gm);
          const uint64_t g_lo = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(g >> 32), lo) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)g, lo);
          const uint64_t g_hi = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(g >> 32), hi) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)g, hi);
          const uint32_t j_lo = find_job(jobs, njobs, g_lo);
          const bool one_job = j_lo + 1u >= njobs || jobs[j_lo + 1u].offset > g_hi;
          if (got) {
            j = one_job ? j_lo : find_job(jobs, njobs, g);
            const BatchJob &J = jobs[j];
            const uint64_t local = g - J.offset;
            idx = J.order ? (uint64_t)J.order[local] : local;
            bv = J.bvh;
            start_ray<SPLIT>(bv, J.rays, idx, J.extent, r, q, s);
            if (SPLIT && deferred_to_exact(r)) {
              __atomic_store_n(deferred, 1ull, __ATOMIC_RELAXED);
            } else if (!deferred_only || deferred_to_exact(r)) {
              has_ray = true;
              trav_begin<SPLIT>(bv, r, q, s, t);
              st.reset();
            }
          }
        }
      }
    }
    if (!__ballot(has_ray) && exhausted) break;
    if (has_ray) {
      if (t.active) trav_step<Rec, SPLIT>(bv, r, q, any_hit, s, t, st);
      if (!t.active) {
        HitsView hv;
        uint32_t jfirst, jtri;
        if (fin_lds) {
          __builtin_memcpy(&hv, job_fin + j * JW, sizeof(HitsView));
          const uint64_t w = job_fin[j * JW + JW - 1];
          jfirst = (uint32_t)w;
          jtri = (uint32_t)(w >> 32);
        } else {
          const BatchJob &J = jobs[j];
          hv = J.hits;
          jfirst = J.first;
          jtri = J.is_tri;
        }
        finish_ray<Rec>(bv, hv, idx, r, s, flags, jfirst, jtri);
        has_ray = false;
      }
    }
  }
}

// Split-precision pre-test records of the triangles in leaf order (TriFilt), from their FP64 records; the FILTER_PAD
// records after the last one are zero.
__global__ void k_tri_filter_records(const TriRec *__restrict__ recs, uint32_t n, TriFilt *out, uint32_t total) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    TriFilt f{};
    if (i < n) {
      const TriRec &T = recs[i];
      double nn[3];
      tri_normal(T.e1, T.e2, nn);  // the FP64 test's n, operation for operation
      double vm = 0.0, sa = 0.0, sb = 0.0;
      for (int k = 0; k < 3; k++) {
        f.v0[k] = (float)T.v0[k];
        f.e1[k] = (float)T.e1[k];
        f.e2[k] = (float)T.e2[k];
        f.n[k] = (float)nn[k];
        vm = fmax(vm, fabs(T.v0[k]));
        sa += fabs(T.e1[k]);
        sb += fabs(T.e2[k]);
      }
      // upper bounds (the FP64 sums and products of non-negative terms are within a few ulps of exact)
      const double up = 1.0 + 0x1p-48, u = 0x1p-24;
      f.vmax = f32_up(vm * up);
      f.kua = f32_up(16.0 * u * sa * up);
      f.kub = f32_up(16.0 * u * sb * up);
      f.ed = f32_up((10.0 * u * sa * sb + 1e-29) * up);
    }
    out[i] = f;
  }
}

// Split-precision pre-test records of the spheres in leaf order (SphFilt): the center rounded to nearest and |r|
// rounded up; the SPH_FILTER_PAD records after the last one are zero.
__global__ void k_sph_filter_records(const SphRec *__restrict__ recs, uint32_t n, SphFilt *out, uint32_t total) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    SphFilt f{};
    if (i < n) {
      const SphRec &S = recs[i];
      for (int k = 0; k < 3; k++) f.c[k] = (float)S.c[k];
      f.r = f32_up(fabs(S.r));
    }
    out[i] = f;
  }
}

// Split-precision pre-test records of the cylinders: each one's bounding sphere, for sph_filter with the window
// [0, inf). The cylinder test (cylinder.h:181-495) accepts any hit at t >= 0, with no distance check, at a point of
// the local frame org + t dir = R^T (o + t d - c) with |z| <= h/2 and, up to the rounding of the test, on or inside
// the elliptic shell of semi-axes a, b; for an orthonormal R that point lies within sqrt(max(a,b)^2 + (h/2)^2) of c.
// The radius is enlarged by 2^-10 of itself and 2^-20 of the coordinate magnitude, far above any rounding of the
// FP64 test, and is infinite (never rejects) when R R^T differs from the identity by more than 1e-6 or an input is
// not finite.
__global__ void k_cyl_filter_records(const CylRec *__restrict__ recs, uint32_t n, SphFilt *out, uint32_t total) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    SphFilt f{};
    if (i < n) {
      const CylRec &C = recs[i];
      double dev = 0.0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          const double d = C.R[3 * a] * C.R[3 * b] + C.R[3 * a + 1] * C.R[3 * b + 1] + C.R[3 * a + 2] * C.R[3 * b + 2];
          dev = fmax(dev, fabs(d - (a == b ? 1.0 : 0.0)));
        }
      const double ab = fmax(fabs(C.a), fabs(C.b)), hh = 0.5 * fabs(C.h);
      const double rb = sqrt(ab * ab + hh * hh);
      const double cm = fmax(fabs(C.c[0]), fmax(fabs(C.c[1]), fabs(C.c[2])));
      double rad = rb * (1.0 + 0x1p-10) + 0x1p-20 * (cm + rb);
      if (!(dev <= 1e-6) || !(rad < 1e300)) rad = __builtin_inf();
      for (int k = 0; k < 3; k++) f.c[k] = (float)C.c[k];
      f.r = f32_up(rad);
    }
    out[i] = f;
  }
}

// Clears a launch's counter group (a fill kernel, not hipMemsetAsync, so that graph captures replay correctly).
__global__ void k_zero_counters(unsigned long long *c) {
  for (uint32_t k = threadIdx.x; k < (uint32_t)COUNTER_GROUP; k += blockDim.x) c[k] = 0ull;
}

// 3 x 9-bit Morton interleave
__device__ __forceinline__ uint32_t spread9(uint32_t x) {
  x &= 0x1ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Sort key of ray i: direction octant (3 bits, the children visit order depends only on it) above a 27-bit Morton
// code of the origin quantised in the scene box (mode 1), or the Morton code alone (mode 2).
__global__ void k_sort_keys(RaysView rays, uint64_t n, double lx, double ly, double lz, double sx, double sy,
                            double sz, uint32_t mode, uint32_t *keys, uint32_t *vals) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double q[3] = {(rays.ox[i] - lx) * sx, (rays.oy[i] - ly) * sy, (rays.oz[i] - lz) * sz};
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double c = q[k] != q[k] ? 0.0 : fmin(fmax(q[k], 0.0), 511.0);
      m |= spread9((uint32_t)c) << k;
    }
    uint32_t oct = (rays.dx[i] < 0.0 ? 1u : 0u) | (rays.dy[i] < 0.0 ? 2u : 0u) | (rays.dz[i] < 0.0 ? 4u : 0u);
    keys[i] = mode == 1u ? ((oct << 27) | m) : m;
    vals[i] = (uint32_t)i;
  }
}

template <class R>
__global__ void k_init_hits(HitsViewT<R> hits, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    hits.t[i] = Real<R>::max;
    if (hits.u) hits.u[i] = R(0);
    if (hits.v) hits.v[i] = R(0);
    hits.prim_id[i] = INVALID_ID;
    if (hits.geom_id) hits.geom_id[i] = INVALID_ID;
    if (hits.nx) {
      hits.nx[i] = R(0);
      hits.ny[i] = R(0);
      hits.nz[i] = R(0);
    }
  }
}

// Blocks of a persistent launch: as many as fit on the device at once, but no more than the rays need.
template <class K>
hipError_t persistent_blocks(K kernel, uint32_t block, size_t lds, uint64_t n, uint64_t &blocks) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kernel, (int)block, lds);
  if (e != hipSuccess) return e;
  blocks = (uint64_t)cus * (uint64_t)(per_cu > 0 ? per_cu : 1);
  const uint64_t need = (n + block - 1) / block;
  if (blocks > need) blocks = need;
  return hipSuccess;
}

// Enqueues one traversal of a geometry type: the counter group is cleared, then the trace kernel runs. A split-
// precision launch is followed by the FP64 launch over the rays it deferred (deferred_to_exact), which returns at
// once when there are none; both run in stream order before the next geometry type, so intersect1's merge order
// holds for the deferred rays too.
template <class Rec, bool SPLIT>
hipError_t launch_kernel(const BvhView &bvh, const RaysViewT<typename Rec::real> &rays,
                         const HitsViewT<typename Rec::real> &hits, uint64_t n, const TraceLaunch &cfg,
                         const SplitParams &sp, uint32_t is_tri, hipStream_t stream) {
  using R = typename Rec::real;
  const uint32_t block = cfg.block ? cfg.block : block_default<Rec, SPLIT>();
  const uint32_t first = cfg.first ? 1u : 0u;
  const uint32_t ring = cfg.ring ? cfg.ring : ring_default<Rec, SPLIT>();
  const size_t lds = ring_lds_bytes<SPLIT, R>(block, ring);
  unsigned long long *grp = cfg.counter;
  if (!grp) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, stream, grp);
  if (cfg.schedule == SCHEDULE_PERSISTENT) {
    uint64_t blocks = 0;
    hipError_t e = persistent_blocks(k_trace_persistent<Rec, SPLIT>, block, lds, n, blocks);
    if (e != hipSuccess) return e;
    const uint32_t refill = cfg.refill ? cfg.refill : refill_default<Rec, SPLIT>();
    hipLaunchKernelGGL((k_trace_persistent<Rec, SPLIT>), dim3((uint32_t)blocks), dim3(block), lds, stream, bvh, rays,
                       hits, n, cfg.flags, first, is_tri, sp.scene_extent, grp + COUNTER_POOL, refill, cfg.order,
                       ring, grp + COUNTER_DEFERRED, cfg.pool ? cfg.pool : POOL_CHUNK);
  } else {
    uint64_t blocks = (n + block - 1) / block;
    if (blocks > (1ull << 30)) blocks = 1ull << 30;
    hipLaunchKernelGGL((k_trace<Rec, SPLIT>), dim3((uint32_t)blocks), dim3(block), lds, stream, bvh, rays, hits, n,
                       cfg.flags, first, is_tri, sp.scene_extent, cfg.order, ring, grp + COUNTER_DEFERRED);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !SPLIT) return e;
  // the FP64 launch over the deferred rays (default launch shape of the exact path)
  const uint32_t ring64 = ring_default<Rec, false>();
  const size_t lds64 = ring_lds_bytes<false, R>(256, ring64);
  uint64_t blocks = 0;
  e = persistent_blocks(k_trace_persistent<Rec, false>, 256, lds64, n, blocks);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_trace_persistent<Rec, false>), dim3((uint32_t)blocks), dim3(256), lds64, stream, bvh, rays,
                     hits, n, cfg.flags | FLAG_DEFERRED_ONLY, first, is_tri, sp.scene_extent,
                     grp + COUNTER_DEFERRED_POOL, refill_default<Rec, false>(), cfg.order, ring64,
                     grp + COUNTER_DEFERRED, POOL_CHUNK);
  return hipGetLastError();
}

template <class Rec, bool SPLIT>
hipError_t launch_batch_kernel(const BatchJob *jobs, uint32_t njobs, uint64_t total, const TraceLaunch &cfg,
                               hipStream_t stream) {
  const uint32_t block = cfg.block ? cfg.block : 256;
  const uint32_t ring = cfg.ring ? cfg.ring : ring_default<Rec, SPLIT>();
  const size_t lds = ring_lds_bytes<SPLIT>(block, ring);
  unsigned long long *grp = cfg.counter;
  if (!grp || njobs == 0) return hipErrorInvalidValue;
  if (total == 0) return hipSuccess;
  uint64_t blocks = 0;
  hipError_t e = persistent_blocks(k_trace_batch<Rec, SPLIT>, block, lds, total, blocks);
  if (e != hipSuccess) return e;
  const uint32_t refill = cfg.refill ? cfg.refill : refill_default<Rec, SPLIT>();
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, stream, grp);
  hipLaunchKernelGGL((k_trace_batch<Rec, SPLIT>), dim3((uint32_t)blocks), dim3(block), lds, stream, jobs, njobs, total,
                     cfg.flags, grp + COUNTER_POOL, refill, ring, grp + COUNTER_DEFERRED, cfg.pool ? cfg.pool : POOL_CHUNK);
  e = hipGetLastError();
  if (e != hipSuccess || !SPLIT) return e;
  const uint32_t ring64 = ring_default<Rec, false>();
  const size_t lds64 = ring_lds_bytes<false>(256, ring64);
  e = persistent_blocks(k_trace_batch<Rec, false>, 256, lds64, total, blocks);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_trace_batch<Rec, false>), dim3((uint32_t)blocks), dim3(256), lds64, stream, jobs, njobs, total,
                     cfg.flags | FLAG_DEFERRED_ONLY, grp + COUNTER_DEFERRED_POOL, refill_default<Rec, false>(),
                     ring64, grp + COUNTER_DEFERRED, POOL_CHUNK);
  return hipGetLastError();
}

template <class Rec>
hipError_t launch_typed(const BvhView &bvh, const RaysView &rays, const HitsView &hits, uint64_t n,
                        const TraceLaunch &cfg, const SplitParams &sp, uint32_t is_tri, hipStream_t stream) {
  if (cfg.precision == PREC_FP64) return launch_kernel<Rec, false>(bvh, rays, hits, n, cfg, sp, is_tri, stream);
  return launch_kernel<Rec, true>(bvh, rays, hits, n, cfg, sp, is_tri, stream);
}

}  // namespace

hipError_t launch_trace(GeomType type, const BvhView &bvh, const RaysView &rays, const HitsView &hits, uint64_t n,
                        const TraceLaunch &cfg, const SplitParams &split, hipStream_t stream) {
  switch (type) {
  case TRIANGLES: return launch_typed<TriRec>(bvh, rays, hits, n, cfg, split, 1u, stream);
  case SPHERES: return launch_typed<SphRec>(bvh, rays, hits, n, cfg, split, cfg.overwrite ? 1u : 0u, stream);
  case PLANES: return launch_typed<PlaRec>(bvh, rays, hits, n, cfg, split, cfg.overwrite ? 1u : 0u, stream);
  default: return launch_typed<CylRec>(bvh, rays, hits, n, cfg, split, cfg.overwrite ? 1u : 0u, stream);
  }
}

hipError_t launch_sph_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream) {
  const uint32_t total = n + SPH_FILTER_PAD;
  uint32_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_sph_filter_records, dim3(blocks), dim3(256), 0, stream, static_cast<const SphRec *>(recs), n,
                     static_cast<SphFilt *>(out), total);
  return hipGetLastError();
}

hipError_t launch_cyl_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream) {
  const uint32_t total = n + SPH_FILTER_PAD;
  uint32_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_cyl_filter_records, dim3(blocks), dim3(256), 0, stream, static_cast<const CylRec *>(recs), n,
                     static_cast<SphFilt *>(out), total);
  return hipGetLastError();
}

hipError_t launch_tri_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream) {
  const uint32_t total = n + FILTER_PAD;
  uint32_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_tri_filter_records, dim3(blocks), dim3(256), 0, stream, static_cast<const TriRec *>(recs), n,
                     static_cast<TriFilt *>(out), total);
  return hipGetLastError();
}

hipError_t sort_workspace_reserve(SortWorkspace &ws, uint64_t n) {
  if (n > 0x7fffffffull) return hipErrorInvalidValue;
  if (ws.capacity >= n && ws.base) return hipSuccess;
  size_t temp = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                    (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 30);
  if (e != hipSuccess) return e;
  if (ws.base) {
    e = hipFree(ws.base);
    ws.base = nullptr;
    ws.capacity = 0;
    if (e != hipSuccess) return e;
  }
  const size_t arr = ((size_t)n * 4 + 255) & ~(size_t)255;
  void *p = nullptr;
  e = hipMalloc(&p, 4 * arr + temp);
  if (e != hipSuccess) return e;
  unsigned char *b = static_cast<unsigned char *>(p);
  ws.base = p;
  ws.keys_in = reinterpret_cast<uint32_t *>(b);
  ws.keys_out = reinterpret_cast<uint32_t *>(b + arr);
  ws.vals_in = reinterpret_cast<uint32_t *>(b + 2 * arr);
  ws.vals_out = reinterpret_cast<uint32_t *>(b + 3 * arr);
  ws.temp = b + 4 * arr;
  ws.temp_bytes = temp;
  ws.capacity = n;
  return hipSuccess;
}

hipError_t sort_workspace_in(SortWorkspace &ws, void *base, size_t bytes, uint64_t n) {
  if (n > 0x7fffffffull) return hipErrorInvalidValue;
  size_t temp = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                    (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 30);
  if (e != hipSuccess) return e;
  const size_t arr = ((size_t)n * 4 + 255) & ~(size_t)255;
  if (4 * arr + temp > bytes) return hipErrorInvalidValue;
  unsigned char *b = static_cast<unsigned char *>(base);
  ws.base = nullptr;  // not owned
  ws.keys_in = reinterpret_cast<uint32_t *>(b);
  ws.keys_out = reinterpret_cast<uint32_t *>(b + arr);
  ws.vals_in = reinterpret_cast<uint32_t *>(b + 2 * arr);
  ws.vals_out = reinterpret_cast<uint32_t *>(b + 3 * arr);
  ws.temp = b + 4 * arr;
  ws.temp_bytes = temp;
  ws.capacity = n;
  return hipSuccess;
}

hipError_t sort_rays(const RaysView &rays, uint64_t n, const double lo[3], const double hi[3], uint32_t mode,
                     SortWorkspace &ws, hipStream_t stream, const uint32_t **order) {
  *order = nullptr;
  if (n == 0) return hipSuccess;
  if (ws.capacity < n) return hipErrorInvalidValue;
  double sc[3];
  for (int k = 0; k < 3; k++) {
    const double ext = hi[k] - lo[k];
    sc[k] = (ext > 0.0 && ext == ext) ? 512.0 / ext : 0.0;
  }
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_sort_keys, dim3((uint32_t)blocks), dim3(256), 0, stream, rays, n, lo[0], lo[1], lo[2], sc[0],
                     sc[1], sc[2], mode, ws.keys_in, ws.vals_in);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t temp = ws.temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(ws.temp, temp, ws.keys_in, ws.keys_out, ws.vals_in, ws.vals_out, (int)n, 0,
                                         30, stream);
  if (e != hipSuccess) return e;
  *order = ws.vals_out;
  return hipSuccess;
}

hipError_t launch_trace_batch(GeomType type, const BatchJob *jobs, uint32_t njobs, uint64_t total,
                              const TraceLaunch &cfg, hipStream_t stream) {
  const bool split = cfg.precision != PREC_FP64;
  switch (type) {
  case TRIANGLES:
    return split ? launch_batch_kernel<TriRec, true>(jobs, njobs, total, cfg, stream)
                 : launch_batch_kernel<TriRec, false>(jobs, njobs, total, cfg, stream);
  case SPHERES:
    return split ? launch_batch_kernel<SphRec, true>(jobs, njobs, total, cfg, stream)
                 : launch_batch_kernel<SphRec, false>(jobs, njobs, total, cfg, stream);
  case PLANES:
    return split ? launch_batch_kernel<PlaRec, true>(jobs, njobs, total, cfg, stream)
                 : launch_batch_kernel<PlaRec, false>(jobs, njobs, total, cfg, stream);
  default:
    return split ? launch_batch_kernel<CylRec, true>(jobs, njobs, total, cfg, stream)
                 : launch_batch_kernel<CylRec, false>(jobs, njobs, total, cfg, stream);
  }
}

template <class R>
hipError_t init_hits(const HitsViewT<R> &hits, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > (1ull << 30)) blocks = 1ull << 30;
  hipLaunchKernelGGL(k_init_hits<R>, dim3((uint32_t)blocks), dim3(256), 0, stream, hits, n);
  return hipGetLastError();
}
hipError_t launch_init_hits(const HitsView &hits, uint64_t n, hipStream_t stream) { return init_hits(hits, n, stream); }
hipError_t launch_init_hits_f32(const HitsViewF &hits, uint64_t n, hipStream_t stream) {
  return init_hits(hits, n, stream);
}

// Scratch memory of brt_intersect_noalloc, for callers that cannot allocate (graph capture, serving loops): each
// GPU that loads this library holds one copy until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol); it uses the device address from noalloc_scratch.
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(unsigned char **ptr) {
  constexpr int MAX_DEVICES = 64;
  static std::mutex mu;
  static unsigned char *addr[MAX_DEVICES] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = static_cast<unsigned char *>(p);
  }
  *ptr = addr[dev];
  return hipSuccess;
}

// Scene<float>: the exact path in FP32 (no split precision: the scene type is already FP32).
hipError_t launch_trace_f32(GeomType type, const BvhView &bvh, const RaysViewF &rays, const HitsViewF &hits,
                            uint64_t n, const TraceLaunch &cfg, hipStream_t stream) {
  const SplitParams sp{0.0};
  const uint32_t ow = cfg.overwrite ? 1u : 0u;
  switch (type) {
  case TRIANGLES: return launch_kernel<TriRecF, false>(bvh, rays, hits, n, cfg, sp, 1u, stream);
  case SPHERES: return launch_kernel<SphRecF, false>(bvh, rays, hits, n, cfg, sp, ow, stream);
  case PLANES: return launch_kernel<PlaRecF, false>(bvh, rays, hits, n, cfg, sp, ow, stream);
  default: return launch_kernel<CylRecF, false>(bvh, rays, hits, n, cfg, sp, ow, stream);
  }
}

}  // namespace brt
