// A caller written like the reference's path tracer (examples/path_tracer/main.cc: light sampling :103-126, the
// reflect / refract / Fresnel helpers :219-240, the shadow ray :255-270), built against blazert_amd/blazert.hpp
// with the blaze vector arithmetic it uses. Prints every computed value as a C99 hex float.
#include <blazert_amd/blazert.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>

using namespace blazert;

#ifndef DROPIN_FT
#define DROPIN_FT double
#endif
using T = DROPIN_FT;

template <typename U>
inline Vec3r<U> reflect(const Vec3r<U> &I, const Vec3r<U> &N) { return {I - U(2.) * dot(I, N) * N}; }

template <typename U>
inline Vec3r<U> refract(const Vec3r<U> &I, const Vec3r<U> &N, U eta) {
  const U NdotI = dot(N, I);
  const U k = static_cast<U>(1.0) - eta * eta * (static_cast<U>(1.0) - NdotI * NdotI);
  if (k < static_cast<U>(0.0))
    return Vec3r<U>{0.};
  else
    return {eta * I - (eta * NdotI + std::sqrt(k)) * N};
}

template <typename U>
inline U pow5(const U val) { return {val * val * val * val * val}; }

template <typename U>
inline U fresnel_schlick(const Vec3r<U> &H, const Vec3r<U> &norm, U n1) {
  const U r0 = n1 * n1;
  return r0 + (static_cast<U>(1.) - r0) * pow5(static_cast<U>(1.) - dot(H, norm));
}

static void put(const char *name, double x) { std::printf("%s %a\n", name, x); }
static void put(const char *name, const Vec3r<T> &v) {
  for (int k = 0; k < 3; k++) std::printf("%s%d %a\n", name, k, (double)v[k]);
}

int main() {
  const Vec3r<T> v0{T(0.1), T(0.2), T(0.3)}, v1{T(1.5), T(-0.25), T(0.75)}, v2{T(-0.5), T(2.0), T(1.25)};
  const Vec3r<T> x{T(0.3), T(-1.0), T(4.0)};
  const T Xi1 = T(0.375), Xi2 = T(0.8125);
  const T Xi1_ = std::sqrt(Xi1);
  const T c0 = static_cast<T>(1.0) - Xi1_;
  const T c1 = Xi1_ * (static_cast<T>(1.0) - Xi2);
  const T c2 = Xi1_ * Xi2;

  Vec3r<T> tmp = cross(v1 - v0, v2 - v0);
  Vec3r<T> norm = normalize(tmp);
  T area = length(tmp) / static_cast<T>(2.0);
  const Vec3r<T> lp{c0 * v0 + c1 * v1 + c2 * v2};
  Vec3r<T> dstDir = lp - x;
  const T dstDist = length(dstDir);
  dstDir = normalize(dstDir);
  const T cosAtLight = std::max(dot(-dstDir, norm), static_cast<T>(0.0));
  const Vec3r<T> ll{T(1.0), T(0.5), T(0.25)};
  const Vec3r<T> dstRadiance = ll * cosAtLight;
  const Vec3r<T> refl = reflect(dstDir, norm);
  const Vec3r<T> refr = refract(dstDir, norm, static_cast<T>(1.0 / 1.5));
  const Vec3r<T> tir = refract(dstDir, norm, static_cast<T>(4.0));  // total internal reflection: the zero vector
  const T F = fresnel_schlick(norm, dstDir, static_cast<T>(0.2));
  Vec3r<T> acc{T(0.5)};
  acc += refl;
  acc *= T(3.0);
  acc -= refr / T(4.0);
  const Vec3r<T> schur = refl * refr;

  static const T ray_eps = T(0.00001);
  const Ray<T> shadow_ray{x, lp - x, ray_eps, dstDist - ray_eps, Ray<T>::CullBackFace::no, Ray<T>::AnyHit::yes};
  const Mat3r<T> R{{T(0.0), T(-1.0), T(0.0)}, {T(1.0), T(0.0), T(0.0)}, {T(0.0), T(0.0), T(1.0)}};
  const Vec3r<T> local = trans(R) * (lp - x);

  put("tmp", tmp);
  put("norm", norm);
  put("area", area);
  put("lp", lp);
  put("dstDist", dstDist);
  put("dstDir", dstDir);
  put("cosAtLight", cosAtLight);
  put("rad", dstRadiance);
  put("refl", refl);
  put("refr", refr);
  put("tir", tir);
  put("F", F);
  put("acc", acc);
  put("schur", schur);
  put("raydir", shadow_ray.direction);
  put("rayinv", shadow_ray.direction_inv);
  for (int k = 0; k < 3; k++) std::printf("raysign%d %u\n", k, shadow_ray.direction_sign[k]);
  put("raytmax", shadow_ray.max_hit_distance);
  put("local", local);
  return 0;
}
