// blazert_amd/blazert.hpp -- C++17 drop-in for blazeRT's scene API on the MI355X engine.
//
// Mirrors the reference's public types so code written against blazert::Scene<T> compiles unchanged for
// T = double and T = float (blazert/blazert.h:15-21, blazert/scene.h:38-116, blazert/ray.h:20-85):
//   blazert::Vec3r<T>, Vec2r<T>, Vec3ui, Mat3r<T>, Vec3rList<T>, Vec3iList, Mat3rList<T>  (blazert/datatypes.h:31-59)
//   blazert::Ray<T>{origin, direction, min_hit_distance, max_hit_distance, cull_back_face, any_hit}
//   blazert::RayHit<T>{normal, uv, hit_distance, prim_id, geom_id}
//   blazert::Scene<T>::add_mesh / add_spheres / add_planes / add_cylinders / commit, build_options
//   bool blazert::intersect1(const Scene<T>&, const Ray<T>&, RayHit<T>&)
// The reference's vector types are blaze's (blazert/datatypes.h:31-59); these are small aggregates with operator[]
// (operator() for matrices) and the blaze arithmetic its callers use (+ - * /, dot, cross, norm, normalize, ...).
// Geometry is copied into the engine at add time. intersect1 is one ray per call (a GPU round trip); batches
// of device-resident rays go through brt_intersect() in blazert_amd.h.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <ostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../blazert_amd.h"

namespace blazert {

template <typename T>
struct Vec3r {
  std::array<T, 3> v{};
  Vec3r() = default;
  Vec3r(T x, T y, T z) : v{x, y, z} {}
  explicit Vec3r(T s) : v{s, s, s} {}
  T &operator[](size_t i) { return v[i]; }
  const T &operator[](size_t i) const { return v[i]; }
  static constexpr size_t size() { return 3; }
  Vec3r &operator+=(const Vec3r &b) { return *this = *this + b; }
  Vec3r &operator-=(const Vec3r &b) { return *this = *this - b; }
  Vec3r &operator*=(T s) { return *this = *this * s; }
  Vec3r &operator/=(T s) { return *this = *this / s; }
  Vec3r &operator*=(const Vec3r &b) { return *this = *this * b; }
};

template <typename T>
struct Vec2r {
  std::array<T, 2> v{};
  Vec2r() = default;
  Vec2r(T x, T y) : v{x, y} {}
  explicit Vec2r(T s) : v{s, s} {}
  T &operator[](size_t i) { return v[i]; }
  const T &operator[](size_t i) const { return v[i]; }
  static constexpr size_t size() { return 2; }
};

struct Vec3ui {
  std::array<unsigned int, 3> v{};
  Vec3ui() = default;
  Vec3ui(unsigned int a, unsigned int b, unsigned int c) : v{a, b, c} {}
  unsigned int &operator[](size_t i) { return v[i]; }
  const unsigned int &operator[](size_t i) const { return v[i]; }
  static constexpr size_t size() { return 3; }
};

// Row-major 3x3 (blaze::rowMajor in the reference), constructible from nested initializer lists.
template <typename T>
struct Mat3r {
  std::array<T, 9> m{};
  Mat3r() = default;
  Mat3r(std::initializer_list<std::initializer_list<T>> rows) {
    size_t i = 0;
    for (const auto &r : rows) {
      size_t j = 0;
      for (T x : r) m[3 * i + j++] = x;
      ++i;
    }
  }
  static Mat3r identity() { return Mat3r{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; }
  T &operator()(size_t i, size_t j) { return m[3 * i + j]; }
  const T &operator()(size_t i, size_t j) const { return m[3 * i + j]; }
  static constexpr size_t rows() { return 3; }
  static constexpr size_t columns() { return 3; }
};

// Vector and matrix arithmetic of the blaze types the reference's callers use (blaze 3.7 semantics, evaluated in
// the expression order the engine and its oracle assume; compile with -ffp-contract=off for the same bits):
// componentwise + - and * between vectors, scalar * / (a division multiplies by the reciprocal), dot with the
// sum ((a0 b0 + a1 b1) + a2 b2), cross, norm / length = sqrt(dot(a, a)), normalize(v) = v * (1 / norm(v)) (a zero
// vector stays zero), componentwise min / max / abs, trans, matrix-vector and matrix-matrix products (row sums
// left to right).
template <typename T>
inline Vec3r<T> operator+(const Vec3r<T> &a, const Vec3r<T> &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
template <typename T>
inline Vec3r<T> operator-(const Vec3r<T> &a, const Vec3r<T> &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
template <typename T>
inline Vec3r<T> operator-(const Vec3r<T> &a) { return {-a[0], -a[1], -a[2]}; }
template <typename T>
inline Vec3r<T> operator*(const Vec3r<T> &a, const Vec3r<T> &b) { return {a[0] * b[0], a[1] * b[1], a[2] * b[2]}; }
template <typename T>
inline Vec3r<T> operator/(const Vec3r<T> &a, const Vec3r<T> &b) { return {a[0] / b[0], a[1] / b[1], a[2] / b[2]}; }
template <typename T>
inline Vec3r<T> operator*(const Vec3r<T> &a, T s) { return {a[0] * s, a[1] * s, a[2] * s}; }
template <typename T>
inline Vec3r<T> operator*(T s, const Vec3r<T> &a) { return {s * a[0], s * a[1], s * a[2]}; }
template <typename T>
inline Vec3r<T> operator/(const Vec3r<T> &a, T s) { return a * (T(1) / s); }
// scalar / vector: componentwise (1 / d in the reference's Ray constructor, blazert/ray.h:60)
template <typename T>
inline Vec3r<T> operator/(T s, const Vec3r<T> &a) { return {s / a[0], s / a[1], s / a[2]}; }
template <typename T>
inline bool operator==(const Vec3r<T> &a, const Vec3r<T> &b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }
template <typename T>
inline bool operator!=(const Vec3r<T> &a, const Vec3r<T> &b) { return !(a == b); }

template <typename T>
inline Vec2r<T> operator+(const Vec2r<T> &a, const Vec2r<T> &b) { return {a[0] + b[0], a[1] + b[1]}; }
template <typename T>
inline Vec2r<T> operator-(const Vec2r<T> &a, const Vec2r<T> &b) { return {a[0] - b[0], a[1] - b[1]}; }
template <typename T>
inline Vec2r<T> operator*(const Vec2r<T> &a, T s) { return {a[0] * s, a[1] * s}; }
template <typename T>
inline Vec2r<T> operator*(T s, const Vec2r<T> &a) { return {s * a[0], s * a[1]}; }

template <typename T>
inline T dot(const Vec3r<T> &a, const Vec3r<T> &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <typename T>
inline Vec3r<T> cross(const Vec3r<T> &a, const Vec3r<T> &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <typename T>
inline T sqrNorm(const Vec3r<T> &a) { return dot(a, a); }
template <typename T>
inline T norm(const Vec3r<T> &a) { return std::sqrt(dot(a, a)); }
template <typename T>
inline T sqrLength(const Vec3r<T> &a) { return sqrNorm(a); }
template <typename T>
inline T length(const Vec3r<T> &a) { return norm(a); }
template <typename T>
inline Vec3r<T> normalize(const Vec3r<T> &a) {
  const T len = norm(a);
  return a * (len != T(0) ? T(1) / len : T(0));
}
template <typename T>
inline Vec3r<T> min(const Vec3r<T> &a, const Vec3r<T> &b) {
  return {std::min(a[0], b[0]), std::min(a[1], b[1]), std::min(a[2], b[2])};
}
template <typename T>
inline Vec3r<T> max(const Vec3r<T> &a, const Vec3r<T> &b) {
  return {std::max(a[0], b[0]), std::max(a[1], b[1]), std::max(a[2], b[2])};
}
template <typename T>
inline Vec3r<T> abs(const Vec3r<T> &a) { return {std::abs(a[0]), std::abs(a[1]), std::abs(a[2])}; }
template <typename T>
inline T min(const Vec3r<T> &a) { return std::min(std::min(a[0], a[1]), a[2]); }
template <typename T>
inline T max(const Vec3r<T> &a) { return std::max(std::max(a[0], a[1]), a[2]); }

template <typename T>
inline Mat3r<T> trans(const Mat3r<T> &a) {
  Mat3r<T> t;
  for (size_t i = 0; i < 3; i++)
    for (size_t j = 0; j < 3; j++) t(i, j) = a(j, i);
  return t;
}
template <typename T>
inline Vec3r<T> operator*(const Mat3r<T> &A, const Vec3r<T> &x) {
  return {A(0, 0) * x[0] + A(0, 1) * x[1] + A(0, 2) * x[2], A(1, 0) * x[0] + A(1, 1) * x[1] + A(1, 2) * x[2],
          A(2, 0) * x[0] + A(2, 1) * x[1] + A(2, 2) * x[2]};
}
template <typename T>
inline Mat3r<T> operator*(const Mat3r<T> &A, const Mat3r<T> &B) {
  Mat3r<T> C;
  for (size_t i = 0; i < 3; i++)
    for (size_t j = 0; j < 3; j++) C(i, j) = A(i, 0) * B(0, j) + A(i, 1) * B(1, j) + A(i, 2) * B(2, j);
  return C;
}
template <typename T>
inline Mat3r<T> operator*(const Mat3r<T> &A, T s) {
  Mat3r<T> C;
  for (size_t k = 0; k < 9; k++) C.m[k] = A.m[k] * s;
  return C;
}
template <typename T>
inline Mat3r<T> operator*(T s, const Mat3r<T> &A) { return A * s; }
template <typename T>
inline Mat3r<T> operator+(const Mat3r<T> &A, const Mat3r<T> &B) {
  Mat3r<T> C;
  for (size_t k = 0; k < 9; k++) C.m[k] = A.m[k] + B.m[k];
  return C;
}

// blaze prints a column vector one element per line in parentheses
template <typename T>
inline std::ostream &operator<<(std::ostream &os, const Vec3r<T> &a) {
  return os << "( " << a[0] << " )\n( " << a[1] << " )\n( " << a[2] << " )\n";
}

template <typename T>
using Vec3rList = std::vector<Vec3r<T>>;
using Vec3iList = std::vector<Vec3ui>;
template <typename T>
using Mat3rList = std::vector<Mat3r<T>>;

// blazert/bvh/options.h:14-28
template <typename T>
struct BVHBuildOptions {
  unsigned int min_leaf_primitives = 4;
  unsigned int max_tree_depth = 28;
  unsigned int bin_size = 64;
};

// blazert/ray.h:20-66: direction is the normalised argument; direction_inv and direction_sign are taken from the
// raw argument, as in the reference's constructor (its parameter shadows the member). The raw argument is kept for
// the engine, which forms the same three fields on the device.
template <typename T>
class Ray {
 public:
  enum class CullBackFace { yes, no };
  enum class AnyHit { yes, no };
  const Vec3r<T> origin;
  const Vec3r<T> direction;
  const Vec3r<T> direction_inv;
  const Vec3ui direction_sign;
  T min_hit_distance;
  T max_hit_distance;
  CullBackFace cull_back_face;
  AnyHit any_hit;
  Ray() = delete;
  Ray(const Vec3r<T> &origin, const Vec3r<T> &direction, T min_hit_distance = T(0.),
      T max_hit_distance = std::numeric_limits<T>::max(), CullBackFace cull_back_face = CullBackFace::no,
      AnyHit any_hit = AnyHit::no)
      : origin(origin), direction(normalize(direction)), direction_inv(T(1) / direction),
        direction_sign(direction[0] < T(0) ? 1u : 0u, direction[1] < T(0) ? 1u : 0u, direction[2] < T(0) ? 1u : 0u),
        min_hit_distance(min_hit_distance), max_hit_distance(max_hit_distance), cull_back_face(cull_back_face),
        any_hit(any_hit), raw_direction_(direction) {}
  const Vec3r<T> &raw_direction() const { return raw_direction_; }

 private:
  Vec3r<T> raw_direction_;
};

// blazert/ray.h:72-85
template <typename T>
struct RayHit {
  Vec3r<T> normal;
  Vec2r<T> uv;
  T hit_distance = std::numeric_limits<T>::max();
  unsigned int prim_id = static_cast<unsigned int>(-1);
  unsigned int geom_id = static_cast<unsigned int>(-1);
};

template <typename T>
class BlazertScene;

namespace detail {
// The C ABI of each scene type: Scene<double> (brt_*) and Scene<float> (brt_*_f32).
template <typename T>
struct Api;
template <>
struct Api<double> {
  using Hit1 = brt_hit1;
  static int create(const brt_build_options *o, brt_scene **s) { return brt_scene_create(o, s); }
  static int add_mesh(brt_scene *s, const double *v, uint32_t nv, const uint32_t *f, uint32_t nf, uint32_t *g) {
    return brt_scene_add_mesh(s, v, nv, f, nf, g);
  }
  static int add_spheres(brt_scene *s, const double *c, const double *r, uint32_t n, uint32_t *g) {
    return brt_scene_add_spheres(s, c, r, n, g);
  }
  static int add_planes(brt_scene *s, const double *c, const double *dx, const double *dy, const double *R,
                        uint32_t n, uint32_t *g) {
    return brt_scene_add_planes(s, c, dx, dy, R, n, g);
  }
  static int add_cylinders(brt_scene *s, const double *c, const double *a, const double *b, const double *h,
                           const double *R, uint32_t n, uint32_t *g) {
    return brt_scene_add_cylinders(s, c, a, b, h, R, n, g);
  }
  static int intersect1(const brt_scene *s, const double *o, const double *d, double tmin, double tmax,
                        uint32_t flags, brt_hit1 *h, int *hit) {
    return brt_intersect1(s, o, d, tmin, tmax, flags, h, hit);
  }
};
template <>
struct Api<float> {
  using Hit1 = brt_hit1_f32;
  static int create(const brt_build_options *o, brt_scene **s) { return brt_scene_create_f32(o, s); }
  static int add_mesh(brt_scene *s, const float *v, uint32_t nv, const uint32_t *f, uint32_t nf, uint32_t *g) {
    return brt_scene_add_mesh_f32(s, v, nv, f, nf, g);
  }
  static int add_spheres(brt_scene *s, const float *c, const float *r, uint32_t n, uint32_t *g) {
    return brt_scene_add_spheres_f32(s, c, r, n, g);
  }
  static int add_planes(brt_scene *s, const float *c, const float *dx, const float *dy, const float *R, uint32_t n,
                        uint32_t *g) {
    return brt_scene_add_planes_f32(s, c, dx, dy, R, n, g);
  }
  static int add_cylinders(brt_scene *s, const float *c, const float *a, const float *b, const float *h,
                           const float *R, uint32_t n, uint32_t *g) {
    return brt_scene_add_cylinders_f32(s, c, a, b, h, R, n, g);
  }
  static int intersect1(const brt_scene *s, const float *o, const float *d, float tmin, float tmax, uint32_t flags,
                        brt_hit1_f32 *h, int *hit) {
    return brt_intersect1_f32(s, o, d, tmin, tmax, flags, h, hit);
  }
};
}  // namespace detail

template <typename T>
bool intersect1(const BlazertScene<T> &scene, const Ray<T> &ray, RayHit<T> &rayhit);

// blazert/scene.h:38-116 for T = double or float. add_* return the geometry id, or unsigned(-1) for a second add of
// a geometry type or any add after commit (blazert/scene.h:230-239).
template <typename T>
class BlazertScene {
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                "the MI355X engine implements Scene<double> and Scene<float>");
  using Api = detail::Api<T>;

 public:
  BVHBuildOptions<T> build_options;
  bool has_been_committed = false;

  BlazertScene() = default;

  unsigned int add_mesh(const Vec3rList<T> &vertices, const Vec3iList &triangles) {
    std::vector<T> v(3 * vertices.size());
    std::vector<uint32_t> f(3 * triangles.size());
    for (size_t i = 0; i < vertices.size(); i++)
      for (int k = 0; k < 3; k++) v[3 * i + k] = vertices[i][k];
    for (size_t i = 0; i < triangles.size(); i++)
      for (int k = 0; k < 3; k++) f[3 * i + k] = triangles[i][k];
    uint32_t gid = BRT_INVALID_ID;
    Api::add_mesh(ensure(), v.data(), (uint32_t)vertices.size(), f.data(), (uint32_t)triangles.size(), &gid);
    return gid;
  }

  unsigned int add_spheres(const Vec3rList<T> &centers, const std::vector<T> &radii) {
    const std::vector<T> c = flat(centers);
    uint32_t gid = BRT_INVALID_ID;
    Api::add_spheres(ensure(), c.data(), radii.data(), (uint32_t)radii.size(), &gid);
    return gid;
  }

  unsigned int add_planes(const Vec3rList<T> &centers, const std::vector<T> &dxs, const std::vector<T> &dys,
                          const Mat3rList<T> &rotations) {
    const std::vector<T> c = flat(centers), r = flat(rotations);
    uint32_t gid = BRT_INVALID_ID;
    Api::add_planes(ensure(), c.data(), dxs.data(), dys.data(), r.data(), (uint32_t)dxs.size(), &gid);
    return gid;
  }

  unsigned int add_cylinders(const Vec3rList<T> &centers, const std::vector<T> &semi_axes_a,
                             const std::vector<T> &semi_axes_b, const std::vector<T> &heights,
                             const Mat3rList<T> &rotations) {
    const std::vector<T> c = flat(centers), r = flat(rotations);
    uint32_t gid = BRT_INVALID_ID;
    Api::add_cylinders(ensure(), c.data(), semi_axes_a.data(), semi_axes_b.data(), heights.data(), r.data(),
                       (uint32_t)semi_axes_a.size(), &gid);
    return gid;
  }

  // blazert/scene.h:91-115 (build_options are read when the first geometry is added)
  bool commit() {
    if (brt_scene_commit(ensure()) != BRT_OK) throw std::runtime_error(brt_last_error());
    has_been_committed = true;
    return has_been_committed;
  }

  brt_scene *native() const { return handle_.get(); }

 private:
  struct Deleter {
    void operator()(brt_scene *s) const { brt_scene_destroy(s); }
  };
  std::unique_ptr<brt_scene, Deleter> handle_;

  brt_scene *ensure() {
    if (!handle_) {
      const brt_build_options o{build_options.min_leaf_primitives, build_options.max_tree_depth,
                                build_options.bin_size};
      brt_scene *s = nullptr;
      if (Api::create(&o, &s) != BRT_OK) throw std::runtime_error(brt_last_error());
      handle_.reset(s);
    }
    return handle_.get();
  }

  static std::vector<T> flat(const Vec3rList<T> &a) {
    std::vector<T> o(3 * a.size());
    for (size_t i = 0; i < a.size(); i++)
      for (int k = 0; k < 3; k++) o[3 * i + k] = a[i][k];
    return o;
  }
  static std::vector<T> flat(const Mat3rList<T> &a) {
    std::vector<T> o(9 * a.size());
    for (size_t i = 0; i < a.size(); i++)
      for (int k = 0; k < 9; k++) o[9 * i + k] = a[i].m[k];
    return o;
  }
};

template <typename T>
using Scene = BlazertScene<T>;

// blazert/scene.h:161-211: rayhit is in/out; returns whether a hit was merged into it.
template <typename T>
bool intersect1(const BlazertScene<T> &scene, const Ray<T> &ray, RayHit<T> &rayhit) {
  const T o[3] = {ray.origin[0], ray.origin[1], ray.origin[2]};
  const Vec3r<T> &raw = ray.raw_direction();
  const T d[3] = {raw[0], raw[1], raw[2]};
  typename detail::Api<T>::Hit1 h{};
  for (int k = 0; k < 3; k++) h.normal[k] = rayhit.normal[k];
  h.u = rayhit.uv[0];
  h.v = rayhit.uv[1];
  h.t = rayhit.hit_distance;
  h.prim_id = rayhit.prim_id;
  h.geom_id = rayhit.geom_id;
  int hit = 0;
  const uint32_t flags = ray.any_hit == Ray<T>::AnyHit::yes ? BRT_FLAG_ANY_HIT : 0u;
  if (detail::Api<T>::intersect1(scene.native(), o, d, ray.min_hit_distance, ray.max_hit_distance, flags, &h, &hit) !=
      BRT_OK)
    throw std::runtime_error(brt_last_error());
  for (int k = 0; k < 3; k++) rayhit.normal[k] = h.normal[k];
  rayhit.uv[0] = h.u;
  rayhit.uv[1] = h.v;
  rayhit.hit_distance = h.t;
  rayhit.prim_id = h.prim_id;
  rayhit.geom_id = h.geom_id;
  return hit != 0;
}

}  // namespace blazert
