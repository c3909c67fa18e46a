// User-defined geometry on the GPU, by example: the counterpart of blazeRT's examples/geometry_template/
// GEOM_TEMPLATE.h. Two collections written the way blazeRT's primitive headers are written (a primitive, an
// intersector, a collection, and the free functions primitive_from_collection / prepare_traversal /
// intersect_primitive / post_traversal), against include/blazert_amd/user_geometry.h. Built by
// blazert_amd.build.build_user_geometry (hipcc --offload-arch=gfx950 -ffp-contract=off, linked to libblazert_amd).
//
// * UserSpheres<T>: the reference's sphere (blazert/primitives/sphere.h:95-150, 336-343, 364-406) re-expressed as a
//   user type; its hits must equal the engine's built-in sphere path bit for bit.
// * UserBoxes: axis-aligned boxes hit at their entry distance, a type the engine does not have; checked against a
//   brute-force minimum over all boxes.
//
// Exported through a small C ABI (ug_*) for the Python tests and the bench.
#include <hip/hip_runtime.h>

#include <utility>

#include "blazert_amd/user_geometry.h"

namespace ug {

using blazert_amd::Ray;
using blazert_amd::RayHit;
using blazert_amd::Vec2r;
using blazert_amd::Vec3r;

// ------------------------------------------------------------------------------------------------ spheres
template <typename T>
struct Sphere {
  Vec3r<T> center;
  T radius;
  unsigned int prim_id;
};

template <typename T>
struct UserSpheres;

template <typename T>
struct SphereIntersector {
  const UserSpheres<T> &collection;
  Vec3r<T> ray_org, ray_dir;
  T min_hit_distance;
  Vec2r<T> uv;
  T hit_distance;
  unsigned int prim_id;
  __device__ explicit SphereIntersector(const UserSpheres<T> &c) : collection(c), prim_id(-1) {}
};

// device view: centers (3 per sphere) and radii in device memory
template <typename T>
struct UserSpheres {
  using intersector = SphereIntersector<T>;
  using primitive_type = Sphere<T>;
  const T *centers;
  const T *radii;
  unsigned int n;
  __device__ Vec3r<T> get_primitive_center(unsigned int i) const {
    return {centers[3ull * i], centers[3ull * i + 1], centers[3ull * i + 2]};
  }
};

template <typename T>
__device__ Sphere<T> primitive_from_collection(const UserSpheres<T> &c, unsigned int i) {
  return {c.get_primitive_center(i), c.radii[i], i};
}

template <typename T>
__device__ void prepare_traversal(SphereIntersector<T> &i, const Ray<T> &ray) {
  i.ray_org = ray.origin;
  i.ray_dir = ray.direction;
  i.min_hit_distance = ray.min_hit_distance;
  i.hit_distance = ray.max_hit_distance;
  i.uv = T(0);
  i.prim_id = static_cast<unsigned int>(-1);
}

template <typename T>
__device__ bool intersect_primitive(SphereIntersector<T> &i, const Sphere<T> &s, const Ray<T> &) {
  const Vec3r<T> v = i.ray_org - s.center;
  const T A = dot(i.ray_dir, i.ray_dir);
  const T B = T(2) * dot(v, i.ray_dir);
  const T C = dot(v, v) - s.radius * s.radius;
  const T D = B * B - T(4) * A * C;
  if (D < T(0)) return false;
  const T Q = brt::exact::dsqrt(D);
  const T rcpA = T(1) / A;
  const T t0 = T(0.5) * rcpA * (-B - Q);
  const T t1 = T(0.5) * rcpA * (-B + Q);
  if (t0 > i.min_hit_distance && t0 < i.hit_distance) {
    i.hit_distance = norm(t0 * i.ray_dir);
    i.prim_id = s.prim_id;
    return true;
  }
  if (t1 > i.min_hit_distance && t1 < i.hit_distance) {
    i.hit_distance = norm(t1 * i.ray_dir);
    i.prim_id = s.prim_id;
    return true;
  }
  return false;
}

template <typename T>
__device__ void post_traversal(SphereIntersector<T> &i, RayHit<T> &rayhit) {
  rayhit.hit_distance = i.hit_distance;
  rayhit.uv = i.uv;
  rayhit.prim_id = i.prim_id;
  const Vec3r<T> center = i.collection.get_primitive_center(i.prim_id);
  rayhit.normal = normalize(i.ray_org + i.hit_distance * i.ray_dir - center);
}

// host side of the same collection, for BVH<T>(collection): boxes center -+ radius (sphere.h:88-99)
template <typename T>
struct HostSpheres {
  const T *centers, *radii;
  unsigned int n;
  unsigned int size() const { return n; }
  Vec3r<T> get_primitive_center(unsigned int i) const {
    return {centers[3ull * i], centers[3ull * i + 1], centers[3ull * i + 2]};
  }
  std::pair<Vec3r<T>, Vec3r<T>> get_primitive_bounding_box(unsigned int i) const {
    const Vec3r<T> c = get_primitive_center(i);
    const T r = radii[i];
    return {Vec3r<T>(c[0] - r, c[1] - r, c[2] - r), Vec3r<T>(c[0] + r, c[1] + r, c[2] + r)};
  }
};

// ------------------------------------------------------------------------------------------------ boxes
struct Box {
  Vec3r<double> lo, hi;
  unsigned int prim_id;
};
struct UserBoxes;
struct BoxIntersector {
  const UserBoxes &collection;
  Vec3r<double> org, inv;
  double min_hit_distance;
  Vec2r<double> uv;
  double hit_distance;
  unsigned int prim_id;
  unsigned int axis;  // entry axis of the current hit
  __device__ explicit BoxIntersector(const UserBoxes &c) : collection(c), prim_id(-1), axis(0) {}
};
struct UserBoxes {
  using intersector = BoxIntersector;
  using primitive_type = Box;
  const double *lo, *hi;  // 3 per box
  unsigned int n;
};

__device__ Box primitive_from_collection(const UserBoxes &c, unsigned int i) {
  return {{c.lo[3ull * i], c.lo[3ull * i + 1], c.lo[3ull * i + 2]},
          {c.hi[3ull * i], c.hi[3ull * i + 1], c.hi[3ull * i + 2]},
          i};
}

__device__ void prepare_traversal(BoxIntersector &i, const Ray<double> &ray) {
  i.org = ray.origin;
  i.inv = ray.direction_inv;
  i.min_hit_distance = ray.min_hit_distance;
  i.hit_distance = ray.max_hit_distance;
  i.uv = 0.0;
  i.prim_id = static_cast<unsigned int>(-1);
}

// entry distance in units of the given direction: t = max_k min(t_lo_k, t_hi_k), hit if t <= exit distance
__device__ bool intersect_primitive(BoxIntersector &i, const Box &b, const Ray<double> &) {
  double tn = -__builtin_inf(), tf = __builtin_inf();
  unsigned int ax = 0;
  for (int k = 0; k < 3; k++) {
    const double a = (b.lo[k] - i.org[k]) * i.inv[k];
    const double c = (b.hi[k] - i.org[k]) * i.inv[k];
    const double lo = a < c ? a : c, hi = a < c ? c : a;
    if (lo > tn) {
      tn = lo;
      ax = k;
    }
    if (hi < tf) tf = hi;
  }
  if (tn <= tf && tn > i.min_hit_distance && tn < i.hit_distance) {
    i.hit_distance = tn;
    i.prim_id = b.prim_id;
    i.axis = ax;
    return true;
  }
  return false;
}

__device__ void post_traversal(BoxIntersector &i, RayHit<double> &rayhit) {
  rayhit.hit_distance = i.hit_distance;
  rayhit.uv = i.uv;
  rayhit.prim_id = i.prim_id;
  rayhit.normal = Vec3r<double>(0.0);
  rayhit.normal[i.axis] = i.inv[i.axis] < 0.0 ? 1.0 : -1.0;
}

template <typename T>
int launch_spheres(brt_bvh *bvh, const T *centers, const T *radii, unsigned int n,
                   const typename blazert_amd::Soa<T>::rays *rays, const typename blazert_amd::Soa<T>::hits *hits,
                   uint64_t nrays, unsigned int any_hit, void *stream) {
  brt_bvh_view v;
  if (int rc = brt_bvh_device_view(bvh, &v)) return rc;
  if (v.real_bytes != sizeof(T) || v.num_prims != n) return BRT_ERR_INVALID_ARGUMENT;
  const UserSpheres<T> coll{centers, radii, n};
  const hipError_t e = blazert_amd::intersect(blazert_amd::DeviceBVH<T>::from_view(v), coll, *rays, *hits, nrays,
                                              any_hit != 0, 0u, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? BRT_OK : BRT_ERR_DEVICE;
}

}  // namespace ug

extern "C" {

// BVH<T>(collection) from host centers / radii through the template builder path
int ug_spheres_build(const double *centers, const double *radii, unsigned int n, const brt_build_options *opts,
                     brt_bvh **out) {
  try {
    blazert_amd::BVH<double> b(ug::HostSpheres<double>{centers, radii, n}, *opts);
    if (b.statistics().num_primitives != n) return BRT_ERR_STATE;
    *out = b.release();  // ownership passes to the caller (brt_bvh_destroy)
    return BRT_OK;
  } catch (...) {
    return BRT_ERR_INVALID_ARGUMENT;
  }
}

int ug_spheres_intersect(brt_bvh *bvh, const double *centers, const double *radii, unsigned int n,
                         const brt_rays *rays, const brt_hits *hits, uint64_t nrays, unsigned int any_hit,
                         void *stream) {
  return ug::launch_spheres<double>(bvh, centers, radii, n, rays, hits, nrays, any_hit, stream);
}

int ug_spheres_intersect_f32(brt_bvh *bvh, const float *centers, const float *radii, unsigned int n,
                             const brt_rays_f32 *rays, const brt_hits_f32 *hits, uint64_t nrays,
                             unsigned int any_hit, void *stream) {
  return ug::launch_spheres<float>(bvh, centers, radii, n, rays, hits, nrays, any_hit, stream);
}

int ug_boxes_intersect(brt_bvh *bvh, const double *lo, const double *hi, unsigned int n, const brt_rays *rays,
                       const brt_hits *hits, uint64_t nrays, void *stream) {
  brt_bvh_view v;
  if (int rc = brt_bvh_device_view(bvh, &v)) return rc;
  if (v.real_bytes != 8 || v.num_prims != n) return BRT_ERR_INVALID_ARGUMENT;
  const hipError_t e = blazert_amd::intersect(blazert_amd::DeviceBVH<double>::from_view(v), ug::UserBoxes{lo, hi, n},
                                              *rays, *hits, nrays, false, 7u, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? BRT_OK : BRT_ERR_DEVICE;
}

}  // extern "C"
