// User-defined primitive types on the GPU: blazeRT's collection concept as a device-callable intersector slot.
//
// In blazeRT a primitive type plugs in through a collection class (examples/geometry_template/GEOM_TEMPLATE.h:
// 19-137): the builder reads size(), get_primitive_bounding_box(i) and get_primitive_center(i)
// (blazert/bvh/builder.h:48-82, binbuffer.h:79-80), and traverse() (blazert/bvh/accel.h:80-120) drives the
// collection's intersector through free functions found by argument-dependent lookup. The same split here:
//
//   host    BVH<T> is built from any collection with those three members through the C ABI (brt_bvh_build,
//           brt_bvh_build_f32): the reference's tree for the collection's boxes and build options.
//   device  the collection is a trivially copyable view (device pointers to the primitives' parameters) with
//             typename Collection::intersector               constructible on the device from the collection
//             primitive_from_collection(collection, idx)     -> the primitive (carrying its prim_id)
//             prepare_traversal(intersector &, const Ray<T> &)
//             intersect_primitive(intersector &, const primitive &, const Ray<T> &) -> bool
//             post_traversal(intersector &, RayHit<T> &)
//           all __device__, the intersector holding hit_distance (the current closest hit), as GEOMIntersector.
//           traverse(bvh, collection, ray, rayhit) visits nodes and primitives in the reference's order with the
//           reference's slab test, so a collection whose functions repeat the reference's arithmetic returns the
//           reference's hits bit for bit; intersect() launches it over structure-of-arrays device rays.
//
// Build the including translation unit with hipcc --offload-arch=gfx950 -ffp-contract=off (so the collection's own
// arithmetic rounds like the reference's separate multiplies and adds) and link libblazert_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../blazert_amd.h"
#include "device/exact.h"
#include "device/nodes.h"

namespace blazert_amd {

// ---------------------------------------------------------------------------------------------- small vectors
// Vec3r<T> / Vec2r<T> / Vec3ui of blazert/datatypes.h:31-59 for device code, with blaze's arithmetic: element-wise
// operators, dot summed left to right, v / s evaluated as v * (1 / s), normalize(v) = v * (1 / |v|) (0 for |v| = 0).
// No operation is fused into a multiply-add, whatever the translation unit's -ffp-contract setting.
template <typename T>
struct Vec3r {
  T v[3];
  __host__ __device__ Vec3r() : v{T(0), T(0), T(0)} {}
  __host__ __device__ Vec3r(T x, T y, T z) : v{x, y, z} {}
  __host__ __device__ explicit Vec3r(T s) : v{s, s, s} {}
  __host__ __device__ T &operator[](int i) { return v[i]; }
  __host__ __device__ const T &operator[](int i) const { return v[i]; }
};

template <typename T>
struct Vec2r {
  T v[2];
  __host__ __device__ Vec2r() : v{T(0), T(0)} {}
  __host__ __device__ Vec2r(T x, T y) : v{x, y} {}
  __host__ __device__ explicit Vec2r(T s) : v{s, s} {}
  __host__ __device__ Vec2r &operator=(T s) {
    v[0] = v[1] = s;
    return *this;
  }
  __host__ __device__ T &operator[](int i) { return v[i]; }
  __host__ __device__ const T &operator[](int i) const { return v[i]; }
};

struct Vec3ui {
  unsigned int v[3];
  __host__ __device__ unsigned int &operator[](int i) { return v[i]; }
  __host__ __device__ const unsigned int &operator[](int i) const { return v[i]; }
};

template <typename T>
__host__ __device__ inline Vec3r<T> operator+(const Vec3r<T> &a, const Vec3r<T> &b) {
#pragma clang fp contract(off)
  return {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator-(const Vec3r<T> &a, const Vec3r<T> &b) {
#pragma clang fp contract(off)
  return {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator-(const Vec3r<T> &a) {
#pragma clang fp contract(off)
  return {-a[0], -a[1], -a[2]};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator*(const Vec3r<T> &a, const Vec3r<T> &b) {
#pragma clang fp contract(off)
  return {a[0] * b[0], a[1] * b[1], a[2] * b[2]};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator*(T s, const Vec3r<T> &a) {
#pragma clang fp contract(off)
  return {s * a[0], s * a[1], s * a[2]};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator*(const Vec3r<T> &a, T s) {
#pragma clang fp contract(off)
  return {a[0] * s, a[1] * s, a[2] * s};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator/(const Vec3r<T> &a, T s) {
#pragma clang fp contract(off)
  const T is = T(1) / s;
  return {a[0] * is, a[1] * is, a[2] * is};
}
template <typename T>
__host__ __device__ inline Vec3r<T> operator/(T s, const Vec3r<T> &a) {
#pragma clang fp contract(off)
  return {s / a[0], s / a[1], s / a[2]};
}
template <typename T>
__host__ __device__ inline T dot(const Vec3r<T> &a, const Vec3r<T> &b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <typename T>
__host__ __device__ inline Vec3r<T> cross(const Vec3r<T> &a, const Vec3r<T> &b) {
#pragma clang fp contract(off)
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <typename T>
__device__ inline T norm(const Vec3r<T> &a) {
#pragma clang fp contract(off)
  return brt::exact::dsqrt(dot(a, a));
}
template <typename T>
__device__ inline T length(const Vec3r<T> &a) {
#pragma clang fp contract(off)
  return norm(a);
}
template <typename T>
__device__ inline Vec3r<T> normalize(const Vec3r<T> &a) {
#pragma clang fp contract(off)
  const T len = norm(a);
  const T il = (len != T(0)) ? T(1) / len : T(0);
  return {a[0] * il, a[1] * il, a[2] * il};
}

// ---------------------------------------------------------------------------------------------- rays and hits
// blazert/ray.h:23-66: direction is normalized, direction_inv = 1 / the direction as given, direction_sign[k] =
// (given direction[k] < 0). `slab` holds the same ray in the form the node test takes.
template <typename T>
struct Ray {
  enum class CullBackFace { yes, no };
  enum class AnyHit { yes, no };
  Vec3r<T> origin;
  Vec3r<T> direction;
  Vec3r<T> direction_inv;
  Vec3ui direction_sign;
  T min_hit_distance;
  T max_hit_distance;
  CullBackFace cull_back_face;
  AnyHit any_hit;
  brt::exact::RayT<T> slab;

  __device__ Ray(const Vec3r<T> &org, const Vec3r<T> &dir, T min_hit = T(0),
                 T max_hit = std::numeric_limits<T>::max(), CullBackFace cull = CullBackFace::no,
                 AnyHit any = AnyHit::no)
      : min_hit_distance(min_hit), max_hit_distance(max_hit), cull_back_face(cull), any_hit(any) {
    brt::exact::setup_ray(org.v, dir.v, min_hit, max_hit, slab);
    for (int k = 0; k < 3; k++) {
      origin[k] = slab.o[k];
      direction[k] = slab.d[k];
      direction_inv[k] = slab.inv[k];
      direction_sign[k] = (slab.sign >> k) & 1u;
    }
  }
};

// blazert/ray.h:70-85
template <typename T>
struct RayHit {
  Vec3r<T> normal;
  Vec2r<T> uv;
  T hit_distance = std::numeric_limits<T>::max();
  unsigned int prim_id = static_cast<unsigned int>(-1);
  unsigned int geom_id = static_cast<unsigned int>(-1);
};

// ---------------------------------------------------------------------------------------------- device tree
// What a kernel needs of a built tree (passed by value). T = double reads brt::Node64 records, T = float Node32.
template <typename T>
struct DeviceBVH {
  using Node = typename std::conditional<sizeof(T) == 8, brt::Node64, brt::Node32>::type;
  const Node *nodes = nullptr;
  const uint32_t *prim_ids = nullptr;  // leaf position -> collection primitive index
  T root[6] = {};                      // root box: min xyz, max xyz
  uint32_t root_ref = 0, root_cnt = 0;
  uint32_t num_prims = 0;
  uint32_t irregular_boxes = 0;  // some node box has a NaN coordinate or min > max

  // from the C ABI's view; throws if the view's real type is not T
  static DeviceBVH from_view(const brt_bvh_view &v) {
    if (v.real_bytes != sizeof(T)) throw std::invalid_argument("blazert_amd::DeviceBVH: tree built for another T");
    DeviceBVH b;
    b.nodes = static_cast<const Node *>(v.nodes);
    b.prim_ids = v.prim_ids;
    for (int k = 0; k < 3; k++) {  // an FP32 tree's box values are exact in double: the narrowing is exact
      b.root[k] = static_cast<T>(v.root_min[k]);
      b.root[3 + k] = static_cast<T>(v.root_max[k]);
    }
    b.root_ref = v.root_ref;
    b.root_cnt = v.root_cnt;
    b.num_prims = v.num_prims;
    b.irregular_boxes = v.irregular_boxes;
    return b;
  }
};

// Deepest traversal stack: trees are limited to depth 63 at upload and the DFS pushes at most one entry per level.
constexpr int kUserStack = 64;

// ---------------------------------------------------------------------------------------------- stacks
// Entries are (child reference, leaf count, key): the key is the entry distance the current hit distance must not
// be below when the entry is popped.
template <typename T>
struct StackArea {  // private memory
  T key[kUserStack];
  uint32_t ref[kUserStack], cnt[kUserStack];
};

// Every entry in private memory: usable from any kernel.
template <typename T>
struct PrivateStack {
  StackArea<T> *a;
  int sp = 0;
  __device__ explicit PrivateStack(StackArea<T> *area) : a(area) {}
  __device__ bool empty() const { return sp == 0; }
  __device__ void push(uint32_t r, uint32_t c, T k) {
    a->key[sp] = k;
    a->ref[sp] = r;
    a->cnt[sp] = c;
    sp++;
  }
  __device__ void pop(uint32_t &r, uint32_t &c, T &k) {
    sp--;
    r = a->ref[sp];
    c = a->cnt[sp];
    k = a->key[sp];
  }
};

// The newest RING entries of a lane in LDS (slot-major, lane-minor, so a wave's accesses spread over the banks),
// older ones spilled to private memory, which only rays deeper than the ring touch. Entries [lo, sp) are in LDS,
// [0, lo) in the spill area. RING is a power of two. The LDS pointers carry the LDS address space: as generic
// pointers the compiler merges the LDS and spill reads of a pop into flat loads.
template <typename T>
using LdsPtr = __attribute__((address_space(3))) T *;

template <typename T, int RING>
struct RingStack {
  LdsPtr<T> lkey;  // this lane's slot 0; slot j is `lanes` elements further
  LdsPtr<uint32_t> lref, lcnt;
  uint32_t lanes;
  StackArea<T> *spill;
  int sp = 0, lo = 0;
  __device__ bool empty() const { return sp == 0; }
  __device__ void push(uint32_t r, uint32_t c, T k) {
    if (sp - lo == RING) {
      const uint32_t j = (uint32_t)(lo & (RING - 1)) * lanes;
      spill->key[lo] = lkey[j];
      spill->ref[lo] = lref[j];
      spill->cnt[lo] = lcnt[j];
      lo++;
    }
    const uint32_t j = (uint32_t)(sp & (RING - 1)) * lanes;
    lkey[j] = k;
    lref[j] = r;
    lcnt[j] = c;
    sp++;
  }
  __device__ void pop(uint32_t &r, uint32_t &c, T &k) {
    sp--;
    if (sp < lo) {
      lo = sp;
      r = spill->ref[sp];
      c = spill->cnt[sp];
      k = spill->key[sp];
    } else {
      const uint32_t j = (uint32_t)(sp & (RING - 1)) * lanes;
      r = lref[j];
      c = lcnt[j];
      k = lkey[j];
    }
  }
};

// ---------------------------------------------------------------------------------------------- traversal
// traverse() of blazert/bvh/accel.h:80-120 with an explicit intersector and stack. The reference pops a node, tests
// it against [min_hit_distance, current hit distance] and pushes far then near child; here the near child is tested
// at its parent (the reference tests it at the very next pop, with the same hit distance) and the far child is
// pushed with the hit-independent part of its test done, so the pop-time check reproduces the reference's decision
// exactly. Leaves test their primitives in leaf order (blazert/bvh/intersect.h:66-78); the hit distance used for
// later node tests is the intersector's after a leaf reported a hit.
template <typename T, typename Collection, typename Stack>
__device__ bool traverse_with(const DeviceBVH<T> &bvh, const Collection &collection, const Ray<T> &ray,
                              RayHit<T> &rayhit, typename Collection::intersector &intersector, Stack &stack) {
#pragma clang fp contract(off)
  using brt::exact::slab_exact;
  using Node = typename DeviceBVH<T>::Node;
  const bool any_hit = ray.any_hit == Ray<T>::AnyHit::yes;
  T hit_distance = ray.max_hit_distance;
  prepare_traversal(intersector, ray);
  if (bvh.num_prims) {
    brt::exact::RayT<T> r = ray.slab;
    if (bvh.irregular_boxes) r.sign &= ~brt::exact::RAY_NAN_FREE;
    T key = T(0);
    bool go = slab_exact(bvh.root, r, key) && !(hit_distance < key);
    uint32_t cur = bvh.root_ref, cnt = bvh.root_cnt;
    while (go) {
      if (!(cur & brt::LEAF_BIT)) {
        const Node nd = bvh.nodes[cur];
        const uint32_t axis = nd.cnt[1] >> brt::AXIS_SHIFT;
        const uint32_t nc = (r.sign >> axis) & 1u;  // direction_sign[axis]: child 1 is the near one
        T k0 = T(0), k1 = T(0);
        const bool p0 = slab_exact(nd.b, r, k0);
        const bool p1 = slab_exact(nd.b + 6, r, k1);
        const uint32_t c0 = nd.cnt[0] & brt::CNT_MASK, c1 = nd.cnt[1] & brt::CNT_MASK;
        if (nc ? p0 : p1) stack.push(nc ? nd.ref[0] : nd.ref[1], nc ? c0 : c1, nc ? k0 : k1);
        if ((nc ? p1 : p0) && !(hit_distance < (nc ? k1 : k0))) {
          cur = nc ? nd.ref[1] : nd.ref[0];
          cnt = nc ? c1 : c0;
          continue;
        }
      } else {
        const uint32_t first = cur & ~brt::LEAF_BIT;
        bool hit = false;
        for (uint32_t k = 0; k < cnt; k++) {
          hit |= intersect_primitive(intersector, primitive_from_collection(collection, bvh.prim_ids[first + k]),
                                     ray);
          if (hit && any_hit) break;
        }
        if (hit) {
          hit_distance = intersector.hit_distance;
          if (any_hit) break;
        }
      }
      go = false;
      while (!stack.empty()) {
        stack.pop(cur, cnt, key);
        if (!(hit_distance < key)) {
          go = true;
          break;
        }
      }
    }
  }
  const bool hit = intersector.hit_distance < ray.max_hit_distance;
  if (hit) post_traversal(intersector, rayhit);
  return hit;
}

// The same with a private-memory stack: callable from any kernel.
template <typename T, typename Collection>
__device__ bool traverse(const DeviceBVH<T> &bvh, const Collection &collection, const Ray<T> &ray, RayHit<T> &rayhit,
                         typename Collection::intersector &intersector) {
  StackArea<T> area;
  PrivateStack<T> stack(&area);
  return traverse_with(bvh, collection, ray, rayhit, intersector, stack);
}

// traverse(bvh, ray, rayhit) of blazert/bvh/accel.h:60-63: a fresh intersector of the collection.
template <typename T, typename Collection>
__device__ bool traverse(const DeviceBVH<T> &bvh, const Collection &collection, const Ray<T> &ray,
                         RayHit<T> &rayhit) {
  typename Collection::intersector intersector(collection);
  return traverse(bvh, collection, ray, rayhit, intersector);
}

// ---------------------------------------------------------------------------------------------- batched launch
template <typename T>
struct Soa;
template <>
struct Soa<double> {
  using rays = brt_rays;
  using hits = brt_hits;
};
template <>
struct Soa<float> {
  using rays = brt_rays_f32;
  using hits = brt_hits_f32;
};

// One ray per thread, blocks of kBlock threads, the newest kRing stack entries of each ray in LDS. Hits follow
// brt_hits: a hit writes the RayHit fields and geom_id; a miss writes t = max, u = v = 0, ids BRT_INVALID_ID,
// normal 0. u / v / geom_id / normal pointers may be NULL.
constexpr uint32_t kBlock = 256;
constexpr int kRing = 8;
template <typename T, typename Collection>
__global__ void __launch_bounds__(kBlock)
    intersect_kernel(DeviceBVH<T> bvh, Collection collection, typename Soa<T>::rays rays,
                     typename Soa<T>::hits hits, uint64_t n, uint32_t any_hit, uint32_t geom_id) {
  __shared__ T lkey[kRing * kBlock];
  __shared__ uint32_t lref[kRing * kBlock], lcnt[kRing * kBlock];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Ray<T> ray(Vec3r<T>(rays.ox[i], rays.oy[i], rays.oz[i]), Vec3r<T>(rays.dx[i], rays.dy[i], rays.dz[i]),
                   rays.tmin ? rays.tmin[i] : T(0), rays.tmax ? rays.tmax[i] : std::numeric_limits<T>::max(),
                   Ray<T>::CullBackFace::no, any_hit ? Ray<T>::AnyHit::yes : Ray<T>::AnyHit::no);
  RayHit<T> rh;
  typename Collection::intersector intersector(collection);
  StackArea<T> spill;
  RingStack<T, kRing> stack{(LdsPtr<T>)(lkey + threadIdx.x), (LdsPtr<uint32_t>)(lref + threadIdx.x),
                            (LdsPtr<uint32_t>)(lcnt + threadIdx.x), kBlock, &spill};
  const bool hit = traverse_with(bvh, collection, ray, rh, intersector, stack);
  hits.t[i] = hit ? rh.hit_distance : std::numeric_limits<T>::max();
  if (hits.u) hits.u[i] = hit ? rh.uv[0] : T(0);
  if (hits.v) hits.v[i] = hit ? rh.uv[1] : T(0);
  hits.prim_id[i] = hit ? rh.prim_id : BRT_INVALID_ID;
  if (hits.geom_id) hits.geom_id[i] = hit ? geom_id : BRT_INVALID_ID;
  if (hits.nx) {
    hits.nx[i] = hit ? rh.normal[0] : T(0);
    hits.ny[i] = hit ? rh.normal[1] : T(0);
    hits.nz[i] = hit ? rh.normal[2] : T(0);
  }
}

// Traces n rays (device memory) against the collection; asynchronous on `stream`.
template <typename T, typename Collection>
hipError_t intersect(const DeviceBVH<T> &bvh, const Collection &collection, const typename Soa<T>::rays &rays,
                     const typename Soa<T>::hits &hits, uint64_t n, bool any_hit = false, unsigned int geom_id = 0,
                     hipStream_t stream = nullptr) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((intersect_kernel<T, Collection>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, bvh,
                     collection, rays, hits, n, any_hit ? 1u : 0u, geom_id);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------- host tree
// blazert::BVH<T, Collection> + SAHBinnedBuilder::build: the tree of a collection's boxes. Owns the brt_bvh handle.
template <typename T>
class BVH {
 public:
  // Any collection with size(), get_primitive_bounding_box(i) (a pair of 3-vectors) and get_primitive_center(i).
  template <typename HostCollection>
  explicit BVH(const HostCollection &collection, const brt_build_options &options = brt_build_options{4, 28, 64}) {
    const uint32_t n = static_cast<uint32_t>(collection.size());
    std::vector<T> lo(3ull * n), hi(3ull * n), ctr(3ull * n);
    for (uint32_t i = 0; i < n; i++) {
      const auto box = collection.get_primitive_bounding_box(i);
      const auto c = collection.get_primitive_center(i);
      for (int k = 0; k < 3; k++) {
        lo[3ull * i + k] = box.first[k];
        hi[3ull * i + k] = box.second[k];
        ctr[3ull * i + k] = c[k];
      }
    }
    build(lo.data(), hi.data(), ctr.data(), n, options);
  }
  // From arrays: 3 values per primitive (box minimum, box maximum, center), host memory.
  BVH(const T *box_min, const T *box_max, const T *centers, uint32_t n,
      const brt_build_options &options = brt_build_options{4, 28, 64}) {
    build(box_min, box_max, centers, n, options);
  }
  BVH(const BVH &) = delete;
  BVH &operator=(const BVH &) = delete;
  BVH(BVH &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  ~BVH() {
    if (h_) brt_bvh_destroy(h_);
  }

  // Device view on the current device (the tree is uploaded on the first call).
  DeviceBVH<T> device() const {
    brt_bvh_view v;
    check(brt_bvh_device_view(h_, &v), "brt_bvh_device_view");
    return DeviceBVH<T>::from_view(v);
  }
  brt_build_stats statistics() const {
    brt_build_stats s;
    check(brt_bvh_build_stats(h_, &s), "brt_bvh_build_stats");
    return s;
  }
  brt_bvh *handle() const { return h_; }
  // gives up ownership of the handle (the caller then calls brt_bvh_destroy)
  brt_bvh *release() { return std::exchange(h_, nullptr); }

 private:
  brt_bvh *h_ = nullptr;
  static void check(int rc, const char *what) {
    if (rc != BRT_OK) throw std::runtime_error(std::string(what) + ": " + brt_last_error());
  }
  void build(const T *lo, const T *hi, const T *ctr, uint32_t n, const brt_build_options &o) {
    if constexpr (std::is_same<T, double>::value) check(brt_bvh_build(lo, hi, ctr, n, &o, &h_), "brt_bvh_build");
    else check(brt_bvh_build_f32(lo, hi, ctr, n, &o, &h_), "brt_bvh_build_f32");
  }
};

}  // namespace blazert_amd
