// Device-resident scene layout shared by the host builder (scene.cpp) and the HIP kernels.
//
// One BVH per geometry type, flattened from the reference topology (binary binned-SAH tree numbered in
// pre-order, blazert/bvh/builder.h:84-115) into "children in parent" records: an internal node holds the
// two child boxes, so one record fetch drives both slab tests of a traversal step. The root box is kept
// in the descriptor. Leaf children are encoded in the child reference (LEAF_BIT | first primitive) with
// the primitive count beside it; primitives are stored contiguously in leaf order.
#pragma once
#include <stdint.h>

#include <limits>

#include "../../include/blazert_amd/device/nodes.h"

namespace brt {

// n = cross(e1, e2) of a triangle record, component order of the reference's test (trimesh.h:160)
template <class R>
inline __host__ __device__ void tri_normal(const R *e1, const R *e2, R *n) {
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

// USER_COLLECTION: a tree over a user collection's boxes (brt_bvh_build), with no primitive records
enum GeomType : uint32_t { TRIANGLES = 0, SPHERES = 1, PLANES = 2, CYLINDERS = 3, NUM_TYPES = 4, USER_COLLECTION = 16 };

// FP32 4-wide node for split-precision traversal, 128 B: the binary tree collapsed greedily (starting from a node's
// two children, the internal slot of largest surface area is replaced by its own two children until four slots are
// filled or all are leaves), so one record holds up to four child boxes, rounded outward (min down, max up). Planes are child-minor: lo[axis][child]. A child reference is a
// Node4 index or LEAF_BIT | first primitive with the leaf count beside it; unused slots have cnt = EMPTY_SLOT.
constexpr uint32_t EMPTY_SLOT = 0xFFFFFFFFu;
struct alignas(128) Node4 {
  float lo[3][4];
  float hi[3][4];
  uint32_t ref[4];
  uint32_t cnt[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// Triangle record in leaf order: v0, e1 = v0 - v1, e2 = v2 - v0 (blazert/primitives/trimesh.h:26-27) and the user
// primitive id (prim_ids at this leaf position), which a hit takes from the record it has just read instead of a
// dependent lookup when the ray finishes. The FP64 record does not carry n = cross(e1, e2): the 16 more bytes per
// record cost more time (measured) than forming n in every test, as the reference does.
struct alignas(16) TriRec {
  using real = double;
  static constexpr uint32_t kind = TRIANGLES;
  static constexpr bool has_normal = false;
  double v0[3];
  double e1[3];
  double e2[3];
  uint64_t prim;
};
static_assert(sizeof(TriRec) == 80, "TriRec size");

// Split precision: FP32 pre-test record of a triangle, in leaf order beside the TriRec array: v0, e1, e2 and the
// FP64 test's n = cross(e1, e2), each rounded to nearest, and the constants of the pre-test's error bound, rounded
// up from FP64: vmax >= max |v0_i|, kua = 16 u sum |e1_i|, kub = 16 u sum |e2_i|, ed = 10 u sum |e1_i| sum |e2_i|
// + 1e-29 (u = 2^-24). Built on the device from the TriRec array; FILTER_PAD records of padding follow the last one,
// so a group of FILTER_GROUP records read from any leaf start stays in bounds.
struct alignas(16) TriFilt {
  float v0[3];
  float e1[3];
  float e2[3];
  float n[3];
  float vmax, kua, kub, ed;
};
static_assert(sizeof(TriFilt) == 64, "TriFilt size");
#ifndef BRT_FILTER_GROUP
#define BRT_FILTER_GROUP 2
#endif
constexpr uint32_t FILTER_GROUP = BRT_FILTER_GROUP;  // triangles pre-tested per traversal step
constexpr uint32_t FILTER_PAD = FILTER_GROUP - 1;

// Split precision: FP32 pre-test record of a sphere, in leaf order beside the SphRec array: the center rounded to
// nearest and |r| rounded up; for a cylinder, its bounding sphere, enlarged (k_cyl_filter_records). Built on the
// device from the SphRec / CylRec array; SPH_FILTER_PAD zero records follow the last one, so a group of
// SPH_FILTER_GROUP records read from any leaf start stays in bounds.
struct alignas(16) SphFilt {
  float c[3];
  float r;
};
static_assert(sizeof(SphFilt) == 16, "SphFilt size");
#ifndef BRT_SPH_FILTER_GROUP
#define BRT_SPH_FILTER_GROUP 4
#endif
constexpr uint32_t SPH_FILTER_GROUP = BRT_SPH_FILTER_GROUP;  // spheres pre-tested per traversal step
constexpr uint32_t SPH_FILTER_PAD = SPH_FILTER_GROUP - 1;
// Largest scene extent (and ray origin magnitude) for which the FP32 pre-test is used: below it no product of the
// pre-test or of its error bound can overflow FP32.
constexpr double FILTER_MAX_EXTENT = 1e12;

struct alignas(32) SphRec {
  using real = double;
  static constexpr uint32_t kind = SPHERES;
  double c[3];
  double r;
};
static_assert(sizeof(SphRec) == 32, "SphRec size");

// Plane: center, extents and the row-major rotation (blazert/primitives/plane.h:19-39).
struct alignas(128) PlaRec {
  using real = double;
  static constexpr uint32_t kind = PLANES;
  double c[3];
  double dx, dy;
  double R[9];
  double pad[2];
};
static_assert(sizeof(PlaRec) == 128, "PlaRec size");

// Cylinder: center, semi axes a/b, height and the row-major rotation (blazert/primitives/cylinder.h:19-40).
struct alignas(128) CylRec {
  using real = double;
  static constexpr uint32_t kind = CYLINDERS;
  double c[3];
  double a, b, h;
  double R[9];
  double pad;
};
static_assert(sizeof(CylRec) == 128, "CylRec size");

// Scene<float> (the reference instantiated with T = float): Node32 records (FP32 boxes) and FP32 primitive
// records, traversed and tested in FP32 throughout.
// The FP32 record keeps n = cross(e1, e2) in what would be padding, formed at build time with the expression of the
// reference's test (trimesh.h:160), so its bits are the ones the test would form.
struct alignas(16) TriRecF {
  using real = float;
  static constexpr uint32_t kind = TRIANGLES;
  static constexpr bool has_normal = true;
  float v0[3];
  float e1[3];
  float e2[3];
  float n[3];
};
static_assert(sizeof(TriRecF) == 48, "TriRecF size");

struct alignas(16) SphRecF {
  using real = float;
  static constexpr uint32_t kind = SPHERES;
  float c[3];
  float r;
};
static_assert(sizeof(SphRecF) == 16, "SphRecF size");

struct alignas(64) PlaRecF {
  using real = float;
  static constexpr uint32_t kind = PLANES;
  float c[3];
  float dx, dy;
  float R[9];
  float pad[2];
};
static_assert(sizeof(PlaRecF) == 64, "PlaRecF size");

struct alignas(64) CylRecF {
  using real = float;
  static constexpr uint32_t kind = CYLINDERS;
  float c[3];
  float a, b, h;
  float R[9];
  float pad;
};
static_assert(sizeof(CylRecF) == 64, "CylRecF size");

// Kernel-side view of one uploaded geometry BVH (passed by value).
struct BvhView {
  const Node64 *nodes64;
  const Node4 *nodes4;
  const void *prims;          // TriRec / SphRec / PlaRec / CylRec in leaf order
  const uint32_t *prim_ids;   // leaf order -> user primitive id
  const uint32_t *faces;      // triangles only: user faces (3 per primitive id)
  const double *vnormals;     // triangles only: vertex normals (3 per vertex)
  const double *vertices;     // triangles only: user vertices (3 per vertex), for mirror tracing and its adjoint
  const Node32 *nodes32;      // Scene<float>: FP32 nodes (nodes64 / nodes4 unused)
  const TriFilt *tri_filter;  // split precision, triangles: FP32 pre-test records (NULL: every test in FP64)
  const SphFilt *sph_filter;  // split precision, spheres and cylinders: FP32 (bounding) spheres (NULL: all in FP64)
  const float *vnormals32;    // Scene<float> triangles: vertex normals
  double root_min[3], root_max[3];
  float root_min32[3], root_max32[3];
  uint32_t root_ref, root_cnt;
  uint32_t root_ref4;         // root as a Node4 reference (index 0, or the same leaf as root_ref)
  uint32_t geom_id;
  uint32_t n_prims;
  uint32_t n_nodes;
  uint32_t depth;             // deepest level of the built tree
  uint32_t irregular_boxes;   // some node box has a NaN coordinate or min > max (disables the NaN-free slab path)
};

// Structure-of-arrays rays / hits in device memory (mirrors brt_rays / brt_hits of the C ABI, and their FP32
// forms for Scene<float>). tmin / tmax may be NULL: every ray then uses tmin0 / tmax0 (0 and the type's max).
template <class R>
struct RaysViewT {
  const R *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
  R tmin0 = R(0);
  R tmax0 = std::numeric_limits<R>::max();
};
template <class R>
struct HitsViewT {
  R *t, *u, *v;
  uint32_t *prim_id, *geom_id;
  R *nx, *ny, *nz;
  uint32_t *merged;  // optional: set to 1 where a geometry type's hit was merged (intersect1's return value)
};
using RaysView = RaysViewT<double>;
using HitsView = HitsViewT<double>;
using RaysViewF = RaysViewT<float>;
using HitsViewF = HitsViewT<float>;

}  // namespace brt
