// Host-side interface between the C ABI (scene.cpp) and the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/blazert_amd.h"
#include "layout.h"

namespace brt {

// Slots of the Node4 record rooted at binary internal node b (pre-order tree T): start from b's two children and,
// while fewer than four slots are filled, replace the internal slot of largest surface area (the first of equal
// ones) by its two children in place, so the slots stay in the binary tree's left-to-right order. The internal slots
// are the roots of further records. Shared by the host flatten and the device build, which give identical records.
// A record advances a root-to-leaf path by at least one binary level and pushes at most three stack entries.
__host__ __device__ inline int collapse4(const brt_tree_node *T, uint32_t b, uint32_t slot[4]) {
#pragma clang fp contract(off)
  slot[0] = T[b].children[0];
  slot[1] = T[b].children[1];
  int m = 2;
  while (m < 4) {
    int best = -1;
    double area = 0.0;
    for (int c = 0; c < m; c++) {
      const brt_tree_node &x = T[slot[c]];
      if (x.leaf) continue;
      const double dx = x.max[0] - x.min[0], dy = x.max[1] - x.min[1], dz = x.max[2] - x.min[2];
      const double a = dx * dy + dy * dz + dz * dx;
      if (best < 0 || a > area) {
        best = c;
        area = a;
      }
    }
    if (best < 0) break;
    const uint32_t e = slot[best];
    for (int c = m; c > best + 1; c--) slot[c] = slot[c - 1];
    slot[best] = T[e].children[0];
    slot[best + 1] = T[e].children[1];
    m++;
  }
  return m;
}

// Device counters of one traversal launch, in a group cleared by the launch itself: the persistent schedule's ray
// pool; a flag a split-precision launch sets when it defers rays to the exact path; the ray pool of the FP64 launch
// that traces them. Launches on one stream run in order, so one group per stream serves all of them.
// The ray pools take claims from CLAIM_COUNTERS counters, each on its own 128-byte line (COUNTER_STRIDE slots apart):
// slot k * COUNTER_STRIDE + COUNTER_POOL for the launch, + COUNTER_DEFERRED_POOL for its deferred-ray launch.
constexpr int COUNTER_POOL = 0;
constexpr int COUNTER_DEFERRED = 1;
constexpr int COUNTER_DEFERRED_POOL = 2;
constexpr int CLAIM_COUNTERS = 8;
constexpr int COUNTER_STRIDE = 16;
constexpr int COUNTER_GROUP = CLAIM_COUNTERS * COUNTER_STRIDE;  // slots per group (1 KiB)

constexpr uint32_t SCHEDULE_PERSISTENT = 0;  // persistent waves with per-lane ray refill (default)
constexpr uint32_t SCHEDULE_PER_THREAD = 1;  // one ray per thread, grid-stride

struct TraceLaunch {
  uint32_t precision;  // BRT_PRECISION_*
  uint32_t flags;      // BRT_FLAG_*
  uint32_t block;      // threads per block
  bool first;          // first geometry type of the scene: initialises every hit record
  uint32_t schedule;   // SCHEDULE_*
  uint32_t refill;     // persistent: refill a wave once this many lanes are idle (1..64)
  unsigned long long *counter;  // COUNTER_GROUP device counters for this launch (cleared on the stream)
  const uint32_t *order;        // optional ray permutation (sorted launches); kernels trace rays[order[k]]
  uint32_t ring = 0;            // LDS stack slots per lane: 4, 8 or 16 (0 = default of the precision)
  uint32_t pool = 0;            // persistent: ray indices per claim, 32, 64 or 128 (0 = 64)
  bool overwrite = false;       // every hit replaces the record (traverse semantics), not only closer ones
};

// One (scene, geometry type, ray batch) of a multi-scene launch; jobs are sorted by offset.
struct BatchJob {
  BvhView bvh;
  RaysView rays;
  HitsView hits;
  uint64_t n;       // rays of this job
  uint64_t offset;  // prefix sum of n over the preceding jobs
  uint32_t first;   // this geometry type is the first one of the job's scene
  uint32_t is_tri;
  double extent;    // split precision: scene extent
  const uint32_t *order;
};

// Enqueues one persistent launch over all jobs of one geometry type (total = sum of n).
hipError_t launch_trace_batch(GeomType type, const BatchJob *jobs, uint32_t njobs, uint64_t total,
                              const TraceLaunch &cfg, hipStream_t stream);

// Device buffers of one ray sort (keys, values, radix-sort temporary storage).
struct SortWorkspace {
  void *base = nullptr;
  uint32_t *keys_in = nullptr, *keys_out = nullptr, *vals_in = nullptr, *vals_out = nullptr;
  void *temp = nullptr;
  size_t temp_bytes = 0;
  uint64_t capacity = 0;
};

// Grows ws to hold n rays (frees the previous buffers; not capturable in a graph).
hipError_t sort_workspace_reserve(SortWorkspace &ws, uint64_t n);
// Lays out a workspace for n rays in caller memory [base, base + bytes) without allocating; hipErrorInvalidValue
// when it does not fit.
hipError_t sort_workspace_in(SortWorkspace &ws, void *base, size_t bytes, uint64_t n);

// Device scratch of the allocation-free intersect (brt_intersect_noalloc): a file-scope __device__ array, one copy
// per GPU, holding a launch counter group in its first 256 bytes and then the workspace of a ray sort of up to
// NOALLOC_MAX_SORTED_RAYS rays.
constexpr uint64_t NOALLOC_MAX_SORTED_RAYS = 1ull << 20;
constexpr size_t NOALLOC_SCRATCH_BYTES = 24ull << 20;
// This device's address of the scratch array (looked up once per device with hipGetSymbolAddress).
hipError_t noalloc_scratch(unsigned char **ptr);
// Computes the sort keys (mode 1: direction octant + origin Morton code, mode 2: Morton only) and radix-sorts the
// ray indices on `stream`; *order points into ws afterwards.
hipError_t sort_rays(const RaysView &rays, uint64_t n, const double lo[3], const double hi[3], uint32_t mode,
                     SortWorkspace &ws, hipStream_t stream, const uint32_t **order);

// Per-scene constants for the split-precision path: the largest finite |coordinate| of any node box.
struct SplitParams {
  double scene_extent;
};

// Split precision: writes the FP32 pre-test records (TriFilt, n + FILTER_PAD of them) of n leaf-order TriRecs.
hipError_t launch_tri_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream);
// Split precision: writes the FP32 pre-test records (SphFilt, n + SPH_FILTER_PAD of them) of n leaf-order SphRecs.
hipError_t launch_sph_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream);
// Split precision: writes the bounding-sphere pre-test records (SphFilt, n + SPH_FILTER_PAD) of n leaf-order CylRecs.
hipError_t launch_cyl_filter_records(const void *recs, uint32_t n, void *out, hipStream_t stream);

// Enqueues the traversal of one geometry type over n rays on `stream` (asynchronous).
hipError_t launch_trace(GeomType type, const BvhView &bvh, const RaysView &rays, const HitsView &hits, uint64_t n,
                        const TraceLaunch &cfg, const SplitParams &split, hipStream_t stream);

// Mirror tracing: the next bounce's rays from this bounce's hits (positions o + t normalize(d), mirror directions;
// NaN for misses). pos / dir are x, y, z arrays of n doubles.
hipError_t launch_mirror_step(const RaysView &rays, const HitsView &hits, const double *vertices,
                              const uint32_t *faces, uint64_t n, double *const pos[3], double *const dir[3],
                              hipStream_t stream);

// Upstream gradients in, parameter gradients out, of the mirror-tracing adjoint (any pointer may be NULL).
struct MirrorGrads {
  const double *gt;              // dL/dt, bounces * n
  const double *gpx, *gpy, *gpz;  // dL/d(hit position), bounces * n each
  double *gv;                    // dL/dvertex, 3 per vertex, accumulated
  double *gox, *goy, *goz;       // dL/d(ray origin), n each, written
  double *gdx, *gdy, *gdz;       // dL/d(raw ray direction), n each, written
};

// Adjoint of `bounces` (1..8) mirror bounces recomputed from the stored face ids prims[b * n + i]; a ray's chain
// ends at its first id >= num_faces (a miss).
hipError_t launch_mirror_adjoint(const RaysView &rays, const uint32_t *prims, uint32_t bounces,
                                 const double *vertices, const uint32_t *faces, uint32_t num_faces, uint64_t n,
                                 const MirrorGrads &g, hipStream_t stream);

// Writes the default RayHit (t = DBL_MAX, ids invalid, u = v = 0, normal 0) for n rays.
hipError_t launch_init_hits(const HitsView &hits, uint64_t n, hipStream_t stream);

// Scene<float>: traversal of one geometry type in FP32 (bvh.nodes32 and FP32 records), and the FP32 default RayHit.
hipError_t launch_trace_f32(GeomType type, const BvhView &bvh, const RaysViewF &rays, const HitsViewF &hits,
                            uint64_t n, const TraceLaunch &cfg, hipStream_t stream);
hipError_t launch_init_hits_f32(const HitsViewF &hits, uint64_t n, hipStream_t stream);

// Distance queries (distance.hip). inv maps primitive ids to leaf positions (spheres, planes, cylinders; built by
// launch_inverse_order from the view's prim_ids); ids >= n_prims give NaN.
hipError_t launch_inverse_order(const uint32_t *order, uint32_t n, uint32_t *inv, hipStream_t stream);
hipError_t launch_distance_pairs(GeomType type, const BvhView &bvh, const uint32_t *inv, const double *px,
                                 const double *py, const double *pz, const uint32_t *ids, uint64_t n, double *out,
                                 hipStream_t stream);
// All geometry types of a scene for the closest-surface query; extent = largest |coordinate| of a type's boxes.
struct ClosestViews {
  BvhView view[NUM_TYPES];
  double extent[NUM_TYPES];
  uint32_t present[NUM_TYPES];
  const double *leaf_vertices;  // triangles in leaf order: v0, v1, v2 (9 doubles each)
};
// Gathers the triangles' vertices in leaf order (9 doubles per triangle) for the closest-surface query.
hipError_t launch_leaf_vertices(const uint32_t *order, const uint32_t *faces, const double *vertices, uint32_t n,
                                double *out, hipStream_t stream);
hipError_t launch_closest_surface(const ClosestViews &views, const double *px, const double *py, const double *pz,
                                  uint64_t n, double *dist, uint32_t *prim, uint32_t *geom, hipStream_t stream);

// Device build of one geometry type's BVH (build.hip): the host builder's tree, primitive order and flattened
// records, bit for bit, provided no primitive box coordinate is NaN or -0.0 (the caller checks) and
// device_build_supported() holds for the options.
struct DeviceBuildIn {
  GeomType type;
  uint32_t n;  // primitives (> 0)
  uint32_t min_leaf, max_depth, nbins;
  const double *vertices;  // triangles: device vertices and faces (the boxes are computed on the device)
  const uint32_t *faces;
  const double *box_lo, *box_hi, *box_ctr;  // other types: host boxes and centers, 3 per primitive
  const double *centers, *radii, *dxs, *dys, *rot, *sa, *sb, *heights;  // other types: host record inputs
};
// Results are device allocations owned by the caller afterwards.
struct DeviceBuildOut {
  brt_tree_node *tree;  // reference-form nodes in pre-order
  uint32_t num_nodes;
  uint32_t *order;      // leaf order -> primitive id
  Node64 *n64;
  uint32_t num_n64;
  Node4 *n4;
  uint32_t num_n4;
  void *recs;           // primitive records in leaf order
  size_t rec_size;
  brt_tree_node root;   // host copy of node 0
  uint32_t max_depth, leaves, branches, bad_splits, max_prims, min_prims;
  uint32_t failed_line;  // on error: source line (build.hip) of the failing call
};
bool device_build_supported(uint32_t nbins, uint32_t max_depth);
// Builds on `stream` and waits for it.
hipError_t device_build(const DeviceBuildIn &in, DeviceBuildOut &out, hipStream_t stream);

}  // namespace brt
