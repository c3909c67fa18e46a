/*
 * blazert_amd.h -- C ABI of the MI355X closest-hit ray-query engine.
 *
 * Drop-in boundary for blazeRT's scene API (cstatz/blazert, header-only C++17):
 *   blazert::Scene<T> = BlazertScene<T>               blazert/blazert.h:15-21, blazert/scene.h:38-116
 *   Scene::add_mesh / add_spheres / add_planes / add_cylinders   blazert/scene.h:75-81, 227-346
 *   Scene::commit()                                   blazert/scene.h:91-115
 *   intersect1(scene, ray, rayhit)                    blazert/scene.h:161-211
 *   Ray<T> / RayHit<T>                                blazert/ray.h:20-85
 *   BVHBuildOptions / BVHBuildStatistics              blazert/bvh/options.h:14-28, blazert/bvh/statistics.h:10-37
 *
 * Plain C types only. All functions return BRT_OK (0) or a negative BRT_ERR_* code; brt_last_error()
 * returns a message for the calling thread's last failure. Geometry is copied at add time (the reference
 * keeps references to caller-owned vectors; here the caller may free its arrays after the call).
 * After brt_scene_commit() the scene is immutable and every intersect entry point is reentrant.
 */
#ifndef BLAZERT_AMD_H
#define BLAZERT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRT_OK 0
#define BRT_ERR_INVALID_ARGUMENT (-1)
#define BRT_ERR_STATE (-2)        /* e.g. second add of a type, add after commit, intersect before commit */
#define BRT_ERR_DEVICE (-3)       /* HIP runtime failure */
#define BRT_ERR_NO_DEVICE (-4)    /* no usable gfx950 device */
#define BRT_ERR_OUT_OF_MEMORY (-5)
#define BRT_ERR_UNSUPPORTED (-6)  /* e.g. tree deeper than the device stack */

#define BRT_INVALID_ID 0xFFFFFFFFu

/* geometry types; intersect1 evaluates them in this fixed order (blazert/scene.h:169-209) */
#define BRT_GEOM_TRIANGLES 0u
#define BRT_GEOM_SPHERES 1u
#define BRT_GEOM_PLANES 2u
#define BRT_GEOM_CYLINDERS 3u

/* traversal precision (launch option) */
#define BRT_PRECISION_FP64 0u              /* drop-in: FP64 everywhere, bit-identical to the reference */
#define BRT_PRECISION_FP32_TRAV_FP64_HIT 1u /* FP32 conservative slab tests, FP64 primitive tests */

/* launch flags */
#define BRT_FLAG_HITS_INOUT 1u  /* hits hold a prior RayHit: its t is the merge threshold, misses leave it */
#define BRT_FLAG_ANY_HIT 2u     /* Ray::AnyHit::yes (blazert/ray.h:24,50; blazert/bvh/accel.h:109-110) */

typedef struct brt_scene brt_scene;

/* BVHBuildOptions (blazert/bvh/options.h:21-27): defaults 4 / 28 / 64 */
typedef struct brt_build_options {
  uint32_t min_leaf_primitives;
  uint32_t max_tree_depth;
  uint32_t bin_size;
} brt_build_options;

/* BVHBuildStatistics (blazert/bvh/statistics.h:10-37) plus the flattened sizes */
typedef struct brt_build_stats {
  uint32_t max_tree_depth;
  uint32_t leaf_nodes;
  uint32_t max_num_prims_per_leaf;
  uint32_t min_num_prims_per_leaf;
  uint32_t branch_nodes;
  uint32_t bad_splits;
  double build_time_s;
  uint32_t num_primitives;
  uint32_t geom_id;
  uint64_t device_bytes; /* nodes (FP64 + FP32) + primitive records + ids + post-processing data */
  uint32_t built_on_device; /* 1: built by brt_scene_commit_device's GPU builder, 0: by the host builder */
  uint32_t reserved;
} brt_build_stats;

/* Reference-form tree node (pre-order numbering as built by blazert/bvh/builder.h:84-115). */
typedef struct brt_tree_node {
  double min[3], max[3];
  uint32_t leaf, axis;
  uint32_t children[2];
  uint32_t prim_begin, prim_count; /* range into the exported primitive order */
} brt_tree_node;

/* Device rays, structure of arrays (device pointers). tmin/tmax may be NULL: 0 and DBL_MAX
 * (blazert/ray.h:172-173). Directions need not be unit length: like Ray<T>, the engine normalizes them for the
 * primitive tests and uses 1/raw direction for the slab tests (blazert/ray.h:56-65). */
typedef struct brt_rays {
  const double *ox, *oy, *oz;
  const double *dx, *dy, *dz;
  const double *tmin, *tmax;
} brt_rays;

/* Device hits, structure of arrays. nx/ny/nz may be NULL (no normals computed). Without
 * BRT_FLAG_HITS_INOUT every ray is written: a miss is t = DBL_MAX, ids = BRT_INVALID_ID, u = v = 0. */
typedef struct brt_hits {
  double *t, *u, *v;
  uint32_t *prim_id, *geom_id;
  double *nx, *ny, *nz;
} brt_hits;

/* work distribution of a launch */
#define BRT_SCHEDULE_PERSISTENT 0u /* persistent waves; idle lanes pull new rays from a device counter */
#define BRT_SCHEDULE_PER_THREAD 1u /* one ray per thread */

/* ray order of a launch: sorting groups rays that take the same child order (direction octant) and start close
 * together (Morton code of the origin in the scene box). Uses a per-(scene, stream) device workspace that is grown
 * with hipMalloc on the first sorted launch of a given size, so that first launch must not be graph-captured. */
#define BRT_SORT_NONE 0u
#define BRT_SORT_OCTANT_MORTON 1u
#define BRT_SORT_MORTON 2u

typedef struct brt_launch {
  uint32_t precision;  /* BRT_PRECISION_* */
  uint32_t flags;      /* BRT_FLAG_* */
  uint32_t block_size; /* threads per block: 64, 128, 192 or 256; 0 = library default (256; 128 for FP64 spheres) */
  uint32_t schedule;   /* BRT_SCHEDULE_* */
  uint32_t refill;     /* persistent: refill a wave once this many of its 64 lanes are idle; 0 = default (32 for
                          triangles on the exact paths, 16 otherwise) */
  uint32_t sort;       /* BRT_SORT_*: trace rays in a sorted order (results are written in input order) */
  uint32_t stack_ring; /* traversal-stack entries per lane kept in LDS: 4, 8 or 16; 0 = 16 for split-precision
                          triangles, 8 otherwise.
                          Deeper entries spill to private memory; a deeper ring spills less but takes more LDS */
  uint32_t pool_size;  /* persistent: ray indices a wave claims at once: 32, 64 or 128; 0 = 64. Smaller claims
                          balance image-order (coherent) rays better, larger ones cost fewer atomics on the launch's
                          counters; results are the same */
} brt_launch;

/* One ray / hit in host memory, for the single-ray drop-in (RayHit<T>, blazert/ray.h:72-85). */
typedef struct brt_hit1 {
  double normal[3];
  double u, v;
  double t;
  uint32_t prim_id, geom_id;
} brt_hit1;

const char *brt_last_error(void);
const char *brt_version(void);

/* Number of usable gfx950 devices (0 when none; never fails). */
int brt_device_count(void);

/* BlazertScene<T>() with build_options (blazert/scene.h:42,73); opts may be NULL for the defaults. */
int brt_scene_create(const brt_build_options *opts, brt_scene **out);
/* Releases host and device memory of the scene. */
int brt_scene_destroy(brt_scene *scene);

/* add_mesh (blazert/scene.h:227-240): vertices xyz (3*nv doubles), faces (3*nf uint32).
 * geom_id receives the geometry id, or BRT_INVALID_ID with BRT_ERR_STATE on a second add / add after commit. */
int brt_scene_add_mesh(brt_scene *scene, const double *vertices, uint32_t num_vertices, const uint32_t *faces,
                       uint32_t num_faces, uint32_t *geom_id);
/* add_spheres (blazert/scene.h:258-271) */
int brt_scene_add_spheres(brt_scene *scene, const double *centers, const double *radii, uint32_t n,
                          uint32_t *geom_id);
/* add_planes (blazert/scene.h:293-307); rotations are row-major 3x3 (9 doubles each) */
int brt_scene_add_planes(brt_scene *scene, const double *centers, const double *dxs, const double *dys,
                         const double *rotations, uint32_t n, uint32_t *geom_id);
/* add_cylinders (blazert/scene.h:331-346) */
int brt_scene_add_cylinders(brt_scene *scene, const double *centers, const double *semi_axes_a,
                            const double *semi_axes_b, const double *heights, const double *rotations, uint32_t n,
                            uint32_t *geom_id);

/* commit (blazert/scene.h:91-115): builds every BVH on the host (binned SAH, reference topology) and
 * flattens it. Device upload happens on the first intersect, or explicitly with brt_scene_upload. */
int brt_scene_commit(brt_scene *scene);
/* Copies the committed scene to the current HIP device (idempotent). */
int brt_scene_upload(brt_scene *scene);

/* commit (blazert/scene.h:91-115) with the BVHs built on the current HIP device, enqueued on `stream` and waited
 * for; leaves the scene uploaded. Extension: the reference builds on the host (blazert/bvh/builder.h:46-226). The
 * trees, leaf order and device records are those of brt_scene_commit, bit for bit: a geometry type whose boxes
 * hold a NaN, an infinity or a negative zero, or whose options have more than 64 bins or a depth limit above 62,
 * is built by the host builder instead (brt_build_stats.built_on_device tells which). Rebuilds by the autotuner
 * stay on the device. */
int brt_scene_commit_device(brt_scene *scene, void *stream);

/* Device arrays of one committed geometry type, for inspection (brt_scene_read_device_array). */
#define BRT_ARRAY_NODES64 0u    /* 128-byte FP64 records, children in parent */
#define BRT_ARRAY_NODES4 1u     /* 128-byte FP32 4-wide records (split precision) */
#define BRT_ARRAY_PRIMITIVES 2u /* primitive records in leaf order */
#define BRT_ARRAY_PRIM_IDS 3u   /* uint32 leaf order -> primitive id */
/* Copies one device array to host memory dst (capacity bytes); *bytes receives its size (dst may be NULL to
 * query it). Uploads a host-committed scene first. Extension (the reference keeps no device copy). */
int brt_scene_read_device_array(const brt_scene *scene, uint32_t geom_type, uint32_t which, void *dst,
                                uint64_t capacity, uint64_t *bytes);

/* Statistics of one geometry type's build (BRT_ERR_INVALID_ARGUMENT if the type is absent). */
int brt_scene_build_stats(const brt_scene *scene, uint32_t geom_type, brt_build_stats *stats);

/* Reference-form export of one built tree (for inspection and parity tests). Call with nodes == NULL to get
 * *num_nodes; prim_order receives num_primitives ids in leaf order. */
int brt_scene_export_tree(const brt_scene *scene, uint32_t geom_type, brt_tree_node *nodes, uint32_t capacity,
                          uint32_t *prim_order, uint32_t *num_nodes);

/* Device points, structure of arrays (device pointers). */
typedef struct brt_points {
  const double *x, *y, *z;
} brt_points;

/* distance_to_surface(primitive_from_collection(collection, prim_ids[i]), point i) of the reference for n pairs on
 * the device (sphere.h:179-182, plane.h:272-284, cylinder.h:677-759): out[i] in FP64, bit-identical to the reference
 * arithmetic; NaN for an id >= the type's primitive count. Triangles: the distance to the closest point of
 * trimesh_distance.h:176-221 on the face (the reference's triangle overload, trimesh.h:183-191, does not compile
 * when instantiated). Asynchronous on `stream`. */
int brt_distance_to_surface(const brt_scene *scene, uint32_t geom_type, const brt_points *points,
                            const uint32_t *prim_ids, uint64_t n, double *out, void *stream);

/* Closest surface of the whole scene for n device points: the smallest distance_to_surface over every primitive,
 * ties to the earlier geometry type (triangles, spheres, planes, cylinders) and then the lower primitive id; +inf
 * and BRT_INVALID_ID when no primitive gives a distance. Extension: the reference has no scene-level distance query
 * (docs/source/roadmap.rst). Asynchronous on `stream`, except that the first call on a scene with triangles
 * allocates (hipMalloc, 72 bytes per triangle) and fills a leaf-order copy of the triangles' vertices, then waits
 * for it; do not graph-capture that first call. */
int brt_closest_surface(const brt_scene *scene, const brt_points *points, uint64_t n, double *distance,
                        uint32_t *prim_id, uint32_t *geom_id, void *stream);

/* Batched intersect1 (blazert/scene.h:161-211) for n rays resident on the device, enqueued on `stream`
 * (a hipStream_t, NULL = default stream). launch may be NULL (FP64, no flags, persistent). Asynchronous.
 * The launches keep a few device counters per scene and stream, allocated on the first launch on that stream
 * (do not graph-capture that first launch); launches on one stream run in order, so any number may be in flight.
 * With BRT_PRECISION_FP32_TRAV_FP64_HIT, rays whose reference slab test can produce NaN terms (a zero, non-finite
 * or reciprocal-overflowing direction component, a non-finite origin, a NaN tmin) are traced by the FP64 path. */
int brt_intersect(const brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                  const brt_launch *launch, void *stream);

/* brt_intersect that allocates nothing. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same arguments and
 * hits as brt_intersect, bit for bit. Its launch counters and its ray-sort workspace live in one static device
 * array of the library (24 MiB per GPU that loads the library, held until the process ends) instead of per-stream
 * and per-scene allocations. Limits, each refused with an error before any device work: the scene must already be
 * on the current device (BRT_ERR_STATE; brt_scene_upload, brt_scene_commit_device or a brt_intersect put it
 * there), and a sorted launch takes at most 2^20 rays (BRT_ERR_UNSUPPORTED); unsorted launches take any n. All
 * calls on one GPU share that array: calls that overlap in time on one GPU (on different streams, or from
 * different threads) race and must be serialised by the caller. */
int brt_intersect_noalloc(const brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                          const brt_launch *launch, void *stream);

/* Low-level traverse of one geometry type's BVH (blazert/bvh/accel.h:60-120): like the reference, a ray that hits
 * gets its record overwritten by post_traversal (t, u, v, prim_id, normal; geom_id is set to that geometry's id,
 * which the reference leaves to intersect1), and a ray that misses keeps the record the caller passed in. launch may be
 * NULL; flags may hold BRT_FLAG_ANY_HIT only; sort must be BRT_SORT_NONE. Asynchronous on `stream`. */
int brt_traverse(const brt_scene *scene, uint32_t geom_type, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                 const brt_launch *launch, void *stream);

/* ---- many scenes in one launch ------------------------------------------------------------------------------
 * A batch binds (scene, device rays, device hits, n) items; every geometry type present in any scene is traced by
 * one persistent launch over the concatenated rays of all items, with each hit merged into its own item exactly as
 * brt_intersect would. The batch keeps the device pointers: they must stay valid while it is launched.
 * brt_batch_capture records the launch sequence on `stream` (a non-default stream) into a hipGraph owned by the
 * batch; brt_batch_replay launches that graph. No allocation happens after brt_batch_create. A batch owns one set
 * of launch counters and writes the same hits buffers every time, so its launches and replays must not overlap:
 * issue them on one stream, or order them with events. */
typedef struct brt_batch brt_batch;

typedef struct brt_batch_item {
  const brt_scene *scene;
  brt_rays rays;
  brt_hits hits;
  uint64_t n;
} brt_batch_item;

/* launch may be NULL; its sort field must be BRT_SORT_NONE. Scenes are uploaded to the current device. */
int brt_batch_create(const brt_batch_item *items, uint32_t count, const brt_launch *launch, brt_batch **out);
int brt_batch_launch(brt_batch *batch, void *stream);
int brt_batch_capture(brt_batch *batch, void *stream);
int brt_batch_replay(brt_batch *batch, void *stream);
int brt_batch_destroy(brt_batch *batch);

/* ---- differentiable mirror tracing ----------------------------------------------------------------------------
 * No counterpart in the reference: an extension of intersect1 (blazert/scene.h:161-211) on a scene whose only
 * geometry is one triangle mesh. Bounce 0 traces `rays`; bounce b >= 1 starts at bounce b-1's hit position and
 * follows the mirror direction of normalize(d) about the unit face normal normalize(cross(v1 - v0, v2 - v0))
 * (blazert/primitives/trimesh.h:125-129), with tmin = bounce_tmin and tmax = DBL_MAX. A miss ends the ray: t =
 * DBL_MAX, prim_id = BRT_INVALID_ID, position and direction NaN for that and every later bounce.
 * Every array is bounces * n long, indexed [b * n + i], device memory. The traversal is the exact FP64 path, so
 * each bounce's (t, u, v, prim_id) equals intersect1 on that bounce's ray. */
typedef struct brt_mirror_out {
  double *t;             /* hit distance along the unit direction */
  double *u, *v;         /* barycentric hit coordinates; both NULL or both set */
  uint32_t *prim_id;     /* face id */
  double *px, *py, *pz;  /* hit position: the next bounce's origin */
  double *rx, *ry, *rz;  /* mirror direction: the next bounce's direction */
} brt_mirror_out;

/* launch may be NULL; flags must be 0 and sort NONE. With BRT_PRECISION_FP32_TRAV_FP64_HIT each bounce is traced
 * by the split-precision path (hit ids as FP64 except near-ties; a tie resolved the other way reflects off the other
 * face, so that ray's later bounces differ). Asynchronous on `stream`. */
int brt_mirror_trace(const brt_scene *scene, const brt_rays *rays, uint64_t n, uint32_t bounces, double bounce_tmin,
                     const brt_launch *launch, const brt_mirror_out *out, void *stream);

/* Gradients of a loss L(t, positions) of brt_mirror_trace's outputs. In: dL/dt and dL/dposition per bounce (either
 * may be NULL = 0). Out: dL/dvertex (3 per mesh vertex, ADDED to `vertices`, so clear it first), dL/d(origin) and
 * dL/d(raw direction) per ray (written; NULL = not wanted). FP64 throughout; vertex gradients are scatter-added with
 * atomics, so their last bits depend on the summation order. */
typedef struct brt_mirror_grads {
  const double *gt;
  const double *gpx, *gpy, *gpz;
  double *vertices;
  double *gox, *goy, *goz;
  double *gdx, *gdy, *gdz;
} brt_mirror_grads;

/* Recomputes every bounce of every ray from the stored face ids (prim_id from brt_mirror_trace, bounces * n) and the
 * committed vertices; no other intermediate of the forward pass is needed. Asynchronous on `stream`. */
int brt_mirror_adjoint(const brt_scene *scene, const brt_rays *rays, uint64_t n, uint32_t bounces,
                       const uint32_t *prim_id, const brt_mirror_grads *grads, void *stream);

/* ---- autotuning ---------------------------------------------------------------------------------------------
 * No counterpart in the reference. Times launch configurations (and BVH leaf sizes) on the caller's device rays with
 * HIP events and reports the fastest. Every candidate computes the same hits (the FP64 path bit for bit, split
 * precision on every ray that is not a tie), so tuning changes speed only. `hits` is scratch output. */
typedef struct brt_tune_result {
  brt_launch launch;             /* fastest configuration (precision and flags as requested) */
  uint32_t min_leaf_primitives;  /* brt_scene_tune_leaf: the leaf size chosen, else 0 */
  uint32_t candidates;           /* configurations timed */
  double best_ms;                /* mean launch time of the fastest configuration */
  double default_ms;             /* mean launch time of the library defaults (leaf size 4 for the leaf tuner) */
} brt_tune_result;

/* block_size {64,128,192,256} x refill {4,8,16,32} x stack_ring {4,8,16} x sort {NONE, OCTANT_MORTON}, plus the
 * defaults; `repeats` timed launches each (0 = 3) after one warm-up. Synchronizes `stream`. flags must not hold
 * BRT_FLAG_HITS_INOUT. */
int brt_autotune_launch(const brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                        uint32_t precision, uint32_t flags, uint32_t repeats, brt_tune_result *out, void *stream);
/* The same over a batch (block_size x refill x stack_ring; batches do not sort). The batch keeps the fastest
 * configuration for later launches and captures, so tune before brt_batch_capture. */
int brt_batch_autotune(brt_batch *batch, uint32_t repeats, brt_tune_result *out, void *stream);
/* BVH leaf size (BVHBuildOptions::min_leaf_primitives, blazert/bvh/options.h:21-27): rebuilds the scene with each
 * size in leaf_sizes (other build options kept), times a default launch of `precision` on the rays, and leaves the
 * scene committed with the fastest. Rebuilding discards the scene's device copy. */
int brt_scene_tune_leaf(brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n, uint32_t precision,
                        const uint32_t *leaf_sizes, uint32_t count, uint32_t repeats, brt_tune_result *out,
                        void *stream);

/* ---- blazert::Scene<float> ----
 * A scene made by brt_scene_create_f32 is the reference's Scene<float>: FP32 geometry, FP32 builder arithmetic
 * (binning, SAH), FP32 traversal and primitive tests, FP32 hits, bit-identical to the reference's T = float
 * instantiation. It takes the _f32 add functions and is traversed with brt_intersect_f32 / brt_intersect1_f32;
 * commit, upload, build statistics, tree export and destroy are shared with Scene<double>. Split precision,
 * batches, the device build, distance queries and mirror tracing are for Scene<double> only. */
int brt_scene_create_f32(const brt_build_options *opts, brt_scene **out);
int brt_scene_add_mesh_f32(brt_scene *scene, const float *vertices, uint32_t num_vertices, const uint32_t *faces,
                           uint32_t num_faces, uint32_t *geom_id);
int brt_scene_add_spheres_f32(brt_scene *scene, const float *centers, const float *radii, uint32_t n,
                              uint32_t *geom_id);
int brt_scene_add_planes_f32(brt_scene *scene, const float *centers, const float *dxs, const float *dys,
                             const float *rotations, uint32_t n, uint32_t *geom_id);
int brt_scene_add_cylinders_f32(brt_scene *scene, const float *centers, const float *semi_axes_a,
                                const float *semi_axes_b, const float *heights, const float *rotations, uint32_t n,
                                uint32_t *geom_id);

/* FP32 device rays / hits (Ray<float>, RayHit<float>); same conventions as brt_rays / brt_hits. */
typedef struct brt_rays_f32 {
  const float *ox, *oy, *oz;
  const float *dx, *dy, *dz;
  const float *tmin, *tmax;
} brt_rays_f32;
typedef struct brt_hits_f32 {
  float *t, *u, *v;
  uint32_t *prim_id, *geom_id;
  float *nx, *ny, *nz;
} brt_hits_f32;
typedef struct brt_hit1_f32 {
  float normal[3];
  float u, v;
  float t;
  uint32_t prim_id, geom_id;
} brt_hit1_f32;

/* Batched intersect1 for a Scene<float> (blazert/scene.h:161-211 with T = float). launch->precision must be
 * BRT_PRECISION_FP64 (0, "the scene's own type") and sort BRT_SORT_NONE. Asynchronous on `stream`. */
int brt_intersect_f32(const brt_scene *scene, const brt_rays_f32 *rays, const brt_hits_f32 *hits, uint64_t n,
                      const brt_launch *launch, void *stream);
/* Single-ray intersect1 for a Scene<float>: host memory in/out, blocking. */
int brt_intersect1_f32(const brt_scene *scene, const float org[3], const float dir[3], float tmin, float tmax,
                       uint32_t flags, brt_hit1_f32 *hit, int *hit_out);

/* Single-ray drop-in for intersect1(scene, Ray<T>{org, dir, tmin, tmax}, rayhit): host memory in/out, blocking.
 * hit is in/out exactly like RayHit<T>; *hit_out receives the bool result. */
int brt_intersect1(const brt_scene *scene, const double org[3], const double dir[3], double tmin, double tmax,
                   uint32_t flags, brt_hit1 *hit, int *hit_out);

/* ---- user-defined collections ----
 * blazert::BVH<T, Collection> for a primitive type of the caller's (examples/geometry_template/GEOM_TEMPLATE.h:
 * 19-137, blazert/bvh/accel.h:32-52): SAHBinnedBuilder::build (blazert/bvh/builder.h:48-82) over the collection's
 * per-primitive boxes (get_primitive_bounding_box) and centers (get_primitive_center), 3 values per primitive, host
 * memory, copied. The tree is the reference's for those boxes and options; the primitives stay with the caller,
 * whose device code traverses the tree with its own intersector (include/blazert_amd/user_geometry.h, the
 * counterpart of traverse(), blazert/bvh/accel.h:80-120). _f32 builds BVH<float, Collection>. */
typedef struct brt_bvh brt_bvh;

/* Device view of a built tree: node records (brt::Node64 for an FP64 tree, brt::Node32 for FP32; layout in
 * include/blazert_amd/device/nodes.h), the leaf order (leaf position -> collection primitive index), the root box
 * and the root as a child reference. nodes / prim_ids are NULL for an empty collection (num_prims == 0). */
typedef struct brt_bvh_view {
  const void *nodes;
  const uint32_t *prim_ids;
  double root_min[3], root_max[3];
  uint32_t root_ref, root_cnt;
  uint32_t num_nodes, num_prims;
  uint32_t depth;
  uint32_t irregular_boxes; /* some node box of the tree has a NaN coordinate or min > max */
  uint32_t real_bytes; /* 8: FP64 tree, 4: FP32 tree */
  uint32_t reserved;
} brt_bvh_view;

int brt_bvh_build(const double *box_min, const double *box_max, const double *centers, uint32_t n,
                  const brt_build_options *opts, brt_bvh **out);
int brt_bvh_build_f32(const float *box_min, const float *box_max, const float *centers, uint32_t n,
                      const brt_build_options *opts, brt_bvh **out);
/* Uploads the tree to the current device on the first call (BRT_ERR_STATE if it was uploaded to another one,
 * BRT_ERR_UNSUPPORTED for a tree deeper than the 64-entry device stack) and returns its device view. */
int brt_bvh_device_view(brt_bvh *bvh, brt_bvh_view *view);
int brt_bvh_build_stats(const brt_bvh *bvh, brt_build_stats *stats);
/* Reference-form tree and leaf order, as brt_scene_export_tree. */
int brt_bvh_export_tree(const brt_bvh *bvh, brt_tree_node *nodes, uint32_t capacity, uint32_t *prim_order,
                        uint32_t *num_nodes);
int brt_bvh_destroy(brt_bvh *bvh);

#ifdef __cplusplus
}
#endif
#endif /* BLAZERT_AMD_H */
