/*
 * blazert_oracle.h -- CPU restatement of blazeRT's closest-hit path (TEST INFRASTRUCTURE ONLY).
 *
 * This is the parity oracle. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline. The product path
 * (blazert_amd/) never links, loads or calls anything in this directory.
 *
 * Every function restates one reference function (file:line under the blazert repository):
 *   ray precompute          blazert/ray.h:56-65
 *   slab test               blazert/bvh/intersect.h:32-51
 *   leaf loop               blazert/bvh/intersect.h:66-78
 *   traversal               blazert/bvh/accel.h:79-120
 *   binned SAH builder      blazert/bvh/builder.h:46-226, blazert/bvh/binbuffer.h:15-167, blazert/bvh/bbox.h:27-39
 *   triangles               blazert/primitives/trimesh.h:19-181
 *   spheres                 blazert/primitives/sphere.h:60-177
 *   planes                  blazert/primitives/plane.h:71-270
 *   cylinders               blazert/primitives/cylinder.h:74-495
 *   scene merge             blazert/scene.h:161-211
 *   distance_to_surface     distance_oracle.c (sphere.h, plane.h, cylinder.h, trimesh_distance.h)
 *
 * Arithmetic conventions (blaze 3.7 is not available, so these are restated from its published
 * semantics and documented as unpinned at the bit level):
 *   dot(a,b) = (a0*b0 + a1*b1) + a2*b2; norm(a) = sqrt(dot(a,a));
 *   normalize(a) = a * (len != 0 ? 1/len : 0);  vec / scalar = vec * (1/scalar);
 *   mat*vec row i = (R[i0]*v0 + R[i1]*v1) + R[i2]*v2; cross as in the textbook component order.
 * Compiled with -ffp-contract=off (no FMA contraction), like the HIP product kernels.
 */
#ifndef BLAZERT_ORACLE_H
#define BLAZERT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TRIANGLES = 0, ORC_SPHERES = 1, ORC_PLANES = 2, ORC_CYLINDERS = 3, ORC_NUM_TYPES = 4 };

/* The scene type T of blazert::Scene<T>: double, or float in the liboracle_f32 build (-DORC_FLOAT), which holds the
 * intersect path only (no mirror tracing or distance queries). */
#ifdef ORC_FLOAT
typedef float orc_real;
#else
typedef double orc_real;
#endif

/* reference Stack<unsigned, BLAZERT_MAX_TREE_DEPTH> capacity (blazert/defines.h:5-7, blazert/bvh/accel.h:83) */
#define ORC_REF_STACK_CAPACITY 28u

typedef struct orc_ray {
  orc_real org[3];
  orc_real dir[3]; /* normalize(raw direction) */
  orc_real inv[3]; /* 1 / raw direction */
  uint32_t sign[3]; /* raw direction < 0 */
  orc_real tmin, tmax;
  int any_hit;
} orc_ray;

typedef struct orc_hit {
  orc_real normal[3];
  orc_real u, v;
  orc_real t;
  uint32_t prim_id, geom_id;
} orc_hit;

/* per-ray traversal trace used for tie exclusion and algorithmic byte counts */
typedef struct orc_trace {
  uint32_t node_visits;  /* nodes popped and slab-tested */
  uint32_t prim_tests;   /* primitive intersection calls */
  uint32_t max_stack;    /* largest reference stack size reached */
  uint32_t overflow;     /* reference stack would exceed its 28 slots (undefined behaviour there) */
  uint32_t tie;          /* two distinct tested candidates within 1e-12 relative of the final t */
  uint32_t multi_cyl;    /* >= 2 cylinder accepts (order dependent: last one wins) */
} orc_trace;

typedef struct orc_node {
  orc_real min[3], max[3];
  uint32_t leaf, axis;
  uint32_t children[2];
  uint32_t prim_begin, prim_count; /* range into orc_bvh.prim_order */
} orc_node;

typedef struct orc_stats {
  uint32_t max_tree_depth, leaf_nodes, max_num_prims_per_leaf, min_num_prims_per_leaf, branch_nodes, bad_splits;
} orc_stats;

typedef struct orc_scene orc_scene;

/* scene construction mirrors blazert/scene.h:227-346; returns geom id or 0xFFFFFFFF */
orc_scene *orc_scene_new(uint32_t min_leaf_primitives, uint32_t max_tree_depth, uint32_t bin_size);
void orc_scene_free(orc_scene *s);
uint32_t orc_add_mesh(orc_scene *s, const orc_real *vertices, uint32_t nv, const uint32_t *faces, uint32_t nf);
uint32_t orc_add_spheres(orc_scene *s, const orc_real *centers, const orc_real *radii, uint32_t n);
uint32_t orc_add_planes(orc_scene *s, const orc_real *centers, const orc_real *dxs, const orc_real *dys,
                        const orc_real *rot9, uint32_t n);
uint32_t orc_add_cylinders(orc_scene *s, const orc_real *centers, const orc_real *a, const orc_real *b,
                           const orc_real *heights, const orc_real *rot9, uint32_t n);
int orc_commit(orc_scene *s);

/* collection precompute accessors (boxes / centers) for known-answer tests */
int orc_prim_box(const orc_scene *s, int type, uint32_t prim, orc_real bmin[3], orc_real bmax[3]);
int orc_prim_center(const orc_scene *s, int type, uint32_t prim, orc_real c[3]);

/* primitive parameters in user order (triangle v0 v1 v2; sphere c r; plane c dx dy R; cylinder c a b h R) */
int orc_prim_params(const orc_scene *s, int type, uint32_t prim, orc_real *out);
uint32_t orc_prim_count(const orc_scene *s, int type);
uint32_t orc_geom_id(const orc_scene *s, int type);

/* built tree export (reference pre-order layout) */
orc_scene *orc_collection_scene(uint32_t min_leaf_primitives, uint32_t max_tree_depth, uint32_t bin_size,
                                const orc_real *box_min, const orc_real *box_max, const orc_real *centers, uint32_t n);
/* test hook: the builder's std::partition restatement on idx[0..n) with centers[3 * i + axis] < cut */
uint32_t orc_partition_test(const orc_real *centers, uint32_t *idx, uint32_t n, uint32_t axis, orc_real cut);
uint32_t orc_tree_size(const orc_scene *s, int type);
int orc_tree_export(const orc_scene *s, int type, orc_node *nodes, uint32_t *prim_order, orc_stats *stats);

void orc_ray_init(orc_ray *r, const orc_real o[3], const orc_real d[3], orc_real tmin, orc_real tmax, int any_hit);
int orc_intersect_node(orc_real *min_distance, orc_real *max_distance, const orc_real bmin[3], const orc_real bmax[3],
                       const orc_ray *ray);

/* single primitive call sequence prepare -> intersect(prim) -> post (the reference tests' helper) */
int orc_intersect_single(const orc_scene *s, int type, uint32_t prim, const orc_ray *ray, orc_hit *hit);
/* low-level traverse of one geometry type (blazert/bvh/accel.h:60-120) */
int orc_traverse(const orc_scene *s, int type, const orc_ray *ray, orc_hit *hit, orc_trace *trace);
/* intersect1 (blazert/scene.h:161-211); hit is in/out like the reference RayHit */
int orc_intersect1(const orc_scene *s, const orc_ray *ray, orc_hit *hit, orc_trace *trace);

/* batched intersect1 over SoA rays, OpenMP over rays. hits start as a default RayHit.
 * trace arrays may be NULL. Returns number of hits. */
uint64_t orc_intersect_batch(const orc_scene *s, uint64_t n, const orc_real *ox, const orc_real *oy, const orc_real *oz,
                             const orc_real *dx, const orc_real *dy, const orc_real *dz, const orc_real *tmin,
                             const orc_real *tmax, orc_real *t, orc_real *u, orc_real *v, uint32_t *prim_id,
                             uint32_t *geom_id, orc_real *nx, orc_real *ny, orc_real *nz, orc_trace *trace, int nthreads,
                             int any_hit);

/* brute-force closest candidate per type (no BVH): every primitive tested from a fresh intersector */
int orc_bruteforce_type(const orc_scene *s, int type, const orc_ray *ray, orc_hit *hit);
uint64_t orc_bruteforce_batch(const orc_scene *s, int type, uint64_t n, const orc_real *ox, const orc_real *oy,
                              const orc_real *oz, const orc_real *dx, const orc_real *dy, const orc_real *dz,
                              const orc_real *tmin, const orc_real *tmax, orc_real *t, uint32_t *prim_id, int nthreads);

/* ---- multi-bounce mirror tracing (mirror_oracle.c; no counterpart in the reference) ----
 * Arrays of per-bounce outputs and upstream gradients are bounces * n long, indexed [b * n + i]; per-ray gradient
 * outputs go / gd are 3 * n long, [c * n + i]. gv (3 per vertex) is accumulated. */
void orc_mirror_trace(const orc_scene *s, const double *V, const uint32_t *F, uint64_t n, const double *ox,
                      const double *oy, const double *oz, const double *dx, const double *dy, const double *dz,
                      const double *tmin, const double *tmax, uint32_t bounces, double bounce_tmin, double *t,
                      double *u, double *v, uint32_t *prim, double *px, double *py, double *pz, double *rx,
                      double *ry, double *rz, int nthreads);
/* per-ray loss sum_b gt.t_b + gp.x_b of the chain with fixed face ids (t from the face plane) */
void orc_mirror_loss(const double *V, const uint32_t *F, uint64_t n, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz, uint32_t bounces,
                     const uint32_t *prims, const double *gt, const double *gpx, const double *gpy,
                     const double *gpz, double *loss, int nthreads);
/* analytic gradient of the summed orc_mirror_loss (explicit Jacobians, barycentric dt/dvertex) */
void orc_mirror_grad(const double *V, const uint32_t *F, uint64_t n, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz, uint32_t bounces,
                     const uint32_t *prims, const double *gt, const double *gpx, const double *gpy,
                     const double *gpz, double *gv, double *go, double *gd, int nthreads);

/* ---- distance_to_surface (distance_oracle.c) ----
 * sphere.h:179-182, plane.h:272-284, cylinder.h:507-759, trimesh_distance.h:36-221 (the triangle distance is
 * norm(p - closest_point_on_triangle); the reference's own triangle overload cannot be instantiated). */
double orc_sphere_distance(const double c[3], double r, const double p[3]);
double orc_plane_distance(const double c[3], double dx, double dy, const double R[9], const double p[3]);
double orc_cylinder_distance(const double c[3], double a, double b, double h, const double R[9], const double p[3]);
void orc_closest_point_on_triangle(const double v0[3], const double v1[3], const double v2[3], const double p[3],
                                   double q[3]);
double orc_triangle_distance(const double v0[3], const double v1[3], const double v2[3], const double p[3]);
/* distance from point i to primitive prims[i] (user id) of one geometry type; NaN for an id out of range */
void orc_distance_batch(const orc_scene *s, int type, uint64_t n, const double *px, const double *py,
                        const double *pz, const uint32_t *prims, double *out, int nthreads);
/* brute-force closest surface over every primitive of the scene: smallest distance, ties to the earlier geometry
 * type then the lower primitive id; +inf and invalid ids when every distance is NaN */
void orc_closest_batch(const orc_scene *s, uint64_t n, const double *px, const double *py, const double *pz,
                       double *dist, uint32_t *prim, uint32_t *geom, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
