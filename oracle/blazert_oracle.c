/*
 * blazert_oracle.c -- CPU restatement of blazeRT's intersect1 path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp). Citations are file:line in the
 * blazert repository. The arithmetic conventions for blaze operations are listed in the header.
 */
#include "blazert_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* One source, two oracles: FP64 (Scene<double>, the default) and, built with -DORC_FLOAT, FP32 (Scene<float>).
 * Every arithmetic literal is of the scene's type, as the reference writes them (static_cast<T>, integer literals
 * or exactly representable constants), so the float build evaluates in float throughout. */
typedef orc_real real;
#ifdef ORC_FLOAT
#define R_(x) x##f
#define REAL_MAX FLT_MAX
#define REAL_MIN FLT_MIN
#define REAL_EPSILON FLT_EPSILON
#define SQRT sqrtf
#define FABS fabsf
#define ROUND roundf
#define TIE_REL 1e-5f /* near-tie flag of the trace (test exclusion only) */
#else
#define R_(x) x
#define REAL_MAX DBL_MAX
#define REAL_MIN DBL_MIN
#define REAL_EPSILON DBL_EPSILON
#define SQRT sqrt
#define FABS fabs
#define ROUND round
#define TIE_REL 1e-12
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------------ */
/* blaze vector conventions                                                                          */
/* ------------------------------------------------------------------------------------------------ */
static inline real dot3(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(real *r, const real *a, const real *b) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
static inline void cross3(real *r, const real *a, const real *b) {
  const real x = a[1] * b[2] - a[2] * b[1];
  const real y = a[2] * b[0] - a[0] * b[2];
  const real z = a[0] * b[1] - a[1] * b[0];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
static inline real norm3(const real *a) { return SQRT(dot3(a, a)); }
static inline void normalize3(real *r, const real *a) {
  const real len = norm3(a);
  const real il = (len != 0) ? R_(1.0) / len : R_(0.0);
  r[0] = a[0] * il;
  r[1] = a[1] * il;
  r[2] = a[2] * il;
}
static inline void scale3(real *r, real s, const real *a) {
  r[0] = s * a[0];
  r[1] = s * a[1];
  r[2] = s * a[2];
}
/* vec / scalar evaluated as vec * (1/scalar) */
static inline void divs3(real *r, const real *a, real s) {
  const real is = R_(1.0) / s;
  r[0] = a[0] * is;
  r[1] = a[1] * is;
  r[2] = a[2] * is;
}
/* row-major R * v */
static inline void matvec(real *r, const real *R, const real *v) {
  const real x = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  const real y = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  const real z = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
/* trans(R) * v */
static inline void matTvec(real *r, const real *R, const real *v) {
  const real x = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  const real y = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  const real z = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
/* std::min(a,b) == (b < a) ? b : a ; std::max(a,b) == (a < b) ? b : a */
static inline real std_min(real a, real b) { return (b < a) ? b : a; }
static inline real std_max(real a, real b) { return (a < b) ? b : a; }

/* blazert/datatypes.h:79-85 */
static inline real product_sign(real x, real y) {
#ifdef ORC_FLOAT
  uint32_t ux, uy;
  memcpy(&ux, &x, 4);
  memcpy(&uy, &y, 4);
  ux ^= (uy & UINT32_C(0x80000000));
#else
  uint64_t ux, uy;
  memcpy(&ux, &x, 8);
  memcpy(&uy, &y, 8);
  ux ^= (uy & UINT64_C(0x8000000000000000));
#endif
  real r;
  memcpy(&r, &ux, sizeof r);
  return r;
}

/* blazert/datatypes.h:87-93 unity(min_, max_, min, max) */
static inline void unity(real *min_, real *max_, const real *mn, const real *mx) {
  for (int k = 0; k < 3; k++) {
    min_[k] = std_min(mn[k], min_[k]);
    max_[k] = std_max(mx[k], max_[k]);
  }
}

/* std::min({...}) / std::max({...}) over an initializer list: first element wins ties */
static inline real list_min6(const real *v) {
  real m = v[0];
  for (int i = 1; i < 6; i++)
    if (v[i] < m) m = v[i];
  return m;
}
static inline real list_max6(const real *v) {
  real m = v[0];
  for (int i = 1; i < 6; i++)
    if (m < v[i]) m = v[i];
  return m;
}

/* ------------------------------------------------------------------------------------------------ */
/* scene and collections                                                                             */
/* ------------------------------------------------------------------------------------------------ */
typedef struct orc_coll {
  uint32_t n;
  real *box_min, *box_max, *centers; /* 3n each */
  /* triangles (blazert/primitives/trimesh.h:61-131) */
  real *vertices; uint32_t nv; uint32_t *faces; real *vertex_normals;
  /* spheres / planes / cylinders */
  real *radii, *dxs, *dys, *rot, *sa, *sb, *heights;
  /* built tree */
  orc_node *nodes; uint32_t n_nodes, cap_nodes; uint32_t *prim_order; orc_stats stats;
  int present;
} orc_coll;

struct orc_scene {
  uint32_t min_leaf, max_depth, bins;
  uint32_t geometries;
  uint32_t geom_id[ORC_NUM_TYPES];
  int committed;
  orc_coll c[ORC_NUM_TYPES];
};

static real *dup_d(const real *p, size_t n) {
  real *r = (real *)malloc(n * sizeof(real) + 8);
  if (n) memcpy(r, p, n * sizeof(real));
  return r;
}

orc_scene *orc_scene_new(uint32_t min_leaf_primitives, uint32_t max_tree_depth, uint32_t bin_size) {
  orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
  s->min_leaf = min_leaf_primitives;
  s->max_depth = max_tree_depth;
  s->bins = bin_size;
  for (int i = 0; i < ORC_NUM_TYPES; i++) s->geom_id[i] = 0xFFFFFFFFu;
  return s;
}

void orc_scene_free(orc_scene *s) {
  if (!s) return;
  for (int i = 0; i < ORC_NUM_TYPES; i++) {
    orc_coll *c = &s->c[i];
    free(c->box_min); free(c->box_max); free(c->centers); free(c->vertices); free(c->faces);
    free(c->vertex_normals); free(c->radii); free(c->dxs); free(c->dys); free(c->rot); free(c->sa);
    free(c->sb); free(c->heights); free(c->nodes); free(c->prim_order);
  }
  free(s);
}

static int can_add(orc_scene *s, int type) { return !s->c[type].present && !s->committed; }

static void alloc_boxes(orc_coll *c, uint32_t n) {
  c->n = n;
  c->box_min = (real *)calloc(3 * (size_t)n + 1, sizeof(real));
  c->box_max = (real *)calloc(3 * (size_t)n + 1, sizeof(real));
  c->centers = (real *)calloc(3 * (size_t)n + 1, sizeof(real));
}

/* blazert/primitives/trimesh.h:78-130 */
uint32_t orc_add_mesh(orc_scene *s, const real *vertices, uint32_t nv, const uint32_t *faces, uint32_t nf) {
  if (!can_add(s, ORC_TRIANGLES)) return 0xFFFFFFFFu;
  orc_coll *c = &s->c[ORC_TRIANGLES];
  alloc_boxes(c, nf);
  c->nv = nv;
  c->vertices = dup_d(vertices, 3 * (size_t)nv);
  c->faces = (uint32_t *)malloc(3 * (size_t)nf * sizeof(uint32_t) + 4);
  memcpy(c->faces, faces, 3 * (size_t)nf * sizeof(uint32_t));
  c->vertex_normals = (real *)calloc(3 * (size_t)nv + 1, sizeof(real));
  for (uint32_t f = 0; f < nf; f++) {
    const real *v0 = &c->vertices[3 * (size_t)faces[3 * f + 0]];
    const real *v1 = &c->vertices[3 * (size_t)faces[3 * f + 1]];
    const real *v2 = &c->vertices[3 * (size_t)faces[3 * f + 2]];
    /* pre_compute_center: (v0 + v1 + v2) / 3 */
    real sum[3];
    for (int k = 0; k < 3; k++) sum[k] = (v0[k] + v1[k]) + v2[k];
    divs3(&c->centers[3 * (size_t)f], sum, 3.0);
    /* pre_compute_bounding_box: start at v0, unity with v1 and v2 */
    real mn[3] = {v0[0], v0[1], v0[2]}, mx[3] = {v0[0], v0[1], v0[2]};
    unity(mn, mx, v1, v1);
    unity(mn, mx, v2, v2);
    memcpy(&c->box_min[3 * (size_t)f], mn, sizeof mn);
    memcpy(&c->box_max[3 * (size_t)f], mx, sizeof mx);
    /* pre_compute_face_normal: normalize(cross(v1 - v0, v2 - v0)); vertex normal += fn / 3 */
    real e1[3], e2[3], cr[3], fn[3], fn3[3];
    sub3(e2, v2, v0);
    sub3(e1, v1, v0);
    cross3(cr, e1, e2);
    normalize3(fn, cr);
    divs3(fn3, fn, 3.0);
    for (int j = 0; j < 3; j++) {
      real *vn = &c->vertex_normals[3 * (size_t)faces[3 * f + j]];
      vn[0] = vn[0] + fn3[0];
      vn[1] = vn[1] + fn3[1];
      vn[2] = vn[2] + fn3[2];
    }
  }
  c->present = 1;
  s->geom_id[ORC_TRIANGLES] = s->geometries++;
  return s->geom_id[ORC_TRIANGLES];
}

/* blazert/primitives/sphere.h:301-329 */
uint32_t orc_add_spheres(orc_scene *s, const real *centers, const real *radii, uint32_t n) {
  if (!can_add(s, ORC_SPHERES)) return 0xFFFFFFFFu;
  orc_coll *c = &s->c[ORC_SPHERES];
  alloc_boxes(c, n);
  c->radii = dup_d(radii, n);
  memcpy(c->centers, centers, 3 * (size_t)n * sizeof(real));
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      c->box_min[3 * (size_t)i + k] = centers[3 * (size_t)i + k] - radii[i];
      c->box_max[3 * (size_t)i + k] = centers[3 * (size_t)i + k] + radii[i];
    }
  c->present = 1;
  s->geom_id[ORC_SPHERES] = s->geometries++;
  return s->geom_id[ORC_SPHERES];
}

/* blazert/primitives/plane.h:89-150 */
uint32_t orc_add_planes(orc_scene *s, const real *centers, const real *dxs, const real *dys, const real *rot9,
                        uint32_t n) {
  if (!can_add(s, ORC_PLANES)) return 0xFFFFFFFFu;
  orc_coll *c = &s->c[ORC_PLANES];
  alloc_boxes(c, n);
  c->dxs = dup_d(dxs, n);
  c->dys = dup_d(dys, n);
  c->rot = dup_d(rot9, 9 * (size_t)n);
  memcpy(c->centers, centers, 3 * (size_t)n * sizeof(real));
  const real thickness = REAL_MIN; /* std::numeric_limits<real>::min() */
  for (uint32_t i = 0; i < n; i++) {
    const real *ctr = &centers[3 * (size_t)i];
    const real *R = &rot9[9 * (size_t)i];
    const real dx = dxs[i], dy = dys[i];
    const real tmp[6][3] = {{-dx / 2, -dy / 2, 0}, {dx / 2, -dy / 2, 0}, {-dx / 2, dy / 2, 0},
                              {dx / 2, dy / 2, 0},   {0, 0, -thickness},   {0, 0, thickness}};
    real pts[6][3];
    for (int p = 0; p < 6; p++) {
      real rv[3];
      matvec(rv, R, tmp[p]);
      for (int k = 0; k < 3; k++) pts[p][k] = ctr[k] + rv[k];
    }
    for (int k = 0; k < 3; k++) {
      real col[6];
      for (int p = 0; p < 6; p++) col[p] = pts[p][k];
      c->box_min[3 * (size_t)i + k] = list_min6(col);
      c->box_max[3 * (size_t)i + k] = list_max6(col);
    }
  }
  c->present = 1;
  s->geom_id[ORC_PLANES] = s->geometries++;
  return s->geom_id[ORC_PLANES];
}

/* blazert/primitives/cylinder.h:91-147 */
uint32_t orc_add_cylinders(orc_scene *s, const real *centers, const real *a, const real *b,
                           const real *heights, const real *rot9, uint32_t n) {
  if (!can_add(s, ORC_CYLINDERS)) return 0xFFFFFFFFu;
  orc_coll *c = &s->c[ORC_CYLINDERS];
  alloc_boxes(c, n);
  c->sa = dup_d(a, n);
  c->sb = dup_d(b, n);
  c->heights = dup_d(heights, n);
  c->rot = dup_d(rot9, 9 * (size_t)n);
  memcpy(c->centers, centers, 3 * (size_t)n * sizeof(real));
  for (uint32_t i = 0; i < n; i++) {
    const real *ctr = &centers[3 * (size_t)i];
    const real *R = &rot9[9 * (size_t)i];
    const real a1t[3] = {a[i], 0.0, 0.0}, b1t[3] = {0.0, b[i], 0.0}, h1t[3] = {0.0, 0.0, heights[i] / 2};
    real ra[3], rb[3], rh[3];
    matvec(ra, R, a1t);
    matvec(rb, R, b1t);
    matvec(rh, R, h1t);
    real pts[6][3];
    for (int k = 0; k < 3; k++) {
      pts[0][k] = ctr[k] + ra[k]; /* a1 */
      pts[1][k] = ctr[k] + rb[k]; /* b1 */
      pts[2][k] = ctr[k] - ra[k]; /* a2 */
      pts[3][k] = ctr[k] - rb[k]; /* b2 */
      pts[4][k] = ctr[k] + rh[k]; /* h1 */
      pts[5][k] = ctr[k] - rh[k]; /* h2 */
    }
    for (int k = 0; k < 3; k++) {
      real col[6];
      for (int p = 0; p < 6; p++) col[p] = pts[p][k];
      c->box_min[3 * (size_t)i + k] = list_min6(col);
      c->box_max[3 * (size_t)i + k] = list_max6(col);
    }
  }
  c->present = 1;
  s->geom_id[ORC_CYLINDERS] = s->geometries++;
  return s->geom_id[ORC_CYLINDERS];
}

int orc_prim_box(const orc_scene *s, int type, uint32_t prim, real bmin[3], real bmax[3]) {
  const orc_coll *c = &s->c[type];
  if (!c->present || prim >= c->n) return -1;
  memcpy(bmin, &c->box_min[3 * (size_t)prim], 3 * sizeof(real));
  memcpy(bmax, &c->box_max[3 * (size_t)prim], 3 * sizeof(real));
  return 0;
}

int orc_prim_center(const orc_scene *s, int type, uint32_t prim, real ctr[3]) {
  const orc_coll *c = &s->c[type];
  if (!c->present || prim >= c->n) return -1;
  memcpy(ctr, &c->centers[3 * (size_t)prim], 3 * sizeof(real));
  return 0;
}

/* parameters of one primitive (user order): triangle v0 v1 v2 (9); sphere c r (4); plane c dx dy R (14);
 * cylinder c a b h R (15) */
int orc_prim_params(const orc_scene *s, int type, uint32_t prim, real *out) {
  const orc_coll *c = &s->c[type];
  if (!c->present || prim >= c->n) return -1;
  const size_t i = prim;
  switch (type) {
  case ORC_TRIANGLES:
    for (int j = 0; j < 3; j++) memcpy(out + 3 * j, &c->vertices[3 * (size_t)c->faces[3 * i + j]], 3 * sizeof(real));
    break;
  case ORC_SPHERES:
    memcpy(out, &c->centers[3 * i], 3 * sizeof(real));
    out[3] = c->radii[i];
    break;
  case ORC_PLANES:
    memcpy(out, &c->centers[3 * i], 3 * sizeof(real));
    out[3] = c->dxs[i];
    out[4] = c->dys[i];
    memcpy(out + 5, &c->rot[9 * i], 9 * sizeof(real));
    break;
  default:
    memcpy(out, &c->centers[3 * i], 3 * sizeof(real));
    out[3] = c->sa[i];
    out[4] = c->sb[i];
    out[5] = c->heights[i];
    memcpy(out + 6, &c->rot[9 * i], 9 * sizeof(real));
    break;
  }
  return 0;
}

uint32_t orc_prim_count(const orc_scene *s, int type) { return s->c[type].present ? s->c[type].n : 0u; }
uint32_t orc_geom_id(const orc_scene *s, int type) { return s->geom_id[type]; }

/* ------------------------------------------------------------------------------------------------ */
/* binned SAH builder                                                                                */
/* ------------------------------------------------------------------------------------------------ */
typedef struct bin_t { real min[3], max[3]; uint32_t count; real cost; } bin_t;

/* blazert/bvh/binbuffer.h:15-19 */
static inline real box_surface(const real *mn, const real *mx) {
  real b[3];
  for (int k = 0; k < 3; k++) b[k] = FABS(mx[k] - mn[k]);
  return R_(2.0) * (b[0] * b[1] + b[1] * b[2] + b[2] * b[0]);
}

/* blazert/bvh/bbox.h:27-39 */
static void compute_bounding_box(const orc_coll *c, const uint32_t *idx, uint32_t n, real *mn, real *mx) {
  for (int k = 0; k < 3; k++) {
    mn[k] = REAL_MAX;
    mx[k] = -REAL_MAX;
  }
  for (uint32_t i = 0; i < n; i++) unity(mn, mx, &c->box_min[3 * (size_t)idx[i]], &c->box_max[3 * (size_t)idx[i]]);
}

/* unsigned(std::round(x)) as compiled by g++ on x86-64 (64-bit truncating convert, low 32 bits) */
static inline uint32_t round_to_unsigned(real x) {
  const real r = ROUND(x);
  if (r != r) return 0u;
  if (r >= 9.2e18 || r <= -9.2e18) return 0u;
  return (uint32_t)(uint64_t)(int64_t)r;
}

/* blazert/bvh/binbuffer.h:66-95 and 114-167 */
static uint32_t find_best_split_binned(const orc_scene *s, const orc_coll *c, const uint32_t *idx, uint32_t n,
                                       const real *mn, const real *mx, real *cut_pos, bin_t *bins) {
  const uint32_t nb = s->bins;
  for (uint32_t i = 0; i < 3 * nb; i++) {
    for (int k = 0; k < 3; k++) {
      bins[i].min[k] = REAL_MAX;
      bins[i].max[k] = -REAL_MAX;
    }
    bins[i].count = 0;
    bins[i].cost = 0.0;
  }
  real size[3], inv_size[3];
  for (int k = 0; k < 3; k++) {
    size[k] = mx[k] - mn[k];
    inv_size[k] = (size[k] > 0) ? R_(1.0) / size[k] : R_(0.0);
  }
  const real scale = (real)(nb - 1);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t p = idx[i];
    const real *bmin = &c->box_min[3 * (size_t)p], *bmax = &c->box_max[3 * (size_t)p];
    const real *ctr = &c->centers[3 * (size_t)p];
    for (uint32_t j = 0; j < 3; j++) {
      const real nc = ((ctr[j] - mn[j]) * inv_size[j]) * scale;
      uint32_t r = round_to_unsigned(nc);
      uint32_t id = (0u < r) ? r : 0u; /* std::max(0u, r) */
      if (nb - 1 < id) id = nb - 1;    /* std::min(size-1, id) */
      bin_t *b = &bins[j * nb + id];
      b->count++;
      unity(b->min, b->max, bmin, bmax);
    }
  }
  real min_cost[3] = {REAL_MAX, REAL_MAX, REAL_MAX};
  for (uint32_t j = 0; j < 3; j++) {
    uint64_t count = 0;
    real mn_[3] = {REAL_MAX, REAL_MAX, REAL_MAX}, mx_[3] = {-REAL_MAX, -REAL_MAX, -REAL_MAX};
    for (uint32_t i = nb - 1; i > 0; i--) {
      bin_t *b = &bins[j * nb + i];
      unity(mn_, mx_, b->min, b->max);
      count += b->count;
      b->cost = (real)count * box_surface(mn_, mx_);
    }
    count = 0;
    for (int k = 0; k < 3; k++) {
      mn_[k] = REAL_MAX;
      mx_[k] = -REAL_MAX;
    }
    uint32_t min_bin = 1;
    for (uint32_t i = 0; i < nb - 1; i++) {
      bin_t *b = &bins[j * nb + i];
      bin_t *nx = &bins[j * nb + i + 1];
      unity(mn_, mx_, b->min, b->max);
      count += b->count;
      const real cost = (real)count * box_surface(mn_, mx_) + nx->cost;
      if (cost < min_cost[j]) {
        min_cost[j] = cost;
        min_bin = i + 1;
      }
    }
    cut_pos[j] = (real)min_bin * ((mx[j] - mn[j]) / (real)nb) + mn[j];
  }
  uint32_t axis = 0;
  if (min_cost[0] > min_cost[1]) axis = 1;
  if (min_cost[axis] > min_cost[2]) axis = 2;
  return axis;
}

/* libstdc++ std::partition for bidirectional iterators (bits/stl_algo.h, __partition); the reference
 * calls std::partition on a std::vector<unsigned>::iterator (blazert/bvh/builder.h:213-215). */
static uint32_t partition_idx(const orc_coll *c, uint32_t *first, uint32_t *last, uint32_t axis, real cut) {
  uint32_t *const begin = first;
  for (;;) {
    for (;;) {
      if (first == last) return (uint32_t)(first - begin);
      if (c->centers[3 * (size_t)(*first) + axis] < cut) ++first;
      else break;
    }
    --last;
    for (;;) {
      if (first == last) return (uint32_t)(first - begin);
      if (!(c->centers[3 * (size_t)(*last) + axis] < cut)) --last;
      else break;
    }
    const uint32_t tmp = *first;
    *first = *last;
    *last = tmp;
    ++first;
  }
}

typedef struct build_ctx {
  const orc_scene *s;
  orc_coll *c;
  uint32_t *idx;
  bin_t *bins;
} build_ctx;

static uint32_t push_node(orc_coll *c) {
  if (c->n_nodes == c->cap_nodes) {
    c->cap_nodes = c->cap_nodes ? 2 * c->cap_nodes : 64;
    c->nodes = (orc_node *)realloc(c->nodes, c->cap_nodes * sizeof(orc_node));
  }
  memset(&c->nodes[c->n_nodes], 0, sizeof(orc_node));
  c->nodes[c->n_nodes].children[0] = c->nodes[c->n_nodes].children[1] = 0xFFFFFFFFu;
  c->nodes[c->n_nodes].axis = 0xFFFFFFFFu;
  return c->n_nodes++;
}

/* blazert/bvh/builder.h:84-115 (leaf 133-148, branch 167-183, split 201-226) */
static uint32_t build_recursive(build_ctx *b, uint32_t begin, uint32_t end, uint32_t depth) {
  orc_coll *c = b->c;
  const uint32_t offset = push_node(c);
  if (c->stats.max_tree_depth < depth) c->stats.max_tree_depth = depth;
  const uint32_t n = end - begin;
  real mn[3], mx[3];
  compute_bounding_box(c, b->idx + begin, n, mn, mx);
  if (n <= b->s->min_leaf || depth >= b->s->max_depth) {
    orc_node *nd = &c->nodes[offset];
    memcpy(nd->min, mn, 3 * sizeof(real));
    memcpy(nd->max, mx, 3 * sizeof(real));
    nd->leaf = 1;
    nd->prim_begin = begin;
    nd->prim_count = n;
    if (c->stats.max_num_prims_per_leaf < n) c->stats.max_num_prims_per_leaf = n;
    if (n < c->stats.min_num_prims_per_leaf) c->stats.min_num_prims_per_leaf = n;
    c->stats.leaf_nodes++;
    return offset;
  }
  real cut_pos[3];
  uint32_t cut_axis = find_best_split_binned(b->s, c, b->idx + begin, n, mn, mx, cut_pos, b->bins);
  uint32_t mid = begin;
  for (uint32_t axis_try = 0; axis_try < 3; axis_try++) {
    mid = begin + partition_idx(c, b->idx + begin, b->idx + end, cut_axis, cut_pos[cut_axis]);
    if (mid == begin || mid == end) {
      c->stats.bad_splits++;
      mid = begin + (n >> 1);
      cut_axis = (cut_axis + 1) % 3;
    } else {
      break;
    }
  }
  {
    orc_node *nd = &c->nodes[offset];
    nd->leaf = 0;
    memcpy(nd->min, mn, 3 * sizeof(real));
    memcpy(nd->max, mx, 3 * sizeof(real));
    nd->axis = cut_axis;
  }
  c->stats.branch_nodes++;
  const uint32_t left = build_recursive(b, begin, mid, depth + 1);
  c->nodes[offset].children[0] = left;
  const uint32_t right = build_recursive(b, mid, end, depth + 1);
  c->nodes[offset].children[1] = right;
  return offset;
}

static void build_type(orc_scene *s, int type) {
  orc_coll *c = &s->c[type];
  c->stats.min_num_prims_per_leaf = 0xFFFFFFFFu;
  c->prim_order = (uint32_t *)malloc(((size_t)c->n + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < c->n; i++) c->prim_order[i] = i;
  c->cap_nodes = 2 * c->n + 1;
  c->nodes = (orc_node *)malloc(c->cap_nodes * sizeof(orc_node));
  c->n_nodes = 0;
  build_ctx b = {s, c, c->prim_order, (bin_t *)malloc(3 * (size_t)s->bins * sizeof(bin_t))};
  build_recursive(&b, 0, c->n, 0);
  free(b.bins);
}

/* blazert/scene.h:91-115 */
int orc_commit(orc_scene *s) {
  for (int t = 0; t < ORC_NUM_TYPES; t++)
    if (s->c[t].present) build_type(s, t);
  s->committed = 1;
  return 1;
}

/* blazert/bvh/builder.h:48-82 on a caller-defined collection (examples/geometry_template/GEOM_TEMPLATE.h:60-96):
 * the builder sees only the collection's boxes and centers, given here as arrays. The result is a committed scene
 * whose sphere slot (type 1) holds the tree; only its tree export is meaningful. */
orc_scene *orc_collection_scene(uint32_t min_leaf_primitives, uint32_t max_tree_depth, uint32_t bin_size,
                                const real *box_min, const real *box_max, const real *centers, uint32_t n) {
  orc_scene *s = orc_scene_new(min_leaf_primitives, max_tree_depth, bin_size);
  orc_coll *c = &s->c[ORC_SPHERES];
  alloc_boxes(c, n);
  memcpy(c->box_min, box_min, 3 * (size_t)n * sizeof(real));
  memcpy(c->box_max, box_max, 3 * (size_t)n * sizeof(real));
  memcpy(c->centers, centers, 3 * (size_t)n * sizeof(real));
  c->present = 1;
  s->geom_id[ORC_SPHERES] = s->geometries++;
  build_type(s, ORC_SPHERES);
  s->committed = 1;
  return s;
}

/* Test hook: partition_idx on a caller's index range with the predicate centers[3 * i + axis] < cut. Returns the
 * number of indices in the first part; idx is permuted in place. */
uint32_t orc_partition_test(const real *centers, uint32_t *idx, uint32_t n, uint32_t axis, real cut) {
  orc_coll c;
  memset(&c, 0, sizeof c);
  c.centers = (real *)centers;
  return partition_idx(&c, idx, idx + n, axis, cut);
}

uint32_t orc_tree_size(const orc_scene *s, int type) { return s->c[type].present ? s->c[type].n_nodes : 0; }

int orc_tree_export(const orc_scene *s, int type, orc_node *nodes, uint32_t *prim_order, orc_stats *stats) {
  const orc_coll *c = &s->c[type];
  if (!c->present || !s->committed) return -1;
  if (nodes) memcpy(nodes, c->nodes, c->n_nodes * sizeof(orc_node));
  if (prim_order) memcpy(prim_order, c->prim_order, c->n * sizeof(uint32_t));
  if (stats) *stats = c->stats;
  return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* rays, slab test                                                                                   */
/* ------------------------------------------------------------------------------------------------ */
/* blazert/ray.h:56-65: direction is normalized, but direction_inv and direction_sign are computed
 * from the raw constructor argument (the parameter shadows the member in the mem-initializers). */
void orc_ray_init(orc_ray *r, const real o[3], const real d[3], real tmin, real tmax, int any_hit) {
  for (int k = 0; k < 3; k++) r->org[k] = o[k];
  normalize3(r->dir, d);
  for (int k = 0; k < 3; k++) {
    r->inv[k] = R_(1.0) / d[k];
    r->sign[k] = d[k] < 0 ? 1u : 0u;
  }
  r->tmin = tmin;
  r->tmax = tmax;
  r->any_hit = any_hit;
}

/* blazert/bvh/intersect.h:32-51 */
int orc_intersect_node(real *min_distance, real *max_distance, const real bmin[3], const real bmax[3],
                       const orc_ray *ray) {
  const real l1 = R_(1.0) + R_(4.0) * REAL_EPSILON;
  for (int i = 0; i < 3; i++) {
    real mn = ray->sign[i] ? bmax[i] : bmin[i];
    real mx = ray->sign[i] ? bmin[i] : bmax[i];
    mn = (mn - ray->org[i]) * ray->inv[i];
    mx = ((mx - ray->org[i]) * ray->inv[i]) * l1;
    *min_distance = std_max(mn, *min_distance);
    *max_distance = std_min(mx, *max_distance);
  }
  return *min_distance <= *max_distance;
}

/* ------------------------------------------------------------------------------------------------ */
/* intersectors                                                                                      */
/* ------------------------------------------------------------------------------------------------ */
typedef struct isect {
  real min_hit_distance, hit_distance;
  real u, v;
  real normal[3];
  uint32_t prim_id;
  /* oracle bookkeeping for tie detection */
  real cand_t[2];
  uint32_t cand_p[2];
  uint32_t ncand;
  uint32_t cyl_accepts;
} isect;

static void cand_add(isect *i, real t, uint32_t prim) {
  if (!(t == t)) return;
  for (uint32_t k = 0; k < i->ncand && k < 2; k++)
    if (i->cand_p[k] == prim) {
      if (t < i->cand_t[k]) i->cand_t[k] = t;
      goto sort;
    }
  if (i->ncand < 2) {
    i->cand_t[i->ncand] = t;
    i->cand_p[i->ncand] = prim;
    i->ncand++;
  } else if (t < i->cand_t[1]) {
    i->cand_t[1] = t;
    i->cand_p[1] = prim;
  }
sort:
  if (i->ncand == 2 && i->cand_t[1] < i->cand_t[0]) {
    real tt = i->cand_t[0];
    uint32_t pp = i->cand_p[0];
    i->cand_t[0] = i->cand_t[1];
    i->cand_p[0] = i->cand_p[1];
    i->cand_t[1] = tt;
    i->cand_p[1] = pp;
  }
}

/* prepare_traversal, e.g. blazert/primitives/trimesh.h:145-151 (identical for all four types) */
static void prepare(isect *i, const orc_ray *ray) {
  memset(i, 0, sizeof *i);
  i->min_hit_distance = ray->tmin;
  i->hit_distance = ray->tmax;
  i->u = i->v = 0.0;
  i->prim_id = 0xFFFFFFFFu;
}

/* blazert/primitives/trimesh.h:153-181 with Triangle{a=v0, b=v2-v0, c=v0-v1} (trimesh.h:26-27) */
static int isect_triangle(isect *i, const orc_coll *c, uint32_t prim, const orc_ray *ray) {
  const real tolerance = R_(4.0) * REAL_EPSILON;
  const uint32_t *f = &c->faces[3 * (size_t)prim];
  const real *a = &c->vertices[3 * (size_t)f[0]];
  const real *v1 = &c->vertices[3 * (size_t)f[1]];
  const real *v2 = &c->vertices[3 * (size_t)f[2]];
  real e2[3], e1[3];
  sub3(e2, v2, a); /* tri.b */
  sub3(e1, a, v1); /* tri.c */
  real cc[3], r[3], n[3];
  sub3(cc, a, ray->org);
  cross3(r, ray->dir, cc);
  cross3(n, e1, e2);
  const real det = dot3(n, ray->dir);
  const real abs_det = FABS(det);
  const real u = product_sign(dot3(r, e2), det);
  const real v = product_sign(dot3(r, e1), det);
  const real w = abs_det - u - v;
  if (u >= -tolerance && v >= -tolerance && w >= -tolerance) {
    const real t = product_sign(dot3(n, cc), det);
    if (t >= abs_det * i->min_hit_distance) cand_add(i, t * (R_(1.0) / abs_det), prim);
    if (t >= abs_det * i->min_hit_distance && abs_det * i->hit_distance > t) {
      const real inv_det = R_(1.0) / abs_det;
      i->hit_distance = t * inv_det;
      i->u = u * inv_det;
      i->v = v * inv_det;
      i->prim_id = prim;
      return 1;
    }
  }
  return 0;
}

/* blazert/primitives/sphere.h:364-406 */
static int isect_sphere(isect *i, const orc_coll *c, uint32_t prim, const orc_ray *ray) {
  const real *org = ray->org, *dir = ray->dir;
  const real *center = &c->centers[3 * (size_t)prim];
  const real radius = c->radii[prim];
  real v[3];
  sub3(v, org, center);
  const real A = dot3(dir, dir);
  const real B = R_(2.0) * dot3(v, dir);
  const real C = dot3(v, v) - radius * radius;
  const real D = B * B - R_(4.0) * A * C;
  if (D < 0) return 0;
  const real Q = SQRT(D);
  const real rcpA = R_(1.0) / A;
  const real t0 = R_(0.5) * rcpA * (-B - Q);
  const real t1 = R_(0.5) * rcpA * (-B + Q);
  real tv[3];
  if (t0 > i->min_hit_distance) {
    scale3(tv, t0, dir);
    cand_add(i, norm3(tv), prim);
  } else if (t1 > i->min_hit_distance) {
    scale3(tv, t1, dir);
    cand_add(i, norm3(tv), prim);
  }
  if ((t0 > i->min_hit_distance) && (t0 < i->hit_distance)) {
    scale3(tv, t0, dir);
    i->hit_distance = norm3(tv);
    i->prim_id = prim;
    return 1;
  }
  if ((t1 > i->min_hit_distance) && (t1 < i->hit_distance)) {
    scale3(tv, t1, dir);
    i->hit_distance = norm3(tv);
    i->prim_id = prim;
    return 1;
  }
  return 0;
}

/* blazert/primitives/plane.h:180-270 */
static int isect_plane(isect *i, const orc_coll *c, uint32_t prim, const orc_ray *ray) {
  const real *center = &c->centers[3 * (size_t)prim];
  const real dx = c->dxs[prim], dy = c->dys[prim];
  const real *R = &c->rot[9 * (size_t)prim];
  real oc[3], org[3], dir[3];
  sub3(oc, ray->org, center);
  matTvec(org, R, oc);
  matTvec(dir, R, ray->dir);
  const real t1 = -org[2] / dir[2];
  real intercept[3];
  for (int k = 0; k < 3; k++) intercept[k] = org[k] + t1 * dir[k];
  const real tnear = i->min_hit_distance, tfar = i->hit_distance;
  const real x_min = -dx / 2, x_max = dx / 2, y_min = -dy / 2, y_max = dy / 2;
  const real X = intercept[0], Y = intercept[1];
  const int tok = (t1 > tnear) && (t1 < tfar);
  if ((t1 > tnear) && (X > x_min) && (X < x_max) && (Y > y_min) && (Y < y_max)) {
    real tv[3];
    scale3(tv, t1, dir);
    cand_add(i, norm3(tv), prim);
  }
  /* branch selection: interior, 4 edges, 4 corners, in the reference order */
  real lz[3] = {0.0, 0.0, 0.0};
  int branch = 0;
  const real sgn = org[2] / FABS(org[2]);
  if (tok && (X > x_min) && (X < x_max) && (Y > y_min) && (Y < y_max)) branch = 1;
  else if (tok && (X == x_min) && (X < x_max) && (Y > y_min) && (Y < y_max)) { branch = 2; lz[0] = -1.0; lz[1] = 0.0; }
  else if (tok && (X > x_min) && (X == x_max) && (Y > y_min) && (Y < y_max)) { branch = 3; lz[0] = 1.0; lz[1] = 0.0; }
  else if (tok && (X > x_min) && (X < x_max) && (Y == y_min) && (Y < y_max)) { branch = 4; lz[0] = 0.0; lz[1] = -1.0; }
  else if (tok && (X > x_min) && (X < x_max) && (Y > y_min) && (Y == y_max)) { branch = 5; lz[0] = 0.0; lz[1] = 1.0; }
  else if (tok && (X == x_min) && (X < x_max) && (Y == y_min) && (Y < y_max)) { branch = 6; lz[0] = -1.0; lz[1] = -1.0; }
  else if (tok && (X == x_min) && (X < x_max) && (Y > y_min) && (Y == y_max)) { branch = 7; lz[0] = -1.0; lz[1] = 1.0; }
  else if (tok && (X > x_min) && (X == x_max) && (Y == y_min) && (Y < y_max)) { branch = 8; lz[0] = 1.0; lz[1] = -1.0; }
  else if (tok && (X > x_min) && (X == x_max) && (Y > y_min) && (Y == y_max)) { branch = 9; lz[0] = 1.0; lz[1] = 1.0; }
  if (!branch) return 0;
  if (branch == 1) {
    const real ez[3] = {0.0, 0.0, 1.0};
    real rz[3];
    matvec(rz, R, ez);
    scale3(i->normal, sgn, rz);
  } else {
    lz[2] = sgn;
    matvec(i->normal, R, lz);
  }
  real tv[3];
  scale3(tv, t1, dir);
  i->hit_distance = norm3(tv);
  i->prim_id = prim;
  return 1;
}

/* Ng = rotation * normal / norm(normal)  (vec / scalar == vec * (1/scalar)) */
static void cyl_ng(real *out, const real *R, const real *nrm) {
  real rn[3];
  matvec(rn, R, nrm);
  divs3(out, rn, norm3(nrm));
}

static int cyl_accept(isect *i, real t, const real *dir, uint32_t prim) {
  real tv[3];
  scale3(tv, t, dir);
  i->hit_distance = norm3(tv);
  i->prim_id = prim;
  i->cyl_accepts++;
  cand_add(i, i->hit_distance, prim);
  return 1;
}

/* blazert/primitives/cylinder.h:181-495. The shell quadratic is shared by all regions. */
static int isect_cylinder(isect *i, const orc_coll *c, uint32_t prim, const orc_ray *ray) {
  const real *center = &c->centers[3 * (size_t)prim];
  const real *R = &c->rot[9 * (size_t)prim];
  real oc[3], org[3], dir[3];
  sub3(oc, ray->org, center);
  matTvec(org, R, oc);
  matTvec(dir, R, ray->dir);
  const real x0 = org[0], y0 = org[1], z0 = org[2];
  const real j = dir[0], k = dir[1], l = dir[2];
  const real a = c->sa[prim], b = c->sb[prim], h = c->heights[prim];

  /* areas 1 and 2: above the top cap / below the bottom cap (cylinder.h:228-332) */
  if (z0 > h / 2 || z0 < -h / 2) {
    const int top = z0 > h / 2;
    if (l == 0) return 0;
    const real t0 = top ? (h / 2 - z0) / l : (-h / 2 - z0) / l;
    if (t0 < 0) return 0;
    real ic[3];
    for (int q = 0; q < 3; q++) ic[q] = org[q] + t0 * dir[q];
    if (b * b * ic[0] * ic[0] + a * a * ic[1] * ic[1] <= a * a * b * b) {
      const real nz[3] = {0.0, 0.0, top ? 1.0 : -1.0};
      cyl_ng(i->normal, R, nz);
      return cyl_accept(i, t0, dir, prim);
    } else {
      const real A = a * a * k * k + b * b * j * j;
      const real B = a * a * k * y0 + b * b * j * x0;
      const real C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
      if (C < 0) return 0;
      const real s0 = R_(1.0) / A * (-B + a * b * SQRT(C));
      const real s1 = R_(-1.0) / A * (B + a * b * SQRT(C));
      real i0[3], i1[3];
      for (int q = 0; q < 3; q++) {
        i0[q] = org[q] + dir[q] * s0;
        i1[q] = org[q] + dir[q] * s1;
      }
      if ((s1 > 0) && (s1 < s0) && (i1[2] <= h / 2) && (i1[2] >= -h / 2)) {
        const real nn[3] = {2 * i1[0] / (a * a), 2 * i1[1] / (b * b), 0};
        cyl_ng(i->normal, R, nn);
        return cyl_accept(i, s1, dir, prim);
      }
      if ((s0 > 0) && (s0 < s1) && (i0[2] <= h / 2) && (i0[2] >= -h / 2)) {
        const real nn[3] = {2 * i0[0] / (a * a), 2 * i0[1] / (b * b), 0};
        cyl_ng(i->normal, R, nn);
        return cyl_accept(i, s0, dir, prim);
      }
    }
  }
  /* area 3: beside the shell (cylinder.h:334-366) */
  else if ((z0 > -h / 2) && (z0 < h / 2) && (x0 * x0 * b * b + y0 * y0 * a * a > a * a * b * b)) {
    const real A = a * a * k * k + b * b * j * j;
    const real B = a * a * k * y0 + b * b * j * x0;
    const real C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
    if (C < 0) return 0;
    const real s0 = R_(1.0) / A * (-B + a * b * SQRT(C));
    const real s1 = R_(-1.0) / A * (B + a * b * SQRT(C));
    real i0[3], i1[3];
    for (int q = 0; q < 3; q++) {
      i0[q] = org[q] + dir[q] * s0;
      i1[q] = org[q] + dir[q] * s1;
    }
    if ((s1 > 0) && (s1 < s0) && (i1[2] <= h / 2) && (i1[2] >= -h / 2)) {
      const real nn[3] = {2 * i1[0] / (a * a), 2 * i1[1] / (b * b), 0};
      cyl_ng(i->normal, R, nn);
      return cyl_accept(i, s1, dir, prim);
    }
    if ((s0 > 0) && (s0 < s1) && (i0[2] <= h / 2) && (i0[2] >= -h / 2)) {
      const real nn[3] = {2 * i0[0] / (a * a), 2 * i0[1] / (b * b), 0};
      cyl_ng(i->normal, R, nn);
      return cyl_accept(i, s0, dir, prim);
    }
  }
  /* area 4: inside (cylinder.h:368-492); not an else-branch in the reference */
  if ((z0 >= -h / 2) && (z0 <= h / 2) && (x0 * x0 * b * b + y0 * y0 * a * a <= a * a * b * b)) {
    if (l == 0) {
      if ((z0 == -h / 2) || (z0 == h / 2)) return 0;
      const real A = a * a * k * k + b * b * j * j;
      const real B = a * a * k * y0 + b * b * j * x0;
      const real C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
      if (C < 0) return 0;
      const real s0 = R_(1.0) / A * (-B + a * b * SQRT(C));
      const real s1 = R_(-1.0) / A * (B + a * b * SQRT(C));
      real i0[3], i1[3];
      for (int q = 0; q < 3; q++) {
        i0[q] = org[q] + dir[q] * s0;
        i1[q] = org[q] + dir[q] * s1;
      }
      if ((s1 > 0) && (i1[2] <= h / 2) && (i1[2] >= -h / 2)) {
        const real nn[3] = {2 * i1[0] / (a * a), 2 * i1[1] / (b * b), 0};
        cyl_ng(i->normal, R, nn);
        return cyl_accept(i, s1, dir, prim);
      }
      if ((s0 > 0) && (i0[2] <= h / 2) && (i0[2] >= -h / 2)) {
        const real nn[3] = {2 * i0[0] / (a * a), 2 * i0[1] / (b * b), 0};
        cyl_ng(i->normal, R, nn);
        return cyl_accept(i, s0, dir, prim);
      }
    } else if ((j == 0) && (k == 0)) {
      if (b * b * x0 * x0 + a * a * y0 * y0 == a * a * b * b) return 0;
      if (l < 0) {
        const real t0 = (-h / 2 - z0) / l;
        const real nz[3] = {0.0, 0.0, -1.0};
        cyl_ng(i->normal, R, nz);
        return cyl_accept(i, t0, dir, prim);
      }
      if (l > 0) {
        const real t0 = (h / 2 - z0) / l;
        const real nz[3] = {0.0, 0.0, 1.0};
        cyl_ng(i->normal, R, nz);
        return cyl_accept(i, t0, dir, prim);
      }
    } else {
      const real A = a * a * k * k + b * b * j * j;
      const real B = a * a * k * y0 + b * b * j * x0;
      const real C = a * a * k * k + b * b * j * j - j * j * y0 * y0 + 2 * j * k * x0 * y0 - k * k * x0 * x0;
      if (C < 0) return 0;
      const real s0 = R_(1.0) / A * (-B + a * b * SQRT(C));
      const real s1 = R_(-1.0) / A * (B + a * b * SQRT(C));
      real i0[3], i1[3];
      for (int q = 0; q < 3; q++) {
        i0[q] = org[q] + dir[q] * s0;
        i1[q] = org[q] + dir[q] * s1;
      }
      if ((s1 > 0) && (i1[2] <= h / 2) && (i1[2] >= -h / 2)) {
        const real nn[3] = {2 * i1[0] / (a * a), 2 * i1[1] / (b * b), 0};
        cyl_ng(i->normal, R, nn);
        return cyl_accept(i, s1, dir, prim);
      }
      if ((s0 > 0) && (i0[2] <= h / 2) && (i0[2] >= -h / 2)) {
        /* reference quirk (cylinder.h:462-463): the un-rotated normal is stored */
        const real nn[3] = {2 * i0[0] / (a * a), 2 * i0[1] / (b * b), 0};
        memcpy(i->normal, nn, 3 * sizeof(real));
        return cyl_accept(i, s0, dir, prim);
      }
      if (l < 0) {
        const real t0 = (-h / 2 - z0) / l;
        const real nz[3] = {0.0, 0.0, -1.0};
        cyl_ng(i->normal, R, nz);
        return cyl_accept(i, t0, dir, prim);
      }
      if (l > 0) {
        const real t0 = (h / 2 - z0) / l;
        const real nz[3] = {0.0, 0.0, 1.0};
        cyl_ng(i->normal, R, nz);
        return cyl_accept(i, t0, dir, prim);
      }
    }
  }
  return 0;
}

static int isect_prim(isect *i, const orc_coll *c, int type, uint32_t prim, const orc_ray *ray) {
  switch (type) {
  case ORC_TRIANGLES: return isect_triangle(i, c, prim, ray);
  case ORC_SPHERES: return isect_sphere(i, c, prim, ray);
  case ORC_PLANES: return isect_plane(i, c, prim, ray);
  default: return isect_cylinder(i, c, prim, ray);
  }
}

/* post_traversal: trimesh.h:133-143, sphere.h:336-343, plane.h:153-159, cylinder.h:154-160 */
static void post(const isect *i, const orc_coll *c, int type, const orc_ray *ray, orc_hit *hit) {
  hit->t = i->hit_distance;
  hit->u = i->u;
  hit->v = i->v;
  hit->prim_id = i->prim_id;
  if (type == ORC_TRIANGLES) {
    const uint32_t *f = &c->faces[3 * (size_t)i->prim_id];
    const real *n0 = &c->vertex_normals[3 * (size_t)f[0]];
    const real *n1 = &c->vertex_normals[3 * (size_t)f[1]];
    const real *n2 = &c->vertex_normals[3 * (size_t)f[2]];
    const real w = (R_(1.0) - i->u) - i->v;
    real s[3];
    for (int k = 0; k < 3; k++) s[k] = (w * n0[k] + i->u * n1[k]) + i->v * n2[k];
    normalize3(hit->normal, s);
  } else if (type == ORC_SPHERES) {
    const real *ctr = &c->centers[3 * (size_t)i->prim_id];
    real p[3];
    for (int k = 0; k < 3; k++) p[k] = (ray->org[k] + i->hit_distance * ray->dir[k]) - ctr[k];
    normalize3(hit->normal, p);
  } else {
    normalize3(hit->normal, i->normal);
  }
}

int orc_intersect_single(const orc_scene *s, int type, uint32_t prim, const orc_ray *ray, orc_hit *hit) {
  const orc_coll *c = &s->c[type];
  if (!c->present || prim >= c->n) return -1;
  isect i;
  prepare(&i, ray);
  const int h = isect_prim(&i, c, type, prim, ray);
  if (h) post(&i, c, type, ray, hit);
  return h;
}

/* blazert/bvh/accel.h:79-120 with the leaf loop of blazert/bvh/intersect.h:66-78 */
static int traverse_impl(const orc_scene *s, int type, const orc_ray *ray, orc_hit *hit, orc_trace *tr, isect *i) {
  const orc_coll *c = &s->c[type];
  uint32_t stack[512];
  uint32_t ss = 0;
  stack[ss++] = 0;
  if (tr && tr->max_stack < ss) tr->max_stack = ss;
  real hit_distance = ray->tmax;
  prepare(i, ray);
  if (c->n == 0) return 0; /* empty collection: the reference would read an empty node vector */
  while (ss > 0) {
    real max_hit_distance = hit_distance;
    real min_hit_distance = ray->tmin;
    const orc_node *node = &c->nodes[stack[ss - 1]];
    ss--;
    if (tr) tr->node_visits++;
    if (orc_intersect_node(&min_hit_distance, &max_hit_distance, node->min, node->max, ray)) {
      if (!node->leaf) {
        const uint32_t order_near = ray->sign[node->axis];
        /* Stack<unsigned,28>::pop_back writes stack[ss] (blazert/datatypes.h:66): once 28 entries are
         * held, the next pop writes past the array, so such rays are flagged and excluded. */
        if (ss + 2 >= ORC_REF_STACK_CAPACITY && tr) tr->overflow = 1;
        stack[ss++] = node->children[1 - order_near];
        stack[ss++] = node->children[order_near];
        if (tr && tr->max_stack < ss) tr->max_stack = ss;
        if (ss + 2 > 510) return -1;
      } else {
        int leaf_hit = 0;
        for (uint32_t p = 0; p < node->prim_count; p++) {
          if (tr) tr->prim_tests++;
          leaf_hit += isect_prim(i, c, type, c->prim_order[node->prim_begin + p], ray);
          if (leaf_hit && ray->any_hit) break;
        }
        if (leaf_hit) {
          hit_distance = i->hit_distance;
          if (ray->any_hit) break;
        }
      }
    }
  }
  const int h = i->hit_distance < ray->tmax;
  if (h) post(i, c, type, ray, hit);
  return h;
}

static void trace_close(orc_trace *tr, const isect *i, real final_t) {
  if (!tr) return;
  if (i->cyl_accepts >= 2) tr->multi_cyl = 1;
  if (i->ncand >= 2) {
    const real d = FABS(i->cand_t[1] - i->cand_t[0]);
    if (d <= TIE_REL * FABS(i->cand_t[0]) || d <= TIE_REL * FABS(final_t)) tr->tie = 1;
  }
}

int orc_traverse(const orc_scene *s, int type, const orc_ray *ray, orc_hit *hit, orc_trace *trace) {
  isect i;
  if (!s->c[type].present || !s->committed) return 0;
  const int h = traverse_impl(s, type, ray, hit, trace, &i);
  trace_close(trace, &i, i.hit_distance);
  return h;
}

/* blazert/scene.h:161-211 */
int orc_intersect1(const orc_scene *s, const orc_ray *ray, orc_hit *rayhit, orc_trace *trace) {
  orc_hit tmp;
  memset(&tmp, 0, sizeof tmp);
  tmp.t = REAL_MAX;
  tmp.prim_id = tmp.geom_id = 0xFFFFFFFFu;
  int hit = 0;
  isect i;
  isect all;
  memset(&all, 0, sizeof all);
  for (int type = 0; type < ORC_NUM_TYPES; type++) {
    if (!s->c[type].present) continue;
    const int h = traverse_impl(s, type, ray, &tmp, trace, &i);
    for (uint32_t k = 0; k < i.ncand; k++) cand_add(&all, i.cand_t[k], i.cand_p[k] | ((uint32_t)type << 30));
    all.cyl_accepts += i.cyl_accepts;
    if (h) {
      if (type == ORC_TRIANGLES || tmp.t < rayhit->t) {
        *rayhit = tmp;
        rayhit->geom_id = s->geom_id[type];
        hit = 1;
      }
    }
  }
  trace_close(trace, &all, rayhit->t);
  return hit;
}

uint64_t orc_intersect_batch(const orc_scene *s, uint64_t n, const real *ox, const real *oy, const real *oz,
                             const real *dx, const real *dy, const real *dz, const real *tmin,
                             const real *tmax, real *t, real *u, real *v, uint32_t *prim_id,
                             uint32_t *geom_id, real *nx, real *ny, real *nz, orc_trace *trace, int nthreads,
                             int any_hit) {
  uint64_t hits = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : hits)
#endif
  for (int64_t r = 0; r < (int64_t)n; r++) {
    orc_ray ray;
    const real o[3] = {ox[r], oy[r], oz[r]}, d[3] = {dx[r], dy[r], dz[r]};
    orc_ray_init(&ray, o, d, tmin ? tmin[r] : R_(0.0), tmax ? tmax[r] : REAL_MAX, any_hit);
    orc_hit h;
    memset(&h, 0, sizeof h);
    h.t = REAL_MAX;
    h.prim_id = h.geom_id = 0xFFFFFFFFu;
    orc_trace *tr = trace ? &trace[r] : NULL;
    if (tr) memset(tr, 0, sizeof *tr);
    hits += (uint64_t)orc_intersect1(s, &ray, &h, tr);
    t[r] = h.t;
    if (u) u[r] = h.u;
    if (v) v[r] = h.v;
    prim_id[r] = h.prim_id;
    geom_id[r] = h.geom_id;
    if (nx) nx[r] = h.normal[0];
    if (ny) ny[r] = h.normal[1];
    if (nz) nz[r] = h.normal[2];
  }
  return hits;
}

int orc_bruteforce_type(const orc_scene *s, int type, const orc_ray *ray, orc_hit *hit) {
  const orc_coll *c = &s->c[type];
  if (!c->present) return 0;
  int found = 0;
  real best = ray->tmax;
  for (uint32_t p = 0; p < c->n; p++) {
    isect i;
    prepare(&i, ray);
    if (isect_prim(&i, c, type, p, ray) && i.hit_distance < best) {
      best = i.hit_distance;
      post(&i, c, type, ray, hit);
      found = 1;
    }
  }
  return found;
}

/* orc_bruteforce_type over SoA rays (OpenMP over rays): t and user prim id of the closest accepted primitive of one
 * type per ray, DBL_MAX / 0xFFFFFFFF where none accepts. */
uint64_t orc_bruteforce_batch(const orc_scene *s, int type, uint64_t n, const real *ox, const real *oy, const real *oz,
                              const real *dx, const real *dy, const real *dz, const real *tmin, const real *tmax,
                              real *t, uint32_t *prim_id, int nthreads) {
  uint64_t hits = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : hits)
#endif
  for (int64_t r = 0; r < (int64_t)n; r++) {
    orc_ray ray;
    const real o[3] = {ox[r], oy[r], oz[r]}, d[3] = {dx[r], dy[r], dz[r]};
    orc_ray_init(&ray, o, d, tmin ? tmin[r] : R_(0.0), tmax ? tmax[r] : REAL_MAX, 0);
    orc_hit h;
    memset(&h, 0, sizeof h);
    h.t = REAL_MAX;
    h.prim_id = h.geom_id = 0xFFFFFFFFu;
    const int hit = orc_bruteforce_type(s, type, &ray, &h);
    hits += (uint64_t)hit;
    t[r] = hit ? h.t : REAL_MAX;
    prim_id[r] = hit ? h.prim_id : 0xFFFFFFFFu;
  }
  return hits;
}
