// HIP kernels for distance queries: blazeRT's distance_to_surface per (point, primitive) pair, and the closest
// surface of a whole scene per point.
//
// Per-primitive distances follow the reference's operation order in FP64 (contraction off):
//   spheres    blazert/primitives/sphere.h:179-182   |norm(center - p) - radius|
//   planes     blazert/primitives/plane.h:272-284    |local.z| when local.x, local.y lie within the plane's
//              world-space center +- half extent (the reference mixes the two frames), else DBL_MAX
//   cylinders  blazert/primitives/cylinder.h:507-759 (Newton iteration on the ellipse, region logic)
//   triangles  norm(p - closest_point_on_triangle(v0, v1, v2, p)), blazert/primitives/trimesh_distance.h:36-221.
//              The reference's triangle overload (trimesh.h:183-191) cannot be instantiated (it returns the point
//              as a scalar and rebuilds the vertices from the wrong edges); this is its evident intent.
//
// Closest surface: the smallest distance over every primitive of the scene, ties going to the earlier geometry
// type (triangles, spheres, planes, cylinders) and then to the lower primitive id -- the brute-force minimum.
// Triangles and spheres are searched through their BVHs: a child box is skipped when its distance from the point,
// less a margin of 1e-10 (|p| + scene extent), exceeds the best distance found, which no primitive inside it can
// then reach or tie. Planes and cylinders are scanned in full: the reference's plane distance can be smaller than
// the distance to the plane's box, and its cylinder box can miss part of a rotated cylinder.
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "../../include/blazert_amd/user_geometry.h"  // RingStack: LDS stack ring with a private spill area
#include "engine.h"

namespace brt {

namespace {

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// trans(R) * v for row-major R
__device__ __forceinline__ void mat_t_vec(const double *R, const double *v, double *r) {
  const double x = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  const double y = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  const double z = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
__device__ __forceinline__ double min2(double a, double b) { return (b < a) ? b : a; }  // std::min(a, b)

// ------------------------------------------------------------------------------------------------ primitives
__device__ double sphere_distance(const SphRec &S, const double *p) {
  const double v[3] = {S.c[0] - p[0], S.c[1] - p[1], S.c[2] - p[2]};
  return fabs(__builtin_sqrt(dot3(v, v)) - S.r);
}

__device__ double plane_distance(const PlaRec &P, const double *p) {
  const double v[3] = {p[0] - P.c[0], p[1] - P.c[1], p[2] - P.c[2]};
  double l[3];
  mat_t_vec(P.R, v, l);
  const double dx = P.dx / 2, dy = P.dy / 2;
  const double cx = P.c[0], cy = P.c[1];
  if ((l[0] >= cx - dx) && (l[0] <= cx + dx) && (l[1] >= cy - dy) && (l[1] <= cy + dy)) return fabs(l[2]);
  return DBL_MAX;
}

enum : int { UPPER_X = 0, LOWER_X = 1, UPPER_Y = 2, LOWER_Y = 3 };

// cylinder.h:517-545
__device__ double distance_derivative(int kind, double a, double b, double xq, double yq, double val) {
  if (kind == UPPER_X || kind == LOWER_X) {
    const double s = (val * val > a * a) ? __builtin_sqrt(val * val - a * a) : __builtin_sqrt(a * a - val * val);
    if (kind == UPPER_X) return -2 * (xq - val) - 2 * b * val * (b / a * s - yq) / (a * s);
    return -2 * (xq - val) - 2 * b * val * (b / a * s + yq) / (a * s);
  }
  const double s = (val * val > b * b) ? __builtin_sqrt(val * val - b * b) : __builtin_sqrt(b * b - val * val);
  if (kind == UPPER_Y) return -2 * (yq - val) - 2 * a * val * (a / b * s + xq) / (b * s);
  return -2 * (yq - val) - 2 * a * val * (a / b * s - xq) / (b * s);
}

// cylinder.h:548-583
__device__ double single_newton_iteration(int kind, double a, double b, double xq, double yq, double val) {
  if (kind == UPPER_X || kind == LOWER_X) {
    const double s = (val * val > a * a) ? __builtin_sqrt(val * val - a * a) : __builtin_sqrt(a * a - val * val);
    if (kind == UPPER_X)
      return ((a * a * a * xq - a * val * val * xq) * s + b * val * val * val * yq) /
             ((a * a * a - a * b * b - a * val * val + b * b * val * val / a) * s + a * a * b * yq);
    return ((-a * a * a * xq + a * val * val * xq) * s + b * val * val * val * yq) /
           ((-a * a * a + a * b * b + a * val * val - b * b * val * val / a) * s + a * a * b * yq);
  }
  const double s = (val * val > b * b) ? __builtin_sqrt(val * val - b * b) : __builtin_sqrt(b * b - val * val);
  if (kind == UPPER_Y)
    return ((b * b * b * yq - b * val * val * yq) * s + a * val * val * val * xq) /
           ((b * b * b - b * a * a - b * val * val + a * a * val * val / b) * s + b * b * a * xq);
  return ((-b * b * b * yq + b * val * val * yq) * s + a * val * val * val * xq) /
         ((-b * b * b + b * a * a + b * val * val - a * a * val * val / b) * s + b * b * a * xq);
}

// cylinder.h:586-607: at most 100 steps, so every thread leaves the loop
__device__ double newton_iteration(int kind, double a, double b, double xq, double yq, double x0) {
  double val0 = x0;
  double cnt = 0;
  const double epsilon = 1e-7;
  while ((fabs(distance_derivative(kind, a, b, xq, yq, val0)) >= epsilon) && (cnt < 100)) {
    const double val = single_newton_iteration(kind, a, b, xq, yq, val0);
    val0 = val;
    cnt++;
    if (fabs(val0) == a) val0 += DBL_EPSILON;
  }
  return val0;
}

// cylinder.h:615-665
__device__ double min_distance_to_ellipse(double a, double b, double xq, double yq) {
  const double percentage = 0.95;
  if (((xq >= 0) && (yq >= (b / a) * xq)) || ((xq <= 0) && (yq >= (-b / a) * xq))) {
    const double x0 = xq > 0 ? percentage * a : -percentage * a;
    const double x_min = newton_iteration(UPPER_X, a, b, xq, yq, x0);
    const double s = __builtin_sqrt(a * a - x_min * x_min);
    const double y_min = b / a * s;
    return __builtin_sqrt((xq - x_min) * (xq - x_min) + (yq - y_min) * (yq - y_min));
  } else if (((xq >= 0) && (yq <= (-b / a) * xq)) || ((xq <= 0) && (yq <= (b / a) * xq))) {
    const double x0 = xq > 0 ? percentage * a : -percentage * a;
    const double x_min = newton_iteration(LOWER_X, a, b, xq, yq, x0);
    const double s = __builtin_sqrt(a * a - x_min * x_min);
    const double y_min = -b / a * s;
    return __builtin_sqrt((xq - x_min) * (xq - x_min) + (yq - y_min) * (yq - y_min));
  } else if ((xq >= 0) && (yq <= (b / a) * xq) && (yq >= (-b / a) * xq)) {
    const double y0 = yq > 0 ? percentage * b : -percentage * b;
    const double y_min = newton_iteration(UPPER_Y, a, b, xq, yq, y0);
    const double s = __builtin_sqrt(b * b - y_min * y_min);
    const double x_min = a / b * s;
    return __builtin_sqrt((yq - y_min) * (yq - y_min) + (xq - x_min) * (xq - x_min));
  } else {
    const double y0 = yq > 0 ? percentage * b : -percentage * b;
    const double y_min = newton_iteration(LOWER_Y, a, b, xq, yq, y0);
    const double s = __builtin_sqrt(b * b - y_min * y_min);
    const double x_min = -a / b * s;
    return __builtin_sqrt((yq - y_min) * (yq - y_min) + (xq - x_min) * (xq - x_min));
  }
}

// cylinder.h:677-759
__device__ double cylinder_distance(const CylRec &Cy, const double *p) {
  const double v[3] = {p[0] - Cy.c[0], p[1] - Cy.c[1], p[2] - Cy.c[2]};
  double l[3];
  mat_t_vec(Cy.R, v, l);
  const double a = Cy.a, b = Cy.b, h = Cy.h;
  const double dist_top = fabs(l[2] - h / 2);
  const double dist_bottom = fabs(l[2] + h / 2);
  const double xq = l[0], yq = l[1], zq = l[2];
  if (xq == 0 && yq == 0) {  // isZero(point_xy); std::min({...}) keeps the first smallest
    if ((zq > h / 2) || (zq < -h / 2)) return (dist_bottom < dist_top) ? dist_bottom : dist_top;
    double m = a;
    if (b < m) m = b;
    if (dist_top < m) m = dist_top;
    if (dist_bottom < m) m = dist_bottom;
    return m;
  }
  const double ellipse = (xq * xq) / (a * a) + (yq * yq) / (b * b) - 1;
  if (fabs(ellipse) < DBL_EPSILON) {
    if ((zq > h / 2) || (zq < -h / 2)) return min2(dist_top, dist_bottom);
    return 0.0;
  }
  if (ellipse < 0) {
    if (zq >= h / 2) return dist_top;
    if (zq <= -h / 2) return dist_bottom;
    double m = min_distance_to_ellipse(a, b, xq, yq);
    if (dist_bottom < m) m = dist_bottom;
    if (dist_top < m) m = dist_top;
    return m;
  }
  if (ellipse > 0) {
    if ((zq < h / 2) && (zq > -h / 2)) return min_distance_to_ellipse(a, b, xq, yq);
    const double ds = min_distance_to_ellipse(a, b, xq, yq);
    const double d1 = __builtin_sqrt(ds * ds + dist_top * dist_top);
    const double d2 = __builtin_sqrt(ds * ds + dist_bottom * dist_bottom);
    return (d2 < d1) ? d2 : d1;
  }
  return -1;
}

// trimesh_distance.h:176-221 with the region helpers r0-r6 (:36-171) in the same order
__device__ double triangle_distance(const double *v0, const double *v1, const double *v2, const double *p) {
  double E0[3], E1[3], D[3];
  for (int k = 0; k < 3; k++) {
    E0[k] = v1[k] - v0[k];
    E1[k] = v2[k] - v0[k];
    D[k] = v0[k] - p[k];
  }
  const double a = dot3(E0, E0), b = dot3(E0, E1), c = dot3(E1, E1), d = dot3(E0, D), e = dot3(E1, D);
  const double det = a * c - b * b;
  double s = b * e - c * d;
  double t = b * d - a * e;
  if (s + t <= det) {
    if (s < 0) {
      if (t < 0) {  // r4
        if (d < 0) {
          t = 0;
          s = (-d >= a) ? 1 : -d / a;
        } else {
          s = 0;
          if (e >= 0) t = 0;
          else t = (-e >= c) ? 1 : -e / c;
        }
      } else {  // r3
        s = 0;
        if (e >= 0) t = 0;
        else t = (-e >= c) ? 1 : -e / c;
      }
    } else if (t < 0) {  // r5
      t = 0;
      if (d >= 0) s = 0;
      else s = (-d >= a) ? 1 : -d / a;
    } else {  // r0
      const double inv_det = 1. / det;
      s *= inv_det;
      t *= inv_det;
    }
  } else {
    if (s < 0) {  // r2
      const double tmp0 = b + d, tmp1 = c + e;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2. * b + c;
        s = (numer >= denom) ? 1 : numer / denom;
        t = 1. - s;
      } else {
        s = 0;
        if (tmp1 <= 0) t = 1;
        else if (e >= 0) t = 0;
        else t = -e / c;
      }
    } else if (t < 0) {  // r6
      const double tmp0 = b + e, tmp1 = a + d;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2. * b + c;
        t = (numer >= denom) ? 1 : numer / denom;
        s = 1. - t;
      } else {
        t = 0;
        if (tmp1 <= 0) s = 1;
        else if (d >= 0) s = 0;
        else s = -d / a;
      }
    } else {  // r1
      const double numer = c + e - b - d;
      if (numer <= 0) {
        s = 0;
      } else {
        const double denom = a - 2. * b + c;
        s = (numer >= denom) ? 1 : numer / denom;
      }
      t = 1. - s;
    }
  }
  double w[3];
  for (int k = 0; k < 3; k++) w[k] = p[k] - (v0[k] + s * E0[k] + t * E1[k]);
  return __builtin_sqrt(dot3(w, w));
}

__device__ __forceinline__ double tri_distance_by_id(const BvhView &bvh, uint32_t id, const double *p) {
  const double *V = bvh.vertices;
  const uint32_t *F = bvh.faces + 3ull * id;
  return triangle_distance(V + 3ull * F[0], V + 3ull * F[1], V + 3ull * F[2], p);
}

// ------------------------------------------------------------------------------------------------ pairs
template <int TYPE>
__global__ void k_distance_pairs(BvhView bvh, const uint32_t *__restrict__ inv, const double *__restrict__ px,
                                 const double *__restrict__ py, const double *__restrict__ pz,
                                 const uint32_t *__restrict__ ids, uint64_t n, double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {px[i], py[i], pz[i]};
  const uint32_t id = ids[i];
  double d = __builtin_nan("");
  if (id < bvh.n_prims) {
    if (TYPE == TRIANGLES) {
      d = tri_distance_by_id(bvh, id, p);
    } else if (TYPE == SPHERES) {
      d = sphere_distance(reinterpret_cast<const SphRec *>(bvh.prims)[inv[id]], p);
    } else if (TYPE == PLANES) {
      d = plane_distance(reinterpret_cast<const PlaRec *>(bvh.prims)[inv[id]], p);
    } else {
      d = cylinder_distance(reinterpret_cast<const CylRec *>(bvh.prims)[inv[id]], p);
    }
  }
  out[i] = d;
}

__global__ void k_leaf_vertices(const uint32_t *__restrict__ order, const uint32_t *__restrict__ F,
                                const double *__restrict__ V, uint32_t n, double *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *f = F + 3ull * order[i];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) out[9ull * i + 3 * c + k] = V[3ull * f[c] + k];
}

__global__ void k_inverse_order(const uint32_t *__restrict__ order, uint32_t n, uint32_t *__restrict__ inv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[order[i]] = i;
}

// ------------------------------------------------------------------------------------------------ closest
struct Best {
  double d;
  uint32_t id;
};

// within one geometry type the smaller distance wins, then the lower id (NaN never wins)
__device__ __forceinline__ void consider(Best &b, double d, uint32_t id) {
  if (d < b.d || (d == b.d && id < b.id)) {
    b.d = d;
    b.id = id;
  }
}

__device__ __forceinline__ double box_gap(const double *lo, const double *hi, const double *p) {
  double s = 0.0;
  for (int k = 0; k < 3; k++) {
    double g = 0.0;
    if (lo[k] - p[k] > g) g = lo[k] - p[k];
    if (p[k] - hi[k] > g) g = p[k] - hi[k];
    s += g * g;
  }
  return __builtin_sqrt(s);
}

// Traversal stack of the closest-surface query: the newest CLOSEST_RING entries of each point in LDS, deeper ones
// in a private spill area (trees are at most 63 levels deep on the device, so 64 entries always suffice).
constexpr int CLOSEST_RING = 8;
constexpr uint32_t CLOSEST_BLOCK = 128;
using ClosestStack = blazert_amd::RingStack<double, CLOSEST_RING>;

// Depth-first search of one BVH, nearer child first. `bound` is the best distance of earlier geometry types.
template <int TYPE>
__device__ void closest_in_bvh(const BvhView &bvh, const double *leaf_vertices, const double *p, double margin,
                               double bound, Best &best, ClosestStack &st) {
  st.sp = st.lo = 0;
  auto limit = [&]() { return best.d < bound ? best.d : bound; };
  if (box_gap(bvh.root_min, bvh.root_max, p) - margin > limit()) return;
  uint32_t ref = bvh.root_ref, cnt = bvh.root_cnt;
  while (true) {
    if (ref & LEAF_BIT) {
      const uint32_t first = ref & ~LEAF_BIT;
      for (uint32_t k = 0; k < cnt; k++) {
        // triangles read their vertices from the leaf-order copy: one gather instead of id -> face -> vertices
        const uint32_t id = bvh.prim_ids[first + k];
        const double *lv = leaf_vertices + 9ull * (first + k);
        const double d = (TYPE == TRIANGLES) ? triangle_distance(lv, lv + 3, lv + 6, p)
                                             : sphere_distance(reinterpret_cast<const SphRec *>(bvh.prims)[first + k], p);
        consider(best, d, id);
      }
    } else {
      const Node64 &nd = bvh.nodes64[ref];
      const double g0 = box_gap(nd.b, nd.b + 3, p), g1 = box_gap(nd.b + 6, nd.b + 9, p);
      const uint32_t r0 = nd.ref[0], r1 = nd.ref[1], c0 = nd.cnt[0], c1 = nd.cnt[1] & CNT_MASK;
      const double lim = limit();
      const bool ok0 = !(g0 - margin > lim), ok1 = !(g1 - margin > lim);
      if (ok0 && ok1) {
        const bool first0 = !(g1 < g0);
        st.push(first0 ? r1 : r0, first0 ? c1 : c0, first0 ? g1 : g0);
        ref = first0 ? r0 : r1;
        cnt = first0 ? c0 : c1;
        continue;
      }
      if (ok0 || ok1) {
        ref = ok0 ? r0 : r1;
        cnt = ok0 ? c0 : c1;
        continue;
      }
    }
    // pop the next entry still within reach
    bool found = false;
    while (!st.empty()) {
      double key;
      st.pop(ref, cnt, key);
      if (!(key - margin > limit())) {
        found = true;
        break;
      }
    }
    if (!found) return;
  }
}

__device__ __forceinline__ double rec_distance(const SphRec &r, const double *p) { return sphere_distance(r, p); }
__device__ __forceinline__ double rec_distance(const PlaRec &r, const double *p) { return plane_distance(r, p); }
__device__ __forceinline__ double rec_distance(const CylRec &r, const double *p) { return cylinder_distance(r, p); }

__device__ void closest_scan_triangles(const BvhView &bvh, const double *leaf_vertices, const double *p, Best &best) {
  for (uint32_t k = 0; k < bvh.n_prims; k++) {
    const double *lv = leaf_vertices + 9ull * k;
    consider(best, triangle_distance(lv, lv + 3, lv + 6, p), bvh.prim_ids[k]);
  }
}

template <class Rec>
__device__ void closest_scan(const BvhView &bvh, const double *p, Best &best) {
  const Rec *recs = reinterpret_cast<const Rec *>(bvh.prims);
  for (uint32_t k = 0; k < bvh.n_prims; k++) consider(best, rec_distance(recs[k], p), bvh.prim_ids[k]);
}

__global__ void k_closest_surface(ClosestViews V, const double *__restrict__ px, const double *__restrict__ py,
                                  const double *__restrict__ pz, uint64_t n, double *__restrict__ dist,
                                  uint32_t *__restrict__ prim, uint32_t *__restrict__ geom) {
  __shared__ double lkey[CLOSEST_RING * CLOSEST_BLOCK];
  __shared__ uint32_t lref[CLOSEST_RING * CLOSEST_BLOCK], lcnt[CLOSEST_RING * CLOSEST_BLOCK];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  blazert_amd::StackArea<double> spill;
  ClosestStack st{(blazert_amd::LdsPtr<double>)(lkey + threadIdx.x), (blazert_amd::LdsPtr<uint32_t>)(lref + threadIdx.x),
                  (blazert_amd::LdsPtr<uint32_t>)(lcnt + threadIdx.x), CLOSEST_BLOCK, &spill};
  const double p[3] = {px[i], py[i], pz[i]};
  const double pm = fmax(fabs(p[0]), fmax(fabs(p[1]), fabs(p[2])));
  double gd = __builtin_inf();
  uint32_t gp = INVALID_ID, gg = INVALID_ID;
  for (int t = 0; t < NUM_TYPES; t++) {
    const BvhView &bvh = V.view[t];
    if (!V.present[t] || bvh.n_prims == 0) continue;
    Best b{__builtin_inf(), INVALID_ID};
    const double margin = 1e-10 * (pm + V.extent[t]);
    // box pruning needs boxes that contain their primitives: a tree with an inverted or NaN node box (a sphere with a
    // negative radius) is scanned in full instead
    if (t == TRIANGLES && !bvh.irregular_boxes) closest_in_bvh<TRIANGLES>(bvh, V.leaf_vertices, p, margin, gd, b, st);
    else if (t == TRIANGLES) closest_scan_triangles(bvh, V.leaf_vertices, p, b);
    else if (t == SPHERES && !bvh.irregular_boxes) closest_in_bvh<SPHERES>(bvh, nullptr, p, margin, gd, b, st);
    else if (t == SPHERES) closest_scan<SphRec>(bvh, p, b);
    else if (t == PLANES) closest_scan<PlaRec>(bvh, p, b);
    else closest_scan<CylRec>(bvh, p, b);
    if (b.d < gd) {  // a later type replaces only when strictly closer
      gd = b.d;
      gp = b.id;
      gg = bvh.geom_id;
    }
  }
  dist[i] = gd;
  prim[i] = gp;
  geom[i] = gg;
}

}  // namespace

hipError_t launch_inverse_order(const uint32_t *order, uint32_t n, uint32_t *inv, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_inverse_order<<<(n + 255) / 256, 256, 0, stream>>>(order, n, inv);
  return hipGetLastError();
}

hipError_t launch_leaf_vertices(const uint32_t *order, const uint32_t *faces, const double *vertices, uint32_t n,
                                double *out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_leaf_vertices<<<(n + 255) / 256, 256, 0, stream>>>(order, faces, vertices, n, out);
  return hipGetLastError();
}

hipError_t launch_distance_pairs(GeomType type, const BvhView &bvh, const uint32_t *inv, const double *px,
                                 const double *py, const double *pz, const uint32_t *ids, uint64_t n, double *out,
                                 hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)((n + 255) / 256);
  switch (type) {
  case TRIANGLES: k_distance_pairs<TRIANGLES><<<blocks, 256, 0, stream>>>(bvh, inv, px, py, pz, ids, n, out); break;
  case SPHERES: k_distance_pairs<SPHERES><<<blocks, 256, 0, stream>>>(bvh, inv, px, py, pz, ids, n, out); break;
  case PLANES: k_distance_pairs<PLANES><<<blocks, 256, 0, stream>>>(bvh, inv, px, py, pz, ids, n, out); break;
  default: k_distance_pairs<CYLINDERS><<<blocks, 256, 0, stream>>>(bvh, inv, px, py, pz, ids, n, out); break;
  }
  return hipGetLastError();
}

hipError_t launch_closest_surface(const ClosestViews &views, const double *px, const double *py, const double *pz,
                                  uint64_t n, double *dist, uint32_t *prim, uint32_t *geom, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_closest_surface<<<(uint32_t)((n + CLOSEST_BLOCK - 1) / CLOSEST_BLOCK), CLOSEST_BLOCK, 0, stream>>>(
      views, px, py, pz, n, dist, prim, geom);
  return hipGetLastError();
}

}  // namespace brt
