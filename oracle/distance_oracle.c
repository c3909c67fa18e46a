/*
 * distance_oracle.c -- CPU oracle of blazeRT's distance_to_surface functions (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in the reference's operation order (contraction off, FP64):
 *   spheres    blazert/primitives/sphere.h:179-182   |norm(center - p) - radius|
 *   planes     blazert/primitives/plane.h:272-284    local = trans(R) (p - c); |local.z| when local.x lies in
 *              [c.x - dx/2, c.x + dx/2] and local.y in [c.y - dy/2, c.y + dy/2] (the reference compares the local
 *              coordinates with the plane's world-space center), else DBL_MAX
 *   cylinders  blazert/primitives/cylinder.h:507-759 (ellipse_equation, distance_derivative,
 *              single_newton_iteration, newton_iteration, find_minimum_distance_to_ellipse, distance_to_surface)
 *   triangles  blazert/primitives/trimesh_distance.h:36-221 (closest_point_on_triangle, Eberly's regions r0-r6);
 *              distance = norm(p - closest point). The reference's triangle distance_to_surface
 *              (blazert/primitives/trimesh.h:183-191) cannot be instantiated: it returns the closest point as a
 *              scalar, and rebuilds the vertices as (v0, v2, -v1) from the Triangle's stored edges. This oracle
 *              follows its evident intent with the face's own vertices.
 * The scene-level query (closest surface over every primitive of the scene) has no counterpart in the reference;
 * here it is the brute-force minimum, ties going to the earlier geometry type, then to the lower primitive id.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "blazert_oracle.h"

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* trans(R) * v for row-major R */
static inline void mat_t_vec(double *r, const double *R, const double *v) {
  const double x = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  const double y = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  const double z = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
static inline double min2(double a, double b) { return (b < a) ? b : a; } /* std::min(a, b) */
/* std::min({...}): the first smallest element */
static inline double min_list(const double *v, int n) {
  double m = v[0];
  for (int i = 1; i < n; i++)
    if (v[i] < m) m = v[i];
  return m;
}

/* ------------------------------------------------------------------------------------------------ spheres */
double orc_sphere_distance(const double c[3], double r, const double p[3]) {
  const double v[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
  return fabs(sqrt(dot3(v, v)) - r);
}

/* ------------------------------------------------------------------------------------------------ planes */
double orc_plane_distance(const double c[3], double dx_full, double dy_full, const double R[9], const double p[3]) {
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  double l[3];
  mat_t_vec(l, R, v);
  const double dx = dx_full / 2, dy = dy_full / 2;
  const double cx = c[0], cy = c[1];
  if ((l[0] >= cx - dx) && (l[0] <= cx + dx) && (l[1] >= cy - dy) && (l[1] <= cy + dy)) return fabs(l[2]);
  return DBL_MAX;
}

/* ------------------------------------------------------------------------------------------------ cylinders */
enum { UPPER_X = 0, LOWER_X = 1, UPPER_Y = 2, LOWER_Y = 3 };

/* cylinder.h:507-512 */
static double ellipse_equation(double a, double b, double x, double y) { return (x * x) / (a * a) + (y * y) / (b * b) - 1; }

/* cylinder.h:517-545 */
static double distance_derivative(int kind, double a, double b, double xq, double yq, double val) {
  if (kind == UPPER_X || kind == LOWER_X) {
    const double s = (val * val > a * a) ? sqrt(val * val - a * a) : sqrt(a * a - val * val);
    if (kind == UPPER_X) return -2 * (xq - val) - 2 * b * val * (b / a * s - yq) / (a * s);
    return -2 * (xq - val) - 2 * b * val * (b / a * s + yq) / (a * s);
  }
  const double s = (val * val > b * b) ? sqrt(val * val - b * b) : sqrt(b * b - val * val);
  if (kind == UPPER_Y) return -2 * (yq - val) - 2 * a * val * (a / b * s + xq) / (b * s);
  return -2 * (yq - val) - 2 * a * val * (a / b * s - xq) / (b * s);
}

/* cylinder.h:548-583: x - R'(x) / R''(x) */
static double single_newton_iteration(int kind, double a, double b, double xq, double yq, double val) {
  if (kind == UPPER_X || kind == LOWER_X) {
    const double s = (val * val > a * a) ? sqrt(val * val - a * a) : sqrt(a * a - val * val);
    if (kind == UPPER_X)
      return ((a * a * a * xq - a * val * val * xq) * s + b * val * val * val * yq) /
             ((a * a * a - a * b * b - a * val * val + b * b * val * val / a) * s + a * a * b * yq);
    return ((-a * a * a * xq + a * val * val * xq) * s + b * val * val * val * yq) /
           ((-a * a * a + a * b * b + a * val * val - b * b * val * val / a) * s + a * a * b * yq);
  }
  const double s = (val * val > b * b) ? sqrt(val * val - b * b) : sqrt(b * b - val * val);
  if (kind == UPPER_Y)
    return ((b * b * b * yq - b * val * val * yq) * s + a * val * val * val * xq) /
           ((b * b * b - b * a * a - b * val * val + a * a * val * val / b) * s + b * b * a * xq);
  return ((-b * b * b * yq + b * val * val * yq) * s + a * val * val * val * xq) /
         ((-b * b * b + b * a * a + b * val * val - a * a * val * val / b) * s + b * b * a * xq);
}

/* cylinder.h:586-607 (stop at |R'| < 1e-7 or after 100 steps; |x| == a is nudged by one epsilon) */
static double newton_iteration(int kind, double a, double b, double xq, double yq, double x0) {
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

/* cylinder.h:615-665: distance from (xq, yq) to the ellipse x^2/a^2 + y^2/b^2 = 1 */
static double min_distance_to_ellipse(double a, double b, double xq, double yq) {
  const double percentage = 0.95;
  if (((xq >= 0) && (yq >= (b / a) * xq)) || ((xq <= 0) && (yq >= (-b / a) * xq))) {
    const double x0 = xq > 0 ? percentage * a : -percentage * a;
    const double x_min = newton_iteration(UPPER_X, a, b, xq, yq, x0);
    const double s = sqrt(a * a - x_min * x_min);
    const double y_min = b / a * s;
    return sqrt((xq - x_min) * (xq - x_min) + (yq - y_min) * (yq - y_min));
  } else if (((xq >= 0) && (yq <= (-b / a) * xq)) || ((xq <= 0) && (yq <= (b / a) * xq))) {
    const double x0 = xq > 0 ? percentage * a : -percentage * a;
    const double x_min = newton_iteration(LOWER_X, a, b, xq, yq, x0);
    const double s = sqrt(a * a - x_min * x_min);
    const double y_min = -b / a * s;
    return sqrt((xq - x_min) * (xq - x_min) + (yq - y_min) * (yq - y_min));
  } else if ((xq >= 0) && (yq <= (b / a) * xq) && (yq >= (-b / a) * xq)) {
    const double y0 = yq > 0 ? percentage * b : -percentage * b;
    const double y_min = newton_iteration(UPPER_Y, a, b, xq, yq, y0);
    const double s = sqrt(b * b - y_min * y_min);
    const double x_min = a / b * s;
    return sqrt((yq - y_min) * (yq - y_min) + (xq - x_min) * (xq - x_min));
  } else {
    const double y0 = yq > 0 ? percentage * b : -percentage * b;
    const double y_min = newton_iteration(LOWER_Y, a, b, xq, yq, y0);
    const double s = sqrt(b * b - y_min * y_min);
    const double x_min = -a / b * s;
    return sqrt((yq - y_min) * (yq - y_min) + (xq - x_min) * (xq - x_min));
  }
}

/* cylinder.h:677-759 */
double orc_cylinder_distance(const double c[3], double a, double b, double h, const double R[9], const double p[3]) {
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  double l[3];
  mat_t_vec(l, R, v);
  const double dist_top = fabs(l[2] - h / 2);
  const double dist_bottom = fabs(l[2] + h / 2);
  const double xq = l[0], yq = l[1], zq = l[2];
  if (xq == 0 && yq == 0) { /* isZero(point_xy) */
    if ((zq > h / 2) || (zq < -h / 2)) {
      const double m[2] = {dist_top, dist_bottom};
      return min_list(m, 2);
    }
    const double m[4] = {a, b, dist_top, dist_bottom};
    return min_list(m, 4);
  }
  const double ellipse = ellipse_equation(a, b, xq, yq);
  if (fabs(ellipse) < DBL_EPSILON) {
    if ((zq > h / 2) || (zq < -h / 2)) return min2(dist_top, dist_bottom);
    return 0.0;
  }
  if (ellipse < 0) {
    if (zq >= h / 2) return dist_top;
    if (zq <= -h / 2) return dist_bottom;
    const double m[3] = {min_distance_to_ellipse(a, b, xq, yq), dist_bottom, dist_top};
    return min_list(m, 3);
  }
  if (ellipse > 0) {
    if ((zq < h / 2) && (zq > -h / 2)) return min_distance_to_ellipse(a, b, xq, yq);
    const double ds = min_distance_to_ellipse(a, b, xq, yq);
    const double m[2] = {sqrt(ds * ds + dist_top * dist_top), sqrt(ds * ds + dist_bottom * dist_bottom)};
    return min_list(m, 2);
  }
  return -1;
}

/* ------------------------------------------------------------------------------------------------ triangles */
/* trimesh_distance.h:176-221 with the region helpers r0-r6 (:36-171) inlined in the same order */
void orc_closest_point_on_triangle(const double v0[3], const double v1[3], const double v2[3], const double p[3],
                                   double q[3]) {
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
      if (t < 0) { /* r4 */
        if (d < 0) {
          t = 0;
          s = (-d >= a) ? 1 : -d / a;
        } else {
          s = 0;
          if (e >= 0) t = 0;
          else t = (-e >= c) ? 1 : -e / c;
        }
      } else { /* r3 */
        s = 0;
        if (e >= 0) t = 0;
        else t = (-e >= c) ? 1 : -e / c;
      }
    } else if (t < 0) { /* r5 */
      t = 0;
      if (d >= 0) s = 0;
      else s = (-d >= a) ? 1 : -d / a;
    } else { /* r0 */
      const double inv_det = 1. / det;
      s *= inv_det;
      t *= inv_det;
    }
  } else {
    if (s < 0) { /* r2 */
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
    } else if (t < 0) { /* r6 */
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
    } else { /* r1 */
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
  for (int k = 0; k < 3; k++) q[k] = v0[k] + s * E0[k] + t * E1[k];
}

double orc_triangle_distance(const double v0[3], const double v1[3], const double v2[3], const double p[3]) {
  double q[3];
  orc_closest_point_on_triangle(v0, v1, v2, p, q);
  const double w[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  return sqrt(dot3(w, w));
}

/* ------------------------------------------------------------------------------------------------ batches */
static double prim_distance(const orc_scene *s, int type, uint32_t prim, const double p[3]) {
  double prm[18];
  if (orc_prim_params(s, type, prim, prm) != 0) return NAN;
  switch (type) {
  case ORC_TRIANGLES: return orc_triangle_distance(prm, prm + 3, prm + 6, p);
  case ORC_SPHERES: return orc_sphere_distance(prm, prm[3], p);
  case ORC_PLANES: return orc_plane_distance(prm, prm[3], prm[4], prm + 5, p);
  default: return orc_cylinder_distance(prm, prm[3], prm[4], prm[5], prm + 6, p);
  }
}

void orc_distance_batch(const orc_scene *s, int type, uint64_t n, const double *px, const double *py,
                        const double *pz, const uint32_t *prims, double *out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const double p[3] = {px[i], py[i], pz[i]};
    out[i] = prim_distance(s, type, prims[i], p);
  }
}

void orc_closest_batch(const orc_scene *s, uint64_t n, const double *px, const double *py, const double *pz,
                       double *dist, uint32_t *prim, uint32_t *geom, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const double p[3] = {px[i], py[i], pz[i]};
    double best = INFINITY;
    uint32_t bp = 0xFFFFFFFFu, bg = 0xFFFFFFFFu;
    for (int type = 0; type < ORC_NUM_TYPES; type++) {
      const uint32_t count = orc_prim_count(s, type);
      const uint32_t gid = orc_geom_id(s, type);
      for (uint32_t k = 0; k < count; k++) {
        const double d = prim_distance(s, type, k, p);
        if (d < best) { /* ascending ids and types: the first of equal distances stays */
          best = d;
          bp = k;
          bg = gid;
        }
      }
    }
    dist[i] = best;
    prim[i] = bp;
    geom[i] = bg;
  }
}
