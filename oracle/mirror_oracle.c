/*
 * mirror_oracle.c -- CPU oracle of multi-bounce mirror tracing and its gradient (TEST INFRASTRUCTURE ONLY).
 *
 * The reference has no differentiable path; this file extends its intersect1 (blazert/scene.h:161-211, restated in
 * blazert_oracle.c) with the mirror bounce that blazert_amd's mirror tracing implements, and gives an analytic FP64
 * gradient written independently of the HIP adjoint: explicit Jacobian matrices of each bounce, with the hit
 * distance's vertex derivative in its barycentric closed form (dt/dp_j = beta_j nh / (nh.dh)) instead of the
 * reverse sweep through t = n.(p0 - o) / n.dh that the device kernel uses.
 *
 * Bounce (state o, raw d; face p0, p1, p2):
 *   dh = normalize(d), nh = normalize(cross(p1 - p0, p2 - p0)) (the reference's face normal,
 *   blazert/primitives/trimesh.h:125-129), x = o + t dh, d' = dh - 2 (dh.nh) nh.
 * Forward tracing takes t from intersect1 (bit-exact with the reference path); the differentiable chain used by the
 * gradient and by mirror_loss takes t from the plane of the stored face.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "blazert_oracle.h"

#define INV 0xFFFFFFFFu

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(double *r, const double *a, const double *b) {
  const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  r[0] = x;
  r[1] = y;
  r[2] = z;
}
/* normalize(a) = a * (len != 0 ? 1/len : 0) */
static inline double normalize3(double *r, const double *a) {
  const double len = sqrt(dot3(a, a));
  const double il = (len != 0.0) ? 1.0 / len : 0.0;
  for (int k = 0; k < 3; k++) r[k] = a[k] * il;
  return len;
}
static inline void vert(double *p, const double *V, uint32_t v) {
  for (int k = 0; k < 3; k++) p[k] = V[3ull * v + k];
}

/* mirror step after a hit at distance t on face f (same operation order as the device kernel) */
static void mirror_step(const double *V, const uint32_t *F, uint32_t f, const double *o, const double *d, double t,
                        double *x, double *r) {
  double dh[3], p0[3], p1[3], p2[3], a[3], b[3], n[3], nh[3];
  normalize3(dh, d);
  for (int k = 0; k < 3; k++) x[k] = o[k] + t * dh[k];
  vert(p0, V, F[3ull * f]);
  vert(p1, V, F[3ull * f + 1]);
  vert(p2, V, F[3ull * f + 2]);
  for (int k = 0; k < 3; k++) {
    a[k] = p1[k] - p0[k];
    b[k] = p2[k] - p0[k];
  }
  cross3(n, a, b);
  normalize3(nh, n);
  const double k2 = 2.0 * dot3(dh, nh);
  for (int k = 0; k < 3; k++) r[k] = dh[k] - k2 * nh[k];
}

void orc_mirror_trace(const orc_scene *s, const double *V, const uint32_t *F, uint64_t n, const double *ox,
                      const double *oy, const double *oz, const double *dx, const double *dy, const double *dz,
                      const double *tmin, const double *tmax, uint32_t bounces, double bounce_tmin, double *t,
                      double *u, double *v, uint32_t *prim, double *px, double *py, double *pz, double *rx,
                      double *ry, double *rz, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
    double lo = tmin ? tmin[i] : 0.0, hi = tmax ? tmax[i] : DBL_MAX;
    int alive = 1;
    for (uint32_t b = 0; b < bounces; b++) {
      const uint64_t k = (uint64_t)b * n + (uint64_t)i;
      orc_hit h;
      memset(&h, 0, sizeof h);
      h.t = DBL_MAX;
      h.prim_id = h.geom_id = INV;
      if (alive) {
        orc_ray ray;
        orc_ray_init(&ray, o, d, lo, hi, 0);
        alive = orc_intersect1(s, &ray, &h, NULL);
      }
      t[k] = h.t;
      if (u) u[k] = h.u;
      if (v) v[k] = h.v;
      prim[k] = h.prim_id;
      if (!alive) {
        px[k] = py[k] = pz[k] = rx[k] = ry[k] = rz[k] = NAN;
        continue;
      }
      double x[3], r[3];
      mirror_step(V, F, h.prim_id, o, d, h.t, x, r);
      px[k] = x[0];
      py[k] = x[1];
      pz[k] = x[2];
      rx[k] = r[0];
      ry[k] = r[1];
      rz[k] = r[2];
      memcpy(o, x, sizeof o);
      memcpy(d, r, sizeof d);
      lo = bounce_tmin;
      hi = DBL_MAX;
    }
  }
}

/* ---- differentiable chain with fixed face ids --------------------------------------------------------------- */
typedef struct bounce_state {
  double o[3], d[3];          /* state entering the bounce (raw direction) */
  double p[3][3];             /* face corners */
  double dh[3], s;            /* unit direction and |d| */
  double nh[3], nl;           /* unit face normal and |cross| */
  double k, t, x[3], dn[3];   /* dh.nh, plane distance, hit point, mirror direction */
  double beta[3];             /* barycentric coordinates of x */
} bounce_state;

static void chain_bounce(const double *V, const uint32_t *F, uint32_t f, const double *o, const double *d,
                         bounce_state *B) {
  memcpy(B->o, o, sizeof B->o);
  memcpy(B->d, d, sizeof B->d);
  for (int j = 0; j < 3; j++) vert(B->p[j], V, F[3ull * f + j]);
  B->s = normalize3(B->dh, d);
  double a[3], b[3], n[3], c[3];
  for (int k = 0; k < 3; k++) {
    a[k] = B->p[1][k] - B->p[0][k];
    b[k] = B->p[2][k] - B->p[0][k];
    c[k] = B->p[0][k] - o[k];
  }
  cross3(n, a, b);
  B->nl = normalize3(B->nh, n);
  B->k = dot3(B->dh, B->nh);
  B->t = dot3(B->nh, c) / B->k;
  for (int k = 0; k < 3; k++) {
    B->x[k] = o[k] + B->t * B->dh[k];
    B->dn[k] = B->dh[k] - 2.0 * B->k * B->nh[k];
  }
  /* x - p0 = b1 a + b2 b (Gram system) */
  double w[3];
  for (int k = 0; k < 3; k++) w[k] = B->x[k] - B->p[0][k];
  const double aa = dot3(a, a), ab = dot3(a, b), bb = dot3(b, b), aw = dot3(a, w), bw = dot3(b, w);
  const double det = aa * bb - ab * ab;
  B->beta[1] = (aw * bb - bw * ab) / det;
  B->beta[2] = (aa * bw - ab * aw) / det;
  B->beta[0] = 1.0 - B->beta[1] - B->beta[2];
}

/* number of leading bounces with a valid face id */
static uint32_t active_bounces(const uint32_t *prims, uint64_t n, uint64_t i, uint32_t bounces) {
  uint32_t nb = 0;
  while (nb < bounces && prims[(uint64_t)nb * n + i] != INV) nb++;
  return nb;
}

void orc_mirror_loss(const double *V, const uint32_t *F, uint64_t n, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz, uint32_t bounces,
                     const uint32_t *prims, const double *gt, const double *gpx, const double *gpy,
                     const double *gpz, double *loss, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
    const uint32_t nb = active_bounces(prims, n, (uint64_t)i, bounces);
    double L = 0.0;
    for (uint32_t b = 0; b < nb; b++) {
      const uint64_t k = (uint64_t)b * n + (uint64_t)i;
      bounce_state B;
      chain_bounce(V, F, prims[k], o, d, &B);
      if (gt) L += gt[k] * B.t;
      if (gpx) L += gpx[k] * B.x[0] + gpy[k] * B.x[1] + gpz[k] * B.x[2];
      memcpy(o, B.x, sizeof o);
      memcpy(d, B.dn, sizeof d);
    }
    loss[i] = L;
  }
}

/* 3x3 helpers: y = M^T v */
static inline void mtv(double *y, const double M[3][3], const double *v) {
  for (int c = 0; c < 3; c++) y[c] = M[0][c] * v[0] + M[1][c] * v[1] + M[2][c] * v[2];
}
static inline void matmul(double R[3][3], const double A[3][3], const double B[3][3]) {
  double T[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) T[r][c] = A[r][0] * B[0][c] + A[r][1] * B[1][c] + A[r][2] * B[2][c];
  memcpy(R, T, sizeof T);
}
/* [a]_x with [a]_x v = a x v */
static inline void skew(double S[3][3], const double *a) {
  S[0][0] = 0.0;
  S[0][1] = -a[2];
  S[0][2] = a[1];
  S[1][0] = a[2];
  S[1][1] = 0.0;
  S[1][2] = -a[0];
  S[2][0] = -a[1];
  S[2][1] = a[0];
  S[2][2] = 0.0;
}

void orc_mirror_grad(const double *V, const uint32_t *F, uint64_t n, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz, uint32_t bounces,
                     const uint32_t *prims, const double *gt, const double *gpx, const double *gpy,
                     const double *gpz, double *gv, double *go, double *gd, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    bounce_state st[8];
    const uint32_t nb = active_bounces(prims, n, (uint64_t)i, bounces > 8 ? 8 : bounces);
    double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
    for (uint32_t b = 0; b < nb; b++) {
      chain_bounce(V, F, prims[(uint64_t)b * n + (uint64_t)i], o, d, &st[b]);
      memcpy(o, st[b].x, sizeof o);
      memcpy(d, st[b].dn, sizeof d);
    }
    double ax[3] = {0, 0, 0}, ad[3] = {0, 0, 0}; /* adjoints of the next bounce's origin / raw direction */
    for (int b = (int)nb - 1; b >= 0; b--) {
      const bounce_state *B = &st[b];
      const uint64_t k = (uint64_t)b * n + (uint64_t)i;
      const double g_t = gt ? gt[k] : 0.0;
      double gx[3] = {ax[0], ax[1], ax[2]};
      if (gpx) {
        gx[0] += gpx[k];
        gx[1] += gpy[k];
        gx[2] += gpz[k];
      }
      /* Jacobians */
      double Jt_o[3], Jt_dh[3], Jt_p[3][3];
      for (int c = 0; c < 3; c++) {
        Jt_o[c] = -B->nh[c] / B->k;
        Jt_dh[c] = -B->t * B->nh[c] / B->k;
        for (int j = 0; j < 3; j++) Jt_p[j][c] = B->beta[j] * B->nh[c] / B->k;
      }
      double Jx_o[3][3], Jx_dh[3][3], H[3][3], P[3][3], Q[3][3], Dn[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          const double I = (r == c) ? 1.0 : 0.0;
          Jx_o[r][c] = I + B->dh[r] * Jt_o[c];
          Jx_dh[r][c] = B->t * I + B->dh[r] * Jt_dh[c];
          H[r][c] = I - 2.0 * B->nh[r] * B->nh[c];
          P[r][c] = (I - B->dh[r] * B->dh[c]) / B->s;
          Q[r][c] = (I - B->nh[r] * B->nh[c]) / B->nl;
          Dn[r][c] = -2.0 * (B->k * I + B->nh[r] * B->dh[c]);
        }
      /* dn/dp_j = skew(e_j) with e_0 = p2 - p1, e_1 = p0 - p2, e_2 = p1 - p0 (n = (p1 - p0) x (p2 - p0)) */
      double E[3][3];
      for (int c = 0; c < 3; c++) {
        E[0][c] = B->p[2][c] - B->p[1][c];
        E[1][c] = B->p[0][c] - B->p[2][c];
        E[2][c] = B->p[1][c] - B->p[0][c];
      }
      /* origin and direction */
      double go_b[3], gdh[3], tmp[3];
      mtv(go_b, Jx_o, gx);
      mtv(gdh, Jx_dh, gx);
      mtv(tmp, H, ad);
      for (int c = 0; c < 3; c++) {
        go_b[c] += g_t * Jt_o[c];
        gdh[c] += g_t * Jt_dh[c] + tmp[c];
      }
      double gd_b[3];
      mtv(gd_b, P, gdh);
      /* vertices */
      if (gv) {
        const double dxg = dot3(B->dh, gx); /* Jx_p = dh Jt_p^T, so Jx_p^T gx = (dh.gx) Jt_p */
        for (int j = 0; j < 3; j++) {
          double S[3][3], M[3][3], gpj[3];
          skew(S, E[j]);
          matmul(M, Q, S);
          matmul(M, Dn, M);
          mtv(gpj, M, ad);
          const uint32_t vid = F[3ull * prims[k] + j];
          for (int c = 0; c < 3; c++) {
            const double val = (dxg + g_t) * Jt_p[j][c] + gpj[c];
#ifdef _OPENMP
#pragma omp atomic
#endif
            gv[3ull * vid + c] += val;
          }
        }
      }
      memcpy(ax, go_b, sizeof ax);
      memcpy(ad, gd_b, sizeof ad);
    }
    if (go) {
      go[i] = ax[0];
      go[n + i] = ax[1];
      go[2 * n + i] = ax[2];
    }
    if (gd) {
      gd[i] = ad[0];
      gd[n + i] = ad[1];
      gd[2 * n + i] = ad[2];
    }
  }
}
