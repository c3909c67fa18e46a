// HIP kernels for differentiable mirror tracing on a triangle mesh (gfx950).
//
// Forward: bounce b traces its rays with the closest-hit kernels (exact FP64 drop-in path), then k_mirror_step
// turns each hit into the next bounce's ray: origin = hit position o + t * normalize(d), direction = the mirror
// image of normalize(d) about the unit face normal normalize(cross(v1 - v0, v2 - v0)) (the reference's face normal,
// blazert/primitives/trimesh.h:125-129). A miss ends the ray: its position and direction become NaN, which every
// later bounce rejects at the root box.
//
// Adjoint: one thread per ray re-derives the whole bounce chain from the stored face ids alone (t from the plane
// of the stored face, no intermediates kept between passes), then runs the chain backwards in FP64: dL/d(origin,
// raw direction) per ray, and dL/dvertex scatter-added per face corner. The scatter goes through a per-block LDS
// table (open addressing keyed by vertex) so that vertices hit by many rays of a block (large walls) cost one global
// atomic per block instead of one per ray; corners whose vertex finds no slot are added to global memory by the
// wave together, three lanes per vertex (x, y, z in one request).
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "engine.h"

namespace brt {

namespace {

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return d3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ D3 vertex(const double *V, uint32_t v) {
  return d3(V[3ull * v], V[3ull * v + 1], V[3ull * v + 2]);
}

// normalize(a) = a * (len != 0 ? 1/len : 0), len = sqrt(a.a)
__device__ __forceinline__ D3 normalized(D3 a, double &len) {
  len = __builtin_sqrt(dot(a, a));
  const double il = (len != 0.0) ? 1.0 / len : 0.0;
  return d3(a.x * il, a.y * il, a.z * il);
}

__global__ void __launch_bounds__(256) k_mirror_step(RaysView rays, HitsView hits, const double *__restrict__ V,
                                                     const uint32_t *__restrict__ F, uint64_t n, double *px,
                                                     double *py, double *pz, double *rx, double *ry, double *rz) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t f = hits.prim_id[i];
    if (f == INVALID_ID) {
      const double nan = __builtin_nan("");
      px[i] = py[i] = pz[i] = nan;
      rx[i] = ry[i] = rz[i] = nan;
      continue;
    }
    double len;
    const D3 o = d3(rays.ox[i], rays.oy[i], rays.oz[i]);
    const D3 dh = normalized(d3(rays.dx[i], rays.dy[i], rays.dz[i]), len);
    const double t = hits.t[i];
    const D3 x = o + t * dh;
    const D3 p0 = vertex(V, F[3ull * f]), p1 = vertex(V, F[3ull * f + 1]), p2 = vertex(V, F[3ull * f + 2]);
    const D3 nh = normalized(cross(p1 - p0, p2 - p0), len);
    const double k2 = 2.0 * dot(dh, nh);
    const D3 r = dh - k2 * nh;
    px[i] = x.x;
    py[i] = x.y;
    pz[i] = x.z;
    rx[i] = r.x;
    ry[i] = r.y;
    rz[i] = r.z;
  }
}

// Differentiable form of one bounce from state (o, raw d) on face (p0, p1, p2):
//   dh = d / |d|, n = (p1 - p0) x (p2 - p0), t = n.(p0 - o) / n.dh, x = o + t dh,
//   nh = n / |n|, d' = dh - 2 (dh.nh) nh.
struct Bounce {
  D3 dh, n, c, nh;
  double s, t, den, nl, k;
};

__device__ __forceinline__ void bounce_forward(D3 o, D3 d, D3 p0, D3 p1, D3 p2, Bounce &B, D3 &x, D3 &dn) {
  B.dh = normalized(d, B.s);
  B.n = cross(p1 - p0, p2 - p0);
  B.c = p0 - o;
  B.den = dot(B.n, B.dh);
  B.t = dot(B.n, B.c) / B.den;
  x = o + B.t * B.dh;
  B.nh = normalized(B.n, B.nl);
  B.k = dot(B.dh, B.nh);
  dn = B.dh - (2.0 * B.k) * B.nh;
}

// Reverse of bounce_forward: upstream gt (dL/dt), gx (dL/dx), gdn (dL/dd'); returns dL/do, dL/dd and dL/dp0..2.
__device__ __forceinline__ void bounce_backward(const Bounce &B, D3 p0, D3 p1, D3 p2, double gt, D3 gx, D3 gdn,
                                                D3 &go, D3 &gd, D3 &g0, D3 &g1, D3 &g2) {
  // d' = dh - 2 k nh, k = dh.nh
  const double gk = -2.0 * dot(gdn, B.nh);
  D3 gdh = gdn + gk * B.nh;
  D3 gnh = (-2.0 * B.k) * gdn + gk * B.dh;
  // nh = n / |n|
  D3 gn = (1.0 / B.nl) * (gnh - dot(gnh, B.nh) * B.nh);
  // x = o + t dh
  go = gx;
  const double gtt = gt + dot(gx, B.dh);
  gdh = gdh + B.t * gx;
  // t = num / den, num = n.c, den = n.dh, c = p0 - o
  const double gnum = gtt / B.den;
  const double gden = -gtt * B.t / B.den;
  gn = gn + gnum * B.c + gden * B.dh;
  const D3 gc = gnum * B.n;
  gdh = gdh + gden * B.n;
  g0 = gc;
  go = go - gc;
  // n = a x b, a = p1 - p0, b = p2 - p0
  const D3 a = p1 - p0, b = p2 - p0;
  const D3 ga = cross(b, gn), gb = cross(gn, a);
  g1 = ga;
  g2 = gb;
  g0 = g0 - (ga + gb);
  // dh = d / s
  gd = (1.0 / B.s) * (gdh - dot(gdh, B.dh) * B.dh);
}

constexpr uint32_t VTAB = 1024;  // LDS vertex slots per block: 128 buckets of 8
constexpr uint32_t VTAB_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t ADJ_WAVES = 4;          // waves of a 256-thread block
constexpr uint32_t ADJ_STAGE = 64u * 3u;   // staged corners per wave and bounce: three per lane

__device__ __forceinline__ uint32_t vtab_bucket(uint32_t v) { return (v * 2654435761u) >> 25; }  // top 7 bits

// Orders the LDS accesses of one wave: what its lanes wrote before is visible to its lanes after.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Every wave takes 64 consecutive rays per iteration, so each wave runs the loop body with all its lanes (rays past
// n are idle lanes): the corner scatter below is a wave-wide operation.
template <int NB>
__global__ void __launch_bounds__(256) k_mirror_adjoint(RaysView rays, const uint32_t *__restrict__ prims,
                                                        const double *__restrict__ V, const uint32_t *__restrict__ F,
                                                        uint32_t nf, uint64_t n, MirrorGrads g) {
  __shared__ __align__(16) uint32_t keys[VTAB];
  __shared__ double acc[3 * VTAB];
  __shared__ uint32_t stage_v[ADJ_WAVES][ADJ_STAGE];
  __shared__ double stage_g[ADJ_WAVES][ADJ_STAGE][3];
  for (uint32_t s = threadIdx.x; s < VTAB; s += blockDim.x) {
    keys[s] = VTAB_EMPTY;
    acc[3 * s] = acc[3 * s + 1] = acc[3 * s + 2] = 0.0;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  // Adds the corner gradient to the vertex's LDS slot in its bucket of 8 (read as two 16-byte loads), claiming the
  // bucket's first free slot for a vertex not yet there; false when the bucket is full of other vertices. Slots are
  // only ever claimed, so a failed claim (another lane took the slot) is retried on the fuller bucket and the loop
  // ends within 8 claims.
  auto to_table = [&](uint32_t v, D3 gvec) -> bool {
    const uint32_t b0 = vtab_bucket(v) * 8u;
    for (int attempt = 0; attempt < 8; attempt++) {
      const uint4 ka = *reinterpret_cast<const uint4 *>(&keys[b0]);
      const uint4 kb = *reinterpret_cast<const uint4 *>(&keys[b0 + 4]);
      const uint32_t ks[8] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y, kb.z, kb.w};
      int found = -1, free_slot = -1;
#pragma unroll
      for (int j = 7; j >= 0; j--) {
        if (ks[j] == v) found = j;
        if (ks[j] == VTAB_EMPTY) free_slot = j;
      }
      if (found < 0) {
        if (free_slot < 0) return false;
        const uint32_t k = atomicCAS(&keys[b0 + free_slot], VTAB_EMPTY, v);
        if (k != VTAB_EMPTY && k != v) continue;
        found = free_slot;
      }
      const uint32_t slot = b0 + (uint32_t)found;
      unsafeAtomicAdd(&acc[3 * slot], gvec.x);
      unsafeAtomicAdd(&acc[3 * slot + 1], gvec.y);
      unsafeAtomicAdd(&acc[3 * slot + 2], gvec.z);
      return true;
    }
    return false;
  };
  // One bounce's three corner gradients of every active lane. Corners without an LDS slot are staged in the wave's
  // LDS area and added to global memory with three adjacent lanes carrying x, y and z of one vertex: one atomic
  // instruction then covers 21 vertices, each in one request, where a lane per vertex needs three instructions and
  // three requests per vertex (random FP64 atomics are bound by requests, not by adds).
  auto scatter3 = [&](bool act, const uint32_t *vs, const D3 *gs) {
    bool need[3];
#pragma unroll
    for (int c = 0; c < 3; c++) need[c] = act && !to_table(vs[c], gs[c]);
    uint32_t base = 0;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long m = __ballot(need[c]);
      if (need[c]) {
        const uint32_t slot = base + (uint32_t)__popcll(m & below);
        stage_v[wave][slot] = vs[c];
        stage_g[wave][slot][0] = gs[c].x;
        stage_g[wave][slot][1] = gs[c].y;
        stage_g[wave][slot][2] = gs[c].z;
      }
      base += (uint32_t)__popcll(m);
    }
    wave_sync();
    const uint32_t comp = lane % 3u, rec = lane / 3u;
    for (uint32_t r0 = 0; r0 < base; r0 += 21u) {
      const uint32_t r = r0 + rec;
      if (lane < 63u && r < base) unsafeAtomicAdd(&g.gv[3ull * stage_v[wave][r] + comp], stage_g[wave][r][comp]);
    }
    wave_sync();  // the area is refilled by the next bounce
  };
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += stride) {
    const uint64_t i = w0 + lane;
    const bool live = i < n;
    D3 o[NB], d[NB];
    uint32_t f[NB];
    int nb = 0;
    if (live) {
      o[0] = d3(rays.ox[i], rays.oy[i], rays.oz[i]);
      d[0] = d3(rays.dx[i], rays.dy[i], rays.dz[i]);
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      f[b] = live ? prims[(uint64_t)b * n + i] : INVALID_ID;
      if (nb == b && f[b] < nf) {  // the chain ends at the first miss (or any id that is not a face)
        nb = b + 1;
        if (b + 1 < NB) {
          const uint32_t *fc = F + 3ull * f[b];
          Bounce B;
          bounce_forward(o[b], d[b], vertex(V, fc[0]), vertex(V, fc[1]), vertex(V, fc[2]), B, o[b + 1], d[b + 1]);
        }
      }
    }
    D3 gx_next = d3(0, 0, 0), gd_next = d3(0, 0, 0);
#pragma unroll
    for (int b = NB - 1; b >= 0; b--) {
      const bool act = b < nb;
      uint32_t vs[3] = {0u, 0u, 0u};
      D3 gs[3];
      if (act) {
        const uint32_t *fc = F + 3ull * f[b];
        vs[0] = fc[0];
        vs[1] = fc[1];
        vs[2] = fc[2];
        const D3 p0 = vertex(V, vs[0]), p1 = vertex(V, vs[1]), p2 = vertex(V, vs[2]);
        Bounce B;
        D3 x, dn;
        bounce_forward(o[b], d[b], p0, p1, p2, B, x, dn);
        const uint64_t k = (uint64_t)b * n + i;
        const double gt = g.gt ? g.gt[k] : 0.0;
        const D3 gp = g.gpx ? d3(g.gpx[k], g.gpy[k], g.gpz[k]) : d3(0, 0, 0);
        D3 go, gd;
        bounce_backward(B, p0, p1, p2, gt, gp + gx_next, gd_next, go, gd, gs[0], gs[1], gs[2]);
        gx_next = go;
        gd_next = gd;
      }
      if (g.gv) scatter3(act, vs, gs);
    }
    if (live && g.gox) {
      g.gox[i] = gx_next.x;
      g.goy[i] = gx_next.y;
      g.goz[i] = gx_next.z;
    }
    if (live && g.gdx) {
      g.gdx[i] = gd_next.x;
      g.gdy[i] = gd_next.y;
      g.gdz[i] = gd_next.z;
    }
  }
  __syncthreads();
  // flush: three adjacent lanes add x, y and z of one slot's vertex (21 slots per wave and pass)
  if (g.gv) {
    const uint32_t comp = lane % 3u;
    for (uint32_t s0 = wave * 21u; s0 < VTAB; s0 += ADJ_WAVES * 21u) {
      const uint32_t s = s0 + lane / 3u;
      if (lane < 63u && s < VTAB) {
        const uint32_t v = keys[s];
        if (v != VTAB_EMPTY) unsafeAtomicAdd(&g.gv[3ull * v + comp], acc[3 * s + comp]);
      }
    }
  }
}

template <int NB>
hipError_t launch_adjoint_nb(const RaysView &rays, const uint32_t *prims, const double *V, const uint32_t *F,
                             uint32_t nf, uint64_t n, const MirrorGrads &g, uint32_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_mirror_adjoint<NB>, dim3(blocks), dim3(256), 0, stream, rays, prims, V, F, nf, n, g);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_mirror_step(const RaysView &rays, const HitsView &hits, const double *vertices,
                              const uint32_t *faces, uint64_t n, double *const pos[3], double *const dir[3],
                              hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_mirror_step, dim3((uint32_t)blocks), dim3(256), 0, stream, rays, hits, vertices, faces, n,
                     pos[0], pos[1], pos[2], dir[0], dir[1], dir[2]);
  return hipGetLastError();
}

hipError_t launch_mirror_adjoint(const RaysView &rays, const uint32_t *prims, uint32_t bounces,
                                 const double *vertices, const uint32_t *faces, uint32_t num_faces, uint64_t n,
                                 const MirrorGrads &g, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  // a few resident blocks per CU, each looping over many rays so that its LDS table aggregates
  uint64_t blocks = (uint64_t)(cus > 0 ? cus : 1) * 8ull;
  const uint64_t need = (n + 255) / 256;
  if (blocks > need) blocks = need;
  switch (bounces) {
  case 1: return launch_adjoint_nb<1>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 2: return launch_adjoint_nb<2>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 3: return launch_adjoint_nb<3>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 4: return launch_adjoint_nb<4>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 5: return launch_adjoint_nb<5>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 6: return launch_adjoint_nb<6>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 7: return launch_adjoint_nb<7>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  case 8: return launch_adjoint_nb<8>(rays, prims, vertices, faces, num_faces, n, g, (uint32_t)blocks, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace brt
