// Device build of one geometry type's BVH (commit() on the GPU).
//
// Produces the host builder's tree (scene.cpp, itself blazert/bvh/builder.h:46-226 with binbuffer.h:15-167 and
// bbox.h:27-39), the same primitive order and the same flattened Node64 / Node4 / primitive-record arrays, bit for
// bit, whenever no primitive box coordinate is NaN or a negative zero. Under that condition every min / max the
// builder takes has one result whatever the evaluation order, so node bounds and bins are reduced with atomics.
// The order-dependent steps are reproduced exactly:
//   * the SAH sweeps of find_best_split_binned (binbuffer.h:115-166) become prefix / suffix scans over the bins,
//     one lane per bin, with the same products and sums; the first minimum wins as in the sequential loop;
//   * the partition is libstdc++'s std::partition on random-access iterators (bidirectional algorithm). It swaps
//     the i-th element of the left part [begin, begin + T) that fails the predicate with the i-th element of the
//     right part that passes it, counted from the right end, and leaves every other element in place (T = number
//     of passing elements). Ranks from prefix counts therefore give its exact result.
// Nodes with more than SMALL primitives are processed level by level, several blocks per node; each subtree of at
// most SMALL primitives is then built depth first by one wave out of LDS. The host numbers the nodes in pre-order
// (subtree sizes are known by then), and prefix scans flatten the tree into the traversal records.
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "engine.h"
#include "f32_round.h"

namespace brt {

namespace {

typedef unsigned long long u64;

constexpr uint32_t SMALL = 256;          // subtrees of at most this many primitives are built by one wave
constexpr uint32_t CHUNK = 2048;         // primitives per block in the level passes
constexpr uint32_t BLOCK = 256;          // threads per block in the level passes
constexpr uint32_t STACK = 64;           // depth-first stack of the subtree waves (max_depth <= 62)
constexpr uint32_t SMALL_REF = 0x80000000u;  // host-side child tag: a small subtree (else a large node)

// Order-preserving unsigned key of a double (NaN excluded): min / max of keys = min / max of values.
__device__ __forceinline__ u64 okey(double x) {
  const u64 b = (u64)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_value(u64 k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

// unsigned(std::round(x)) as the reference's x86-64 build evaluates it, clamped to the last bin
// (blazert/bvh/binbuffer.h:86-88)
__device__ __forceinline__ uint32_t bin_of(double x, uint32_t last) {
  const double r = round(x);
  uint32_t u = 0;
  if (r == r && r < 9.2e18 && r > -9.2e18) u = (uint32_t)(uint64_t)(int64_t)r;
  return u < last ? u : last;
}

// blazert/bvh/bbox.h:27-39
__device__ __forceinline__ double surface(const double *lo, const double *hi) {
  const double b0 = fabs(hi[0] - lo[0]), b1 = fabs(hi[1] - lo[1]), b2 = fabs(hi[2] - lo[2]);
  return 2.0 * (b0 * b1 + b1 * b2 + b2 * b0);
}

// Bins of one node: bin t = axis * nb + b holds a count and the min / max keys of the boxes binned there
// (3 keys each, at 3t + k).
struct Bins {
  uint32_t *cnt;
  u64 *klo, *khi;
};

// SAH sweep of one node over its bins (find_best_split_binned, binbuffer.h:115-166): all 64 lanes of a wave call
// it with the same arguments (nb <= 64); returns the split axis and writes the three cut positions.
__device__ uint32_t sah_sweep(const uint32_t *cnt, const u64 *klo, const u64 *khi, uint32_t nb, const double *lo,
                              const double *hi, double *cut) {
  const int lane = threadIdx.x & 63;
  double min_cost[3];
  for (int j = 0; j < 3; j++) {
    const bool in = lane < (int)nb;
    const uint32_t t = j * nb + lane;
    const uint64_t c = in ? cnt[t] : 0;
    double bl[3], bh[3];
    for (int k = 0; k < 3; k++) {
      bl[k] = in ? okey_value(klo[3 * t + k]) : DBL_MAX;
      bh[k] = in ? okey_value(khi[3 * t + k]) : -DBL_MAX;
    }
    // right side: bins lane .. nb-1 (the right-to-left sweep)
    double sl[3] = {bl[0], bl[1], bl[2]}, sh[3] = {bh[0], bh[1], bh[2]};
    uint64_t sc = c;
    for (int off = 1; off < 64; off <<= 1) {
      const bool ok = lane + off < 64;
      const uint64_t oc = __shfl_down(sc, off);
      if (ok) sc += oc;
      for (int k = 0; k < 3; k++) {
        const double ol = __shfl_down(sl[k], off), oh = __shfl_down(sh[k], off);
        if (ok) {
          sl[k] = dmin(sl[k], ol);
          sh[k] = dmax(sh[k], oh);
        }
      }
    }
    const double rcost = (double)sc * surface(sl, sh);
    // left side: bins 0 .. lane (the left-to-right sweep)
    double pl[3] = {bl[0], bl[1], bl[2]}, ph[3] = {bh[0], bh[1], bh[2]};
    uint64_t pc = c;
    for (int off = 1; off < 64; off <<= 1) {
      const bool ok = lane >= off;
      const uint64_t oc = __shfl_up(pc, off);
      if (ok) pc += oc;
      for (int k = 0; k < 3; k++) {
        const double ol = __shfl_up(pl[k], off), oh = __shfl_up(ph[k], off);
        if (ok) {
          pl[k] = dmin(pl[k], ol);
          ph[k] = dmax(ph[k], oh);
        }
      }
    }
    const double rnext = __shfl_down(rcost, 1);
    const double cost = (double)pc * surface(pl, ph) + rnext;
    // the sequential loop keeps the first cost below its running minimum (which starts at DBL_MAX): the first
    // lane holding the smallest cost below DBL_MAX, NaN never taken
    const bool valid = lane + 1 < (int)nb && cost < DBL_MAX;
    double m = valid ? cost : __builtin_inf();
    for (int off = 32; off > 0; off >>= 1) m = dmin(m, __shfl_xor(m, off));
    uint32_t min_bin = 1;
    min_cost[j] = DBL_MAX;
    if (m < DBL_MAX) {
      const u64 mask = __ballot(valid && cost == m);
      min_bin = (uint32_t)__ffsll((long long)mask);  // lowest such lane + 1
      min_cost[j] = m;
    }
    cut[j] = (double)min_bin * ((hi[j] - lo[j]) / (double)nb) + lo[j];
  }
  uint32_t axis = 0;
  if (min_cost[0] > min_cost[1]) axis = 1;
  if (min_cost[axis] > min_cost[2]) axis = 2;
  return axis;
}

// Exclusive rank of `flag` among the block's threads (thread order) and the block's total; every thread of the
// block must call it.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t *wave_tot, uint32_t &total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const u64 m = __ballot(flag);
  const uint32_t in_wave = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t w = 0; w < nw; w++) {
    const uint32_t c = wave_tot[w];
    before += (w < wave) ? c : 0u;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return before + in_wave;
}

// ------------------------------------------------------------------------------------------------ inputs
// Triangle boxes and centers exactly as the host collection precompute (blazert/primitives/trimesh.h:78-130).
__global__ void k_tri_boxes(const double *__restrict__ V, const uint32_t *__restrict__ F, uint32_t n,
                            double *__restrict__ lo, double *__restrict__ hi, double *__restrict__ ctr) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double *v0 = V + 3ull * F[3ull * f], *v1 = V + 3ull * F[3ull * f + 1], *v2 = V + 3ull * F[3ull * f + 2];
  const double third = 1.0 / 3.0;
  for (int k = 0; k < 3; k++) {
    double l = v0[k], h = v0[k];
    l = (l < v1[k]) ? l : v1[k];
    h = (v1[k] < h) ? h : v1[k];
    l = (l < v2[k]) ? l : v2[k];
    h = (v2[k] < h) ? h : v2[k];
    lo[3ull * f + k] = l;
    hi[3ull * f + k] = h;
    ctr[3ull * f + k] = ((v0[k] + v1[k]) + v2[k]) * third;
  }
}

__global__ void k_iota(uint32_t *idx, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) idx[i] = i;
}

// ------------------------------------------------------------------------------------------------ large nodes
// One node of the current level (nodes with more than SMALL primitives).
struct LNode {
  uint32_t begin, n, depth, leaf;
  uint32_t axis, mid, bad, split;
  uint32_t blk0, nblk, misplaced, pad;
  uint32_t cnt3[4];
  u64 klo[3], khi[3];
  double lo[3], hi[3], cut[3];
};

struct Level {
  LNode *nodes;
  const uint32_t *blk_node, *blk_start;  // block -> level node, first position of the block
  uint32_t *blk_l, *blk_r, *off_l, *off_r;
  Bins bins;
};

__device__ __forceinline__ uint32_t block_stop(const LNode &nd, uint32_t start) {
  const uint32_t end = nd.begin + nd.n;
  return (end - start < CHUNK) ? end : start + CHUNK;
}

__global__ void k_lv_init(Level L, uint32_t nb) {
  LNode &nd = L.nodes[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  if (tid < 3) {
    nd.klo[tid] = okey(DBL_MAX);
    nd.khi[tid] = okey(-DBL_MAX);
    nd.cnt3[tid] = 0;
  }
  const size_t base = (size_t)blockIdx.x * 3 * nb;
  for (uint32_t t = tid; t < 3 * nb; t += blockDim.x) {
    L.bins.cnt[base + t] = 0;
    for (int k = 0; k < 3; k++) {
      L.bins.klo[3 * (base + t) + k] = okey(DBL_MAX);
      L.bins.khi[3 * (base + t) + k] = okey(-DBL_MAX);
    }
  }
}

// Node bounds (the reference's compute_bounding_box over the node's primitive boxes, bbox.h:27-39).
__global__ void k_lv_bounds(Level L, const uint32_t *__restrict__ idx, const double *__restrict__ lo,
                            const double *__restrict__ hi) {
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  LNode &nd = L.nodes[node];
  const uint32_t stop = block_stop(nd, start);
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (uint32_t pos = start + threadIdx.x; pos < stop; pos += blockDim.x) {
    const uint32_t p = idx[pos];
    for (int k = 0; k < 3; k++) {
      mn[k] = dmin(mn[k], lo[3ull * p + k]);
      mx[k] = dmax(mx[k], hi[3ull * p + k]);
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; k++) {
      mn[k] = dmin(mn[k], __shfl_xor(mn[k], off));
      mx[k] = dmax(mx[k], __shfl_xor(mx[k], off));
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; k++) {
      atomicMin(&nd.klo[k], okey(mn[k]));
      atomicMax(&nd.khi[k], okey(mx[k]));
    }
}

__global__ void k_lv_prep(Level L, uint32_t m, uint32_t min_leaf, uint32_t max_depth) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LNode &nd = L.nodes[i];
  for (int k = 0; k < 3; k++) {
    nd.lo[k] = okey_value(nd.klo[k]);
    nd.hi[k] = okey_value(nd.khi[k]);
  }
  nd.leaf = (nd.n <= min_leaf || nd.depth >= max_depth) ? 1u : 0u;
}

// Binning (sort_collection_into_bins, binbuffer.h:67-91): per-block bins in LDS, merged into the node's bins.
__global__ void k_lv_bin(Level L, uint32_t nb, const uint32_t *__restrict__ idx, const double *__restrict__ lo,
                         const double *__restrict__ hi, const double *__restrict__ ctr) {
  __shared__ uint32_t scnt[3 * 64];
  __shared__ u64 sklo[3 * 64 * 3], skhi[3 * 64 * 3];
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  const LNode &nd = L.nodes[node];
  if (nd.leaf) return;
  const uint32_t stop = block_stop(nd, start);
  for (uint32_t t = threadIdx.x; t < 3 * nb; t += blockDim.x) {
    scnt[t] = 0;
    for (int k = 0; k < 3; k++) {
      sklo[3 * t + k] = okey(DBL_MAX);
      skhi[3 * t + k] = okey(-DBL_MAX);
    }
  }
  __syncthreads();
  double inv[3];
  for (int k = 0; k < 3; k++) {
    const double size = nd.hi[k] - nd.lo[k];
    inv[k] = (size > 0.0) ? 1.0 / size : 0.0;
  }
  const double scale = (double)(nb - 1);
  for (uint32_t pos = start + threadIdx.x; pos < stop; pos += blockDim.x) {
    const uint32_t p = idx[pos];
    u64 kl[3], kh[3];
    for (int k = 0; k < 3; k++) {
      kl[k] = okey(lo[3ull * p + k]);
      kh[k] = okey(hi[3ull * p + k]);
    }
    for (uint32_t j = 0; j < 3; j++) {
      const uint32_t t = j * nb + bin_of(((ctr[3ull * p + j] - nd.lo[j]) * inv[j]) * scale, nb - 1);
      atomicAdd(&scnt[t], 1u);
      for (int k = 0; k < 3; k++) {
        atomicMin(&sklo[3 * t + k], kl[k]);
        atomicMax(&skhi[3 * t + k], kh[k]);
      }
    }
  }
  __syncthreads();
  const size_t base = (size_t)node * 3 * nb;
  for (uint32_t t = threadIdx.x; t < 3 * nb; t += blockDim.x) {
    if (!scnt[t]) continue;
    atomicAdd(&L.bins.cnt[base + t], scnt[t]);
    for (int k = 0; k < 3; k++) {
      atomicMin(&L.bins.klo[3 * (base + t) + k], sklo[3 * t + k]);
      atomicMax(&L.bins.khi[3 * (base + t) + k], skhi[3 * t + k]);
    }
  }
}

// One wave per node.
__global__ void k_lv_sweep(Level L, uint32_t nb) {
  LNode &nd = L.nodes[blockIdx.x];
  if (nd.leaf) return;
  const size_t base = (size_t)blockIdx.x * 3 * nb;
  double cut[3];
  const uint32_t axis = sah_sweep(L.bins.cnt + base, L.bins.klo + 3 * base, L.bins.khi + 3 * base, nb, nd.lo, nd.hi,
                                  cut);
  if (threadIdx.x == 0) {
    nd.axis = axis;
    for (int k = 0; k < 3; k++) nd.cut[k] = cut[k];
  }
}

// Elements left of each axis's cut (the partition predicate center[axis] < cut[axis], builder.h:213-215).
__global__ void k_lv_count(Level L, const uint32_t *__restrict__ idx, const double *__restrict__ ctr) {
  __shared__ uint32_t sc[3];
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  LNode &nd = L.nodes[node];
  if (nd.leaf) return;
  const uint32_t stop = block_stop(nd, start);
  if (threadIdx.x < 3) sc[threadIdx.x] = 0;
  __syncthreads();
  uint32_t c[3] = {0, 0, 0};
  for (uint32_t pos = start + threadIdx.x; pos < stop; pos += blockDim.x) {
    const uint32_t p = idx[pos];
    for (int j = 0; j < 3; j++) c[j] += (ctr[3ull * p + j] < nd.cut[j]) ? 1u : 0u;
  }
  for (int j = 0; j < 3; j++)
    if (c[j]) atomicAdd(&sc[j], c[j]);
  __syncthreads();
  if (threadIdx.x < 3 && sc[threadIdx.x]) atomicAdd(&nd.cnt3[threadIdx.x], sc[threadIdx.x]);
}

// The builder's three partition attempts (blazert/bvh/builder.h:210-223): a partition that leaves one side
// empty changes nothing, falls back to the middle and moves on to the next axis.
__global__ void k_lv_decide(Level L, uint32_t m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LNode &nd = L.nodes[i];
  if (nd.leaf) return;
  uint32_t axis = nd.axis, mid = nd.begin, bad = 0, split = 0;
  for (int a = 0; a < 3; a++) {
    const uint32_t T = nd.cnt3[axis];
    if (T == 0 || T == nd.n) {
      bad++;
      mid = nd.begin + (nd.n >> 1);
      axis = (axis + 1) % 3;
    } else {
      mid = nd.begin + T;
      split = 1;
      break;
    }
  }
  nd.axis = axis;
  nd.mid = mid;
  nd.bad = bad;
  nd.split = split;
}

__device__ __forceinline__ bool left_of_cut(const LNode &nd, const uint32_t *idx, const double *ctr, uint32_t pos) {
  return ctr[3ull * idx[pos] + nd.axis] < nd.cut[nd.axis];
}

// Per block: elements on the wrong side (left part failing, right part passing).
__global__ void k_lv_rank1(Level L, const uint32_t *__restrict__ idx, const double *__restrict__ ctr) {
  __shared__ uint32_t s[2];
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  const LNode &nd = L.nodes[node];
  if (nd.leaf || !nd.split) return;
  const uint32_t stop = block_stop(nd, start);
  if (threadIdx.x < 2) s[threadIdx.x] = 0;
  __syncthreads();
  uint32_t cl = 0, cr = 0;
  for (uint32_t pos = start + threadIdx.x; pos < stop; pos += blockDim.x) {
    const bool pass = left_of_cut(nd, idx, ctr, pos);
    cl += (pos < nd.mid && !pass) ? 1u : 0u;
    cr += (pos >= nd.mid && pass) ? 1u : 0u;
  }
  if (cl) atomicAdd(&s[0], cl);
  if (cr) atomicAdd(&s[1], cr);
  __syncthreads();
  if (threadIdx.x == 0) {
    L.blk_l[blockIdx.x] = s[0];
    L.blk_r[blockIdx.x] = s[1];
  }
}

// Per node: block offsets of the misplaced counts.
__global__ void k_lv_rank2(Level L, uint32_t m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LNode &nd = L.nodes[i];
  if (nd.leaf || !nd.split) return;
  uint32_t sl = 0, sr = 0;
  for (uint32_t b = nd.blk0; b < nd.blk0 + nd.nblk; b++) {
    L.off_l[b] = sl;
    L.off_r[b] = sr;
    sl += L.blk_l[b];
    sr += L.blk_r[b];
  }
  nd.misplaced = sr;
}

// Positions of the right part's passing elements, indexed by their rank from the right end.
__global__ void k_lv_rank3(Level L, const uint32_t *__restrict__ idx, const double *__restrict__ ctr,
                           uint32_t *__restrict__ pos_r) {
  __shared__ uint32_t wt[BLOCK / 64];
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  const LNode &nd = L.nodes[node];
  if (nd.leaf || !nd.split) return;
  const uint32_t stop = block_stop(nd, start);
  uint32_t run = L.off_r[blockIdx.x];
  for (uint32_t base = start; base < stop; base += blockDim.x) {
    const uint32_t pos = base + threadIdx.x;
    const bool f = pos < stop && pos >= nd.mid && left_of_cut(nd, idx, ctr, pos);
    uint32_t tot;
    const uint32_t r = block_rank(f, wt, tot);
    if (f) pos_r[nd.begin + (nd.misplaced - 1u - (run + r))] = pos;
    run += tot;
  }
}

// The swaps: the r-th failing element of the left part trades places with the r-th passing element of the right
// part counted from the right end. Only left-part positions are read here, each by the one thread that swaps it.
__global__ void k_lv_swap(Level L, uint32_t *__restrict__ idx, const double *__restrict__ ctr,
                          const uint32_t *__restrict__ pos_r) {
  __shared__ uint32_t wt[BLOCK / 64];
  const uint32_t node = L.blk_node[blockIdx.x], start = L.blk_start[blockIdx.x];
  const LNode &nd = L.nodes[node];
  if (nd.leaf || !nd.split) return;
  const uint32_t stop = block_stop(nd, start);
  uint32_t run = L.off_l[blockIdx.x];
  for (uint32_t base = start; base < stop; base += blockDim.x) {
    const uint32_t pos = base + threadIdx.x;
    const bool f = pos < stop && pos < nd.mid && !left_of_cut(nd, idx, ctr, pos);
    uint32_t tot;
    const uint32_t r = block_rank(f, wt, tot);
    if (f) {
      const uint32_t partner = pos_r[nd.begin + run + r];
      const uint32_t a = idx[pos];
      idx[pos] = idx[partner];
      idx[partner] = a;
    }
    run += tot;
  }
}

// ------------------------------------------------------------------------------------------------ small subtrees
struct SmallRoot {
  uint32_t begin, n, depth, pad;
};

struct BuildStatsDev {
  uint32_t max_depth, leaves, branches, bad, max_prims, min_prims;
};

// One wave builds the whole subtree of a small root depth first (build_recursive, blazert/bvh/builder.h:85-115), its
// primitives' boxes and centers staged in LDS and permuted through an LDS index array. Nodes are written in the
// subtree's own pre-order to tmp[2 * begin ...] (a subtree of n primitives has at most 2n - 1 nodes) with
// subtree-local child indices; sizes[s] receives the node count.
__global__ void __launch_bounds__(64) k_small(const SmallRoot *__restrict__ roots, uint32_t *__restrict__ idx,
                                              const double *__restrict__ lo, const double *__restrict__ hi,
                                              const double *__restrict__ ctr, uint32_t nb, uint32_t min_leaf,
                                              uint32_t max_depth, brt_tree_node *__restrict__ tmp,
                                              uint32_t *__restrict__ sizes,
                                              BuildStatsDev *__restrict__ stats) {
  __shared__ uint32_t e_idx[SMALL];
  __shared__ double e_lo[3][SMALL], e_hi[3][SMALL], e_ctr[3][SMALL];
  __shared__ uint16_t perm[SMALL];
  __shared__ uint16_t pos_r[SMALL / 2];
  __shared__ uint32_t bcnt[3 * 64];
  __shared__ u64 bklo[3 * 64 * 3], bkhi[3 * 64 * 3];
  __shared__ uint32_t stk[STACK][4];

  const SmallRoot R = roots[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  for (uint32_t k = lane; k < R.n; k += 64) {
    const uint32_t p = idx[R.begin + k];
    e_idx[k] = p;
    for (int c = 0; c < 3; c++) {
      e_lo[c][k] = lo[3ull * p + c];
      e_hi[c][k] = hi[3ull * p + c];
      e_ctr[c][k] = ctr[3ull * p + c];
    }
    perm[k] = (uint16_t)k;
  }
  brt_tree_node *out = tmp + 2ull * R.begin;
  BuildStatsDev st{0, 0, 0, 0, 0, 0xFFFFFFFFu};
  if (lane == 0) {
    stk[0][0] = 0;
    stk[0][1] = R.n;
    stk[0][2] = R.depth;
    stk[0][3] = INVALID_ID;
  }
  __syncthreads();
  uint32_t sp = 1, count = 0;
  while (sp > 0) {
    sp--;
    const uint32_t b = stk[sp][0], e = stk[sp][1], d = stk[sp][2], parent = stk[sp][3];
    __syncthreads();
    const uint32_t self = count++;
    const uint32_t n = e - b;
    // bounds
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t k = b + lane; k < e; k += 64) {
      const uint32_t q = perm[k];
      for (int c = 0; c < 3; c++) {
        mn[c] = dmin(mn[c], e_lo[c][q]);
        mx[c] = dmax(mx[c], e_hi[c][q]);
      }
    }
    for (int off = 32; off > 0; off >>= 1)
      for (int c = 0; c < 3; c++) {
        mn[c] = dmin(mn[c], __shfl_xor(mn[c], off));
        mx[c] = dmax(mx[c], __shfl_xor(mx[c], off));
      }
    st.max_depth = d > st.max_depth ? d : st.max_depth;
    brt_tree_node node;
    for (int c = 0; c < 3; c++) {
      node.min[c] = mn[c];
      node.max[c] = mx[c];
    }
    node.children[0] = node.children[1] = INVALID_ID;
    node.prim_begin = node.prim_count = 0;
    if (n <= min_leaf || d >= max_depth) {
      node.leaf = 1;
      node.axis = INVALID_ID;
      node.prim_begin = R.begin + b;
      node.prim_count = n;
      st.leaves++;
      st.max_prims = n > st.max_prims ? n : st.max_prims;
      st.min_prims = n < st.min_prims ? n : st.min_prims;
      if (lane == 0) {
        out[self] = node;
        if (parent != INVALID_ID) out[parent].children[1] = self;
      }
      continue;
    }
    // binning
    for (uint32_t t = lane; t < 3 * nb; t += 64) {
      bcnt[t] = 0;
      for (int k = 0; k < 3; k++) {
        bklo[3 * t + k] = okey(DBL_MAX);
        bkhi[3 * t + k] = okey(-DBL_MAX);
      }
    }
    __syncthreads();
    double inv[3];
    for (int k = 0; k < 3; k++) {
      const double size = mx[k] - mn[k];
      inv[k] = (size > 0.0) ? 1.0 / size : 0.0;
    }
    const double scale = (double)(nb - 1);
    for (uint32_t k = b + lane; k < e; k += 64) {
      const uint32_t q = perm[k];
      for (uint32_t j = 0; j < 3; j++) {
        const uint32_t t = j * nb + bin_of(((e_ctr[j][q] - mn[j]) * inv[j]) * scale, nb - 1);
        atomicAdd(&bcnt[t], 1u);
        for (int c = 0; c < 3; c++) {
          atomicMin(&bklo[3 * t + c], okey(e_lo[c][q]));
          atomicMax(&bkhi[3 * t + c], okey(e_hi[c][q]));
        }
      }
    }
    __syncthreads();
    double cut[3];
    uint32_t axis = sah_sweep(bcnt, bklo, bkhi, nb, mn, mx, cut);
    // partition attempts
    uint32_t mid = b;
    bool split = false;
    for (int a = 0; a < 3; a++) {
      uint32_t T = 0;
      for (uint32_t base = b; base < e; base += 64) {
        const uint32_t k = base + lane;
        T += (uint32_t)__popcll(__ballot(k < e && e_ctr[axis][perm[k]] < cut[axis]));
      }
      if (T == 0 || T == n) {
        st.bad++;
        mid = b + (n >> 1);
        axis = (axis + 1) % 3;
      } else {
        mid = b + T;
        split = true;
        break;
      }
    }
    if (split) {
      const double c = cut[axis];
      uint32_t run = 0;
      for (uint32_t base = mid; base < e; base += 64) {
        const uint32_t k = base + lane;
        const bool f = k < e && e_ctr[axis][perm[k]] < c;
        const u64 msk = __ballot(f);
        if (f) pos_r[run + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull))] = (uint16_t)k;
        run += (uint32_t)__popcll(msk);
      }
      __syncthreads();
      const uint32_t misplaced = run;
      run = 0;
      for (uint32_t base = b; base < mid; base += 64) {
        const uint32_t k = base + lane;
        const bool f = k < mid && !(e_ctr[axis][perm[k]] < c);
        const u64 msk = __ballot(f);
        if (f) {
          const uint32_t r = run + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull));
          const uint32_t partner = pos_r[misplaced - 1u - r];
          const uint16_t x = perm[k];
          perm[k] = perm[partner];
          perm[partner] = x;
        }
        run += (uint32_t)__popcll(msk);
      }
      __syncthreads();
    }
    node.leaf = 0;
    node.axis = axis;
    node.children[0] = self + 1;
    st.branches++;
    if (lane == 0) {
      out[self] = node;
      if (parent != INVALID_ID) out[parent].children[1] = self;
      stk[sp][0] = mid;
      stk[sp][1] = e;
      stk[sp][2] = d + 1;
      stk[sp][3] = self;
      stk[sp + 1][0] = b;
      stk[sp + 1][1] = mid;
      stk[sp + 1][2] = d + 1;
      stk[sp + 1][3] = INVALID_ID;
    }
    sp += 2;
    __syncthreads();
  }
  for (uint32_t k = lane; k < R.n; k += 64) idx[R.begin + k] = e_idx[perm[k]];
  if (lane == 0) {
    sizes[blockIdx.x] = count;
    atomicMax(&stats->max_depth, st.max_depth);
    atomicAdd(&stats->leaves, st.leaves);
    atomicAdd(&stats->branches, st.branches);
    atomicAdd(&stats->bad, st.bad);
    atomicMax(&stats->max_prims, st.max_prims);
    atomicMin(&stats->min_prims, st.min_prims);
  }
}

// ------------------------------------------------------------------------------------------------ assembly
struct TopPut {
  brt_tree_node node;
  uint32_t pre, pad;
};

__global__ void k_put_top(const TopPut *__restrict__ tops, uint32_t m, brt_tree_node *__restrict__ tree) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  tree[tops[i].pre] = tops[i].node;
}

// Moves each small subtree from its scratch slot to its pre-order position (base[s]), making child indices global.
__global__ void k_reloc(const SmallRoot *__restrict__ roots, const uint32_t *__restrict__ sizes,
                        const uint32_t *__restrict__ bases, const brt_tree_node *__restrict__ tmp,
                        brt_tree_node *__restrict__ tree) {
  const uint32_t s = blockIdx.x, base = bases[s];
  const size_t from = 2ull * roots[s].begin;
  for (uint32_t j = threadIdx.x; j < sizes[s]; j += blockDim.x) {
    brt_tree_node nd = tmp[from + j];
    if (!nd.leaf) {
      nd.children[0] += base;
      nd.children[1] += base;
    }
    tree[base + j] = nd;
  }
}

// ------------------------------------------------------------------------------------------------ flatten
// Every internal node becomes one Node64 (numbered in pre-order among internal nodes). The root and every internal
// slot of a record (collapse4) become one Node4 each, numbered as the host's depth-first allocation does: a Node4's
// internal slots get consecutive indices when it is processed, and Node4s are processed in pre-order, so slot r of
// node P gets 1 + (internal slots of all Node4s before P) + r.

// is_root[i] = 1 when internal node i roots a Node4. A node's status depends on the records above it, so each
// thread walks from the root down to its node: in pre-order the left subtree of v is [v + 1, right child), and on
// the way the current record's slots tell whether the next node starts a record or lies inside the current one.
__global__ void k_roots4(const brt_tree_node *__restrict__ T, uint32_t M, uint32_t *__restrict__ is_root) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  uint32_t root = 0;
  if (i > 0 && !T[i].leaf) {
    uint32_t slot[4];
    int m = collapse4(T, 0, slot);
    uint32_t v = 0;
    // every step descends one level towards i, which lies below v, so v never is a leaf here and the loop ends
    // within the tree's depth (at most 63); the bound and the leaf test only guard the reads
    for (int level = 0; level < 64 && !T[v].leaf; level++) {
      const uint32_t right = T[v].children[1];
      v = (i >= right) ? right : T[v].children[0];
      bool is_slot = false;
      for (int c = 0; c < m; c++) is_slot = is_slot || slot[c] == v;
      if (v == i) {
        root = is_slot ? 1u : 0u;
        break;
      }
      if (is_slot) m = collapse4(T, v, slot);
    }
  } else if (i == 0) {
    root = T[0].leaf ? 0u : 1u;
  }
  is_root[i] = root;
}

__global__ void k_flags(const brt_tree_node *__restrict__ T, const uint32_t *__restrict__ is_root, uint32_t M,
                        uint32_t *__restrict__ internal, uint32_t *__restrict__ islots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  internal[i] = T[i].leaf ? 0u : 1u;
  uint32_t s = 0;
  if (is_root[i]) {
    uint32_t slot[4];
    const int m = collapse4(T, i, slot);
    for (int c = 0; c < m; c++) s += T[slot[c]].leaf ? 0u : 1u;
  }
  islots[i] = s;
}

__global__ void k_assign4(const brt_tree_node *__restrict__ T, const uint32_t *__restrict__ is_root, uint32_t M,
                          const uint32_t *__restrict__ S, uint32_t *__restrict__ k4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M || !is_root[i]) return;
  if (i == 0) k4[0] = 0;
  uint32_t slot[4];
  const int m = collapse4(T, i, slot);
  uint32_t r = 0;
  for (int c = 0; c < m; c++)
    if (!T[slot[c]].leaf) k4[slot[c]] = 1u + S[i] + r++;
}

__global__ void k_node64(const brt_tree_node *__restrict__ T, uint32_t M, const uint32_t *__restrict__ ni,
                         Node64 *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M || T[i].leaf) return;
  Node64 a;
  for (int c = 0; c < 2; c++) {
    const uint32_t chi = T[i].children[c];
    const brt_tree_node &ch = T[chi];
    for (int k = 0; k < 3; k++) {
      a.b[6 * c + k] = ch.min[k];
      a.b[6 * c + 3 + k] = ch.max[k];
    }
    a.ref[c] = ch.leaf ? (LEAF_BIT | ch.prim_begin) : ni[chi];
    a.cnt[c] = ch.leaf ? ch.prim_count : 0u;
  }
  a.cnt[1] |= T[i].axis << AXIS_SHIFT;
  for (int k = 0; k < 4; k++) a.pad[k] = 0;
  out[ni[i]] = a;
}

__global__ void k_node4(const brt_tree_node *__restrict__ T, const uint32_t *__restrict__ is_root, uint32_t M,
                        const uint32_t *__restrict__ k4, Node4 *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M || !is_root[i]) return;
  uint32_t slot[4];
  const int m = collapse4(T, i, slot);
  Node4 nd;
  for (int c = 0; c < 4; c++) {
    for (int k = 0; k < 3; k++) {
      nd.lo[k][c] = __builtin_inff();
      nd.hi[k][c] = -__builtin_inff();
    }
    nd.ref[c] = 0;
    nd.cnt[c] = EMPTY_SLOT;
  }
  for (int c = 0; c < m; c++) {
    const brt_tree_node &sn = T[slot[c]];
    for (int k = 0; k < 3; k++) {
      nd.lo[k][c] = f32_down(sn.min[k]);
      nd.hi[k][c] = f32_up(sn.max[k]);
    }
    nd.ref[c] = sn.leaf ? (LEAF_BIT | sn.prim_begin) : k4[slot[c]];
    nd.cnt[c] = sn.leaf ? sn.prim_count : 0u;
  }
  out[k4[i]] = nd;
}

// Primitive records in leaf order (the host flatten's record loops).
__global__ void k_tri_recs(const uint32_t *__restrict__ order, const double *__restrict__ V,
                           const uint32_t *__restrict__ F, uint32_t n, TriRec *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  const double *v0 = V + 3ull * F[3ull * p], *v1 = V + 3ull * F[3ull * p + 1], *v2 = V + 3ull * F[3ull * p + 2];
  TriRec r;
  for (int k = 0; k < 3; k++) {
    r.v0[k] = v0[k];
    r.e1[k] = v0[k] - v1[k];
    r.e2[k] = v2[k] - v0[k];
  }
  r.prim = p;
  out[i] = r;
}

__global__ void k_sph_recs(const uint32_t *__restrict__ order, const double *__restrict__ c,
                           const double *__restrict__ rad, uint32_t n, SphRec *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  SphRec r;
  for (int k = 0; k < 3; k++) r.c[k] = c[3ull * p + k];
  r.r = rad[p];
  out[i] = r;
}

__global__ void k_pla_recs(const uint32_t *__restrict__ order, const double *__restrict__ c,
                           const double *__restrict__ dx, const double *__restrict__ dy,
                           const double *__restrict__ R, uint32_t n, PlaRec *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  PlaRec r;
  for (int k = 0; k < 3; k++) r.c[k] = c[3ull * p + k];
  r.dx = dx[p];
  r.dy = dy[p];
  for (int k = 0; k < 9; k++) r.R[k] = R[9ull * p + k];
  r.pad[0] = r.pad[1] = 0.0;
  out[i] = r;
}

__global__ void k_cyl_recs(const uint32_t *__restrict__ order, const double *__restrict__ c,
                           const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ h,
                           const double *__restrict__ R, uint32_t n, CylRec *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  CylRec r;
  for (int k = 0; k < 3; k++) r.c[k] = c[3ull * p + k];
  r.a = a[p];
  r.b = b[p];
  r.h = h[p];
  for (int k = 0; k < 9; k++) r.R[k] = R[9ull * p + k];
  r.pad = 0.0;
  out[i] = r;
}

// ------------------------------------------------------------------------------------------------ host driver
inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Device allocations of one build; temporaries are released by the destructor (after the build's stream has
// drained, also on an error path), results are handed over.
struct Arena {
  hipStream_t stream;
  std::vector<void *> ptrs;
  hipError_t err = hipSuccess;
  explicit Arena(hipStream_t s) : stream(s) {}
  template <class T>
  T *get(size_t count) {
    if (err != hipSuccess) return nullptr;
    void *p = nullptr;
    err = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (err != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
  void release(void *p) {
    for (auto &q : ptrs)
      if (q == p) q = nullptr;
  }
  ~Arena() {
    (void)hipStreamSynchronize(stream);
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
};

template <class T>
hipError_t to_device(Arena &A, const T *host, size_t count, const T **dst, hipStream_t st) {
  *dst = nullptr;
  if (!count) return hipSuccess;
  T *d = A.get<T>(count);
  if (!d) return A.err;
  *dst = d;
  return hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, st);
}

struct HostTop {
  uint32_t begin, n, depth, leaf, axis;
  uint32_t child[2];
  double lo[3], hi[3];
};

}  // namespace

// Each failing step records its source line in out.failed_line.
#define BRT_CHECK(x)                 \
  do {                               \
    const hipError_t e_ = (x);       \
    if (e_ != hipSuccess) {          \
      out.failed_line = __LINE__;    \
      return e_;                     \
    }                                \
  } while (0)
#define BRT_LAUNCH(...)           \
  do {                            \
    __VA_ARGS__;                  \
    BRT_CHECK(hipGetLastError()); \
  } while (0)

bool device_build_supported(uint32_t nbins, uint32_t max_depth) { return nbins >= 2 && nbins <= 64 && max_depth <= 62; }

hipError_t device_build(const DeviceBuildIn &in, DeviceBuildOut &out, hipStream_t st) {
  out = DeviceBuildOut{};
  const uint32_t n = in.n, nb = in.nbins;
  if (n == 0 || !device_build_supported(nb, in.max_depth)) return hipErrorInvalidValue;
  BRT_CHECK(hipGetLastError());  // an error left by earlier work is reported as such, not as this build's
  Arena A(st);
  // primitive boxes and centers
  double *lo = A.get<double>(3ull * n), *hi = A.get<double>(3ull * n), *ctr = A.get<double>(3ull * n);
  uint32_t *idx = A.get<uint32_t>(n), *pos_r = A.get<uint32_t>(n);
  BRT_CHECK(A.err);
  if (in.type == TRIANGLES) {
    BRT_LAUNCH(k_tri_boxes<<<blocks_for(n, 256), 256, 0, st>>>(in.vertices, in.faces, n, lo, hi, ctr));
  } else {
    BRT_CHECK(hipMemcpyAsync(lo, in.box_lo, 3ull * n * 8, hipMemcpyHostToDevice, st));
    BRT_CHECK(hipMemcpyAsync(hi, in.box_hi, 3ull * n * 8, hipMemcpyHostToDevice, st));
    BRT_CHECK(hipMemcpyAsync(ctr, in.box_ctr, 3ull * n * 8, hipMemcpyHostToDevice, st));
  }
  BRT_LAUNCH(k_iota<<<blocks_for(n, 256), 256, 0, st>>>(idx, n));

  // ---- large nodes, level by level
  const uint32_t cap_nodes = n / SMALL + 2;                 // large nodes of one level are disjoint, > SMALL each
  const uint32_t cap_blocks = n / CHUNK + cap_nodes + 1;
  LNode *d_nodes = A.get<LNode>(cap_nodes);
  uint32_t *d_map = A.get<uint32_t>(2ull * cap_blocks);
  uint32_t *d_cnt4 = A.get<uint32_t>(4ull * cap_blocks);
  uint32_t *bcnt = A.get<uint32_t>((size_t)cap_nodes * 3 * nb);
  u64 *bklo = A.get<u64>((size_t)cap_nodes * 9 * nb), *bkhi = A.get<u64>((size_t)cap_nodes * 9 * nb);
  BRT_CHECK(A.err);
  Level L{d_nodes, d_map, d_map + cap_blocks, d_cnt4, d_cnt4 + cap_blocks, d_cnt4 + 2 * cap_blocks,
          d_cnt4 + 3 * cap_blocks, Bins{bcnt, bklo, bkhi}};

  BuildStatsDev hs{0, 0, 0, 0, 0, 0xFFFFFFFFu};
  std::vector<HostTop> top;
  std::vector<SmallRoot> small;
  auto add_child = [&](uint32_t begin, uint32_t cnt, uint32_t depth, std::vector<uint32_t> &next) -> uint32_t {
    if (cnt > SMALL) {
      top.push_back(HostTop{begin, cnt, depth, 0, INVALID_ID, {INVALID_ID, INVALID_ID}, {0, 0, 0}, {0, 0, 0}});
      next.push_back((uint32_t)top.size() - 1);
      return (uint32_t)top.size() - 1;
    }
    small.push_back(SmallRoot{begin, cnt, depth, 0});
    return SMALL_REF | ((uint32_t)small.size() - 1);
  };
  std::vector<uint32_t> level;
  const uint32_t root_ref = add_child(0, n, 0, level);
  std::vector<LNode> hn;
  std::vector<uint32_t> hmap;
  while (!level.empty()) {
    const uint32_t m = (uint32_t)level.size();
    hn.assign(m, LNode{});
    hmap.clear();
    uint32_t nblocks = 0;
    for (uint32_t i = 0; i < m; i++) {
      const HostTop &t = top[level[i]];
      LNode &x = hn[i];
      x.begin = t.begin;
      x.n = t.n;
      x.depth = t.depth;
      x.blk0 = nblocks;
      x.nblk = blocks_for(t.n, CHUNK);
      nblocks += x.nblk;
    }
    hmap.resize(2ull * cap_blocks);
    for (uint32_t i = 0; i < m; i++)
      for (uint32_t b = 0; b < hn[i].nblk; b++) {
        hmap[hn[i].blk0 + b] = i;
        hmap[cap_blocks + hn[i].blk0 + b] = hn[i].begin + b * CHUNK;
      }
    BRT_CHECK(hipMemcpyAsync(d_nodes, hn.data(), m * sizeof(LNode), hipMemcpyHostToDevice, st));
    BRT_CHECK(hipMemcpyAsync(d_map, hmap.data(), 2ull * cap_blocks * 4, hipMemcpyHostToDevice, st));
    BRT_LAUNCH(k_lv_init<<<m, 256, 0, st>>>(L, nb));
    BRT_LAUNCH(k_lv_bounds<<<nblocks, BLOCK, 0, st>>>(L, idx, lo, hi));
    BRT_LAUNCH(k_lv_prep<<<blocks_for(m, 64), 64, 0, st>>>(L, m, in.min_leaf, in.max_depth));
    BRT_LAUNCH(k_lv_bin<<<nblocks, BLOCK, 0, st>>>(L, nb, idx, lo, hi, ctr));
    BRT_LAUNCH(k_lv_sweep<<<m, 64, 0, st>>>(L, nb));
    BRT_LAUNCH(k_lv_count<<<nblocks, BLOCK, 0, st>>>(L, idx, ctr));
    BRT_LAUNCH(k_lv_decide<<<blocks_for(m, 64), 64, 0, st>>>(L, m));
    BRT_LAUNCH(k_lv_rank1<<<nblocks, BLOCK, 0, st>>>(L, idx, ctr));
    BRT_LAUNCH(k_lv_rank2<<<blocks_for(m, 64), 64, 0, st>>>(L, m));
    BRT_LAUNCH(k_lv_rank3<<<nblocks, BLOCK, 0, st>>>(L, idx, ctr, pos_r));
    BRT_LAUNCH(k_lv_swap<<<nblocks, BLOCK, 0, st>>>(L, idx, ctr, pos_r));
    BRT_CHECK(hipMemcpyAsync(hn.data(), d_nodes, m * sizeof(LNode), hipMemcpyDeviceToHost, st));
    BRT_CHECK(hipStreamSynchronize(st));
    std::vector<uint32_t> next;
    for (uint32_t i = 0; i < m; i++) {
      const LNode &x = hn[i];
      const uint32_t ti = level[i];
      for (int k = 0; k < 3; k++) {
        top[ti].lo[k] = x.lo[k];
        top[ti].hi[k] = x.hi[k];
      }
      hs.max_depth = std::max(hs.max_depth, x.depth);
      if (x.leaf) {
        top[ti].leaf = 1;
        hs.leaves++;
        hs.max_prims = std::max(hs.max_prims, x.n);
        hs.min_prims = std::min(hs.min_prims, x.n);
        continue;
      }
      hs.branches++;
      hs.bad += x.bad;
      top[ti].axis = x.axis;
      const uint32_t c0 = add_child(x.begin, x.mid - x.begin, x.depth + 1, next);
      const uint32_t c1 = add_child(x.mid, x.begin + x.n - x.mid, x.depth + 1, next);
      top[ti].child[0] = c0;
      top[ti].child[1] = c1;
    }
    level.swap(next);
  }

  // ---- small subtrees
  const uint32_t ns = (uint32_t)small.size();
  brt_tree_node *tmp = A.get<brt_tree_node>(2ull * n);
  uint32_t *sizes = A.get<uint32_t>(ns), *bases = A.get<uint32_t>(ns);
  SmallRoot *d_small = A.get<SmallRoot>(ns);
  BuildStatsDev *d_stats = A.get<BuildStatsDev>(1);
  BRT_CHECK(A.err);
  BuildStatsDev init{0, 0, 0, 0, 0, 0xFFFFFFFFu};
  BRT_CHECK(hipMemcpyAsync(d_stats, &init, sizeof init, hipMemcpyHostToDevice, st));
  BRT_CHECK(hipMemcpyAsync(d_small, small.data(), ns * sizeof(SmallRoot), hipMemcpyHostToDevice, st));
  if (ns)
    BRT_LAUNCH(k_small<<<ns, 64, 0, st>>>(d_small, idx, lo, hi, ctr, nb, in.min_leaf, in.max_depth, tmp, sizes,
                                          d_stats));
  std::vector<uint32_t> hsizes(ns), hbases(ns);
  BuildStatsDev ds{};
  BRT_CHECK(hipMemcpyAsync(hsizes.data(), sizes, ns * 4, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipMemcpyAsync(&ds, d_stats, sizeof ds, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipStreamSynchronize(st));

  // ---- pre-order numbering of the large nodes and placement of the small subtrees
  std::vector<uint32_t> pre_top(top.size(), INVALID_ID);
  uint32_t M = 0;
  {
    std::vector<uint32_t> stack{root_ref};
    while (!stack.empty()) {
      const uint32_t r = stack.back();
      stack.pop_back();
      if (r & SMALL_REF) {
        const uint32_t s = r & ~SMALL_REF;
        hbases[s] = M;
        M += hsizes[s];
      } else {
        pre_top[r] = M++;
        if (!top[r].leaf) {
          stack.push_back(top[r].child[1]);
          stack.push_back(top[r].child[0]);
        }
      }
    }
  }
  auto final_index = [&](uint32_t r) { return (r & SMALL_REF) ? hbases[r & ~SMALL_REF] : pre_top[r]; };
  std::vector<TopPut> puts(top.size());
  for (size_t t = 0; t < top.size(); t++) {
    TopPut &p = puts[t];
    const HostTop &h = top[t];
    for (int k = 0; k < 3; k++) {
      p.node.min[k] = h.lo[k];
      p.node.max[k] = h.hi[k];
    }
    p.node.leaf = h.leaf;
    p.node.axis = h.leaf ? INVALID_ID : h.axis;
    p.node.children[0] = h.leaf ? INVALID_ID : final_index(h.child[0]);
    p.node.children[1] = h.leaf ? INVALID_ID : final_index(h.child[1]);
    p.node.prim_begin = h.leaf ? h.begin : 0u;
    p.node.prim_count = h.leaf ? h.n : 0u;
    p.pre = pre_top[t];
    p.pad = 0;
  }
  brt_tree_node *tree = A.get<brt_tree_node>(M);
  const TopPut *d_puts = nullptr;
  BRT_CHECK(A.err);
  BRT_CHECK(to_device(A, puts.data(), puts.size(), &d_puts, st));
  BRT_CHECK(hipMemcpyAsync(bases, hbases.data(), ns * 4, hipMemcpyHostToDevice, st));
  if (!puts.empty())
    BRT_LAUNCH(k_put_top<<<blocks_for(puts.size(), 64), 64, 0, st>>>(d_puts, (uint32_t)puts.size(), tree));
  if (ns) BRT_LAUNCH(k_reloc<<<ns, 256, 0, st>>>(d_small, sizes, bases, tmp, tree));

  // ---- flatten
  uint32_t *internal = A.get<uint32_t>(M), *islots = A.get<uint32_t>(M), *ni = A.get<uint32_t>(M),
           *S = A.get<uint32_t>(M), *k4 = A.get<uint32_t>(M), *is_root = A.get<uint32_t>(M);
  BRT_CHECK(A.err);
  BRT_LAUNCH(k_roots4<<<blocks_for(M, 256), 256, 0, st>>>(tree, M, is_root));
  BRT_LAUNCH(k_flags<<<blocks_for(M, 256), 256, 0, st>>>(tree, is_root, M, internal, islots));
  size_t tb1 = 0, tb2 = 0;
  BRT_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, internal, ni, (int)M, st));
  BRT_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, islots, S, (int)M, st));
  unsigned char *tstore = A.get<unsigned char>(std::max(tb1, tb2));
  BRT_CHECK(A.err);
  BRT_CHECK(hipcub::DeviceScan::ExclusiveSum(tstore, tb1, internal, ni, (int)M, st));
  BRT_CHECK(hipcub::DeviceScan::ExclusiveSum(tstore, tb2, islots, S, (int)M, st));
  uint32_t last[4] = {0, 0, 0, 0};
  BRT_CHECK(hipMemcpyAsync(&last[0], ni + M - 1, 4, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipMemcpyAsync(&last[1], internal + M - 1, 4, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipMemcpyAsync(&last[2], S + M - 1, 4, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipMemcpyAsync(&last[3], islots + M - 1, 4, hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipMemcpyAsync(&out.root, tree, sizeof(brt_tree_node), hipMemcpyDeviceToHost, st));
  BRT_CHECK(hipStreamSynchronize(st));
  const uint32_t n64 = last[0] + last[1];
  const uint32_t n4 = out.root.leaf ? 0u : 1u + last[2] + last[3];
  Node64 *nodes64 = A.get<Node64>(n64);
  Node4 *nodes4 = A.get<Node4>(n4);
  size_t rec_size = 0;
  switch (in.type) {
  case TRIANGLES: rec_size = sizeof(TriRec); break;
  case SPHERES: rec_size = sizeof(SphRec); break;
  case PLANES: rec_size = sizeof(PlaRec); break;
  default: rec_size = sizeof(CylRec); break;
  }
  unsigned char *recs = A.get<unsigned char>((size_t)n * rec_size);
  BRT_CHECK(A.err);
  BRT_LAUNCH(k_assign4<<<blocks_for(M, 256), 256, 0, st>>>(tree, is_root, M, S, k4));
  BRT_LAUNCH(k_node64<<<blocks_for(M, 256), 256, 0, st>>>(tree, M, ni, nodes64));
  BRT_LAUNCH(k_node4<<<blocks_for(M, 256), 256, 0, st>>>(tree, is_root, M, k4, nodes4));
  const uint32_t g = blocks_for(n, 256);
  switch (in.type) {
  case TRIANGLES:
    BRT_LAUNCH(k_tri_recs<<<g, 256, 0, st>>>(idx, in.vertices, in.faces, n, reinterpret_cast<TriRec *>(recs)));
    break;
  case SPHERES: {
    const double *c = nullptr, *r = nullptr;
    BRT_CHECK(to_device(A, in.centers, 3ull * n, &c, st));
    BRT_CHECK(to_device(A, in.radii, n, &r, st));
    BRT_LAUNCH(k_sph_recs<<<g, 256, 0, st>>>(idx, c, r, n, reinterpret_cast<SphRec *>(recs)));
    break;
  }
  case PLANES: {
    const double *c = nullptr, *dx = nullptr, *dy = nullptr, *R = nullptr;
    BRT_CHECK(to_device(A, in.centers, 3ull * n, &c, st));
    BRT_CHECK(to_device(A, in.dxs, n, &dx, st));
    BRT_CHECK(to_device(A, in.dys, n, &dy, st));
    BRT_CHECK(to_device(A, in.rot, 9ull * n, &R, st));
    BRT_LAUNCH(k_pla_recs<<<g, 256, 0, st>>>(idx, c, dx, dy, R, n, reinterpret_cast<PlaRec *>(recs)));
    break;
  }
  default: {
    const double *c = nullptr, *a = nullptr, *b = nullptr, *h = nullptr, *R = nullptr;
    BRT_CHECK(to_device(A, in.centers, 3ull * n, &c, st));
    BRT_CHECK(to_device(A, in.sa, n, &a, st));
    BRT_CHECK(to_device(A, in.sb, n, &b, st));
    BRT_CHECK(to_device(A, in.heights, n, &h, st));
    BRT_CHECK(to_device(A, in.rot, 9ull * n, &R, st));
    BRT_LAUNCH(k_cyl_recs<<<g, 256, 0, st>>>(idx, c, a, b, h, R, n, reinterpret_cast<CylRec *>(recs)));
    break;
  }
  }
  BRT_CHECK(hipStreamSynchronize(st));

  out.tree = tree;
  out.num_nodes = M;
  out.order = idx;
  out.n64 = nodes64;
  out.num_n64 = n64;
  out.n4 = nodes4;
  out.num_n4 = n4;
  out.recs = recs;
  out.rec_size = rec_size;
  out.max_depth = std::max(hs.max_depth, ds.max_depth);
  out.leaves = hs.leaves + ds.leaves;
  out.branches = hs.branches + ds.branches;
  out.bad_splits = hs.bad + ds.bad;
  out.max_prims = std::max(hs.max_prims, ds.max_prims);
  out.min_prims = std::min(hs.min_prims, ds.min_prims);
  for (void *p : {(void *)tree, (void *)idx, (void *)nodes64, (void *)nodes4, (void *)recs}) A.release(p);
  return hipSuccess;
}

}  // namespace brt
