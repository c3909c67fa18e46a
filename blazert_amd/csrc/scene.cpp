// Host side of the engine: scene construction, binned-SAH build, flattening, device upload and the C ABI.
//
// The builder reproduces the reference tree exactly (blazert/bvh/builder.h:46-226, binbuffer.h:15-167,
// bbox.h:27-39): same binning rounding, same sweep, same cut position, same std::partition (libstdc++) and
// the same bad-split fallback, numbered in pre-order. Identical topology is what makes the GPU's hit ids and
// tie order match the reference bit for bit. Independent subtrees are built on separate threads; each thread
// numbers its subtree locally and the results are spliced in pre-order, so the output does not depend on the
// thread count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/blazert_amd.h"
#include "engine.h"
#include "internal.h"
#include "layout.h"

namespace brt {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

namespace {

// ------------------------------------------------------------------------------------------------ vec helpers
// blaze conventions: dot = (a0 b0 + a1 b1) + a2 b2, normalize = v * (len != 0 ? 1/len : 0),
// vec / s = vec * (1/s), matrix-vector rows summed left to right.
// Every helper is generic over the scene type R (double, or float for Scene<float>), with the reference's
// operation order in either type.
template <class R>
struct V3T {
  R x[3];
  R &operator[](int i) { return x[i]; }
  R operator[](int i) const { return x[i]; }
};
using V3 = V3T<double>;
template <class T>
struct Id {
  using type = T;
};
template <class R>
inline V3T<R> sub(const V3T<R> &a, const V3T<R> &b) {
  return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}};
}
template <class R>
inline R dot(const V3T<R> &a, const V3T<R> &b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class R>
inline V3T<R> cross(const V3T<R> &a, const V3T<R> &b) {
  return {{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}
template <class R>
inline V3T<R> scale_by_inverse(const V3T<R> &a, typename Id<R>::type s) {
  const R is = R(1) / s;
  return {{a[0] * is, a[1] * is, a[2] * is}};
}
template <class R>
inline V3T<R> normalized(const V3T<R> &a) {
  const R len = std::sqrt(dot(a, a));
  const R il = (len != R(0)) ? R(1) / len : R(0);
  return {{a[0] * il, a[1] * il, a[2] * il}};
}
template <class R>
inline V3T<R> rot_apply(const R *M, const V3T<R> &v) {
  return {{M[0] * v[0] + M[1] * v[1] + M[2] * v[2], M[3] * v[0] + M[4] * v[1] + M[5] * v[2],
           M[6] * v[0] + M[7] * v[1] + M[8] * v[2]}};
}
template <class R>
inline R ref_min(R a, R b) { return (b < a) ? b : a; }  // std::min(a, b)
template <class R>
inline R ref_max(R a, R b) { return (a < b) ? b : a; }  // std::max(a, b)

// ------------------------------------------------------------------------------------------------ collections
// Per-primitive bounding boxes and centers in structure-of-arrays form (the collection precompute of
// trimesh.h:78-130, sphere.h:301-329, plane.h:89-150, cylinder.h:91-147).
template <class R>
struct BoxesT {
  std::vector<R> lo, hi, ctr;  // 3 per primitive
  uint32_t n = 0;
  // every box coordinate is finite and none is -0.0: then every min / max the builder takes has one result
  // whatever the evaluation order, which the device builder relies on
  bool device_buildable = false;
  // every box has min <= max on each axis (then so has every node box, a union of them)
  bool ordered = false;
  void resize(uint32_t m) {
    n = m;
    lo.assign(3ull * m, R(0));
    hi.assign(3ull * m, R(0));
    ctr.assign(3ull * m, R(0));
  }
  void classify() {
    auto plain = [](R x) { return std::isfinite(x) && !(x == R(0) && std::signbit(x)); };
    device_buildable = std::all_of(lo.begin(), lo.end(), plain) && std::all_of(hi.begin(), hi.end(), plain);
    ordered = true;
    for (size_t i = 0; i < lo.size(); i++) ordered = ordered && lo[i] <= hi[i];
  }
};

using Boxes = BoxesT<double>;

// Primitive records are built in a byte buffer and accessed as TriRec / SphRec / PlaRec / CylRec (or their FP32
// forms); the buffer is aligned like the most aligned of them (128 B), which std::vector<unsigned char> is not.
template <class T, size_t Align>
struct AlignedAllocator {
  using value_type = T;
  template <class U>
  struct rebind {
    using other = AlignedAllocator<U, Align>;
  };
  AlignedAllocator() = default;
  template <class U>
  AlignedAllocator(const AlignedAllocator<U, Align> &) {}
  T *allocate(size_t n) { return static_cast<T *>(::operator new(n * sizeof(T), std::align_val_t(Align))); }
  void deallocate(T *p, size_t) { ::operator delete(p, std::align_val_t(Align)); }
  template <class U>
  bool operator==(const AlignedAllocator<U, Align> &) const { return true; }
  template <class U>
  bool operator!=(const AlignedAllocator<U, Align> &) const { return false; }
};
using RecordBytes = std::vector<unsigned char, AlignedAllocator<unsigned char, 128>>;

template <class R>
struct GeometryT {
  bool present = false;
  uint32_t geom_id = INVALID_ID;
  BoxesT<R> boxes;
  // inputs (copied)
  std::vector<R> vertices;  // triangles
  std::vector<uint32_t> faces;
  std::vector<R> vnormals;
  std::vector<R> centers, radii, dxs, dys, rot, sa, sb, heights;
  // build output
  std::vector<brt_tree_node> tree;  // reference form, pre-order
  std::vector<uint32_t> order;      // primitive ids in leaf order
  brt_build_stats stats{};
  // flattened
  std::vector<Node64> n64;  // double scenes
  std::vector<Node4> n4;
  std::vector<Node32> n32;  // Scene<float>
  RecordBytes recs;
  size_t rec_size = 0;
  uint32_t num_n4 = 0;
  BvhView view{};
  double extent = 0.0;  // max finite |coordinate| of all boxes (split-precision error bound)
  bool extent_ok = true;
  // device build: the reference-form tree stays on the device (downloaded on export)
  bool built_on_device = false;
  const brt_tree_node *dtree = nullptr;
  uint32_t dtree_nodes = 0;
};
using Geometry = GeometryT<double>;
using GeometryF = GeometryT<float>;

template <class R>
void mesh_precompute(GeometryT<R> &g, uint32_t nv) {
  const uint32_t nf = (uint32_t)(g.faces.size() / 3);
  g.boxes.resize(nf);
  g.vnormals.assign(3ull * nv, R(0));
  auto vert = [&](uint32_t i) {
    return V3T<R>{{g.vertices[3ull * i], g.vertices[3ull * i + 1], g.vertices[3ull * i + 2]}};
  };
  for (uint32_t f = 0; f < nf; f++) {
    const uint32_t i0 = g.faces[3ull * f], i1 = g.faces[3ull * f + 1], i2 = g.faces[3ull * f + 2];
    const V3T<R> v0 = vert(i0), v1 = vert(i1), v2 = vert(i2);
    const V3T<R> sum{{(v0[0] + v1[0]) + v2[0], (v0[1] + v1[1]) + v2[1], (v0[2] + v1[2]) + v2[2]}};
    const V3T<R> c = scale_by_inverse(sum, R(3));
    V3T<R> lo = v0, hi = v0;
    for (const V3T<R> *p : {&v1, &v2})
      for (int k = 0; k < 3; k++) {
        lo[k] = ref_min((*p)[k], lo[k]);
        hi[k] = ref_max((*p)[k], hi[k]);
      }
    for (int k = 0; k < 3; k++) {
      g.boxes.lo[3ull * f + k] = lo[k];
      g.boxes.hi[3ull * f + k] = hi[k];
      g.boxes.ctr[3ull * f + k] = c[k];
    }
    // face normal normalize(cross(v1 - v0, v2 - v0)); each vertex accumulates fn / 3 in face order
    const V3T<R> fn3 = scale_by_inverse(normalized(cross(sub(v1, v0), sub(v2, v0))), R(3));
    for (uint32_t vi : {i0, i1, i2})
      for (int k = 0; k < 3; k++) g.vnormals[3ull * vi + k] = g.vnormals[3ull * vi + k] + fn3[k];
  }
}

template <class R>
void sphere_precompute(GeometryT<R> &g) {
  const uint32_t n = (uint32_t)g.radii.size();
  g.boxes.resize(n);
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      const R c = g.centers[3ull * i + k];
      g.boxes.lo[3ull * i + k] = c - g.radii[i];
      g.boxes.hi[3ull * i + k] = c + g.radii[i];
      g.boxes.ctr[3ull * i + k] = c;
    }
}

// Box of six points: first element wins ties as in std::min / std::max over an initializer list.
template <class R>
void six_point_box(const V3T<R> *pts, R *lo, R *hi) {
  for (int k = 0; k < 3; k++) {
    R mn = pts[0][k], mx = pts[0][k];
    for (int p = 1; p < 6; p++) {
      if (pts[p][k] < mn) mn = pts[p][k];
      if (mx < pts[p][k]) mx = pts[p][k];
    }
    lo[k] = mn;
    hi[k] = mx;
  }
}

template <class R>
void plane_precompute(GeometryT<R> &g) {
  const uint32_t n = (uint32_t)g.dxs.size();
  g.boxes.resize(n);
  const R th = std::numeric_limits<R>::min();  // plane.h:29
  for (uint32_t i = 0; i < n; i++) {
    const R *M = &g.rot[9ull * i];
    const V3T<R> c{{g.centers[3ull * i], g.centers[3ull * i + 1], g.centers[3ull * i + 2]}};
    const R dx = g.dxs[i], dy = g.dys[i];
    const V3T<R> loc[6] = {{{-dx / 2, -dy / 2, R(0)}}, {{dx / 2, -dy / 2, R(0)}}, {{-dx / 2, dy / 2, R(0)}},
                           {{dx / 2, dy / 2, R(0)}},   {{R(0), R(0), -th}},      {{R(0), R(0), th}}};
    V3T<R> pts[6];
    for (int p = 0; p < 6; p++) {
      const V3T<R> r = rot_apply(M, loc[p]);
      pts[p] = {{c[0] + r[0], c[1] + r[1], c[2] + r[2]}};
    }
    six_point_box(pts, &g.boxes.lo[3ull * i], &g.boxes.hi[3ull * i]);
    for (int k = 0; k < 3; k++) g.boxes.ctr[3ull * i + k] = c[k];
  }
}

template <class R>
void cylinder_precompute(GeometryT<R> &g) {
  const uint32_t n = (uint32_t)g.sa.size();
  g.boxes.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    const R *M = &g.rot[9ull * i];
    const V3T<R> c{{g.centers[3ull * i], g.centers[3ull * i + 1], g.centers[3ull * i + 2]}};
    const V3T<R> ra = rot_apply(M, V3T<R>{{g.sa[i], R(0), R(0)}});
    const V3T<R> rb = rot_apply(M, V3T<R>{{R(0), g.sb[i], R(0)}});
    const V3T<R> rh = rot_apply(M, V3T<R>{{R(0), R(0), g.heights[i] / 2}});
    V3T<R> pts[6];
    for (int k = 0; k < 3; k++) {
      pts[0][k] = c[k] + ra[k];  // a1
      pts[1][k] = c[k] + rb[k];  // b1
      pts[2][k] = c[k] - ra[k];  // a2
      pts[3][k] = c[k] - rb[k];  // b2
      pts[4][k] = c[k] + rh[k];  // h1
      pts[5][k] = c[k] - rh[k];  // h2
    }
    six_point_box(pts, &g.boxes.lo[3ull * i], &g.boxes.hi[3ull * i]);
    for (int k = 0; k < 3; k++) g.boxes.ctr[3ull * i + k] = c[k];
  }
}

// ------------------------------------------------------------------------------------------------ builder
struct BuildCounters {
  std::atomic<uint32_t> max_depth{0}, leaves{0}, branches{0}, bad_splits{0}, max_prims{0};
  std::atomic<uint32_t> min_prims{0xFFFFFFFFu};
  void atomic_max(std::atomic<uint32_t> &a, uint32_t v) {
    uint32_t cur = a.load();
    while (cur < v && !a.compare_exchange_weak(cur, v)) {
    }
  }
  void atomic_min(std::atomic<uint32_t> &a, uint32_t v) {
    uint32_t cur = a.load();
    while (v < cur && !a.compare_exchange_weak(cur, v)) {
    }
  }
};

template <class R>
struct BuilderT {
  const BoxesT<R> &B;
  uint32_t *idx;
  uint32_t min_leaf, max_depth, nbins;
  BuildCounters cnt;

  struct BinAcc {
    R lo[3], hi[3];
    uint32_t count;
    R cost;
  };

  void bounds(uint32_t begin, uint32_t end, R *lo, R *hi) const {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::numeric_limits<R>::max();
      hi[k] = -std::numeric_limits<R>::max();
    }
    for (uint32_t i = begin; i < end; i++) {
      const uint32_t p = idx[i];
      for (int k = 0; k < 3; k++) {
        lo[k] = ref_min(B.lo[3ull * p + k], lo[k]);
        hi[k] = ref_max(B.hi[3ull * p + k], hi[k]);
      }
    }
  }

  static R surface(const R *lo, const R *hi) {
    const R b0 = std::fabs(hi[0] - lo[0]), b1 = std::fabs(hi[1] - lo[1]), b2 = std::fabs(hi[2] - lo[2]);
    return R(2) * (b0 * b1 + b1 * b2 + b2 * b0);
  }

  // unsigned(std::round(x)) as the reference's x86-64 build evaluates it (64-bit truncation, low 32 bits)
  static uint32_t bin_of(R x, uint32_t last) {
    const R r = std::round(x);
    uint32_t u = 0;
    if (r == r && (double)r < 9.2e18 && (double)r > -9.2e18) u = (uint32_t)(uint64_t)(int64_t)r;
    return u < last ? u : last;
  }

  // find_best_split_binned: returns the axis and writes the three cut positions
  uint32_t best_split(uint32_t begin, uint32_t end, const R *lo, const R *hi, R *cut,
                      std::vector<BinAcc> &bins) const {
    const R MX = std::numeric_limits<R>::max();
    bins.assign(3ull * nbins, BinAcc{{MX, MX, MX}, {-MX, -MX, -MX}, 0u, R(0)});
    R inv[3];
    for (int k = 0; k < 3; k++) {
      const R size = hi[k] - lo[k];
      inv[k] = (size > R(0)) ? R(1) / size : R(0);
    }
    const R scale = (R)(nbins - 1);
    for (uint32_t i = begin; i < end; i++) {
      const uint32_t p = idx[i];
      for (uint32_t j = 0; j < 3; j++) {
        const uint32_t b = bin_of(((B.ctr[3ull * p + j] - lo[j]) * inv[j]) * scale, nbins - 1);
        BinAcc &bin = bins[j * nbins + b];
        bin.count++;
        for (int k = 0; k < 3; k++) {
          bin.lo[k] = ref_min(B.lo[3ull * p + k], bin.lo[k]);
          bin.hi[k] = ref_max(B.hi[3ull * p + k], bin.hi[k]);
        }
      }
    }
    R min_cost[3] = {MX, MX, MX};
    for (uint32_t j = 0; j < 3; j++) {
      BinAcc *ax = &bins[j * nbins];
      R alo[3] = {MX, MX, MX}, ahi[3] = {-MX, -MX, -MX};
      uint64_t count = 0;
      for (uint32_t i = nbins - 1; i > 0; i--) {  // right-to-left sweep: cost of the right side
        for (int k = 0; k < 3; k++) {
          alo[k] = ref_min(ax[i].lo[k], alo[k]);
          ahi[k] = ref_max(ax[i].hi[k], ahi[k]);
        }
        count += ax[i].count;
        ax[i].cost = (R)count * surface(alo, ahi);
      }
      for (int k = 0; k < 3; k++) {
        alo[k] = MX;
        ahi[k] = -MX;
      }
      count = 0;
      uint32_t min_bin = 1;
      for (uint32_t i = 0; i + 1 < nbins; i++) {  // left-to-right sweep: full cost, first minimum wins
        for (int k = 0; k < 3; k++) {
          alo[k] = ref_min(ax[i].lo[k], alo[k]);
          ahi[k] = ref_max(ax[i].hi[k], ahi[k]);
        }
        count += ax[i].count;
        const R cost = (R)count * surface(alo, ahi) + ax[i + 1].cost;
        if (cost < min_cost[j]) {
          min_cost[j] = cost;
          min_bin = i + 1;
        }
      }
      cut[j] = (R)min_bin * ((hi[j] - lo[j]) / (R)nbins) + lo[j];
    }
    uint32_t axis = 0;
    if (min_cost[0] > min_cost[1]) axis = 1;
    if (min_cost[axis] > min_cost[2]) axis = 2;
    return axis;
  }

  // Builds [begin, end) at `depth` into `out` (pre-order, indices local to `out` starting at out.size()).
  void build(std::vector<brt_tree_node> &out, uint32_t begin, uint32_t end, uint32_t depth, int par_levels) {
    std::vector<BinAcc> bins;
    build_rec(out, begin, end, depth, par_levels, bins);
  }

  uint32_t build_rec(std::vector<brt_tree_node> &out, uint32_t begin, uint32_t end, uint32_t depth, int par_levels,
                     std::vector<BinAcc> &bins) {
    const uint32_t self = (uint32_t)out.size();
    out.emplace_back();
    cnt.atomic_max(cnt.max_depth, depth);
    const uint32_t n = end - begin;
    R lo[3], hi[3];
    bounds(begin, end, lo, hi);
    brt_tree_node node{};
    for (int k = 0; k < 3; k++) {  // a float box widens exactly
      node.min[k] = lo[k];
      node.max[k] = hi[k];
    }
    node.children[0] = node.children[1] = INVALID_ID;
    if (n <= min_leaf || depth >= max_depth) {
      node.leaf = 1;
      node.axis = INVALID_ID;
      node.prim_begin = begin;
      node.prim_count = n;
      out[self] = node;
      cnt.leaves++;
      cnt.atomic_max(cnt.max_prims, n);
      cnt.atomic_min(cnt.min_prims, n);
      return self;
    }
    R cut[3];
    uint32_t axis = best_split(begin, end, lo, hi, cut, bins);
    uint32_t mid = begin;
    for (int attempt = 0; attempt < 3; attempt++) {
      const R c = cut[axis];
      const R *ctr = B.ctr.data();
      const uint32_t ax = axis;
      uint32_t *m = std::partition(idx + begin, idx + end,
                                   [ctr, ax, c](uint32_t p) { return ctr[3ull * p + ax] < c; });
      mid = (uint32_t)(m - idx);
      if (mid == begin || mid == end) {
        cnt.bad_splits++;
        mid = begin + (n >> 1);
        axis = (axis + 1) % 3;
      } else {
        break;
      }
    }
    node.leaf = 0;
    node.axis = axis;
    out[self] = node;
    cnt.branches++;
    if (par_levels > 0 && n >= 32768) {
      // left subtree on another thread into its own vector, right subtree here; splice in pre-order
      std::vector<brt_tree_node> left;
      left.reserve(2ull * (mid - begin) / 3 + 16);
      std::thread th([&] { build(left, begin, mid, depth + 1, par_levels - 1); });
      std::vector<brt_tree_node> right;
      right.reserve(2ull * (end - mid) / 3 + 16);
      build(right, mid, end, depth + 1, par_levels - 1);
      th.join();
      const uint32_t lbase = (uint32_t)out.size();
      const uint32_t rbase = lbase + (uint32_t)left.size();
      for (auto &x : left)
        for (auto &ch : x.children)
          if (ch != INVALID_ID) ch += lbase;
      for (auto &x : right)
        for (auto &ch : x.children)
          if (ch != INVALID_ID) ch += rbase;
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
      out[self].children[0] = lbase;
      out[self].children[1] = rbase;
    } else {
      const uint32_t l = build_rec(out, begin, mid, depth + 1, 0, bins);
      out[self].children[0] = l;
      const uint32_t r = build_rec(out, mid, end, depth + 1, 0, bins);
      out[self].children[1] = r;
    }
    return self;
  }
};

float f32_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -INFINITY);
  return f;
}
float f32_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafter(f, INFINITY);
  return f;
}

// Flattens the reference-form tree into children-in-parent records and leaf-order primitive records.
void flatten(Geometry &g, GeomType type) {
  const auto &T = g.tree;
  std::vector<uint32_t> internal(T.size(), INVALID_ID);
  uint32_t ni = 0;
  for (size_t i = 0; i < T.size(); i++)
    if (!T[i].leaf) internal[i] = ni++;
  g.n64.assign(ni, Node64{});
  double extent = 0.0;
  bool ok = true, irregular = false;
  auto note = [&](double v) {
    if (v != v) irregular = true;
    if (!std::isfinite(v)) ok = false;
    else extent = std::max(extent, std::fabs(v));
  };
  auto child_ref = [&](uint32_t c, uint32_t &ref, uint32_t &cnt) {
    if (T[c].leaf) {
      ref = LEAF_BIT | T[c].prim_begin;
      cnt = T[c].prim_count;
    } else {
      ref = internal[c];
      cnt = 0;
    }
  };
  for (size_t i = 0; i < T.size(); i++) {
    if (T[i].leaf) continue;
    Node64 &a = g.n64[internal[i]];
    for (int c = 0; c < 2; c++) {
      const brt_tree_node &ch = T[T[i].children[c]];
      for (int k = 0; k < 3; k++) {
        a.b[6 * c + k] = ch.min[k];
        a.b[6 * c + 3 + k] = ch.max[k];
        note(ch.min[k]);
        note(ch.max[k]);
        if (ch.min[k] > ch.max[k]) irregular = true;
      }
      child_ref(T[i].children[c], a.ref[c], a.cnt[c]);
    }
    a.cnt[1] |= T[i].axis << AXIS_SHIFT;
  }
  // 4-wide FP32 records: every Node4 stands for one binary internal node, with up to four slots taken from its
  // subtree by collapse4 (largest internal slot opened first). Depth-first allocation keeps subtrees together.
  g.n4.clear();
  uint32_t root4 = INVALID_ID;
  if (!T.empty() && !T[0].leaf) {
    std::vector<std::pair<uint32_t, uint32_t>> work;  // (binary internal node, its Node4 index)
    g.n4.push_back(Node4{});
    root4 = 0;
    work.emplace_back(0u, 0u);
    while (!work.empty()) {
      const auto [bi, k4] = work.back();
      work.pop_back();
      uint32_t slot[4];
      const int m = collapse4(T.data(), bi, slot);
      Node4 nd{};
      for (int c = 0; c < 4; c++) {
        for (int k = 0; k < 3; k++) {
          nd.lo[k][c] = __builtin_inff();
          nd.hi[k][c] = -__builtin_inff();
        }
        nd.ref[c] = 0;
        nd.cnt[c] = EMPTY_SLOT;
      }
      std::vector<std::pair<uint32_t, uint32_t>> children;
      for (int c = 0; c < m; c++) {
        const brt_tree_node &sn = T[slot[c]];
        for (int k = 0; k < 3; k++) {
          nd.lo[k][c] = f32_down(sn.min[k]);
          nd.hi[k][c] = f32_up(sn.max[k]);
        }
        if (sn.leaf) {
          nd.ref[c] = LEAF_BIT | sn.prim_begin;
          nd.cnt[c] = sn.prim_count;
        } else {
          nd.ref[c] = (uint32_t)g.n4.size();
          nd.cnt[c] = 0;
          g.n4.push_back(Node4{});
          children.emplace_back(slot[c], nd.ref[c]);
        }
      }
      g.n4[k4] = nd;
      for (auto it = children.rbegin(); it != children.rend(); ++it) work.push_back(*it);
    }
  }
  g.num_n4 = (uint32_t)g.n4.size();
  BvhView &v = g.view;
  v = BvhView{};
  for (int k = 0; k < 3; k++) {
    v.root_min[k] = T[0].min[k];
    v.root_max[k] = T[0].max[k];
    v.root_min32[k] = f32_down(T[0].min[k]);
    v.root_max32[k] = f32_up(T[0].max[k]);
    note(T[0].min[k]);
    note(T[0].max[k]);
    if (T[0].min[k] > T[0].max[k]) irregular = true;  // a root leaf's box is not among the children checked above
  }
  child_ref(0, v.root_ref, v.root_cnt);
  v.root_ref4 = (root4 == INVALID_ID) ? v.root_ref : root4;
  v.geom_id = g.geom_id;
  v.n_prims = g.boxes.n;
  v.n_nodes = ni;
  v.depth = g.stats.max_tree_depth;
  v.irregular_boxes = irregular ? 1u : 0u;
  g.extent = extent;
  g.extent_ok = ok && extent < 1e30;

  const uint32_t n = g.boxes.n;
  switch (type) {
  case TRIANGLES: {
    g.rec_size = sizeof(TriRec);
    g.recs.assign((size_t)n * sizeof(TriRec), 0);
    TriRec *r = reinterpret_cast<TriRec *>(g.recs.data());
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t p = g.order[i];
      const double *v0 = &g.vertices[3ull * g.faces[3ull * p]];
      const double *v1 = &g.vertices[3ull * g.faces[3ull * p + 1]];
      const double *v2 = &g.vertices[3ull * g.faces[3ull * p + 2]];
      for (int k = 0; k < 3; k++) {
        r[i].v0[k] = v0[k];
        r[i].e1[k] = v0[k] - v1[k];  // Triangle::c = a - b
        r[i].e2[k] = v2[k] - v0[k];  // Triangle::b = c - a
      }
      r[i].prim = p;
    }
    break;
  }
  case SPHERES: {
    g.rec_size = sizeof(SphRec);
    g.recs.assign((size_t)n * sizeof(SphRec), 0);
    SphRec *r = reinterpret_cast<SphRec *>(g.recs.data());
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t p = g.order[i];
      for (int k = 0; k < 3; k++) r[i].c[k] = g.centers[3ull * p + k];
      r[i].r = g.radii[p];
    }
    break;
  }
  case PLANES: {
    g.rec_size = sizeof(PlaRec);
    g.recs.assign((size_t)n * sizeof(PlaRec), 0);
    PlaRec *r = reinterpret_cast<PlaRec *>(g.recs.data());
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t p = g.order[i];
      for (int k = 0; k < 3; k++) r[i].c[k] = g.centers[3ull * p + k];
      r[i].dx = g.dxs[p];
      r[i].dy = g.dys[p];
      for (int k = 0; k < 9; k++) r[i].R[k] = g.rot[9ull * p + k];
    }
    break;
  }
  case USER_COLLECTION:  // a user collection's tree: the primitives stay with the user
    g.rec_size = 0;
    g.recs.clear();
    break;
  default: {
    g.rec_size = sizeof(CylRec);
    g.recs.assign((size_t)n * sizeof(CylRec), 0);
    CylRec *r = reinterpret_cast<CylRec *>(g.recs.data());
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t p = g.order[i];
      for (int k = 0; k < 3; k++) r[i].c[k] = g.centers[3ull * p + k];
      r[i].a = g.sa[p];
      r[i].b = g.sb[p];
      r[i].h = g.heights[p];
      for (int k = 0; k < 9; k++) r[i].R[k] = g.rot[9ull * p + k];
    }
    break;
  }
  }
}

unsigned hw_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return h ? std::min(h, 64u) : 4u;
}

// Scene<float>: the same children-in-parent layout with FP32 boxes, and FP32 records in leaf order.
void flatten_f32(GeometryF &g, GeomType type) {
  const auto &T = g.tree;
  std::vector<uint32_t> internal(T.size(), INVALID_ID);
  uint32_t ni = 0;
  for (size_t i = 0; i < T.size(); i++)
    if (!T[i].leaf) internal[i] = ni++;
  g.n32.assign(ni, Node32{});
  auto child_ref = [&](uint32_t c, uint32_t &ref, uint32_t &cnt) {
    if (T[c].leaf) {
      ref = LEAF_BIT | T[c].prim_begin;
      cnt = T[c].prim_count;
    } else {
      ref = internal[c];
      cnt = 0;
    }
  };
  for (size_t i = 0; i < T.size(); i++) {
    if (T[i].leaf) continue;
    Node32 &a = g.n32[internal[i]];
    for (int c = 0; c < 2; c++) {
      const brt_tree_node &ch = T[T[i].children[c]];
      for (int k = 0; k < 3; k++) {  // the tree holds FP32 values widened to FP64: narrowing is exact
        a.b[6 * c + k] = (float)ch.min[k];
        a.b[6 * c + 3 + k] = (float)ch.max[k];
      }
      child_ref(T[i].children[c], a.ref[c], a.cnt[c]);
    }
    a.cnt[1] |= T[i].axis << AXIS_SHIFT;
  }
  BvhView &v = g.view;
  v = BvhView{};
  bool irregular = false;
  for (const auto &x : T)
    for (int k = 0; k < 3; k++) irregular = irregular || x.min[k] != x.min[k] || x.max[k] != x.max[k] || x.min[k] > x.max[k];
  for (int k = 0; k < 3; k++) {
    v.root_min[k] = T[0].min[k];
    v.root_max[k] = T[0].max[k];
  }
  child_ref(0, v.root_ref, v.root_cnt);
  v.root_ref4 = v.root_ref;
  v.geom_id = g.geom_id;
  v.n_prims = g.boxes.n;
  v.n_nodes = ni;
  v.depth = g.stats.max_tree_depth;
  v.irregular_boxes = irregular ? 1u : 0u;
  g.extent_ok = false;  // no split precision for FP32 scenes
  const uint32_t n = g.boxes.n;
  auto fill = [&](auto *r, auto &&one) {
    for (uint32_t i = 0; i < n; i++) one(r[i], g.order[i]);
  };
  switch (type) {
  case TRIANGLES:
    g.rec_size = sizeof(TriRecF);
    g.recs.assign((size_t)n * sizeof(TriRecF), 0);
    fill(reinterpret_cast<TriRecF *>(g.recs.data()), [&](TriRecF &r, uint32_t p) {
      const float *v0 = &g.vertices[3ull * g.faces[3ull * p]];
      const float *v1 = &g.vertices[3ull * g.faces[3ull * p + 1]];
      const float *v2 = &g.vertices[3ull * g.faces[3ull * p + 2]];
      for (int k = 0; k < 3; k++) {
        r.v0[k] = v0[k];
        r.e1[k] = v0[k] - v1[k];  // Triangle::c = a - b
        r.e2[k] = v2[k] - v0[k];  // Triangle::b = c - a
      }
      tri_normal(r.e1, r.e2, r.n);
    });
    break;
  case SPHERES:
    g.rec_size = sizeof(SphRecF);
    g.recs.assign((size_t)n * sizeof(SphRecF), 0);
    fill(reinterpret_cast<SphRecF *>(g.recs.data()), [&](SphRecF &r, uint32_t p) {
      for (int k = 0; k < 3; k++) r.c[k] = g.centers[3ull * p + k];
      r.r = g.radii[p];
    });
    break;
  case PLANES:
    g.rec_size = sizeof(PlaRecF);
    g.recs.assign((size_t)n * sizeof(PlaRecF), 0);
    fill(reinterpret_cast<PlaRecF *>(g.recs.data()), [&](PlaRecF &r, uint32_t p) {
      for (int k = 0; k < 3; k++) r.c[k] = g.centers[3ull * p + k];
      r.dx = g.dxs[p];
      r.dy = g.dys[p];
      for (int k = 0; k < 9; k++) r.R[k] = g.rot[9ull * p + k];
    });
    break;
  case USER_COLLECTION:
    g.rec_size = 0;
    g.recs.clear();
    break;
  default:
    g.rec_size = sizeof(CylRecF);
    g.recs.assign((size_t)n * sizeof(CylRecF), 0);
    fill(reinterpret_cast<CylRecF *>(g.recs.data()), [&](CylRecF &r, uint32_t p) {
      for (int k = 0; k < 3; k++) r.c[k] = g.centers[3ull * p + k];
      r.a = g.sa[p];
      r.b = g.sb[p];
      r.h = g.heights[p];
      for (int k = 0; k < 9; k++) r.R[k] = g.rot[9ull * p + k];
    });
    break;
  }
}

template <class R>
void build_geometry(GeometryT<R> &g, GeomType type, const brt_build_options &o) {
  const auto t0 = std::chrono::steady_clock::now();
  g.built_on_device = false;
  g.dtree = nullptr;
  g.dtree_nodes = 0;
  const uint32_t n = g.boxes.n;
  g.order.resize(n);
  std::iota(g.order.begin(), g.order.end(), 0u);
  BuilderT<R> b{g.boxes, g.order.data(), o.min_leaf_primitives, o.max_tree_depth, o.bin_size};
  int par = 0;
  for (unsigned t = hw_threads(); t > 1; t >>= 1) par++;
  g.tree.clear();
  g.tree.reserve(n ? 2ull * n : 1);
  b.build(g.tree, 0, n, 0, par);
  g.tree.shrink_to_fit();
  brt_build_stats &s = g.stats;
  s.max_tree_depth = b.cnt.max_depth;
  s.leaf_nodes = b.cnt.leaves;
  s.branch_nodes = b.cnt.branches;
  s.bad_splits = b.cnt.bad_splits;
  s.max_num_prims_per_leaf = b.cnt.max_prims;
  s.min_num_prims_per_leaf = b.cnt.min_prims;
  s.num_primitives = n;
  s.geom_id = g.geom_id;
  s.built_on_device = 0;
  if constexpr (std::is_same<R, double>::value) flatten(g, type);
  else flatten_f32(g, type);
  s.build_time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  s.device_bytes = g.n64.size() * sizeof(Node64) + g.n4.size() * sizeof(Node4) + g.n32.size() * sizeof(Node32) +
                   g.recs.size() + g.order.size() * 4 + g.faces.size() * 4 + (g.vnormals.size() + g.vertices.size()) * sizeof(R);
}

// ------------------------------------------------------------------------------------------------ device
struct DeviceScene {
  int device = -1;
  std::vector<void *> allocs;
  BvhView views[NUM_TYPES]{};
  // One group of launch counters per stream (COUNTER_GROUP): launches on one stream run in order and each clears its
  // group first, so they can share it; launches on different streams never share one.
  std::mutex counter_mu;
  std::vector<std::pair<hipStream_t, unsigned long long *>> stream_counters;
  std::mutex inv_mu;
  uint32_t *inv_order[NUM_TYPES]{};  // primitive id -> leaf position, made on the first distance query
  double *leaf_vertices = nullptr;    // triangles' vertices in leaf order, made on the first closest-surface query
  std::mutex sort_mu;
  std::vector<std::pair<hipStream_t, std::unique_ptr<SortWorkspace>>> sorts;  // one ray-sort workspace per stream
  ~DeviceScene() {
    for (auto &w : sorts)
      if (w.second->base) (void)hipFree(w.second->base);
    for (auto &c : stream_counters) (void)hipFree(c.second);
    for (void *p : allocs) (void)hipFree(p);
  }
  // The counter group of `st`, allocated on the stream's first launch (NULL if that allocation fails).
  unsigned long long *counter_group(hipStream_t st) {
    std::lock_guard<std::mutex> lk(counter_mu);
    for (auto &c : stream_counters)
      if (c.first == st) return c.second;
    void *p = nullptr;
    if (hipMalloc(&p, COUNTER_GROUP * sizeof(unsigned long long)) != hipSuccess) return nullptr;
    stream_counters.emplace_back(st, static_cast<unsigned long long *>(p));
    return stream_counters.back().second;
  }
  SortWorkspace *sort_workspace(hipStream_t st) {
    std::lock_guard<std::mutex> lk(sort_mu);
    for (auto &w : sorts)
      if (w.first == st) return w.second.get();
    sorts.emplace_back(st, std::make_unique<SortWorkspace>());
    return sorts.back().second.get();
  }
  template <class T, class A>
  hipError_t put(const std::vector<T, A> &v, const T **dst) {
    *dst = nullptr;
    if (v.empty()) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    allocs.push_back(p);
    e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    *dst = static_cast<const T *>(p);
    return e;
  }
};

}  // namespace
}  // namespace brt

namespace brt {

int check_launch(const brt_launch &L) {
  if (L.precision > BRT_PRECISION_FP32_TRAV_FP64_HIT) return fail(BRT_ERR_INVALID_ARGUMENT, "unknown precision");
  if (L.schedule > BRT_SCHEDULE_PER_THREAD) return fail(BRT_ERR_INVALID_ARGUMENT, "unknown schedule");
  if (L.refill > 64) return fail(BRT_ERR_INVALID_ARGUMENT, "refill must be 0..64");
  if (L.sort > BRT_SORT_MORTON) return fail(BRT_ERR_INVALID_ARGUMENT, "unknown sort mode");
  if (L.block_size && (L.block_size % 64 || L.block_size > 256))
    return fail(BRT_ERR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
  if (L.stack_ring && L.stack_ring != 4 && L.stack_ring != 8 && L.stack_ring != 16)
    return fail(BRT_ERR_INVALID_ARGUMENT, "stack_ring must be 0, 4, 8 or 16");
  if (L.pool_size && L.pool_size != 32 && L.pool_size != 64 && L.pool_size != 128)
    return fail(BRT_ERR_INVALID_ARGUMENT, "pool_size must be 0, 32, 64 or 128");
  if (L.flags & ~(BRT_FLAG_HITS_INOUT | BRT_FLAG_ANY_HIT)) return fail(BRT_ERR_INVALID_ARGUMENT, "unknown launch flags");
  return BRT_OK;
}

}  // namespace brt

using namespace brt;

struct brt_scene {
  brt_build_options opts{4, 28, 64};
  bool committed = false;
  bool device_commit = false;  // committed with brt_scene_commit_device (rebuilds build on the device too)
  bool f32 = false;            // blazert::Scene<float>: geometry in geoF, FP32 throughout
  uint32_t geometries = 0;
  Geometry geo[NUM_TYPES];
  GeometryF geoF[NUM_TYPES];
  std::unique_ptr<DeviceScene> dev;
  std::mutex upload_mu;
};

namespace {

template <class R>
GeometryT<R> *geo_of(brt_scene *s) {
  if constexpr (std::is_same<R, double>::value) return s->geo;
  else return s->geoF;
}

template <class R>
int add_common(brt_scene *s, uint32_t type, uint32_t *geom_id) {
  if (geom_id) *geom_id = BRT_INVALID_ID;
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (s->f32 != std::is_same<R, float>::value)
    return fail(BRT_ERR_INVALID_ARGUMENT, s->f32 ? "Scene<float>: use the _f32 add functions"
                                                 : "Scene<double>: the _f32 add functions need brt_scene_create_f32");
  if (s->committed) return fail(BRT_ERR_STATE, "geometry added after commit");
  if (geo_of<R>(s)[type].present) return fail(BRT_ERR_STATE, "geometry type already added (one add per type)");
  return BRT_OK;
}

template <class R>
void finish_add(brt_scene *s, uint32_t type, uint32_t *geom_id) {
  GeometryT<R> &g = geo_of<R>(s)[type];
  g.present = true;
  g.geom_id = s->geometries++;
  if (geom_id) *geom_id = g.geom_id;
}

int double_only(const brt_scene *s, const char *what) {
  if (s && s->f32)
    return fail(BRT_ERR_UNSUPPORTED, std::string(what) + " is not available for Scene<float> scenes "
                                     "(add, commit, intersect and intersect1 are)");
  return BRT_OK;
}

hipError_t new_device_scene(int dev, std::unique_ptr<DeviceScene> &out) {
  auto d = std::make_unique<DeviceScene>();
  d->device = dev;
  out = std::move(d);
  return hipSuccess;
}

// Condition of the split-precision path: finite coordinates below 1e30 (its error bound).
int split_supported(const Geometry &g) {
  if (!g.extent_ok) return fail(BRT_ERR_UNSUPPORTED, "split precision needs finite scene coordinates below 1e30");
  return BRT_OK;
}

// Precision a geometry type is traced in. The split path's conservative FP32 boxes bound the primitives only where
// the tree's boxes do; a tree with an inverted or NaN box (a negative sphere radius, NaN input) culls primitives
// its boxes do not contain, which the split path would reach. Such a type is traced by the exact FP64 path, so a
// split-precision request returns the FP64 result for it.
uint32_t trace_precision(uint32_t requested, const BvhView &v) {
  return (requested == BRT_PRECISION_FP32_TRAV_FP64_HIT && v.irregular_boxes) ? BRT_PRECISION_FP64 : requested;
}

int device_error(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? BRT_ERR_OUT_OF_MEMORY : BRT_ERR_DEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

// Split precision: the FP32 pre-test records of a triangle (TriFilt) or sphere (SphFilt) geometry, made on the
// device from its uploaded FP64 records and waited for; none for other types, or when the scene's extent is above
// FILTER_MAX_EXTENT. BLAZERT_AMD_SPLIT_PRETEST=0 in the environment at upload time leaves them out (every
// split-precision primitive test then runs in FP64; the hits are the same, which the tests check).
hipError_t make_prim_filter(DeviceScene &d, BvhView &v, uint32_t type, double extent, hipStream_t st,
                            uint64_t *device_bytes) {
  v.tri_filter = nullptr;
  v.sph_filter = nullptr;
  if (type != TRIANGLES && type != SPHERES && type != CYLINDERS) return hipSuccess;
  if (!(extent <= FILTER_MAX_EXTENT) || v.n_prims == 0 || !v.prims) return hipSuccess;
  const char *env = std::getenv("BLAZERT_AMD_SPLIT_PRETEST");
  if (env && env[0] == '0') return hipSuccess;
  const size_t bytes = type == TRIANGLES ? (size_t)(v.n_prims + FILTER_PAD) * sizeof(TriFilt)
                                         : (size_t)(v.n_prims + SPH_FILTER_PAD) * sizeof(SphFilt);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return e;
  d.allocs.push_back(p);
  e = type == TRIANGLES ? launch_tri_filter_records(v.prims, v.n_prims, p, st)
      : type == SPHERES  ? launch_sph_filter_records(v.prims, v.n_prims, p, st)
                         : launch_cyl_filter_records(v.prims, v.n_prims, p, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (type == TRIANGLES) v.tri_filter = static_cast<const TriFilt *>(p);
  else v.sph_filter = static_cast<const SphFilt *>(p);
  *device_bytes += bytes;
  return e;
}

// Copies one host-built geometry to the device.
int put_geometry(DeviceScene &d, Geometry &g, uint32_t t) {
  if (g.stats.max_tree_depth >= 64) return fail(BRT_ERR_UNSUPPORTED, "tree deeper than the 64-entry device stack");
  BvhView v = g.view;
  hipError_t e = d.put(g.n64, &v.nodes64);
  if (e == hipSuccess) e = d.put(g.n4, &v.nodes4);
  const unsigned char *recs = nullptr;
  if (e == hipSuccess) e = d.put(g.recs, &recs);
  v.prims = recs;
  if (e == hipSuccess) e = d.put(g.order, &v.prim_ids);
  if (t == TRIANGLES) {
    if (e == hipSuccess) e = d.put(g.faces, &v.faces);
    if (e == hipSuccess) e = d.put(g.vnormals, &v.vnormals);
    if (e == hipSuccess) e = d.put(g.vertices, &v.vertices);
  }
  if (e == hipSuccess) e = make_prim_filter(d, v, t, g.extent, nullptr, &g.stats.device_bytes);
  if (e != hipSuccess) return device_error(e, "upload failed");
  d.views[t] = v;
  return BRT_OK;
}

// Copies one host-built Scene<float> geometry to the device.
int put_geometry_f32(DeviceScene &d, GeometryF &g, uint32_t t) {
  if (g.stats.max_tree_depth >= 64) return fail(BRT_ERR_UNSUPPORTED, "tree deeper than the 64-entry device stack");
  BvhView v = g.view;
  hipError_t e = d.put(g.n32, &v.nodes32);
  const unsigned char *recs = nullptr;
  if (e == hipSuccess) e = d.put(g.recs, &recs);
  v.prims = recs;
  if (e == hipSuccess) e = d.put(g.order, &v.prim_ids);
  if (t == TRIANGLES) {
    if (e == hipSuccess) e = d.put(g.faces, &v.faces);
    if (e == hipSuccess) e = d.put(g.vnormals, &v.vnormals32);
  }
  if (e != hipSuccess) return device_error(e, "upload failed");
  d.views[t] = v;
  return BRT_OK;
}

int upload(brt_scene *s) {
  std::lock_guard<std::mutex> lk(s->upload_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRT_ERR_NO_DEVICE, "no HIP device");
  if (s->dev) {
    if (s->dev->device != dev) return fail(BRT_ERR_STATE, "scene was uploaded to another device");
    return BRT_OK;
  }
  std::unique_ptr<DeviceScene> d;
  if (hipError_t e = new_device_scene(dev, d)) return device_error(e, "counter allocation");
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    if (s->f32) {
      GeometryF &g = s->geoF[t];
      if (!g.present || g.boxes.n == 0) continue;
      if (int rc = put_geometry_f32(*d, g, t)) return rc;
      continue;
    }
    Geometry &g = s->geo[t];
    if (!g.present || g.boxes.n == 0) continue;
    if (int rc = put_geometry(*d, g, t)) return rc;
  }
  s->dev = std::move(d);
  return BRT_OK;
}

// Builds every geometry type on the current device (brt_scene_commit_device); types whose boxes or build options
// the device builder does not reproduce bit for bit are built on the host and uploaded. Leaves the scene uploaded.
int commit_on_device(brt_scene *s, hipStream_t st) {
  std::lock_guard<std::mutex> lk(s->upload_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRT_ERR_NO_DEVICE, "no HIP device");
  std::unique_ptr<DeviceScene> d;
  if (hipError_t e = new_device_scene(dev, d)) return device_error(e, "counter allocation");
  const brt_build_options &o = s->opts;
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    Geometry &g = s->geo[t];
    if (!g.present) continue;
    g.built_on_device = false;
    g.dtree = nullptr;
    g.dtree_nodes = 0;
    if (g.boxes.n == 0) {
      build_geometry(g, (GeomType)t, o);
      continue;
    }
    if (!g.boxes.device_buildable || !device_build_supported(o.bin_size, o.max_tree_depth)) {
      build_geometry(g, (GeomType)t, o);
      if (int rc = put_geometry(*d, g, t)) return rc;
      continue;
    }
    const auto t0 = std::chrono::steady_clock::now();
    BvhView v{};
    DeviceBuildIn in{};
    in.type = (GeomType)t;
    in.n = g.boxes.n;
    in.min_leaf = o.min_leaf_primitives;
    in.max_depth = o.max_tree_depth;
    in.nbins = o.bin_size;
    if (t == TRIANGLES) {
      hipError_t e = d->put(g.faces, &v.faces);
      if (e == hipSuccess) e = d->put(g.vnormals, &v.vnormals);
      if (e == hipSuccess) e = d->put(g.vertices, &v.vertices);
      if (e != hipSuccess) return device_error(e, "upload failed");
      in.vertices = v.vertices;
      in.faces = v.faces;
    } else {
      in.box_lo = g.boxes.lo.data();
      in.box_hi = g.boxes.hi.data();
      in.box_ctr = g.boxes.ctr.data();
      in.centers = g.centers.data();
      in.radii = g.radii.data();
      in.dxs = g.dxs.data();
      in.dys = g.dys.data();
      in.rot = g.rot.data();
      in.sa = g.sa.data();
      in.sb = g.sb.data();
      in.heights = g.heights.data();
    }
    DeviceBuildOut r{};
    hipError_t e = device_build(in, r, st);
    if (e != hipSuccess)
      return device_error(e, ("device build failed (build.hip:" + std::to_string(r.failed_line) + ")").c_str());
    for (void *p : {(void *)r.tree, (void *)r.order, (void *)r.n64, (void *)r.n4, r.recs}) d->allocs.push_back(p);
    g.tree.clear();
    g.order.clear();
    g.n64.clear();
    g.n4.clear();
    g.recs.clear();
    g.rec_size = r.rec_size;
    g.num_n4 = r.num_n4;
    g.built_on_device = true;
    g.dtree = r.tree;
    g.dtree_nodes = r.num_nodes;
    brt_build_stats &bs = g.stats;
    bs = brt_build_stats{};
    bs.max_tree_depth = r.max_depth;
    bs.leaf_nodes = r.leaves;
    bs.branch_nodes = r.branches;
    bs.bad_splits = r.bad_splits;
    bs.max_num_prims_per_leaf = r.max_prims;
    bs.min_num_prims_per_leaf = r.min_prims;
    bs.num_primitives = g.boxes.n;
    bs.geom_id = g.geom_id;
    bs.built_on_device = 1;
    // root descriptor as the host flatten writes it; all node boxes lie inside the (finite) root box
    const brt_tree_node &root = r.root;
    double extent = 0.0;
    for (int k = 0; k < 3; k++) {
      v.root_min[k] = root.min[k];
      v.root_max[k] = root.max[k];
      v.root_min32[k] = f32_down(root.min[k]);
      v.root_max32[k] = f32_up(root.max[k]);
      extent = std::max(extent, std::max(std::fabs(root.min[k]), std::fabs(root.max[k])));
    }
    v.root_ref = root.leaf ? (LEAF_BIT | root.prim_begin) : 0u;
    v.root_cnt = root.leaf ? root.prim_count : 0u;
    v.root_ref4 = v.root_ref;
    v.geom_id = g.geom_id;
    v.n_prims = g.boxes.n;
    v.n_nodes = r.num_n64;
    v.depth = r.max_depth;
    v.irregular_boxes = g.boxes.ordered ? 0u : 1u;  // finite boxes (device_buildable): only the order can fail
    g.view = v;
    g.extent = extent;
    g.extent_ok = extent < 1e30;
    if (r.max_depth >= 64) return fail(BRT_ERR_UNSUPPORTED, "tree deeper than the 64-entry device stack");
    v.nodes64 = r.n64;
    v.nodes4 = r.n4;
    v.prims = r.recs;
    v.prim_ids = r.order;
    bs.device_bytes = (uint64_t)r.num_n64 * sizeof(Node64) + (uint64_t)r.num_n4 * sizeof(Node4) +
                      (uint64_t)g.boxes.n * (r.rec_size + 4) + g.faces.size() * 4 + g.vnormals.size() * 8 +
                      g.vertices.size() * 8;
    if (hipError_t fe = make_prim_filter(*d, v, t, g.extent, st, &bs.device_bytes))
      return device_error(fe, "pre-test records");
    d->views[t] = v;
    bs.build_time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  s->dev = std::move(d);
  return BRT_OK;
}

// Reference-form tree and leaf order of a device-built geometry, fetched once for export.
int download_tree(brt_scene *s, Geometry &g, uint32_t t) {
  if (!g.built_on_device || !g.tree.empty()) return BRT_OK;
  if (hipSetDevice(s->dev->device) != hipSuccess) return fail(BRT_ERR_DEVICE, "hipSetDevice failed");
  std::vector<brt_tree_node> tree(g.dtree_nodes);
  std::vector<uint32_t> order(g.boxes.n);
  hipError_t e = hipMemcpy(tree.data(), g.dtree, tree.size() * sizeof(brt_tree_node), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(order.data(), s->dev->views[t].prim_ids, order.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return device_error(e, "tree download failed");
  g.tree.swap(tree);
  g.order.swap(order);
  return BRT_OK;
}

int mirror_check(brt_scene *s, uint64_t n, uint32_t bounces) {
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (s->f32) return fail(BRT_ERR_UNSUPPORTED, "mirror tracing is not available for Scene<float> scenes");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (bounces < 1 || bounces > 8) return fail(BRT_ERR_INVALID_ARGUMENT, "bounces must be 1..8");
  const Geometry &tri = s->geo[TRIANGLES];
  if (!tri.present || tri.boxes.n == 0) return fail(BRT_ERR_UNSUPPORTED, "mirror tracing needs a triangle mesh");
  for (uint32_t t = SPHERES; t < NUM_TYPES; t++)
    if (s->geo[t].present && s->geo[t].boxes.n) return fail(BRT_ERR_UNSUPPORTED, "mirror tracing supports triangle meshes only");
  if (n > 0xFFFFFFFFull / 8) return fail(BRT_ERR_UNSUPPORTED, "too many rays");
  return upload(s);
}



template <class R>
int add_mesh_t(brt_scene *s, const R *vertices, uint32_t nv, const uint32_t *faces, uint32_t nf, uint32_t *geom_id) {
  int rc = add_common<R>(s, TRIANGLES, geom_id);
  if (rc) return rc;
  if ((nv && !vertices) || (nf && !faces)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL vertex or face array");
  for (uint64_t i = 0; i < 3ull * nf; i++)
    if (faces[i] >= nv) return fail(BRT_ERR_INVALID_ARGUMENT, "face index out of range");
  if (nf >= (1u << 30)) return fail(BRT_ERR_UNSUPPORTED, "too many triangles");
  GeometryT<R> &g = geo_of<R>(s)[TRIANGLES];
  g.vertices.assign(vertices, vertices + 3ull * nv);
  g.faces.assign(faces, faces + 3ull * nf);
  mesh_precompute(g, nv);
  g.boxes.classify();
  finish_add<R>(s, TRIANGLES, geom_id);
  return BRT_OK;
}

template <class R>
int add_spheres_t(brt_scene *s, const R *centers, const R *radii, uint32_t n, uint32_t *geom_id) {
  int rc = add_common<R>(s, SPHERES, geom_id);
  if (rc) return rc;
  if (n && (!centers || !radii)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL sphere array");
  if (n >= (1u << 30)) return fail(BRT_ERR_UNSUPPORTED, "too many spheres");
  GeometryT<R> &g = geo_of<R>(s)[SPHERES];
  g.centers.assign(centers, centers + 3ull * n);
  g.radii.assign(radii, radii + n);
  sphere_precompute(g);
  g.boxes.classify();
  finish_add<R>(s, SPHERES, geom_id);
  return BRT_OK;
}

template <class R>
int add_planes_t(brt_scene *s, const R *centers, const R *dxs, const R *dys, const R *rotations, uint32_t n,
                 uint32_t *geom_id) {
  int rc = add_common<R>(s, PLANES, geom_id);
  if (rc) return rc;
  if (n && (!centers || !dxs || !dys || !rotations)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL plane array");
  if (n >= (1u << 30)) return fail(BRT_ERR_UNSUPPORTED, "too many planes");
  GeometryT<R> &g = geo_of<R>(s)[PLANES];
  g.centers.assign(centers, centers + 3ull * n);
  g.dxs.assign(dxs, dxs + n);
  g.dys.assign(dys, dys + n);
  g.rot.assign(rotations, rotations + 9ull * n);
  plane_precompute(g);
  g.boxes.classify();
  finish_add<R>(s, PLANES, geom_id);
  return BRT_OK;
}

template <class R>
int add_cylinders_t(brt_scene *s, const R *centers, const R *a, const R *b, const R *heights, const R *rotations,
                    uint32_t n, uint32_t *geom_id) {
  int rc = add_common<R>(s, CYLINDERS, geom_id);
  if (rc) return rc;
  if (n && (!centers || !a || !b || !heights || !rotations))
    return fail(BRT_ERR_INVALID_ARGUMENT, "NULL cylinder array");
  if (n >= (1u << 30)) return fail(BRT_ERR_UNSUPPORTED, "too many cylinders");
  GeometryT<R> &g = geo_of<R>(s)[CYLINDERS];
  g.centers.assign(centers, centers + 3ull * n);
  g.sa.assign(a, a + n);
  g.sb.assign(b, b + n);
  g.heights.assign(heights, heights + n);
  g.rot.assign(rotations, rotations + 9ull * n);
  cylinder_precompute(g);
  g.boxes.classify();
  finish_add<R>(s, CYLINDERS, geom_id);
  return BRT_OK;
}


template <class R, class H>
int intersect1_t(const brt_scene *s, const R org[3], const R dir[3], R tmin, R tmax, uint32_t flags, H *hit,
                 int *hit_out) {
  if (!s || !org || !dir || !hit) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  // one ray through the batched kernels, keeping RayHit's in/out semantics: 8 ray values, 6 hit values,
  // then prim id, geom id and the merged flag
  R host[14] = {org[0], org[1], org[2], dir[0], dir[1], dir[2], tmin, tmax,
                     hit->t, hit->u, hit->v, hit->normal[0], hit->normal[1], hit->normal[2]};
  uint32_t ids[3] = {hit->prim_id, hit->geom_id, 0u};
  char *d = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&d), sizeof host + sizeof ids);
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  R *dd = reinterpret_cast<R *>(d);
  uint32_t *di = reinterpret_cast<uint32_t *>(d + sizeof host);
  e = hipMemcpy(d, host, sizeof host, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(di, ids, sizeof ids, hipMemcpyHostToDevice);
  int rc = BRT_OK;
  brt_scene *ms = const_cast<brt_scene *>(s);
  if (e == hipSuccess && !ms->committed) rc = fail(BRT_ERR_STATE, "scene not committed");
  if (e == hipSuccess && rc == BRT_OK) rc = upload(ms);
  if (e == hipSuccess && rc == BRT_OK) {
    RaysViewT<R> rv{dd, dd + 1, dd + 2, dd + 3, dd + 4, dd + 5, dd + 6, dd + 7};
    HitsViewT<R> hv{dd + 8, dd + 9, dd + 10, di, di + 1, dd + 11, dd + 12, dd + 13, di + 2};
    for (uint32_t t = 0; t < NUM_TYPES && e == hipSuccess; t++) {
      const GeometryT<R> &g = geo_of<R>(ms)[t];
      if (!g.present || g.boxes.n == 0) continue;
      TraceLaunch cfg{BRT_PRECISION_FP64, flags | BRT_FLAG_HITS_INOUT, 64, false, SCHEDULE_PER_THREAD, 0,
                      ms->dev->counter_group(nullptr), nullptr};
      if (!cfg.counter) {
        e = hipErrorOutOfMemory;
        break;
      }
      if constexpr (std::is_same<R, float>::value)
        e = launch_trace_f32((GeomType)t, ms->dev->views[t], rv, hv, 1, cfg, nullptr);
      else
        e = launch_trace((GeomType)t, ms->dev->views[t], rv, hv, 1, cfg, SplitParams{g.extent}, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(host, d, sizeof host, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ids, di, sizeof ids, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
      hit->t = host[8];
      hit->u = host[9];
      hit->v = host[10];
      hit->normal[0] = host[11];
      hit->normal[1] = host[12];
      hit->normal[2] = host[13];
      hit->prim_id = ids[0];
      hit->geom_id = ids[1];
      if (hit_out) *hit_out = ids[2] ? 1 : 0;
    }
  }
  (void)hipFree(d);
  if (rc != BRT_OK) return rc;
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("intersect1: ") + hipGetErrorString(e));
  return BRT_OK;
}

}  // namespace

extern "C" {

const char *brt_last_error(void) { return g_error.c_str(); }
const char *brt_version(void) { return "blazert_amd 0.1 (gfx950)"; }

int brt_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int brt_scene_create(const brt_build_options *opts, brt_scene **out) {
  if (!out) return fail(BRT_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  brt_scene *s = new (std::nothrow) brt_scene();
  if (!s) return fail(BRT_ERR_OUT_OF_MEMORY, "scene allocation failed");
  if (opts) {
    if (opts->bin_size < 2 || opts->min_leaf_primitives == 0) {
      delete s;
      return fail(BRT_ERR_INVALID_ARGUMENT, "bin_size must be >= 2 and min_leaf_primitives >= 1");
    }
    s->opts = *opts;
  }
  *out = s;
  return BRT_OK;
}

int brt_scene_destroy(brt_scene *scene) {
  delete scene;
  return BRT_OK;
}

int brt_scene_add_mesh(brt_scene *s, const double *vertices, uint32_t nv, const uint32_t *faces, uint32_t nf,
                       uint32_t *geom_id) {
  return add_mesh_t(s, vertices, nv, faces, nf, geom_id);
}
int brt_scene_add_spheres(brt_scene *s, const double *centers, const double *radii, uint32_t n, uint32_t *geom_id) {
  return add_spheres_t(s, centers, radii, n, geom_id);
}
int brt_scene_add_planes(brt_scene *s, const double *centers, const double *dxs, const double *dys,
                         const double *rotations, uint32_t n, uint32_t *geom_id) {
  return add_planes_t(s, centers, dxs, dys, rotations, n, geom_id);
}
int brt_scene_add_cylinders(brt_scene *s, const double *centers, const double *a, const double *b,
                            const double *heights, const double *rotations, uint32_t n, uint32_t *geom_id) {
  return add_cylinders_t(s, centers, a, b, heights, rotations, n, geom_id);
}

int brt_scene_create_f32(const brt_build_options *opts, brt_scene **out) {
  const int rc = brt_scene_create(opts, out);
  if (rc == BRT_OK) (*out)->f32 = true;
  return rc;
}
int brt_scene_add_mesh_f32(brt_scene *s, const float *vertices, uint32_t nv, const uint32_t *faces, uint32_t nf,
                           uint32_t *geom_id) {
  return add_mesh_t(s, vertices, nv, faces, nf, geom_id);
}
int brt_scene_add_spheres_f32(brt_scene *s, const float *centers, const float *radii, uint32_t n, uint32_t *geom_id) {
  return add_spheres_t(s, centers, radii, n, geom_id);
}
int brt_scene_add_planes_f32(brt_scene *s, const float *centers, const float *dxs, const float *dys,
                             const float *rotations, uint32_t n, uint32_t *geom_id) {
  return add_planes_t(s, centers, dxs, dys, rotations, n, geom_id);
}
int brt_scene_add_cylinders_f32(brt_scene *s, const float *centers, const float *a, const float *b,
                                const float *heights, const float *rotations, uint32_t n, uint32_t *geom_id) {
  return add_cylinders_t(s, centers, a, b, heights, rotations, n, geom_id);
}

int brt_scene_commit(brt_scene *s) {
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (s->committed) return BRT_OK;
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    if (s->f32 && s->geoF[t].present) build_geometry(s->geoF[t], (GeomType)t, s->opts);
    if (!s->f32 && s->geo[t].present) build_geometry(s->geo[t], (GeomType)t, s->opts);
  }
  s->committed = true;
  return BRT_OK;
}

int brt_scene_commit_device(brt_scene *s, void *stream) {
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (int rc = double_only(s, "brt_scene_commit_device")) return rc;
  if (s->committed) return BRT_OK;
  if (int rc = commit_on_device(s, static_cast<hipStream_t>(stream))) return rc;
  s->committed = true;
  s->device_commit = true;
  return BRT_OK;
}

int brt_scene_read_device_array(const brt_scene *cs, uint32_t type, uint32_t which, void *dst, uint64_t capacity,
                                uint64_t *bytes) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (int rc = double_only(s, "brt_scene_read_device_array")) return rc;
  if (!s || type >= NUM_TYPES || !bytes || which > BRT_ARRAY_PRIM_IDS)
    return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  const Geometry &g = s->geo[type];
  if (!g.present) return fail(BRT_ERR_INVALID_ARGUMENT, "geometry type not in scene");
  *bytes = 0;
  if (g.boxes.n == 0) return BRT_OK;
  if (int rc = upload(s)) return rc;
  const BvhView &v = s->dev->views[type];
  const void *src = nullptr;
  switch (which) {
  case BRT_ARRAY_NODES64: src = v.nodes64; *bytes = (uint64_t)v.n_nodes * sizeof(Node64); break;
  case BRT_ARRAY_NODES4: src = v.nodes4; *bytes = (uint64_t)g.num_n4 * sizeof(Node4); break;
  case BRT_ARRAY_PRIMITIVES: src = v.prims; *bytes = (uint64_t)g.boxes.n * g.rec_size; break;
  default: src = v.prim_ids; *bytes = (uint64_t)g.boxes.n * 4; break;
  }
  if (!dst || *bytes == 0) return BRT_OK;
  if (capacity < *bytes) return fail(BRT_ERR_INVALID_ARGUMENT, "capacity too small");
  if (hipSetDevice(s->dev->device) != hipSuccess) return fail(BRT_ERR_DEVICE, "hipSetDevice failed");
  hipError_t e = hipMemcpy(dst, src, *bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return device_error(e, "read failed");
  return BRT_OK;
}

int brt_distance_to_surface(const brt_scene *cs, uint32_t type, const brt_points *points, const uint32_t *prim_ids,
                            uint64_t n, double *out, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (int rc = double_only(s, "brt_distance_to_surface")) return rc;
  if (!s || !points || type >= NUM_TYPES) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (!s->geo[type].present) return fail(BRT_ERR_INVALID_ARGUMENT, "the scene holds no geometry of this type");
  if (n == 0) return BRT_OK;
  if (!points->x || !points->y || !points->z || !prim_ids || !out)
    return fail(BRT_ERR_INVALID_ARGUMENT, "point, id and output arrays are required");
  if (int rc = upload(s)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Geometry &g = s->geo[type];
  BvhView v = s->dev->views[type];
  v.n_prims = g.boxes.n;  // an empty geometry has no device arrays: every id is out of range
  const uint32_t *inv = nullptr;
  if (type != TRIANGLES && g.boxes.n) {
    std::lock_guard<std::mutex> lk(s->dev->inv_mu);
    uint32_t *&slot = s->dev->inv_order[type];
    if (!slot) {
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, (size_t)g.boxes.n * 4);
      if (e != hipSuccess) return device_error(e, "inverse order allocation");
      s->dev->allocs.push_back(p);
      e = launch_inverse_order(v.prim_ids, g.boxes.n, static_cast<uint32_t *>(p), st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);  // complete before a query on another stream reads it
      if (e != hipSuccess) return device_error(e, "launch failed");
      slot = static_cast<uint32_t *>(p);
    }
    inv = slot;
  }
  hipError_t e = launch_distance_pairs((GeomType)type, v, inv, points->x, points->y, points->z, prim_ids, n, out, st);
  if (e != hipSuccess) return device_error(e, "launch failed");
  return BRT_OK;
}

int brt_closest_surface(const brt_scene *cs, const brt_points *points, uint64_t n, double *distance,
                        uint32_t *prim_id, uint32_t *geom_id, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (int rc = double_only(s, "brt_closest_surface")) return rc;
  if (!s || !points) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (n == 0) return BRT_OK;
  if (!points->x || !points->y || !points->z || !distance || !prim_id || !geom_id)
    return fail(BRT_ERR_INVALID_ARGUMENT, "point and output arrays are required");
  if (int rc = upload(s)) return rc;
  ClosestViews cv{};
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    const Geometry &g = s->geo[t];
    cv.present[t] = (g.present && g.boxes.n) ? 1u : 0u;
    if (!cv.present[t]) continue;
    cv.view[t] = s->dev->views[t];
    cv.extent[t] = g.extent_ok ? g.extent : INFINITY;
  }
  if (cv.present[TRIANGLES]) {
    std::lock_guard<std::mutex> lk(s->dev->inv_mu);
    if (!s->dev->leaf_vertices) {
      const Geometry &g = s->geo[TRIANGLES];
      const BvhView &v = s->dev->views[TRIANGLES];
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, (size_t)g.boxes.n * 9 * sizeof(double));
      if (e != hipSuccess) return device_error(e, "leaf-order vertex allocation");
      s->dev->allocs.push_back(p);
      const hipStream_t st = static_cast<hipStream_t>(stream);
      e = launch_leaf_vertices(v.prim_ids, v.faces, v.vertices, g.boxes.n, static_cast<double *>(p), st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);  // complete before a query on another stream reads it
      if (e != hipSuccess) return device_error(e, "launch failed");
      s->dev->leaf_vertices = static_cast<double *>(p);
    }
    cv.leaf_vertices = s->dev->leaf_vertices;
  }
  hipError_t e = launch_closest_surface(cv, points->x, points->y, points->z, n, distance, prim_id, geom_id,
                                        static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return device_error(e, "launch failed");
  return BRT_OK;
}

int brt_scene_upload(brt_scene *s) {
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  return upload(s);
}

int brt_scene_build_stats(const brt_scene *s, uint32_t type, brt_build_stats *stats) {
  if (!s || !stats || type >= NUM_TYPES) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (s->f32) {
    if (!s->geoF[type].present) return fail(BRT_ERR_INVALID_ARGUMENT, "geometry type not in scene");
    *stats = s->geoF[type].stats;
    return BRT_OK;
  }
  if (!s->geo[type].present) return fail(BRT_ERR_INVALID_ARGUMENT, "geometry type not in scene");
  *stats = s->geo[type].stats;
  return BRT_OK;
}

int brt_scene_export_tree(const brt_scene *s, uint32_t type, brt_tree_node *nodes, uint32_t capacity,
                          uint32_t *prim_order, uint32_t *num_nodes) {
  if (!s || type >= NUM_TYPES || !num_nodes) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (s->f32) {
    const GeometryF &g = s->geoF[type];
    if (!g.present) return fail(BRT_ERR_INVALID_ARGUMENT, "geometry type not in scene");
    *num_nodes = (uint32_t)g.tree.size();
    if (nodes) {
      if (capacity < g.tree.size()) return fail(BRT_ERR_INVALID_ARGUMENT, "capacity too small");
      std::memcpy(nodes, g.tree.data(), g.tree.size() * sizeof(brt_tree_node));
    }
    if (prim_order && !g.order.empty()) std::memcpy(prim_order, g.order.data(), g.order.size() * 4);
    return BRT_OK;
  }
  Geometry &g = const_cast<brt_scene *>(s)->geo[type];
  if (!g.present) return fail(BRT_ERR_INVALID_ARGUMENT, "geometry type not in scene");
  if (int rc = download_tree(const_cast<brt_scene *>(s), g, type)) return rc;
  *num_nodes = (uint32_t)g.tree.size();
  if (nodes) {
    if (capacity < g.tree.size()) return fail(BRT_ERR_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(nodes, g.tree.data(), g.tree.size() * sizeof(brt_tree_node));
  }
  if (prim_order && !g.order.empty()) std::memcpy(prim_order, g.order.data(), g.order.size() * 4);
  return BRT_OK;
}

// brt_intersect, and with `noalloc` brt_intersect_noalloc: the launch counters and the ray-sort workspace then come
// from the library's static device scratch (noalloc_scratch) instead of per-stream and per-scene allocations, and a
// scene not yet on the device is an error rather than uploaded.
static int intersect_impl(brt_scene *s, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                          const brt_launch *launch, hipStream_t st, bool noalloc) {
  if (!s || !rays || !hits) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (int rc = double_only(s, "brt_intersect (use brt_intersect_f32)")) return rc;
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (n == 0) return BRT_OK;
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "ray origin / direction arrays are required");
  if (!hits->t || !hits->prim_id || !hits->geom_id)
    return fail(BRT_ERR_INVALID_ARGUMENT, "hit t / prim_id / geom_id arrays are required");
  if ((hits->nx != nullptr) != (hits->ny != nullptr) || (hits->nx != nullptr) != (hits->nz != nullptr))
    return fail(BRT_ERR_INVALID_ARGUMENT, "normal arrays must be all set or all NULL");
  brt_launch L{};
  L.precision = BRT_PRECISION_FP64;
  L.schedule = BRT_SCHEDULE_PERSISTENT;
  if (launch) L = *launch;
  int rc = check_launch(L);
  if (rc) return rc;
  unsigned char *scratch = nullptr;
  if (noalloc) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(BRT_ERR_NO_DEVICE, "no HIP device");
    if (!s->dev || s->dev->device != dev)
      return fail(BRT_ERR_STATE, "brt_intersect_noalloc needs the scene on the current device already "
                                 "(brt_scene_upload, brt_scene_commit_device or an earlier brt_intersect allocate)");
    if (L.sort != BRT_SORT_NONE && n > NOALLOC_MAX_SORTED_RAYS)
      return fail(BRT_ERR_UNSUPPORTED, "brt_intersect_noalloc sorts at most " +
                                           std::to_string(NOALLOC_MAX_SORTED_RAYS) +
                                           " rays (its sort workspace is fixed device scratch); got " + std::to_string(n));
    if (hipError_t e = noalloc_scratch(&scratch)) return device_error(e, "scratch lookup");
  } else {
    rc = upload(s);
    if (rc) return rc;
  }
  RaysView rv{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  HitsView hv{hits->t, hits->u, hits->v, hits->prim_id, hits->geom_id, hits->nx, hits->ny, hits->nz, nullptr};
  const uint32_t *order = nullptr;
  if (L.sort != BRT_SORT_NONE) {
    if (n > 0x7fffffffull) return fail(BRT_ERR_UNSUPPORTED, "sorted launches are limited to 2^31-1 rays");
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t t = 0; t < NUM_TYPES; t++) {
      const Geometry &g = s->geo[t];
      if (!g.present || g.boxes.n == 0) continue;
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], g.view.root_min[k]);
        hi[k] = std::max(hi[k], g.view.root_max[k]);
      }
    }
    SortWorkspace fixed;
    SortWorkspace *ws = &fixed;
    hipError_t e = hipSuccess;
    if (noalloc) {
      constexpr size_t cbytes = COUNTER_GROUP * sizeof(unsigned long long);  // the launch counters come first
      e = sort_workspace_in(fixed, scratch + cbytes, NOALLOC_SCRATCH_BYTES - cbytes, n);
      if (e == hipErrorInvalidValue) return fail(BRT_ERR_UNSUPPORTED, "sort workspace does not fit the scratch");
    } else {
      ws = s->dev->sort_workspace(st);
      e = sort_workspace_reserve(*ws, n);
    }
    if (e == hipSuccess) e = sort_rays(rv, n, lo, hi, L.sort, *ws, st, &order);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("ray sort failed: ") + hipGetErrorString(e));
  }
  bool first = true;
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    const Geometry &g = s->geo[t];
    if (!g.present || g.boxes.n == 0) continue;
    if (L.precision == BRT_PRECISION_FP32_TRAV_FP64_HIT)
      if (int rc = split_supported(g)) return rc;
    TraceLaunch cfg{trace_precision(L.precision, s->dev->views[t]), L.flags, L.block_size, first, L.schedule,
                    L.refill, nullptr, order};
    cfg.ring = L.stack_ring;
    cfg.pool = L.pool_size;
    cfg.counter = noalloc ? reinterpret_cast<unsigned long long *>(scratch) : s->dev->counter_group(st);
    if (!cfg.counter) return fail(BRT_ERR_OUT_OF_MEMORY, "launch counters");
    SplitParams sp{g.extent};
    hipError_t e = launch_trace((GeomType)t, s->dev->views[t], rv, hv, n, cfg, sp, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
    first = false;
  }
  if (first && !(L.flags & BRT_FLAG_HITS_INOUT)) {
    hipError_t e = launch_init_hits(hv, n, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  }
  return BRT_OK;
}

int brt_intersect(const brt_scene *cs, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                  const brt_launch *launch, void *stream) {
  return intersect_impl(const_cast<brt_scene *>(cs), rays, hits, n, launch, static_cast<hipStream_t>(stream), false);
}

int brt_intersect_noalloc(const brt_scene *cs, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                          const brt_launch *launch, void *stream) {
  return intersect_impl(const_cast<brt_scene *>(cs), rays, hits, n, launch, static_cast<hipStream_t>(stream), true);
}

int brt_traverse(const brt_scene *cs, uint32_t geom_type, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                 const brt_launch *launch, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (!s || !rays || !hits) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (int rc = double_only(s, "brt_traverse")) return rc;
  if (geom_type >= NUM_TYPES) return fail(BRT_ERR_INVALID_ARGUMENT, "unknown geometry type");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (!s->geo[geom_type].present) return fail(BRT_ERR_INVALID_ARGUMENT, "the scene holds no geometry of this type");
  if (n == 0) return BRT_OK;
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "ray origin / direction arrays are required");
  if (!hits->t || !hits->prim_id || !hits->geom_id)
    return fail(BRT_ERR_INVALID_ARGUMENT, "hit t / prim_id / geom_id arrays are required");
  if ((hits->nx != nullptr) != (hits->ny != nullptr) || (hits->nx != nullptr) != (hits->nz != nullptr))
    return fail(BRT_ERR_INVALID_ARGUMENT, "normal arrays must be all set or all NULL");
  brt_launch L{};
  if (launch) L = *launch;
  if (int rc = check_launch(L)) return rc;
  if (L.sort != BRT_SORT_NONE) return fail(BRT_ERR_UNSUPPORTED, "traverse keeps the input ray order");
  if (L.flags & ~BRT_FLAG_ANY_HIT) return fail(BRT_ERR_INVALID_ARGUMENT, "traverse takes only BRT_FLAG_ANY_HIT");
  const Geometry &g = s->geo[geom_type];
  if (g.boxes.n == 0) return BRT_OK;
  if (L.precision == BRT_PRECISION_FP32_TRAV_FP64_HIT)
    if (int rc = split_supported(g)) return rc;
  if (int rc = upload(s)) return rc;
  RaysView rv{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  HitsView hv{hits->t, hits->u, hits->v, hits->prim_id, hits->geom_id, hits->nx, hits->ny, hits->nz, nullptr};
  TraceLaunch cfg{trace_precision(L.precision, s->dev->views[geom_type]), L.flags | BRT_FLAG_HITS_INOUT,
                  L.block_size, true, L.schedule, L.refill, nullptr, nullptr};
  cfg.ring = L.stack_ring;
  cfg.pool = L.pool_size;
  cfg.overwrite = true;
  cfg.counter = s->dev->counter_group(static_cast<hipStream_t>(stream));
  if (!cfg.counter) return fail(BRT_ERR_OUT_OF_MEMORY, "launch counters");
  SplitParams sp{g.extent};
  hipError_t e = launch_trace((GeomType)geom_type, s->dev->views[geom_type], rv, hv, n, cfg, sp,
                              static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  return BRT_OK;
}

int brt_intersect1(const brt_scene *s, const double org[3], const double dir[3], double tmin, double tmax,
                   uint32_t flags, brt_hit1 *hit, int *hit_out) {
  if (int rc = double_only(s, "brt_intersect1 (use brt_intersect1_f32)")) return rc;
  return intersect1_t(s, org, dir, tmin, tmax, flags, hit, hit_out);
}

int brt_intersect1_f32(const brt_scene *s, const float org[3], const float dir[3], float tmin, float tmax,
                       uint32_t flags, brt_hit1_f32 *hit, int *hit_out) {
  if (s && !s->f32) return fail(BRT_ERR_INVALID_ARGUMENT, "brt_intersect1_f32 needs a Scene<float>");
  return intersect1_t(s, org, dir, tmin, tmax, flags, hit, hit_out);
}

int brt_intersect_f32(const brt_scene *cs, const brt_rays_f32 *rays, const brt_hits_f32 *hits, uint64_t n,
                      const brt_launch *launch, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (!s || !rays || !hits) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!s->f32) return fail(BRT_ERR_INVALID_ARGUMENT, "brt_intersect_f32 needs a Scene<float>");
  if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
  if (n == 0) return BRT_OK;
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "ray origin / direction arrays are required");
  if (!hits->t || !hits->prim_id || !hits->geom_id)
    return fail(BRT_ERR_INVALID_ARGUMENT, "hit t / prim_id / geom_id arrays are required");
  if ((hits->nx != nullptr) != (hits->ny != nullptr) || (hits->nx != nullptr) != (hits->nz != nullptr))
    return fail(BRT_ERR_INVALID_ARGUMENT, "normal arrays must be all set or all NULL");
  brt_launch L{};
  L.schedule = BRT_SCHEDULE_PERSISTENT;
  if (launch) L = *launch;
  if (int rc = check_launch(L)) return rc;
  if (L.precision != BRT_PRECISION_FP64)
    return fail(BRT_ERR_UNSUPPORTED, "a Scene<float> is traversed in FP32: precision must be BRT_PRECISION_FP64 (0)");
  if (L.sort != BRT_SORT_NONE) return fail(BRT_ERR_UNSUPPORTED, "Scene<float> launches keep the input ray order");
  if (int rc = upload(s)) return rc;
  RaysViewF rv{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  HitsViewF hv{hits->t, hits->u, hits->v, hits->prim_id, hits->geom_id, hits->nx, hits->ny, hits->nz, nullptr};
  hipStream_t st = static_cast<hipStream_t>(stream);
  bool first = true;
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    const GeometryF &g = s->geoF[t];
    if (!g.present || g.boxes.n == 0) continue;
    TraceLaunch cfg{L.precision, L.flags, L.block_size, first, L.schedule, L.refill, nullptr, nullptr};
    cfg.ring = L.stack_ring;
    cfg.pool = L.pool_size;
    cfg.counter = s->dev->counter_group(st);
    if (!cfg.counter) return fail(BRT_ERR_OUT_OF_MEMORY, "launch counters");
    hipError_t e = launch_trace_f32((GeomType)t, s->dev->views[t], rv, hv, n, cfg, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
    first = false;
  }
  if (first && !(L.flags & BRT_FLAG_HITS_INOUT)) {
    hipError_t e = launch_init_hits_f32(hv, n, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  }
  return BRT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------ user collections
// blazert::BVH<T, Collection> over a user-defined primitive type (blazert/bvh/accel.h:32-52): the reference's
// SAHBinnedBuilder run on the collection's boxes and centers, flattened to the device node records. The primitives
// stay with the user; device code traverses the tree with its own intersector (include/blazert_amd/user_geometry.h).
struct brt_bvh {
  bool f32 = false;
  brt_build_options opts{4, 28, 64};
  Geometry g;
  GeometryF gF;
  std::mutex mu;
  int device = -1;  // device the tree was uploaded to (-1: not yet)
  std::vector<void *> allocs;
  brt_bvh_view view{};
  ~brt_bvh() {
    for (void *p : allocs) (void)hipFree(p);
  }
};

namespace {

template <class R>
int bvh_build_t(const R *box_min, const R *box_max, const R *centers, uint32_t n, const brt_build_options *opts,
                brt_bvh **out) {
  if (!out) return fail(BRT_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (n && (!box_min || !box_max || !centers)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL box or center array");
  if (opts && (opts->bin_size < 2 || opts->min_leaf_primitives == 0))
    return fail(BRT_ERR_INVALID_ARGUMENT, "bin_size must be >= 2 and min_leaf_primitives >= 1");
  std::unique_ptr<brt_bvh> b(new (std::nothrow) brt_bvh());
  if (!b) return fail(BRT_ERR_OUT_OF_MEMORY, "bvh allocation failed");
  if (opts) b->opts = *opts;
  b->f32 = std::is_same<R, float>::value;
  GeometryT<R> *g;
  if constexpr (std::is_same<R, double>::value) g = &b->g;
  else g = &b->gF;
  g->present = true;
  g->geom_id = 0;
  g->boxes.resize(n);
  std::copy(box_min, box_min + 3ull * n, g->boxes.lo.begin());
  std::copy(box_max, box_max + 3ull * n, g->boxes.hi.begin());
  std::copy(centers, centers + 3ull * n, g->boxes.ctr.begin());
  g->boxes.classify();
  build_geometry(*g, USER_COLLECTION, b->opts);
  *out = b.release();
  return BRT_OK;
}

template <class R>
int bvh_upload(brt_bvh *b, GeometryT<R> &g) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRT_ERR_NO_DEVICE, "no HIP device");
  if (b->device >= 0) {
    if (b->device != dev) return fail(BRT_ERR_STATE, "the tree was uploaded to another device");
    return BRT_OK;
  }
  if (g.stats.max_tree_depth >= 64) return fail(BRT_ERR_UNSUPPORTED, "tree deeper than the 64-entry device stack");
  brt_bvh_view v{};
  v.real_bytes = sizeof(R);
  v.num_prims = g.boxes.n;
  v.depth = g.stats.max_tree_depth;
  if (g.boxes.n) {
    auto put = [&](const void *src, size_t bytes, const void **dst) -> hipError_t {
      *dst = nullptr;
      if (!bytes) return hipSuccess;
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) return e;
      b->allocs.push_back(p);
      *dst = p;
      return hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
    };
    const void *nodes = nullptr, *ids = nullptr;
    hipError_t e;
    if constexpr (std::is_same<R, double>::value) e = put(g.n64.data(), g.n64.size() * sizeof(Node64), &nodes);
    else e = put(g.n32.data(), g.n32.size() * sizeof(Node32), &nodes);
    if (e == hipSuccess) e = put(g.order.data(), g.order.size() * 4, &ids);
    if (e != hipSuccess) return device_error(e, "upload failed");
    v.nodes = nodes;
    v.prim_ids = static_cast<const uint32_t *>(ids);
    for (int k = 0; k < 3; k++) {
      v.root_min[k] = g.view.root_min[k];
      v.root_max[k] = g.view.root_max[k];
    }
    v.root_ref = g.view.root_ref;
    v.root_cnt = g.view.root_cnt;
    v.num_nodes = g.view.n_nodes;
    v.irregular_boxes = g.view.irregular_boxes;
  }
  b->view = v;
  b->device = dev;
  return BRT_OK;
}

}  // namespace

extern "C" {

int brt_bvh_build(const double *box_min, const double *box_max, const double *centers, uint32_t n,
                  const brt_build_options *opts, brt_bvh **out) {
  return bvh_build_t(box_min, box_max, centers, n, opts, out);
}

int brt_bvh_build_f32(const float *box_min, const float *box_max, const float *centers, uint32_t n,
                      const brt_build_options *opts, brt_bvh **out) {
  return bvh_build_t(box_min, box_max, centers, n, opts, out);
}

int brt_bvh_device_view(brt_bvh *b, brt_bvh_view *view) {
  if (!b || !view) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  std::lock_guard<std::mutex> lk(b->mu);
  const int rc = b->f32 ? bvh_upload(b, b->gF) : bvh_upload(b, b->g);
  if (rc) return rc;
  *view = b->view;
  return BRT_OK;
}

int brt_bvh_build_stats(const brt_bvh *b, brt_build_stats *stats) {
  if (!b || !stats) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  *stats = b->f32 ? b->gF.stats : b->g.stats;
  return BRT_OK;
}

int brt_bvh_export_tree(const brt_bvh *b, brt_tree_node *nodes, uint32_t capacity, uint32_t *prim_order,
                        uint32_t *num_nodes) {
  if (!b || !num_nodes) return fail(BRT_ERR_INVALID_ARGUMENT, "bad argument");
  const std::vector<brt_tree_node> &tree = b->f32 ? b->gF.tree : b->g.tree;
  const std::vector<uint32_t> &order = b->f32 ? b->gF.order : b->g.order;
  *num_nodes = (uint32_t)tree.size();
  if (nodes) {
    if (capacity < tree.size()) return fail(BRT_ERR_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(nodes, tree.data(), tree.size() * sizeof(brt_tree_node));
  }
  if (prim_order && !order.empty()) std::memcpy(prim_order, order.data(), order.size() * 4);
  return BRT_OK;
}

int brt_bvh_destroy(brt_bvh *b) {
  delete b;
  return BRT_OK;
}

// ---------------------------------------------------------------------------------------------- batches
}  // extern "C"

struct brt_batch {
  int device = -1;
  brt_launch L{};
  std::vector<brt_batch_item> items;
  void *d_jobs[NUM_TYPES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t njobs[NUM_TYPES] = {0, 0, 0, 0};
  uint64_t total[NUM_TYPES] = {0, 0, 0, 0};
  unsigned long long *counters = nullptr;
  std::vector<size_t> empty_items;  // items whose scene holds no primitive: their hits are initialised
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~brt_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : d_jobs)
      if (p) (void)hipFree(p);
    if (counters) (void)hipFree(counters);
  }
};

namespace {

int batch_enqueue(brt_batch *b, hipStream_t st) {
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    if (!b->njobs[t]) continue;
    TraceLaunch cfg{b->L.precision, b->L.flags, b->L.block_size, false, SCHEDULE_PERSISTENT, b->L.refill,
                    b->counters + (size_t)t * COUNTER_GROUP, nullptr};
    cfg.ring = b->L.stack_ring;
    cfg.pool = b->L.pool_size;
    hipError_t e = launch_trace_batch((GeomType)t, static_cast<const BatchJob *>(b->d_jobs[t]), b->njobs[t],
                                      b->total[t], cfg, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("batch launch failed: ") + hipGetErrorString(e));
  }
  if (!(b->L.flags & BRT_FLAG_HITS_INOUT))
    for (size_t i : b->empty_items) {
      const brt_hits &h = b->items[i].hits;
      HitsView hv{h.t, h.u, h.v, h.prim_id, h.geom_id, h.nx, h.ny, h.nz, nullptr};
      hipError_t e = launch_init_hits(hv, b->items[i].n, st);
      if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
    }
  return BRT_OK;
}

}  // namespace

extern "C" {

int brt_batch_create(const brt_batch_item *items, uint32_t count, const brt_launch *launch, brt_batch **out) {
  if (!out || (count && !items)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  auto b = std::make_unique<brt_batch>();
  b->L.precision = BRT_PRECISION_FP64;
  if (launch) b->L = *launch;
  if (int rc = check_launch(b->L)) return rc;
  if (b->L.sort != BRT_SORT_NONE) return fail(BRT_ERR_UNSUPPORTED, "batches trace rays in input order");
  if (hipGetDevice(&b->device) != hipSuccess) return fail(BRT_ERR_NO_DEVICE, "no HIP device");
  b->items.assign(items, items + count);
  std::vector<BatchJob> jobs[NUM_TYPES];
  for (uint32_t i = 0; i < count; i++) {
    const brt_batch_item &it = items[i];
    brt_scene *s = const_cast<brt_scene *>(it.scene);
    if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "item without a scene");
    if (int rc = double_only(s, "brt_batch_create")) return rc;
    if (!s->committed) return fail(BRT_ERR_STATE, "scene not committed");
    if (it.n && (!it.rays.ox || !it.rays.oy || !it.rays.oz || !it.rays.dx || !it.rays.dy || !it.rays.dz ||
                 !it.hits.t || !it.hits.prim_id || !it.hits.geom_id))
      return fail(BRT_ERR_INVALID_ARGUMENT, "item without ray or hit arrays");
    if ((it.hits.nx != nullptr) != (it.hits.ny != nullptr) || (it.hits.nx != nullptr) != (it.hits.nz != nullptr))
      return fail(BRT_ERR_INVALID_ARGUMENT, "normal arrays must be all set or all NULL");
    int rc = upload(s);
    if (rc) return rc;
    if (it.n == 0) continue;
    bool first = true;
    for (uint32_t t = 0; t < NUM_TYPES; t++) {
      const Geometry &g = s->geo[t];
      if (!g.present || g.boxes.n == 0) continue;
      if (b->L.precision == BRT_PRECISION_FP32_TRAV_FP64_HIT) {
        if (int rc = split_supported(g)) return rc;
        if (s->dev->views[t].irregular_boxes)
          return fail(BRT_ERR_UNSUPPORTED, "split-precision batches need trees without inverted or NaN boxes "
                                           "(such a scene is traced by the FP64 path in brt_intersect)");
      }
      BatchJob j{};
      j.bvh = s->dev->views[t];
      j.rays = RaysView{it.rays.ox, it.rays.oy, it.rays.oz, it.rays.dx, it.rays.dy, it.rays.dz, it.rays.tmin,
                        it.rays.tmax};
      j.hits = HitsView{it.hits.t,  it.hits.u,  it.hits.v,  it.hits.prim_id, it.hits.geom_id,
                        it.hits.nx, it.hits.ny, it.hits.nz, nullptr};
      j.n = it.n;
      j.offset = b->total[t];
      j.first = first ? 1u : 0u;
      j.is_tri = t == TRIANGLES ? 1u : 0u;
      j.extent = g.extent;
      j.order = nullptr;
      b->total[t] += it.n;
      jobs[t].push_back(j);
      first = false;
    }
    if (first) b->empty_items.push_back(i);
  }
  for (uint32_t t = 0; t < NUM_TYPES; t++) {
    b->njobs[t] = (uint32_t)jobs[t].size();
    if (jobs[t].empty()) continue;
    hipError_t e = hipMalloc(&b->d_jobs[t], jobs[t].size() * sizeof(BatchJob));
    if (e == hipSuccess)
      e = hipMemcpy(b->d_jobs[t], jobs[t].data(), jobs[t].size() * sizeof(BatchJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("batch table: ") + hipGetErrorString(e));
  }
  if (hipMalloc(reinterpret_cast<void **>(&b->counters), NUM_TYPES * COUNTER_GROUP * sizeof(unsigned long long)) !=
      hipSuccess)
    return fail(BRT_ERR_OUT_OF_MEMORY, "batch counters");
  *out = b.release();
  return BRT_OK;
}

int brt_batch_launch(brt_batch *b, void *stream) {
  if (!b) return fail(BRT_ERR_INVALID_ARGUMENT, "batch is NULL");
  return batch_enqueue(b, static_cast<hipStream_t>(stream));
}

int brt_batch_capture(brt_batch *b, void *stream) {
  if (!b || !stream) return fail(BRT_ERR_INVALID_ARGUMENT, "batch and a non-default stream are required");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("begin capture: ") + hipGetErrorString(e));
  const int rc = batch_enqueue(b, st);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(st, &g);
  if (rc != BRT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("end capture: ") + hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
  return BRT_OK;
}

int brt_batch_replay(brt_batch *b, void *stream) {
  if (!b || !b->exec) return fail(BRT_ERR_STATE, "batch has no captured graph");
  hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
  return BRT_OK;
}

int brt_batch_destroy(brt_batch *b) {
  delete b;
  return BRT_OK;
}

// ------------------------------------------------------------------------------------------------ mirror tracing
int brt_mirror_trace(const brt_scene *cs, const brt_rays *rays, uint64_t n, uint32_t bounces, double bounce_tmin,
                     const brt_launch *launch, const brt_mirror_out *out, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (!rays || !out) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "ray origin / direction arrays are required");
  if (!out->t || !out->prim_id || !out->px || !out->py || !out->pz || !out->rx || !out->ry || !out->rz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "t, prim_id, position and direction arrays are required");
  if ((out->u != nullptr) != (out->v != nullptr)) return fail(BRT_ERR_INVALID_ARGUMENT, "u and v must both be set or NULL");
  if (!(bounce_tmin >= 0.0)) return fail(BRT_ERR_INVALID_ARGUMENT, "bounce_tmin must be >= 0");
  brt_launch L{};
  if (launch) L = *launch;
  if (int rc = check_launch(L)) return rc;
  if (L.flags) return fail(BRT_ERR_INVALID_ARGUMENT, "mirror tracing takes no launch flags");
  if (L.sort != BRT_SORT_NONE) return fail(BRT_ERR_UNSUPPORTED, "mirror tracing does not sort rays");
  int rc = mirror_check(s, n, bounces);
  if (rc) return rc;
  if (n == 0) return BRT_OK;
  if (L.precision == BRT_PRECISION_FP32_TRAV_FP64_HIT)
    if (int rc2 = split_supported(s->geo[TRIANGLES])) return rc2;
  const BvhView &bvh = s->dev->views[TRIANGLES];
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (uint32_t b = 0; b < bounces; b++) {
    const uint64_t off = (uint64_t)b * n;
    RaysView rv{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    if (b > 0) {
      const uint64_t prev = off - n;
      rv = RaysView{out->px + prev, out->py + prev, out->pz + prev, out->rx + prev, out->ry + prev, out->rz + prev,
                    nullptr, nullptr};
      rv.tmin0 = bounce_tmin;
    }
    HitsView hv{out->t + off, out->u ? out->u + off : nullptr, out->v ? out->v + off : nullptr, out->prim_id + off,
                nullptr, nullptr, nullptr, nullptr, nullptr};
    TraceLaunch cfg{trace_precision(L.precision, bvh), 0u, L.block_size, true, L.schedule, L.refill, nullptr,
                    nullptr};
    cfg.ring = L.stack_ring;
    cfg.pool = L.pool_size;
    cfg.counter = s->dev->counter_group(st);
    if (!cfg.counter) return fail(BRT_ERR_OUT_OF_MEMORY, "launch counters");
    SplitParams sp{s->geo[TRIANGLES].extent};
    hipError_t e = launch_trace(TRIANGLES, bvh, rv, hv, n, cfg, sp, st);
    double *pos[3] = {out->px + off, out->py + off, out->pz + off};
    double *dir[3] = {out->rx + off, out->ry + off, out->rz + off};
    if (e == hipSuccess) e = launch_mirror_step(rv, hv, bvh.vertices, bvh.faces, n, pos, dir, st);
    if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  }
  return BRT_OK;
}

int brt_mirror_adjoint(const brt_scene *cs, const brt_rays *rays, uint64_t n, uint32_t bounces,
                       const uint32_t *prim_id, const brt_mirror_grads *grads, void *stream) {
  brt_scene *s = const_cast<brt_scene *>(cs);
  if (!rays || !prim_id || !grads) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
    return fail(BRT_ERR_INVALID_ARGUMENT, "ray origin / direction arrays are required");
  const bool gp = grads->gpx || grads->gpy || grads->gpz, go = grads->gox || grads->goy || grads->goz,
             gd = grads->gdx || grads->gdy || grads->gdz;
  if ((gp && !(grads->gpx && grads->gpy && grads->gpz)) || (go && !(grads->gox && grads->goy && grads->goz)) ||
      (gd && !(grads->gdx && grads->gdy && grads->gdz)))
    return fail(BRT_ERR_INVALID_ARGUMENT, "x, y, z gradient arrays must be all set or all NULL");
  int rc = mirror_check(s, n, bounces);
  if (rc) return rc;
  if (n == 0) return BRT_OK;
  const BvhView &bvh = s->dev->views[TRIANGLES];
  RaysView rv{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, nullptr, nullptr};
  MirrorGrads g{grads->gt,  grads->gpx, grads->gpy, grads->gpz, grads->vertices, grads->gox,
                grads->goy, grads->goz, grads->gdx, grads->gdy, grads->gdz};
  hipError_t e = launch_mirror_adjoint(rv, prim_id, bounces, bvh.vertices, bvh.faces, bvh.n_prims, n, g,
                                       static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(BRT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  return BRT_OK;
}

}  // extern "C"

namespace brt {

int scene_rebuild(brt_scene *s, const brt_build_options &opts) {
  if (!s) return fail(BRT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (int rc = double_only(s, "leaf-size tuning")) return rc;
  if (opts.min_leaf_primitives == 0 || opts.bin_size == 0)
    return fail(BRT_ERR_INVALID_ARGUMENT, "leaf size and bin count must be positive");
  {
    std::lock_guard<std::mutex> lk(s->upload_mu);
    s->dev.reset();
    s->opts = opts;
  }
  if (s->device_commit) {
    s->committed = false;
    if (int rc = commit_on_device(s, nullptr)) return rc;
  } else {
    for (uint32_t t = 0; t < NUM_TYPES; t++)
      if (s->geo[t].present) build_geometry(s->geo[t], (GeomType)t, s->opts);
  }
  s->committed = true;
  return BRT_OK;
}

brt_build_options scene_options(const brt_scene *s) { return s->opts; }

brt_launch batch_launch_config(const brt_batch *b) { return b->L; }

int batch_set_launch_config(brt_batch *b, const brt_launch &L) {
  if (b->exec) return fail(BRT_ERR_STATE, "the batch has been captured; tune it before brt_batch_capture");
  if (int rc = check_launch(L)) return rc;
  if (L.sort != BRT_SORT_NONE) return fail(BRT_ERR_UNSUPPORTED, "batches trace rays in input order");
  const uint32_t precision = b->L.precision, flags = b->L.flags;
  b->L = L;
  b->L.precision = precision;
  b->L.flags = flags;
  return BRT_OK;
}

}  // namespace brt
