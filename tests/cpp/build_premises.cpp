// CPU check of the two order-dependent steps the device builder (blazert_amd/csrc/build.hip) computes in
// parallel, against their sequential forms:
//   1. libstdc++'s std::partition on random-access iterators equals "swap the i-th failing element of the left
//      part with the i-th passing element of the right part counted from the end" (the left part being the first
//      T positions, T = number of passing elements);
//   2. the SAH sweep of find_best_split_binned (blazert/bvh/binbuffer.h:115-166), run as prefix / suffix unions
//      with "first cost below DBL_MAX that equals the minimum", picks the same bin and axis as the sequential loop,
//      empty bins (NaN costs) and ties included.
// Prints one line per check and exits non-zero on the first mismatch.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <random>
#include <vector>

namespace {

struct Bin {
  double lo[3], hi[3];
  uint32_t count;
};

double surface(const double *lo, const double *hi) {
  const double b0 = std::fabs(hi[0] - lo[0]), b1 = std::fabs(hi[1] - lo[1]), b2 = std::fabs(hi[2] - lo[2]);
  return 2.0 * (b0 * b1 + b1 * b2 + b2 * b0);
}

// the sequential sweep of one axis (as the host builder runs it)
uint32_t sweep_sequential(const std::vector<Bin> &ax, uint32_t nb, double &min_cost) {
  std::vector<double> rc(nb, 0.0);
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  uint64_t count = 0;
  for (uint32_t i = nb - 1; i > 0; i--) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(ax[i].lo[k], lo[k]);
      hi[k] = std::max(ax[i].hi[k], hi[k]);
    }
    count += ax[i].count;
    rc[i] = (double)count * surface(lo, hi);
  }
  for (int k = 0; k < 3; k++) {
    lo[k] = DBL_MAX;
    hi[k] = -DBL_MAX;
  }
  count = 0;
  min_cost = DBL_MAX;
  uint32_t min_bin = 1;
  for (uint32_t i = 0; i + 1 < nb; i++) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(ax[i].lo[k], lo[k]);
      hi[k] = std::max(ax[i].hi[k], hi[k]);
    }
    count += ax[i].count;
    const double cost = (double)count * surface(lo, hi) + rc[i + 1];
    if (cost < min_cost) {
      min_cost = cost;
      min_bin = i + 1;
    }
  }
  return min_bin;
}

// the device form: lane l holds bin l (identity beyond nb), inclusive suffix and prefix unions, then the first
// lane whose cost is below DBL_MAX and equal to the smallest such cost
uint32_t sweep_lanes(const std::vector<Bin> &ax, uint32_t nb, double &min_cost) {
  const int L = 64;
  std::vector<double> rc(L), cost(L);
  for (int l = 0; l < L; l++) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    uint64_t c = 0;
    for (int i = l; i < (int)nb; i++) {
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], ax[i].lo[k]);
        hi[k] = std::max(hi[k], ax[i].hi[k]);
      }
      c += ax[i].count;
    }
    rc[l] = (double)c * surface(lo, hi);
  }
  double m = INFINITY;
  for (int l = 0; l < L; l++) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    uint64_t c = 0;
    for (int i = 0; i <= l && i < (int)nb; i++) {
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], ax[i].lo[k]);
        hi[k] = std::max(hi[k], ax[i].hi[k]);
      }
      c += ax[i].count;
    }
    cost[l] = (double)c * surface(lo, hi) + rc[l + 1 < L ? l + 1 : l];
    if (l + 1 < (int)nb && cost[l] < DBL_MAX && cost[l] < m) m = cost[l];
  }
  min_cost = DBL_MAX;
  if (!(m < DBL_MAX)) return 1;
  for (int l = 0; l + 1 < (int)nb; l++)
    if (cost[l] < DBL_MAX && cost[l] == m) {
      min_cost = m;
      return (uint32_t)l + 1;
    }
  return 1;
}

}  // namespace

int main() {
  std::mt19937_64 rng(12345);
  for (int trial = 0; trial < 50000; trial++) {
    const int n = 1 + (int)(rng() % 300);
    std::vector<int> a(n), key(n);
    for (int i = 0; i < n; i++) {
      a[i] = i;
      key[i] = (int)(rng() % 7);
    }
    const int cut = (int)(rng() % 8);
    auto pred = [&](int x) { return key[x] < cut; };
    std::vector<int> ref = a;
    const auto mid = std::partition(ref.begin(), ref.end(), pred);
    int T = 0;
    for (int x : a) T += pred(x) ? 1 : 0;
    std::vector<int> b = a, pos_r;
    for (int i = T; i < n; i++)
      if (pred(a[i])) pos_r.push_back(i);
    const int mis = (int)pos_r.size();
    int r = 0;
    for (int i = 0; i < T; i++)
      if (!pred(a[i])) std::swap(b[i], b[pos_r[mis - 1 - r++]]);
    if (mid - ref.begin() != T || r != mis || b != ref) {
      std::printf("partition mismatch: trial %d, n %d\n", trial, n);
      return 1;
    }
  }
  std::printf("partition: rank pairing equals std::partition on 50000 random cases\n");

  std::uniform_real_distribution<double> U(-4.0, 4.0);
  int ties = 0, empties = 0;
  for (int trial = 0; trial < 20000; trial++) {
    const uint32_t nb = 2 + (uint32_t)(rng() % 63);
    std::vector<Bin> ax(nb);
    const bool coarse = rng() % 3 == 0;  // few distinct coordinates: equal costs
    for (uint32_t i = 0; i < nb; i++) {
      Bin &x = ax[i];
      const bool empty = rng() % 3 == 0;
      x.count = empty ? 0u : 1u + (uint32_t)(rng() % 9);
      for (int k = 0; k < 3; k++) {
        double p = U(rng), q = U(rng);
        if (coarse) {
          p = std::round(p);
          q = std::round(q);
        }
        x.lo[k] = empty ? DBL_MAX : std::min(p, q);
        x.hi[k] = empty ? -DBL_MAX : std::max(p, q);
      }
      empties += empty ? 1 : 0;
    }
    double c1 = 0.0, c2 = 0.0;
    const uint32_t b1 = sweep_sequential(ax, nb, c1), b2 = sweep_lanes(ax, nb, c2);
    if (b1 != b2 || !(c1 == c2)) {
      std::printf("sweep mismatch: trial %d, nb %u: bin %u vs %u, cost %.17g vs %.17g\n", trial, nb, b1, b2, c1, c2);
      return 1;
    }
    ties += coarse ? 1 : 0;
  }
  std::printf("sweep: lane form equals the sequential loop on 20000 random bin sets (%d coarse, %d empty bins)\n",
              ties, empties);
  return 0;
}
