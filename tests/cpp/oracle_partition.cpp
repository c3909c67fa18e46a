// The oracle builder's partition (oracle/blazert_oracle.c, partition_idx) against libstdc++'s std::partition,
// which the reference builder calls on a std::vector<unsigned>::iterator range with the predicate
// "center[axis] < cut" (blazert/bvh/builder.h:213-215). Both run on identical random index ranges, centers and
// cuts; the resulting permutations and split points must be identical. Cases include many equal centers, cuts
// equal to a center, all-pass and all-fail ranges, empty and one-element ranges, and sub-ranges of a larger index
// array. Prints one line and exits non-zero on the first mismatch.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" uint32_t orc_partition_test(const double *centers, uint32_t *idx, uint32_t n, uint32_t axis, double cut);

int main() {
  std::mt19937_64 rng(20240611);
  long cases = 0, all_pass = 0, all_fail = 0, mixed = 0, equal_cut = 0;
  for (int trial = 0; trial < 200000; trial++) {
    const uint32_t nprim = (uint32_t)(rng() % 97);
    const int levels = 1 + (int)(rng() % 9);  // few distinct coordinates: many equal centers
    std::vector<double> centers(3 * (size_t)nprim + 3);
    for (auto &c : centers) c = (double)(int)(rng() % levels) * 0.25 - 1.0 + ((rng() % 4 == 0) ? 1e-9 * (double)(rng() % 5) : 0.0);
    std::vector<uint32_t> idx(nprim);
    for (uint32_t i = 0; i < nprim; i++) idx[i] = i;
    std::shuffle(idx.begin(), idx.end(), rng);
    // a sub-range [b, e) of the index array, as a node's range inside the builder's index vector
    const uint32_t b = nprim ? (uint32_t)(rng() % (nprim + 1)) : 0u;
    const uint32_t e = b + (uint32_t)(rng() % (nprim - b + 1));
    const uint32_t axis = (uint32_t)(rng() % 3);
    double cut;
    const uint32_t kind = (uint32_t)(rng() % 8);
    if (kind < 4 && e > b)
      cut = centers[3 * (size_t)idx[b + rng() % (e - b)] + axis];  // equal to a center
    else if (kind == 4)
      cut = -2.0;  // all fail
    else if (kind == 5)
      cut = 2.0;  // all pass
    else
      cut = -1.25 + 2.5 * (double)(rng() % 1000) / 1000.0;
    std::vector<uint32_t> ref = idx, mine = idx;
    const auto mid = std::partition(ref.begin() + b, ref.begin() + e,
                                    [&](uint32_t i) { return centers[3 * (size_t)i + axis] < cut; });
    const uint32_t m = orc_partition_test(centers.data(), mine.data() + b, e - b, axis, cut);
    if ((uint32_t)(mid - (ref.begin() + b)) != m || mine != ref) {
      std::printf("partition mismatch: trial %d, range [%u, %u) of %u, axis %u, cut %.17g\n", trial, b, e, nprim, axis,
                  cut);
      return 1;
    }
    cases++;
    all_pass += (e > b && m == e - b) ? 1 : 0;
    all_fail += (e > b && m == 0) ? 1 : 0;
    mixed += (m > 0 && m < e - b) ? 1 : 0;
    for (uint32_t i = b; i < e; i++) equal_cut += (centers[3 * (size_t)idx[i] + axis] == cut) ? 1 : 0;
  }
  std::printf("oracle partition equals std::partition on %ld random ranges (%ld mixed, %ld all-pass, %ld all-fail, "
              "%ld centers equal to the cut)\n", cases, mixed, all_pass, all_fail, equal_cut);
  return 0;
}
