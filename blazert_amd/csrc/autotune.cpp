// Autotuner: chooses the launch parameters (threads per block, persistent refill threshold, LDS stack depth, ray
// sort key, ray-pool claim size) and the BVH leaf size by timing candidates on the caller's own device rays with HIP events. Every
// candidate computes the same hits (bit for bit on the FP64 path; on every non-tie ray with split precision), so
// tuning changes speed only. Results are returned to the caller, who may keep them (the Python layer caches them
// per device and workload).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/blazert_amd.h"
#include "internal.h"

using brt::fail;

namespace {

constexpr uint32_t BLOCKS[] = {64, 128, 192, 256};
constexpr uint32_t REFILLS[] = {4, 8, 16, 32};
constexpr uint32_t RINGS[] = {4, 8, 16};
constexpr uint32_t SORTS[] = {BRT_SORT_NONE, BRT_SORT_OCTANT_MORTON};
constexpr uint32_t POOLS[] = {32, 128};  // tried with the best shape found (the default is 64)

// Mean time of `repeats` enqueues of `run` on `st` after one untimed warm-up; < 0 with *rc set on failure.
double time_ms(const std::function<int()> &run, hipStream_t st, uint32_t repeats, int *rc) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    *rc = fail(BRT_ERR_DEVICE, "hipEventCreate failed");
    return -1.0;
  }
  double ms = -1.0;
  *rc = run();
  if (*rc == BRT_OK && hipEventRecord(a, st) == hipSuccess) {
    for (uint32_t i = 0; i < repeats && *rc == BRT_OK; i++) *rc = run();
    float e = 0.0f;
    if (*rc == BRT_OK) {
      if (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
          hipEventElapsedTime(&e, a, b) != hipSuccess)
        *rc = fail(BRT_ERR_DEVICE, "timing the candidate failed");
      else
        ms = (double)e / repeats;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

brt_launch default_launch(uint32_t precision, uint32_t flags) {
  brt_launch L{};
  L.precision = precision;
  L.flags = flags;
  L.schedule = BRT_SCHEDULE_PERSISTENT;
  return L;
}

}  // namespace

extern "C" {

int brt_autotune_launch(const brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n,
                        uint32_t precision, uint32_t flags, uint32_t repeats, brt_tune_result *out, void *stream) {
  if (!scene || !rays || !hits || !out) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n == 0) return fail(BRT_ERR_INVALID_ARGUMENT, "no rays to time");
  if (flags & BRT_FLAG_HITS_INOUT) return fail(BRT_ERR_INVALID_ARGUMENT, "tuning rewrites the hits: no HITS_INOUT");
  if (repeats == 0) repeats = 3;
  hipStream_t st = static_cast<hipStream_t>(stream);
  brt_launch L = default_launch(precision, flags);
  if (int rc = brt::check_launch(L)) return rc;
  int rc = BRT_OK;
  auto timed = [&](const brt_launch &c) {
    return time_ms([&] { return brt_intersect(scene, rays, hits, n, &c, stream); }, st, repeats, &rc);
  };
  const double base = timed(L);
  if (rc) return rc;
  brt_tune_result r{};
  r.launch = L;
  r.best_ms = r.default_ms = base;
  r.candidates = 1;
  for (uint32_t block : BLOCKS)
    for (uint32_t refill : REFILLS)
      for (uint32_t ring : RINGS)
        for (uint32_t sort : SORTS) {
          brt_launch c = L;
          c.block_size = block;
          c.refill = refill;
          c.stack_ring = ring;
          c.sort = sort;
          const double ms = timed(c);
          if (rc) return rc;
          r.candidates++;
          if (ms < r.best_ms) {
            r.best_ms = ms;
            r.launch = c;
          }
        }
  // then the ray-pool claim size with the best shape
  for (uint32_t pool : POOLS) {
    brt_launch c = r.launch;
    c.pool_size = pool;
    const double ms = timed(c);
    if (rc) return rc;
    r.candidates++;
    if (ms < r.best_ms) {
      r.best_ms = ms;
      r.launch = c;
    }
  }
  *out = r;
  return BRT_OK;
}

int brt_batch_autotune(brt_batch *batch, uint32_t repeats, brt_tune_result *out, void *stream) {
  if (!batch || !out) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (repeats == 0) repeats = 3;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const brt_launch orig = brt::batch_launch_config(batch);
  int rc = BRT_OK;
  auto timed = [&](const brt_launch &c) {
    rc = brt::batch_set_launch_config(batch, c);
    if (rc) return -1.0;
    return time_ms([&] { return brt_batch_launch(batch, stream); }, st, repeats, &rc);
  };
  brt_tune_result r{};
  r.launch = default_launch(orig.precision, orig.flags);
  r.default_ms = r.best_ms = timed(r.launch);
  r.candidates = 1;
  for (uint32_t block : BLOCKS)
    for (uint32_t refill : REFILLS)
      for (uint32_t ring : RINGS) {
        if (rc) break;
        brt_launch c = r.launch;
        c.block_size = block;
        c.refill = refill;
        c.stack_ring = ring;
        c.sort = BRT_SORT_NONE;
        const double ms = timed(c);
        if (rc) break;
        r.candidates++;
        if (ms < r.best_ms) {
          r.best_ms = ms;
          r.launch = c;
        }
      }
  for (uint32_t pool : POOLS) {
    if (rc) break;
    brt_launch c = r.launch;
    c.pool_size = pool;
    const double ms = timed(c);
    if (rc) break;
    r.candidates++;
    if (ms < r.best_ms) {
      r.best_ms = ms;
      r.launch = c;
    }
  }
  if (rc) {
    (void)brt::batch_set_launch_config(batch, orig);
    return rc;
  }
  r.launch.precision = orig.precision;
  r.launch.flags = orig.flags;
  if ((rc = brt::batch_set_launch_config(batch, r.launch))) return rc;
  *out = r;
  return BRT_OK;
}

int brt_scene_tune_leaf(brt_scene *scene, const brt_rays *rays, const brt_hits *hits, uint64_t n, uint32_t precision,
                        const uint32_t *leaf_sizes, uint32_t count, uint32_t repeats, brt_tune_result *out,
                        void *stream) {
  if (!scene || !rays || !hits || !out || (count && !leaf_sizes)) return fail(BRT_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n == 0 || count == 0) return fail(BRT_ERR_INVALID_ARGUMENT, "no rays or no leaf sizes to time");
  if (repeats == 0) repeats = 3;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const brt_launch L = default_launch(precision, 0);
  if (int rc = brt::check_launch(L)) return rc;
  brt_build_options opts = brt::scene_options(scene);
  brt_tune_result r{};
  r.launch = L;
  r.best_ms = -1.0;
  int rc = BRT_OK;
  uint32_t best_leaf = 0;
  for (uint32_t i = 0; i < count; i++) {
    opts.min_leaf_primitives = leaf_sizes[i];
    if ((rc = brt::scene_rebuild(scene, opts))) return rc;
    const double ms = time_ms([&] { return brt_intersect(scene, rays, hits, n, &L, stream); }, st, repeats, &rc);
    if (rc) return rc;
    r.candidates++;
    if (leaf_sizes[i] == 4) r.default_ms = ms;
    if (r.best_ms < 0.0 || ms < r.best_ms) {
      r.best_ms = ms;
      best_leaf = leaf_sizes[i];
    }
  }
  opts.min_leaf_primitives = best_leaf;
  if ((rc = brt::scene_rebuild(scene, opts))) return rc;
  r.min_leaf_primitives = best_leaf;
  *out = r;
  return BRT_OK;
}

}  // extern "C"
