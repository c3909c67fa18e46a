// Host-side interface between the C ABI (scene.cpp) and the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

namespace brt {

constexpr uint32_t SCHEDULE_PERSISTENT = 0;  // persistent waves with per-lane ray refill (default)
constexpr uint32_t SCHEDULE_PER_THREAD = 1;  // one ray per thread, grid-stride

struct TraceLaunch {
  uint32_t precision;  // BRT_PRECISION_*
  uint32_t flags;      // BRT_FLAG_*
  uint32_t block;      // threads per block
  bool first;          // first geometry type of the scene: initialises every hit record
  uint32_t schedule;   // SCHEDULE_*
  uint32_t refill;     // persistent: refill a wave once this many lanes are idle (1..64)
  unsigned long long *counter;  // persistent: device ray counter for this launch (zeroed on the stream)
};

// Per-scene constants for the split-precision path: the largest finite |coordinate| of any node box.
struct SplitParams {
  double scene_extent;
};

// Enqueues the traversal of one geometry type over n rays on `stream` (asynchronous).
hipError_t launch_trace(GeomType type, const BvhView &bvh, const RaysView &rays, const HitsView &hits, uint64_t n,
                        const TraceLaunch &cfg, const SplitParams &split, hipStream_t stream);

// Writes the default RayHit (t = DBL_MAX, ids invalid, u = v = 0, normal 0) for n rays.
hipError_t launch_init_hits(const HitsView &hits, uint64_t n, hipStream_t stream);

}  // namespace brt
