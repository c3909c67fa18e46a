// Microbenchmarks that measure the ceilings the traversal and adjoint kernels are compared with (gfx950). Not part of
// the ray-query library: built into libbrt_microbench.so and run by `bench.py --microbench`.
//
//   brt_mb_line_chase   lanes that each follow a chain of dependent random 128-byte lines (the traversal's access
//                       pattern: the next record's address comes from the current one)
//   brt_mb_line_gather  independent random 128-byte line reads, with plain or non-temporal loads
//   brt_mb_fp64_atomics random FP64 atomic adds into an (n, 3) array: one component per lane, or three adjacent
//                       lanes adding x, y and z of one row (the adjoint's vertex scatter)
//
// Each function allocates its buffers, times 5 launches after one warm-up with HIP events and frees everything.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

using u32x4 = unsigned int __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t mix(uint64_t x) {
  x *= 0x9E3779B97F4A7C15ull;
  x ^= x >> 31;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 29;
  return x;
}

__global__ void k_chase(const u32x4 *buf, uint64_t lines, uint32_t steps, unsigned *sink) {
  uint64_t l = mix(blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) % lines;
  unsigned acc = 0;
  for (uint32_t s = 0; s < steps; s++) {
    const u32x4 *p = buf + l * 8;
    u32x4 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = p[k];
#pragma unroll
    for (int k = 0; k < 8; k++) acc += v[k].x;
    l = mix(l ^ acc) % lines;  // the next line depends on the data just read
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads; practically never true
}

template <bool NT>
__global__ void k_gather(const u32x4 *buf, uint64_t lines, uint64_t n, unsigned *sink) {
  unsigned acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const u32x4 *p = buf + (mix(i + 7) % lines) * 8;
    u32x4 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = NT ? __builtin_nontemporal_load(p + k) : p[k];
#pragma unroll
    for (int k = 0; k < 8; k++) acc ^= v[k].x ^ v[k].w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;
}

// rows: number of (x, y, z) rows; adds: FP64 adds per launch (a multiple of 3 when triples)
__global__ void k_atomics(double *buf, uint64_t rows, uint64_t adds, int triples, uint64_t seed) {
  const uint32_t lane = threadIdx.x & 63u;
  if (!triples) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < adds;
         i += (uint64_t)gridDim.x * blockDim.x)
      unsafeAtomicAdd(&buf[3 * (mix(i + seed) % rows) + i % 3], 1.0);  // a random row per lane
    return;
  }
  // three adjacent lanes per row, 21 rows per wave instruction
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t nrows = adds / 3;
  for (uint64_t base = wave * 21; base < nrows; base += waves * 21) {
    const uint64_t r = base + lane / 3;
    if (lane < 63 && r < nrows) unsafeAtomicAdd(&buf[3 * (mix(r + seed) % rows) + lane % 3], 1.0);
  }
}

template <class F>
hipError_t time5(F launch, double *ms) {
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipEventRecord(a, 0);
  for (int r = 0; r < 5 && e == hipSuccess; r++) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(b, 0);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  *ms = t / 5.0;
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  return e;
}

}  // namespace

extern "C" {

// Chains of dependent random 128-byte lines in a buffer of `bytes`: `lanes` chains of `steps` links each.
// Writes GB/s of lines read.
int brt_mb_line_chase(uint64_t bytes, uint32_t lanes, uint32_t steps, double *gbs) {
  const uint64_t lines = bytes / 128;
  if (lines == 0 || lanes == 0 || lanes % 256 != 0 || !gbs) return -1;
  u32x4 *buf = nullptr;
  unsigned *sink = nullptr;
  hipError_t e = hipMalloc(&buf, lines * 128);
  if (e == hipSuccess) e = hipMalloc(&sink, 64);
  if (e == hipSuccess) e = hipMemset(buf, 1, lines * 128);
  double ms = 0.0;
  if (e == hipSuccess)
    e = time5([&] { hipLaunchKernelGGL(k_chase, dim3(lanes / 256), dim3(256), 0, 0, buf, lines, steps, sink); }, &ms);
  if (e == hipSuccess) *gbs = 128.0 * lanes * (double)steps / (ms * 1e6);
  if (buf) (void)hipFree(buf);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? 0 : -(int)e;
}

// `gathers` independent random 128-byte line reads from a buffer of `bytes`, plain or non-temporal loads.
int brt_mb_line_gather(uint64_t bytes, uint64_t gathers, int nontemporal, double *gbs) {
  const uint64_t lines = bytes / 128;
  if (lines == 0 || gathers == 0 || !gbs) return -1;
  u32x4 *buf = nullptr;
  unsigned *sink = nullptr;
  hipError_t e = hipMalloc(&buf, lines * 128);
  if (e == hipSuccess) e = hipMalloc(&sink, 64);
  if (e == hipSuccess) e = hipMemset(buf, 1, lines * 128);
  double ms = 0.0;
  if (e == hipSuccess)
    e = time5([&] {
      if (nontemporal) hipLaunchKernelGGL(k_gather<true>, dim3(8192), dim3(256), 0, 0, buf, lines, gathers, sink);
      else hipLaunchKernelGGL(k_gather<false>, dim3(8192), dim3(256), 0, 0, buf, lines, gathers, sink);
    }, &ms);
  if (e == hipSuccess) *gbs = 128.0 * (double)gathers / (ms * 1e6);
  if (buf) (void)hipFree(buf);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? 0 : -(int)e;
}

// `adds` random FP64 atomic adds into a (rows, 3) array; triples = 1 puts x, y, z of one row in three adjacent
// lanes of one instruction. Writes G adds/s.
int brt_mb_fp64_atomics(uint64_t rows, uint64_t adds, int triples, double *gadds) {
  if (rows == 0 || adds == 0 || !gadds) return -1;
  double *buf = nullptr;
  hipError_t e = hipMalloc(&buf, rows * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(buf, 0, rows * 3 * sizeof(double));
  double ms = 0.0;
  if (e == hipSuccess)
    e = time5([&] { hipLaunchKernelGGL(k_atomics, dim3(2048), dim3(256), 0, 0, buf, rows, adds, triples, 11ull); },
              &ms);
  if (e == hipSuccess) *gadds = (double)adds / (ms * 1e6);
  if (buf) (void)hipFree(buf);
  return e == hipSuccess ? 0 : -(int)e;
}

}  // extern "C"
