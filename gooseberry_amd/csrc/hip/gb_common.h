// gb_common.h -- shared device helpers for the gfx950 kernels.
#ifndef GB_COMMON_H
#define GB_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../gb_device.h"

#define GB_WAVE 64
#define GB_BLOCK 256
#define GB_WAVES_PER_BLOCK (GB_BLOCK / GB_WAVE)

// Record a HIP failure and map it to GOOSEBERRY_ERROR_DEVICE.
int gb_hip_fail(hipError_t e, const char *what, const char *file, int line);
// Record a refusal's reason as the last error text (stderr too).
void gb_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
#define GB_HIP(call)                                                        \
  do {                                                                      \
    hipError_t gb_e_ = (call);                                              \
    if (gb_e_ != hipSuccess)                                                \
      return gb_hip_fail(gb_e_, #call, __FILE__, __LINE__);                 \
  } while (0)
// the last launch's error, or an injected launch failure (test hook)
hipError_t gb_launch_status();
#define GB_FAULT_ALLOC 0
#define GB_FAULT_LAUNCH 1
#define GB_FAULT_KINDS 2
#define GB_LAUNCH_CHECK() GB_HIP(gb_launch_status())

hipStream_t gb_stream();
int gb_ensure_init();

// Device memory from the device's pool, ordered on the library stream: every
// use of library memory is on that stream, so a block freed here is reused
// only after the work queued before the free, and freed blocks stay in the
// pool (no device-wide synchronisation, no trip to the driver per request).
hipError_t gb_dev_alloc(void **p, size_t bytes);
hipError_t gb_dev_free(void *p);
bool gb_graphs_enabled();  // false when GOOSEBERRY_NO_GRAPHS is set
uint32_t gb_num_cus();     // compute units of the device

static inline uint32_t gb_div_up(uint64_t a, uint64_t b) {
  return (uint32_t)((a + b - 1) / b);
}

// Butterfly sum over the 64 lanes of a wavefront; every lane gets the same
// value and the combine order is fixed, so the result is deterministic.
template <typename T>
__device__ __forceinline__ T gb_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, GB_WAVE);
  return v;
}

// Sum over a 256-thread block in a fixed order: wave butterflies, then the
// four wave totals added in wave order by thread 0.  Valid on thread 0.
template <typename T>
__device__ __forceinline__ T gb_block_sum(T v, T *lds4) {
  v = gb_wave_sum(v);
  const int lane = threadIdx.x & (GB_WAVE - 1);
  const int wave = threadIdx.x / GB_WAVE;
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    s = lds4[0];
#pragma unroll
    for (int w = 1; w < GB_WAVES_PER_BLOCK; ++w) s += lds4[w];
  }
  return s;
}

// Reads a device flag once per workgroup (thread 0) and hands the same
// value to every thread, so a flag that another workgroup of the same
// launch may be writing still gives one decision per workgroup.
__device__ __forceinline__ int gb_block_flag(const int32_t *flag) {
  __shared__ int32_t v;
  if (threadIdx.x == 0) v = *(const volatile int32_t *)flag;
  __syncthreads();
  int32_t r = v;
  __syncthreads();
  return r;
}

// Value of v in lane j (j uniform across the wavefront).
__device__ __forceinline__ double gb_readlane(double v, uint32_t j) {
  union { double d; int i[2]; } u;
  u.d = v;
  u.i[0] = __builtin_amdgcn_readlane(u.i[0], j);
  u.i[1] = __builtin_amdgcn_readlane(u.i[1], j);
  return u.d;
}
__device__ __forceinline__ float gb_readlane(float v, uint32_t j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Inclusive scan over a wavefront.
__device__ __forceinline__ uint32_t gb_wave_scan_incl(uint32_t v) {
  const int lane = threadIdx.x & (GB_WAVE - 1);
#pragma unroll
  for (int off = 1; off < GB_WAVE; off <<= 1) {
    uint32_t y = __shfl_up(v, off, GB_WAVE);
    if (lane >= off) v += y;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the thread's
// exclusive prefix, *total receives the block total.
__device__ __forceinline__ uint32_t gb_block_scan_excl(uint32_t v,
                                                       uint32_t *lds4,
                                                       uint32_t *total) {
  const int lane = threadIdx.x & (GB_WAVE - 1);
  const int wave = threadIdx.x / GB_WAVE;
  uint32_t inc = gb_wave_scan_incl(v);
  if (lane == GB_WAVE - 1) lds4[wave] = inc;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < GB_WAVES_PER_BLOCK; ++w) {
    uint32_t t = lds4[w];
    if (w < wave) woff += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Map b to a logical index so that each XCD walks one
// contiguous range of [0, n): neighbouring row blocks then share one XCD's
// L2 for the x gather.  A bijection on [0, n) for any n.
__device__ __forceinline__ uint32_t gb_xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t nx = 8;
  if (n < 2 * nx) return b;
  const uint32_t q = n / nx, r = n % nx;
  const uint32_t x = b % nx, local = b / nx;
  return x * q + (x < r ? x : r) + local;
}

// Chunked variant: chunks of C consecutive indices go round-robin to the
// XCDs, each XCD walking its chunks in order.  Neighbouring indices share an
// XCD (and its L2) within a chunk, while the eight XCDs still work on
// nearby addresses.  A bijection on [0, n): the tail past the last whole
// round of 8 chunks keeps its order.
__device__ __forceinline__ uint32_t gb_xcd_chunk_remap(uint32_t b, uint32_t n, uint32_t C) {
  const uint32_t nx = 8;
  const uint32_t full = (n / (nx * C)) * (nx * C);
  if (b >= full) return b;
  const uint32_t x = b % nx, l = b / nx;
  return ((l / C) * nx + x) * C + l % C;
}

#endif
