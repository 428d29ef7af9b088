// gb_runtime.hip -- device selection, the library stream, memory, events.
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <list>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>

#include "gb_common.h"

static hipStream_t g_stream = nullptr;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_init_rv = GOOSEBERRY_SUCCESS;
static char g_last_error[512] = "";
static size_t g_cache_cap = 0;  // bytes the large-block cache may hold
static bool g_graphs = true;     // replay captured step graphs

int gb_hip_fail(hipError_t e, const char *what, const char *file, int line) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%s:%d)",
           hipGetErrorString(e), what, file, line);
  fprintf(stderr, "gooseberry: HIP error %s\n", g_last_error);
  if (e == hipErrorOutOfMemory) return GOOSEBERRY_ERROR_NOMEMORY;
  return GOOSEBERRY_ERROR_DEVICE;
}

void gb_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  fprintf(stderr, "gooseberry: %s\n", g_last_error);
}

static void gb_do_init() {
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    g_init_rv = gb_hip_fail(e != hipSuccess ? e : hipErrorNoDevice,
                            "hipGetDeviceCount (no GPU visible)", __FILE__,
                            __LINE__);
    return;
  }
  e = hipSetDevice(0);
  if (e != hipSuccess) {
    g_init_rv = gb_hip_fail(e, "hipSetDevice(0)", __FILE__, __LINE__);
    return;
  }
  // default priority, non-blocking with respect to the null stream
  e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_init_rv = gb_hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    return;
  }
  // keep freed blocks in the pool instead of returning them at every
  // synchronisation
  hipMemPool_t pool;
  e = hipDeviceGetDefaultMemPool(&pool, 0);
  if (e == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  if (e != hipSuccess) {
    g_init_rv = gb_hip_fail(e, "default memory pool", __FILE__, __LINE__);
    return;
  }
  // the block cache holds at most an eighth of the device, and 32 GiB
  size_t fr = 0, tot = 0;
  e = hipMemGetInfo(&fr, &tot);
  if (e != hipSuccess) {
    g_init_rv = gb_hip_fail(e, "hipMemGetInfo", __FILE__, __LINE__);
    return;
  }
  g_cache_cap = tot / 8 < (32ull << 30) ? tot / 8 : (32ull << 30);
  // GOOSEBERRY_NO_GRAPHS (set, not "0"): launch CG and batch steps one by
  // one instead of replaying captured graphs, for tracing tools that do not
  // follow graph launches
  const char *ng = getenv("GOOSEBERRY_NO_GRAPHS");
  g_graphs = !(ng && *ng && strcmp(ng, "0") != 0);
}

bool gb_graphs_enabled() { return g_graphs; }

static uint32_t g_num_cus = 0;

uint32_t gb_num_cus() {
  if (!g_num_cus) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        v <= 0)
      v = 256;
    g_num_cus = (uint32_t)v;
  }
  return g_num_cus;
}

// Large blocks (4 MiB and up, sizes rounded to 2 MiB) that the library
// frees are kept here and handed back to a later allocation of the same
// rounded size, the most recently freed first.  The default pool gave such blocks back to the driver at
// free despite its release threshold: a free of a few hundred MiB took
// 0.65-2.5 ms and the next allocation of that size 50-350 us, several
// milliseconds per sparse product.  Every library launch and copy runs on
// g_stream, so a block's reuse is ordered after its last use exactly as a
// pool block's is.  The cache holds at most g_cache_cap bytes (oldest
// blocks go back first), is emptied when an allocation finds the device
// out of memory, and its bytes count as free in gbd_device_memory.
#define GB_CACHE_MIN (4ull << 20)
#define GB_CACHE_ROUND (2ull << 20)

static std::mutex g_cache_mu;
static std::unordered_map<void *, size_t> g_cache_live;  // handed out: size
static std::list<std::pair<size_t, void *>> g_cache_fifo;  // free, oldest first
static std::multimap<size_t, std::list<std::pair<size_t, void *>>::iterator> g_cache_free;
static size_t g_cache_bytes = 0;

static void gb_cache_evict_oldest() {  // g_cache_mu held
  auto &front = g_cache_fifo.front();
  auto range = g_cache_free.equal_range(front.first);
  for (auto it = range.first; it != range.second; ++it)
    if (it->second == g_cache_fifo.begin()) {
      g_cache_free.erase(it);
      break;
    }
  g_cache_bytes -= front.first;
  (void)hipFreeAsync(front.second, g_stream);
  g_cache_fifo.pop_front();
}

static void gb_cache_trim() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  while (!g_cache_fifo.empty()) gb_cache_evict_oldest();
}

size_t gb_cache_cached_bytes() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  return g_cache_bytes;
}

// Injected faults (gbd_inject_fault, a test hook): with a countdown c >= 0
// armed for a kind, the c+1-th device allocation (or launch check) from then
// on reports a failure instead, once; c < 0 disarms.  An injected launch
// failure is reported after the kernel was launched normally.
static std::atomic<long long> g_fault[GB_FAULT_KINDS] = {{-1}, {-1}};

static bool gb_fault_fires(int kind) {
  long long v = g_fault[kind].load();
  while (v >= 0)
    if (g_fault[kind].compare_exchange_weak(v, v - 1)) return v == 0;
  return false;
}

hipError_t gb_launch_status() {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && gb_fault_fires(GB_FAULT_LAUNCH)) e = hipErrorLaunchFailure;
  return e;
}

hipError_t gb_dev_alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (gb_fault_fires(GB_FAULT_ALLOC)) return hipErrorOutOfMemory;
  if (bytes < GB_CACHE_MIN) return hipMallocAsync(p, bytes ? bytes : 16, g_stream);
  const size_t r = (bytes + GB_CACHE_ROUND - 1) & ~(GB_CACHE_ROUND - 1);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    // the block of this size freed last (its lines are the likeliest to
    // be in the caches still)
    auto range = g_cache_free.equal_range(r);
    if (range.first != range.second) {
      auto it = std::prev(range.second);
      *p = it->second->second;
      g_cache_fifo.erase(it->second);
      g_cache_free.erase(it);
      g_cache_bytes -= r;
      g_cache_live[*p] = r;
      return hipSuccess;
    }
  }
  hipError_t e = hipMallocAsync(p, r, g_stream);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    gb_cache_trim();
    e = hipMallocAsync(p, r, g_stream);
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache_live[*p] = r;
  }
  return e;
}

hipError_t gb_dev_free(void *p) {
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_cache_live.find(p);
    if (it != g_cache_live.end()) {
      const size_t r = it->second;
      g_cache_live.erase(it);
      if (r <= g_cache_cap) {
        while (g_cache_bytes + r > g_cache_cap) gb_cache_evict_oldest();
        g_cache_fifo.emplace_back(r, p);
        g_cache_free.emplace(r, std::prev(g_cache_fifo.end()));
        g_cache_bytes += r;
        return hipSuccess;
      }
    }
  }
  return hipFreeAsync(p, g_stream);
}

int gb_ensure_init() {
  pthread_once(&g_once, gb_do_init);
  return g_init_rv;
}

hipStream_t gb_stream() { return g_stream; }

__global__ void gb_fill_zero_kernel(uint32_t *p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

__global__ void gb_fill_zero_bytes_kernel(uint8_t *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void gb_scale_kernel(double *__restrict__ d, size_t n, double s) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] *= s;
}

__global__ void gb_add_scalar_kernel(double *__restrict__ d, size_t n, double s) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] += s;
}

extern "C" {

int gbd_add_scalar(double *d, size_t n, double s) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!n) return GOOSEBERRY_SUCCESS;
  uint32_t grid = gb_div_up(n, GB_BLOCK);
  if (grid > 65536u) grid = 65536u;
  gb_add_scalar_kernel<<<grid, GB_BLOCK, 0, g_stream>>>(d, n, s);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_scale(double *d, size_t n, double s) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!n) return GOOSEBERRY_SUCCESS;
  uint32_t grid = gb_div_up(n, GB_BLOCK);
  if (grid > 65536u) grid = 65536u;
  gb_scale_kernel<<<grid, GB_BLOCK, 0, g_stream>>>(d, n, s);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_init(void) { return gb_ensure_init(); }

const char *gbd_last_error(void) { return g_last_error; }

void gbd_set_error(const char *msg) { gb_set_error("%s", msg); }

int gbd_inject_fault(int kind, long long after) {
  if (kind < 0 || kind >= GB_FAULT_KINDS) return GOOSEBERRY_ERROR_INVALIDINPUT;
  g_fault[kind].store(after < 0 ? -1 : after);
  return GOOSEBERRY_SUCCESS;
}

void *gbd_stream(void) {
  if (gb_ensure_init() != GOOSEBERRY_SUCCESS) return nullptr;
  return (void *)g_stream;
}

int gbd_malloc(void **d_ptr, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  *d_ptr = nullptr;
  GB_HIP(gb_dev_alloc(d_ptr, bytes ? bytes : 16));
  return GOOSEBERRY_SUCCESS;
}

int gbd_free(void *d_ptr) {
  if (!d_ptr) return GOOSEBERRY_SUCCESS;
  GB_HIP(gb_dev_free(d_ptr));
  return GOOSEBERRY_SUCCESS;
}

int gbd_host_alloc(void **h_ptr, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  GB_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 16, hipHostMallocDefault));
  return GOOSEBERRY_SUCCESS;
}

int gbd_host_free(void *h_ptr) {
  if (!h_ptr) return GOOSEBERRY_SUCCESS;
  GB_HIP(hipHostFree(h_ptr));
  return GOOSEBERRY_SUCCESS;
}

int gbd_htod(void *d_dst, const void *h_src, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!bytes) return GOOSEBERRY_SUCCESS;
  GB_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, g_stream));
  GB_HIP(hipStreamSynchronize(g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_dtoh(void *h_dst, const void *d_src, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!bytes) return GOOSEBERRY_SUCCESS;
  GB_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, g_stream));
  GB_HIP(hipStreamSynchronize(g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_dtod(void *d_dst, const void *d_src, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!bytes) return GOOSEBERRY_SUCCESS;
  GB_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_zero(void *d_dst, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!bytes) return GOOSEBERRY_SUCCESS;
  if (((uintptr_t)d_dst & 3u) == 0 && (bytes & 3u) == 0) {
    size_t nw = bytes / 4;
    uint32_t grid = gb_div_up(nw, GB_BLOCK);
    if (grid > 65536u) grid = 65536u;
    gb_fill_zero_kernel<<<grid, GB_BLOCK, 0, g_stream>>>((uint32_t *)d_dst, nw);
  } else {
    gb_fill_zero_bytes_kernel<<<gb_div_up(bytes, GB_BLOCK), GB_BLOCK, 0,
                                g_stream>>>((uint8_t *)d_dst, bytes);
  }
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_sync(void) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  GB_HIP(hipStreamSynchronize(g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_device_memory(size_t *free_bytes, size_t *total_bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  size_t f = 0, t = 0;
  GB_HIP(hipMemGetInfo(&f, &t));
  f += gb_cache_cached_bytes();  // cached blocks are given back on demand
  if (f > t) f = t;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return GOOSEBERRY_SUCCESS;
}

int gbd_event_create(void **ev) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipEvent_t e;
  GB_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return GOOSEBERRY_SUCCESS;
}

int gbd_event_destroy(void *ev) {
  if (!ev) return GOOSEBERRY_SUCCESS;
  GB_HIP(hipEventDestroy((hipEvent_t)ev));
  return GOOSEBERRY_SUCCESS;
}

int gbd_event_record(void *ev) {
  GB_HIP(hipEventRecord((hipEvent_t)ev, g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_event_elapsed_ms(void *start, void *stop, float *ms) {
  GB_HIP(hipEventSynchronize((hipEvent_t)stop));
  GB_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return GOOSEBERRY_SUCCESS;
}

}  // extern "C"
