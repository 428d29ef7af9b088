// gb_runtime.hip -- device selection, the library stream, memory, events.
#include <pthread.h>
#include <sched.h>
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
#include "gb_stage.h"

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

// The CPUs the process could run on when the library was loaded.  Threads
// inherit their creator's CPU binding: a caller bound to fewer CPUs than
// there are staging threads (the initial thread of an OpenMP program run
// with OMP_PROC_BIND is bound to one core) would put every staging thread
// on its own core(s), and so would the HIP runtime's threads, which start
// with the first HIP call.  The library's set-up therefore runs on these
// CPUs (the caller's binding is restored afterwards), and a bound caller's
// staging threads run on them too.
static cpu_set_t g_load_cpus;
static bool g_load_cpus_ok = false;

__attribute__((constructor)) static void gb_record_load_cpus() {
  g_load_cpus_ok = sched_getaffinity(0, sizeof(g_load_cpus), &g_load_cpus) == 0;
}

static void gb_do_init_on_device() {
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

static void gb_do_init() {
  cpu_set_t caller;
  const bool widen = g_load_cpus_ok && sched_getaffinity(0, sizeof(caller), &caller) == 0 &&
                     CPU_COUNT(&caller) < CPU_COUNT(&g_load_cpus) &&
                     sched_setaffinity(0, sizeof(g_load_cpus), &g_load_cpus) == 0;
  gb_do_init_on_device();
  if (widen) (void)sched_setaffinity(0, sizeof(caller), &caller);
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

// ----------------------------------------------------- staged host copies
// Large copies between pageable host memory and the device run through two
// pinned staging buffers: host threads fill (or drain) one buffer while the
// DMA engine copies the other.  The runtime's own path for pageable memory
// stages through one thread.  The buffers are kept from the first staged
// copy to the end of the process: a freshly pinned page costs a page fault
// on its first write and a mapping on its first copy to the device, several
// times the copy itself.  Staged copies hold g_stage_mu, so copies from
// several host threads take turns.
#ifndef GB_STAGE_MAX
#define GB_STAGE_MAX (16ull << 20)  // bytes per staging buffer
#endif
#ifndef GB_STAGE_THREADS
#define GB_STAGE_THREADS 4          // host threads filling or draining one buffer
#endif
// smaller copies go straight to the runtime, whose own copies from pageable
// memory beat staging below about 12 MiB to the device (33-36 GB/s against
// 28-35) but not to the host (16-18 GB/s against 24-34)
#ifndef GB_STAGE_MIN_HTOD
#define GB_STAGE_MIN_HTOD (16ull << 20)
#endif
#ifndef GB_STAGE_MIN_DTOH
#define GB_STAGE_MIN_DTOH (4ull << 20)
#endif
#define GB_STAGE_PART (1ull << 20)  // a buffer's share of one thread, at least

static pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;
static char *g_stage[2];
static hipEvent_t g_stage_ev[2];
static uint64_t g_stage_bytes;

struct gb_stage_job {
  gb_stage_fn fn;
  void *ctx;
  uint64_t off, len;
  char *buf;
  int rv;
};

static void *gb_stage_job_run(void *p) {
  gb_stage_job *j = (gb_stage_job *)p;
  j->rv = j->fn(j->ctx, j->off, j->len, j->buf);
  return nullptr;
}

// fn over bytes [off, off + len) and buf[0, len), split at 4 KiB boundaries
// over up to GB_STAGE_THREADS threads (a part whose thread cannot be started
// runs on this one)
static int gb_stage_parallel(gb_stage_fn fn, void *ctx, uint64_t off, uint64_t len, char *buf) {
  gb_stage_job job[GB_STAGE_THREADS];
  pthread_t th[GB_STAGE_THREADS];
  bool started[GB_STAGE_THREADS] = {};
  int nt = (int)(len / GB_STAGE_PART);
  if (nt < 1) nt = 1;
  if (nt > GB_STAGE_THREADS) nt = GB_STAGE_THREADS;
  for (int k = 0; k < nt; ++k) {
    const uint64_t a = k ? (len * k / nt) & ~4095ull : 0;
    const uint64_t b = k + 1 < nt ? (len * (k + 1) / nt) & ~4095ull : len;
    job[k] = {fn, ctx, off + a, b - a, buf + a, 0};
  }
  // A caller bound to fewer CPUs than there are parts: each part gets a CPU
  // of its own from those the process had at load time and the caller is
  // not bound to, the highest-numbered first.  Threads that merely may run
  // anywhere start on their creator's CPU and, living about a millisecond,
  // mostly finish there.
  int spread_cpu[GB_STAGE_THREADS];
  bool spread = false;
  if (nt > 1 && g_load_cpus_ok) {
    cpu_set_t cur;
    if (sched_getaffinity(0, sizeof(cur), &cur) == 0 && CPU_COUNT(&cur) < nt) {
      int found = 0;
      for (int c = CPU_SETSIZE - 1; c >= 0 && found < nt; --c)
        if (CPU_ISSET(c, &g_load_cpus) && !CPU_ISSET(c, &cur)) spread_cpu[found++] = c;
      for (int k = found; k < nt && found > 0; ++k) spread_cpu[k] = spread_cpu[k % found];
      spread = found > 0;
    }
  }
  // a bound caller's own CPU takes no part
  const int k0 = spread ? 0 : 1;
  for (int k = k0; k < nt; ++k) {
    pthread_attr_t attr;
    bool own = false;
    if (spread && pthread_attr_init(&attr) == 0) {
      cpu_set_t one;
      CPU_ZERO(&one);
      CPU_SET(spread_cpu[k], &one);
      own = true;
      if (pthread_attr_setaffinity_np(&attr, sizeof(one), &one) != 0) {
        pthread_attr_destroy(&attr);
        own = false;
      }
    }
    started[k] = pthread_create(&th[k], own ? &attr : nullptr, gb_stage_job_run, &job[k]) == 0;
    if (own) pthread_attr_destroy(&attr);
  }
  int rv = GOOSEBERRY_SUCCESS;
  for (int k = 0; k < nt; ++k) {
    if (started[k]) pthread_join(th[k], nullptr);
    else gb_stage_job_run(&job[k]);
    if (job[k].rv) rv = job[k].rv;
  }
  return rv;
}

// staging buffers of at least `want` bytes (g_stage_mu held, no copy from
// or to them in flight)
static int gb_stage_reserve(uint64_t want) {
  hipError_t he = hipSuccess;
  for (int j = 0; j < 2 && he == hipSuccess; ++j)
    if (!g_stage_ev[j]) he = hipEventCreateWithFlags(&g_stage_ev[j], hipEventDisableTiming);
  if (he != hipSuccess) return gb_hip_fail(he, "hipEventCreateWithFlags", __FILE__, __LINE__);
  if (g_stage_bytes >= want) return GOOSEBERRY_SUCCESS;
  for (int j = 0; j < 2; ++j) {
    if (g_stage[j]) (void)hipHostFree(g_stage[j]);
    g_stage[j] = nullptr;
  }
  g_stage_bytes = 0;
  for (int j = 0; j < 2 && he == hipSuccess; ++j)
    he = hipHostMalloc((void **)&g_stage[j], want, hipHostMallocDefault);
  if (he != hipSuccess) {
    for (int j = 0; j < 2; ++j) {
      if (g_stage[j]) (void)hipHostFree(g_stage[j]);
      g_stage[j] = nullptr;
    }
    return gb_hip_fail(he, "hipHostMalloc (staging)", __FILE__, __LINE__);
  }
  g_stage_bytes = want;
  return GOOSEBERRY_SUCCESS;
}

// chunk size of a staged copy: four chunks or more where that keeps them at
// 2 MiB or more, so that filling and copying overlap for most of the copy
static uint64_t gb_stage_chunk(uint64_t bytes) {
  uint64_t c = (bytes / 4 + (1ull << 20) - 1) & ~((1ull << 20) - 1);
  if (c < (2ull << 20)) c = 2ull << 20;
  return c < GB_STAGE_MAX ? c : GB_STAGE_MAX;
}

// One pipeline over several copies to the device: the buffers keep filling
// from one copy to the next, so only the first chunk's fill and the last
// chunk's DMA go unoverlapped, once for all of them.
struct gb_stage_seg {
  char *d;
  gb_stage_fn fill;
  void *ctx;
  uint64_t bytes;
};

static int gb_stage_htod_segs(const gb_stage_seg *seg, uint32_t nseg) {
  uint64_t total = 0;
  for (uint32_t k = 0; k < nseg; ++k) total += seg[k].bytes;
  if (total == 0) return GOOSEBERRY_SUCCESS;
  hipStream_t s = g_stream;
  bool used[2] = {false, false};
  int i = 0;
  hipError_t he = hipSuccess;
  pthread_mutex_lock(&g_stage_mu);
  int rv = gb_stage_reserve(GB_STAGE_MAX);
  const uint64_t stage = gb_stage_chunk(total);
  for (uint32_t k = 0; k < nseg && rv == GOOSEBERRY_SUCCESS; ++k) {
    for (uint64_t copied = 0; copied < seg[k].bytes;) {
      const uint64_t want = seg[k].bytes - copied < stage ? seg[k].bytes - copied : stage;
      if (used[i]) {  // the copy two chunks back has left this buffer
        he = hipEventSynchronize(g_stage_ev[i]);
        if (he != hipSuccess) {
          rv = gb_hip_fail(he, "hipEventSynchronize", __FILE__, __LINE__);
          break;
        }
      }
      rv = gb_stage_parallel(seg[k].fill, seg[k].ctx, copied, want, g_stage[i]);
      if (rv != GOOSEBERRY_SUCCESS) break;
      he = hipMemcpyAsync(seg[k].d + copied, g_stage[i], want, hipMemcpyHostToDevice, s);
      if (he == hipSuccess) he = hipEventRecord(g_stage_ev[i], s);
      if (he != hipSuccess) {
        rv = gb_hip_fail(he, "hipMemcpyAsync (staged)", __FILE__, __LINE__);
        break;
      }
      used[i] = true;
      copied += want;
      i ^= 1;
    }
  }
  // the staging buffers are idle again before another copy may use them
  he = hipStreamSynchronize(s);
  if (he != hipSuccess && rv == GOOSEBERRY_SUCCESS)
    rv = gb_hip_fail(he, "hipStreamSynchronize", __FILE__, __LINE__);
  pthread_mutex_unlock(&g_stage_mu);
  return rv;
}

int gb_stage_htod(char *d_dst, uint64_t bytes, gb_stage_fn fill, void *ctx) {
  const gb_stage_seg seg = {d_dst, fill, ctx, bytes};
  return gb_stage_htod_segs(&seg, 1);
}

int gb_stage_dtoh(uint64_t bytes, const char *d_src, gb_stage_fn drain, void *ctx) {
  if (bytes == 0) return GOOSEBERRY_SUCCESS;
  hipStream_t s = g_stream;
  hipError_t he = hipSuccess;
  pthread_mutex_lock(&g_stage_mu);
  int rv = gb_stage_reserve(GB_STAGE_MAX);
  const uint64_t stage = gb_stage_chunk(bytes);
  const uint64_t nchunks = rv == GOOSEBERRY_SUCCESS ? (bytes + stage - 1) / stage : 0;
  // chunk c goes through buffer c & 1; chunk c + 1's copy is queued before
  // chunk c is drained, so the DMA engine and the host threads overlap
  for (uint64_t c = 0; c <= nchunks && rv == GOOSEBERRY_SUCCESS; ++c) {
    if (c < nchunks) {  // buffer c & 1 was drained in step c - 1
      const uint64_t o = c * stage;
      const uint64_t want = bytes - o < stage ? bytes - o : stage;
      he = hipMemcpyAsync(g_stage[c & 1], d_src + o, want, hipMemcpyDeviceToHost, s);
      if (he == hipSuccess) he = hipEventRecord(g_stage_ev[c & 1], s);
      if (he != hipSuccess) {
        rv = gb_hip_fail(he, "hipMemcpyAsync (staged)", __FILE__, __LINE__);
        break;
      }
    }
    if (c == 0) continue;
    const uint64_t p = c - 1, o = p * stage;
    const uint64_t want = bytes - o < stage ? bytes - o : stage;
    he = hipEventSynchronize(g_stage_ev[p & 1]);
    if (he != hipSuccess) {
      rv = gb_hip_fail(he, "hipEventSynchronize", __FILE__, __LINE__);
      break;
    }
    rv = gb_stage_parallel(drain, ctx, o, want, g_stage[p & 1]);
  }
  he = hipStreamSynchronize(s);
  if (he != hipSuccess && rv == GOOSEBERRY_SUCCESS)
    rv = gb_hip_fail(he, "hipStreamSynchronize", __FILE__, __LINE__);
  pthread_mutex_unlock(&g_stage_mu);
  return rv;
}

static int gb_stage_from_mem(void *ctx, uint64_t off, uint64_t len, char *buf) {
  memcpy(buf, (const char *)ctx + off, len);
  return GOOSEBERRY_SUCCESS;
}

static int gb_stage_to_mem(void *ctx, uint64_t off, uint64_t len, char *buf) {
  memcpy((char *)ctx + off, buf, len);
  return GOOSEBERRY_SUCCESS;
}

// whether h is host memory the runtime has pinned (then the DMA engine
// reads or writes it directly and staging would only add a copy)
static bool gb_host_pinned(const void *h) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, h) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory: an error, not sticky
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

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
  if (bytes >= GB_STAGE_MIN_HTOD && !gb_host_pinned(h_src))
    return gb_stage_htod((char *)d_dst, bytes, gb_stage_from_mem, (void *)h_src);
  GB_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, g_stream));
  GB_HIP(hipStreamSynchronize(g_stream));
  return GOOSEBERRY_SUCCESS;
}

int gbd_htod_many(uint32_t n, void *const *d_dst, const void *const *h_src,
                  const size_t *bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  uint64_t total = 0;
  bool pinned = false;
  for (uint32_t k = 0; k < n; ++k) {
    total += bytes[k];
    if (bytes[k] && gb_host_pinned(h_src[k])) pinned = true;
  }
  if (pinned || n > 16) {
    for (uint32_t k = 0; k < n && rv == GOOSEBERRY_SUCCESS; ++k)
      rv = gbd_htod(d_dst[k], h_src[k], bytes[k]);
    return rv;
  }
  if (total < GB_STAGE_MIN_HTOD) {
    // the runtime's copies, one synchronisation (also after a failed one:
    // no copy from the caller's arrays is left in flight)
    hipError_t he = hipSuccess;
    for (uint32_t k = 0; k < n && he == hipSuccess; ++k)
      if (bytes[k])
        he = hipMemcpyAsync(d_dst[k], h_src[k], bytes[k], hipMemcpyHostToDevice, g_stream);
    const hipError_t hs = hipStreamSynchronize(g_stream);
    if (he != hipSuccess) return gb_hip_fail(he, "hipMemcpyAsync", __FILE__, __LINE__);
    if (hs != hipSuccess) return gb_hip_fail(hs, "hipStreamSynchronize", __FILE__, __LINE__);
    return GOOSEBERRY_SUCCESS;
  }
  gb_stage_seg seg[16];
  for (uint32_t k = 0; k < n; ++k)
    seg[k] = {(char *)d_dst[k], gb_stage_from_mem, (void *)h_src[k], bytes[k]};
  return gb_stage_htod_segs(seg, n);
}

int gbd_dtoh(void *h_dst, const void *d_src, size_t bytes) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!bytes) return GOOSEBERRY_SUCCESS;
  if (bytes >= GB_STAGE_MIN_DTOH && !gb_host_pinned(h_dst))
    return gb_stage_dtoh(bytes, (const char *)d_src, gb_stage_to_mem, h_dst);
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
