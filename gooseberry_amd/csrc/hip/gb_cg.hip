// gb_cg.hip -- device side of the conjugate-gradient solver.
//
// Reference: src/cgd.c:40-129.  One iteration is three kernels on the
// library stream, all scalars kept on the device (no host round trip):
//   K1 gb_spmv_kernel<DOT>: ap = A p and the p.ap partials
//   K2 gb_cg_update_r:      alpha = dr / p.ap; r -= alpha ap; r.r partials
//   K3 gb_cg_update_xp:     beta = r.r / dr, convergence test;
//                           ps = alpha p; x += ps; p = r + (beta/alpha) ps
// which is the reference's scale-then-add order (src/cgd.c:92-105) with
// the same separately rounded products and sums (-ffp-contract=off); only
// the order of the dot-product sums differs.  x += ps sits in K3, which
// reads p anyway (one vector read less than in K2; x does not feed r).
// K2 and K3 give each workgroup one contiguous chunk of rows; each
// workgroup adds the previous kernel's partials itself (the same values in
// the same order in every workgroup, so every workgroup gets the same
// scalar) while the loads of its chunk are in flight.
// K2 and K3 read the device `done` flag and return when the solve has
// stopped, so the host can enqueue iterations in batches (or replay them
// from a hipGraph) and still stop exactly where the reference stops.
// `done` is written only by the state's set-up and by K2's workgroup 0: K3
// records its stop in `status`, which every K3 workgroup computes for
// itself, so all of them still add their share of x; the next K2 turns a
// nonzero status into `done`.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gb_matrix.h"

#define GB_ELEM_GRID_MAX 2048u
#define GB_CG_NGRAPH 7  // captured batches of 1, 2, 4, ..., 64 iterations

struct gbd_cg {
  const gbd_matrix *m;
  uint32_t n;
  double *r, *p, *ap;
  double *partials;   // SpMV p.Ap partials (one per SpMV workgroup)
  double *partials2;  // r.r partials (one per elementwise workgroup)
  uint32_t npartials;
  gbd_cg_state *st;
  gbd_cg_state *h_st;  // pinned
  double *trace;
  uint32_t trace_cap;
  double *h_trace;     // pinned, trace_cap entries
  const double *b;
  double *x;
  uint32_t grid_elem;  // workgroups of the set-up kernel
  uint32_t grid_vec;   // workgroups of the r and x/p updates
  uint32_t chunk;      // rows per workgroup of the r and x/p updates
  hipGraph_t graph[GB_CG_NGRAPH];       // graph[i]: 2^i iterations
  hipGraphExec_t gexec[GB_CG_NGRAPH];
  double *graph_x;                      // x the graphs were captured with
  bool noalloc;  // the workspace in `scratch`: single launches, no graphs
};

// Fixed-order sum of n partials, computed identically by every workgroup
// that calls it and returned to all of its threads.
__device__ __forceinline__ double gb_sum_partials(const double *partials,
                                                  uint32_t n, double *red) {
  // all loads of a chunk issued before the adds: one memory round trip per
  // 2048 partials instead of one per 256
  double s = 0.0;
  for (uint32_t base = 0; base < n; base += GB_BLOCK * 8) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      uint32_t i = base + j * GB_BLOCK + threadIdx.x;
      v[j] = i < n ? partials[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  s = gb_wave_sum(s);
  const int lane = threadIdx.x & (GB_WAVE - 1);
  const int wave = threadIdx.x / GB_WAVE;
  if (lane == 0) red[wave] = s;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < GB_WAVES_PER_BLOCK; ++w) t += red[w];
  __syncthreads();
  return t;
}

// src/cgd.c:81-85: x = 0; r = b; p = r; partial sums of r.r
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_init(const double *__restrict__ b, double *__restrict__ x,
           double *__restrict__ r, double *__restrict__ p, uint32_t n,
           double *__restrict__ partials) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  double d = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double v = b[i];
    x[i] = 0.0;
    r[i] = v;
    p[i] = v;
    d += v * v;
  }
  double bs = gb_block_sum(d, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// dr = r.r, ||b||, the absolute threshold; one workgroup
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_init_state(const double *__restrict__ partials, uint32_t np,
                 gbd_cg_state *st, double tol, int rel,
                 unsigned long long maxiter) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  double s = gb_sum_partials(partials, np, red);
  if (threadIdx.x == 0) {
    st->dr = s;
    st->dr_prev = s;
    st->rnorm0 = sqrt(s);
    st->rnorm = st->rnorm0;
    st->tol = rel ? tol * st->rnorm0 : tol;
    st->pap = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->iter = 0;
    st->iter_next = 0;
    st->maxiter = maxiter;
    st->status = 0;
    st->done = 0;
    if (s == 0.0) {  // b == 0: x = 0 is the solution
      st->done = 1;
      st->status = 1;
    } else if (!isfinite(s)) {
      st->done = 1;
      st->status = 3;
    }
  }
}

// src/cgd.c:90, 96-97: alpha = dr / p.Ap; r -= alpha Ap; partial sums of
// r.r.  Each workgroup owns one contiguous chunk of rows and first issues
// the loads of its r and Ap values, then adds the SpMV's p.Ap partials
// itself -- every workgroup the same values in the same fixed order, so
// all of them get the same alpha and no separate reduction launch is
// needed -- while those loads are in flight.  Workgroup 0 records alpha
// and turns a stop recorded by the previous iteration's x/p update into
// `done`.
// Launch shape of K2 and K3: GB_CGV_THREADS threads, each with GB_CGV_U
// rows per pass, at most one workgroup per CU.  On the 160^3 Laplacian
// (n = 4,096,000) every workgroup then covers its 16,000 rows in one pass,
// with all of its loads issued at once.
#define GB_CGV_THREADS 512
#define GB_CGV_GRID 256u
#define GB_CGV_U 32
// loads per thread of the partial sums: K2 adds up to 16K SpMV partials in
// one round, K3 the r update's (at most GB_CGV_GRID)
#define GB_CGV_RED1 32
#define GB_CGV_RED2 1

// Fixed-order sum of n partials by a workgroup of T threads: each thread
// issues R loads before adding (one memory round trip per T R partials);
// the result is returned to every thread.
template <int T, int R>
__device__ __forceinline__ double gb_cgv_sum(const double *__restrict__ part,
                                             uint32_t n, double *red) {
  double s = 0.0;
  for (uint32_t base = 0; base < n; base += T * R) {
    double v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      uint32_t i = base + j * T + threadIdx.x;
      v[j] = i < n ? part[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) s += v[j];
  }
  s = gb_wave_sum(s);
  if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = s;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < T / GB_WAVE; ++w) t += red[w];
  __syncthreads();
  return t;
}

// Sum of one value per thread over a workgroup of T threads (wave
// butterflies, then the wave totals in wave order); valid on thread 0.
template <int T>
__device__ __forceinline__ double gb_cgv_block_sum(double v, double *red) {
  v = gb_wave_sum(v);
  if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    t = red[0];
#pragma unroll
    for (int w = 1; w < T / GB_WAVE; ++w) t += red[w];
  }
  return t;
}

template <int T, int U>
__global__ void __launch_bounds__(T)
gb_cg_update_r(double *__restrict__ r, const double *__restrict__ ap, uint32_t n,
               uint32_t chunk, const double *__restrict__ pap_partials, uint32_t npap,
               double *__restrict__ partials, gbd_cg_state *st) {
  __shared__ double red[T / GB_WAVE];
  __shared__ int32_t s_stop, s_bad;
  __shared__ double s_alpha;
  const uint32_t tid = threadIdx.x;
  const uint32_t i0 = blockIdx.x * chunk;  // < n: the grid is ceil(n / chunk)
  const uint32_t len = n - i0 < chunk ? n - i0 : chunk;
  double rv[U], av[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const uint32_t o = j * T + tid;
    if (o < len) {
      const uint32_t i = i0 + o;
      rv[j] = r[i];
      av[j] = ap[i];
    }
  }
  if (tid == 0) {
    // a stop recorded by the previous iteration's x/p update becomes `done`
    // (workgroups that read `done` while workgroup 0 sets it see the
    // nonzero status as well)
    const int32_t dn = *(volatile int32_t *)&st->done;
    const int32_t stop = dn || *(volatile int32_t *)&st->status;
    if (stop && !dn && blockIdx.x == 0) st->done = 1;
    s_stop = stop;
  }
  const double pap = gb_cgv_sum<T, GB_CGV_RED1>(pap_partials, npap, red);  // barriers inside
  if (s_stop) return;
  if (tid == 0) {
    const double dr = st->dr;
    const double alpha = dr / pap;
    const bool bad = !(pap != 0.0) || !isfinite(alpha);
    if (blockIdx.x == 0) {
      st->pap = pap;
      st->alpha = alpha;
      if (bad) {
        st->done = 1;
        st->status = 3;
      } else {
        st->dr_prev = dr;
        st->iter_next = st->iter + 1;
      }
    }
    s_alpha = bad ? 0.0 : alpha;
    s_bad = bad;
  }
  __syncthreads();
  if (s_bad) return;
  const double alpha = s_alpha;
  double d = 0.0;
  for (uint32_t done = 0;;) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t o = done + j * T + tid;
      if (o < len) {
        const uint32_t i = i0 + o;
        const double aps = alpha * av[j];
        const double rn = rv[j] - aps;
        r[i] = rn;
        d += rn * rn;
      }
    }
    if (len - done <= U * T) break;
    done += U * T;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t o = done + j * T + tid;
      if (o < len) {
        const uint32_t i = i0 + o;
        rv[j] = r[i];
        av[j] = ap[i];
      }
    }
  }
  const double bs = gb_cgv_block_sum<T>(d, red);
  if (tid == 0) partials[blockIdx.x] = bs;
}

// src/cgd.c:92-94, 100-119: beta = r.r / dr; stop when sqrt(r.r) <= tol or
// at the iteration cap; ps = alpha p; x += ps; unless stopping,
// p = nr + (beta/alpha) ps.  The workgroup's x, p and r values are loaded
// first; the r.r partials of the r update are added (the same order in
// every workgroup) while they are in flight.
template <int T, int U>
__global__ void __launch_bounds__(T)
gb_cg_update_xp(double *__restrict__ x, const double *__restrict__ r,
                double *__restrict__ p, uint32_t n, uint32_t chunk,
                const double *__restrict__ dnr_partials, uint32_t ndnr,
                gbd_cg_state *st, double *__restrict__ trace, uint32_t trace_cap) {
  __shared__ double red[T / GB_WAVE];
  __shared__ int32_t s_done;
  const uint32_t tid = threadIdx.x;
  const uint32_t i0 = blockIdx.x * chunk;  // < n: the grid is ceil(n / chunk)
  const uint32_t len = n - i0 < chunk ? n - i0 : chunk;
  double xv[U], pv[U], rv[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const uint32_t o = j * T + tid;
    if (o < len) {
      const uint32_t i = i0 + o;
      xv[j] = x[i];
      pv[j] = p[i];
      rv[j] = r[i];
    }
  }
  // `done` is written only by the r update and the set-up, never during
  // this kernel
  if (tid == 0) s_done = *(volatile int32_t *)&st->done;
  const double dnr = gb_cgv_sum<T, GB_CGV_RED2>(dnr_partials, ndnr, red);  // barriers inside
  if (s_done) return;
  const double alpha = st->alpha;
  const double beta = dnr / st->dr_prev;
  const double rnorm = sqrt(dnr);
  const unsigned long long it1 = st->iter_next;
  int status = 0;
  if (rnorm <= st->tol) status = 1;
  else if (!isfinite(dnr)) status = 3;
  else if (st->maxiter && it1 >= st->maxiter) status = 2;
  if (blockIdx.x == 0 && tid == 0) {
    st->beta = beta;
    st->dr = dnr;
    st->rnorm = rnorm;
    if (trace_cap) trace[(it1 - 1) % trace_cap] = rnorm;
    st->iter = it1;
    // not `done`: the other workgroups of this launch may not have read it
    if (status) st->status = status;
  }
  const double boa = beta / alpha;
  for (uint32_t done = 0;;) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t o = done + j * T + tid;
      if (o < len) {
        const uint32_t i = i0 + o;
        const double ps = alpha * pv[j];
        x[i] = xv[j] + ps;
        if (!status) {
          const double t = ps * boa;
          p[i] = rv[j] + t;
        }
      }
    }
    if (len - done <= U * T) break;
    done += U * T;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t o = done + j * T + tid;
      if (o < len) {
        const uint32_t i = i0 + o;
        xv[j] = x[i];
        pv[j] = p[i];
        rv[j] = r[i];
      }
    }
  }
}

static int gb_cg_enqueue_raw(gbd_cg *cg, uint32_t k) {
  hipStream_t s = gb_stream();
  const uint32_t n = cg->n;
  const uint32_t g1 = gb_spmv_npartials(cg->m);
  for (uint32_t it = 0; it < k; ++it) {
    int rv = gb_spmv_launch(cg->m, cg->p, cg->ap, 1, cg->partials, cg->st);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    gb_cg_update_r<GB_CGV_THREADS, GB_CGV_U><<<cg->grid_vec, GB_CGV_THREADS, 0, s>>>(
        cg->r, cg->ap, n, cg->chunk, cg->partials, g1, cg->partials2, cg->st);
    GB_LAUNCH_CHECK();
    gb_cg_update_xp<GB_CGV_THREADS, GB_CGV_U><<<cg->grid_vec, GB_CGV_THREADS, 0, s>>>(
        cg->x, cg->r, cg->p, n, cg->chunk, cg->partials2, cg->grid_vec, cg->st,
        cg->trace, cg->trace_cap);
    GB_LAUNCH_CHECK();
  }
  return GOOSEBERRY_SUCCESS;
}

static void gb_cg_drop_graph(gbd_cg *cg) {
  for (int i = 0; i < GB_CG_NGRAPH; ++i) {
    if (cg->gexec[i]) (void)hipGraphExecDestroy(cg->gexec[i]);
    if (cg->graph[i]) (void)hipGraphDestroy(cg->graph[i]);
    cg->gexec[i] = nullptr;
    cg->graph[i] = nullptr;
  }
  cg->graph_x = nullptr;
}

extern "C" {

// ------------------------------------------------- allocation-free solves
// The workspace of gbd_cg_scratch solves (r, p, Ap, the two partial arrays
// and the state) lives in this file-scope device array: such a solve makes
// no device allocation.  Each device holds its own copy at its own address,
// looked up once per device with hipGetSymbolAddress; host code uses that
// address, never the array's name.  Every such solve on a device shares the
// array, so solves that overlap on one device would race (the library lock
// serialises the public entry points of one process).
#define GB_CG_SCRATCH_BYTES (128ull << 20)
#define GB_CG_SCRATCH_MAXDEV 64
static __device__ __align__(256) unsigned char scratch[GB_CG_SCRATCH_BYTES];
static void *g_scratch_addr[GB_CG_SCRATCH_MAXDEV];
static gbd_cg g_scratch_cg;            // host bookkeeping of the shared workspace
static gbd_cg_state g_scratch_hst;     // its polled state (pageable host memory)

// Launch geometry and partial-array sizes of a solve with m.
static void gb_cg_geometry(gbd_cg *cg, const gbd_matrix *m) {
  cg->m = m;
  cg->n = m->nrows;
  uint32_t ge = gb_div_up(cg->n ? cg->n : 1, GB_BLOCK);
  cg->grid_elem = ge > GB_ELEM_GRID_MAX ? GB_ELEM_GRID_MAX : ge;
  // the vector updates: at most one workgroup per CU, each a contiguous
  // chunk of a multiple of 64 rows (aligned wavefront accesses)
  const uint64_t nn = cg->n ? cg->n : 1;
  uint32_t gv = gb_div_up(nn, GB_CGV_THREADS);
  if (gv > GB_CGV_GRID) gv = GB_CGV_GRID;
  cg->chunk = gb_div_up(gb_div_up(nn, gv), GB_WAVE) * GB_WAVE;
  cg->grid_vec = gb_div_up(nn, cg->chunk);
  uint32_t gs = gb_spmv_npartials(m);
  cg->npartials = gs > cg->grid_elem ? gs : cg->grid_elem;
}

static uint32_t gb_cg_np2(const gbd_cg *cg) {
  return cg->grid_elem > cg->grid_vec ? cg->grid_elem : cg->grid_vec;
}

int gbd_cg_create(const gbd_matrix *m, gbd_cg **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (m->nrows != m->ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_cg *cg = (gbd_cg *)calloc(1, sizeof(gbd_cg));
  if (!cg) return GOOSEBERRY_ERROR_NOMEMORY;
  gb_cg_geometry(cg, m);
  size_t nb = (size_t)(cg->n ? cg->n : 1) * sizeof(double);
  hipError_t e = gb_dev_alloc((void **)&cg->r, nb);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&cg->p, nb);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&cg->ap, nb);
  if (e == hipSuccess)
    e = gb_dev_alloc((void **)&cg->partials, (size_t)cg->npartials * sizeof(double));
  if (e == hipSuccess)
    e = gb_dev_alloc((void **)&cg->partials2, (size_t)gb_cg_np2(cg) * sizeof(double));
  if (e == hipSuccess) e = gb_dev_alloc((void **)&cg->st, sizeof(gbd_cg_state));
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&cg->h_st, sizeof(gbd_cg_state), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(cg->st, 0, sizeof(gbd_cg_state), gb_stream());
  if (e != hipSuccess) {
    gbd_cg_destroy(cg);
    return gb_hip_fail(e, "gbd_cg_create", __FILE__, __LINE__);
  }
  *out = cg;
  return GOOSEBERRY_SUCCESS;
}

void gbd_cg_destroy(gbd_cg *cg) {
  if (!cg || cg->noalloc) return;
  (void)hipStreamSynchronize(gb_stream());
  gb_cg_drop_graph(cg);
  (void)gb_dev_free(cg->r);
  (void)gb_dev_free(cg->p);
  (void)gb_dev_free(cg->ap);
  (void)gb_dev_free(cg->partials);
  (void)gb_dev_free(cg->partials2);
  (void)gb_dev_free(cg->st);
  (void)gb_dev_free(cg->trace);
  if (cg->h_st) (void)hipHostFree(cg->h_st);
  if (cg->h_trace) (void)hipHostFree(cg->h_trace);
  free(cg);
}

int gbd_cg_start(gbd_cg *cg, const double *d_b, double *d_x, double tol,
                 int rel, uint64_t maxiter, uint32_t trace_cap) {
  hipStream_t s = gb_stream();
  if (trace_cap != cg->trace_cap) {
    GB_HIP(hipStreamSynchronize(s));
    gb_cg_drop_graph(cg);
    if (cg->trace) { GB_HIP(gb_dev_free(cg->trace)); cg->trace = nullptr; }
    if (cg->h_trace) { GB_HIP(hipHostFree(cg->h_trace)); cg->h_trace = nullptr; }
    if (trace_cap) {
      GB_HIP(gb_dev_alloc((void **)&cg->trace, (size_t)trace_cap * sizeof(double)));
      GB_HIP(hipHostMalloc((void **)&cg->h_trace, (size_t)trace_cap * sizeof(double),
                           hipHostMallocDefault));
    }
    cg->trace_cap = trace_cap;
  }
  if (cg->x != d_x) gb_cg_drop_graph(cg);
  cg->b = d_b;
  cg->x = d_x;
  gb_cg_init<<<cg->grid_elem, GB_BLOCK, 0, s>>>(d_b, d_x, cg->r, cg->p, cg->n,
                                                cg->partials2);
  GB_LAUNCH_CHECK();
  gb_cg_init_state<<<1, GB_BLOCK, 0, s>>>(cg->partials2, cg->grid_elem, cg->st,
                                          tol, rel, (unsigned long long)maxiter);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_cg_enqueue(gbd_cg *cg, uint32_t k) { return gb_cg_enqueue_raw(cg, k); }

// k iterations as replays of captured graphs of 2^i iterations (one graph
// per power of two, captured on first use and kept while x stays the same).
int gbd_cg_enqueue_graph(gbd_cg *cg, uint32_t k) {
  hipStream_t s = gb_stream();
  if (!gb_graphs_enabled() || cg->noalloc) return gb_cg_enqueue_raw(cg, k);
  if (cg->graph_x != cg->x) {
    gb_cg_drop_graph(cg);
    cg->graph_x = cg->x;
  }
  for (int i = GB_CG_NGRAPH - 1; i >= 0 && k; --i) {
    const uint32_t len = 1u << i;
    while (k >= len) {
      if (!cg->gexec[i]) {
        GB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rv = gb_cg_enqueue_raw(cg, len);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rv != GOOSEBERRY_SUCCESS) {
          if (g) (void)hipGraphDestroy(g);
          return rv;
        }
        GB_HIP(e);
        cg->graph[i] = g;
        GB_HIP(hipGraphInstantiate(&cg->gexec[i], g, nullptr, nullptr, 0));
      }
      GB_HIP(hipGraphLaunch(cg->gexec[i], s));
      k -= len;
    }
  }
  return GOOSEBERRY_SUCCESS;
}

int gbd_cg_poll(gbd_cg *cg, gbd_cg_state *st, double *h_trace) {
  hipStream_t s = gb_stream();
  GB_HIP(hipMemcpyAsync(cg->h_st, cg->st, sizeof(gbd_cg_state),
                        hipMemcpyDeviceToHost, s));
  if (h_trace && cg->trace_cap)
    GB_HIP(hipMemcpyAsync(cg->h_trace, cg->trace,
                          (size_t)cg->trace_cap * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  *st = *cg->h_st;
  // the last K4 of a stopped solve records its status before any K2 has
  // turned it into `done`
  if (st->status) st->done = 1;
  if (h_trace && cg->trace_cap)
    memcpy(h_trace, cg->h_trace, (size_t)cg->trace_cap * sizeof(double));
  return GOOSEBERRY_SUCCESS;
}

size_t gbd_cg_scratch_max() { return GB_CG_SCRATCH_BYTES; }

int gbd_cg_scratch(const gbd_matrix *m, gbd_cg **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (m->nrows != m->ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_cg *cg = &g_scratch_cg;
  memset(cg, 0, sizeof(*cg));
  gb_cg_geometry(cg, m);
  auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  const uint64_t nb = al((uint64_t)(cg->n ? cg->n : 1) * sizeof(double));
  const uint64_t o_p = nb, o_ap = 2 * nb, o_pa = 3 * nb;
  const uint64_t o_pa2 = o_pa + al((uint64_t)cg->npartials * sizeof(double));
  const uint64_t o_st = o_pa2 + al((uint64_t)gb_cg_np2(cg) * sizeof(double));
  const uint64_t need = o_st + al(sizeof(gbd_cg_state));
  if (need > GB_CG_SCRATCH_BYTES) {
    gb_set_error("the allocation-free CG needs %llu bytes of workspace for this matrix "
                 "(%u rows); its device array holds %llu",
                 (unsigned long long)need, m->nrows, (unsigned long long)GB_CG_SCRATCH_BYTES);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  int dev = 0;
  GB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= GB_CG_SCRATCH_MAXDEV) {
    gb_set_error("device %d is beyond the %d devices the scratch table covers", dev,
                 GB_CG_SCRATCH_MAXDEV);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (!g_scratch_addr[dev]) {
    void *a = nullptr;
    GB_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    g_scratch_addr[dev] = a;
  }
  unsigned char *base = (unsigned char *)g_scratch_addr[dev];
  cg->r = (double *)base;
  cg->p = (double *)(base + o_p);
  cg->ap = (double *)(base + o_ap);
  cg->partials = (double *)(base + o_pa);
  cg->partials2 = (double *)(base + o_pa2);
  cg->st = (gbd_cg_state *)(base + o_st);
  cg->h_st = &g_scratch_hst;
  cg->noalloc = true;
  *out = cg;
  return GOOSEBERRY_SUCCESS;
}

}  // extern "C"
