// gb_cg.hip -- device side of the conjugate-gradient solver.
//
// Reference: src/cgd.c:40-129.  One iteration is four kernels on the
// library stream, all scalars kept on the device (no host round trip):
//   K1 gb_spmv_kernel<DOT>: ap = A p and the p.ap partials
//   K2 gb_cg_alpha:         alpha = dr / p.ap
//   K3 gb_cg_update_r:      r -= alpha ap; r.r partials
//   K4 gb_cg_update_xp:     beta = r.r / dr, convergence test;
//                           ps = alpha p; x += ps; p = r + (beta/alpha) ps
// which is the reference's scale-then-add order (src/cgd.c:92-105) with
// the same separately rounded products and sums (-ffp-contract=off); only
// the order of the dot-product sums differs.  x += ps sits in K4, which
// reads p anyway (one vector read less than in K3; x does not feed r).
// K2-K4 first read the device `done` flag and return at once when the
// solve has stopped, so the host can enqueue iterations in batches (or
// replay them from a hipGraph) and still stop exactly where the reference
// stops.  `done` is written only by the state's set-up and by K2 (one
// workgroup): K4 records its stop in `status`, which every K4 workgroup
// computes for itself, so all of them still add their share of x; the next
// K2 turns a nonzero status into `done`.
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
  uint32_t grid_elem;
  hipGraph_t graph[GB_CG_NGRAPH];       // graph[i]: 2^i iterations
  hipGraphExec_t gexec[GB_CG_NGRAPH];
  double *graph_x;                      // x the graphs were captured with
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

// src/cgd.c:90: alpha = dr / p.Ap from the SpMV's per-row-block partials,
// added in a fixed order by one 1024-thread workgroup; each thread issues
// all of its loads before adding (one memory round trip per 16K partials),
// and the first round's loads go out before the stop flag is known.
#define GB_RED_THREADS 1024
#define GB_RED_UNROLL 16
__global__ void __launch_bounds__(GB_RED_THREADS)
gb_cg_alpha(const double *__restrict__ pap_partials, uint32_t npap,
            gbd_cg_state *st) {
  __shared__ double red[GB_RED_THREADS / GB_WAVE];
  __shared__ int32_t stop;
  double v[GB_RED_UNROLL];
#pragma unroll
  for (int j = 0; j < GB_RED_UNROLL; ++j) {
    uint32_t i = j * GB_RED_THREADS + threadIdx.x;
    v[j] = i < npap ? pap_partials[i] : 0.0;
  }
  if (threadIdx.x == 0) {
    // a stop recorded by the previous iteration's K4 becomes `done` here
    stop = st->done || st->status;
    if (stop && !st->done) st->done = 1;
  }
  __syncthreads();
  if (stop) return;
  double s = 0.0;
  for (uint32_t base = 0;;) {
#pragma unroll
    for (int j = 0; j < GB_RED_UNROLL; ++j) s += v[j];
    base += GB_RED_THREADS * GB_RED_UNROLL;
    if (base >= npap) break;
#pragma unroll
    for (int j = 0; j < GB_RED_UNROLL; ++j) {
      uint32_t i = base + j * GB_RED_THREADS + threadIdx.x;
      v[j] = i < npap ? pap_partials[i] : 0.0;
    }
  }
  s = gb_wave_sum(s);
  if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double pap = red[0];
    for (int w = 1; w < GB_RED_THREADS / GB_WAVE; ++w) pap += red[w];
    const double dr = st->dr;
    const double alpha = dr / pap;
    st->pap = pap;
    st->alpha = alpha;
    if (!(pap != 0.0) || !isfinite(alpha)) {
      st->done = 1;
      st->status = 3;
    } else {
      st->dr_prev = dr;
      st->iter_next = st->iter + 1;
    }
  }
}

// src/cgd.c:96-97: r -= alpha Ap; partial sums of r.r
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_update_r(double *__restrict__ r, const double *__restrict__ ap,
               uint32_t n, double *__restrict__ partials,
               const gbd_cg_state *st) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  if (gb_block_flag(&st->done)) return;
  const double alpha = st->alpha;
  double d = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double aps = alpha * ap[i];
    double rn = r[i] - aps;
    r[i] = rn;
    d += rn * rn;
  }
  double bs = gb_block_sum(d, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// src/cgd.c:92-94, 100-119: beta = r.r / dr; stop when sqrt(r.r) <= tol or
// at the iteration cap; ps = alpha p; x += ps; unless stopping,
// p = nr + (beta/alpha) ps
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_update_xp(double *__restrict__ x, const double *__restrict__ r,
               double *__restrict__ p,
               uint32_t n, const double *__restrict__ dnr_partials,
               uint32_t ndnr, gbd_cg_state *st, double *__restrict__ trace,
               uint32_t trace_cap) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  if (gb_block_flag(&st->done)) return;
  const double dnr = gb_sum_partials(dnr_partials, ndnr, red);
  const double alpha = st->alpha;
  const double beta = dnr / st->dr_prev;
  const double rnorm = sqrt(dnr);
  const unsigned long long it1 = st->iter_next;
  int status = 0;
  if (rnorm <= st->tol) status = 1;
  else if (!isfinite(dnr)) status = 3;
  else if (st->maxiter && it1 >= st->maxiter) status = 2;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->beta = beta;
    st->dr = dnr;
    st->rnorm = rnorm;
    if (trace_cap) trace[(it1 - 1) % trace_cap] = rnorm;
    st->iter = it1;
    // not `done`: the other workgroups of this launch may not have read it
    if (status) st->status = status;
  }
  const double boa = beta / alpha;
  if (status) {
    for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
         i += gridDim.x * GB_BLOCK)
      x[i] = x[i] + alpha * p[i];
    return;
  }
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double ps = alpha * p[i];
    x[i] = x[i] + ps;
    double t = ps * boa;
    p[i] = r[i] + t;
  }
}

static int gb_cg_enqueue_raw(gbd_cg *cg, uint32_t k) {
  hipStream_t s = gb_stream();
  const uint32_t n = cg->n;
  const uint32_t g1 = gb_spmv_npartials(cg->m);
  for (uint32_t it = 0; it < k; ++it) {
    int rv = gb_spmv_launch(cg->m, cg->p, cg->ap, 1, cg->partials, cg->st);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    gb_cg_alpha<<<1, GB_RED_THREADS, 0, s>>>(cg->partials, g1, cg->st);
    GB_LAUNCH_CHECK();
    gb_cg_update_r<<<cg->grid_elem, GB_BLOCK, 0, s>>>(
        cg->r, cg->ap, n, cg->partials2, cg->st);
    GB_LAUNCH_CHECK();
    gb_cg_update_xp<<<cg->grid_elem, GB_BLOCK, 0, s>>>(
        cg->x, cg->r, cg->p, n, cg->partials2, cg->grid_elem, cg->st, cg->trace,
        cg->trace_cap);
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

int gbd_cg_create(const gbd_matrix *m, gbd_cg **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (m->nrows != m->ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_cg *cg = (gbd_cg *)calloc(1, sizeof(gbd_cg));
  if (!cg) return GOOSEBERRY_ERROR_NOMEMORY;
  cg->m = m;
  cg->n = m->nrows;
  size_t nb = (size_t)(cg->n ? cg->n : 1) * sizeof(double);
  uint32_t ge = gb_div_up(cg->n ? cg->n : 1, GB_BLOCK);
  cg->grid_elem = ge > GB_ELEM_GRID_MAX ? GB_ELEM_GRID_MAX : ge;
  uint32_t gs = gb_spmv_npartials(m);
  cg->npartials = gs > cg->grid_elem ? gs : cg->grid_elem;
  hipError_t e = gb_dev_alloc((void **)&cg->r, nb);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&cg->p, nb);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&cg->ap, nb);
  if (e == hipSuccess)
    e = gb_dev_alloc((void **)&cg->partials, (size_t)cg->npartials * sizeof(double));
  if (e == hipSuccess)
    e = gb_dev_alloc((void **)&cg->partials2, (size_t)cg->grid_elem * sizeof(double));
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
  if (!cg) return;
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
  if (!gb_graphs_enabled()) return gb_cg_enqueue_raw(cg, k);
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

}  // extern "C"
