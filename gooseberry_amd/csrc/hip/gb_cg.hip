// gb_cg.hip -- device side of the conjugate-gradient solver.
//
// Reference: src/cgd.c:40-129.  One iteration is three kernels on the
// library stream, all scalars kept on the device (no host round trip):
//   K1 gb_spmv_kernel<DOT>: ap = A p and p.ap -> alpha = dr / p.ap
//   K2 gb_cg_update_xr:     ps = alpha p; x += ps; r -= alpha ap; r.r
//                           -> beta = r.r / dr, convergence test
//   K3 gb_cg_update_p:      p = r + (beta/alpha) (alpha p)
// which is the reference's scale-then-add order (src/cgd.c:92-105) with
// the same separately rounded products and sums (-ffp-contract=off); only
// the order of the dot-product sums differs.  Every kernel first reads the
// device `done` flag and returns at once when the solve has stopped, so
// the host can enqueue iterations in batches (or replay them from a
// hipGraph) and still stop exactly where the reference stops.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gb_matrix.h"

#define GB_ELEM_GRID_MAX 2048u

struct gbd_cg {
  const gbd_matrix *m;
  uint32_t n;
  double *r, *p, *ap;
  double *partials;
  uint32_t npartials;
  gbd_cg_state *st;
  gbd_cg_state *h_st;  // pinned
  double *trace;
  uint32_t trace_cap;
  double *h_trace;     // pinned, trace_cap entries
  const double *b;
  double *x;
  uint32_t grid_elem;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  uint32_t graph_k;
  double *graph_x;
};

// Single-pass grid reduction: block partial -> partials[blockIdx]; the
// last block (ticket) sums partials in block order.  Returns true on
// thread 0 of the last block with *total set.
__device__ __forceinline__ bool gb_grid_sum(double v, double *partials,
                                            uint32_t *ticket, double *red,
                                            uint32_t *is_last,
                                            double *total) {
  double bs = gb_block_sum(v, red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = bs;
    __threadfence();
    uint32_t t = atomicAdd(ticket, 1u);
    *is_last = (t == gridDim.x - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!*is_last) return false;
  __threadfence();
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += GB_BLOCK) s += partials[i];
  s = gb_block_sum(s, red);
  if (threadIdx.x == 0) {
    *ticket = 0;
    *total = s;
    return true;
  }
  return false;
}

// src/cgd.c:81-85: x = 0; r = b; p = r; dr = r.r
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_init(const double *__restrict__ b, double *__restrict__ x,
           double *__restrict__ r, double *__restrict__ p, uint32_t n,
           double *__restrict__ partials, gbd_cg_state *st, double tol,
           int rel, unsigned long long maxiter) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  __shared__ uint32_t is_last;
  double d = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double v = b[i];
    x[i] = 0.0;
    r[i] = v;
    p[i] = v;
    d += v * v;
  }
  double s;
  if (gb_grid_sum(d, partials, &st->ticket[1], red, &is_last, &s)) {
    st->dr = s;
    st->rnorm0 = sqrt(s);
    st->rnorm = st->rnorm0;
    st->tol = rel ? tol * st->rnorm0 : tol;
    st->pap = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->iter = 0;
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

// src/cgd.c:92-101: ps = alpha p; nx = x + ps; nr = r - alpha ap; dnr = nr.nr
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_update_xr(double *__restrict__ x, double *__restrict__ r,
                const double *__restrict__ p, const double *__restrict__ ap,
                uint32_t n, double *__restrict__ partials, gbd_cg_state *st,
                double *__restrict__ trace, uint32_t trace_cap) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  __shared__ uint32_t is_last;
  if (st->done) return;
  const double alpha = st->alpha;
  double d = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double ps = alpha * p[i];
    x[i] = x[i] + ps;
    double aps = alpha * ap[i];
    double rn = r[i] - aps;
    r[i] = rn;
    d += rn * rn;
  }
  double dnr;
  if (gb_grid_sum(d, partials, &st->ticket[1], red, &is_last, &dnr)) {
    // src/cgd.c:100-119
    st->beta = dnr / st->dr;
    st->dr = dnr;
    double rnorm = sqrt(dnr);
    st->rnorm = rnorm;
    unsigned long long it = st->iter;
    if (trace_cap) trace[it % trace_cap] = rnorm;
    st->iter = it + 1;
    if (rnorm <= st->tol) {
      st->done = 1;
      st->status = 1;
    } else if (!isfinite(dnr)) {
      st->done = 1;
      st->status = 3;
    } else if (st->maxiter && it + 1 >= st->maxiter) {
      st->done = 1;
      st->status = 2;
    }
  }
}

// src/cgd.c:104-105: p = nr + (beta/alpha) * (alpha p)
__global__ void __launch_bounds__(GB_BLOCK)
gb_cg_update_p(const double *__restrict__ r, double *__restrict__ p,
               uint32_t n, const gbd_cg_state *st) {
  if (st->done) return;
  const double alpha = st->alpha;
  const double boa = st->beta / alpha;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK) {
    double ps = alpha * p[i];
    double t = ps * boa;
    p[i] = r[i] + t;
  }
}

static int gb_cg_enqueue_raw(gbd_cg *cg, uint32_t k) {
  hipStream_t s = gb_stream();
  const uint32_t n = cg->n;
  for (uint32_t it = 0; it < k; ++it) {
    int rv = gb_spmv_launch(cg->m, cg->p, cg->ap, 1, cg->partials, cg->st);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    gb_cg_update_xr<<<cg->grid_elem, GB_BLOCK, 0, s>>>(
        cg->x, cg->r, cg->p, cg->ap, n, cg->partials, cg->st, cg->trace,
        cg->trace_cap);
    GB_LAUNCH_CHECK();
    gb_cg_update_p<<<cg->grid_elem, GB_BLOCK, 0, s>>>(cg->r, cg->p, n, cg->st);
    GB_LAUNCH_CHECK();
  }
  return GOOSEBERRY_SUCCESS;
}

static void gb_cg_drop_graph(gbd_cg *cg) {
  if (cg->gexec) hipGraphExecDestroy(cg->gexec);
  if (cg->graph) hipGraphDestroy(cg->graph);
  cg->gexec = nullptr;
  cg->graph = nullptr;
  cg->graph_k = 0;
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
  uint32_t gs = gb_spmv_grid(m);
  cg->npartials = gs > cg->grid_elem ? gs : cg->grid_elem;
  hipError_t e = hipMalloc((void **)&cg->r, nb);
  if (e == hipSuccess) e = hipMalloc((void **)&cg->p, nb);
  if (e == hipSuccess) e = hipMalloc((void **)&cg->ap, nb);
  if (e == hipSuccess)
    e = hipMalloc((void **)&cg->partials, (size_t)cg->npartials * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&cg->st, sizeof(gbd_cg_state));
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
  hipStreamSynchronize(gb_stream());
  gb_cg_drop_graph(cg);
  hipFree(cg->r);
  hipFree(cg->p);
  hipFree(cg->ap);
  hipFree(cg->partials);
  hipFree(cg->st);
  hipFree(cg->trace);
  if (cg->h_st) hipHostFree(cg->h_st);
  if (cg->h_trace) hipHostFree(cg->h_trace);
  free(cg);
}

int gbd_cg_start(gbd_cg *cg, const double *d_b, double *d_x, double tol,
                 int rel, uint64_t maxiter, uint32_t trace_cap) {
  hipStream_t s = gb_stream();
  if (trace_cap != cg->trace_cap) {
    GB_HIP(hipStreamSynchronize(s));
    gb_cg_drop_graph(cg);
    if (cg->trace) { GB_HIP(hipFree(cg->trace)); cg->trace = nullptr; }
    if (cg->h_trace) { GB_HIP(hipHostFree(cg->h_trace)); cg->h_trace = nullptr; }
    if (trace_cap) {
      GB_HIP(hipMalloc((void **)&cg->trace, (size_t)trace_cap * sizeof(double)));
      GB_HIP(hipHostMalloc((void **)&cg->h_trace, (size_t)trace_cap * sizeof(double),
                           hipHostMallocDefault));
    }
    cg->trace_cap = trace_cap;
  }
  if (cg->x != d_x) gb_cg_drop_graph(cg);
  cg->b = d_b;
  cg->x = d_x;
  gb_cg_init<<<cg->grid_elem, GB_BLOCK, 0, s>>>(d_b, d_x, cg->r, cg->p, cg->n,
                                                cg->partials, cg->st, tol, rel,
                                                (unsigned long long)maxiter);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_cg_enqueue(gbd_cg *cg, uint32_t k) { return gb_cg_enqueue_raw(cg, k); }

int gbd_cg_enqueue_graph(gbd_cg *cg, uint32_t k) {
  hipStream_t s = gb_stream();
  if (!cg->gexec || cg->graph_k != k || cg->graph_x != cg->x) {
    gb_cg_drop_graph(cg);
    GB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rv = gb_cg_enqueue_raw(cg, k);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rv != GOOSEBERRY_SUCCESS) {
      if (g) hipGraphDestroy(g);
      return rv;
    }
    GB_HIP(e);
    cg->graph = g;
    GB_HIP(hipGraphInstantiate(&cg->gexec, g, nullptr, nullptr, 0));
    cg->graph_k = k;
    cg->graph_x = cg->x;
  }
  GB_HIP(hipGraphLaunch(cg->gexec, s));
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
  if (h_trace && cg->trace_cap)
    memcpy(h_trace, cg->h_trace, (size_t)cg->trace_cap * sizeof(double));
  return GOOSEBERRY_SUCCESS;
}

}  // extern "C"
