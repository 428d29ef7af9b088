/*
 * gb_batch_api.c -- batched solves, generated systems and the continuous-
 * batching solve server (include/gooseberry.h), on top of gb_device.h.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gooseberry.h"
#include "../gb_device.h"
#include "gb_host.h"

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------- batch */

struct gooseberry_batch_s {
  gbd_batch *b;
  uint32_t max;
  gbd_bstate *states;
  uint8_t *active;
  int dirty;
};

int gooseberry_batch_create(uint32_t max_systems, double delta,
    gooseberry_batch_t *r_batch)
{
  GB_API_GUARD();
  if (!r_batch || max_systems == 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  struct gooseberry_batch_s *B = calloc(1, sizeof(*B));
  if (!B) return GOOSEBERRY_ERROR_NOMEMORY;
  B->max = max_systems;
  B->states = calloc(max_systems, sizeof(gbd_bstate));
  B->active = calloc(max_systems, 1);
  if (!B->states || !B->active) {
    free(B->states); free(B->active); free(B);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  int rv = gbd_batch_create(max_systems, delta, &B->b);
  if (rv != GOOSEBERRY_SUCCESS) {
    free(B->states); free(B->active); free(B);
    return rv;
  }
  *r_batch = B;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_batch_free(gooseberry_batch_t B)
{
  GB_API_GUARD();
  if (!B) return GOOSEBERRY_SUCCESS;
  gbd_batch_destroy(B->b);
  free(B->states);
  free(B->active);
  free(B);
  return GOOSEBERRY_SUCCESS;
}

size_t gooseberry_batch_workspace_bytes(gooseberry_matrix_t mat, int precision)
{
  GB_API_GUARD();
  return mat ? gbd_batch_workspace_bytes(gb_matrix_dev(mat), precision) : 0;
}

int gooseberry_batch_add(gooseberry_batch_t B, gooseberry_matrix_t mat,
    const real_t *d_b, real_t *d_x, real_t tol, int tol_mode, size_t niter,
    int precision, uint32_t *r_slot)
{
  GB_API_GUARD();
  if (!B || !mat || !d_b || !d_x || !r_slot) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (tol_mode != GOOSEBERRY_TOL_ABSOLUTE && tol_mode != GOOSEBERRY_TOL_RELATIVE)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_matrix *m = gb_matrix_dev(mat);
  int rv;
  if (precision == GOOSEBERRY_PRECISION_MIXED &&
      (rv = gbd_matrix_make_f32(m)) != GOOSEBERRY_SUCCESS)
    return rv;
  if (precision == GOOSEBERRY_PRECISION_MIXED_JACOBI &&
      (rv = gbd_matrix_make_f32_jacobi(m)) != GOOSEBERRY_SUCCESS)
    return rv;
  uint32_t slot;
  rv = gbd_batch_add(B->b, m, d_b, d_x, tol, tol_mode == GOOSEBERRY_TOL_RELATIVE,
      niter, precision, &slot);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  B->active[slot] = 1;
  memset(&B->states[slot], 0, sizeof(gbd_bstate));
  B->dirty = 1;
  *r_slot = slot;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_batch_remove(gooseberry_batch_t B, uint32_t slot)
{
  GB_API_GUARD();
  if (!B || slot >= B->max) return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv = gbd_batch_remove(B->b, slot);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (B->active[slot]) B->dirty = 1;
  B->active[slot] = 0;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_batch_run(gooseberry_batch_t B, uint32_t steps,
    uint32_t *r_running)
{
  GB_API_GUARD();
  if (!B) return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv;
  uint32_t running = 0;
  for (uint32_t i = 0; i < B->max; ++i) running += B->active[i];
  if (running && steps) {
    if (B->dirty) {
      if ((rv = gbd_batch_commit(B->b, B->active)) != GOOSEBERRY_SUCCESS) return rv;
      B->dirty = 0;
    }
    if ((rv = gbd_batch_step(B->b, steps)) != GOOSEBERRY_SUCCESS) return rv;
  }
  if ((rv = gbd_batch_poll(B->b, B->states)) != GOOSEBERRY_SUCCESS) return rv;
  running = 0;
  for (uint32_t i = 0; i < B->max; ++i) {
    if (!B->active[i]) continue;
    if (B->states[i].done) {
      B->active[i] = 0;  /* finished systems leave the work lists */
      B->dirty = 1;
    } else {
      ++running;
    }
  }
  if (r_running) *r_running = running;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_batch_solve(gooseberry_batch_t B)
{
  GB_API_GUARD();
  uint32_t running = 1;
  int rv = GOOSEBERRY_SUCCESS;
  uint32_t k = 8;
  while (running && rv == GOOSEBERRY_SUCCESS) {
    rv = gooseberry_batch_run(B, k, &running);
    if (k < 32) k *= 2;
  }
  return rv;
}

int gooseberry_batch_info(gooseberry_batch_t B, uint32_t slot,
    gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  if (!B || !info || slot >= B->max || !gbd_batch_slot_used(B->b, slot))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  const gbd_bstate *s = &B->states[slot];
  memset(info, 0, sizeof(*info));
  info->niter = s->iter;
  info->nrefine = s->nrefine;
  info->bnorm = s->rnorm0;
  info->status = s->status;
  info->converged = s->status == 1;
  if (s->precision != GOOSEBERRY_PRECISION_FP64) {
    info->rnorm = s->rnorm64;
    info->true_rel_residual = s->rnorm0 > 0 ? s->rnorm64 / s->rnorm0 : 0.0;
  } else {
    info->rnorm = s->rnorm;
    info->true_rel_residual = NAN;
  }
  return GOOSEBERRY_SUCCESS;
}

/* ------------------------------------------------------ generated inputs */

int gooseberry_matrix_generate_spd(dim_t n, uint64_t seed, int law,
    double zipf_s, dim_t dmin, dim_t dmax, const gooseberry_plan_t *plan,
    gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  if (!r_mat) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t *rp = NULL, *ri = NULL, nnz = 0;
  double *rv_ = NULL;
  int rv = gbd_generate_spd(n, seed, law, zipf_s, dmin, dmax, &rp, &ri, &rv_, &nnz);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return gb_matrix_adopt(n, n, rp, ri, rv_, nnz, plan, r_mat);
}

int gooseberry_matrix_generate_sdd(dim_t n, uint64_t seed, dim_t dmin,
    dim_t dmax, const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  return gooseberry_matrix_generate_spd(n, seed, GOOSEBERRY_DEGREE_UNIFORM,
      0.0, dmin, dmax, plan, r_mat);
}

int gooseberry_matrix_download(gooseberry_matrix_t mat, ind_t *rowptr,
    dim_t *rowind, real_t *rowval)
{
  GB_API_GUARD();
  if (!mat || !rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_matrix *m = gb_matrix_dev(mat);
  uint32_t nrows;
  uint64_t nnz;
  gbd_matrix_info(m, &nrows, NULL, &nnz, NULL, NULL, NULL, NULL, NULL);
  if (gbd_matrix_nparts(m)) return gb_parts_unsupported();
  int rv = gbd_dtoh(rowptr, gbd_matrix_rowptr(m), ((size_t)nrows + 1) * 4);
  if (rv == GOOSEBERRY_SUCCESS && nnz && rowind)
    rv = gbd_dtoh(rowind, gbd_matrix_rowind(m), nnz * 4);
  if (rv == GOOSEBERRY_SUCCESS && nnz && rowval)
    rv = gbd_dtoh(rowval, gbd_matrix_rowval(m), nnz * 8);
  return rv;
}

int gooseberry_generate_uniform_device(uint64_t seed, uint64_t stream,
    size_t n, real_t lo, real_t hi, real_t *d_out)
{
  GB_API_GUARD();
  if (n && !d_out) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gbd_fill_uniform(d_out, n, seed, stream, lo, hi);
}

/* ----------------------------------------------------------------- server */

enum { REQ_QUEUED = 0, REQ_ACTIVE = 1, REQ_DONE = 2 };

typedef struct {
  uint64_t id;
  int generated;
  dim_t n, dmin, dmax;
  uint64_t seed;
  double tol;
  gooseberry_matrix_t mat;  /* owned when generated */
  const real_t *d_b;
  real_t *d_x;
  real_t *own_b, *own_x;    /* generated requests own b and x */
  size_t footprint;
  uint32_t slot;
  int state;
  gooseberry_request_result_t res;
} req_t;

struct gooseberry_server_s {
  gooseberry_batch_t batch;
  uint32_t max_active;
  int precision;
  int keep;
  size_t budget, in_use;
  req_t **queue;            /* FIFO of queued requests */
  size_t qhead, qtail, qcap;
  req_t **slots;            /* active request per batch slot */
  req_t **kept;             /* finished generated requests holding x */
  size_t nkept, capkept;
  uint32_t nactive;
  double timeout;           /* seconds after submission; <= 0: none */
  gooseberry_request_result_t *pending;  /* ended outside a batch step */
  size_t npending, cappending;
};

/* Device bytes a request holds while admitted: its workspace, and for a
 * generated request an upper bound of its matrix (nnz <= n (2 dmax - 1)),
 * b and x. */
static size_t req_footprint(const struct gooseberry_server_s *S, const req_t *r)
{
  const size_t n = r->n;
  const int mixed = S->precision != GOOSEBERRY_PRECISION_FP64;
  size_t vec = (mixed ? 8 + 16 : 24) * n;
  size_t ws = vec + 64 * 1024 + n / 4;   /* vectors + partials and slack */
  if (!r->generated) return ws;
  size_t nnz = n * (2 * (size_t)r->dmax - 1);
  size_t mat = nnz * (4 + 8 + (mixed ? 4 : 0))
             + (S->precision == GOOSEBERRY_PRECISION_MIXED_JACOBI ? 8 * n : 0)
             + 4 * (n + 1)       /* rowptr */
             + 4 * n + 8 * (n / 8 + 16); /* plan: row bins, blocks */
  return ws + mat + 16 * n;      /* b and x */
}

/* Device bytes a generated request needs on top of its footprint while it
 * is generated (edge lists and the counting sort's buffers, 30 bytes per
 * drawn edge at most, all freed before admission ends). */
static size_t req_transient(const req_t *r)
{
  if (!r->generated) return 0;
  const size_t m = (size_t)r->n * ((size_t)r->dmax - 1);
  return 30 * m + 12 * (size_t)r->n + (1u << 20);
}

int gooseberry_server_create(size_t mem_budget, uint32_t max_active,
    int precision, double delta, int keep_results, gooseberry_server_t *r)
{
  GB_API_GUARD();
  if (!r || max_active == 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (precision != GOOSEBERRY_PRECISION_FP64 && precision != GOOSEBERRY_PRECISION_MIXED &&
      precision != GOOSEBERRY_PRECISION_MIXED_JACOBI)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  struct gooseberry_server_s *S = calloc(1, sizeof(*S));
  if (!S) return GOOSEBERRY_ERROR_NOMEMORY;
  int rv;
  if (!mem_budget) {
    size_t fr = 0, tot = 0;
    if ((rv = gbd_device_memory(&fr, &tot)) != GOOSEBERRY_SUCCESS) { free(S); return rv; }
    mem_budget = (size_t)(0.9 * (double)fr);
  }
  S->budget = mem_budget;
  S->max_active = max_active;
  S->precision = precision;
  S->keep = keep_results;
  S->slots = calloc(max_active, sizeof(req_t *));
  if (!S->slots) { free(S); return GOOSEBERRY_ERROR_NOMEMORY; }
  rv = gooseberry_batch_create(max_active, delta, &S->batch);
  if (rv != GOOSEBERRY_SUCCESS) { free(S->slots); free(S); return rv; }
  *r = S;
  return GOOSEBERRY_SUCCESS;
}

static void req_release(req_t *r)
{
  if (r->generated) {
    gooseberry_matrix_free(r->mat);
    gbd_free(r->own_b);
    gbd_free(r->own_x);
  }
  free(r);
}

int gooseberry_server_free(gooseberry_server_t S)
{
  GB_API_GUARD();
  if (!S) return GOOSEBERRY_SUCCESS;
  gooseberry_device_synchronize();
  /* the batch first: its slots hold the active requests' matrices */
  gooseberry_batch_free(S->batch);
  for (size_t i = S->qhead; i < S->qtail; ++i) free(S->queue[i]);
  for (uint32_t i = 0; i < S->max_active; ++i)
    if (S->slots[i]) req_release(S->slots[i]);
  for (size_t i = 0; i < S->nkept; ++i) req_release(S->kept[i]);
  free(S->queue);
  free(S->slots);
  free(S->kept);
  free(S->pending);
  free(S);
  return GOOSEBERRY_SUCCESS;
}

/* record the result of a request that ends outside a batch step */
static int push_pending(struct gooseberry_server_s *S, const req_t *r, int status)
{
  if (S->npending == S->cappending) {
    size_t nc = S->cappending ? 2 * S->cappending : 64;
    gooseberry_request_result_t *np = realloc(S->pending, nc * sizeof(*np));
    if (!np) return GOOSEBERRY_ERROR_NOMEMORY;
    S->pending = np;
    S->cappending = nc;
  }
  gooseberry_request_result_t res = r->res;
  res.status = status;
  res.done_s = now_s();
  S->pending[S->npending++] = res;
  return GOOSEBERRY_SUCCESS;
}

/* take request i of the queue out (order of the others kept) */
static void queue_remove(struct gooseberry_server_s *S, size_t i)
{
  memmove(S->queue + i, S->queue + i + 1, (S->qtail - i - 1) * sizeof(req_t *));
  S->qtail--;
}

/* end the running request in batch slot i (its x keeps the last folded
 * iterate) */
static void drop_active(struct gooseberry_server_s *S, uint32_t i)
{
  req_t *r = S->slots[i];
  gooseberry_batch_remove(S->batch, i);
  S->slots[i] = NULL;
  S->nactive--;
  S->in_use -= r->footprint;
  if (r->generated) req_release(r);
  else free(r);
}

int gooseberry_server_cancel(gooseberry_server_t S, uint64_t id)
{
  GB_API_GUARD();
  if (!S) return GOOSEBERRY_ERROR_INVALIDINPUT;
  for (size_t i = S->qhead; i < S->qtail; ++i) {
    req_t *r = S->queue[i];
    if (r->id != id) continue;
    int rv = push_pending(S, r, GOOSEBERRY_REQUEST_CANCELLED);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    queue_remove(S, i);
    free(r);
    return GOOSEBERRY_SUCCESS;
  }
  for (uint32_t i = 0; i < S->max_active; ++i) {
    req_t *r = S->slots[i];
    if (!r || r->id != id) continue;
    int rv = push_pending(S, r, GOOSEBERRY_REQUEST_CANCELLED);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    drop_active(S, i);
    return GOOSEBERRY_SUCCESS;
  }
  return GOOSEBERRY_ERROR_INVALIDINPUT;
}

int gooseberry_server_set_timeout(gooseberry_server_t S, double seconds)
{
  GB_API_GUARD();
  if (!S || seconds != seconds) return GOOSEBERRY_ERROR_INVALIDINPUT;
  S->timeout = seconds;
  return GOOSEBERRY_SUCCESS;
}

/* cancel what is past the timeout, queued or running */
static int expire(struct gooseberry_server_s *S)
{
  if (!(S->timeout > 0)) return GOOSEBERRY_SUCCESS;
  const double t = now_s();
  int rv;
  for (size_t i = S->qhead; i < S->qtail;) {
    req_t *r = S->queue[i];
    if (t - r->res.submit_s <= S->timeout) { ++i; continue; }
    if ((rv = push_pending(S, r, GOOSEBERRY_REQUEST_TIMEDOUT)) != GOOSEBERRY_SUCCESS)
      return rv;
    queue_remove(S, i);
    free(r);
  }
  for (uint32_t i = 0; i < S->max_active; ++i) {
    req_t *r = S->slots[i];
    if (!r || t - r->res.submit_s <= S->timeout) continue;
    if ((rv = push_pending(S, r, GOOSEBERRY_REQUEST_TIMEDOUT)) != GOOSEBERRY_SUCCESS)
      return rv;
    drop_active(S, i);
  }
  return GOOSEBERRY_SUCCESS;
}

static int enqueue(struct gooseberry_server_s *S, req_t *r)
{
  if (S->qtail == S->qcap) {
    if (S->qhead > 0) {
      memmove(S->queue, S->queue + S->qhead, (S->qtail - S->qhead) * sizeof(req_t *));
      S->qtail -= S->qhead;
      S->qhead = 0;
    }
    if (S->qtail == S->qcap) {
      size_t nc = S->qcap ? 2 * S->qcap : 256;
      req_t **nq = realloc(S->queue, nc * sizeof(req_t *));
      if (!nq) return GOOSEBERRY_ERROR_NOMEMORY;
      S->queue = nq;
      S->qcap = nc;
    }
  }
  r->res.submit_s = now_s();
  r->res.id = r->id;
  r->res.n = r->n;
  r->footprint = req_footprint(S, r);
  S->queue[S->qtail++] = r;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_server_submit_sdd(gooseberry_server_t S, uint64_t id, dim_t n,
    uint64_t seed, dim_t dmin, dim_t dmax, real_t tol)
{
  GB_API_GUARD();
  if (!S || n < 2 || dmin < 1 || dmax < dmin || dmax > 65535 || !(tol > 0))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  req_t *r = calloc(1, sizeof(*r));
  if (!r) return GOOSEBERRY_ERROR_NOMEMORY;
  r->id = id;
  r->generated = 1;
  r->n = n;
  r->seed = seed;
  r->dmin = dmin;
  r->dmax = dmax;
  r->tol = tol;
  int rv = enqueue(S, r);
  if (rv != GOOSEBERRY_SUCCESS) free(r);
  return rv;
}

int gooseberry_server_submit(gooseberry_server_t S, uint64_t id,
    gooseberry_matrix_t mat, const real_t *d_b, real_t *d_x, real_t tol)
{
  GB_API_GUARD();
  if (!S || !mat || !d_b || !d_x || !(tol > 0)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gooseberry_matrix_info_t mi;
  int rv = gooseberry_matrix_info(mat, &mi);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (mi.nrows != mi.ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  req_t *r = calloc(1, sizeof(*r));
  if (!r) return GOOSEBERRY_ERROR_NOMEMORY;
  r->id = id;
  r->n = mi.nrows;
  r->tol = tol;
  r->mat = mat;
  r->d_b = d_b;
  r->d_x = d_x;
  r->res.nnz = mi.nnz;
  rv = enqueue(S, r);
  if (rv != GOOSEBERRY_SUCCESS) free(r);
  return rv;
}

/* materialise (generated requests) and add to the batch */
static int admit(struct gooseberry_server_s *S, req_t *r)
{
  int rv;
  if (r->generated) {
    rv = gooseberry_matrix_generate_sdd(r->n, r->seed, r->dmin, r->dmax, NULL, &r->mat);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    gooseberry_matrix_info_t mi;
    gooseberry_matrix_info(r->mat, &mi);
    r->res.nnz = mi.nnz;
    if ((rv = gbd_malloc((void **)&r->own_b, (size_t)r->n * 8)) != GOOSEBERRY_SUCCESS ||
        (rv = gbd_malloc((void **)&r->own_x, (size_t)r->n * 8)) != GOOSEBERRY_SUCCESS ||
        (rv = gbd_fill_uniform(r->own_b, r->n, r->seed, 1, -1.0, 1.0)) != GOOSEBERRY_SUCCESS)
      goto fail;
    r->d_b = r->own_b;
    r->d_x = r->own_x;
  }
  rv = gooseberry_batch_add(S->batch, r->mat, r->d_b, r->d_x, r->tol,
      GOOSEBERRY_TOL_RELATIVE, 0, S->precision, &r->slot);
  if (rv != GOOSEBERRY_SUCCESS) goto fail;
  r->state = REQ_ACTIVE;
  r->res.admit_s = now_s();
  S->slots[r->slot] = r;
  S->nactive++;
  return GOOSEBERRY_SUCCESS;
fail:
  if (r->generated) {
    gooseberry_matrix_free(r->mat);
    gbd_free(r->own_b);
    gbd_free(r->own_x);
    r->mat = NULL;
    r->own_b = r->own_x = NULL;
  }
  return rv;
}

int gooseberry_server_progress(gooseberry_server_t S, uint32_t steps,
    gooseberry_request_result_t *done, uint32_t max_done, uint32_t *r_ndone)
{
  GB_API_GUARD();
  if (!S) return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv;
  /* 0. requests past the timeout leave */
  if ((rv = expire(S)) != GOOSEBERRY_SUCCESS) return rv;
  /* 1. admission: first in, first out, while the head fits */
  while (S->qhead < S->qtail && S->nactive < S->max_active) {
    req_t *r = S->queue[S->qhead];
    if (S->in_use + r->footprint + req_transient(r) > S->budget) {
      if (S->nactive == 0 && S->nkept == 0) {
        /* can never fit: fail the request instead of blocking the queue */
        S->qhead++;
        r->res.status = 0;
        r->state = REQ_DONE;
        free(r);
        return GOOSEBERRY_ERROR_NOMEMORY;
      }
      break;
    }
    if ((rv = admit(S, r)) != GOOSEBERRY_SUCCESS) {
      if (rv != GOOSEBERRY_ERROR_INVALIDINPUT) return rv;
      /* the input itself is refused: report the request, go on */
      if ((rv = push_pending(S, r, GOOSEBERRY_REQUEST_REJECTED)) != GOOSEBERRY_SUCCESS)
        return rv;
      S->qhead++;
      free(r);
      continue;
    }
    S->in_use += r->footprint;
    S->qhead++;
  }
  /* 2. one batch of steps for everything admitted */
  uint32_t running = 0;
  if ((rv = gooseberry_batch_run(S->batch, steps, &running)) != GOOSEBERRY_SUCCESS)
    return rv;
  /* 3. report what ended outside the steps, then retire finished solves */
  uint32_t nd = 0;
  for (size_t i = 0; i < S->npending; ++i) {
    if (done && nd < max_done) done[nd] = S->pending[i];
    ++nd;
  }
  S->npending = 0;
  double t = now_s();
  for (uint32_t i = 0; i < S->max_active; ++i) {
    req_t *r = S->slots[i];
    if (!r) continue;
    gooseberry_cgd_info_t info;
    gooseberry_batch_info(S->batch, i, &info);
    if (!info.status) continue;
    r->res.status = info.status;
    r->res.niter = info.niter;
    r->res.nrefine = info.nrefine;
    r->res.rel_residual = info.bnorm > 0 ? info.rnorm / info.bnorm : 0.0;
    r->res.done_s = t;
    if (done && nd < max_done) done[nd] = r->res;
    ++nd;
    gooseberry_batch_remove(S->batch, i);
    S->slots[i] = NULL;
    S->nactive--;
    r->state = REQ_DONE;
    if (r->generated && S->keep) {
      /* keep x (and its footprint) until take_result */
      if (S->nkept == S->capkept) {
        size_t nc = S->capkept ? 2 * S->capkept : 64;
        req_t **nk = realloc(S->kept, nc * sizeof(req_t *));
        if (!nk) return GOOSEBERRY_ERROR_NOMEMORY;
        S->kept = nk;
        S->capkept = nc;
      }
      S->kept[S->nkept++] = r;
    } else {
      S->in_use -= r->footprint;
      if (r->generated) req_release(r);
      else free(r);
    }
  }
  if (r_ndone) *r_ndone = nd;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_server_status(gooseberry_server_t S, uint32_t *r_active,
    uint32_t *r_queued, size_t *r_mem)
{
  GB_API_GUARD();
  if (!S) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (r_active) *r_active = S->nactive;
  if (r_queued) *r_queued = (uint32_t)(S->qtail - S->qhead);
  if (r_mem) *r_mem = S->in_use;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_server_take_result(gooseberry_server_t S, uint64_t id,
    real_t *h_x, size_t n)
{
  GB_API_GUARD();
  if (!S) return GOOSEBERRY_ERROR_INVALIDINPUT;
  for (size_t i = 0; i < S->nkept; ++i) {
    req_t *r = S->kept[i];
    if (r->id != id) continue;
    if (n != r->n) return GOOSEBERRY_ERROR_INVALIDINPUT;
    int rv = h_x ? gbd_dtoh(h_x, r->own_x, n * 8) : GOOSEBERRY_SUCCESS;
    S->in_use -= r->footprint;
    S->kept[i] = S->kept[--S->nkept];
    req_release(r);
    return rv;
  }
  return GOOSEBERRY_ERROR_INVALIDINPUT;
}

/* ------------------------------------------------ explicitly typed FP32 */

int gooseberry_matrix_spmv_f32(gooseberry_matrix_t mat, const float *d_x,
    float *d_y)
{
  GB_API_GUARD();
  if (!mat || !d_x || !d_y) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_matrix *m = gb_matrix_dev(mat);
  int rv = gbd_matrix_make_f32(m);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return gbd_spmv_f32(m, d_x, d_y);
}

int gooseberry_matrix_cgd_mixed(gooseberry_matrix_t mat, const real_t *d_b,
    real_t *d_x, real_t tol, int tol_mode, size_t niter,
    gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  if (!mat || !d_b || !d_x) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gooseberry_batch_t B = NULL;
  uint32_t slot = 0;
  int rv = gooseberry_batch_create(1, 0.0, &B);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gooseberry_batch_add(B, mat, d_b, d_x, tol, tol_mode, niter,
        GOOSEBERRY_PRECISION_MIXED, &slot);
  if (rv == GOOSEBERRY_SUCCESS) rv = gooseberry_batch_solve(B);
  gooseberry_cgd_info_t local;
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gooseberry_batch_info(B, slot, info ? info : &local);
  if (B) gooseberry_batch_free(B);
  return rv;
}

/* the reference's float build (real_t = float): the SpMV in FP32 with
 * host pointers, B column-major, C row-major, as gooseberry_spmult */
int gooseberry_spmult_f32(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const float *arowval,
    const float *bcolval, float *crowval, const dim_t *blocks, dim_t nblocks)
{
  GB_API_GUARD();
  dim_t rlo = 0, rhi = anrows;
  if (blocks) {
    for (dim_t c = 0; c < nblocks; ++c)
      if (blocks[c + 1] < blocks[c]) return GOOSEBERRY_ERROR_INVALIDINPUT;
    if (nblocks && blocks[nblocks] > anrows) return GOOSEBERRY_ERROR_INVALIDINPUT;
    rlo = nblocks ? blocks[0] : 0;
    rhi = nblocks ? blocks[nblocks] : 0;
  }
  if (anrows == 0 || bncols == 0 || rlo >= rhi) return GOOSEBERRY_SUCCESS;
  if (!arowptr || !bcolval || !crowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* rows [rlo, rhi) only, row pointers rebased to 0, one-call plan: as
   * gooseberry_spmult */
  const dim_t nr = rhi - rlo;
  const ind_t base = arowptr[rlo];
  if (arowptr[rhi] < base) return GOOSEBERRY_ERROR_INVALIDINPUT;
  const size_t nnz = arowptr[rhi] - base;
  if (nnz && !arowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* FP64 copies of FP32 values are exact, and so is their FP32 copy back */
  real_t *v64 = malloc((nnz ? nnz : 1) * sizeof(real_t));
  float *col = malloc((size_t)nr * sizeof(float));
  ind_t *rp = malloc(((size_t)nr + 1) * sizeof(ind_t));
  if (!v64 || !col || !rp) {
    free(v64); free(col); free(rp);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  int rv = GOOSEBERRY_SUCCESS;
  for (dim_t i = 0; i <= nr; ++i) {
    if (arowptr[rlo + i] < base) rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    rp[i] = arowptr[rlo + i] - base;
  }
  for (size_t k = 0; k < nnz; ++k) v64[k] = arowval[base + k];
  gooseberry_plan_t once;
  memset(&once, 0, sizeof(once));
  once.wide_index = 1;
  once.wide_values = 1;
  gooseberry_matrix_t mat = NULL;
  void *dx = NULL, *dy = NULL;
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gooseberry_matrix_create(nr, ancols, rp, arowind ? arowind + base : NULL,
        v64, &once, &mat);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dx, (size_t)ancols * sizeof(float));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dy, (size_t)nr * sizeof(float));
  for (dim_t j = 0; rv == GOOSEBERRY_SUCCESS && j < bncols; ++j) {
    rv = gbd_htod(dx, bcolval + (size_t)j * ancols, (size_t)ancols * sizeof(float));
    if (rv == GOOSEBERRY_SUCCESS) rv = gooseberry_matrix_spmv_f32(mat, dx, dy);
    if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(col, dy, (size_t)nr * sizeof(float));
    for (dim_t i = 0; rv == GOOSEBERRY_SUCCESS && i < nr; ++i)
      crowval[((size_t)rlo + i) * bncols + j] = col[i];
  }
  gbd_free(dx);
  gbd_free(dy);
  gooseberry_matrix_free(mat);
  free(v64);
  free(col);
  free(rp);
  return rv;
}
