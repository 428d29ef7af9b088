/*
 * gb_api.c -- public C entry points of libgooseberry (include/gooseberry.h)
 * on top of the thin device ABI (gb_device.h).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <inttypes.h>
#include <stdlib.h>
#include <string.h>

#include "gooseberry.h"
#include "../gb_device.h"
#include "gb_host.h"

struct gooseberry_matrix_s {
  gbd_matrix *m;
  gbd_cg *cg;
};

static pthread_mutex_t g_api_mu;
static pthread_once_t g_api_once = PTHREAD_ONCE_INIT;

static void gb_api_mu_init(void)
{
  pthread_mutexattr_t a;
  pthread_mutexattr_init(&a);
  pthread_mutexattr_settype(&a, PTHREAD_MUTEX_RECURSIVE);
  pthread_mutex_init(&g_api_mu, &a);
  pthread_mutexattr_destroy(&a);
}

void gb_api_lock(void)
{
  pthread_once(&g_api_once, gb_api_mu_init);
  pthread_mutex_lock(&g_api_mu);
}

void gb_api_unlock(int *scope)
{
  (void)scope;
  pthread_mutex_unlock(&g_api_mu);
}

void gooseberry_free(void *ptr) { free(ptr); }

const char *gooseberry_strerror(int err)
{
  switch (err) {
    case GOOSEBERRY_SUCCESS: return "success";
    case GOOSEBERRY_ERROR_UNIMPLEMENTED: return "unimplemented";
    case GOOSEBERRY_ERROR_INVALIDINPUT: return "invalid input";
    case GOOSEBERRY_ERROR_READ: return "read error";
    case GOOSEBERRY_ERROR_WRITE: return "write error";
    case GOOSEBERRY_ERROR_FILENOTFOUND: return "file not found";
    case GOOSEBERRY_ERROR_PERMISSIONDENIED: return "permission denied";
    case GOOSEBERRY_ERROR_UNKNOWN: return "unknown error";
    case GOOSEBERRY_ERROR_DEVICE: return "device (HIP) error";
    case GOOSEBERRY_ERROR_NOMEMORY: return "out of device memory";
    case GOOSEBERRY_ERROR_BREAKDOWN: return "CG breakdown";
    default: return "unrecognised error code";
  }
}

const char *gooseberry_last_device_error(void) { return gbd_last_error(); }

int gooseberry_inject_fault(int kind, long long after)
{
  return gbd_inject_fault(kind, after);
}

/* --------------------------------------------------------- device memory */

int gooseberry_device_alloc(void **d_ptr, size_t bytes)
{
  GB_API_GUARD();
  if (!d_ptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gbd_malloc(d_ptr, bytes);
}

int gooseberry_device_free(void *d_ptr)
{
  GB_API_GUARD();
  return gbd_free(d_ptr);
}

int gooseberry_memcpy_htod(void *d_dst, const void *h_src, size_t bytes)
{
  GB_API_GUARD();
  return gbd_htod(d_dst, h_src, bytes);
}

int gooseberry_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes)
{
  GB_API_GUARD();
  return gbd_dtoh(h_dst, d_src, bytes);
}

int gooseberry_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes)
{
  GB_API_GUARD();
  return gbd_dtod(d_dst, d_src, bytes);
}

int gooseberry_device_synchronize(void)
{
  GB_API_GUARD();
  return gbd_sync();
}

int gooseberry_device_memory(size_t *free_bytes, size_t *total_bytes)
{
  GB_API_GUARD();
  return gbd_device_memory(free_bytes, total_bytes);
}

int gooseberry_event_create(void **ev)
{
  GB_API_GUARD();
  return gbd_event_create(ev);
}
int gooseberry_event_destroy(void *ev)
{
  GB_API_GUARD();
  return gbd_event_destroy(ev);
}
int gooseberry_event_record(void *ev)
{
  GB_API_GUARD();
  return gbd_event_record(ev);
}
int gooseberry_event_elapsed_ms(void *a, void *b, float *ms)
{
  GB_API_GUARD();
  return gbd_event_elapsed_ms(a, b, ms);
}

/* upload a host array into a fresh device buffer */
int gb_upload(void **d, const void *h, size_t bytes)
{
  int rv = gbd_malloc(d, bytes);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  rv = gbd_htod(*d, h, bytes);
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(*d);
    *d = NULL;
  }
  return rv;
}

/* --------------------------------------------------------------- scale */

int gooseberry_scale(ind_t n, real_t *val, real_t s)
{
  GB_API_GUARD();
  if (n == 0) return GOOSEBERRY_SUCCESS;
  if (!val) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *d = NULL;
  int rv = gb_upload(&d, val, (size_t)n * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_scale(d, n, s);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(val, d, (size_t)n * sizeof(real_t));
  gbd_free(d);
  return rv;
}

int gooseberry_add_scalar(ind_t n, real_t *val, real_t s)
{
  GB_API_GUARD();
  if (n == 0) return GOOSEBERRY_SUCCESS;
  if (!val) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *d = NULL;
  int rv = gb_upload(&d, val, (size_t)n * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_add_scalar(d, n, s);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(val, d, (size_t)n * sizeof(real_t));
  gbd_free(d);
  return rv;
}

/* ------------------------------------------------------------ matrices */

static void plan_to_dev(const gooseberry_plan_t *p, const gbd_matrix *m,
    gbd_plan_params *out, int *have)
{
  *have = 0;
  if (!p) return;
  gbd_plan_params d;
  gbd_plan_defaults(NULL, 0, 0, &d);
  (void)m;
  if (p->short_max) d.short_max = p->short_max;
  if (p->wave_max) d.wave_max = p->wave_max;
  if (p->window) d.window = p->window;
  if (p->rowcap) d.rowcap = p->rowcap;
  d.exact = p->exact ? 1 : 0;
  if (p->grid) d.grid = p->grid;
  d.xcd_remap = p->xcd_remap;
  d.csplit = p->csplit > 1 ? p->csplit : 0;
  d.wide_index = p->wide_index ? 1 : 0;
  d.wide_values = p->wide_values ? 1 : 0;
  d.xtile = p->xtile;
  *out = d;
  *have = 1;
}

/* validate a host CSR: monotone rowptr starting at 0, indices < ncols */
int gb_check_csr(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind)
{
  if (!rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (rowptr[0] != 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  for (dim_t i = 0; i < nrows; ++i) {
    if (rowptr[i + 1] < rowptr[i]) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  ind_t nnz = rowptr[nrows];
  if (nnz && !rowind) return GOOSEBERRY_ERROR_INVALIDINPUT;
  for (ind_t k = 0; k < nnz; ++k) {
    if (rowind[k] >= ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  return GOOSEBERRY_SUCCESS;
}

/* what a matrix held as row parts (64-bit row pointers, more than 2^31
 * entries) does not take */
int gb_parts_unsupported(void)
{
  gbd_set_error("this operation does not take a matrix held as row parts (64-bit row "
                "pointers)");
  return GOOSEBERRY_ERROR_UNIMPLEMENTED;
}

static int wrap_matrix(gbd_matrix *m, gooseberry_matrix_t *r_mat)
{
  struct gooseberry_matrix_s *h = calloc(1, sizeof(*h));
  if (!h) {
    gbd_matrix_destroy(m);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  h->m = m;
  *r_mat = h;
  return GOOSEBERRY_SUCCESS;
}

/* wrap a CSR the library built itself on the device (valid by
 * construction, nnz known; no copy); the arrays are freed here when the
 * matrix cannot be made */
int gb_matrix_adopt(dim_t nrows, dim_t ncols, ind_t *d_rowptr,
    dim_t *d_rowind, real_t *d_rowval, ind_t nnz,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  gbd_plan_params pp;
  int have;
  plan_to_dev(plan, NULL, &pp, &have);
  gbd_matrix *m = NULL;
  int rv = gbd_matrix_create_trusted(nrows, ncols, d_rowptr, d_rowind,
      d_rowval, nnz, have ? &pp : NULL, &m);
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(d_rowptr); gbd_free(d_rowind); gbd_free(d_rowval);
    return rv;
  }
  return wrap_matrix(m, r_mat);
}

/* A matrix from host CSR arrays: the three arrays, and one more array when
 * extra_bytes is nonzero (left in *d_extra), uploaded in one staged
 * pipeline into fresh device buffers.  Only rowptr[0] is checked here; the
 * rest of the structure (non-decreasing row pointers, columns below ncols)
 * is checked on the device once the arrays are there: a scan of the host
 * arrays took about as long as their upload. */
static int matrix_from_host(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval, const gooseberry_plan_t *plan,
    const void *extra, size_t extra_bytes, void **d_extra,
    gooseberry_matrix_t *r_mat)
{
  if (!r_mat || !rowptr || rowptr[0] != 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  size_t nnz = rowptr[nrows];
  if (nnz && (!rowind || !rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *d[4] = {NULL, NULL, NULL, NULL};
  const void *h[4] = {rowptr, rowind, rowval, extra};
  size_t b[4] = {((size_t)nrows + 1) * sizeof(ind_t), nnz * sizeof(dim_t),
                 nnz * sizeof(real_t), extra_bytes};
  uint32_t n = extra_bytes ? 4 : 3, k;
  int rv = GOOSEBERRY_SUCCESS;
  for (k = 0; k < n && rv == GOOSEBERRY_SUCCESS; ++k) rv = gbd_malloc(&d[k], b[k]);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_htod_many(n, d, h, b);
  gbd_matrix *m = NULL;
  if (rv == GOOSEBERRY_SUCCESS) {
    gbd_plan_params pp;
    int have;
    plan_to_dev(plan, NULL, &pp, &have);
    rv = gbd_matrix_create(nrows, ncols, d[0], d[1], d[2], have ? &pp : NULL, &m);
  }
  if (rv != GOOSEBERRY_SUCCESS) {
    for (k = 0; k < 4; ++k) gbd_free(d[k]);
    return rv;
  }
  rv = wrap_matrix(m, r_mat);  /* frees the matrix and its arrays on failure */
  if (rv != GOOSEBERRY_SUCCESS || !d_extra) gbd_free(d[3]);
  else *d_extra = d[3];
  return rv;
}

int gooseberry_matrix_create(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval, const gooseberry_plan_t *plan,
    gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  return matrix_from_host(nrows, ncols, rowptr, rowind, rowval, plan, NULL, 0, NULL,
      r_mat);
}

int gooseberry_matrix_create_device(dim_t nrows, dim_t ncols,
    const ind_t *d_rowptr, const dim_t *d_rowind, const real_t *d_rowval,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  if (!r_mat || !d_rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  ind_t nnz = 0;
  int rv = gbd_dtoh(&nnz, d_rowptr + nrows, sizeof(ind_t));
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  void *dp = NULL, *di = NULL, *dv = NULL;
  rv = gbd_malloc(&dp, ((size_t)nrows + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&di, (size_t)nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dv, (size_t)nnz * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtod(dp, d_rowptr, ((size_t)nrows + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtod(di, d_rowind, (size_t)nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtod(dv, d_rowval, (size_t)nnz * sizeof(real_t));
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(dp); gbd_free(di); gbd_free(dv);
    return rv;
  }
  gbd_plan_params pp;
  int have;
  plan_to_dev(plan, NULL, &pp, &have);
  gbd_matrix *m = NULL;
  rv = gbd_matrix_create(nrows, ncols, dp, di, dv, have ? &pp : NULL, &m);
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(dp); gbd_free(di); gbd_free(dv);
    return rv;
  }
  return wrap_matrix(m, r_mat);
}

int gooseberry_matrix_free(gooseberry_matrix_t mat)
{
  GB_API_GUARD();
  if (!mat) return GOOSEBERRY_SUCCESS;
  /* a batch or server still reads the matrix: remove its solve first */
  if (gbd_matrix_holders(mat->m)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_cg_destroy(mat->cg);
  gbd_matrix_destroy(mat->m);
  free(mat);
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_matrix_replan(gooseberry_matrix_t mat,
    const gooseberry_plan_t *plan)
{
  GB_API_GUARD();
  if (!mat) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (gbd_matrix_holders(mat->m)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_plan_params pp;
  int have;
  plan_to_dev(plan, mat->m, &pp, &have);
  /* the CG workspace is sized by the launch grid: rebuild it lazily */
  gbd_cg_destroy(mat->cg);
  mat->cg = NULL;
  return gbd_matrix_replan(mat->m, have ? &pp : NULL);
}

int gooseberry_matrix_info(gooseberry_matrix_t mat,
    gooseberry_matrix_info_t *info)
{
  GB_API_GUARD();
  if (!mat || !info) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_plan_params p;
  memset(info, 0, sizeof(*info));
  gbd_matrix_info(mat->m, &info->nrows, &info->ncols, &info->nnz, &p,
      &info->nblocks_stream, &info->nrows_wave, &info->nrows_block,
      info->hist);
  info->plan.short_max = p.short_max;
  info->plan.wave_max = p.wave_max;
  info->plan.window = p.window;
  info->plan.rowcap = p.rowcap;
  info->plan.exact = p.exact;
  info->plan.grid = p.grid;
  info->plan.xcd_remap = p.xcd_remap;
  info->plan.csplit = p.csplit;
  info->plan.wide_index = p.wide_index;
  info->plan.wide_values = p.wide_values;
  info->plan.xtile = p.xtile;
  info->index16 = gbd_matrix_index16(mat->m);
  info->values8 = gbd_matrix_values8(mat->m);
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_matrix_spmv(gooseberry_matrix_t mat, const real_t *d_x,
    real_t *d_y)
{
  GB_API_GUARD();
  if (!mat || !d_x || !d_y) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gbd_spmv(mat->m, d_x, d_y);
}

int gooseberry_matrix_spmm(gooseberry_matrix_t mat, dim_t bncols,
    const real_t *d_b, real_t *d_c)
{
  GB_API_GUARD();
  if (!mat || !d_b || !d_c) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gbd_spmm(mat->m, bncols, d_b, d_c);
}

gbd_matrix *gb_matrix_dev(gooseberry_matrix_t mat) { return mat->m; }

/* -------------------------------------------------------------- spmult */

/* the plan of a matrix made for one host-pointer product: no 16-bit
 * offsets and no value table, whose build costs more than one SpMV saves */
static void spmult_once_plan(gooseberry_plan_t *p)
{
  memset(p, 0, sizeof(*p));
  p->wide_index = 1;
  p->wide_values = 1;
}

/* C rows = A B for a matrix made for this one call (freed here, with db):
 * B (ancols x bncols, column-major, uploaded with A into db) in, the nr
 * rows of C (row-major) out */
static int spmult_once(gooseberry_matrix_t mat, dim_t nr, dim_t bncols, void *db,
    real_t *crows)
{
  void *dc = NULL;
  size_t cbytes = (size_t)nr * bncols * sizeof(real_t);
  int rv = gbd_malloc(&dc, cbytes);
  if (rv == GOOSEBERRY_SUCCESS) {
    rv = (bncols == 1) ? gbd_spmv(mat->m, db, dc)
                       : gbd_spmm(mat->m, bncols, db, dc);
  }
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(crows, dc, cbytes);
  gbd_free(db);
  gbd_free(dc);
  gooseberry_matrix_free(mat);
  return rv;
}

/* blocks (nblocks + 1 non-decreasing row boundaries, the last at most
 * anrows) to the row range the reference computes */
static int spmult_rows(dim_t anrows, const dim_t *blocks, dim_t nblocks,
    dim_t *rlo, dim_t *rhi)
{
  *rlo = 0;
  *rhi = anrows;
  if (!blocks) return GOOSEBERRY_SUCCESS;
  for (dim_t c = 0; c < nblocks; ++c) {
    if (blocks[c + 1] < blocks[c]) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (nblocks && blocks[nblocks] > anrows) return GOOSEBERRY_ERROR_INVALIDINPUT;
  *rlo = nblocks ? blocks[0] : 0;
  *rhi = nblocks ? blocks[nblocks] : 0;
  return GOOSEBERRY_SUCCESS;
}

/* src/blas.c:552-606 through the GPU.  With blocks, the reference computes
 * only the rows inside [blocks[0], blocks[nblocks]) (the union of the
 * blocks of a non-decreasing blocks array); other rows of crowval are left
 * untouched, as there.  Only those rows are uploaded and planned, their row
 * pointers rebased to start at 0 (so a rowptr that does not start at 0 is
 * read as the reference reads it).  The matrix lives for this one call, so
 * its plan skips the 16-bit column offsets and the value table: building
 * them costs more than one SpMV saves (the products are the same). */
int gooseberry_spmult(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const real_t *bcolval, real_t *crowval, const dim_t *blocks,
    dim_t nblocks)
{
  GB_API_GUARD();
  dim_t rlo, rhi;
  if (spmult_rows(anrows, blocks, nblocks, &rlo, &rhi) != GOOSEBERRY_SUCCESS)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (anrows == 0 || bncols == 0 || rlo >= rhi) return GOOSEBERRY_SUCCESS;
  if (!arowptr || !bcolval || !crowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  const dim_t nr = rhi - rlo;
  const ind_t base = arowptr[rlo];
  const ind_t *rp = arowptr + rlo;
  ind_t *rebased = NULL;
  if (base) {
    rebased = malloc(((size_t)nr + 1) * sizeof(ind_t));
    if (!rebased) return GOOSEBERRY_ERROR_NOMEMORY;
    for (dim_t i = 0; i <= nr; ++i) {
      if (arowptr[rlo + i] < base) {
        free(rebased);
        return GOOSEBERRY_ERROR_INVALIDINPUT;
      }
      rebased[i] = arowptr[rlo + i] - base;
    }
    rp = rebased;
  }
  gooseberry_plan_t once;
  spmult_once_plan(&once);
  gooseberry_matrix_t mat = NULL;
  void *db = NULL;
  /* A and B in one upload */
  int rv = matrix_from_host(nr, ancols, rp,
      arowind ? arowind + base : NULL, arowval ? arowval + base : NULL,
      &once, bcolval, (size_t)ancols * bncols * sizeof(real_t), &db, &mat);
  free(rebased);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return spmult_once(mat, nr, bncols, db, crowval + (size_t)rlo * bncols);
}

/* ------------------------------------------------------ sparse x sparse */

/* blocks given: monotone row ranges; the dense product writes rows
 * [blocks[0], blocks[nblocks]) as the reference's block loop does */
static int spgemm_blocks(dim_t anrows, const dim_t *blocks, dim_t nblocks,
    dim_t *rlo, dim_t *rhi)
{
  *rlo = 0;
  *rhi = anrows;
  if (!blocks) return GOOSEBERRY_SUCCESS;
  for (dim_t c = 0; c < nblocks; ++c) {
    if (blocks[c + 1] < blocks[c]) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (nblocks && blocks[nblocks] > anrows) return GOOSEBERRY_ERROR_INVALIDINPUT;
  *rlo = nblocks ? blocks[0] : 0;
  *rhi = nblocks ? blocks[nblocks] : 0;
  return GOOSEBERRY_SUCCESS;
}

/* A (anrows x ancols) and B (ancols x bncols) checked and uploaded */
static int spgemm_upload(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const ind_t *browptr, const dim_t *browind, const real_t *browval,
    void *d[6])
{
  int rv = gb_check_csr(anrows, ancols, arowptr, arowind);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gb_check_csr(ancols, bncols, browptr, browind);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  const size_t annz = arowptr[anrows], bnnz = browptr[ancols];
  if ((annz && !arowval) || (bnnz && !browval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  rv = gb_upload(&d[0], arowptr, ((size_t)anrows + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&d[1], arowind, annz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&d[2], arowval, annz * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gb_upload(&d[3], browptr, ((size_t)ancols + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&d[4], browind, bnnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&d[5], browval, bnnz * sizeof(real_t));
  return rv;
}

static void spgemm_free(void *d[6])
{
  for (int i = 0; i < 6; ++i) gbd_free(d[i]);
}

int gooseberry_spmultsp(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const ind_t *browptr, const dim_t *browind, const real_t *browval,
    real_t *crowval, const dim_t *blocks, dim_t nblocks)
{
  GB_API_GUARD();
  dim_t rlo, rhi;
  int rv = spgemm_blocks(anrows, blocks, nblocks, &rlo, &rhi);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (anrows == 0 || bncols == 0 || rlo >= rhi) return GOOSEBERRY_SUCCESS;
  if (!crowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *d[6] = {NULL, NULL, NULL, NULL, NULL, NULL}, *dc = NULL;
  rv = spgemm_upload(anrows, ancols, bncols, arowptr, arowind, arowval,
      browptr, browind, browval, d);
  const size_t cbytes = (size_t)anrows * bncols * sizeof(real_t);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dc, cbytes);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_spgemm(anrows, ancols, bncols, d[0], d[1], d[2], d[3], d[4],
        d[5], dc, NULL, NULL, NULL, NULL);
  if (rv == GOOSEBERRY_SUCCESS) {
    const size_t off = (size_t)rlo * bncols;
    rv = gbd_dtoh(crowval + off, (real_t *)dc + off,
        (size_t)(rhi - rlo) * bncols * sizeof(real_t));
  }
  gbd_free(dc);
  spgemm_free(d);
  return rv;
}

int gooseberry_spmultsp_sp(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const ind_t *browptr, const dim_t *browind, const real_t *browval,
    ind_t **r_crowptr, dim_t **r_crowind, real_t **r_crowval,
    const dim_t *blocks, dim_t nblocks)
{
  GB_API_GUARD();
  if (!r_crowptr || !r_crowind || !r_crowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* the reference fills every row whatever the blocks, and its row
   * pointers are only whole when the blocks partition the rows */
  if (blocks) {
    dim_t rlo, rhi;
    int brv = spgemm_blocks(anrows, blocks, nblocks, &rlo, &rhi);
    if (brv != GOOSEBERRY_SUCCESS) return brv;
    if (rlo != 0 || rhi != anrows) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  void *d[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  uint32_t *dptr = NULL, *dind = NULL;
  double *dval = NULL;
  uint32_t cnnz = 0;
  ind_t *cptr = NULL;
  dim_t *cind = NULL;
  real_t *cval = NULL;
  int rv = spgemm_upload(anrows, ancols, bncols, arowptr, arowind, arowval,
      browptr, browind, browval, d);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_spgemm(anrows, ancols, bncols, d[0], d[1], d[2], d[3], d[4],
        d[5], NULL, &dptr, &dind, &dval, &cnnz);
  if (rv == GOOSEBERRY_SUCCESS) {
    cptr = malloc(((size_t)anrows + 1) * sizeof(ind_t));
    cind = malloc((cnnz ? cnnz : 1) * sizeof(dim_t));
    cval = malloc((cnnz ? cnnz : 1) * sizeof(real_t));
    if (!cptr || !cind || !cval) rv = GOOSEBERRY_ERROR_NOMEMORY;
  }
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtoh(cptr, dptr, ((size_t)anrows + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(cind, dind, (size_t)cnnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(cval, dval, (size_t)cnnz * sizeof(real_t));
  gbd_free(dptr);
  gbd_free(dind);
  gbd_free(dval);
  spgemm_free(d);
  if (rv != GOOSEBERRY_SUCCESS) {
    free(cptr);
    free(cind);
    free(cval);
    return rv;
  }
  *r_crowptr = cptr;
  *r_crowind = cind;
  *r_crowval = cval;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_matrix_spmultsp(gooseberry_matrix_t a, gooseberry_matrix_t b,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_c)
{
  GB_API_GUARD();
  if (!a || !b || !r_c) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t anr, anc, bnr, bnc;
  gbd_matrix_info(a->m, &anr, &anc, NULL, NULL, NULL, NULL, NULL, NULL);
  gbd_matrix_info(b->m, &bnr, &bnc, NULL, NULL, NULL, NULL, NULL, NULL);
  if (bnr < anc) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (gbd_matrix_nparts(a->m) || gbd_matrix_nparts(b->m)) return gb_parts_unsupported();
  uint32_t *dptr = NULL, *dind = NULL, cnnz = 0;
  double *dval = NULL;
  int rv = gbd_spgemm(anr, anc, bnc, gbd_matrix_rowptr(a->m),
      gbd_matrix_rowind(a->m), gbd_matrix_rowval(a->m),
      gbd_matrix_rowptr(b->m), gbd_matrix_rowind(b->m),
      gbd_matrix_rowval(b->m), NULL, &dptr, &dind, &dval, &cnnz);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  /* without a plan: the defaults but the value index, whose build reads
   * C's values twice (0.36 ms of a 3.7 ms product on the 150^3 Laplacian's
   * square) and pays off only over many SpMVs; a replan adds it */
  gooseberry_plan_t product;
  if (!plan) {
    memset(&product, 0, sizeof(product));
    product.wide_values = 1;
    plan = &product;
  }
  return gb_matrix_adopt(anr, bnc, dptr, dind, dval, cnnz, plan, r_c);
}

int gooseberry_matrix_spmultsp_dense(gooseberry_matrix_t a,
    gooseberry_matrix_t b, real_t *d_c)
{
  GB_API_GUARD();
  if (!a || !b) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t anr, anc, bnr, bnc;
  gbd_matrix_info(a->m, &anr, &anc, NULL, NULL, NULL, NULL, NULL, NULL);
  gbd_matrix_info(b->m, &bnr, &bnc, NULL, NULL, NULL, NULL, NULL, NULL);
  if (bnr < anc) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if ((size_t)anr * bnc == 0) return GOOSEBERRY_SUCCESS;
  if (!d_c) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (gbd_matrix_nparts(a->m) || gbd_matrix_nparts(b->m)) return gb_parts_unsupported();
  return gbd_spgemm(anr, anc, bnc, gbd_matrix_rowptr(a->m),
      gbd_matrix_rowind(a->m), gbd_matrix_rowval(a->m),
      gbd_matrix_rowptr(b->m), gbd_matrix_rowind(b->m),
      gbd_matrix_rowval(b->m), d_c, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------- Cuthill-McKee */

static int cm_blocks_ok(dim_t nrows, const dim_t *blocks, dim_t nblocks)
{
  if (!blocks) return 1;
  for (dim_t c = 0; c < nblocks; ++c)
    if (blocks[c + 1] < blocks[c]) return 0;
  return nblocks == 0 || blocks[nblocks] <= nrows;
}

int gooseberry_permute_cuthillmckee(dim_t nrows, dim_t ncols,
    const ind_t *rowptr, const dim_t *rowind, const real_t *rowval,
    const dim_t *blocks, dim_t nblocks, dim_t *perm)
{
  GB_API_GUARD();
  (void)rowval;
  if (nrows != ncols) {
    fprintf(stderr, "Cuthill-McKee requires a structurally symmetric matrix: "
        "Given %" PRIu32 "x%" PRIu32 "\n", nrows, ncols);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (!perm || !cm_blocks_ok(nrows, blocks, nblocks))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv = gb_check_csr(nrows, ncols, rowptr, rowind);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (nrows == 0) return GOOSEBERRY_SUCCESS;
  void *drp = NULL, *dri = NULL, *dp = NULL;
  const size_t nnz = rowptr[nrows];
  rv = gb_upload(&drp, rowptr, ((size_t)nrows + 1) * sizeof(ind_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&dri, rowind, nnz * sizeof(dim_t));
  /* positions outside the blocks keep the caller's values */
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&dp, perm, (size_t)nrows * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_cuthillmckee(nrows, drp, dri, blocks, nblocks, dp);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtoh(perm, dp, (size_t)nrows * sizeof(dim_t));
  gbd_free(drp);
  gbd_free(dri);
  gbd_free(dp);
  return rv;
}

int gooseberry_matrix_cuthillmckee(gooseberry_matrix_t mat,
    const dim_t *blocks, dim_t nblocks, dim_t *d_perm)
{
  GB_API_GUARD();
  if (!mat || !d_perm) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t nrows, ncols;
  gbd_matrix_info(mat->m, &nrows, &ncols, NULL, NULL, NULL, NULL, NULL, NULL);
  if (nrows != ncols || !cm_blocks_ok(nrows, blocks, nblocks))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (gbd_matrix_nparts(mat->m)) return gb_parts_unsupported();
  return gbd_cuthillmckee(nrows, gbd_matrix_rowptr(mat->m),
      gbd_matrix_rowind(mat->m), blocks, nblocks, d_perm);
}

/* ------------------------------------------------------------------- CG */

/* Iterations are enqueued in batches replayed from captured hipGraphs and
 * the host polls the device state between batches.  The device stops by
 * itself at the reference's stopping point; iterations launched after it
 * cost time, so the next batch is sized from the residual decrease seen in
 * the last one (a little short of the predicted remainder). */
#define GB_CG_FIRST 8
#define GB_CG_MAXBATCH 64
#define GB_CG_TRACE 64

static uint32_t gb_cg_next_batch(uint32_t k, uint64_t it0, double r0,
    const gbd_cg_state *st)
{
  uint64_t done_its = st->iter - it0;
  if (done_its == 0 || !(r0 > 0) || !(st->rnorm > 0) || !(st->rnorm < r0) ||
      !(st->tol > 0)) {
    k *= 2;
    return k > GB_CG_MAXBATCH ? GB_CG_MAXBATCH : k;
  }
  double rate = log(st->rnorm / r0) / (double)done_its;   /* < 0 */
  double need = log(st->tol / st->rnorm) / rate;          /* > 0 */
  double want = 0.8 * need;
  if (!(want >= 1.0)) return 1;
  if (want > GB_CG_MAXBATCH) return GB_CG_MAXBATCH;
  return (uint32_t)want;
}

/* The solve loop on a workspace: start, then batches of iterations (graph
 * replays, or single launches for a workspace in the scratch array) and a
 * poll after each. */
static int cgd_loop(gbd_cg *cg, const real_t *d_b, real_t *d_x, real_t tol,
    int tol_mode, size_t niter, FILE *flog, gooseberry_cgd_info_t *info)
{
  /* the per-iteration residuals are kept on the device only for a log */
  int rv = gbd_cg_start(cg, d_b, d_x, tol,
      tol_mode == GOOSEBERRY_TOL_RELATIVE, niter, flog ? GB_CG_TRACE : 0);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  gbd_cg_state st;
  double trace[GB_CG_TRACE];
  uint64_t printed = 0;
  rv = gbd_cg_poll(cg, &st, NULL);
  uint32_t k = GB_CG_FIRST;
  while (rv == GOOSEBERRY_SUCCESS && !st.done) {
    const uint64_t it0 = st.iter;
    const double r0 = st.rnorm;
    rv = gbd_cg_enqueue_graph(cg, k);
    if (rv != GOOSEBERRY_SUCCESS) break;
    rv = gbd_cg_poll(cg, &st, flog ? trace : NULL);
    if (rv != GOOSEBERRY_SUCCESS) break;
    if (flog) {
      for (; printed < st.iter; ++printed) {
        /* src/cgd.c:116 */
        fprintf(flog, "ITER: %zu RMSE: %lf\n", (size_t)printed,
            trace[printed % GB_CG_TRACE]);
      }
    }
    k = gb_cg_next_batch(k, it0, r0, &st);
  }
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (info) {
    memset(info, 0, sizeof(*info));
    info->niter = st.iter;
    info->rnorm = st.rnorm;
    info->bnorm = st.rnorm0;
    info->converged = st.status == 1;
    info->status = st.status;
    info->true_rel_residual = NAN;
  }
  return st.status == 3 ? GOOSEBERRY_ERROR_BREAKDOWN : GOOSEBERRY_SUCCESS;
}

static int cgd_args_ok(gooseberry_matrix_t mat, const real_t *d_b,
    const real_t *d_x, int tol_mode)
{
  if (!mat || !d_b || !d_x) return 0;
  if (tol_mode != GOOSEBERRY_TOL_ABSOLUTE && tol_mode != GOOSEBERRY_TOL_RELATIVE)
    return 0;
  uint32_t nrows, ncols;
  gbd_matrix_info(mat->m, &nrows, &ncols, NULL, NULL, NULL, NULL, NULL, NULL);
  return nrows == ncols;
}

int gooseberry_matrix_cgd(gooseberry_matrix_t mat, const real_t *d_b,
    real_t *d_x, real_t tol, int tol_mode, size_t niter, void *log,
    gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  if (!cgd_args_ok(mat, d_b, d_x, tol_mode)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!mat->cg) {
    int rv = gbd_cg_create(mat->m, &mat->cg);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  return cgd_loop(mat->cg, d_b, d_x, tol, tol_mode, niter, (FILE *)log, info);
}

int gooseberry_matrix_cgd_noalloc(gooseberry_matrix_t mat, const real_t *d_b,
    real_t *d_x, real_t tol, int tol_mode, size_t niter,
    gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  if (!cgd_args_ok(mat, d_b, d_x, tol_mode)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_cg *cg = NULL;
  int rv = gbd_cg_scratch(mat->m, &cg);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return cgd_loop(cg, d_b, d_x, tol, tol_mode, niter, NULL, info);
}

size_t gooseberry_cgd_noalloc_max_bytes(void)
{
  return gbd_cg_scratch_max();
}

/* the host-array CG on a matrix made for this call (freed here) */
static int cgd_resident_host(gooseberry_matrix_t mat, dim_t nrows,
    const real_t *b, real_t *x, real_t error, size_t niter,
    gooseberry_cgd_info_t *info)
{
  void *db = NULL, *dx = NULL;
  size_t vb = (size_t)nrows * sizeof(real_t);
  int rv = gb_upload(&db, b, vb);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dx, vb);
  int rs = GOOSEBERRY_SUCCESS;
  if (rv == GOOSEBERRY_SUCCESS) {
    rs = gooseberry_matrix_cgd(mat, db, dx, error, GOOSEBERRY_TOL_ABSOLUTE,
        niter, NULL, info);
    if (rs == GOOSEBERRY_SUCCESS || rs == GOOSEBERRY_ERROR_BREAKDOWN)
      rv = gbd_dtoh(x, dx, vb);
    else
      rv = rs;
  }
  gbd_free(db);
  gbd_free(dx);
  gooseberry_matrix_free(mat);
  return rv != GOOSEBERRY_SUCCESS ? rv : rs;
}

int gooseberry_cgd(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval, const real_t *b, real_t *x,
    real_t error, size_t niter, gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  /* src/cgd.c:51-65 shape checks (b and x are nrows/ncols long here) */
  if (nrows != ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!b || !x) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gooseberry_matrix_t mat = NULL;
  int rv = gooseberry_matrix_create(nrows, ncols, rowptr, rowind, rowval, NULL,
      &mat);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return cgd_resident_host(mat, nrows, b, x, error, niter, info);
}

/* --------------------------------------------------- index work (host) */

int gooseberry_countingsort_kv(const dim_t *keys, const dim_t *vals,
    dim_t kmin, dim_t kmax, size_t n, dim_t *out, dim_t *dist)
{
  GB_API_GUARD();
  if (kmax < kmin || (n && (!keys || !out))) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *dk = NULL, *dv = NULL, *dout = NULL, *dd = NULL;
  size_t nd = (size_t)kmax - kmin + 2;
  int rv = gb_upload(&dk, keys, n * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS && vals) rv = gb_upload(&dv, vals, n * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dout, n * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS && dist) rv = gbd_malloc(&dd, nd * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_countingsort_kv(dk, dv, kmin, kmax, n, dout, dd);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(out, dout, n * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS && dist) rv = gbd_dtoh(dist, dd, nd * sizeof(dim_t));
  gbd_free(dk); gbd_free(dv); gbd_free(dout); gbd_free(dd);
  return rv;
}

static int check_perm(const dim_t *perm, dim_t n)
{
  if (!perm) return GOOSEBERRY_SUCCESS;
  unsigned char *seen = calloc(n ? n : 1, 1);
  if (!seen) return GOOSEBERRY_ERROR_NOMEMORY;
  int rv = GOOSEBERRY_SUCCESS;
  for (dim_t i = 0; i < n; ++i) {
    if (perm[i] >= n || seen[perm[i]]) { rv = GOOSEBERRY_ERROR_INVALIDINPUT; break; }
    seen[perm[i]] = 1;
  }
  free(seen);
  return rv;
}

int gooseberry_permute_sparse(dim_t nrows, dim_t ncols, ind_t *rowptr,
    dim_t *rowind, real_t *rowval, const dim_t *rowperm, const dim_t *colperm)
{
  GB_API_GUARD();
  int rv = gb_check_csr(nrows, ncols, rowptr, rowind);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if ((rv = check_perm(rowperm, nrows)) != GOOSEBERRY_SUCCESS) return rv;
  if ((rv = check_perm(colperm, ncols)) != GOOSEBERRY_SUCCESS) return rv;
  if (!rowperm && !colperm) return GOOSEBERRY_SUCCESS;
  size_t nnz = rowptr[nrows];
  size_t pb = ((size_t)nrows + 1) * sizeof(ind_t);
  void *dp = NULL, *di = NULL, *dv = NULL, *drp = NULL, *dcp = NULL;
  void *op = NULL, *oi = NULL, *ov = NULL;
  rv = gb_upload(&dp, rowptr, pb);
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&di, rowind, nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gb_upload(&dv, rowval, nnz * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS && rowperm)
    rv = gb_upload(&drp, rowperm, (size_t)nrows * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS && colperm)
    rv = gb_upload(&dcp, colperm, (size_t)ncols * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&op, pb);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&oi, nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&ov, nnz * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_permute_sparse(nrows, ncols, dp, di, dv, drp, dcp, op, oi, ov);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(rowptr, op, pb);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(rowind, oi, nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(rowval, ov, nnz * sizeof(real_t));
  gbd_free(dp); gbd_free(di); gbd_free(dv); gbd_free(drp); gbd_free(dcp);
  gbd_free(op); gbd_free(oi); gbd_free(ov);
  return rv;
}

int gooseberry_permute_dense(dim_t nrows, dim_t ncols, real_t *rowval,
    const dim_t *rowperm, const dim_t *colperm, int rev)
{
  GB_API_GUARD();
  int rv;
  if ((rv = check_perm(rowperm, nrows)) != GOOSEBERRY_SUCCESS) return rv;
  if ((rv = check_perm(colperm, ncols)) != GOOSEBERRY_SUCCESS) return rv;
  size_t tot = (size_t)nrows * ncols;
  if ((!rowperm && !colperm) || tot == 0) return GOOSEBERRY_SUCCESS;
  if (!rowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *din = NULL, *dout = NULL, *drp = NULL, *dcp = NULL;
  rv = gb_upload(&din, rowval, tot * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dout, tot * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS && rowperm)
    rv = gb_upload(&drp, rowperm, (size_t)nrows * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS && colperm)
    rv = gb_upload(&dcp, colperm, (size_t)ncols * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_permute_dense(nrows, ncols, din, dout, drp, dcp, rev ? 1 : 0);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(rowval, dout, tot * sizeof(real_t));
  gbd_free(din); gbd_free(dout); gbd_free(drp); gbd_free(dcp);
  return rv;
}

/* -------------------------------------------------------------- PageRank */

int gooseberry_matrix_pagerank(gooseberry_matrix_t mat, real_t damping,
    size_t niter, real_t minerror, real_t *d_out,
    gooseberry_pagerank_info_t *info)
{
  GB_API_GUARD();
  if (!mat || !d_out) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint64_t it = 0;
  double rmse = 0;
  uint32_t ns = 0;
  int rv = gbd_pagerank(mat->m, damping, niter, minerror, d_out, &it, &rmse, &ns);
  if (rv == GOOSEBERRY_SUCCESS && info) {
    info->niter = it;
    info->rmse = rmse;
    info->nsinks = ns;
  }
  return rv;
}

int gooseberry_pagerank(dim_t n, const ind_t *rowptr, const dim_t *rowind,
    const real_t *rowval, real_t damping, size_t niter, real_t minerror,
    real_t *out, gooseberry_pagerank_info_t *info)
{
  GB_API_GUARD();
  if (!out || n == 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gooseberry_matrix_t m = NULL;
  void *d = NULL;
  int rv = gooseberry_matrix_create(n, n, rowptr, rowind, rowval, NULL, &m);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&d, (size_t)n * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gooseberry_matrix_pagerank(m, damping, niter, minerror, d, info);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtoh(out, d, (size_t)n * sizeof(real_t));
  gbd_free(d);
  gooseberry_matrix_free(m);
  return rv;
}

/* ------------------------------------------------- 64-bit row pointers */
/* The reference's GOOSEBERRY_64BIT_INDICES build (include/gooseberry.h:
 * 65-71) gives ind_t 64 bits.  Resident matrices of more than 2^31 entries
 * are held as row parts (gbd_matrix_create64); the readers, writers and
 * sparse products keep 32-bit storage and narrow or widen at the boundary. */

/* GOOSEBERRY_PART_NNZ: entries per row part (0 or unset: the default) */
static uint64_t part_nnz_env(void)
{
  const char *e = getenv("GOOSEBERRY_PART_NNZ");
  if (!e || !*e) return 0;
  char *end = NULL;
  unsigned long long v = strtoull(e, &end, 10);
  return (end && *end == '\0') ? (uint64_t)v : 0;
}

/* 64-bit row pointers narrowed to a new 32-bit array (NULL: too many
 * entries, or no memory: *rv says which) */
static ind_t *narrow_rowptr(const uint64_t *rp, dim_t nrows, int *rv)
{
  *rv = GOOSEBERRY_SUCCESS;
  if (rp[nrows] > UINT32_MAX) {
    gbd_set_error("more than 2^32 - 1 entries: this operation keeps 32-bit row pointers");
    *rv = GOOSEBERRY_ERROR_UNIMPLEMENTED;
    return NULL;
  }
  ind_t *out = malloc(((size_t)nrows + 1) * sizeof(ind_t));
  if (!out) {
    *rv = GOOSEBERRY_ERROR_NOMEMORY;
    return NULL;
  }
  for (size_t i = 0; i <= nrows; ++i) {
    if (rp[i] > UINT32_MAX) {
      free(out);
      *rv = GOOSEBERRY_ERROR_INVALIDINPUT;
      return NULL;
    }
    out[i] = (ind_t)rp[i];
  }
  return out;
}

/* takes the three device arrays (freed here on failure) */
static int create64_owned(dim_t nrows, dim_t ncols, void *dp, void *di,
    void *dv, const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  gbd_plan_params pp;
  int have;
  plan_to_dev(plan, NULL, &pp, &have);
  gbd_matrix *m = NULL;
  int rv = gbd_matrix_create64(nrows, ncols, dp, di, dv, part_nnz_env(),
      have ? &pp : NULL, &m);
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(dp); gbd_free(di); gbd_free(dv);
    return rv;
  }
  return wrap_matrix(m, r_mat);
}

int gooseberry_matrix_create_i64(dim_t nrows, dim_t ncols,
    const uint64_t *rowptr, const dim_t *rowind, const real_t *rowval,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  if (!r_mat || !rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  const uint64_t nnz = rowptr[nrows];
  if (nnz && (!rowind || !rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* the three arrays in one staged upload */
  void *d[3] = {NULL, NULL, NULL};
  const void *h[3] = {rowptr, rowind, rowval};
  size_t b[3] = {((size_t)nrows + 1) * sizeof(uint64_t), nnz * sizeof(dim_t),
                 nnz * sizeof(real_t)};
  int rv = GOOSEBERRY_SUCCESS;
  for (int k = 0; k < 3 && rv == GOOSEBERRY_SUCCESS; ++k) rv = gbd_malloc(&d[k], b[k]);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_htod_many(3, d, h, b);
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(d[0]); gbd_free(d[1]); gbd_free(d[2]);
    return rv;
  }
  return create64_owned(nrows, ncols, d[0], d[1], d[2], plan, r_mat);
}

int gooseberry_matrix_create_device_i64(dim_t nrows, dim_t ncols,
    const uint64_t *d_rowptr, const dim_t *d_rowind, const real_t *d_rowval,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  GB_API_GUARD();
  if (!r_mat || !d_rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint64_t nnz = 0;
  int rv = gbd_dtoh(&nnz, d_rowptr + nrows, sizeof(uint64_t));
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (nnz && (!d_rowind || !d_rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  void *dp = NULL, *di = NULL, *dv = NULL;
  rv = gbd_malloc(&dp, ((size_t)nrows + 1) * sizeof(uint64_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&di, nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&dv, nnz * sizeof(real_t));
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_dtod(dp, d_rowptr, ((size_t)nrows + 1) * sizeof(uint64_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtod(di, d_rowind, nnz * sizeof(dim_t));
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_dtod(dv, d_rowval, nnz * sizeof(real_t));
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(dp); gbd_free(di); gbd_free(dv);
    return rv;
  }
  return create64_owned(nrows, ncols, dp, di, dv, plan, r_mat);
}

int gooseberry_matrix_parts(gooseberry_matrix_t mat, dim_t *r_nparts)
{
  GB_API_GUARD();
  if (!mat || !r_nparts) return GOOSEBERRY_ERROR_INVALIDINPUT;
  *r_nparts = gbd_matrix_nparts(mat->m);
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_spmult_i64(dim_t anrows, dim_t ancols, dim_t bncols,
    const uint64_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const real_t *bcolval, real_t *crowval, const dim_t *blocks,
    dim_t nblocks)
{
  GB_API_GUARD();
  dim_t rlo, rhi;
  if (spmult_rows(anrows, blocks, nblocks, &rlo, &rhi) != GOOSEBERRY_SUCCESS)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (anrows == 0 || bncols == 0 || rlo >= rhi) return GOOSEBERRY_SUCCESS;
  if (!arowptr || !bcolval || !crowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* rows [rlo, rhi) with row pointers rebased to 0, as gooseberry_spmult */
  const dim_t nr = rhi - rlo;
  const uint64_t base = arowptr[rlo];
  uint64_t *rp = malloc(((size_t)nr + 1) * sizeof(uint64_t));
  if (!rp) return GOOSEBERRY_ERROR_NOMEMORY;
  for (dim_t i = 0; i <= nr; ++i) {
    if (arowptr[rlo + i] < base) {
      free(rp);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
    }
    rp[i] = arowptr[rlo + i] - base;
  }
  gooseberry_plan_t once;
  spmult_once_plan(&once);
  gooseberry_matrix_t mat = NULL;
  int rv = gooseberry_matrix_create_i64(nr, ancols, rp,
      arowind ? arowind + base : NULL, arowval ? arowval + base : NULL,
      &once, &mat);
  free(rp);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  void *db = NULL;
  rv = gb_upload(&db, bcolval, (size_t)ancols * bncols * sizeof(real_t));
  if (rv != GOOSEBERRY_SUCCESS) {
    gooseberry_matrix_free(mat);
    return rv;
  }
  return spmult_once(mat, nr, bncols, db, crowval + (size_t)rlo * bncols);
}

int gooseberry_cgd_i64(dim_t nrows, dim_t ncols, const uint64_t *rowptr,
    const dim_t *rowind, const real_t *rowval, const real_t *b, real_t *x,
    real_t error, size_t niter, gooseberry_cgd_info_t *info)
{
  GB_API_GUARD();
  if (nrows != ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!b || !x) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gooseberry_matrix_t mat = NULL;
  int rv = gooseberry_matrix_create_i64(nrows, ncols, rowptr, rowind, rowval,
      NULL, &mat);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  return cgd_resident_host(mat, nrows, b, x, error, niter, info);
}

/* n values in chunks the 32-bit forms take */
int gooseberry_scale_i64(uint64_t n, real_t *val, real_t s)
{
  GB_API_GUARD();
  const uint64_t step = 1ull << 30;
  for (uint64_t off = 0; off < n; off += step) {
    const uint64_t k = n - off < step ? n - off : step;
    int rv = gooseberry_scale((ind_t)k, val + off, s);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_add_scalar_i64(uint64_t n, real_t *val, real_t s)
{
  GB_API_GUARD();
  const uint64_t step = 1ull << 30;
  for (uint64_t off = 0; off < n; off += step) {
    const uint64_t k = n - off < step ? n - off : step;
    int rv = gooseberry_add_scalar((ind_t)k, val + off, s);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_spmultsp_i64(dim_t anrows, dim_t ancols, dim_t bncols,
    const uint64_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const uint64_t *browptr, const dim_t *browind, const real_t *browval,
    real_t *crowval, const dim_t *blocks, dim_t nblocks)
{
  GB_API_GUARD();
  if (!arowptr || !browptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv;
  ind_t *a32 = narrow_rowptr(arowptr, anrows, &rv);
  if (!a32) return rv;
  ind_t *b32 = narrow_rowptr(browptr, ancols, &rv);
  if (!b32) {
    free(a32);
    return rv;
  }
  rv = gooseberry_spmultsp(anrows, ancols, bncols, a32, arowind, arowval, b32,
      browind, browval, crowval, blocks, nblocks);
  free(a32);
  free(b32);
  return rv;
}

int gooseberry_spmultsp_sp_i64(dim_t anrows, dim_t ancols, dim_t bncols,
    const uint64_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const uint64_t *browptr, const dim_t *browind, const real_t *browval,
    uint64_t **crowptr, dim_t **crowind, real_t **crowval,
    const dim_t *blocks, dim_t nblocks)
{
  GB_API_GUARD();
  if (!arowptr || !browptr || !crowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  int rv;
  ind_t *a32 = narrow_rowptr(arowptr, anrows, &rv);
  if (!a32) return rv;
  ind_t *b32 = narrow_rowptr(browptr, ancols, &rv);
  if (!b32) {
    free(a32);
    return rv;
  }
  ind_t *c32 = NULL;
  rv = gooseberry_spmultsp_sp(anrows, ancols, bncols, a32, arowind, arowval,
      b32, browind, browval, &c32, crowind, crowval, blocks, nblocks);
  free(a32);
  free(b32);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  uint64_t *c64 = malloc(((size_t)anrows + 1) * sizeof(uint64_t));
  if (!c64) {
    free(c32);
    free(*crowind);
    free(*crowval);
    *crowind = NULL;
    *crowval = NULL;
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  for (size_t i = 0; i <= anrows; ++i) c64[i] = c32[i];
  free(c32);
  *crowptr = c64;
  return GOOSEBERRY_SUCCESS;
}
