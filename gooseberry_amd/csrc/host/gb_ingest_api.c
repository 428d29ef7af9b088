/*
 * gb_ingest_api.c -- sparse matrix files straight into device memory.
 * CSR text is parsed on the GPU (gb_ingest.hip) with the semantics of the
 * reference's reader (src/io.c:369-498); the binary CSR format is copied
 * through pinned staging buffers and checked on the GPU.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>

#include "gooseberry.h"
#include "gb_host.h"

static int read_device_bcsr(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **d_rowptr, dim_t **d_rowind, real_t **d_rowval, ind_t *r_nnz)
{
  FILE *f = fopen(file, "rb");
  if (!f) {
    switch (errno) {
      case ENOENT: case ENOTDIR: return GOOSEBERRY_ERROR_FILENOTFOUND;
      case EACCES: case EPERM: return GOOSEBERRY_ERROR_PERMISSIONDENIED;
      default: return GOOSEBERRY_ERROR_UNKNOWN;
    }
  }
  dim_t nr = 0, nc = 0;
  uint64_t nnz = 0;
  int rv = gb_bcsr_header(f, &nr, &nc, &nnz);
  fclose(f);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  void *rp = NULL, *ri = NULL, *rvl = NULL;
  const uint64_t o1 = 24, o2 = o1 + ((uint64_t)nr + 1) * 4, o3 = o2 + nnz * 4;
  rv = gbd_malloc(&rp, ((size_t)nr + 1) * 4);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&ri, (nnz ? nnz : 1) * 4);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_malloc(&rvl, (nnz ? nnz : 1) * 8);
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_load_file_range(file, o1, ((uint64_t)nr + 1) * 4, rp);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_load_file_range(file, o2, nnz * 4, ri);
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_load_file_range(file, o3, nnz * 8, rvl);
  uint32_t got = 0;
  if (rv == GOOSEBERRY_SUCCESS) rv = gbd_csr_check(nr, nc, rp, ri, &got);
  if (rv == GOOSEBERRY_SUCCESS && got != nnz) rv = GOOSEBERRY_ERROR_INVALIDINPUT;
  if (rv != GOOSEBERRY_SUCCESS) {
    gbd_free(rp); gbd_free(ri); gbd_free(rvl);
    return rv;
  }
  *nrows = nr;
  *ncols = nc;
  *d_rowptr = rp;
  *d_rowind = ri;
  *d_rowval = rvl;
  *r_nnz = (ind_t)nnz;
  return GOOSEBERRY_SUCCESS;
}

static int read_device(int type, const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **d_rowptr, dim_t **d_rowind, real_t **d_rowval, ind_t *r_nnz)
{
  if (!file || !nrows || !ncols || !d_rowptr || !d_rowind || !d_rowval)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (type == GOOSEBERRY_FORMAT_BCSR)
    return read_device_bcsr(file, nrows, ncols, d_rowptr, d_rowind, d_rowval,
        r_nnz);
  if (type != GOOSEBERRY_FORMAT_CSR) {
    fprintf(stderr, "Device reading of sparse format %d is not provided by "
        "this build\n", type);
    return GOOSEBERRY_ERROR_UNIMPLEMENTED;
  }
  char *d_txt = NULL;
  uint64_t len = 0;
  int rv = gbd_load_file(file, &d_txt, &len);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  rv = gbd_parse_csr(d_txt, len, nrows, ncols, r_nnz, d_rowptr, d_rowind,
      d_rowval);
  gbd_free(d_txt);
  return rv;
}

int gooseberry_read_sparse_matrix_device(int type, const char *file,
    dim_t *nrows, dim_t *ncols, ind_t **d_rowptr, dim_t **d_rowind,
    real_t **d_rowval)
{
  ind_t nnz;
  return read_device(type, file, nrows, ncols, d_rowptr, d_rowind, d_rowval,
      &nnz);
}

int gooseberry_parse_csr_device(const char *d_text, size_t len, dim_t *nrows,
    dim_t *ncols, ind_t **d_rowptr, dim_t **d_rowind, real_t **d_rowval)
{
  if ((len && !d_text) || !nrows || !ncols || !d_rowptr || !d_rowind ||
      !d_rowval)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  ind_t nnz;
  return gbd_parse_csr(d_text, len, nrows, ncols, &nnz, d_rowptr, d_rowind,
      d_rowval);
}

int gooseberry_matrix_read(int type, const char *file,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat)
{
  if (!r_mat) return GOOSEBERRY_ERROR_INVALIDINPUT;
  dim_t nrows = 0, ncols = 0;
  ind_t *rp = NULL, nnz = 0;
  dim_t *ri = NULL;
  real_t *rv = NULL;
  int err = read_device(type, file, &nrows, &ncols, &rp, &ri, &rv, &nnz);
  if (err != GOOSEBERRY_SUCCESS) return err;
  /* columns up to 2^32 - 1 above a column 0 wrap ncols to 0 (as the
   * reference's dim_t arithmetic does); no matrix can hold them */
  if (nnz && ncols == 0) {
    gbd_free(rp); gbd_free(ri); gbd_free(rv);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  return gb_matrix_adopt(nrows, ncols, rp, ri, rv, nnz, plan, r_mat);
}
