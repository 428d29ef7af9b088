/*
 * gb_ingest_api.c -- sparse matrix files straight into device memory.
 * CSR text is parsed on the GPU (gb_ingest.hip) with the semantics of the
 * reference's reader (src/io.c:369-498), and so is a MatrixMarket file's
 * body (src/io.c:1060-1145; its header lines are read here); the binary CSR
 * format is copied through pinned staging buffers and checked on the GPU.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gooseberry.h"
#include "gb_host.h"

static ssize_t strip_eol(char *line, size_t n)
{
  while (n > 0 && (line[n - 1] == '\n' || line[n - 1] == '\r')) line[--n] = '\0';
  return (ssize_t)n;
}

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

/* src/io.c:921-1059 (as the host reader restates it): the first line is
 * either "nrows ncols nnz" (no banner: real general) or the banner
 * "%%MatrixMarket matrix coordinate real general|symmetric"; after a banner
 * blank and comment lines are skipped and the size line follows.  Leaves the
 * byte offset of the body in *off. */
static int mtx_header(const char *file, dim_t *nrows, dim_t *ncols, int *sym,
    uint64_t *off)
{
  FILE *f = fopen(file, "rb");
  if (!f) {
    switch (errno) {
      case ENOENT: case ENOTDIR: return GOOSEBERRY_ERROR_FILENOTFOUND;
      case EACCES: case EPERM: return GOOSEBERRY_ERROR_PERMISSIONDENIED;
      default: return GOOSEBERRY_ERROR_UNKNOWN;
    }
  }
  char *line = NULL;
  size_t cap = 0;
  ssize_t got;
  uint64_t pos = 0;
  int rv = GOOSEBERRY_ERROR_INVALIDINPUT;
  /* a line: up to '\n', trailing '\r' and '\n' removed */
#define MTX_LINE()                                                        \
  ((got = getline(&line, &cap, f)) < 0 ? -1 : (pos += (uint64_t)got,      \
    strip_eol(line, (size_t)got)))
  ssize_t ll = MTX_LINE();
  if (ll < 1) goto done;
  {
    unsigned a, b, c;
    if (sscanf(line, "%u %u %u", &a, &b, &c) == 3) {
      *nrows = a;
      *ncols = b;
      *sym = 0;
    } else {
      static const char ws[] = " \t";
      char *save = NULL;
      char *tok = strtok_r(line, ws, &save);
      if (!tok || strcmp(tok, "%%MatrixMarket") != 0) goto done;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "matrix") != 0) goto done;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "coordinate") != 0) goto done;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "real") != 0) goto done;
      tok = strtok_r(NULL, ws, &save);
      if (!tok) goto done;
      if (strcmp(tok, "general") == 0) *sym = 0;
      else if (strcmp(tok, "symmetric") == 0) *sym = 1;
      else goto done;
      while ((ll = MTX_LINE()) >= 0) {
        if (ll == 0 || (line[0] != '\0' && strchr("#%'\"/", line[0]) != NULL)) continue;
        break;
      }
      if (ll < 0) goto done;
      char *sptr = line, *eptr;
      *nrows = (dim_t)strtoull(sptr, &eptr, 10);
      sptr = eptr;
      *ncols = (dim_t)strtoull(sptr, &eptr, 10);
      sptr = eptr;
      (void)strtoull(sptr, &eptr, 10);
      if (sptr == eptr) goto done;
    }
  }
#undef MTX_LINE
  *off = pos;
  rv = GOOSEBERRY_SUCCESS;
done:
  free(line);
  fclose(f);
  return rv;
}

static int read_device(int type, const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **d_rowptr, dim_t **d_rowind, real_t **d_rowval, ind_t *r_nnz)
{
  if (!file || !nrows || !ncols || !d_rowptr || !d_rowind || !d_rowval)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (type == GOOSEBERRY_FORMAT_BCSR)
    return read_device_bcsr(file, nrows, ncols, d_rowptr, d_rowind, d_rowval,
        r_nnz);
  if (type != GOOSEBERRY_FORMAT_CSR && type != GOOSEBERRY_FORMAT_MATRIXMARKET) {
    fprintf(stderr, "Device reading of sparse format %d is not provided by "
        "this build\n", type);
    return GOOSEBERRY_ERROR_UNIMPLEMENTED;
  }
  dim_t mr = 0, mc = 0;
  int sym = 0;
  uint64_t off = 0;
  int rv;
  if (type == GOOSEBERRY_FORMAT_MATRIXMARKET) {
    rv = mtx_header(file, &mr, &mc, &sym, &off);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  char *d_txt = NULL;
  uint64_t len = 0;
  rv = gbd_load_file(file, &d_txt, &len);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (type == GOOSEBERRY_FORMAT_CSR) {
    rv = gbd_parse_csr(d_txt, len, nrows, ncols, r_nnz, d_rowptr, d_rowind,
        d_rowval);
  } else {
    rv = gbd_parse_mtx_body(d_txt, len, off, mr, mc, sym, r_nnz, d_rowptr,
        d_rowind, d_rowval);
    if (rv == GOOSEBERRY_SUCCESS) {
      *nrows = mr;
      *ncols = mc;
    }
  }
  gbd_free(d_txt);
  return rv;
}

int gooseberry_read_sparse_matrix_device(int type, const char *file,
    dim_t *nrows, dim_t *ncols, ind_t **d_rowptr, dim_t **d_rowind,
    real_t **d_rowval)
{
  GB_API_GUARD();
  ind_t nnz;
  return read_device(type, file, nrows, ncols, d_rowptr, d_rowind, d_rowval,
      &nnz);
}

int gooseberry_parse_csr_device(const char *d_text, size_t len, dim_t *nrows,
    dim_t *ncols, ind_t **d_rowptr, dim_t **d_rowind, real_t **d_rowval)
{
  GB_API_GUARD();
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
  GB_API_GUARD();
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
