/*
 * gb_io.c -- CSR construction from text and vector I/O for the blas/cgd
 * path: the .csr, .mtx and grid readers and the grid/.csr writers of
 * src/io.c, with the same parsing semantics (identical integer arrays and
 * values for identical input text).
 *
 * The whole file is read into memory once and split into lines in place;
 * a line is the text up to '\n' with trailing '\r' removed.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "gooseberry.h"
#include "gb_host.h"

/* src/io.c:63-69 */
static int is_comment(char c)
{
  return c == '#' || c == '%' || c == '\'' || c == '"' || c == '/';
}

static int errno_to_rv(int e)
{
  switch (e) {
    case ENOENT: case ENOTDIR: return GOOSEBERRY_ERROR_FILENOTFOUND;
    case EACCES: case EPERM: return GOOSEBERRY_ERROR_PERMISSIONDENIED;
    case EINVAL: case ENAMETOOLONG: return GOOSEBERRY_ERROR_INVALIDINPUT;
    default: return GOOSEBERRY_ERROR_UNKNOWN;
  }
}

typedef struct {
  char *buf;
  size_t len;
  size_t pos;
} text_t;

static int text_load(const char *file, text_t *t)
{
  memset(t, 0, sizeof(*t));
  if (!file) return GOOSEBERRY_ERROR_INVALIDINPUT;
  FILE *f = fopen(file, "rb");
  if (!f) return errno_to_rv(errno);
  size_t cap = 1 << 16, len = 0;
  char *buf = malloc(cap + 1);
  if (!buf) { fclose(f); return GOOSEBERRY_ERROR_UNKNOWN; }
  for (;;) {
    if (len == cap) {
      cap *= 2;
      char *nb = realloc(buf, cap + 1);
      if (!nb) { free(buf); fclose(f); return GOOSEBERRY_ERROR_UNKNOWN; }
      buf = nb;
    }
    size_t got = fread(buf + len, 1, cap - len, f);
    len += got;
    if (got == 0) break;
  }
  int err = ferror(f);
  fclose(f);
  if (err) { free(buf); return GOOSEBERRY_ERROR_READ; }
  buf[len] = '\0';
  t->buf = buf;
  t->len = len;
  t->pos = 0;
  return GOOSEBERRY_SUCCESS;
}

/* next line as a NUL-terminated string; returns its length or -1 at end */
static ssize_t text_line(text_t *t, char **line)
{
  if (t->pos >= t->len) return -1;
  char *s = t->buf + t->pos;
  char *nl = memchr(s, '\n', t->len - t->pos);
  size_t ll;
  if (nl) {
    ll = (size_t)(nl - s);
    *nl = '\0';
    t->pos += ll + 1;
  } else {
    ll = t->len - t->pos;
    t->pos = t->len;
  }
  while (ll > 0 && (s[ll - 1] == '\r' || s[ll - 1] == '\n')) s[--ll] = '\0';
  *line = s;
  return (ssize_t)ll;
}

typedef struct { void *p; size_t n, cap, esz; } arr_t;

static int arr_push(arr_t *a, const void *e)
{
  if (a->n == a->cap) {
    size_t nc = a->cap ? 2 * a->cap : 1024;
    void *np = realloc(a->p, nc * a->esz);
    if (!np) return 0;
    a->p = np;
    a->cap = nc;
  }
  memcpy((char *)a->p + a->n * a->esz, e, a->esz);
  a->n++;
  return 1;
}

/* src/io.c:369-498 */
static int read_csr(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  arr_t ptr = {0, 0, 0, sizeof(ind_t)}, ind = {0, 0, 0, sizeof(dim_t)},
        val = {0, 0, 0, sizeof(real_t)};
  dim_t mincol = 1, maxcol = 0, n = 0;
  ind_t nnz = 0;
  int ok = arr_push(&ptr, &nnz);
  char *line;
  ssize_t ll;
  while (ok && (ll = text_line(&t, &line)) >= 0) {
    if (ll != 0 && is_comment(line[0])) continue;
    char *sptr = line, *eptr;
    dim_t c = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    real_t v = strtod(sptr, &eptr);
    while (ok && eptr != sptr) {
      if (c > maxcol) maxcol = c;
      if (c < mincol) mincol = c;
      ok = arr_push(&ind, &c) && arr_push(&val, &v);
      ++nnz;
      sptr = eptr;
      c = (dim_t)strtoull(sptr, &eptr, 10);
      sptr = eptr;
      v = strtod(sptr, &eptr);
    }
    ok = ok && arr_push(&ptr, &nnz);
    ++n;
  }
  free(t.buf);
  if (!ok) {
    free(ptr.p); free(ind.p); free(val.p);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  dim_t *ri = ind.p;
  for (size_t k = 0; k < ind.n; ++k) ri[k] -= mincol; /* mincol <= 1 */
  *r_rowptr = ptr.p;
  *r_rowind = ind.n ? ind.p : calloc(1, sizeof(dim_t));
  *r_rowval = val.n ? val.p : calloc(1, sizeof(real_t));
  *nrows = n;
  *ncols = maxcol - mincol + 1;
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:921-1190 */
static int read_mtx(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  char *line;
  ssize_t ll;
  int sym = 0, wgts = 0;
  dim_t nrows = 0, ncols = 0;
  ind_t *rowptr = NULL;
  dim_t *rowind = NULL;
  real_t *rowval = NULL;
  rv = GOOSEBERRY_ERROR_INVALIDINPUT;
  if ((ll = text_line(&t, &line)) < 1) goto ERR;
  {
    unsigned a, b, c;
    if (sscanf(line, "%u %u %u", &a, &b, &c) == 3) {
      nrows = a; ncols = b; sym = 0; /* no header: real general */
    } else {
      static const char ws[] = " \t";
      char *save = NULL;
      char *tok = strtok_r(line, ws, &save);
      if (!tok || strcmp(tok, "%%MatrixMarket") != 0) goto ERR;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "matrix") != 0) goto ERR;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "coordinate") != 0) goto ERR;
      tok = strtok_r(NULL, ws, &save);
      if (!tok || strcmp(tok, "real") != 0) goto ERR;
      tok = strtok_r(NULL, ws, &save);
      if (!tok) goto ERR;
      if (strcmp(tok, "general") == 0) sym = 0;
      else if (strcmp(tok, "symmetric") == 0) sym = 1;
      else goto ERR;
      while ((ll = text_line(&t, &line)) >= 0) {
        if (ll == 0 || is_comment(line[0])) continue;
        break;
      }
      if (ll < 0) goto ERR;
      char *sptr = line, *eptr;
      nrows = (dim_t)strtoull(sptr, &eptr, 10); sptr = eptr;
      ncols = (dim_t)strtoull(sptr, &eptr, 10); sptr = eptr;
      (void)strtoull(sptr, &eptr, 10);
      if (sptr == eptr) goto ERR;
    }
  }
  rowptr = calloc((size_t)nrows + 2, sizeof(ind_t));
  if (!rowptr) { rv = GOOSEBERRY_ERROR_UNKNOWN; goto ERR; }
  /* pass 1: parse the coordinates and count entries per row in
   * rowptr[i+1] (src/io.c:1060-1112) */
  arr_t ti = {0, 0, 0, sizeof(dim_t)}, tj = {0, 0, 0, sizeof(dim_t)},
        tw = {0, 0, 0, sizeof(real_t)};
  size_t nnz = 0, nent = 0;
  int ok = 1;
  while (ok && (ll = text_line(&t, &line)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    char *sptr = line, *eptr;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10) - 1;
    if (eptr == sptr) goto ERRT;
    sptr = eptr;
    dim_t j = (dim_t)strtoull(sptr, &eptr, 10) - 1;
    if (eptr == sptr) goto ERRT;
    sptr = eptr;
    real_t w = strtod(sptr, &eptr);
    if (eptr != sptr) {
      if (nent == 0) wgts = 1;
      else if (!wgts) goto ERRT;
    } else if (wgts) {
      goto ERRT;
    }
    if (i >= nrows || j >= ncols || (sym && j >= nrows)) goto ERRT;
    ok = arr_push(&ti, &i) && arr_push(&tj, &j) && (!wgts || arr_push(&tw, &w));
    ++nent;
    if (sym) { ++rowptr[j + 1]; ++nnz; }
    ++rowptr[i + 1];
    ++nnz;
  }
  if (!ok || nnz > 0xffffffffull) goto ERRT;
  {
    ind_t s = 0;
    for (size_t k = 1; k <= nrows; ++k) { ind_t c = rowptr[k]; rowptr[k] = s; s += c; }
  }
  rowind = malloc(sizeof(dim_t) * (nnz ? nnz : 1));
  rowval = malloc(sizeof(real_t) * (nnz ? nnz : 1));
  if (!rowind || !rowval) goto ERRT;
  /* pass 2: fill in file order, mirrored entry first (src/io.c:1115-1145) */
  for (size_t e = 0; e < nent; ++e) {
    dim_t i = ((dim_t *)ti.p)[e], j = ((dim_t *)tj.p)[e];
    real_t w = wgts ? ((real_t *)tw.p)[e] : 1;
    ind_t k;
    if (sym) { k = rowptr[j + 1]++; rowind[k] = i; rowval[k] = w; }
    k = rowptr[i + 1]++;
    rowind[k] = j;
    rowval[k] = w;
  }
  free(ti.p); free(tj.p); free(tw.p);
  free(t.buf);
  *r_nrows = nrows;
  *r_ncols = ncols;
  *r_rowptr = rowptr;
  *r_rowind = rowind;
  *r_rowval = rowval;
  return GOOSEBERRY_SUCCESS;
ERRT:
  free(ti.p); free(tj.p); free(tw.p);
ERR:
  free(t.buf);
  free(rowptr);
  free(rowind);
  free(rowval);
  return rv;
}

/* src/io.c:145-239 */
static int read_grid(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  arr_t val = {0, 0, 0, sizeof(real_t)};
  dim_t nrows = 0, ncols = 0;
  char *line;
  ssize_t ll;
  int ok = 1;
  while (ok && (ll = text_line(&t, &line)) >= 0) {
    if (is_comment(line[0])) continue;
    dim_t j = 0;
    char *sptr = line, *eptr;
    real_t v = strtod(sptr, &eptr);
    while (ok && eptr != sptr) {
      ok = arr_push(&val, &v);
      ++j;
      sptr = eptr;
      v = strtod(sptr, &eptr);
    }
    if (ncols == 0) {
      ncols = j;
    } else if (j != ncols) {
      free(t.buf);
      free(val.p);
      fprintf(stderr, "Invalid line in grid matrix file: %u\n", nrows + 1);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
    }
    ++nrows;
  }
  free(t.buf);
  if (!ok) { free(val.p); return GOOSEBERRY_ERROR_UNKNOWN; }
  if ((size_t)nrows * ncols != val.n) {
    free(val.p);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  *r_nrows = nrows;
  *r_ncols = ncols;
  *r_rowval = val.n ? val.p : calloc(1, sizeof(real_t));
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:1195-1229 */
static int write_grid(const char *file, dim_t nrows, dim_t ncols,
    const real_t *rowval)
{
  FILE *f = fopen(file, "w");
  if (!f) return errno_to_rv(errno);
  static char big[1 << 20];
  setvbuf(f, big, _IOFBF, sizeof(big));
  for (dim_t i = 0; i < nrows; ++i) {
    fprintf(f, "%g", rowval[(size_t)i * ncols]);
    for (dim_t j = 1; j < ncols; ++j) fprintf(f, " %g", rowval[(size_t)i * ncols + j]);
    fputc('\n', f);
  }
  int bad = ferror(f);
  if (fclose(f) != 0) bad = 1;
  return bad ? GOOSEBERRY_ERROR_WRITE : GOOSEBERRY_SUCCESS;
}

/* src/io.c:1233-1274 */
static int write_csr(const char *file, dim_t nrows, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval)
{
  FILE *f = fopen(file, "w");
  if (!f) return errno_to_rv(errno);
  static char big[1 << 20];
  setvbuf(f, big, _IOFBF, sizeof(big));
  for (dim_t i = 0; i < nrows; ++i) {
    if (rowptr[i] < rowptr[i + 1]) {
      fprintf(f, "%" PRIu32 " %g", rowind[rowptr[i]] + 1, rowval[rowptr[i]]);
      for (ind_t j = rowptr[i] + 1; j < rowptr[i + 1]; ++j)
        fprintf(f, " %" PRIu32 " %g", rowind[j] + 1, rowval[j]);
    }
    fputc('\n', f);
  }
  int bad = ferror(f);
  if (fclose(f) != 0) bad = 1;
  return bad ? GOOSEBERRY_ERROR_WRITE : GOOSEBERRY_SUCCESS;
}

/* Binary CSR: the 8-byte magic, uint32 nrows, uint32 ncols, uint64 nnz,
 * then rowptr, rowind, rowval as stored in memory (little-endian) */
static const unsigned char BCSR_MAGIC[8] = {'G', 'B', 'C', 'S', 'R', 0, 1, 0};

int gb_bcsr_header(FILE *f, dim_t *nrows, dim_t *ncols, uint64_t *nnz)
{
  unsigned char h[24];
  if (fread(h, 1, 24, f) != 24 || memcmp(h, BCSR_MAGIC, 8) != 0)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  memcpy(nrows, h + 8, 4);
  memcpy(ncols, h + 12, 4);
  memcpy(nnz, h + 16, 8);
  if (*nnz > 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* the file must hold exactly the arrays the header announces */
  struct stat st;
  if (fstat(fileno(f), &st) != 0) return GOOSEBERRY_ERROR_READ;
  uint64_t want = 24 + ((uint64_t)*nrows + 1) * 4 + *nnz * 12;
  if ((uint64_t)st.st_size != want) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return GOOSEBERRY_SUCCESS;
}

static int write_bcsr(const char *file, dim_t nrows, dim_t ncols,
    const ind_t *rowptr, const dim_t *rowind, const real_t *rowval)
{
  FILE *f = fopen(file, "wb");
  if (!f) return errno_to_rv(errno);
  uint64_t nnz = rowptr[nrows];
  unsigned char h[24];
  memcpy(h, BCSR_MAGIC, 8);
  memcpy(h + 8, &nrows, 4);
  memcpy(h + 12, &ncols, 4);
  memcpy(h + 16, &nnz, 8);
  int bad = fwrite(h, 1, 24, f) != 24;
  bad = bad || fwrite(rowptr, 4, (size_t)nrows + 1, f) != (size_t)nrows + 1;
  bad = bad || (nnz && fwrite(rowind, 4, nnz, f) != nnz);
  bad = bad || (nnz && fwrite(rowval, 8, nnz, f) != nnz);
  if (fclose(f) != 0) bad = 1;
  return bad ? GOOSEBERRY_ERROR_WRITE : GOOSEBERRY_SUCCESS;
}

static int read_bcsr(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  FILE *f = fopen(file, "rb");
  if (!f) return errno_to_rv(errno);
  dim_t nrows, ncols;
  uint64_t nnz;
  int rv = gb_bcsr_header(f, &nrows, &ncols, &nnz);
  ind_t *rp = NULL;
  dim_t *ri = NULL;
  real_t *rv8 = NULL;
  if (rv == GOOSEBERRY_SUCCESS) {
    rp = malloc(((size_t)nrows + 1) * 4);
    ri = malloc((nnz ? nnz : 1) * 4);
    rv8 = malloc((nnz ? nnz : 1) * 8);
    if (!rp || !ri || !rv8) rv = GOOSEBERRY_ERROR_UNKNOWN;
  }
  if (rv == GOOSEBERRY_SUCCESS &&
      (fread(rp, 4, (size_t)nrows + 1, f) != (size_t)nrows + 1 ||
       (nnz && fread(ri, 4, nnz, f) != nnz) ||
       (nnz && fread(rv8, 8, nnz, f) != nnz)))
    rv = GOOSEBERRY_ERROR_READ;
  fclose(f);
  if (rv == GOOSEBERRY_SUCCESS &&
      (rp[nrows] != nnz || gb_check_csr(nrows, ncols, rp, ri) != GOOSEBERRY_SUCCESS))
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
  if (rv != GOOSEBERRY_SUCCESS) {
    free(rp); free(ri); free(rv8);
    return rv;
  }
  *r_nrows = nrows;
  *r_ncols = ncols;
  *r_rowptr = rp;
  *r_rowind = ri;
  *r_rowval = rv8;
  return GOOSEBERRY_SUCCESS;
}

/* ---- the other sparse formats of src/io.c (METIS/Chaco graph, CLUTO,
 * libsvm, coordinate points), read on the host.  Each follows its reference
 * reader statement for statement.  The reference trusts every count a header
 * gives and every 1-based index a line gives; where it would then write past
 * its arrays or store an index wrapped below zero or beyond the matrix, these
 * readers stop with GOOSEBERRY_ERROR_INVALIDINPUT and say why. */

typedef struct {
  ind_t *rowptr;
  dim_t *rowind;
  real_t *rowval;
} csr_bufs_t;

static void csr_bufs_free(csr_bufs_t *b)
{
  free(b->rowptr);
  free(b->rowind);
  free(b->rowval);
  memset(b, 0, sizeof(*b));
}

static int csr_bufs_alloc(csr_bufs_t *b, size_t nrows, size_t nnz)
{
  b->rowptr = malloc((nrows + 1) * sizeof(ind_t));
  b->rowind = malloc((nnz ? nnz : 1) * sizeof(dim_t));
  b->rowval = malloc((nnz ? nnz : 1) * sizeof(real_t));
  if (!b->rowptr || !b->rowind || !b->rowval) {
    csr_bufs_free(b);
    return 0;
  }
  return 1;
}

/* the first non-empty line that is not a comment, or NULL */
static char *header_line(text_t *t)
{
  char *line;
  ssize_t ll;
  while ((ll = text_line(t, &line)) >= 0)
    if (ll != 0 && !is_comment(line[0])) return line;
  return NULL;
}

static int refuse(text_t *t, csr_bufs_t *b, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  free(t->buf);
  t->buf = NULL;
  csr_bufs_free(b);
  return GOOSEBERRY_ERROR_INVALIDINPUT;
}

/* columns must fall inside the matrix the reader reports */
static int columns_in_range(const dim_t *rowind, ind_t nnz, dim_t ncols)
{
  for (ind_t k = 0; k < nnz; ++k)
    if (rowind[k] >= ncols) return 0;
  return 1;
}

/* src/io.c:242-367: header "n m [fmt]" (2m entries, weights when fmt is
 * a nonzero number), then one line per vertex -- an empty line is a vertex
 * without edges -- listing 1-based neighbours, each followed by its weight
 * when there are weights; rows = lines read, ncols = nrows */
static int read_graph(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  csr_bufs_t b = {0, 0, 0};
  char *line = header_line(&t), *sptr, *eptr;
  dim_t hn = 0;
  ind_t cap = 0;
  int do_wgt = 0;
  if (line) {
    sptr = line;
    hn = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    cap = (ind_t)strtoull(sptr, &eptr, 10) * 2;
    sptr = eptr;
    do_wgt = (int)strtoull(sptr, &eptr, 10);
    if (sptr == eptr) do_wgt = 0;
  }
  if (hn == 0 && cap == 0)
    return refuse(&t, &b, "Unable to read graph file or empty graph specified\n");
  if (!csr_bufs_alloc(&b, hn, cap)) {
    free(t.buf);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  dim_t n = 0;
  ind_t j = 0;
  ssize_t ll;
  b.rowptr[0] = 0;
  while ((ll = text_line(&t, &line)) >= 0) {
    if (ll != 0 && is_comment(line[0])) continue;
    if (n == hn)
      return refuse(&t, &b, "Graph file has more vertex lines than its header's %" PRIu32 "\n", hn);
    real_t f = 0.0;
    sptr = line;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10);
    if (do_wgt) {
      sptr = eptr;
      f = strtod(sptr, &eptr);
    }
    while (eptr != sptr) {
      if (j == cap)
        return refuse(&t, &b, "Graph file has more edges than its header's %" PRIu32 "\n", cap / 2);
      if (i == 0)
        return refuse(&t, &b, "Graph vertex numbers start at 1 (line of vertex %" PRIu32 ")\n",
            n + 1);
      b.rowind[j] = i - 1;
      b.rowval[j] = do_wgt ? f : 1.0;
      ++j;
      sptr = eptr;
      i = (dim_t)strtoull(sptr, &eptr, 10);
      if (do_wgt) {
        sptr = eptr;
        f = strtod(sptr, &eptr);
      }
    }
    b.rowptr[++n] = j;
  }
  if (!columns_in_range(b.rowind, j, n))
    return refuse(&t, &b, "Graph file names a vertex beyond its %" PRIu32 " vertex lines\n", n);
  free(t.buf);
  *r_rowptr = b.rowptr;
  *r_rowind = b.rowind;
  *r_rowval = b.rowval;
  *nrows = n;
  *ncols = n;
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:501-611: header "nrows ncols nnz", then one line per row (empty
 * lines and comments skipped) of 1-based "col val" pairs; ncols from the
 * header */
static int read_clu(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  csr_bufs_t b = {0, 0, 0};
  char *line = header_line(&t), *sptr, *eptr;
  dim_t hn = 0, maxcol = 0;
  ind_t cap = 0;
  if (!line) return refuse(&t, &b, "CLUTO file has no header line\n");
  /* fields the header lacks stay 0, as in the reference */
  (void)sscanf(line, "%" SCNu32 " %" SCNu32 " %" SCNu32, &hn, &maxcol, &cap);
  if (!csr_bufs_alloc(&b, hn, cap)) {
    free(t.buf);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  dim_t n = 0;
  ind_t j = 0;
  ssize_t ll;
  b.rowptr[0] = 0;
  while ((ll = text_line(&t, &line)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    if (n == hn)
      return refuse(&t, &b, "CLUTO file has more rows than its header's %" PRIu32 "\n", hn);
    sptr = line;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    real_t f = strtod(sptr, &eptr);
    while (eptr != sptr) {
      if (j == cap)
        return refuse(&t, &b, "CLUTO file has more entries than its header's %" PRIu32 "\n", cap);
      if (i == 0 || i > maxcol)
        return refuse(&t, &b, "CLUTO column %" PRIu32 " is outside 1..%" PRIu32 "\n", i, maxcol);
      b.rowind[j] = i - 1;
      b.rowval[j] = f;
      ++j;
      sptr = eptr;
      i = (dim_t)strtoull(sptr, &eptr, 10);
      sptr = eptr;
      f = strtod(sptr, &eptr);
    }
    b.rowptr[++n] = j;
  }
  free(t.buf);
  *r_rowptr = b.rowptr;
  *r_rowind = b.rowind;
  *r_rowval = b.rowval;
  *nrows = n;
  *ncols = maxcol;
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:612-745: one line per row (empty lines and comments skipped):
 * a label, then 1-based "index:value" items; the one character after an
 * index is skipped whatever it is.  ncols = the largest index.  Where an
 * index ends its line the reference steps past the line's end before
 * reading a value; here the row ends there. */
static int read_svm(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  csr_bufs_t b = {0, 0, 0};
  arr_t ptr = {0, 0, 0, sizeof(ind_t)}, ind = {0, 0, 0, sizeof(dim_t)},
        val = {0, 0, 0, sizeof(real_t)};
  dim_t n = 0, maxcol = 0;
  ind_t nnz = 0;
  int ok = arr_push(&ptr, &nnz);
  char *line, *sptr, *eptr;
  ssize_t ll;
  while (ok && (ll = text_line(&t, &line)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    sptr = line;
    (void)strtoll(sptr, &eptr, 10); /* the label */
    sptr = eptr;
    for (;;) {
      dim_t i = (dim_t)strtoull(sptr, &eptr, 10);
      if (*eptr == '\0') break;
      sptr = eptr + 1;
      real_t f = strtod(sptr, &eptr);
      if (eptr == sptr) break;
      if (i == 0) {
        free(ptr.p); free(ind.p); free(val.p);
        return refuse(&t, &b, "libsvm indices start at 1 (row %" PRIu32 ")\n", n + 1);
      }
      if (i > maxcol) maxcol = i;
      const dim_t c = i - 1;
      ok = arr_push(&ind, &c) && arr_push(&val, &f);
      if (!ok) break;
      ++nnz;
      sptr = eptr;
    }
    ok = ok && arr_push(&ptr, &nnz);
    ++n;
  }
  free(t.buf);
  if (!ok) {
    free(ptr.p); free(ind.p); free(val.p);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  *r_rowptr = ptr.p;
  *r_rowind = ind.n ? ind.p : calloc(1, sizeof(dim_t));
  *r_rowval = val.n ? val.p : calloc(1, sizeof(real_t));
  *nrows = n;
  *ncols = maxcol;
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:746-920: "row col value" lines (empty lines and comments
 * skipped) in any order.  The index bases follow the reference's rule: a
 * row number that is not a new maximum becomes minrow (likewise columns),
 * nrows = max + 1 - minrow; entries are grouped by row in file order. */
static int read_point(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  csr_bufs_t b = {0, 0, 0};
  arr_t ri = {0, 0, 0, sizeof(dim_t)}, ci = {0, 0, 0, sizeof(dim_t)},
        wv = {0, 0, 0, sizeof(real_t)};
  dim_t nrows = 0, ncols = 0, minrow = 1, mincol = 1;
  ind_t nnz = 0;
  int ok = 1;
  char *line, *sptr, *eptr;
  ssize_t ll;
  while (ok && (ll = text_line(&t, &line)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    sptr = line;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    dim_t j = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    real_t w = strtod(sptr, &eptr);
    if (eptr == sptr) {
      free(ri.p); free(ci.p); free(wv.p);
      return refuse(&t, &b, "Error in point list at line %" PRIu32 "\n", nnz);
    }
    if (i > nrows) nrows = i;
    else minrow = i;
    if (j > ncols) ncols = j;
    else mincol = j;
    ok = arr_push(&ri, &i) && arr_push(&ci, &j) && arr_push(&wv, &w);
    ++nnz;
  }
  free(t.buf);
  t.buf = NULL;
  nrows += 1 - minrow;
  ncols += 1 - mincol;
  const dim_t *pi = ri.p, *pj = ci.p;
  for (ind_t k = 0; ok && k < nnz; ++k) {
    if ((dim_t)(pi[k] - minrow) >= nrows || (dim_t)(pj[k] - mincol) >= ncols) {
      free(ri.p); free(ci.p); free(wv.p);
      fprintf(stderr, "Point list entry %llu falls outside the %" PRIu32 " x %" PRIu32
          " matrix its index bases give\n", (unsigned long long)k + 1, nrows, ncols);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
    }
  }
  if (!ok || !csr_bufs_alloc(&b, nrows, nnz)) {
    free(ri.p); free(ci.p); free(wv.p);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  memset(b.rowptr, 0, ((size_t)nrows + 1) * sizeof(ind_t));
  for (ind_t k = 0; k < nnz; ++k) ++b.rowptr[pi[k] - minrow + 1];
  for (dim_t r = 0; r < nrows; ++r) b.rowptr[r + 1] += b.rowptr[r];
  const real_t *pw = wv.p;
  for (ind_t k = 0; k < nnz; ++k) {
    const ind_t at = b.rowptr[pi[k] - minrow]++;
    b.rowind[at] = pj[k] - mincol;
    b.rowval[at] = pw[k];
  }
  for (dim_t r = nrows; r > 0; --r) b.rowptr[r] = b.rowptr[r - 1];
  b.rowptr[0] = 0;
  free(ri.p); free(ci.p); free(wv.p);
  *r_rowptr = b.rowptr;
  *r_rowind = b.rowind;
  *r_rowval = b.rowval;
  *r_nrows = nrows;
  *r_ncols = ncols;
  return GOOSEBERRY_SUCCESS;
}

/* src/io.c:1277-1366: loops left out (the edge count drops by one per row
 * holding one), weights written when any two consecutive values among the
 * first nnz differ, integral weights as integers */
static int write_graph(const char *file, dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval)
{
  if (nrows != ncols) {
    fprintf(stderr, "Graphs must have an equal number of rows and columns: %" PRIu32
        "x%" PRIu32 "\n", nrows, ncols);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  FILE *f = fopen(file, "w");
  if (!f) return errno_to_rv(errno);
  ind_t nnz = rowptr[nrows];
  for (dim_t i = 0; i < nrows; ++i)
    for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
      if (rowind[j] == i) {
        --nnz;
        break;
      }
  int ewgts = 0;
  if (rowval)
    for (ind_t j = 1; j < nnz; ++j)
      if (rowval[j] != rowval[j - 1]) {
        ewgts = 1;
        break;
      }
  if (ewgts) {
    fprintf(f, "%" PRIu32 " %" PRIu32 " 01\n", ncols, nnz / 2);
    for (dim_t i = 0; i < nrows; ++i) {
      for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j) {
        if (rowind[j] == i) continue;
        const real_t v = rowval[j];
        /* the reference's v == (real_t)(long long)v, with the conversion
         * defined: integral values inside the range of long long */
        if (v >= -9223372036854775808.0 && v < 9223372036854775808.0 &&
            v == (real_t)(long long)v)
          fprintf(f, " %" PRIu32 " %lld", rowind[j] + 1, (long long)v);
        else
          fprintf(f, " %" PRIu32 " %g", rowind[j] + 1, v);
      }
      fprintf(f, "\n");
    }
  } else {
    fprintf(f, "%" PRIu32 " %" PRIu32 "\n", ncols, nnz / 2);
    for (dim_t i = 0; i < nrows; ++i) {
      for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
        if (rowind[j] != i) fprintf(f, " %" PRIu32, rowind[j] + 1);
      fprintf(f, "\n");
    }
  }
  return fclose(f) == 0 ? GOOSEBERRY_SUCCESS : GOOSEBERRY_ERROR_UNKNOWN;
}

/* src/io.c:1367-1418 (the banner names the reference's version, 0.1.7) */
static int write_mtx(const char *file, dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval)
{
  FILE *f = fopen(file, "w");
  if (!f) return errno_to_rv(errno);
  fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  fprintf(f, "%%\n");
  fprintf(f, "%% Written by Gooseberry version %d.%d.%d\n", GOOSEBERRY_VER_MAJOR,
      GOOSEBERRY_VER_MINOR, GOOSEBERRY_VER_SUBMINOR);
  fprintf(f, "%%   http://cs.umn.edu/~lasalle/gooseberry\n");
  fprintf(f, "%%\n");
  fprintf(f, "%" PRIu32 " %" PRIu32 " %" PRIu32 "\n", nrows, ncols, rowptr[nrows]);
  for (dim_t i = 0; i < nrows; ++i)
    for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
      fprintf(f, "%" PRIu32 "\t%" PRIu32 "\t%lf\n", i + 1, rowind[j] + 1, rowval[j]);
  return fclose(f) == 0 ? GOOSEBERRY_SUCCESS : GOOSEBERRY_ERROR_UNKNOWN;
}

/* src/io.c:1419-1462 */
static int write_point(const char *file, dim_t nrows, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval)
{
  FILE *f = fopen(file, "w");
  if (!f) return errno_to_rv(errno);
  for (dim_t i = 0; i < nrows; ++i)
    for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
      fprintf(f, "%" PRIu32 "\t%" PRIu32 "\t%lf\n", i + 1, rowind[j] + 1, rowval[j]);
  return fclose(f) == 0 ? GOOSEBERRY_SUCCESS : GOOSEBERRY_ERROR_UNKNOWN;
}

/* src/io.c:1464-1549 dispatch */
int gooseberry_read_sparse_matrix(int type, const char *file, dim_t *nrows,
    dim_t *ncols, ind_t **rowptr, dim_t **rowind, real_t **rowval)
{
  if (!nrows || !ncols || !rowptr || !rowind || !rowval)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  switch (type) {
    case GOOSEBERRY_FORMAT_CSR:
      return read_csr(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_MATRIXMARKET:
      return read_mtx(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_BCSR:
      return read_bcsr(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_GRAPH:
      return read_graph(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_SVM:
      return read_svm(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_CLU:
      return read_clu(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_POINT:
      return read_point(file, nrows, ncols, rowptr, rowind, rowval);
    default:
      fprintf(stderr, "Unrecognized sparse matrix format '%d'\n", type);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
}

int gooseberry_read_dense_matrix(int type, const char *file, dim_t *nrows,
    dim_t *ncols, real_t **rowval)
{
  if (!nrows || !ncols || !rowval) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (type != GOOSEBERRY_FORMAT_GRID) {
    fprintf(stderr, "Unrecognized dense matrix format '%d'\n", type);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  return read_grid(file, nrows, ncols, rowval);
}

int gooseberry_write_dense_matrix(int type, const char *file, dim_t nrows,
    dim_t ncols, const real_t *rowval)
{
  if (type != GOOSEBERRY_FORMAT_GRID) {
    fprintf(stderr, "Unrecognized dense matrix format '%d'\n", type);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (!file || (nrows && ncols && !rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return write_grid(file, nrows, ncols, rowval);
}

int gooseberry_write_sparse_matrix(int type, const char *file, dim_t nrows,
    dim_t ncols, const ind_t *rowptr, const dim_t *rowind,
    const real_t *rowval)
{
  if (!file || !rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  switch (type) {
    case GOOSEBERRY_FORMAT_CSR:
      return write_csr(file, nrows, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_BCSR:
      return write_bcsr(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_GRAPH:
      return write_graph(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_MATRIXMARKET:
      if (rowptr[nrows] && (!rowind || !rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
      return write_mtx(file, nrows, ncols, rowptr, rowind, rowval);
    case GOOSEBERRY_FORMAT_POINT:
      if (rowptr[nrows] && (!rowind || !rowval)) return GOOSEBERRY_ERROR_INVALIDINPUT;
      return write_point(file, nrows, rowptr, rowind, rowval);
    default:
      fprintf(stderr, "Unrecognized sparse matrix format '%d'\n", type);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
}

/* src/io.c:1552-1639 */
int gooseberry_read_labels(const char *file, dim_t *r_nrows, dim_t **r_labels)
{
  text_t t;
  int rv = text_load(file, &t);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  char *line;
  ssize_t ll;
  unsigned l;
  ll = text_line(&t, &line);
  while (ll > 0 && is_comment(line[0])) ll = text_line(&t, &line);
  size_t first_pos = 0;
  char *first_line = line;
  ssize_t first_ll = ll;
  (void)first_pos;
  dim_t nl;
  arr_t lab = {0, 0, 0, sizeof(dim_t)};
  dim_t cap = (r_nrows && *r_nrows) ? *r_nrows : (dim_t)-1;
  nl = 0;
  line = first_line;
  ll = first_ll;
  int ok = 1;
  while (ok && ll > 0 && sscanf(line, "%u", &l) == 1) {
    if (nl == cap) {
      fprintf(stderr, "Found excess of rows in permutation file\n");
      break;
    }
    dim_t v = l;
    ok = arr_push(&lab, &v);
    ++nl;
    ll = text_line(&t, &line);
  }
  free(t.buf);
  if (!ok) { free(lab.p); return GOOSEBERRY_ERROR_UNKNOWN; }
  if (r_nrows) *r_nrows = nl;
  if (r_labels) *r_labels = lab.n ? lab.p : calloc(1, sizeof(dim_t));
  else free(lab.p);
  return GOOSEBERRY_SUCCESS;
}

/* src/gooseberry.c:253-306 with the stable counting sort on the GPU */
int gooseberry_read_partition(const char *const file, dim_t *const r_nrows,
    dim_t *const r_nparts, dim_t **const r_map, dim_t **const r_perm,
    dim_t **const r_dist)
{
  GB_API_GUARD();
  dim_t nrows = r_nrows ? *r_nrows : 0;
  dim_t *map = NULL;
  int rv = gooseberry_read_labels(file, &nrows, &map);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  dim_t mx = 0;
  for (dim_t i = 0; i < nrows; ++i) if (map[i] > mx) mx = map[i];
  dim_t nparts = nrows ? mx + 1 : 1;
  dim_t *perm = malloc(sizeof(dim_t) * (nrows ? nrows : 1));
  dim_t *dist = malloc(sizeof(dim_t) * ((size_t)nparts + 1));
  if (!perm || !dist) {
    free(perm); free(dist); free(map);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  rv = gooseberry_countingsort_kv(map, NULL, 0, nparts - 1, nrows, perm, dist);
  if (rv != GOOSEBERRY_SUCCESS) {
    free(perm); free(dist); free(map);
    return rv;
  }
  if (r_nrows) *r_nrows = nrows;
  if (r_nparts) *r_nparts = nparts;
  if (r_dist) *r_dist = dist; else free(dist);
  if (r_perm) *r_perm = perm; else free(perm);
  if (r_map) *r_map = map; else free(map);
  return GOOSEBERRY_SUCCESS;
}

/* The readers and writers with 64-bit row pointers (the reference's
 * GOOSEBERRY_64BIT_INDICES build): the files are read and written through
 * the 32-bit forms, the row pointers widened or narrowed at the boundary. */
int gooseberry_read_sparse_matrix_i64(int type, const char *file,
    dim_t *nrows, dim_t *ncols, uint64_t **rowptr, dim_t **rowind,
    real_t **rowval)
{
  if (!nrows || !rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  ind_t *rp32 = NULL;
  int rv = gooseberry_read_sparse_matrix(type, file, nrows, ncols, &rp32,
      rowind, rowval);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  uint64_t *rp = malloc(((size_t)*nrows + 1) * sizeof(uint64_t));
  if (!rp) {
    free(rp32);
    free(*rowind);
    free(*rowval);
    *rowind = NULL;
    *rowval = NULL;
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  for (size_t i = 0; i <= *nrows; ++i) rp[i] = rp32[i];
  free(rp32);
  *rowptr = rp;
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_write_sparse_matrix_i64(int type, const char *file,
    dim_t nrows, dim_t ncols, const uint64_t *rowptr, const dim_t *rowind,
    const real_t *rowval)
{
  if (!file || !rowptr) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (rowptr[nrows] > UINT32_MAX) return GOOSEBERRY_ERROR_UNIMPLEMENTED;
  ind_t *rp = malloc(((size_t)nrows + 1) * sizeof(ind_t));
  if (!rp) return GOOSEBERRY_ERROR_NOMEMORY;
  for (size_t i = 0; i <= nrows; ++i) {
    if (rowptr[i] > rowptr[nrows]) {
      free(rp);
      return GOOSEBERRY_ERROR_INVALIDINPUT;
    }
    rp[i] = (ind_t)rowptr[i];
  }
  int rv = gooseberry_write_sparse_matrix(type, file, nrows, ncols, rp, rowind,
      rowval);
  free(rp);
  return rv;
}
