/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain C restatement of the reference's sparse-solve path.  Built with
 * -ffp-contract=off so that, like the reference's gcc build at the default
 * -march=x86-64 (no FMA), every multiply and add rounds separately.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker or the timed CPU baseline.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef orc_dim_t dim_t;
typedef orc_ind_t ind_t;
typedef orc_real_t real_t;

/* src/base.h:184 */
#define ORC_BIG_BLOCK 256

void orc_free(void *p) { free(p); }

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int t)
{
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* ------------------------------------------------------------------ BLAS */

/* src/blas.c:67-92: one sequential left-to-right sum per (row, column of B),
 * starting from +0.0, multiply and add rounded separately. */
static void sparse_mult_dense_block(dim_t start, dim_t end, dim_t ancols,
    dim_t bncols, const ind_t *arowptr, const dim_t *arowind,
    const real_t *arowval, const real_t *bcolval, real_t *crowval)
{
  for (dim_t i = start; i < end; ++i) {
    for (dim_t j = 0; j < bncols; ++j) {
      real_t x = 0;
      for (ind_t k = arowptr[i]; k < arowptr[i + 1]; ++k) {
        dim_t l = arowind[k];
        x += arowval[k] * bcolval[(size_t)j * ancols + l];
      }
      crowval[(size_t)i * bncols + j] = x;
    }
  }
}

/* src/blas.c:552-606.  With blocks: round-robin over row blocks
 * (:588-594); without: contiguous equal row-count chunks (:597-601).  The
 * result does not depend on the partition because each row sum is
 * sequential. */
int orc_spmult(dim_t anrows, dim_t ancols, dim_t bncols, const ind_t *arowptr,
    const dim_t *arowind, const real_t *arowval, const real_t *bcolval,
    real_t *crowval, const dim_t *blocks, dim_t nblocks)
{
  #pragma omp parallel
  {
    size_t myid = 0, nthreads = 1;
#ifdef _OPENMP
    myid = (size_t)omp_get_thread_num();
    nthreads = (size_t)omp_get_num_threads();
#endif
    if (blocks) {
      for (size_t c = myid; c < nblocks; c += nthreads) {
        sparse_mult_dense_block(blocks[c], blocks[c + 1], ancols, bncols,
            arowptr, arowind, arowval, bcolval, crowval);
      }
    } else {
      size_t base = anrows / nthreads, rem = anrows % nthreads;
      size_t start = myid * base + (myid < rem ? myid : rem);
      size_t size = base + (myid < rem ? 1 : 0);
      sparse_mult_dense_block((dim_t)start, (dim_t)(start + size), ancols,
          bncols, arowptr, arowind, arowval, bcolval, crowval);
    }
  }
  return ORC_SUCCESS;
}

/* src/blas.c:413-430.  At one thread this is the plain sequential sum; at
 * T threads the combine order is OpenMP's, as in the reference. */
real_t orc_dot(dim_t n, const real_t *a, const real_t *b)
{
  real_t x = 0;
  dim_t i;
  #pragma omp parallel for schedule(static, ORC_BIG_BLOCK) reduction(+:x)
  for (i = 0; i < n; ++i) {
    x += a[i] * b[i];
  }
  return x;
}

/* src/blas.c:244-257 */
int orc_scale(ind_t n, real_t *v, real_t s)
{
  ind_t i;
  #pragma omp parallel for schedule(static, ORC_BIG_BLOCK)
  for (i = 0; i < n; ++i) {
    v[i] *= s;
  }
  return ORC_SUCCESS;
}

/* src/blas.c:275-290 */
int orc_add(dim_t anrows, dim_t ancols, const real_t *a, const real_t *b,
    real_t *c)
{
  ind_t i, n = (ind_t)anrows * (ind_t)ancols;
  #pragma omp parallel for schedule(static, ORC_BIG_BLOCK)
  for (i = 0; i < n; ++i) {
    c[i] = a[i] + b[i];
  }
  return ORC_SUCCESS;
}

/* src/blas.c:321-336 */
int orc_sub(dim_t anrows, dim_t ancols, const real_t *a, const real_t *b,
    real_t *c)
{
  ind_t i, n = (ind_t)anrows * (ind_t)ancols;
  #pragma omp parallel for schedule(static, ORC_BIG_BLOCK)
  for (i = 0; i < n; ++i) {
    c[i] = a[i] - b[i];
  }
  return ORC_SUCCESS;
}

/* -------------------------------------------------------------------- CG */

/* src/cgd.c:40-129, operation for operation:
 *   x=0; r=b; p=r; dr=r.r                                   (:81-85)
 *   loop: ap=A p; alpha=dr/(p.ap)                          (:88-90)
 *         p*=alpha; nx=x+p                                  (:92-93)
 *         ap*=alpha; nr=r-ap                                (:96-97)
 *         dnr=nr.nr; beta=dnr/dr                            (:100-101)
 *         p*=beta/alpha; p=nr+p                             (:104-105)
 *         x=nx; r=nr; dr=dnr; rmse=sqrt(dr); print; stop if rmse<=error
 * niter==0 means no iteration cap (:87). */
int orc_cgd(dim_t nrows, dim_t ncols, const ind_t *rowptr, const dim_t *rowind,
    const real_t *rowval, dim_t bnrows, const real_t *b, dim_t xnrows,
    real_t *x, real_t error, size_t niter, FILE *log, double *trace,
    size_t trace_cap, size_t *iters_out, double *rmse_out)
{
  /* shape checks, src/cgd.c:51-65 */
  if (nrows != ncols || nrows != bnrows || ncols != xnrows) {
    return ORC_ERROR_INVALIDINPUT;
  }
  size_t n = nrows;
  real_t *r = malloc(sizeof(real_t) * (n ? n : 1));
  real_t *p = malloc(sizeof(real_t) * (n ? n : 1));
  real_t *nr = malloc(sizeof(real_t) * (n ? n : 1));
  real_t *nx = malloc(sizeof(real_t) * (n ? n : 1));
  real_t *ap = malloc(sizeof(real_t) * (n ? n : 1));
  size_t iter = 0, done = 0;
  real_t alpha, beta, rmse = 0, dr, dnr;

  memset(x, 0, sizeof(real_t) * n);
  memcpy(r, b, sizeof(real_t) * n);
  memcpy(p, r, sizeof(real_t) * n);
  dr = orc_dot(nrows, r, r);

  for (iter = 0; iter < niter || niter == 0; ++iter) {
    orc_spmult(nrows, ncols, 1, rowptr, rowind, rowval, p, ap, NULL, 0);
    alpha = dr / orc_dot(nrows, p, ap);
    orc_scale(nrows, p, alpha);
    orc_add(nrows, 1, x, p, nx);
    orc_scale(nrows, ap, alpha);
    orc_sub(nrows, 1, r, ap, nr);
    dnr = orc_dot(nrows, nr, nr);
    beta = dnr / dr;
    orc_scale(nrows, p, beta / alpha);
    orc_add(nrows, 1, nr, p, p);
    memcpy(x, nx, sizeof(real_t) * n);
    memcpy(r, nr, sizeof(real_t) * n);
    dr = dnr;
    rmse = sqrt(dr);
    if (log) {
      fprintf(log, "ITER: %zu RMSE: %lf\n", iter, rmse);
    }
    if (trace && iter < trace_cap) {
      trace[3 * iter + 0] = alpha;
      trace[3 * iter + 1] = beta;
      trace[3 * iter + 2] = rmse;
    }
    if (rmse <= error) {
      done = 1;
      break;
    }
  }
  if (iters_out) {
    /* number of SpMVs performed = last ITER index + 1 */
    *iters_out = done ? iter + 1 : iter;
  }
  if (rmse_out) {
    *rmse_out = rmse;
  }
  free(r); free(p); free(nr); free(nx); free(ap);
  return ORC_SUCCESS;
}

/* ------------------------------------------------------------------- I/O */

/* src/io.c:63-69 */
static int is_comment(char c)
{
  return c == '#' || c == '%' || c == '\'' || c == '"' || c == '/';
}

/* src/io.c:108-139 (domlib's open mapped to the reference's error codes) */
static int open_file(const char *name, const char *mode, FILE **f)
{
  *f = fopen(name, mode);
  if (*f) return ORC_SUCCESS;
  switch (errno) {
    case ENOENT: case ENOTDIR: return ORC_ERROR_FILENOTFOUND;
    case EACCES: case EPERM: return ORC_ERROR_PERMISSIONDENIED;
    case EINVAL: case ENAMETOOLONG: return ORC_ERROR_INVALIDINPUT;
    default: return ORC_ERROR_UNKNOWN;
  }
}

/* domlib dl_get_next_line: next line without its line terminator; length, or
 * -1 at end of file. */
static ssize_t next_line(FILE *f, char **buf, size_t *cap)
{
  ssize_t ll = getline(buf, cap, f);
  if (ll < 0) return -1;
  while (ll > 0 && ((*buf)[ll - 1] == '\n' || (*buf)[ll - 1] == '\r')) {
    (*buf)[--ll] = '\0';
  }
  return ll;
}

/* growable arrays */
typedef struct { void *p; size_t n, cap, esz; } vec_t;
static void vec_init(vec_t *v, size_t esz) { v->p = NULL; v->n = 0; v->cap = 0; v->esz = esz; }
static void vec_push(vec_t *v, const void *e)
{
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 64;
    v->p = realloc(v->p, v->cap * v->esz);
  }
  memcpy((char *)v->p + v->n * v->esz, e, v->esz);
  v->n++;
}

/* src/io.c:369-498.  Each line is one row ("col val col val ..."); a blank
 * line is an empty row; comment lines are skipped.  Entries keep file
 * order.  The index base is min(1, smallest index); ncols = max-min+1. */
int orc_read_csr(const char *file, dim_t *nrows, dim_t *ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  FILE *f;
  int rv = open_file(file, "r", &f);
  if (rv != ORC_SUCCESS) return rv;
  char *line = NULL;
  size_t cap = 0;
  ssize_t ll;
  vec_t ptr, ind, val;
  vec_init(&ptr, sizeof(ind_t));
  vec_init(&ind, sizeof(dim_t));
  vec_init(&val, sizeof(real_t));
  dim_t mincol = 1, maxcol = 0, n = 0;
  ind_t j = 0;
  vec_push(&ptr, &j);
  while ((ll = next_line(f, &line, &cap)) >= 0) {
    if (ll != 0 && is_comment(line[0])) continue;
    char *sptr = line, *eptr;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10);
    sptr = eptr;
    real_t v = strtod(sptr, &eptr);
    while (eptr != sptr) {
      if (i > maxcol) maxcol = i;
      if (i < mincol) mincol = i;
      vec_push(&ind, &i);
      vec_push(&val, &v);
      ++j;
      sptr = eptr;
      i = (dim_t)strtoull(sptr, &eptr, 10);
      sptr = eptr;
      v = strtod(sptr, &eptr);
    }
    vec_push(&ptr, &j);
    ++n;
  }
  free(line);
  fclose(f);
  dim_t *rowind = ind.p;
  for (size_t k = 0; k < ind.n; ++k) rowind[k] -= mincol;
  *r_rowptr = ptr.p;
  *r_rowind = ind.n ? ind.p : calloc(1, sizeof(dim_t));
  *r_rowval = val.n ? val.p : calloc(1, sizeof(real_t));
  *nrows = n;
  *ncols = maxcol - mincol + 1;
  return ORC_SUCCESS;
}

/* src/io.c:921-1190.  Coordinate real general/symmetric; symmetric files
 * insert the mirrored entry first and duplicate diagonals (:1137-1144);
 * rows are filled by a stable counting sort in file order. */
int orc_read_mtx(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  FILE *f;
  int rv = open_file(file, "r", &f);
  if (rv != ORC_SUCCESS) return rv;
  char *line = NULL;
  size_t cap = 0;
  ssize_t ll;
  int sym = 0, wgts = 0;
  dim_t nrows = 0, ncols = 0;
  unsigned long long nnzh = 0;
  rv = ORC_ERROR_INVALIDINPUT;
  if ((ll = next_line(f, &line, &cap)) < 1) goto ERR;
  {
    unsigned a, b2, c;
    if (sscanf(line, "%u %u %u", &a, &b2, &c) == 3) {
      nrows = a; ncols = b2; nnzh = c; sym = 0;
    } else {
      const char *ws = " \t";
      char *t = strtok(line, ws);
      if (!t || strcmp(t, "%%MatrixMarket") != 0) goto ERR;
      t = strtok(NULL, ws);
      if (!t || strcmp(t, "matrix") != 0) goto ERR;
      t = strtok(NULL, ws);
      if (!t || strcmp(t, "coordinate") != 0) goto ERR;
      t = strtok(NULL, ws);
      if (!t || strcmp(t, "real") != 0) goto ERR;
      t = strtok(NULL, ws);
      if (!t) goto ERR;
      if (strcmp(t, "general") == 0) sym = 0;
      else if (strcmp(t, "symmetric") == 0) sym = 1;
      else goto ERR;
      while ((ll = next_line(f, &line, &cap)) >= 0) {
        if (ll == 0 || is_comment(line[0])) continue;
        break;
      }
      if (ll < 0) goto ERR;
      char *sptr = line, *eptr;
      nrows = (dim_t)strtoull(sptr, &eptr, 10); sptr = eptr;
      ncols = (dim_t)strtoull(sptr, &eptr, 10); sptr = eptr;
      nnzh = strtoull(sptr, &eptr, 10);
      if (sptr == eptr) goto ERR;
    }
  }
  (void)nnzh;
  long mark = ftell(f);
  ind_t *rowptr = calloc((size_t)nrows + 2, sizeof(ind_t));
  size_t nnz = 0;
  /* first pass: count entries per row (src/io.c:1060-1112) */
  while ((ll = next_line(f, &line, &cap)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    char *sptr = line, *eptr;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10) - 1;
    if (eptr == sptr) { free(rowptr); goto ERR; }
    sptr = eptr;
    dim_t j = (dim_t)strtoull(sptr, &eptr, 10) - 1;
    if (eptr == sptr) { free(rowptr); goto ERR; }
    sptr = eptr;
    (void)strtod(sptr, &eptr);
    if (eptr != sptr) {
      if (nnz == 0) wgts = 1;
      else if (!wgts) { free(rowptr); goto ERR; }
    } else if (wgts) { free(rowptr); goto ERR; }
    if (i >= nrows || j >= ncols || (sym && j >= nrows)) { free(rowptr); goto ERR; }
    if (sym) { ++rowptr[j + 1]; ++nnz; }
    ++rowptr[i + 1];
    ++nnz;
  }
  /* exclusive prefix sum over rowptr[1..nrows] (ind_prefixsum_exc) */
  {
    ind_t s = 0;
    for (size_t k = 1; k <= nrows; ++k) { ind_t c = rowptr[k]; rowptr[k] = s; s += c; }
  }
  dim_t *rowind = malloc(sizeof(dim_t) * (nnz ? nnz : 1));
  real_t *rowval = malloc(sizeof(real_t) * (nnz ? nnz : 1));
  fseek(f, mark, SEEK_SET);
  real_t w = 1;
  while ((ll = next_line(f, &line, &cap)) >= 0) {
    if (ll == 0 || is_comment(line[0])) continue;
    char *sptr = line, *eptr;
    dim_t i = (dim_t)strtoull(sptr, &eptr, 10) - 1; sptr = eptr;
    dim_t j = (dim_t)strtoull(sptr, &eptr, 10) - 1;
    if (wgts) { sptr = eptr; w = strtod(sptr, &eptr); }
    ind_t k;
    if (sym) { k = rowptr[j + 1]++; rowind[k] = i; rowval[k] = w; }
    k = rowptr[i + 1]++;
    rowind[k] = j;
    rowval[k] = w;
  }
  free(line);
  fclose(f);
  *r_nrows = nrows;
  *r_ncols = ncols;
  *r_rowptr = rowptr;
  *r_rowind = rowind;
  *r_rowval = rowval;
  return ORC_SUCCESS;
ERR:
  free(line);
  fclose(f);
  return rv;
}

/* src/io.c:145-239: whitespace-separated values, one row per line; the
 * column count is fixed by the first row. */
int orc_read_grid(const char *file, dim_t *r_nrows, dim_t *r_ncols,
    real_t **r_rowval)
{
  FILE *f;
  int rv = open_file(file, "r", &f);
  if (rv != ORC_SUCCESS) return rv;
  char *line = NULL;
  size_t cap = 0;
  ssize_t ll;
  dim_t nrows = 0, ncols = 0;
  vec_t val;
  vec_init(&val, sizeof(real_t));
  while ((ll = next_line(f, &line, &cap)) >= 0) {
    if (is_comment(line[0])) continue;
    dim_t j = 0;
    char *sptr = line, *eptr;
    real_t v = strtod(sptr, &eptr);
    while (eptr != sptr) {
      vec_push(&val, &v);
      ++j;
      sptr = eptr;
      v = strtod(sptr, &eptr);
    }
    if (ncols == 0) {
      ncols = j;
    } else if (j != ncols) {
      free(line); fclose(f); free(val.p);
      return ORC_ERROR_INVALIDINPUT;
    }
    ++nrows;
  }
  free(line);
  fclose(f);
  if ((size_t)nrows * ncols != val.n) {
    /* a leading blank row counted before the column count was known */
    free(val.p);
    return ORC_ERROR_INVALIDINPUT;
  }
  *r_nrows = nrows;
  *r_ncols = ncols;
  *r_rowval = val.n ? val.p : calloc(1, sizeof(real_t));
  return ORC_SUCCESS;
}

/* src/io.c:1195-1229: %g, space separated, newline per row */
int orc_write_grid(const char *file, dim_t nrows, dim_t ncols,
    const real_t *rowval)
{
  FILE *f;
  int rv = open_file(file, "w", &f);
  if (rv != ORC_SUCCESS) return rv;
  for (dim_t i = 0; i < nrows; ++i) {
    fprintf(f, "%g", rowval[(size_t)i * ncols]);
    for (dim_t j = 1; j < ncols; ++j) {
      fprintf(f, " %g", rowval[(size_t)i * ncols + j]);
    }
    fprintf(f, "\n");
  }
  fclose(f);
  return ORC_SUCCESS;
}

/* src/io.c:1233-1274: 1-based "col val" pairs, %g values */
int orc_write_csr(const char *file, dim_t nrows, const ind_t *rowptr,
    const dim_t *rowind, const real_t *rowval)
{
  FILE *f;
  int rv = open_file(file, "w", &f);
  if (rv != ORC_SUCCESS) return rv;
  for (dim_t i = 0; i < nrows; ++i) {
    if (rowptr[i] < rowptr[i + 1]) {
      fprintf(f, "%u %g", rowind[rowptr[i]] + 1, rowval[rowptr[i]]);
      for (ind_t j = rowptr[i] + 1; j < rowptr[i + 1]; ++j) {
        fprintf(f, " %u %g", rowind[j] + 1, rowval[j]);
      }
    }
    fprintf(f, "\n");
  }
  fclose(f);
  return ORC_SUCCESS;
}

/* src/io.c:1552-1639: leading comment lines skipped; one label per line
 * until a blank or unparsable line; with *nrows > 0 at most that many. */
int orc_read_labels(const char *file, dim_t *r_nrows, dim_t **r_labels)
{
  FILE *f;
  int rv = open_file(file, "r", &f);
  if (rv != ORC_SUCCESS) return rv;
  char *line = NULL;
  size_t cap = 0;
  ssize_t ll;
  unsigned l;
  dim_t nl;
  ll = next_line(f, &line, &cap);
  while (ll > 0 && is_comment(line[0])) ll = next_line(f, &line, &cap);
  if (!r_nrows || *r_nrows == 0) {
    nl = 0;
    while (ll > 0 && sscanf(line, "%u", &l) == 1) {
      ++nl;
      ll = next_line(f, &line, &cap);
    }
    rewind(f);
    ll = next_line(f, &line, &cap);
    while (ll > 0 && is_comment(line[0])) ll = next_line(f, &line, &cap);
  } else {
    nl = *r_nrows;
  }
  dim_t *labels = malloc(sizeof(dim_t) * (nl ? nl : 1));
  dim_t rowcap = nl;
  nl = 0;
  while (ll > 0 && sscanf(line, "%u", &l) == 1) {
    if (nl == rowcap) break;
    labels[nl++] = l;
    ll = next_line(f, &line, &cap);
  }
  fclose(f);
  free(line);
  if (r_nrows) *r_nrows = nl;
  if (r_labels) *r_labels = labels; else free(labels);
  return ORC_SUCCESS;
}

/* ------------------------------------------------------- index work */

void orc_countingsort_kv(const dim_t *keys, const dim_t *vals, dim_t kmin,
    dim_t kmax, size_t n, dim_t *out, dim_t *dist)
{
  size_t nk = (size_t)kmax - kmin + 1;
  size_t *start = calloc(nk + 1, sizeof(size_t));
  for (size_t i = 0; i < n; ++i) start[keys[i] - kmin + 1]++;
  for (size_t k = 1; k <= nk; ++k) start[k] += start[k - 1];
  if (dist) {
    for (size_t k = 0; k <= nk; ++k) dist[k] = (dim_t)start[k];
  }
  for (size_t i = 0; i < n; ++i) {
    out[start[keys[i] - kmin]++] = vals[i];
  }
  free(start);
}

/* src/gooseberry.c:253-306: labels -> nparts = max+1 -> stable counting sort
 * of row ids by label -> perm and dist (nparts+1 boundaries). */
int orc_read_partition(const char *file, dim_t *r_nrows, dim_t *r_nparts,
    dim_t **r_map, dim_t **r_perm, dim_t **r_dist)
{
  dim_t nrows = r_nrows ? *r_nrows : 0;
  dim_t *map = NULL;
  int rv = orc_read_labels(file, &nrows, &map);
  if (rv != ORC_SUCCESS) return rv;
  dim_t mx = 0;
  for (dim_t i = 0; i < nrows; ++i) if (map[i] > mx) mx = map[i];
  dim_t nparts = nrows ? mx + 1 : 1;
  dim_t *perm = malloc(sizeof(dim_t) * (nrows ? nrows : 1));
  dim_t *order = malloc(sizeof(dim_t) * (nrows ? nrows : 1));
  dim_t *dist = malloc(sizeof(dim_t) * ((size_t)nparts + 1));
  for (dim_t i = 0; i < nrows; ++i) order[i] = i;
  orc_countingsort_kv(map, order, 0, nparts - 1, nrows, perm, dist);
  free(order);
  if (r_nrows) *r_nrows = nrows;
  if (r_nparts) *r_nparts = nparts;
  if (r_dist) *r_dist = dist; else free(dist);
  if (r_perm) *r_perm = perm; else free(perm);
  if (r_map) *r_map = map; else free(map);
  return ORC_SUCCESS;
}

/* src/permute.c:120-205: new row i = old row rowperm[i]; column c becomes
 * rename[c] where rename[colperm[i]] = i. */
int orc_permute_sparse(dim_t nrows, dim_t ncols, ind_t *rowptr, dim_t *rowind,
    real_t *rowval, const dim_t *rowperm, const dim_t *colperm)
{
  ind_t nnz = rowptr[nrows];
  dim_t *rename = NULL;
  if (colperm) {
    rename = malloc(sizeof(dim_t) * (ncols ? ncols : 1));
    for (dim_t i = 0; i < ncols; ++i) rename[colperm[i]] = i;
  }
  if (rowperm) {
    ind_t *prowptr = malloc(sizeof(ind_t) * ((size_t)nrows + 1));
    dim_t *prowind = malloc(sizeof(dim_t) * (nnz ? nnz : 1));
    real_t *prowval = malloc(sizeof(real_t) * (nnz ? nnz : 1));
    prowptr[0] = 0;
    for (dim_t i = 0; i < nrows; ++i) {
      dim_t v = rowperm[i];
      prowptr[i + 1] = prowptr[i] + (rowptr[v + 1] - rowptr[v]);
    }
    ind_t k = 0;
    for (dim_t i = 0; i < nrows; ++i) {
      dim_t v = rowperm[i];
      for (ind_t j = rowptr[v]; j < rowptr[v + 1]; ++j) {
        prowind[k] = rename ? rename[rowind[j]] : rowind[j];
        prowval[k] = rowval[j];
        ++k;
      }
    }
    memcpy(rowptr, prowptr, sizeof(ind_t) * nrows);
    memcpy(rowind, prowind, sizeof(dim_t) * nnz);
    memcpy(rowval, prowval, sizeof(real_t) * nnz);
    free(prowptr); free(prowind); free(prowval);
  } else if (colperm) {
    for (ind_t j = 0; j < nnz; ++j) rowind[j] = rename[rowind[j]];
  }
  free(rename);
  return ORC_SUCCESS;
}

/* src/permute.c:208-245: new[i][j] = old[rowperm[i]][colperm[j]] */
int orc_permute_dense(dim_t nrows, dim_t ncols, real_t *rowval,
    const dim_t *rowperm, const dim_t *colperm)
{
  if (!rowperm && !colperm) return ORC_SUCCESS;
  size_t tot = (size_t)nrows * ncols;
  real_t *p = malloc(sizeof(real_t) * (tot ? tot : 1));
  for (dim_t i = 0; i < nrows; ++i) {
    dim_t v = rowperm ? rowperm[i] : i;
    for (dim_t j = 0; j < ncols; ++j) {
      dim_t u = colperm ? colperm[j] : j;
      p[(size_t)i * ncols + j] = rowval[(size_t)v * ncols + u];
    }
  }
  memcpy(rowval, p, sizeof(real_t) * tot);
  free(p);
  return ORC_SUCCESS;
}

/* src/permute.c:248-285: old[rowperm[i]][colperm[j]] = new[i][j] */
int orc_permute_dense_rev(dim_t nrows, dim_t ncols, real_t *rowval,
    const dim_t *rowperm, const dim_t *colperm)
{
  if (!rowperm && !colperm) return ORC_SUCCESS;
  size_t tot = (size_t)nrows * ncols;
  real_t *p = malloc(sizeof(real_t) * (tot ? tot : 1));
  for (dim_t i = 0; i < nrows; ++i) {
    dim_t v = rowperm ? rowperm[i] : i;
    for (dim_t j = 0; j < ncols; ++j) {
      dim_t u = colperm ? colperm[j] : j;
      p[(size_t)v * ncols + u] = rowval[(size_t)i * ncols + j];
    }
  }
  memcpy(rowval, p, sizeof(real_t) * tot);
  free(p);
  return ORC_SUCCESS;
}

/* src/blas.c:397-411 */
int orc_add_scalar(ind_t n, real_t *v, real_t s)
{
  #pragma omp parallel for schedule(static, 256)
  for (ind_t i = 0; i < n; ++i) v[i] += s;
  return ORC_SUCCESS;
}

/* src/gooseberry_bin.c:2622-2641 (sparse matrix case) */
int orc_pagerank_prepare(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, real_t *rowval, dim_t *sinks, dim_t *nsinks)
{
  real_t *indeg = calloc(ncols ? ncols : 1, sizeof(real_t));
  if (!indeg) return ORC_ERROR_UNKNOWN;
  dim_t ns = 0;
  for (dim_t k = 0; k < nrows; ++k) {
    if (rowptr[k] == rowptr[k + 1]) {
      sinks[ns++] = k;
    } else {
      for (ind_t l = rowptr[k]; l < rowptr[k + 1]; ++l)
        indeg[rowind[l]] += rowval[l];
    }
  }
  for (dim_t k = 0; k < nrows; ++k)
    for (ind_t l = rowptr[k]; l < rowptr[k + 1]; ++l)
      rowval[l] /= indeg[rowind[l]];
  free(indeg);
  *nsinks = ns;
  return ORC_SUCCESS;
}

static real_t pr_sqerr(dim_t n, const real_t *a, const real_t *b)
{
  real_t e = 0.0, d;
  for (dim_t k = 0; k < n; ++k) {
    d = a[k] - b[k];
    e += d * d;
  }
  return e;
}

/* src/gooseberry_bin.c:2669-2737 */
int orc_pagerank(dim_t n, const ind_t *rowptr, const dim_t *rowind,
    const real_t *nrowval, const dim_t *sinks, dim_t nsinks, real_t damping,
    size_t niter, real_t minerror, real_t *out, size_t *iters, real_t *rmse)
{
  real_t *a = malloc(sizeof(real_t) * (n ? n : 1));
  real_t *b = malloc(sizeof(real_t) * (n ? n : 1));
  if (!a || !b) { free(a); free(b); return ORC_ERROR_UNKNOWN; }
  real_t *rank = a, *o = b, *t;
  real_t error = 0.0, dist;
  size_t i;
  for (dim_t k = 0; k < n; ++k) rank[k] = 1.0 / n;
  for (i = 0; niter == 0 || i < niter; ++i) {
    orc_spmult(n, n, 1, rowptr, rowind, nrowval, rank, o, NULL, 0);
    orc_scale(n, o, damping);
    dist = 0;
    for (dim_t k = 0; k < nsinks; ++k) dist += rank[sinks[k]];
    orc_add_scalar(n, o, ((1.0 - damping) / n) + dist);
    if (i % 10 == 0) {
      error = pr_sqerr(n, rank, o);
      if (error <= minerror * minerror) {
        ++i;
        goto finish;
      }
    }
    t = rank; rank = o; o = t;
  }
  error = pr_sqerr(n, rank, o);
finish:
  *rmse = sqrt(error);
  *iters = i;
  memcpy(out, o, sizeof(real_t) * n);
  free(a);
  free(b);
  return ORC_SUCCESS;
}

/* src/blas.c:94-121 (one row range), driven as src/blas.c:609-648 */
int orc_spmultsp(dim_t anrows, dim_t ancols, dim_t bncols, const ind_t *arowptr,
    const dim_t *arowind, const real_t *arowval, const ind_t *browptr,
    const dim_t *browind, const real_t *browval, real_t *crowval)
{
  (void)ancols;
  #pragma omp parallel for schedule(static)
  for (dim_t i = 0; i < anrows; ++i) {
    real_t *c = crowval + (size_t)i * bncols;
    for (dim_t m = 0; m < bncols; ++m) c[m] = 0;
    for (ind_t j = arowptr[i]; j < arowptr[i + 1]; ++j) {
      const dim_t l = arowind[j];
      const real_t x = arowval[j];
      for (ind_t k = browptr[l]; k < browptr[l + 1]; ++k)
        c[browind[k]] += x * browval[k];
    }
  }
  return ORC_SUCCESS;
}

/* src/blas.c:124-234 (count, fill) driven as src/blas.c:651-791 without
 * blocks; sequential here (the result does not depend on the threads) */
int orc_spmultsp_sp(dim_t anrows, dim_t ancols, dim_t bncols,
    const ind_t *arowptr, const dim_t *arowind, const real_t *arowval,
    const ind_t *browptr, const dim_t *browind, const real_t *browval,
    ind_t **r_crowptr, dim_t **r_crowind, real_t **r_crowval)
{
  (void)ancols;
  ind_t *crowptr = calloc((size_t)anrows + 1, sizeof(ind_t));
  int *row = calloc(bncols ? bncols : 1, sizeof(int));
  dim_t *mark = malloc(sizeof(dim_t) * (bncols ? bncols : 1));
  if (!crowptr || !row || !mark) { free(crowptr); free(row); free(mark); return ORC_ERROR_UNKNOWN; }
  /* count */
  for (dim_t i = 0; i < anrows; ++i) {
    dim_t rowsize = 0;
    for (ind_t j = arowptr[i]; j < arowptr[i + 1]; ++j) {
      const dim_t l = arowind[j];
      for (ind_t k = browptr[l]; k < browptr[l + 1]; ++k) {
        const dim_t m = browind[k];
        if (!row[m]) { row[m] = 1; mark[rowsize++] = m; }
      }
    }
    crowptr[i + 1] = crowptr[i] + rowsize;
    for (dim_t m = 0; m < rowsize; ++m) row[mark[m]] = 0;
  }
  free(row);
  free(mark);
  const ind_t nnz = crowptr[anrows];
  dim_t *crowind = malloc(sizeof(dim_t) * (nnz ? nnz : 1));
  real_t *crowval = malloc(sizeof(real_t) * (nnz ? nnz : 1));
  ind_t *pos = malloc(sizeof(ind_t) * (bncols ? bncols : 1));
  if (!crowind || !crowval || !pos) {
    free(crowptr); free(crowind); free(crowval); free(pos);
    return ORC_ERROR_UNKNOWN;
  }
  const ind_t NONE = (ind_t)-1;
  for (dim_t m = 0; m < bncols; ++m) pos[m] = NONE;
  /* fill */
  for (dim_t i = 0; i < anrows; ++i) {
    ind_t ptr = crowptr[i];
    for (ind_t j = arowptr[i]; j < arowptr[i + 1]; ++j) {
      const dim_t l = arowind[j];
      const real_t x = arowval[j];
      for (ind_t k = browptr[l]; k < browptr[l + 1]; ++k) {
        const dim_t m = browind[k];
        ind_t q = pos[m];
        if (q == NONE) {
          q = ptr++;
          crowind[q] = m;
          crowval[q] = x * browval[k];
          pos[m] = q;
        } else {
          crowval[q] += x * browval[k];
        }
      }
    }
    for (ind_t q = crowptr[i]; q < crowptr[i + 1]; ++q) pos[crowind[q]] = NONE;
  }
  free(pos);
  *r_crowptr = crowptr;
  *r_crowind = crowind;
  *r_crowval = crowval;
  return ORC_SUCCESS;
}

/* ---------------------------------------------------------- Cuthill-McKee */

/* A min-heap of (key, vertex) pairs with membership, standing in for
 * domlib's dd_pq (DLPQ_MIN, src/permute.c:29-39).  domlib is not in the
 * reference tree, so its order among equal keys is not known: here equal
 * keys leave in ascending vertex order. */
typedef struct {
  dim_t *key, *val, *pos;  /* pos[v - lo]: heap slot of v, or NONE */
  dim_t n, lo;
} orc_heap_t;

#define ORC_NONE ((dim_t)-1)

static int hp_less(const orc_heap_t *h, dim_t a, dim_t b)
{
  return h->key[a] < h->key[b] || (h->key[a] == h->key[b] && h->val[a] < h->val[b]);
}

static void hp_swap(orc_heap_t *h, dim_t a, dim_t b)
{
  dim_t k = h->key[a], v = h->val[a];
  h->key[a] = h->key[b];
  h->val[a] = h->val[b];
  h->key[b] = k;
  h->val[b] = v;
  h->pos[h->val[a] - h->lo] = a;
  h->pos[h->val[b] - h->lo] = b;
}

static void hp_up(orc_heap_t *h, dim_t i)
{
  while (i > 0) {
    dim_t p = (i - 1) / 2;
    if (!hp_less(h, i, p)) break;
    hp_swap(h, i, p);
    i = p;
  }
}

static void hp_down(orc_heap_t *h, dim_t i)
{
  for (;;) {
    dim_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->n && hp_less(h, l, m)) m = l;
    if (r < h->n && hp_less(h, r, m)) m = r;
    if (m == i) return;
    hp_swap(h, i, m);
    i = m;
  }
}

static int hp_init(orc_heap_t *h, dim_t lo, dim_t hi)
{
  size_t cap = (size_t)(hi - lo) + 1;
  h->key = malloc(cap * sizeof(dim_t));
  h->val = malloc(cap * sizeof(dim_t));
  h->pos = malloc(cap * sizeof(dim_t));
  h->n = 0;
  h->lo = lo;
  if (!h->key || !h->val || !h->pos) return 0;
  for (size_t i = 0; i < cap; ++i) h->pos[i] = ORC_NONE;
  return 1;
}

static void hp_free(orc_heap_t *h)
{
  free(h->key); free(h->val); free(h->pos);
}

static void hp_push(orc_heap_t *h, dim_t key, dim_t v)
{
  dim_t i = h->n++;
  h->key[i] = key;
  h->val[i] = v;
  h->pos[v - h->lo] = i;
  hp_up(h, i);
}

static dim_t hp_pop(orc_heap_t *h)
{
  dim_t v = h->val[0];
  h->pos[v - h->lo] = ORC_NONE;
  if (--h->n > 0) {
    h->key[0] = h->key[h->n];
    h->val[0] = h->val[h->n];
    h->pos[h->val[0] - h->lo] = 0;
    hp_down(h, 0);
  }
  return v;
}

static void hp_remove(orc_heap_t *h, dim_t v)
{
  dim_t i = h->pos[v - h->lo];
  h->pos[v - h->lo] = ORC_NONE;
  if (i != --h->n) {
    const dim_t mv = h->val[h->n];  /* the last element fills the hole */
    h->key[i] = h->key[h->n];
    h->val[i] = mv;
    h->pos[mv - h->lo] = i;
    hp_up(h, i);
    hp_down(h, h->pos[mv - h->lo]);
  }
}

/* src/permute.c:49-110: Cuthill-McKee on the rows [start, end), counting
 * only the entries whose column lies in the block */
static int cm_block(dim_t start, dim_t end, const ind_t *rowptr,
    const dim_t *rowind, dim_t *perm)
{
  if (start >= end) return ORC_SUCCESS;
  orc_heap_t q, rem;
  dim_t *deg = malloc(sizeof(dim_t) * (size_t)(end - start));
  int ok = deg && hp_init(&q, start, end) && hp_init(&rem, start, end);
  if (!ok) { free(deg); return ORC_ERROR_UNKNOWN; }
  for (dim_t i = start; i < end; ++i) {
    dim_t d = 0;
    for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j) {
      dim_t k = rowind[j];
      if (k < end && k >= start) ++d;
    }
    deg[i - start] = d;
    hp_push(&rem, d, i);
  }
  dim_t sr = start, nordered = start;
  while (rem.n > 0) {
    dim_t i = hp_pop(&rem);
    perm[nordered++] = i;
    while (sr < nordered) {
      i = perm[sr++];
      for (ind_t j = rowptr[i]; j < rowptr[i + 1]; ++j) {
        dim_t k = rowind[j];
        if (k < end && k >= start && rem.pos[k - start] != ORC_NONE) {
          hp_remove(&rem, k);
          hp_push(&q, deg[k - start], k);
        }
      }
      while (q.n > 0) perm[nordered++] = hp_pop(&q);
    }
  }
  hp_free(&q);
  hp_free(&rem);
  free(deg);
  return ORC_SUCCESS;
}

/* src/permute.c:288-332: perm[p] = the row placed at position p */
int orc_cuthillmckee(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind, const dim_t *blocks, dim_t nblocks, dim_t *perm)
{
  if (nrows != ncols) return ORC_ERROR_INVALIDINPUT;
  if (blocks) {
    for (dim_t c = 0; c < nblocks; ++c) {
      int rv = cm_block(blocks[c], blocks[c + 1], rowptr, rowind, perm);
      if (rv != ORC_SUCCESS) return rv;
    }
    return ORC_SUCCESS;
  }
  return cm_block(0, nrows, rowptr, rowind, perm);
}
