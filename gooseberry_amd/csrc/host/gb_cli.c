/*
 * gb_cli.c -- the `gooseberry` command-line program for the sparse-solve
 * path: `gooseberry blas -x multiply ...` and `gooseberry cgd ...`, with the
 * reference's option letters, stdout lines, output formats and exit codes
 * (src/gooseberry_bin.c:1509-1857, 1860-2144, 2842-2877).  The compute runs
 * on the GPU through libgooseberry.
 *
 * Exit codes: 0 success, 1 usage / unknown command, 2 operation failed.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gooseberry.h"

/* ------------------------------------------------------------ options */

typedef enum { OPT_FLAG, OPT_STRING, OPT_INT, OPT_FLOAT, OPT_CHOICE } opt_kind_t;

typedef struct {
  int id;
  char sflag;
  const char *lflag;
  const char *desc;
  opt_kind_t kind;
  const char *const *choices;
} opt_t;

typedef struct {
  int id;
  const char *s;
  long long i;
  double f;
  int o;
} arg_t;

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

typedef struct { double acc, start; int on; } tmr_t;
static void tmr_start(tmr_t *t) { t->start = now(); t->on = 1; }
static void tmr_stop(tmr_t *t) { if (t->on) { t->acc += now() - t->start; t->on = 0; } }

/* Accepted forms: -x value, -xvalue, --name value, --name=value. */
static int parse_args(int argc, char **argv, const opt_t *opts, size_t nopts,
    arg_t *args, size_t *nargs)
{
  size_t n = 0;
  for (int a = 0; a < argc; ++a) {
    const char *s = argv[a];
    const opt_t *o = NULL;
    const char *val = NULL;
    if (s[0] == '-' && s[1] == '-') {
      const char *eq = strchr(s + 2, '=');
      size_t len = eq ? (size_t)(eq - s - 2) : strlen(s + 2);
      for (size_t k = 0; k < nopts; ++k) {
        if (strlen(opts[k].lflag) == len && strncmp(opts[k].lflag, s + 2, len) == 0) {
          o = &opts[k];
          break;
        }
      }
      if (eq) val = eq + 1;
    } else if (s[0] == '-' && s[1]) {
      for (size_t k = 0; k < nopts; ++k) {
        if (opts[k].sflag == s[1]) { o = &opts[k]; break; }
      }
      if (s[2]) val = s + 2;
    }
    if (!o) {
      fprintf(stderr, "Unknown argument '%s'\n", s);
      return 0;
    }
    arg_t *r = &args[n++];
    memset(r, 0, sizeof(*r));
    r->id = o->id;
    if (o->kind == OPT_FLAG) continue;
    if (!val) {
      if (a + 1 >= argc) {
        fprintf(stderr, "Missing value for '%s'\n", s);
        return 0;
      }
      val = argv[++a];
    }
    r->s = val;
    char *end = NULL;
    switch (o->kind) {
      case OPT_INT:
        r->i = strtoll(val, &end, 10);
        if (end == val || *end) { fprintf(stderr, "Bad integer '%s'\n", val); return 0; }
        break;
      case OPT_FLOAT:
        r->f = strtod(val, &end);
        if (end == val || *end) { fprintf(stderr, "Bad number '%s'\n", val); return 0; }
        break;
      case OPT_CHOICE:
        r->o = -1;
        for (int c = 0; o->choices[c]; ++c) {
          if (strcmp(o->choices[c], val) == 0) r->o = c;
        }
        if (r->o < 0) { fprintf(stderr, "Unknown choice '%s'\n", val); return 0; }
        break;
      default:
        break;
    }
  }
  *nargs = n;
  return 1;
}

static void command_usage(const char *prog, const char *cmd, const opt_t *opts,
    size_t nopts, FILE *f)
{
  fprintf(f, "USAGE:\n%s %s [options]\n\nOptions:\n", prog, cmd);
  for (size_t k = 0; k < nopts; ++k) {
    fprintf(f, "  -%c, --%s\n    %s\n", opts[k].sflag, opts[k].lflag, opts[k].desc);
  }
}

/* src/gooseberry_bin.c:440-470: format by file extension */
static int file_type(const char *name)
{
  static const struct { const char *ext; int type; } tbl[] = {
    {"raw", GOOSEBERRY_FORMAT_RAW}, {"mat", GOOSEBERRY_FORMAT_GRID},
    {"grid", GOOSEBERRY_FORMAT_GRID}, {"vec", GOOSEBERRY_FORMAT_GRID},
    {"txt", GOOSEBERRY_FORMAT_GRID}, {"bcsr", GOOSEBERRY_FORMAT_BCSR},
    {"csr", GOOSEBERRY_FORMAT_CSR},
    {"svm", GOOSEBERRY_FORMAT_SVM}, {"libfm", GOOSEBERRY_FORMAT_SVM},
    {"ij", GOOSEBERRY_FORMAT_POINT}, {"point", GOOSEBERRY_FORMAT_POINT},
    {"metis", GOOSEBERRY_FORMAT_GRAPH}, {"chaco", GOOSEBERRY_FORMAT_GRAPH},
    {"graph", GOOSEBERRY_FORMAT_GRAPH}, {"clu", GOOSEBERRY_FORMAT_CLU},
    {"mm", GOOSEBERRY_FORMAT_MATRIXMARKET}, {"mtx", GOOSEBERRY_FORMAT_MATRIXMARKET},
  };
  if (!name) return -1;
  size_t ln = strlen(name);
  for (size_t k = 0; k < sizeof(tbl) / sizeof(tbl[0]); ++k) {
    size_t le = strlen(tbl[k].ext);
    if (ln >= le && strcmp(name + ln - le, tbl[k].ext) == 0) return tbl[k].type;
  }
  return -1;
}

static int is_dense(int t)
{
  return t == GOOSEBERRY_FORMAT_RAW || t == GOOSEBERRY_FORMAT_GRID;
}

static void print_times(const char *names[], const double *vals, int n)
{
  printf("################################ Times ###############################\n");
  for (int i = 0; i < n; ++i) printf(" %s: %0.04lf\n", names[i], vals[i]);
  printf("######################################################################\n");
}

typedef struct {
  int dense;
  dim_t nrows, ncols;
  ind_t *rowptr;
  dim_t *rowind;
  real_t *rowval;
} mat_t;

static void mat_free(mat_t *m)
{
  free(m->rowptr); free(m->rowind); free(m->rowval);
  memset(m, 0, sizeof(*m));
}

/* A dense row-major matrix as CSR of its nonzero entries (row order,
 * columns ascending). */
static int dense_to_csr(mat_t *m)
{
  const size_t nr = m->nrows, nc = m->ncols;
  size_t nnz = 0;
  for (size_t k = 0; k < nr * nc; ++k) nnz += m->rowval[k] != 0;
  if (nnz > 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  ind_t *rp = malloc(sizeof(ind_t) * (nr + 1));
  dim_t *ri = malloc(sizeof(dim_t) * (nnz ? nnz : 1));
  real_t *rv = malloc(sizeof(real_t) * (nnz ? nnz : 1));
  if (!rp || !ri || !rv) {
    free(rp); free(ri); free(rv);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  size_t e = 0;
  rp[0] = 0;
  for (size_t i = 0; i < nr; ++i) {
    for (size_t j = 0; j < nc; ++j) {
      const real_t w = m->rowval[i * nc + j];
      if (w != 0) {
        ri[e] = (dim_t)j;
        rv[e] = w;
        ++e;
      }
    }
    rp[i + 1] = (ind_t)e;
  }
  free(m->rowval);
  m->rowptr = rp;
  m->rowind = ri;
  m->rowval = rv;
  m->dense = 0;
  return GOOSEBERRY_SUCCESS;
}

static int read_input(const char *file, mat_t *m)
{
  int t = file_type(file);
  memset(m, 0, sizeof(*m));
  if (t < 0) {
    fprintf(stderr, "Unknown file format of '%s'\n", file);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  m->dense = is_dense(t);
  if (m->dense)
    return gooseberry_read_dense_matrix(t, file, &m->nrows, &m->ncols, &m->rowval);
  return gooseberry_read_sparse_matrix(t, file, &m->nrows, &m->ncols, &m->rowptr,
      &m->rowind, &m->rowval);
}

/* CSR text, MatrixMarket and binary CSR go straight to a resident matrix:
 * parsed on the GPU with the host reader's semantics (the same arrays, bit
 * for bit) */
static int device_readable(const char *file)
{
  int t = file_type(file);
  return t == GOOSEBERRY_FORMAT_CSR || t == GOOSEBERRY_FORMAT_BCSR ||
      t == GOOSEBERRY_FORMAT_MATRIXMARKET;
}

static int read_resident(const char *file, gooseberry_matrix_t *dmat, mat_t *m)
{
  memset(m, 0, sizeof(*m));
  int err = gooseberry_matrix_read(file_type(file), file, NULL, dmat);
  if (err != GOOSEBERRY_SUCCESS) {
    if (err == GOOSEBERRY_ERROR_INVALIDINPUT)
      fprintf(stderr, "Could not read '%s' as a matrix\n", file);
    return err;
  }
  gooseberry_matrix_info_t info;
  err = gooseberry_matrix_info(*dmat, &info);
  m->nrows = info.nrows;
  m->ncols = info.ncols;
  return err;
}

/* host arrays of a resident matrix (when a host-side step needs them) */
static int download_resident(gooseberry_matrix_t dmat, mat_t *m)
{
  gooseberry_matrix_info_t info;
  int err = gooseberry_matrix_info(dmat, &info);
  if (err != GOOSEBERRY_SUCCESS) return err;
  m->rowptr = malloc(((size_t)info.nrows + 1) * sizeof(ind_t));
  m->rowind = malloc((info.nnz ? info.nnz : 1) * sizeof(dim_t));
  m->rowval = malloc((info.nnz ? info.nnz : 1) * sizeof(real_t));
  if (!m->rowptr || !m->rowind || !m->rowval) return GOOSEBERRY_ERROR_UNKNOWN;
  return gooseberry_matrix_download(dmat, m->rowptr, m->rowind, m->rowval);
}

/* stable counting sort of row ids by label (src/gooseberry_bin.c:2038-2049) */
static int labels_to_perm(const char *file, dim_t n, dim_t **perm)
{
  dim_t cnt = n;
  dim_t *lab = NULL;
  int rv = gooseberry_read_labels(file, &cnt, &lab);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (cnt != n) {
    fprintf(stderr, "Label file '%s' has %u labels for %u rows\n", file, cnt, n);
    free(lab);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  *perm = malloc(sizeof(dim_t) * (n ? n : 1));
  rv = gooseberry_countingsort_kv(lab, NULL, 0, n, n, *perm, NULL);
  free(lab);
  return rv;
}

/* ---------------------------------------------------------------- blas */

enum { BLAS_HELP, BLAS_OPER, BLAS_IN, BLAS_OUT, BLAS_TIME, BLAS_RUNS,
       BLAS_THREADS, BLAS_RPERM, BLAS_CPERM, BLAS_BAND };
static const char *const BLAS_CHOICES[] = {"multiply", NULL};
static const opt_t BLAS_OPTS[] = {
  {BLAS_HELP, 'h', "help", "Display this help page.", OPT_FLAG, NULL},
  {BLAS_OPER, 'x', "operation", "The type of operation to perform.", OPT_CHOICE, BLAS_CHOICES},
  {BLAS_IN, 'i', "infile", "An input matrix/vector file.", OPT_STRING, NULL},
  {BLAS_OUT, 'o', "outfile", "The output matrix/vector file.", OPT_STRING, NULL},
  {BLAS_TIME, 't', "times", "Print timing of the blas routines.", OPT_FLAG, NULL},
  {BLAS_RUNS, 'r', "runs", "Number of repeated runs (only useful for timing purposes).", OPT_INT, NULL},
  {BLAS_THREADS, 'T', "threads", "Number of threads (accepted; the GPU ignores it).", OPT_INT, NULL},
  {BLAS_RPERM, 'R', "rowperm", "Row permutation file.", OPT_STRING, NULL},
  {BLAS_CPERM, 'C', "colperm", "Column permutation file.", OPT_STRING, NULL},
  {BLAS_BAND, 'b', "bandwidthreduce", "Re-order the matrix to reduce bandwidth.", OPT_FLAG, NULL},
};
#define NBLAS_OPTS (sizeof(BLAS_OPTS) / sizeof(BLAS_OPTS[0]))

/* sparse x sparse (src/gooseberry_bin.c:1727-1750) on inputs already
 * permuted: the output file's type picks the dense (blas_spmultsp) or
 * sparse (blas_spmultsp_sp) product; rows are unpermuted afterwards as
 * matrix_unpermute does (src/matrix.c:348-378) */
static int blas_sparse_sparse(mat_t *a, mat_t *b, const char *outfile,
    const dim_t *rperm, const dim_t *blocks, dim_t nblocks, long long runs,
    int times, tmr_t *io, tmr_t *aux, tmr_t *op)
{
  int err = GOOSEBERRY_SUCCESS;
  const int otype = file_type(outfile);
  const dim_t anrows = a->nrows, ancols = a->ncols, bncols = b->ncols;
  real_t *dense = NULL;
  ind_t *cptr = NULL;
  dim_t *cind = NULL, *rren = NULL;
  real_t *cval = NULL;
  if (ancols > b->nrows) {
    fprintf(stderr, "Matrix dimensions do not match for multiplication: "
        "%" PRIu32 "x%" PRIu32 " and %" PRIu32 "x%" PRIu32 "\n",
        anrows, ancols, b->nrows, bncols);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (is_dense(otype)) {
    dense = calloc(((size_t)anrows * bncols) ? (size_t)anrows * bncols : 1,
        sizeof(real_t));
    if (!dense) return GOOSEBERRY_ERROR_NOMEMORY;
  }
  if (times) { tmr_stop(aux); tmr_start(op); }
  for (long long r = 0; r < runs; ++r) {
    if (dense) {
      err = gooseberry_spmultsp(anrows, ancols, bncols, a->rowptr, a->rowind,
          a->rowval, b->rowptr, b->rowind, b->rowval, dense, blocks, nblocks);
    } else {
      free(cptr); free(cind); free(cval);
      cptr = NULL; cind = NULL; cval = NULL;
      err = gooseberry_spmultsp_sp(anrows, ancols, bncols, a->rowptr,
          a->rowind, a->rowval, b->rowptr, b->rowind, b->rowval, &cptr, &cind,
          &cval, blocks, nblocks);
    }
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) { tmr_stop(op); tmr_start(aux); }
  if (rperm) {
    if (dense) {
      err = gooseberry_permute_dense(anrows, bncols, dense, (dim_t *)rperm,
          NULL, 1);
    } else {
      rren = malloc(sizeof(dim_t) * (anrows ? anrows : 1));
      if (!rren) { err = GOOSEBERRY_ERROR_NOMEMORY; goto END; }
      for (dim_t i = 0; i < anrows; ++i) rren[rperm[i]] = i;
      err = gooseberry_permute_sparse(anrows, bncols, cptr, cind, cval, rren,
          NULL);
    }
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) { tmr_stop(aux); tmr_start(io); }
  if (dense) {
    err = gooseberry_write_dense_matrix(GOOSEBERRY_FORMAT_GRID, outfile,
        anrows, bncols, dense);
  } else {
    err = gooseberry_write_sparse_matrix(otype, outfile, anrows, bncols, cptr,
        cind, cval);
  }
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(io);
    const char *names[] = {"I/O", "Auxillary", "Compute"};
    double vals[] = {io->acc, aux->acc, op->acc};
    print_times(names, vals, 3);
  }
END:
  free(dense); free(cptr); free(cind); free(cval); free(rren);
  return err;
}

/* sparse x sparse on resident matrices read on the GPU (no permutations):
 * the product stays on the device until it is written */
static int blas_resident_sparse(gooseberry_matrix_t a, gooseberry_matrix_t b,
    const mat_t *ma, const mat_t *mb, const char *outfile, long long runs,
    int times, tmr_t *aux, tmr_t *op, tmr_t *io)
{
  const dim_t anrows = ma->nrows, ancols = ma->ncols, bncols = mb->ncols;
  if (ancols > mb->nrows) {
    fprintf(stderr, "Matrix dimensions do not match for multiplication: "
        "%" PRIu32 "x%" PRIu32 " and %" PRIu32 "x%" PRIu32 "\n",
        anrows, ancols, mb->nrows, bncols);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  const int otype = file_type(outfile);
  int err = GOOSEBERRY_SUCCESS;
  real_t *dense = NULL;
  void *d_c = NULL;
  gooseberry_matrix_t c = NULL;
  mat_t mc;
  memset(&mc, 0, sizeof(mc));
  const size_t csz = (size_t)anrows * bncols;
  if (is_dense(otype)) {
    dense = calloc(csz ? csz : 1, sizeof(real_t));
    if (!dense) return GOOSEBERRY_ERROR_NOMEMORY;
    err = gooseberry_device_alloc(&d_c, (csz ? csz : 1) * sizeof(real_t));
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) { tmr_stop(aux); tmr_start(op); }
  for (long long r = 0; r < runs; ++r) {
    if (dense) {
      err = gooseberry_matrix_spmultsp_dense(a, b, d_c);
    } else {
      gooseberry_matrix_free(c);
      c = NULL;
      err = gooseberry_matrix_spmultsp(a, b, NULL, &c);
    }
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (dense) err = gooseberry_memcpy_dtoh(dense, d_c, csz * sizeof(real_t));
  else err = download_resident(c, &mc);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) { tmr_stop(op); tmr_start(io); }
  if (dense)
    err = gooseberry_write_dense_matrix(GOOSEBERRY_FORMAT_GRID, outfile, anrows,
        bncols, dense);
  else
    err = gooseberry_write_sparse_matrix(otype, outfile, anrows, bncols,
        mc.rowptr, mc.rowind, mc.rowval);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(io);
    const char *names[] = {"I/O", "Auxillary", "Compute"};
    double vals[] = {io->acc, aux->acc, op->acc};
    print_times(names, vals, 3);
  }
END:
  free(dense);
  gooseberry_device_free(d_c);
  gooseberry_matrix_free(c);
  mat_free(&mc);
  return err;
}

static int cmd_blas(int argc, char **argv)
{
  arg_t args[512];
  size_t nargs = 0;
  if (argc - 2 > 512) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!parse_args(argc - 2, argv + 2, BLAS_OPTS, NBLAS_OPTS, args, &nargs))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  int err = GOOSEBERRY_SUCCESS, times = 0, oper = -1, redband = 0;
  long long runs = 1;
  const char *infiles[256], *outfile = NULL, *rpf = NULL, *cpf = NULL;
  size_t nin = 0;
  mat_t in[2];
  memset(in, 0, sizeof(in));
  dim_t *rperm = NULL, *cperm = NULL, *rdist = NULL, *cdist = NULL, *bperm = NULL;
  real_t *out = NULL;
  void *d_b = NULL, *d_c = NULL;
  gooseberry_matrix_t dmat = NULL, dmatB = NULL;
  tmr_t io = {0}, aux = {0}, op = {0};

  if (nargs < 2) {
    command_usage(argv[0], argv[1], BLAS_OPTS, NBLAS_OPTS, stderr);
    return GOOSEBERRY_SUCCESS;
  }
  for (size_t i = 0; i < nargs; ++i) {
    switch (args[i].id) {
      case BLAS_HELP:
        command_usage(argv[0], argv[1], BLAS_OPTS, NBLAS_OPTS, stdout);
        return GOOSEBERRY_SUCCESS;
      case BLAS_OPER: oper = args[i].o; break;
      case BLAS_IN:
        if (nin == 256) return GOOSEBERRY_ERROR_INVALIDINPUT;
        infiles[nin++] = args[i].s;
        break;
      case BLAS_OUT: outfile = args[i].s; break;
      case BLAS_TIME: times = 1; break;
      case BLAS_RUNS: runs = args[i].i; break;
      case BLAS_THREADS: break;
      case BLAS_RPERM: rpf = args[i].s; break;
      case BLAS_CPERM: cpf = args[i].s; break;
      case BLAS_BAND: redband = 1; break;
    }
  }
  if (nin != 2) {
    fprintf(stderr, "blas needs exactly two input files (matrix, then vector)\n");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (!outfile || file_type(outfile) < 0) {
    fprintf(stderr, "Unknown or missing output file '%s'\n", outfile ? outfile : "");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (times) tmr_start(&io);
  /* without host-side permutations a CSR / BCSR matrix is read on the GPU */
  const int devA = !rpf && !cpf && !redband && device_readable(infiles[0]);
  const int devB = devA && device_readable(infiles[1]);
  err = devA ? read_resident(infiles[0], &dmat, &in[0]) : read_input(infiles[0], &in[0]);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  err = devB ? read_resident(infiles[1], &dmatB, &in[1]) : read_input(infiles[1], &in[1]);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (devA && !devB && !in[1].dense) {
    /* a sparse second input in another format: the host product */
    if ((err = download_resident(dmat, &in[0])) != GOOSEBERRY_SUCCESS) goto END;
    gooseberry_matrix_free(dmat);
    dmat = NULL;
  }
  dim_t nrdist = 1, ncdist = 0;
  if (rpf) {
    dim_t prows = in[0].nrows;
    err = gooseberry_read_partition(rpf, &prows, &nrdist, NULL, &rperm, &rdist);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    printf("Read %" PRIu32 "-way row partition\n", nrdist);
  }
  if (cpf) {
    dim_t prows = in[0].ncols;
    err = gooseberry_read_partition(cpf, &prows, &ncdist, NULL, &cperm, &cdist);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    printf("Read %" PRIu32 "-way column partition\n", ncdist);
  }
  if (times) { tmr_stop(&io); tmr_start(&aux); }
  if (oper != 0) {
    fprintf(stderr, "Unknown operation '%d'\n", oper);
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  if (in[0].dense) {
    fprintf(stderr, "This build multiplies a sparse first matrix only\n");
    err = GOOSEBERRY_ERROR_UNIMPLEMENTED;
    goto END;
  }
  if (!in[1].dense && dmat && dmatB) {
    err = blas_resident_sparse(dmat, dmatB, &in[0], &in[1], outfile, runs, times,
        &aux, &op, &io);
    goto END;
  }
  if (in[1].dense && !is_dense(file_type(outfile))) {
    fprintf(stderr, "A sparse-times-dense product needs a dense output file\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  /* matrix_permute(in[0], rperm, cperm); matrix_permute(in[1], cperm, NULL)
   * (src/gooseberry_bin.c:1655-1656) */
  if (rperm || cperm) {
    err = gooseberry_permute_sparse(in[0].nrows, in[0].ncols, in[0].rowptr,
        in[0].rowind, in[0].rowval, rperm, cperm);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    if (cperm) {
      if (in[1].nrows != in[0].ncols) {
        err = GOOSEBERRY_ERROR_INVALIDINPUT;
        goto END;
      }
      err = in[1].dense
          ? gooseberry_permute_dense(in[1].nrows, in[1].ncols, in[1].rowval, cperm, NULL, 0)
          : gooseberry_permute_sparse(in[1].nrows, in[1].ncols, in[1].rowptr,
                in[1].rowind, in[1].rowval, cperm, NULL);
      if (err != GOOSEBERRY_SUCCESS) goto END;
    }
  }
  /* -b (src/gooseberry_bin.c:1658-1675): Cuthill-McKee within the row
   * partition's blocks, applied to rows and columns of A and rows of B;
   * with both partitions the two permutations are composed with it as the
   * reference composes them, and only rperm undoes the output's order */
  if (redband) {
    const dim_t n = in[0].nrows;
    bperm = malloc(sizeof(dim_t) * (n ? n : 1));
    if (!bperm) { err = GOOSEBERRY_ERROR_NOMEMORY; goto END; }
    err = gooseberry_permute_cuthillmckee(n, in[0].ncols, in[0].rowptr, in[0].rowind,
        in[0].rowval, rpf ? rdist : NULL, rpf ? nrdist : 0, bperm);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    if (rperm && cperm) {
      for (dim_t i = 0; i < n; ++i) rperm[i] = bperm[rperm[i]];
      for (dim_t i = 0; i < n; ++i) cperm[i] = bperm[cperm[i]];
    }
    err = gooseberry_permute_sparse(n, in[0].ncols, in[0].rowptr, in[0].rowind,
        in[0].rowval, bperm, bperm);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    if (in[1].nrows != n) {
      fprintf(stderr, "The second input has %" PRIu32 " rows, the matrix %" PRIu32 "\n",
          in[1].nrows, n);
      err = GOOSEBERRY_ERROR_INVALIDINPUT;
      goto END;
    }
    err = in[1].dense
        ? gooseberry_permute_dense(in[1].nrows, in[1].ncols, in[1].rowval, bperm, NULL, 0)
        : gooseberry_permute_sparse(in[1].nrows, in[1].ncols, in[1].rowptr,
              in[1].rowind, in[1].rowval, bperm, NULL);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (!in[1].dense) {
    err = blas_sparse_sparse(&in[0], &in[1], outfile, rperm, rpf ? rdist : NULL,
        rpf ? nrdist : 0, runs, times, &io, &aux, &op);
    goto END;
  }
  const dim_t anrows = in[0].nrows, ancols = in[0].ncols;
  const dim_t bnrows = in[1].nrows, bncols = in[1].ncols;
  if (ancols > bnrows) {
    fprintf(stderr, "Matrix dimensions do not match for multiplication: "
        "%" PRIu32 "x%" PRIu32 " and %" PRIu32 "x%" PRIu32 "\n",
        anrows, ancols, bnrows, bncols);
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  /* matrix_buildindex: column-major copy of B (src/matrix.c:451-471) */
  size_t bsz = (size_t)bnrows * bncols;
  real_t *bcol = in[1].rowval;
  int bcol_owned = 0;
  if (bncols > 1 && bnrows > 1) {
    bcol = malloc(sizeof(real_t) * bsz);
    bcol_owned = 1;
    for (dim_t i = 0; i < bnrows; ++i)
      for (dim_t j = 0; j < bncols; ++j)
        bcol[(size_t)j * bnrows + i] = in[1].rowval[(size_t)i * bncols + j];
  }
  size_t csz = (size_t)anrows * bncols;
  out = calloc(csz ? csz : 1, sizeof(real_t));
  if (!dmat)
    err = gooseberry_matrix_create(anrows, ancols, in[0].rowptr, in[0].rowind,
        in[0].rowval, NULL, &dmat);
  /* B is read as bcol[j*ancols + l] (src/blas.c:87); keep that bound */
  size_t bneed = bncols ? (size_t)(bncols - 1) * ancols + ancols : 0;
  if (err == GOOSEBERRY_SUCCESS)
    err = gooseberry_device_alloc(&d_b, (bsz > bneed ? bsz : bneed) * sizeof(real_t) + 8);
  if (err == GOOSEBERRY_SUCCESS)
    err = gooseberry_memcpy_htod(d_b, bcol, bsz * sizeof(real_t));
  if (bcol_owned) free(bcol);
  if (err == GOOSEBERRY_SUCCESS) err = gooseberry_device_alloc(&d_c, csz * sizeof(real_t) + 8);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) { tmr_stop(&aux); tmr_start(&op); }
  for (long long r = 0; r < runs; ++r) {
    err = (bncols == 1) ? gooseberry_matrix_spmv(dmat, d_b, d_c)
                        : gooseberry_matrix_spmm(dmat, bncols, d_b, d_c);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  err = gooseberry_memcpy_dtoh(out, d_c, csz * sizeof(real_t));
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) { tmr_stop(&op); tmr_start(&aux); }
  if (rperm) {
    err = gooseberry_permute_dense(anrows, bncols, out, rperm, NULL, 1);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) { tmr_stop(&aux); tmr_start(&io); }
  err = gooseberry_write_dense_matrix(GOOSEBERRY_FORMAT_GRID, outfile, anrows,
      bncols, out);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(&io);
    const char *names[] = {"I/O", "Auxillary", "Compute"};
    double vals[] = {io.acc, aux.acc, op.acc};
    print_times(names, vals, 3);
  }
END:
  gooseberry_device_free(d_b);
  gooseberry_device_free(d_c);
  gooseberry_matrix_free(dmat);
  gooseberry_matrix_free(dmatB);
  free(out);
  free(rperm); free(cperm); free(rdist); free(cdist); free(bperm);
  mat_free(&in[0]);
  mat_free(&in[1]);
  return err;
}

/* ----------------------------------------------------------------- cgd */

enum { CGD_HELP, CGD_IN, CGD_OUT, CGD_ERROR, CGD_NITER, CGD_TIME, CGD_RUNS,
       CGD_THREADS, CGD_RPERM, CGD_CPERM };
static const opt_t CGD_OPTS[] = {
  {CGD_HELP, 'h', "help", "Display this help page.", OPT_FLAG, NULL},
  {CGD_IN, 'i', "infile", "An input matrix/vector file.", OPT_STRING, NULL},
  {CGD_OUT, 'o', "outfile", "The output vector file.", OPT_STRING, NULL},
  {CGD_ERROR, 'e', "error", "The RMSE to achieve before exiting.", OPT_FLOAT, NULL},
  {CGD_NITER, 'I', "iter", "The number of iterations to run before exiting.", OPT_INT, NULL},
  {CGD_TIME, 't', "times", "Print timing of the cgd routines.", OPT_FLAG, NULL},
  {CGD_RUNS, 'r', "runs", "Number of repeated runs (only useful for timing purposes).", OPT_INT, NULL},
  {CGD_THREADS, 'T', "threads", "Number of threads (accepted; the GPU ignores it).", OPT_INT, NULL},
  {CGD_RPERM, 'R', "rowperm", "Row permutation file.", OPT_STRING, NULL},
  {CGD_CPERM, 'C', "colperm", "Column permutation file.", OPT_STRING, NULL},
};
#define NCGD_OPTS (sizeof(CGD_OPTS) / sizeof(CGD_OPTS[0]))

static int cmd_cgd(int argc, char **argv)
{
  arg_t args[512];
  size_t nargs = 0;
  if (argc - 2 > 512) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!parse_args(argc - 2, argv + 2, CGD_OPTS, NCGD_OPTS, args, &nargs))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  int err = GOOSEBERRY_SUCCESS, times = 0;
  double error = 0;
  long long niter = 0, runs = 1;
  const char *matfile = NULL, *vecfile = NULL, *outfile = NULL, *rpf = NULL,
             *cpf = NULL;
  mat_t mat, vec;
  memset(&mat, 0, sizeof(mat));
  memset(&vec, 0, sizeof(vec));
  dim_t *rperm = NULL, *cperm = NULL;
  real_t *out = NULL;
  void *d_b = NULL, *d_x = NULL;
  gooseberry_matrix_t dmat = NULL;
  tmr_t io = {0}, op = {0};

  if (nargs < 2) {
    command_usage(argv[0], argv[1], CGD_OPTS, NCGD_OPTS, stderr);
    return GOOSEBERRY_SUCCESS;
  }
  for (size_t i = 0; i < nargs; ++i) {
    switch (args[i].id) {
      case CGD_HELP:
        command_usage(argv[0], argv[1], CGD_OPTS, NCGD_OPTS, stdout);
        return GOOSEBERRY_SUCCESS;
      case CGD_ERROR: error = args[i].f; break;
      case CGD_NITER: niter = args[i].i; break;
      case CGD_IN:
        if (!matfile) matfile = args[i].s;
        else if (!vecfile) vecfile = args[i].s;
        else {
          fprintf(stderr, "Too many input files specified\n");
          err = GOOSEBERRY_ERROR_INVALIDINPUT;
        }
        break;
      case CGD_OUT: outfile = args[i].s; break;
      case CGD_TIME: times = 1; break;
      case CGD_RUNS: runs = args[i].i; break;
      case CGD_THREADS: break;
      case CGD_RPERM: rpf = args[i].s; break;
      case CGD_CPERM: cpf = args[i].s; break;
    }
  }
  if (!matfile || !vecfile) {
    fprintf(stderr, "You must specify both a matrix input file and a vector "
        "input file (in that order).\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (err == GOOSEBERRY_SUCCESS && !outfile) {
    fprintf(stderr, "You must specify an output file.\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (err == GOOSEBERRY_SUCCESS && niter < 0) err = GOOSEBERRY_ERROR_INVALIDINPUT;
  if (err != GOOSEBERRY_SUCCESS) return err;

  if (times) tmr_start(&io);
  err = (!rpf && !cpf && device_readable(matfile)) ? read_resident(matfile, &dmat, &mat)
                                                  : read_input(matfile, &mat);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if ((err = read_input(vecfile, &vec)) != GOOSEBERRY_SUCCESS) goto END;
  printf("matrix = %" PRIu32 "x%" PRIu32 " vector = %" PRIu32 "x%" PRIu32 "\n",
      mat.nrows, mat.ncols, vec.nrows, vec.ncols);
  if (mat.dense || !vec.dense) {
    fprintf(stderr, "cgd needs a sparse matrix and a dense vector\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  if (rpf && (err = labels_to_perm(rpf, mat.nrows, &rperm)) != GOOSEBERRY_SUCCESS)
    goto END;
  if (cpf && (err = labels_to_perm(cpf, mat.ncols, &cperm)) != GOOSEBERRY_SUCCESS)
    goto END;
  if (times) tmr_stop(&io);
  /* src/cgd.c:51-65 */
  if (mat.nrows != mat.ncols) {
    fprintf(stderr, "Input matrix for CGD is not square: %" PRIu32 "x%" PRIu32 "\n",
        mat.nrows, mat.ncols);
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  if (vec.nrows * (size_t)vec.ncols != mat.nrows || vec.ncols != 1) {
    fprintf(stderr, "Input vector for CGD has wrong number of rows: %" PRIu32 "\n",
        vec.nrows);
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto END;
  }
  if (rperm || cperm) {
    err = gooseberry_permute_sparse(mat.nrows, mat.ncols, mat.rowptr, mat.rowind,
        mat.rowval, rperm, cperm);
    if (err != GOOSEBERRY_SUCCESS) goto END;
    if (rperm) {
      err = gooseberry_permute_dense(vec.nrows, 1, vec.rowval, rperm, NULL, 0);
      if (err != GOOSEBERRY_SUCCESS) goto END;
    }
  }
  const dim_t n = mat.nrows;
  out = calloc(n ? n : 1, sizeof(real_t));
  if (!dmat)
    err = gooseberry_matrix_create(n, n, mat.rowptr, mat.rowind, mat.rowval, NULL,
        &dmat);
  if (err == GOOSEBERRY_SUCCESS) err = gooseberry_device_alloc(&d_b, n * sizeof(real_t));
  if (err == GOOSEBERRY_SUCCESS) err = gooseberry_memcpy_htod(d_b, vec.rowval, n * sizeof(real_t));
  if (err == GOOSEBERRY_SUCCESS) err = gooseberry_device_alloc(&d_x, n * sizeof(real_t));
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) tmr_start(&op);
  for (long long r = 0; r < runs; ++r) {
    gooseberry_cgd_info_t info;
    err = gooseberry_matrix_cgd(dmat, d_b, d_x, error, GOOSEBERRY_TOL_ABSOLUTE,
        (size_t)niter, stdout, &info);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  err = gooseberry_memcpy_dtoh(out, d_x, n * sizeof(real_t));
  if (times) tmr_stop(&op);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (cperm) {
    err = gooseberry_permute_dense(n, 1, out, cperm, NULL, 1);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) tmr_start(&io);
  /* always GRID, whatever the extension (src/gooseberry_bin.c:2102-2103) */
  err = gooseberry_write_dense_matrix(GOOSEBERRY_FORMAT_GRID, outfile, n, 1, out);
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(&io);
    const char *names[] = {"I/O", "Compute"};
    double vals[] = {io.acc, op.acc};
    print_times(names, vals, 2);
  }
END:
  gooseberry_device_free(d_b);
  gooseberry_device_free(d_x);
  gooseberry_matrix_free(dmat);
  free(out);
  free(rperm);
  free(cperm);
  mat_free(&mat);
  mat_free(&vec);
  return err;
}

/* ------------------------------------------------------------ pagerank */

/* src/gooseberry_bin.c:421-444 */
enum { PR_HELP, PR_IN, PR_OUT, PR_TIME, PR_RUNS, PR_NITER, PR_ERROR, PR_DAMPING,
       PR_THREADS, PR_PERM };
static const opt_t PR_OPTS[] = {
  {PR_HELP, 'h', "help", "Display this help page.", OPT_FLAG, NULL},
  {PR_IN, 'i', "infile", "An input matrix file.", OPT_STRING, NULL},
  {PR_OUT, 'o', "outfile", "The output matrix/vector file.", OPT_STRING, NULL},
  {PR_TIME, 't', "times", "Print timing of the pagerank calcuation.", OPT_FLAG, NULL},
  {PR_RUNS, 'r', "runs", "Number of repeated runs (only useful for timing purposes).",
      OPT_INT, NULL},
  {PR_NITER, 'I', "iterations", "Maximum number of iterations.", OPT_INT, NULL},
  {PR_ERROR, 'e', "error", "Error threshold for stopping.", OPT_FLOAT, NULL},
  {PR_DAMPING, 'd', "damping", "Damping factor to use.", OPT_FLOAT, NULL},
  {PR_THREADS, 'T', "threads", "Number of threads.", OPT_INT, NULL},
  {PR_PERM, 'p', "perm", "Row and column permutation file.", OPT_STRING, NULL},
};
static const size_t NPR_OPTS = sizeof(PR_OPTS) / sizeof(PR_OPTS[0]);

/* src/gooseberry_bin.c:2434-2815 */
static int cmd_pagerank(int argc, char **argv)
{
  arg_t args[512];
  size_t nargs = 0;
  if (argc - 2 > 512) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!parse_args(argc - 2, argv + 2, PR_OPTS, NPR_OPTS, args, &nargs))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  int err = GOOSEBERRY_SUCCESS, times = 0;
  double minerror = 0.0, damping = 0.85;
  long long niter = 0, runs = 1;
  const char *matfile = NULL, *outfile = NULL, *pf = NULL;
  mat_t mat;
  memset(&mat, 0, sizeof(mat));
  dim_t *perm = NULL;
  real_t *out = NULL;
  void *d_out = NULL;
  gooseberry_matrix_t dmat = NULL;
  gooseberry_pagerank_info_t info;
  memset(&info, 0, sizeof(info));
  tmr_t io = {0}, pre = {0}, op = {0};

  if (nargs < 2) {
    command_usage(argv[0], argv[1], PR_OPTS, NPR_OPTS, stderr);
    return GOOSEBERRY_SUCCESS;
  }
  for (size_t i = 0; i < nargs; ++i) {
    switch (args[i].id) {
      case PR_HELP:
        command_usage(argv[0], argv[1], PR_OPTS, NPR_OPTS, stdout);
        return GOOSEBERRY_SUCCESS;
      case PR_IN:
        if (!matfile) matfile = args[i].s;
        else {
          fprintf(stderr, "Too many input files specified\n");
          err = GOOSEBERRY_ERROR_INVALIDINPUT;
        }
        break;
      case PR_OUT: outfile = args[i].s; break;
      case PR_TIME: times = 1; break;
      case PR_PERM: pf = args[i].s; break;
      case PR_RUNS: runs = args[i].i; break;
      case PR_NITER: niter = args[i].i; break;
      case PR_DAMPING: damping = args[i].f; break;
      case PR_ERROR: minerror = args[i].f; break;
      case PR_THREADS: break;
    }
  }
  if (!matfile) {
    fprintf(stderr, "You must specify a matrix input file.\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (!outfile) {
    fprintf(stderr, "You must specify an output vector file.\n");
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (err == GOOSEBERRY_SUCCESS && (niter < 0 || runs < 0))
    err = GOOSEBERRY_ERROR_INVALIDINPUT;
  if (err != GOOSEBERRY_SUCCESS) return err;

  if (times) tmr_start(&io);
  if (!pf && device_readable(matfile)) {
    err = read_resident(matfile, &dmat, &mat);
    if (err == GOOSEBERRY_SUCCESS && mat.nrows != mat.ncols) {
      /* squaring (below) works on the host arrays */
      err = download_resident(dmat, &mat);
      gooseberry_matrix_free(dmat);
      dmat = NULL;
    }
  } else {
    err = read_input(matfile, &mat);
  }
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (mat.dense) {
    /* src/gooseberry_bin.c:2561-2565: a dense matrix must be square */
    if (mat.nrows != mat.ncols) {
      fprintf(stderr, "PageRank requires a square matrix, but input matrix '%s' is "
          "%u""x""%u.\n", matfile, (unsigned)mat.nrows, (unsigned)mat.ncols);
      err = GOOSEBERRY_ERROR_INVALIDINPUT;
      goto END;
    }
    /* the weights the reference's dense normalisation counts are the
     * nonzero ones (src/gooseberry_bin.c:2633-2644): run the sparse path on
     * them, in row order with columns ascending */
    if ((err = dense_to_csr(&mat)) != GOOSEBERRY_SUCCESS) goto END;
  }
  /* src/gooseberry_bin.c:2561-2577: a non-square sparse matrix is squared */
  if (mat.nrows > mat.ncols) {
    mat.ncols = mat.nrows;
  } else if (mat.ncols > mat.nrows) {
    ind_t *rp = realloc(mat.rowptr, sizeof(ind_t) * ((size_t)mat.ncols + 1));
    if (!rp) { err = GOOSEBERRY_ERROR_UNKNOWN; goto END; }
    for (dim_t k = mat.nrows + 1; k <= mat.ncols; ++k) rp[k] = rp[mat.nrows];
    mat.rowptr = rp;
    mat.nrows = mat.ncols;
  }
  if (pf && (err = labels_to_perm(pf, mat.nrows, &perm)) != GOOSEBERRY_SUCCESS)
    goto END;
  if (times) {
    tmr_stop(&io);
    tmr_start(&pre);
  }
  const dim_t n = mat.nrows;
  if (perm) {
    err = gooseberry_permute_sparse(n, n, mat.rowptr, mat.rowind, mat.rowval, perm, perm);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  out = calloc(n ? n : 1, sizeof(real_t));
  if (!dmat)
    err = gooseberry_matrix_create(n, n, mat.rowptr, mat.rowind, mat.rowval, NULL, &dmat);
  if (err == GOOSEBERRY_SUCCESS) err = gooseberry_device_alloc(&d_out, (n ? n : 1) * sizeof(real_t));
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(&pre);
    tmr_start(&op);
  }
  for (long long r = 0; r < runs; ++r) {
    err = gooseberry_matrix_pagerank(dmat, damping, (size_t)niter, minerror, d_out, &info);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  err = gooseberry_memcpy_dtoh(out, d_out, n * sizeof(real_t));
  if (err != GOOSEBERRY_SUCCESS) goto END;
  printf("PageRank finished in %zu iterations with an RMSE of %lf\n",
      (size_t)info.niter, info.rmse);
  if (times) tmr_stop(&op);
  if (perm) {
    err = gooseberry_permute_dense(n, 1, out, perm, NULL, 1);
    if (err != GOOSEBERRY_SUCCESS) goto END;
  }
  if (times) tmr_start(&io);
  {
    int t = file_type(outfile);
    if (t < 0) {
      fprintf(stderr, "Unknown file format of '%s'\n", outfile);
      err = GOOSEBERRY_ERROR_INVALIDINPUT;
      goto END;
    }
    if (is_dense(t)) {
      err = gooseberry_write_dense_matrix(t, outfile, n, 1, out);
    } else {
      /* the reference's matrix_sparsify rejects a dense vector
       * (src/matrix.c:185-200), so a sparse output file is an error */
      fprintf(stderr, "Unknown matrix type: %d\n", 0);
      err = GOOSEBERRY_ERROR_INVALIDINPUT;
    }
  }
  if (err != GOOSEBERRY_SUCCESS) goto END;
  if (times) {
    tmr_stop(&io);
    /* this build times the whole power iteration (SpMV included) as one */
    const char *names[] = {"I/O", "Preprocessing", "Compute", "  SpMV"};
    double vals[] = {io.acc, pre.acc, op.acc, op.acc};
    print_times(names, vals, 4);
  }
END:
  gooseberry_device_free(d_out);
  gooseberry_matrix_free(dmat);
  free(out);
  free(perm);
  mat_free(&mat);
  return err;
}

/* ---------------------------------------------------------------- main */

static const struct { const char *name; const char *desc; int provided; } COMMANDS[] = {
  {"help", "Display this help page.", 1},
  {"analysis", "Perform analysis on a matrix.", 0},
  {"transform", "Transform a matrix.", 0},
  {"permute", "Permute a matrix.", 0},
  {"generate", "Generate a matrix/vector.", 0},
  {"blas", "Perform a blas operation.", 1},
  {"cgd", "Perform conjugate gradient descent.", 1},
  {"sgd", "Perform stochastic gradient descent.", 0},
  {"pagerank", "Perform pagerank on a square matrix.", 1},
};
#define NCOMMANDS (sizeof(COMMANDS) / sizeof(COMMANDS[0]))

static void usage(const char *prog, FILE *f)
{
  fprintf(f, "USAGE:\n%s <command> [options]\n\nCommands:\n", prog);
  for (size_t i = 0; i < NCOMMANDS; ++i)
    fprintf(f, "\t%s : %s%s\n", COMMANDS[i].name, COMMANDS[i].desc,
        COMMANDS[i].provided ? "" : " (not provided by this build)");
}

int main(int argc, char **argv)
{
  if (argc < 2) {
    fprintf(stderr, "Must supply a command.\n");
    usage(argv[0], stderr);
    return 1;
  }
  const char *cmd = argv[1];
  int err;
  if (strcmp(cmd, "help") == 0) {
    usage(argv[0], stdout);
    err = GOOSEBERRY_SUCCESS;
  } else if (strcmp(cmd, "blas") == 0) {
    err = cmd_blas(argc, argv);
  } else if (strcmp(cmd, "cgd") == 0) {
    err = cmd_cgd(argc, argv);
  } else if (strcmp(cmd, "pagerank") == 0) {
    err = cmd_pagerank(argc, argv);
  } else {
    size_t i;
    for (i = 0; i < NCOMMANDS; ++i) if (strcmp(cmd, COMMANDS[i].name) == 0) break;
    if (i == NCOMMANDS) {
      fprintf(stderr, "Unrecognized command '%s'.\n", cmd);
      usage(argv[0], stderr);
      return 1;
    }
    fprintf(stderr, "Command '%s' is not provided by this build.\n", cmd);
    err = GOOSEBERRY_ERROR_UNIMPLEMENTED;
  }
  if (err == GOOSEBERRY_SUCCESS) return 0;
  fprintf(stderr, "Operation failed.\n");
  return 2;
}
