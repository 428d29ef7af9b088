/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C (C99 + OpenMP) restatement of the reference's sparse-solve path
 * (gooseberry: src/blas.c, src/cgd.c, src/io.c, src/permute.c,
 * src/gooseberry.c).  It exists so that tests, the smoke check and the CPU
 * baseline leg of the benchmark can check and time the HIP product against
 * the reference's algorithm.  Nothing in the product (gooseberry_amd/, the
 * CLI, libgooseberry.so) includes, links or calls this code.
 *
 * Types follow the reference's FP64 build: dim_t/ind_t are uint32_t
 * (include/gooseberry.h:45-71 without GOOSEBERRY_SIGNED_TYPES/64BIT macros),
 * real_t is double (include/gooseberry.h:73-79 with GOOSEBERRY_DOUBLE_REALS).
 *
 * Pinning status of each function (what checks it besides itself) is listed
 * in tests/test_oracle_pinning.py; a function with no independent check is
 * marked "parity unpinned" below.
 *   - orc_countingsort_kv: the `dist` convention of domlib's
 *     dd_countingsort_kv is not visible (domlib is an empty submodule); the
 *     stable order is pinned by brute force, the dist convention is this
 *     build's documented choice (parity unpinned for dist).
 *   - line splitting (domlib dl_get_next_line) is restated as "strip one
 *     trailing \n and any trailing \r"; parity unpinned for \r handling.
 */
#ifndef GB_ORACLE_H
#define GB_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

typedef uint32_t orc_dim_t;
typedef uint32_t orc_ind_t;
typedef double orc_real_t;

/* error codes: identical values to include/gooseberry.h:85-94 */
enum {
  ORC_SUCCESS = 1,
  ORC_ERROR_UNIMPLEMENTED,
  ORC_ERROR_INVALIDINPUT,
  ORC_ERROR_READ,
  ORC_ERROR_WRITE,
  ORC_ERROR_FILENOTFOUND,
  ORC_ERROR_PERMISSIONDENIED,
  ORC_ERROR_UNKNOWN
};

/* src/blas.c:67-92 + 552-606 */
int orc_spmult(orc_dim_t anrows, orc_dim_t ancols, orc_dim_t bncols,
    const orc_ind_t *arowptr, const orc_dim_t *arowind,
    const orc_real_t *arowval, const orc_real_t *bcolval,
    orc_real_t *crowval, const orc_dim_t *blocks, orc_dim_t nblocks);

/* src/blas.c:413-430 */
orc_real_t orc_dot(orc_dim_t n, const orc_real_t *a, const orc_real_t *b);
/* src/blas.c:244-257 */
int orc_scale(orc_ind_t n, orc_real_t *v, orc_real_t s);
/* src/blas.c:275-290 */
int orc_add(orc_dim_t anrows, orc_dim_t ancols, const orc_real_t *a,
    const orc_real_t *b, orc_real_t *c);
/* src/blas.c:321-336 */
int orc_sub(orc_dim_t anrows, orc_dim_t ancols, const orc_real_t *a,
    const orc_real_t *b, orc_real_t *c);

/* src/cgd.c:40-129.  trace (optional) receives 3 doubles per iteration:
 * alpha, beta, sqrt(r.r).  iters_out = number of SpMVs performed.
 * log (optional) receives the reference's "ITER: i RMSE: v" lines. */
int orc_cgd(orc_dim_t nrows, orc_dim_t ncols, const orc_ind_t *rowptr,
    const orc_dim_t *rowind, const orc_real_t *rowval, orc_dim_t bnrows,
    const orc_real_t *b, orc_dim_t xnrows, orc_real_t *x, orc_real_t error,
    size_t niter, FILE *log, double *trace, size_t trace_cap,
    size_t *iters_out, double *rmse_out);

/* src/io.c:369-498 */
int orc_read_csr(const char *file, orc_dim_t *nrows, orc_dim_t *ncols,
    orc_ind_t **rowptr, orc_dim_t **rowind, orc_real_t **rowval);
/* src/io.c:921-1190 */
int orc_read_mtx(const char *file, orc_dim_t *nrows, orc_dim_t *ncols,
    orc_ind_t **rowptr, orc_dim_t **rowind, orc_real_t **rowval);
/* src/io.c:145-239 */
int orc_read_grid(const char *file, orc_dim_t *nrows, orc_dim_t *ncols,
    orc_real_t **rowval);
/* src/io.c:1195-1229 */
int orc_write_grid(const char *file, orc_dim_t nrows, orc_dim_t ncols,
    const orc_real_t *rowval);
/* src/io.c:1233-1274 */
int orc_write_csr(const char *file, orc_dim_t nrows, const orc_ind_t *rowptr,
    const orc_dim_t *rowind, const orc_real_t *rowval);
/* src/io.c:1552-1639 */
int orc_read_labels(const char *file, orc_dim_t *nrows, orc_dim_t **labels);

/* stable counting sort of vals by keys in [kmin,kmax] as called at
 * src/gooseberry.c:275 and src/gooseberry_bin.c:2038-2049.  dist (optional,
 * kmax-kmin+2 entries) receives the start of each key's run and n at the
 * end. */
void orc_countingsort_kv(const orc_dim_t *keys, const orc_dim_t *vals,
    orc_dim_t kmin, orc_dim_t kmax, size_t n, orc_dim_t *out,
    orc_dim_t *dist);
/* src/gooseberry.c:253-306 */
int orc_read_partition(const char *file, orc_dim_t *nrows, orc_dim_t *nparts,
    orc_dim_t **map, orc_dim_t **perm, orc_dim_t **dist);

/* src/permute.c:120-205 */
int orc_permute_sparse(orc_dim_t nrows, orc_dim_t ncols, orc_ind_t *rowptr,
    orc_dim_t *rowind, orc_real_t *rowval, const orc_dim_t *rowperm,
    const orc_dim_t *colperm);
/* src/permute.c:208-245 */
int orc_permute_dense(orc_dim_t nrows, orc_dim_t ncols, orc_real_t *rowval,
    const orc_dim_t *rowperm, const orc_dim_t *colperm);
/* src/permute.c:248-285 */
int orc_permute_dense_rev(orc_dim_t nrows, orc_dim_t ncols,
    orc_real_t *rowval, const orc_dim_t *rowperm, const orc_dim_t *colperm);

/* src/blas.c:94-121 + 609-648: C (dense, row-major anrows x bncols) = A B
 * with A and B sparse; each C element starts at 0 and adds the products in
 * the scan order of A's row and B's rows. */
int orc_spmultsp(orc_dim_t anrows, orc_dim_t ancols, orc_dim_t bncols,
    const orc_ind_t *arowptr, const orc_dim_t *arowind,
    const orc_real_t *arowval, const orc_ind_t *browptr,
    const orc_dim_t *browind, const orc_real_t *browval,
    orc_real_t *crowval);

/* src/blas.c:124-234 + 651-791: sparse C = A B (two passes, count and
 * fill); the columns of a C row in order of first appearance in the scan,
 * each value the first product plus the later ones in scan order.  The
 * three arrays are malloc'd (free with orc_free). */
int orc_spmultsp_sp(orc_dim_t anrows, orc_dim_t ancols, orc_dim_t bncols,
    const orc_ind_t *arowptr, const orc_dim_t *arowind,
    const orc_real_t *arowval, const orc_ind_t *browptr,
    const orc_dim_t *browind, const orc_real_t *browval,
    orc_ind_t **crowptr, orc_dim_t **crowind, orc_real_t **crowval);

/* src/blas.c:397-411: v[i] += s */
int orc_add_scalar(orc_ind_t n, orc_real_t *v, orc_real_t s);

/* src/gooseberry_bin.c:2622-2641 (sparse case): indeg[c] = sum of the
 * values in column c, added in row order; every value is divided by its
 * column's sum; sinks = the rows without entries, ascending.  rowval is
 * normalised in place; sinks holds nrows entries, *nsinks is set. */
int orc_pagerank_prepare(orc_dim_t nrows, orc_dim_t ncols,
    const orc_ind_t *rowptr, const orc_dim_t *rowind, orc_real_t *rowval,
    orc_dim_t *sinks, orc_dim_t *nsinks);

/* src/gooseberry_bin.c:2669-2737: one run of the power iteration on the
 * normalised matrix (square, n rows).  niter == 0 runs until the error
 * test passes.  out receives the vector the reference writes (its `out`
 * after the last swap), *iters the printed iteration count and *rmse the
 * printed RMSE.  The error sums are sequential (the reference's OpenMP
 * reduction at one thread). */
int orc_pagerank(orc_dim_t n, const orc_ind_t *rowptr, const orc_dim_t *rowind,
    const orc_real_t *nrowval, const orc_dim_t *sinks, orc_dim_t nsinks,
    orc_real_t damping, size_t niter, orc_real_t minerror, orc_real_t *out,
    size_t *iters, orc_real_t *rmse);

/* src/permute.c:49-110,288-332 (Cuthill-McKee; equal degrees in ascending
 * vertex order, see oracle.c) */
int orc_cuthillmckee(orc_dim_t nrows, orc_dim_t ncols, const orc_ind_t *rowptr,
    const orc_dim_t *rowind, const orc_dim_t *blocks, orc_dim_t nblocks,
    orc_dim_t *perm);
void orc_free(void *p);
int orc_num_threads(void);
void orc_set_num_threads(int t);

#endif
