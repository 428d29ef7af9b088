/*
 * gooseberry.h -- public C API of the MI355X (gfx950) sparse-solve library.
 *
 * Drop-in for the sparse-solve path of the reference's include/gooseberry.h:
 * the same types, error and format enums and the same prototypes for the
 * functions on that path, backed by hand-written HIP kernels on one GPU.
 * New entry points (device-resident matrices, the exported CG solver,
 * mixed precision, batched solves and the solve server) are appended.
 *
 * ABI: this library is built with FP64 values (the reference's
 * GOOSEBERRY_DOUBLE_REALS build, include/gooseberry.h:73-79) and unsigned
 * 32-bit indices (include/gooseberry.h:45-71, default widths).
 *
 * Devices and streams: the library drives device 0 of the process through
 * one library stream.  Functions that take host pointers copy in, compute
 * and copy out before they return.  Functions taking device pointers
 * (named d_*) enqueue work on the library stream; results are ready after
 * gooseberry_device_synchronize() or any function that returns host data.
 *
 * Threads: any host thread may call any function.  One library lock
 * serialises the calls (the stream, the block cache, the graphs and the
 * workspaces kept with matrices are shared), so a call from a second
 * thread starts when the running call returns; the work a d_* call enqueues
 * is ordered on the one stream.  Host-pointer calls such as gooseberry_spmult
 * own no state between calls and may be made from several threads at once
 * (they then take turns on the GPU), as the reference's may.  Only the
 * file readers and writers (host work) run without the lock.
 */
#ifndef GOOSEBERRY_H
#define GOOSEBERRY_H

#include <stddef.h>
#include <stdint.h>
#include <inttypes.h>

#define GOOSEBERRY_VER_MAJOR 0
#define GOOSEBERRY_VER_MINOR 1
#define GOOSEBERRY_VER_SUBMINOR 7

/* ---------------------------------------------------------------- types */
/* include/gooseberry.h:43-82: values FP64, dimensions uint32, and row
 * pointers (ind_t) uint32 -- or uint64 when GOOSEBERRY_64BIT_INDICES is
 * defined (include/gooseberry.h:65-71): the library exports both row-pointer
 * widths, and under that macro the names of the functions taking ind_t map
 * onto their _i64 forms (below, "64-bit row pointers"). */
#ifndef BOWSTRING_TYPES_DEFINED
#if defined(GOOSEBERRY_SIGNED_TYPES) || defined(GOOSEBERRY_64BIT_DIMENSIONS)
#error "this library is built with unsigned 32-bit dim_t and unsigned ind_t"
#endif
typedef uint64_t gooseberry_int64_t;
#define PF_BSINT64_T "%" PRIu64
typedef uint32_t gooseberry_int32_t;
#define PF_BSINT32_T "%" PRIu32
typedef gooseberry_int32_t dim_t;
#define PF_DIM_T PF_BSINT32_T
#ifdef GOOSEBERRY_64BIT_INDICES
typedef gooseberry_int64_t ind_t;
#define PF_IND_T PF_BSINT64_T
#else
typedef gooseberry_int32_t ind_t;
#define PF_IND_T PF_BSINT32_T
#endif
#ifndef GOOSEBERRY_DOUBLE_REALS
#define GOOSEBERRY_DOUBLE_REALS 1
#endif
typedef double real_t;
#define PF_REAL_T "%lf"
#define PF_MINREAL_T "%g"
#endif

#ifdef GOOSEBERRY_64BIT_INDICES
/* ind_t is 64-bit: the reference's names resolve to the _i64 entry points,
 * whose prototypes below are the same with ind_t = uint64_t */
#define gooseberry_scale gooseberry_scale_i64
#define gooseberry_add_scalar gooseberry_add_scalar_i64
#define gooseberry_spmult gooseberry_spmult_i64
#define gooseberry_spmultsp gooseberry_spmultsp_i64
#define gooseberry_spmultsp_sp gooseberry_spmultsp_sp_i64
#define gooseberry_read_sparse_matrix gooseberry_read_sparse_matrix_i64
#define gooseberry_write_sparse_matrix gooseberry_write_sparse_matrix_i64
#define gooseberry_matrix_create gooseberry_matrix_create_i64
#define gooseberry_matrix_create_device gooseberry_matrix_create_device_i64
#define gooseberry_cgd gooseberry_cgd_i64
#endif

/* include/gooseberry.h:85-94; codes after GOOSEBERRY_ERROR_UNKNOWN are new */
typedef enum gooseberry_error_t {
  GOOSEBERRY_SUCCESS=1,
  GOOSEBERRY_ERROR_UNIMPLEMENTED,
  GOOSEBERRY_ERROR_INVALIDINPUT,
  GOOSEBERRY_ERROR_READ,
  GOOSEBERRY_ERROR_WRITE,
  GOOSEBERRY_ERROR_FILENOTFOUND,
  GOOSEBERRY_ERROR_PERMISSIONDENIED,
  GOOSEBERRY_ERROR_UNKNOWN,
  GOOSEBERRY_ERROR_DEVICE,      /* a HIP call failed */
  GOOSEBERRY_ERROR_NOMEMORY,    /* device memory or admission budget */
  GOOSEBERRY_ERROR_BREAKDOWN    /* CG produced a non-finite or zero p.Ap */
} gooseberry_error_t;

/* include/gooseberry.h:97-106 */
typedef enum gooseberry_format_t {
  GOOSEBERRY_FORMAT_RAW,
  GOOSEBERRY_FORMAT_GRID,
  GOOSEBERRY_FORMAT_CSR,
  GOOSEBERRY_FORMAT_POINT,
  GOOSEBERRY_FORMAT_SVM,
  GOOSEBERRY_FORMAT_GRAPH,
  GOOSEBERRY_FORMAT_CLU,
  GOOSEBERRY_FORMAT_MATRIXMARKET,
  GOOSEBERRY_FORMAT_BCSR          /* new: binary CSR, see below */
} gooseberry_format_t;

#ifdef __cplusplus
extern "C" {
#endif

/* ================================================== reference entry points */

/* a <- s*a on n values (include/gooseberry.h:134-137; src/blas.c:244-257).
 * Host pointer; computed on the GPU. */
int gooseberry_scale(
    ind_t n,
    real_t * val,
    real_t s);

/* val[i] += s on n values (include/gooseberry.h:150-153; src/blas.c:397-411).
 * Host pointer; computed on the GPU. */
int gooseberry_add_scalar(
    ind_t n,
    real_t * val,
    real_t s);

/* C = A x B (include/gooseberry.h:174-184; src/gooseberry.c:53-67,
 * src/blas.c:552-606).  A is CSR (anrows x ancols), B is column-major
 * (ancols x bncols, i.e. B^t rows), C row-major (anrows x bncols).  All
 * pointers are host pointers owned by the caller.  blocks (nblocks + 1
 * non-decreasing row boundaries, the last at most anrows) select the rows
 * the reference computes: rows [blocks[0], blocks[nblocks]) of C are
 * written and the other rows of crowval are left as they were; NULL
 * computes every row.  Row i is entries arowptr[i]..arowptr[i+1]-1, also
 * when arowptr[0] is not 0, as in the reference's loop.  Each row sum of a
 * row with at most 64 entries is bitwise the reference's left-to-right sum.
 * Every call uploads the rows it computes, builds their SpMV plan (without
 * the 16-bit offsets and the value table, which only pay back over many
 * SpMVs) and frees it again: callers that multiply by the same A repeatedly
 * should keep it resident (gooseberry_matrix_create,
 * gooseberry_matrix_spmv). */
int gooseberry_spmult(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const ind_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const real_t * bcolval,
    real_t * crowval,
    const dim_t * blocks,
    dim_t nblocks);

/* include/gooseberry.h:187-216; src/blas.c:609-791.  C = A B for a sparse
 * A (anrows x ancols) and a sparse B (ancols x bncols, so browptr holds
 * ancols + 1 entries).  Every C element is the sum of its products in the
 * reference's scan order (row i of A entry by entry, each selecting a row
 * of B entry by entry), so results are bitwise the reference's.
 * gooseberry_spmultsp writes the dense row-major C into crowval: with
 * blocks, only rows [blocks[0], blocks[nblocks]) are written (the
 * reference writes exactly the block rows).  gooseberry_spmultsp_sp returns
 * a new CSR C (malloc'd arrays, free with free()), each row's columns in
 * the order of their first product as the reference's fill loop leaves
 * them.  Its blocks, when given, must partition [0, anrows): the reference
 * computes every row whatever the blocks and its row pointers are only
 * whole in that case.  The product count must stay below 2^32 - 1
 * (GOOSEBERRY_ERROR_INVALIDINPUT beyond). */
int gooseberry_spmultsp(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const ind_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const ind_t * browptr,
    const dim_t * browind,
    const real_t * browval,
    real_t * crowval,
    const dim_t * blocks,
    dim_t nblocks);

int gooseberry_spmultsp_sp(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const ind_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const ind_t * browptr,
    const dim_t * browind,
    const real_t * browval,
    ind_t ** crowptr,
    dim_t ** crowind,
    real_t ** crowval,
    const dim_t * blocks,
    dim_t nblocks);

/* src/permute.c:288-332 (not in the reference's public header; the CLI's
 * -b option): the Cuthill-McKee order of a square matrix's rows, per
 * diagonal block when blocks (nblocks + 1 monotone row boundaries) are
 * given; perm[p] = the row placed at position p, and positions outside the
 * blocks keep their values.  Each component starts at its least-degree row
 * and each row places its unplaced neighbours by ascending degree (entries
 * in the block counted); equal degrees are taken in ascending row order.
 * Host arrays; computed on the GPU. */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_permute_cuthillmckee(
    dim_t nrows,
    dim_t ncols,
    const ind_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    const dim_t * blocks,
    dim_t nblocks,
    dim_t * perm);
#endif

/* include/gooseberry.h IO prototypes; src/io.c:1464-1549.  Sparse readers:
 * CSR, MATRIXMARKET, GRAPH, CLU, SVM, POINT and BCSR; sparse writers: CSR,
 * GRAPH, MATRIXMARKET, POINT and BCSR.  Dense: GRID.  Arrays returned
 * through pointers are malloc'd; free with free() or gooseberry_free().
 * Where the reference's readers would write past their arrays or keep an
 * index wrapped below zero or beyond the matrix (more rows or entries than
 * a header announces, a 0 where indices are 1-based, a column past ncols),
 * these return GOOSEBERRY_ERROR_INVALIDINPUT with a message on stderr.
 *
 * GOOSEBERRY_FORMAT_BCSR (new, extension .bcsr) is a binary CSR: the 8
 * bytes "GBCSR\0\1\0", uint32 nrows, uint32 ncols, uint64 nnz, then
 * rowptr (nrows + 1 uint32), rowind (nnz uint32) and rowval (nnz
 * float64), little-endian, nothing else. */
int gooseberry_read_sparse_matrix(
    int type,
    const char * file,
    dim_t * nrows,
    dim_t * ncols,
    ind_t ** rowptr,
    dim_t ** rowind,
    real_t ** rowval);

int gooseberry_read_dense_matrix(
    int type,
    const char * file,
    dim_t * nrows,
    dim_t * ncols,
    real_t ** rowval);

int gooseberry_write_dense_matrix(
    int type,
    const char * file,
    dim_t nrows,
    dim_t ncols,
    const real_t * rowval);

int gooseberry_write_sparse_matrix(
    int type,
    const char * file,
    dim_t nrows,
    dim_t ncols,
    const ind_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval);

/* New: a sparse matrix file straight into device memory.  CSR text and
 * MatrixMarket bodies are parsed on the GPU with the reader semantics of
 * src/io.c:369-498 and :921-1190 -- the arrays are those of
 * gooseberry_read_sparse_matrix, bit for bit -- and BCSR is copied through
 * pinned buffers and checked on the GPU (other formats:
 * GOOSEBERRY_ERROR_UNIMPLEMENTED).  The arrays are device memory (free with
 * gooseberry_device_free). */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_read_sparse_matrix_device(
    int type,
    const char * file,
    dim_t * nrows,
    dim_t * ncols,
    ind_t ** d_rowptr,
    dim_t ** d_rowind,
    real_t ** d_rowval);
#endif

/* The same parse of CSR text already in device memory (len bytes). */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_parse_csr_device(
    const char * d_text,
    size_t len,
    dim_t * nrows,
    dim_t * ncols,
    ind_t ** d_rowptr,
    dim_t ** d_rowind,
    real_t ** d_rowval);
#endif

/* src/io.c:1552-1639 */
int gooseberry_read_labels(
    const char * file,
    dim_t * nrows,
    dim_t ** labels);

/* src/gooseberry.c:253-306: labels -> stable counting sort (on the GPU) ->
 * perm and dist (nparts+1 entries). */
int gooseberry_read_partition(
    const char * const file,
    dim_t * const r_nrows,
    dim_t * const r_nparts,
    dim_t ** const r_map,
    dim_t ** const r_perm,
    dim_t ** const r_dist);

/* ======================================================== new entry points */

void gooseberry_free(void * ptr);
const char * gooseberry_strerror(int err);
/* text of the last device error seen by the library ("" if none) */
const char * gooseberry_last_device_error(void);

/* Fault injection, a test hook with no counterpart in the reference: after
 * `after` more successful device allocations (GOOSEBERRY_FAULT_ALLOC) or
 * kernel launch checks (GOOSEBERRY_FAULT_LAUNCH), the next one reports a
 * failure -- GOOSEBERRY_ERROR_NOMEMORY or GOOSEBERRY_ERROR_DEVICE from the
 * entry point that made it -- once; after < 0 disarms.  An injected launch
 * failure is reported after the kernel has run normally, so no kernel is
 * ever left half-launched.  The library stays usable after an injected
 * failure; an interrupted solve must be started again. */
#define GOOSEBERRY_FAULT_ALLOC 0
#define GOOSEBERRY_FAULT_LAUNCH 1
int gooseberry_inject_fault(
    int kind,
    long long after);

/* ---------------------------------------------------- index work (GPU) */

/* Stable counting sort of vals by keys in [kmin, kmax] (host pointers):
 * out[] receives vals in key order, input order kept within a key; dist
 * (optional) receives kmax-kmin+2 run boundaries.  vals == NULL means
 * 0..n-1.  The reference's counterpart is domlib's dd_countingsort_kv as
 * called at src/gooseberry.c:275 and src/gooseberry_bin.c:2038-2049. */
int gooseberry_countingsort_kv(
    const dim_t * keys,
    const dim_t * vals,
    dim_t kmin,
    dim_t kmax,
    size_t n,
    dim_t * out,
    dim_t * dist);

/* In-place CSR permutation, src/permute.c:120-205 (host pointers): new row
 * i is old row rowperm[i]; column c becomes i where colperm[i] == c. */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_permute_sparse(
    dim_t nrows,
    dim_t ncols,
    ind_t * rowptr,
    dim_t * rowind,
    real_t * rowval,
    const dim_t * rowperm,
    const dim_t * colperm);
#endif

/* In-place dense permutation, src/permute.c:208-245 (rev == 0) and its
 * inverse src/permute.c:248-285 (rev != 0). */
int gooseberry_permute_dense(
    dim_t nrows,
    dim_t ncols,
    real_t * rowval,
    const dim_t * rowperm,
    const dim_t * colperm,
    int rev);

/* -------------------------------------------------------- device memory */
int gooseberry_device_alloc(void ** d_ptr, size_t bytes);
int gooseberry_device_free(void * d_ptr);
int gooseberry_memcpy_htod(void * d_dst, const void * h_src, size_t bytes);
int gooseberry_memcpy_dtoh(void * h_dst, const void * d_src, size_t bytes);
/* device-to-device copy, enqueued on the library stream */
int gooseberry_memcpy_dtod(void * d_dst, const void * d_src, size_t bytes);
int gooseberry_device_synchronize(void);
int gooseberry_device_memory(size_t * free_bytes, size_t * total_bytes);
/* timing events on the library stream */
int gooseberry_event_create(void ** ev);
int gooseberry_event_destroy(void * ev);
int gooseberry_event_record(void * ev);
int gooseberry_event_elapsed_ms(void * start, void * stop, float * ms);

/* ------------------------------------------- device-resident CSR matrices */
typedef struct gooseberry_matrix_s * gooseberry_matrix_t;

/* SpMV plan knobs (0 = choose from the row-length histogram). */
typedef struct gooseberry_plan_t {
  dim_t short_max;  /* rows up to this length: LDS-staged row blocks */
  dim_t wave_max;   /* rows up to this length: one wavefront per row */
  dim_t window;     /* nonzeros per LDS-staged row block */
  dim_t rowcap;     /* rows per LDS-staged row block */
  dim_t exact;      /* 1: rows of any length are summed left to right from
                       +0.0 (bitwise the reference for every row); 0: rows
                       longer than short_max use tree reductions */
  dim_t grid;       /* cap on SpMV workgroups (0: one per row block) */
  dim_t xcd_remap;  /* 1: each XCD walks a contiguous range of blocks;
                       C >= 2: chunks of C consecutive blocks go to the
                       XCDs in turn (same results, another order) */
  dim_t csplit;     /* > 1: split the columns into this many slices, one
                       launch per slice into a partial vector, partials
                       added after; keeps the gathered slice of x in L2 for
                       matrices with scattered columns (not with exact) */
  dim_t wide_index; /* 1: the row blocks read the 32-bit column indices;
                       0: they read each entry's column as a 16-bit offset
                       from the block's first row (2 bytes an entry less),
                       when every offset fits (same results either way) */
  dim_t wide_values; /* 1: the values are read as stored; 0: when the matrix
                       has at most 256 distinct values (and no column split),
                       each entry's value is read as an 8-bit index into a
                       table of them (1 byte an entry instead of 8; CSR-VI,
                       the same values, so the same results) */
  dim_t xtile;      /* C (a power of two, 256..8192): the FP64 SpMV stages x
                       through LDS in tiles of C columns; each workgroup owns
                       a block of rows and keeps its y in LDS, adding one
                       tile's products at a time (matrices with scattered
                       columns; a row's sum is regrouped by tile, as with a
                       column split, and the layout costs 12 bytes an entry
                       more); 0: off.  Not with csplit or exact. */
} gooseberry_plan_t;

/* Upload a host CSR matrix to HBM and build its SpMV plan on the GPU. */
int gooseberry_matrix_create(
    dim_t nrows,
    dim_t ncols,
    const ind_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

/* Same from device arrays (copied; the caller keeps its buffers). */
int gooseberry_matrix_create_device(
    dim_t nrows,
    dim_t ncols,
    const ind_t * d_rowptr,
    const dim_t * d_rowind,
    const real_t * d_rowval,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

/* A resident matrix read from a file as gooseberry_read_sparse_matrix_device
 * reads it, planned with `plan` (NULL: defaults). */
int gooseberry_matrix_read(
    int type,
    const char * file,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * mat);

/* GOOSEBERRY_ERROR_INVALIDINPUT (and nothing freed) while a batch or a
 * server holds the matrix in one of its slots: remove the solve first. */
int gooseberry_matrix_free(gooseberry_matrix_t mat);

/* Rebuild the SpMV plan (NULL: defaults).  INVALIDINPUT while a batch or a
 * server holds the matrix (its slot reads the plan's arrays). */
int gooseberry_matrix_replan(
    gooseberry_matrix_t mat,
    const gooseberry_plan_t * plan);

/* Choose the SpMV plan by timing candidates drawn from the row-length
 * histogram (HIP events over `reps` SpMVs each, 0 = 5; FP64 SpMV, or FP32
 * for GOOSEBERRY_PRECISION_MIXED) and keep the fastest; *r_plan and *r_ms
 * (may be NULL) receive it and its SpMV time.  Results do not depend on the
 * plan beyond the SpMV's rounding tolerance.  INVALIDINPUT while the matrix
 * is in a batch or a server (they hold its plan arrays). */
int gooseberry_matrix_autotune(
    gooseberry_matrix_t mat,
    int precision,
    uint32_t reps,
    gooseberry_plan_t * r_plan,
    double * r_ms);

typedef struct gooseberry_matrix_info_t {
  dim_t nrows, ncols;
  uint64_t nnz;
  gooseberry_plan_t plan;
  dim_t nblocks_stream;   /* LDS-staged row blocks */
  dim_t nrows_wave;       /* rows handled one wavefront per row */
  dim_t nrows_block;      /* rows handled one workgroup per row */
  uint64_t hist[33];      /* rows with floor(log2(len+1)) == b */
  dim_t index16;          /* 1: the row blocks read 16-bit column offsets */
  dim_t values8;          /* distinct values of a value-indexed plan (0: the
                             values are read as stored) */
} gooseberry_matrix_info_t;

int gooseberry_matrix_info(
    gooseberry_matrix_t mat,
    gooseberry_matrix_info_t * info);

/* y = A x, device vectors, enqueued on the library stream. */
int gooseberry_matrix_spmv(
    gooseberry_matrix_t mat,
    const real_t * d_x,
    real_t * d_y);

/* C = A B, device arrays (B column-major, C row-major). */
int gooseberry_matrix_spmm(
    gooseberry_matrix_t mat,
    dim_t bncols,
    const real_t * d_b,
    real_t * d_c);

/* ------------------------------------------------------ conjugate gradient */
#define GOOSEBERRY_TOL_ABSOLUTE 0   /* stop when ||r||_2 <= tol (reference) */
#define GOOSEBERRY_TOL_RELATIVE 1   /* stop when ||r||_2 <= tol * ||b||_2 */

typedef struct gooseberry_cgd_info_t {
  uint64_t niter;      /* SpMVs performed (inner SpMVs for mixed precision) */
  double rnorm;        /* final recurrence residual sqrt(r.r) */
  double bnorm;        /* ||b||_2 */
  int converged;       /* rnorm reached the threshold */
  int status;          /* 1 converged, 2 iteration cap, 3 breakdown */
  uint64_t nrefine;    /* outer refinement steps (mixed precision) */
  double true_rel_residual; /* ||b - A x|| / ||b|| in FP64 (mixed only) */
} gooseberry_cgd_info_t;

/* CG of src/cgd.c:40-129 on a resident matrix: x0 = 0, the reference's
 * update order, stop when sqrt(r.r) <= threshold or after niter SpMVs
 * (niter == 0: no cap).  b == 0 returns x = 0 at once.  d_b, d_x device
 * vectors.  log (a FILE *, may be NULL) receives the reference's
 * "ITER: i RMSE: v" lines. */
int gooseberry_matrix_cgd(
    gooseberry_matrix_t mat,
    const real_t * d_b,
    real_t * d_x,
    real_t tol,
    int tol_mode,
    size_t niter,
    void * log,
    gooseberry_cgd_info_t * info);

/* gooseberry_matrix_cgd without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * work vectors (r, p, Ap: 24 bytes per row), the partial sums and the state
 * live in one file-scope device array of
 * gooseberry_cgd_noalloc_max_bytes() bytes (128 MiB: about 5.5 million
 * rows) that each GPU holds until the process ends; a larger system returns
 * GOOSEBERRY_ERROR_INVALIDINPUT before any launch.  Kernels are launched one
 * by one (no graphs); no log.  Every call on a GPU shares that array, so
 * calls that overlap on one GPU race (the library lock serialises the calls
 * of one process).  Same results, bit for bit, as gooseberry_matrix_cgd. */
int gooseberry_matrix_cgd_noalloc(
    gooseberry_matrix_t mat,
    const real_t * d_b,
    real_t * d_x,
    real_t tol,
    int tol_mode,
    size_t niter,
    gooseberry_cgd_info_t * info);

/* Bytes of the allocation-free CG's device array. */
size_t gooseberry_cgd_noalloc_max_bytes(void);

/* Mixed precision for one system (a one-system batch, see the batched
 * solves below): FP32 inner iterations, FP64 dots, solution and reliable
 * updates; converged on the FP64 residual.  info->niter counts inner
 * iterations, info->nrefine the FP64 refreshes. */
int gooseberry_matrix_cgd_mixed(
    gooseberry_matrix_t mat,
    const real_t * d_b,
    real_t * d_x,
    real_t tol,
    int tol_mode,
    size_t niter,
    gooseberry_cgd_info_t * info);

/* y = A x in FP32 (the matrix's FP32 value copy, made on first use), device
 * vectors; rows of up to 64 entries are the row-sequential FP32 sums of the
 * reference's float build. */
int gooseberry_matrix_spmv_f32(
    gooseberry_matrix_t mat,
    const float * d_x,
    float * d_y);

/* include/gooseberry.h:174-184 for the reference's float build (real_t =
 * float, the default without GOOSEBERRY_DOUBLE_REALS): the same call as
 * gooseberry_spmult with FP32 values and vectors, host pointers. */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_spmult_f32(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const ind_t * arowptr,
    const dim_t * arowind,
    const float * arowval,
    const float * bcolval,
    float * crowval,
    const dim_t * blocks,
    dim_t nblocks);
#endif

/* Host-pointer form of the reference's internal cgd() (src/cgd.h:29-34),
 * absolute threshold `error` as in the reference. */
int gooseberry_cgd(
    dim_t nrows,
    dim_t ncols,
    const ind_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    const real_t * b,
    real_t * x,
    real_t error,
    size_t niter,
    gooseberry_cgd_info_t * info);

/* ------------------------------------------------------- batched solves */
/* Independent CG solves advanced together: one device step advances every
 * active system by one CG iteration (or one FP64 residual refresh) through
 * work lists over all systems, replayed from captured hipGraphs.  Every
 * sum of a system runs over that system's own fixed decomposition, so its
 * result is the same bits alone and in any mix.
 *
 * GOOSEBERRY_PRECISION_FP64: the reference's CG (src/cgd.c:40-129).
 * GOOSEBERRY_PRECISION_MIXED: FP32 matrix values and inner vectors, dot
 * products accumulated in FP64, FP64 solution; when the FP32 recurrence
 * residual has dropped by `delta` (batch parameter) the solution update is
 * folded into the FP64 x and the residual recomputed as b - A x with the FP64
 * matrix (reliable update); convergence is declared on that FP64 residual.
 * GOOSEBERRY_PRECISION_MIXED_JACOBI: MIXED with the inner FP32 iterations on
 * the Jacobi-scaled system D^-1/2 A D^-1/2 (D = diag(A)), i.e. Jacobi-
 * preconditioned CG; the FP64 refreshes and the convergence test use A and
 * b unscaled.  Needs a positive, finite diagonal in every row
 * (GOOSEBERRY_ERROR_INVALIDINPUT otherwise); the matrix keeps the scaled
 * FP32 values (4 bytes per entry) and D^-1/2 (8 bytes per row). */
#define GOOSEBERRY_PRECISION_FP64 0
#define GOOSEBERRY_PRECISION_MIXED 1
#define GOOSEBERRY_PRECISION_MIXED_JACOBI 2

typedef struct gooseberry_batch_s * gooseberry_batch_t;

/* delta <= 0 (or >= 1) selects 0.05: a refresh every 1.3 decades of the
 * FP32 recurrence residual keeps the Krylov direction consistent with the
 * refreshed residual (much smaller deltas let the two drift apart and the
 * FP32 recurrence can stall on ill-conditioned systems). */
int gooseberry_batch_create(
    uint32_t max_systems,
    double delta,
    gooseberry_batch_t * r_batch);

int gooseberry_batch_free(gooseberry_batch_t batch);

/* Add a solve of mat x = b (device vectors, x overwritten; mat, d_b, d_x
 * must stay alive until the solve is removed).  The batch makes the FP32
 * copy of the values for MIXED. */
int gooseberry_batch_add(
    gooseberry_batch_t batch,
    gooseberry_matrix_t mat,
    const real_t * d_b,
    real_t * d_x,
    real_t tol,
    int tol_mode,
    size_t niter,
    int precision,
    uint32_t * r_slot);

int gooseberry_batch_remove(
    gooseberry_batch_t batch,
    uint32_t slot);

/* Advance all unfinished systems by up to `steps` steps; returns the number
 * still running. */
int gooseberry_batch_run(
    gooseberry_batch_t batch,
    uint32_t steps,
    uint32_t * r_running);

/* Run until every system has stopped. */
int gooseberry_batch_solve(gooseberry_batch_t batch);

/* State of one slot after the last run (niter = CG iterations; nrefine =
 * FP64 refreshes for MIXED; rnorm = last FP64 residual norm for MIXED). */
int gooseberry_batch_info(
    gooseberry_batch_t batch,
    uint32_t slot,
    gooseberry_cgd_info_t * info);

/* Device bytes of one solve's workspace (excluding the matrix and the
 * caller's b and x). */
size_t gooseberry_batch_workspace_bytes(
    gooseberry_matrix_t mat,
    int precision);

/* ------------------------------------------------------------ solve server */
/* Continuous batching: requests are queued, admitted first-in first-out
 * while their device-memory footprint fits the budget and a slot is free,
 * advanced together one CG step per device step with every other admitted
 * solve (requests that arrive while others run join at the next step), and
 * retired as they finish. */
typedef struct gooseberry_server_s * gooseberry_server_t;

/* request result status, beside 1 converged, 2 iteration cap, 3 breakdown */
#define GOOSEBERRY_REQUEST_CANCELLED 4  /* gooseberry_server_cancel */
#define GOOSEBERRY_REQUEST_TIMEDOUT 5   /* past the server's timeout */
#define GOOSEBERRY_REQUEST_REJECTED 6   /* its admission refused the input (e.g.
                                           no positive diagonal for
                                           MIXED_JACOBI) */

typedef struct gooseberry_request_result_t {
  uint64_t id;
  int status;            /* 1 converged, 2 iteration cap, 3 breakdown, or
                            one of the GOOSEBERRY_REQUEST_* codes above */
  uint64_t niter;        /* CG (inner) iterations */
  uint64_t nrefine;      /* FP64 refreshes (MIXED) */
  double rel_residual;   /* final residual norm / ||b|| */
  double submit_s;       /* CLOCK_MONOTONIC seconds: submission, */
  double admit_s;        /*   admission, */
  double done_s;         /*   completion seen by the host */
  dim_t n;
  uint64_t nnz;
} gooseberry_request_result_t;

/* mem_budget 0: 90% of the device's free memory at creation.
 * keep_results: keep each generated request's solution on the device
 * until gooseberry_server_take_result. */
int gooseberry_server_create(
    size_t mem_budget,
    uint32_t max_active,
    int precision,
    double delta,
    int keep_results,
    gooseberry_server_t * r_server);

int gooseberry_server_free(gooseberry_server_t server);

/* A request generated on the device at admission: the random symmetric
 * diagonally dominant matrix of gooseberry_matrix_generate_sdd(n, seed,
 * dmin, dmax) (n >= 2) and b uniform in [-1, 1) from (seed, stream 1);
 * relative tolerance tol. */
int gooseberry_server_submit_sdd(
    gooseberry_server_t server,
    uint64_t id,
    dim_t n,
    uint64_t seed,
    dim_t dmin,
    dim_t dmax,
    real_t tol);

/* A request on a resident matrix and device vectors owned by the caller
 * (kept alive until the request is reported done); relative tolerance. */
int gooseberry_server_submit(
    gooseberry_server_t server,
    uint64_t id,
    gooseberry_matrix_t mat,
    const real_t * d_b,
    real_t * d_x,
    real_t tol);

/* Cancel a queued or running request: it leaves the queue or the batch at
 * once (a running request's x holds its last folded iterate) and the next
 * progress call reports it with status GOOSEBERRY_REQUEST_CANCELLED.
 * INVALIDINPUT when no queued or running request has this id. */
int gooseberry_server_cancel(
    gooseberry_server_t server,
    uint64_t id);

/* Requests still queued or running `seconds` after their submission are
 * cancelled by the next progress call and reported with status
 * GOOSEBERRY_REQUEST_TIMEDOUT; seconds <= 0 (the default) disables. */
int gooseberry_server_set_timeout(
    gooseberry_server_t server,
    double seconds);

/* Time out overdue requests, admit what fits, advance every admitted solve
 * by up to `steps` steps, retire the finished ones and report up to
 * max_done of them (*r_ndone counts all).  A request whose admission
 * refuses its input is reported with status GOOSEBERRY_REQUEST_REJECTED;
 * other admission failures (device memory, device errors) return the error
 * and leave the request at the head of the queue. */
int gooseberry_server_progress(
    gooseberry_server_t server,
    uint32_t steps,
    gooseberry_request_result_t * done,
    uint32_t max_done,
    uint32_t * r_ndone);

int gooseberry_server_status(
    gooseberry_server_t server,
    uint32_t * r_active,
    uint32_t * r_queued,
    size_t * r_mem_in_use);

/* Copy a finished generated request's solution (n values) and release it. */
int gooseberry_server_take_result(
    gooseberry_server_t server,
    uint64_t id,
    real_t * h_x,
    size_t n);

/* --------------------------------------------------------------- PageRank */
/* The power iteration of `gooseberry pagerank` (src/gooseberry_bin.c:
 * 2622-2737) on the GPU SpMV.  The matrix's values are divided by their
 * column sums (added in row order), rows without entries are the sinks,
 * rank starts at 1/n, and each iteration computes out = damping * (A rank)
 * + ((1 - damping) / n + sum of rank over the sinks) -- the reference adds
 * the whole sink mass to every entry.  Every tenth iteration (i % 10 == 0)
 * stops when sum (rank - out)^2 <= minerror^2.  niter == 0 runs until that
 * test passes (at most 10^6 iterations here; the reference has no bound).
 * The result is the vector the reference writes: the last iterate when
 * the test passed, else the one before the last. */
typedef struct gooseberry_pagerank_info_t {
  uint64_t niter;   /* the iteration count the reference prints */
  double rmse;      /* the RMSE the reference prints */
  dim_t nsinks;
} gooseberry_pagerank_info_t;

/* The same products on resident matrices (B needs at least as many rows as
 * A has columns; only the first A-ncols rows are used).  The sparse form
 * returns C as a new resident matrix planned with `plan` (NULL: the defaults
 * without the value index, whose build costs about a tenth of a product and
 * pays off only over many SpMVs: gooseberry_matrix_replan(C, NULL) adds it);
 * the dense form writes row-major C (A-nrows x B-ncols) to device memory.
 * The sparse forms stage the short rows' entries in 12 bytes of device
 * memory per product, up to 8 GiB; the environment variable
 * GOOSEBERRY_SPGEMM_STAGE_MAX (bytes, 0 = off) changes that limit, past
 * which the rows are counted and written in two passes (same result). */
int gooseberry_matrix_spmultsp(
    gooseberry_matrix_t a,
    gooseberry_matrix_t b,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * c);

int gooseberry_matrix_spmultsp_dense(
    gooseberry_matrix_t a,
    gooseberry_matrix_t b,
    real_t * d_c);

/* The Cuthill-McKee order of a resident square matrix into device memory
 * d_perm (nrows entries); blocks is a host array. */
int gooseberry_matrix_cuthillmckee(
    gooseberry_matrix_t mat,
    const dim_t * blocks,
    dim_t nblocks,
    dim_t * d_perm);

/* On a resident square matrix; d_out: nrows values on the device.  The
 * matrix itself is not modified. */
int gooseberry_matrix_pagerank(
    gooseberry_matrix_t mat,
    real_t damping,
    size_t niter,
    real_t minerror,
    real_t * d_out,
    gooseberry_pagerank_info_t * info);

/* The same on host CSR arrays (n x n) and a host output vector. */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_pagerank(
    dim_t n,
    const ind_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    real_t damping,
    size_t niter,
    real_t minerror,
    real_t * out,
    gooseberry_pagerank_info_t * info);
#endif

/* ------------------------------------------------------- synthetic inputs */

/* Degree laws of the generated matrices. */
#define GOOSEBERRY_DEGREE_UNIFORM 0   /* d uniform in [dmin, dmax] */
#define GOOSEBERRY_DEGREE_ZIPF 1      /* d = Zipf(s) clipped to [dmin, dmax] */

/* Random symmetric, strictly diagonally dominant SPD matrix generated in
 * HBM (n >= 2).  Row i draws a degree d_i by `law` and d_i - 1 directed
 * off-diagonals at uniform columns != i with values in [-1, 0); the pattern
 * is then symmetrised (each drawn entry also placed at its transpose), and
 * the diagonal, stored last in its row, is 1.01 x the row's absolute
 * off-diagonal sum.  nnz = n + 2 sum(d_i - 1), which must stay below 2^32.
 * Deterministic in (n, seed, law, s, dmin, dmax). */
int gooseberry_matrix_generate_spd(
    dim_t n,
    uint64_t seed,
    int law,
    double zipf_s,
    dim_t dmin,
    dim_t dmax,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

/* gooseberry_matrix_generate_spd with degrees uniform in [dmin, dmax]. */
int gooseberry_matrix_generate_sdd(
    dim_t n,
    uint64_t seed,
    dim_t dmin,
    dim_t dmax,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

/* Copy a resident matrix back to host arrays (nrows+1, nnz, nnz). */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_matrix_download(
    gooseberry_matrix_t mat,
    ind_t * rowptr,
    dim_t * rowind,
    real_t * rowval);
#endif

/* d_out[i] uniform in [lo, hi) from a counter-based hash of (seed, stream,
 * i), filled on the device. */
int gooseberry_generate_uniform_device(
    uint64_t seed,
    uint64_t stream,
    size_t n,
    real_t lo,
    real_t hi,
    real_t * d_out);
/* 2D 5-point / 3D 7-point Dirichlet Laplacians on nx*ny(*nz) grids,
 * lexicographic order, columns ascending within a row (host arrays). */
#ifndef GOOSEBERRY_64BIT_INDICES /* 32-bit row pointers only */
int gooseberry_generate_laplacian(
    int dims,
    dim_t nx, dim_t ny, dim_t nz,
    dim_t * nrows,
    ind_t ** rowptr,
    dim_t ** rowind,
    real_t ** rowval);
#endif

/* n values uniform in [lo, hi) from (seed, stream): xoshiro256** seeded
 * with SplitMix64, u = (x >> 11) * 2^-53. */
int gooseberry_generate_uniform(
    uint64_t seed,
    uint64_t stream,
    size_t n,
    real_t lo,
    real_t hi,
    real_t * out);

/* -------------------------------------------------- 64-bit row pointers */
/* The functions above that take ind_t, with 64-bit row pointers: the
 * reference's GOOSEBERRY_64BIT_INDICES build (include/gooseberry.h:65-71,
 * CMakeLists.txt:59-60).  A program compiled with that macro calls these
 * under the reference's names.  Semantics, errors and results are those of
 * the 32-bit forms.
 *
 * A matrix with at most 2^31 entries is stored as with 32-bit row
 * pointers.  A larger one (up to what HBM holds: about 24 billion entries
 * of FP64 values and 32-bit indices in 288 GB) is held as row parts of at
 * most 2^31 entries, each part with 32-bit row pointers counted from its
 * first entry and an SpMV plan of its own; the SpMV is one launch per part
 * into that part's rows of y, and CG adds p.Ap over the whole vectors.
 * Such a matrix takes the FP64 SpMV, SpMM and CG; the FP32 and mixed paths,
 * batches and servers, PageRank, column splits, the sparse products and
 * Cuthill-McKee return GOOSEBERRY_ERROR_UNIMPLEMENTED for it.  The
 * environment variable GOOSEBERRY_PART_NNZ (read when a matrix is created)
 * sets a smaller part size, down to 1, as a test and tuning knob.  A single
 * row of more entries than a part holds is GOOSEBERRY_ERROR_INVALIDINPUT.
 *
 * The readers and writers and the sparse products keep 32-bit storage
 * inside: their _i64 forms widen or narrow the row pointers and return
 * GOOSEBERRY_ERROR_UNIMPLEMENTED for more than 2^32 - 1 entries. */
int gooseberry_scale_i64(
    uint64_t n,
    real_t * val,
    real_t s);

int gooseberry_add_scalar_i64(
    uint64_t n,
    real_t * val,
    real_t s);

int gooseberry_spmult_i64(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const uint64_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const real_t * bcolval,
    real_t * crowval,
    const dim_t * blocks,
    dim_t nblocks);

int gooseberry_spmultsp_i64(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const uint64_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const uint64_t * browptr,
    const dim_t * browind,
    const real_t * browval,
    real_t * crowval,
    const dim_t * blocks,
    dim_t nblocks);

int gooseberry_spmultsp_sp_i64(
    dim_t anrows,
    dim_t ancols,
    dim_t bncols,
    const uint64_t * arowptr,
    const dim_t * arowind,
    const real_t * arowval,
    const uint64_t * browptr,
    const dim_t * browind,
    const real_t * browval,
    uint64_t ** crowptr,
    dim_t ** crowind,
    real_t ** crowval,
    const dim_t * blocks,
    dim_t nblocks);

int gooseberry_read_sparse_matrix_i64(
    int type,
    const char * file,
    dim_t * r_nrows,
    dim_t * r_ncols,
    uint64_t ** rowptr,
    dim_t ** rowind,
    real_t ** rowval);

int gooseberry_write_sparse_matrix_i64(
    int type,
    const char * file,
    dim_t nrows,
    dim_t ncols,
    const uint64_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval);

int gooseberry_matrix_create_i64(
    dim_t nrows,
    dim_t ncols,
    const uint64_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

int gooseberry_matrix_create_device_i64(
    dim_t nrows,
    dim_t ncols,
    const uint64_t * d_rowptr,
    const dim_t * d_rowind,
    const real_t * d_rowval,
    const gooseberry_plan_t * plan,
    gooseberry_matrix_t * r_mat);

int gooseberry_cgd_i64(
    dim_t nrows,
    dim_t ncols,
    const uint64_t * rowptr,
    const dim_t * rowind,
    const real_t * rowval,
    const real_t * b,
    real_t * x,
    real_t error,
    size_t niter,
    gooseberry_cgd_info_t * info);

/* row parts of a resident matrix (0: one ordinary matrix) */
int gooseberry_matrix_parts(
    gooseberry_matrix_t mat,
    dim_t * r_nparts);

#ifdef __cplusplus
}
#endif

#endif
