/*
 * gb_device.h -- the thin C ABI between the host C code (csrc/host) and
 * the HIP kernels (csrc/hip).  Plain pointers and sizes only.
 *
 * Every gbd_* function returns a gooseberry_error_t value
 * (GOOSEBERRY_SUCCESS == 1).  Pointers named d_* are device (HBM) pointers.
 * All work is enqueued on one library stream per process (created on first
 * use on device 0); functions that hand results to the host synchronise
 * that stream themselves.
 */
#ifndef GB_DEVICE_H
#define GB_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#include "gooseberry.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ runtime */
int gbd_init(void);
int gbd_malloc(void **d_ptr, size_t bytes);
int gbd_free(void *d_ptr);
int gbd_host_alloc(void **h_ptr, size_t bytes); /* pinned host memory */
int gbd_host_free(void *h_ptr);
int gbd_htod(void *d_dst, const void *h_src, size_t bytes);
/* n host arrays to n device buffers in one staged pipeline (n <= 16 share
 * it; more, or arrays in pinned memory, are copied one by one) */
int gbd_htod_many(uint32_t n, void *const *d_dst, const void *const *h_src,
                  const size_t *bytes);
int gbd_dtoh(void *h_dst, const void *d_src, size_t bytes);
int gbd_dtod(void *d_dst, const void *d_src, size_t bytes);
int gbd_zero(void *d_dst, size_t bytes); /* fill kernel, safe under capture */
/* d[i] *= s (src/blas.c:244-257) */
int gbd_scale(double *d, size_t n, double s);
/* d[i] += s (src/blas.c:397-411) */
int gbd_add_scalar(double *d, size_t n, double s);
int gbd_sync(void);
void *gbd_stream(void); /* the library stream as a hipStream_t */
/* Test hook: the (after+1)-th device allocation (kind 0) or kernel launch
 * check (kind 1) from now on reports a failure, once; after < 0 disarms. */
int gbd_inject_fault(int kind, long long after);
int gbd_device_memory(size_t *free_bytes, size_t *total_bytes);
const char *gbd_last_error(void);
/* sets the text gbd_last_error returns (host-side refusals) */
void gbd_set_error(const char *msg);

/* events on the library stream (timing) */
int gbd_event_create(void **ev);
int gbd_event_destroy(void *ev);
int gbd_event_record(void *ev);
int gbd_event_elapsed_ms(void *start, void *stop, float *ms);

/* ------------------------------------------------------ index work */
/* out = exclusive prefix sum of in (n entries); out may alias in.  When
 * d_total is non-NULL it receives the sum of all n entries. */
int gbd_scan_u32(const uint32_t *d_in, uint32_t *d_out, size_t n,
    uint32_t *d_total);
/* Stable counting sort of d_vals by d_keys (keys in [kmin, kmax]).
 * d_out[n] receives vals in key order, stable in input order.  d_dist
 * (optional, kmax-kmin+2 entries) receives the start of each key's run and
 * n at the end.  d_vals == NULL means vals = 0..n-1. */
int gbd_countingsort_kv(const uint32_t *d_keys, const uint32_t *d_vals,
    uint32_t kmin, uint32_t kmax, size_t n, uint32_t *d_out,
    uint32_t *d_dist);
/* The same for keys known to lie in [kmin, kmax] (made by the library
 * itself): no range check, no synchronisation. */
int gbd_countingsort_kv_unchecked(const uint32_t *d_keys,
    const uint32_t *d_vals, uint32_t kmin, uint32_t kmax, size_t n,
    uint32_t *d_out, uint32_t *d_dist);
/* src/permute.c:120-205 semantics on device arrays.  Writes the permuted
 * CSR into d_prowptr/d_prowind/d_prowval (distinct from the inputs). */
int gbd_permute_sparse(uint32_t nrows, uint32_t ncols, const uint32_t *d_rowptr,
    const uint32_t *d_rowind, const double *d_rowval,
    const uint32_t *d_rowperm, const uint32_t *d_colperm,
    uint32_t *d_prowptr, uint32_t *d_prowind, double *d_prowval);
/* src/permute.c:208-285: rev == 0 gathers (new[i][j] =
 * old[rp[i]][cp[j]]), rev == 1 scatters (old[rp[i]][cp[j]] = new[i][j]).
 * d_out is distinct from d_in. */
int gbd_permute_dense(uint32_t nrows, uint32_t ncols, const double *d_in,
    double *d_out, const uint32_t *d_rowperm, const uint32_t *d_colperm,
    int rev);

/* ------------------------------------------------------------- SpMV */
typedef struct gbd_plan_params {
  uint32_t short_max;  /* rows with len <= short_max use LDS-staged blocks */
  uint32_t wave_max;   /* short_max < len <= wave_max: wavefront per row;
                          longer rows: one workgroup per row */
  uint32_t window;     /* nnz window of an LDS-staged row block */
  uint32_t rowcap;     /* max rows per LDS-staged row block */
  uint32_t exact;      /* 1: every row summed left to right (the reference's
                          order) also in the wavefront/workgroup paths */
  uint32_t grid;       /* cap on SpMV workgroups (0: one per work item) */
  uint32_t xcd_remap;  /* 1: renumber row blocks so each XCD walks a
                          contiguous range */
  uint32_t csplit;     /* > 1: the columns are split into this many slices,
                          each multiplied by its own launch into a partial
                          vector, the partials added after (x gathers stay
                          in one slice of x at a time); 0 or 1: off */
  uint32_t wide_index; /* 1: the row blocks read the 32-bit column indices;
                          0: 16-bit offsets from the block's first row when
                          every block's offsets fit in int16 */
  uint32_t wide_values;/* 1: the values are read as stored; 0: as 8-bit
                          indices into a table of the distinct values when
                          there are at most 256 of them (no column split) */
  uint32_t xtile;      /* C > 0 (a power of two, 256..8192): the FP64 SpMV
                          stages x through LDS in tiles of C columns, each
                          workgroup owning a block of rows whose y it keeps
                          in LDS (matrices with scattered columns); 0: off */
} gbd_plan_params;

/* Default parameters for a matrix with the given row-length histogram
 * (hist[b] = number of rows with floor(log2(len+1)) == b, 33 bins). */
void gbd_plan_defaults(const uint64_t *hist, uint64_t nrows, uint64_t nnz,
    gbd_plan_params *p);

typedef struct gbd_matrix gbd_matrix;

/* Structural check of a device CSR: rowptr[0] == 0, non-decreasing, every
 * column index < ncols; *r_nnz = rowptr[nrows]. */
int gbd_csr_check(uint32_t nrows, uint32_t ncols, const uint32_t *d_rowptr,
    const uint32_t *d_rowind, uint32_t *r_nnz);
/* Takes ownership of the device arrays d_rowptr (nrows+1), d_rowind and
 * d_rowval (nnz each; freed by gbd_matrix_destroy) and builds the SpMV plan
 * on the device.  params == NULL selects gbd_plan_defaults(). */
int gbd_matrix_create(uint32_t nrows, uint32_t ncols, uint32_t *d_rowptr,
    uint32_t *d_rowind, double *d_rowval, const gbd_plan_params *params,
    gbd_matrix **out);
/* The same for a CSR the library built itself (valid by construction, nnz
 * known on the host): no read-back and no validation pass. */
int gbd_matrix_create_trusted(uint32_t nrows, uint32_t ncols,
    uint32_t *d_rowptr, uint32_t *d_rowind, double *d_rowval, uint32_t nnz,
    const gbd_plan_params *params, gbd_matrix **out);
/* INVALIDINPUT while a batch slot holds the matrix (gbd_matrix_holders) */
int gbd_matrix_replan(gbd_matrix *m, const gbd_plan_params *params);
/* batch slots currently holding pointers into the matrix's plan */
uint32_t gbd_matrix_holders(const gbd_matrix *m);
void gbd_matrix_destroy(gbd_matrix *m);
int gbd_matrix_info(const gbd_matrix *m, uint32_t *nrows, uint32_t *ncols,
    uint64_t *nnz, gbd_plan_params *params, uint32_t *nblocks_stream,
    uint32_t *nrows_wave, uint32_t *nrows_block, uint64_t *hist33);
/* 1 when the plan's row blocks read 16-bit column offsets */
uint32_t gbd_matrix_index16(const gbd_matrix *m);
/* number of distinct values of a value-indexed plan (0: plain values) */
uint32_t gbd_matrix_values8(const gbd_matrix *m);
/* A CSR with 64-bit row pointers; takes ownership of the three arrays as
 * gbd_matrix_create does.  With at most part_nnz entries (0 or above 2^31:
 * 2^31) the row pointers are narrowed and the matrix is an ordinary one;
 * otherwise it is held as row parts of at most part_nnz entries, each with
 * 32-bit row pointers and a plan of its own (FP64 SpMV, SpMM and CG only;
 * gbd_matrix_rowptr is then NULL). */
int gbd_matrix_create64(uint32_t nrows, uint32_t ncols, uint64_t *d_rowptr64,
    uint32_t *d_rowind, double *d_rowval, uint64_t part_nnz,
    const gbd_plan_params *params, gbd_matrix **out);
/* row parts of a matrix made by gbd_matrix_create64 (0: an ordinary one) */
uint32_t gbd_matrix_nparts(const gbd_matrix *m);
/* NULL for a matrix held as row parts */
const uint32_t *gbd_matrix_rowptr(const gbd_matrix *m);
const uint32_t *gbd_matrix_rowind(const gbd_matrix *m);
const double *gbd_matrix_rowval(const gbd_matrix *m);
/* FP32 copy of the values for the mixed-precision solver (made on first
 * request, kept with the matrix). */
int gbd_matrix_make_f32(gbd_matrix *m);
/* FP32 values of the Jacobi-scaled matrix D^-1/2 A D^-1/2 and D^-1/2
 * (square matrices with a positive, finite diagonal; INVALIDINPUT else). */
int gbd_matrix_make_f32_jacobi(gbd_matrix *m);

/* y = A x (FP64).  bitwise the reference's row-sequential sums for rows in
 * LDS-staged blocks; tree-reduced for longer rows. */
int gbd_spmv(const gbd_matrix *m, const double *d_x, double *d_y);
/* C = A B with B column-major (ncols x bncols) and C row-major
 * (nrows x bncols), src/blas.c:67-92 layout. */
int gbd_spmm(const gbd_matrix *m, uint32_t bncols, const double *d_b,
    double *d_c);
/* y = A x in FP32 values/vectors (FP32 accumulate). */
int gbd_spmv_f32(const gbd_matrix *m, const float *d_x, float *d_y);

/* --------------------------------------------------------------- CG */
typedef struct gbd_cg_state {
  double dr;       /* r.r of the current residual */
  double pap;      /* p.Ap of the current iteration */
  double alpha;
  double beta;
  double dr_prev;  /* dr of the iteration in flight (read by the p update) */
  double rnorm;    /* sqrt(dr) */
  double rnorm0;   /* sqrt(b.b) */
  double tol;      /* absolute stopping threshold on rnorm */
  uint64_t iter;   /* completed iterations (SpMVs) */
  uint64_t iter_next; /* iteration in flight + 1 */
  uint64_t maxiter;/* 0 = unbounded */
  int32_t done;    /* 1: stop (converged, breakdown or maxiter) */
  int32_t status;  /* 0 running, 1 converged, 2 maxiter, 3 breakdown */
  uint32_t ticket[4];
} gbd_cg_state;

typedef struct gbd_cg gbd_cg;

/* Workspace for solving with m (vectors r, p, ap, partials, state). */
int gbd_cg_create(const gbd_matrix *m, gbd_cg **out);
void gbd_cg_destroy(gbd_cg *cg);
/* x = 0, r = b, p = b, dr = b.b; tol is absolute when rel == 0, else
 * relative to ||b||_2.  trace_cap > 0 records rnorm of each iteration in a
 * ring of trace_cap entries (iteration i at slot i % trace_cap). */
int gbd_cg_start(gbd_cg *cg, const double *d_b, double *d_x, double tol,
    int rel, uint64_t maxiter, uint32_t trace_cap);
/* Enqueue k iterations (no host synchronisation). */
int gbd_cg_enqueue(gbd_cg *cg, uint32_t k);
/* Synchronise and copy the state (and, when h_trace != NULL, the trace
 * ring) to the host. */
int gbd_cg_poll(gbd_cg *cg, gbd_cg_state *st, double *h_trace);
/* Capture k iterations into a hipGraph once and replay it for each
 * gbd_cg_enqueue_graph call. */
int gbd_cg_enqueue_graph(gbd_cg *cg, uint32_t k);
/* The workspace of a solve with m in the library's file-scope device array
 * (no device allocation; graphs are not used: gbd_cg_enqueue_graph launches
 * kernel by kernel).  INVALIDINPUT when m needs more than
 * gbd_cg_scratch_max() bytes.  The workspace is shared by every such solve
 * on the device and is not destroyed; start it with trace_cap = 0. */
int gbd_cg_scratch(const gbd_matrix *m, gbd_cg **out);
size_t gbd_cg_scratch_max(void);

/* ----------------------------------------------------- batched solves */
/* Per-system state of a batched solve (device layout, polled to host). */
typedef struct gbd_bstate {
  double dr;        /* r.r of the (inner) recurrence residual */
  double pap;
  double alpha;
  double beta;
  double dr_prev;
  double rnorm;     /* sqrt(dr) */
  double rnorm0;    /* ||b|| */
  double rnorm64;   /* last FP64 residual norm (MIXED refresh steps) */
  double rnorm_ref; /* FP64 residual norm at the last refresh */
  double tol;       /* absolute threshold */
  uint64_t iter;    /* CG iterations (inner iterations for MIXED) */
  uint64_t iter_next;
  uint64_t maxiter;
  uint64_t nrefine; /* FP64 residual refreshes */
  uint64_t inner;   /* inner iterations since the last refresh */
  int32_t done;
  int32_t status;   /* 1 converged, 2 iteration cap, 3 breakdown */
  int32_t mode;     /* step kind of the current step (0 CG, 1 refresh) */
  int32_t mode_next;
  int32_t precision;
  int32_t final_refresh;
} gbd_bstate;

typedef struct gbd_batch gbd_batch;

/* delta: FP32 residual reduction that triggers an FP64 refresh (MIXED). */
int gbd_batch_create(uint32_t max_slots, double delta, gbd_batch **out);
void gbd_batch_destroy(gbd_batch *b);
/* Device bytes the batch allocates for one system solving with m. */
size_t gbd_batch_workspace_bytes(const gbd_matrix *m, int precision);
/* Allocate a slot's workspace and start its solve (x = 0); the slot joins
 * the work lists at the next gbd_batch_commit. */
int gbd_batch_add(gbd_batch *b, const gbd_matrix *m, const double *d_b,
    double *d_x, double tol, int rel, uint64_t maxiter, int precision,
    uint32_t *slot);
int gbd_batch_remove(gbd_batch *b, uint32_t slot);
/* Rebuild the work lists over the used slots with active[slot] != 0
 * (active == NULL: all used slots). */
int gbd_batch_commit(gbd_batch *b, const uint8_t *active);
/* Enqueue k steps of every listed system (graph replays). */
int gbd_batch_step(gbd_batch *b, uint32_t k);
/* Synchronise and copy all max_slots states to the host. */
int gbd_batch_poll(gbd_batch *b, gbd_bstate *states);
int gbd_batch_slot_used(const gbd_batch *b, uint32_t slot);

/* ------------------------------------------------ generated matrices */
/* Random symmetric, strictly diagonally dominant matrix built on the GPU
 * (see gb_gen.hip): degrees drawn by `law` (GOOSEBERRY_DEGREE_*), then
 * symmetrised.  The caller owns the returned device arrays. */
int gbd_generate_spd(uint32_t n, uint64_t seed, int law, double zipf_s,
    uint32_t dmin, uint32_t dmax, uint32_t **d_rowptr, uint32_t **d_rowind,
    double **d_rowval, uint32_t *r_nnz);
/* ------------------------------------------------------------ PageRank */
/* src/gooseberry_bin.c:2622-2737 on a resident square matrix (see
 * gb_pagerank.hip): d_out receives the vector the reference writes. */
int gbd_pagerank(const gbd_matrix *m, double damping, uint64_t niter,
    double minerror, double *d_out, uint64_t *r_iters, double *r_rmse,
    uint32_t *r_nsinks);

/* C = A B for CSR A (anrows x ancols) and CSR B (ancols x bncols), each
 * element summed in the reference's scan order (src/blas.c:94-234).
 * d_cdense non-NULL: C written there row-major (anrows x bncols).
 * Otherwise C returned as new device CSR arrays (free with gbd_free), each
 * row's columns in the order of their first product.  Inputs are trusted
 * (checked by the caller); more than 2^32 - 2 products is INVALIDINPUT. */
int gbd_spgemm(uint32_t anrows, uint32_t ancols, uint32_t bncols,
    const uint32_t *d_arowptr, const uint32_t *d_arowind,
    const double *d_arowval, const uint32_t *d_browptr,
    const uint32_t *d_browind, const double *d_browval, double *d_cdense,
    uint32_t **d_crowptr, uint32_t **d_crowind, double **d_crowval,
    uint32_t *r_cnnz);

/* src/permute.c:49-110,288-332: Cuthill-McKee order of a square CSR's rows
 * (device arrays), per diagonal block when blocks (host array, nblocks + 1
 * monotone entries within [0, nrows]) are given: d_perm[p] = the row at
 * position p.  Positions outside the blocks are left as they are. */
int gbd_cuthillmckee(uint32_t nrows, const uint32_t *d_rowptr,
    const uint32_t *d_rowind, const uint32_t *blocks, uint32_t nblocks,
    uint32_t *d_perm);

/* ------------------------------------------------------------ ingest */
/* The whole file into new device memory (*d_txt, len + 16 bytes; free
 * with gbd_free), through pinned staging buffers. */
int gbd_load_file(const char *file, char **d_txt, uint64_t *r_len);
/* bytes [offset, offset + bytes) of a file into device memory at d_dst */
int gbd_load_file_range(const char *file, uint64_t offset, uint64_t bytes,
    void *d_dst);
/* The body of a MatrixMarket file (the lines after the size line, at
 * d_txt[off, len)) parsed on the device with the semantics of
 * src/io.c:1060-1145; nrows, ncols and sym come from the header. */
int gbd_parse_mtx_body(const char *d_txt, uint64_t len, uint64_t off,
    uint32_t nrows, uint32_t ncols, int sym, uint32_t *r_nnz,
    uint32_t **d_rowptr, uint32_t **d_rowind, double **d_rowval);
/* .csr text at d_txt (len bytes) parsed on the device with the semantics
 * of src/io.c:369-498; new device arrays (free with gbd_free). */
int gbd_parse_csr(const char *d_txt, uint64_t len, uint32_t *r_nrows,
    uint32_t *r_ncols, uint32_t *r_nnz, uint32_t **d_rowptr,
    uint32_t **d_rowind, double **d_rowval);

/* d_out[i] = lo + (hi - lo) u, u from a counter-based hash of (seed,
 * stream, i) */
int gbd_fill_uniform(double *d_out, size_t n, uint64_t seed, uint64_t stream,
    double lo, double hi);

#ifdef __cplusplus
}
#endif

#endif
