// gb_matrix.h -- the device-resident CSR matrix and its SpMV plan.
#ifndef GB_MATRIX_H
#define GB_MATRIX_H

#include "gb_common.h"

// LDS staging buffer of one row block, in values.  window + short_max must
// not exceed it.
#define GB_STREAM_LDS 2304
#define GB_STREAM_ITERS (GB_STREAM_LDS / GB_BLOCK)

struct gbd_matrix {
  uint32_t nrows, ncols;
  uint64_t nnz;
  uint32_t *rowptr;   // nrows + 1
  uint32_t *rowind;   // nnz
  double *val;        // nnz
  float *val32;       // nnz, made on request (mixed precision)
  float *val32j;      // nnz, made on request: FP32 values of the Jacobi-scaled
                      // matrix D^-1/2 A D^-1/2 (D = diag(A))
  double *jscale;     // nrows, with val32j: D^-1/2
  gbd_plan_params params;
  // plan
  uint32_t nstream;     // LDS-staged row blocks
  uint32_t *blk;        // 2 * nstream: [row start, row end) per block
  uint4 *bdesc;         // nstream: a block's first row, end row, first entry
                        // and entry count, read by the SpMV in one load
  uint32_t nwave;       // rows processed one wavefront per row
  uint32_t nblockrows;  // rows processed one workgroup per row
  uint32_t *rowbins;    // stable bin permutation: short | wave | block rows
  int16_t *dcol;        // nnz or null: row-block entries' column - (block's
                        // first row + cshift) (valid only in the row blocks)
  uint32_t cshift;      // 0, or a row part's first row in the whole matrix
                        // (its diagonal's column offset)
  // value-indexed plan (plan wide_values = 0, at most GB_VDICT_MAX distinct
  // values, no column split): each entry's value as an 8-bit index into a
  // table of the distinct values, read instead of val / val32 (vtab32 =
  // the FP32 roundings of vtab)
  uint8_t *vidx;        // nnz or null
  double *vtab;         // GB_VDICT_MAX
  float *vtab32;        // GB_VDICT_MAX, with val32
  uint32_t nvals;
  uint32_t pipe_grid;   // > 0: the SpMV runs the persistent pipelined
                        // row-block kernel on this many workgroups
  // x tiles (params.xtile = C): the entries grouped by (block of xt_R rows,
  // tile of C columns), each group in CSR order; entry j of the layout is
  // xt_idx[j] = (row in block << xt_lc) | (column in tile) with value
  // xt_val[j], group g = b xt_T + t holding [xt_ptr[g], xt_ptr[g + 1])
  uint32_t *xt_ptr;     // xt_W xt_T + 1
  uint32_t *xt_idx;     // nnz
  double *xt_val;       // nnz
  float *xt_val32;      // nnz, with val32: its values in the layout's order
  float *xt_val32j;     // nnz, with val32j: its values in the layout's order
  uint32_t xt_R, xt_W, xt_T, xt_lc;
  uint64_t hist[33];
  // column split (params.csplit = P > 1): P views, view r holding the
  // entries of columns [r W, (r + 1) W), W = ceil(ncols / P), in row order;
  // the views share the arrays below (row pointers of view r start at
  // srowptr + r * nrows) and own only their plans
  uint32_t nsplit;
  gbd_matrix **split;
  uint32_t *srowptr;  // P * nrows + 1
  uint32_t *srowind;  // nnz
  double *sval;       // nnz
  float *sval32;      // nnz, with val32
  float *sval32j;     // nnz, with val32j
  double *ypart;      // P * nrows partial results (FP64, or FP32 in place)
  bool view;          // rowptr/rowind/val belong to a parent matrix
  // row parts (more entries than one part may hold, GB_PART_NNZ by
  // default): part q is rows [prow[q], prow[q + 1]) as a view whose 32-bit
  // row pointers (prowptr + prow[q] + q) count from its first entry
  // pbase[q], with a plan of its own; the parent keeps the 64-bit row
  // pointers and owns the indices and values (rowptr is null)
  bool part;          // this view is a row part (it owns its plan arrays)
  uint32_t nparts;
  gbd_matrix **parts;
  uint32_t *prow;     // host, nparts + 1
  uint64_t *pbase;    // host, nparts + 1
  uint32_t *prowptr;  // nrows + nparts
  uint64_t *rowptr64; // nrows + 1
  uint32_t holders;   // batch slots holding pointers into the plan: while
                      // nonzero the plan must not be rebuilt or freed
};

#define GB_SPLIT_MAX 64
// columns of x an x-tile plan stages in LDS at most
#define GB_XT_CMAX 8192
// entries per row part of a matrix with 64-bit row pointers: every entry
// offset inside a part, and every offset a kernel forms past a part's last
// entry, stays below 2^32
#define GB_PART_NNZ (1ull << 31)
#define GB_VDICT_MAX 256

// One matrix entry's value: val[k], or vtab[vidx[k]] on a value-indexed
// plan (VD).  Plain FP64 values are streamed with non-temporal loads.
template <typename T, bool VD>
__device__ __forceinline__ T gb_ld_val(const T *__restrict__ val,
                                       const uint8_t *__restrict__ vidx,
                                       const T *__restrict__ vtab, uint32_t k) {
  if (VD) return vtab[vidx[k]];
  return __builtin_nontemporal_load(&val[k]);
}

// workgroups of the split sum (fixed, so the partial count is known)
static inline uint32_t gb_split_sum_grid(const gbd_matrix *m) {
  uint32_t g = (m->nrows + GB_BLOCK - 1) / GB_BLOCK;
  return g > 1024u ? 1024u : (g ? g : 1u);
}

// One lane's share of a row: the products val[k] * x[rowind[k]] for
// k = k0 + lane, k0 + lane + STRIDE, ... added in k order from +0.0.  Four
// strides of index and value loads, then their four gathers, are issued
// before the products are added, so a lane keeps four gathers in flight;
// the additions are those of the plain strided loop, in the same order.
template <typename T, uint32_t STRIDE, bool VD = false>
__device__ __forceinline__ T gb_row_lane_sum(const uint32_t *__restrict__ rowind,
                                             const T *__restrict__ val,
                                             const T *__restrict__ x, uint32_t k0,
                                             uint32_t k1, uint32_t lane,
                                             const uint8_t *__restrict__ vidx = nullptr,
                                             const T *__restrict__ vtab = nullptr) {
  T s = 0;
  uint32_t k = k0 + lane;
  for (; (uint64_t)k + 3 * STRIDE < k1; k += 4 * STRIDE) {
    const uint32_t c0 = __builtin_nontemporal_load(&rowind[k]);
    const uint32_t c1 = __builtin_nontemporal_load(&rowind[k + STRIDE]);
    const uint32_t c2 = __builtin_nontemporal_load(&rowind[k + 2 * STRIDE]);
    const uint32_t c3 = __builtin_nontemporal_load(&rowind[k + 3 * STRIDE]);
    const T v0 = gb_ld_val<T, VD>(val, vidx, vtab, k);
    const T v1 = gb_ld_val<T, VD>(val, vidx, vtab, k + STRIDE);
    const T v2 = gb_ld_val<T, VD>(val, vidx, vtab, k + 2 * STRIDE);
    const T v3 = gb_ld_val<T, VD>(val, vidx, vtab, k + 3 * STRIDE);
    const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    s += v0 * x0;
    s += v1 * x1;
    s += v2 * x2;
    s += v3 * x3;
  }
  for (; k < k1; k += STRIDE) {
    const uint32_t c = __builtin_nontemporal_load(&rowind[k]);
    const T v = gb_ld_val<T, VD>(val, vidx, vtab, k);
    s += v * x[c];
  }
  return s;
}

// Launch geometry of one SpMV: stream blocks, then blocks of 4 wave rows,
// then one block per long row.
static inline uint32_t gb_spmv_nwaveblocks(const gbd_matrix *m) {
  return (m->nwave + GB_WAVES_PER_BLOCK - 1) / GB_WAVES_PER_BLOCK;
}
static inline uint32_t gb_spmv_nwork(const gbd_matrix *m) {
  return m->nstream + gb_spmv_nwaveblocks(m) + m->nblockrows;
}
// workgroups of one SpMV launch (the number of p.Ap partials with DOT)
uint32_t gb_spmv_grid(const gbd_matrix *m);

// p.Ap partials written by one SpMV launch with DOT (one per workgroup)
static inline uint32_t gb_spmv_npartials(const gbd_matrix *m) {
  return (m->nsplit || m->nparts || m->xt_ptr) ? gb_split_sum_grid(m) : gb_spmv_grid(m);
}

// Enqueue y = A x; when cg != nullptr also write the per-workgroup partial
// sums of p.Ap (x = p, y = Ap) to partials[gb_spmv_npartials(m)].
int gb_spmv_launch(const gbd_matrix *m, const double *x, double *y,
                   uint32_t ystride, double *partials, const gbd_cg_state *cg);
int gb_spmv_launch_f32(const gbd_matrix *m, const float *x, float *y,
                       double *partials, const gbd_cg_state *cg);

#endif
