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
  gbd_plan_params params;
  // plan
  uint32_t nstream;     // LDS-staged row blocks
  uint32_t *blk;        // 2 * nstream: [row start, row end) per block
  uint32_t nwave;       // rows processed one wavefront per row
  uint32_t nblockrows;  // rows processed one workgroup per row
  uint32_t *rowbins;    // stable bin permutation: short | wave | block rows
  uint64_t hist[33];
};

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

// p.Ap partials written by one SpMV launch with DOT (one per wavefront)
static inline uint32_t gb_spmv_npartials(const gbd_matrix *m) {
  return gb_spmv_grid(m) * GB_WAVES_PER_BLOCK;
}

// Enqueue y = A x; when cg != nullptr also write the per-wavefront partial
// sums of p.Ap (x = p, y = Ap) to partials[gb_spmv_npartials(m)].
int gb_spmv_launch(const gbd_matrix *m, const double *x, double *y,
                   uint32_t ystride, double *partials, const gbd_cg_state *cg);
int gb_spmv_launch_f32(const gbd_matrix *m, const float *x, float *y,
                       double *partials, const gbd_cg_state *cg);

#endif
