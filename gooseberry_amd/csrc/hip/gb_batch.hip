// gb_batch.hip -- many independent CG solves advanced together.
//
// A batch holds up to max_slots systems, each a resident matrix (its own CSR
// and SpMV plan) with its own vectors, partial-sum arrays and state.  One
// device step advances every active system by one step with four launches
// over work lists:
//   K1 gb_b_spmv     (system, SpMV work item): Ap and one p.Ap partial per
//                    item, or in a refresh step the FP64 residual b - A x and
//                    its partials.  A column-split matrix runs over its
//                    slices into partial vectors (FP64 ones in refresh
//                    steps), which K3 adds; its refresh norm is tested in K4.
//                    A matrix with an x-tile plan has no items: its SpMV is a
//                    launch of its own (gb_b_spmv_xt, one partial per row
//                    block) after the list's.
//   K2 gb_b_reduce   (system): alpha, or the refreshed residual norm and the
//                    convergence test
//   K3 gb_b_update   (system, chunk of GB_BCHUNK rows): x/r update and r.r
//                    partials, or the FP32 copy of a refreshed residual
//   K4 gb_b_pupdate  (system, chunk): beta (every chunk adds the system's
//                    partials in the same order), stopping / refresh
//                    decision, p update
// Every sum of a system is taken over that system's own fixed decomposition
// (its SpMV plan, its chunks) in a fixed order, so a system gives the same
// bits alone and in any mix.
//
// Precision per system:
//   FP64  -- the reference's CG (src/cgd.c:40-129) in FP64.
//   MIXED -- FP32 matrix values and inner vectors (d, r, p, Ap), dot
//            products accumulated in FP64, and reliable updates: when the
//            FP32 recurrence residual has dropped by `delta` since the last
//            refresh (or reaches the tolerance), the step folds d into the
//            FP64 solution x and the next step recomputes r = b - A x with
//            the FP64 matrix; the Krylov direction p is kept.  Convergence
//            is only declared on the FP64 residual.
//   MIXED_JACOBI -- MIXED with the inner FP32 solve on the Jacobi-scaled
//            system D^-1/2 A D^-1/2 (D = diag(A)): the inner vectors live in
//            the scaled space (r^ = D^-1/2 r, x = D^-1/2 x^), the FP64
//            refreshes use A itself, and the refresh test compares the
//            scaled recurrence residual with its value at the last refresh.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gb_matrix.h"
#include "gb_xtile.h"

#define GB_BCHUNK 2048           // rows per elementwise work item
#define GB_BCHUNK_ITEMS (GB_BCHUNK / GB_BLOCK)
// The SpMV and elementwise launches get one workgroup per entry of their
// work-list allocation (workgroups past the current count exit at once), so
// items are dispatched in list order, as one launch per matrix would be,
// and a graph stays valid while the lists fit their allocation.  The
// allocation is resized (and the graphs recaptured) when the lists outgrow
// it or shrink below a quarter of it.
#define GB_BGRID_SYS 256u
#define GB_BLIST_MIN 1024u
#define GB_BRED 1024             // threads of the per-system reduction
#define GB_B_NGRAPH 7

enum { GB_MODE_CG = 0, GB_MODE_REFRESH = 1 };

// One SpMV plan as the batched kernels read it: the system's own plan, or
// one column slice of a split matrix (its items numbered from `base`).
struct gb_bplan {
  const uint32_t *rowptr, *rowind, *blk, *wave_rows, *block_rows;
  const int16_t *dcol;  // the row blocks' 16-bit column offsets, or null
  const double *val;
  const float *val32;
  uint32_t nstream, nwave, nwaveblocks, nblockrows, nwork, exact, base, pad;
};

struct gb_bsys {
  gb_bplan pl;      // the matrix's plan (refresh steps, and every step when
                    // the matrix is not split)
  const gb_bplan *slices;  // column slices (nsplit of them) for CG steps
  void *ypart;      // nsplit * n partial products (FP32 for MIXED, else FP64)
  uint32_t nsplit, nwork_split;
  uint32_t n, nchunks;
  const double *b;
  double *x;        // solution (FP64 in both precisions)
  double *r, *p, *ap;  // FP64: residual, direction, A p.  MIXED: r = FP64
                       // residual of refresh steps; p, ap unused
  float *d32, *r32, *p32, *ap32;  // MIXED inner vectors
  double *part1;    // one partial per K1 work item
  double *part2;    // nchunks partials of K3
  double *part3;    // nchunks partials of ||b - A x||^2 (split matrices)
  gbd_bstate *st;
  const double *jscale;  // MIXED_JACOBI: D^-1/2, else null
  // an x-tile plan (xt_ptr non-null): the layout, its FP64 values and the
  // FP32 values of the inner solve (plain or Jacobi-scaled)
  const uint32_t *xt_ptr, *xt_idx;
  const double *xt_val;
  const float *xt_val32;
  uint32_t xt_R, xt_W, xt_T, xt_lc;
  int32_t precision;
  int32_t pad;
};

// The SpMV list is cut into segments, one launch each: segment k holds the
// items of column slice k of every split matrix, and segment 0 also the
// items of the matrices that are not split.  Launch boundaries make the GPU
// finish one slice of x before it starts gathering from the next.
struct gb_blists {
  uint32_t n_spmv, n_elem, n_sys, nseg;
  uint32_t seg[GB_SPLIT_MAX + 1];
};

struct gbd_batch {
  uint32_t max_slots;
  gb_bsys *d_sys;          // device table
  gb_bsys *h_sys;          // host mirror
  gbd_bstate *d_st;        // device states, one per slot
  gbd_bstate *h_st;        // pinned copy
  uint8_t *used;           // host: slot in use
  void **slot_mem;         // host: per-slot workspace allocation
  gbd_matrix **slot_mat;   // host: the matrix each used slot holds
  gb_blists *d_lists;      // device counts
  gb_blists *h_lists;      // pinned
  uint32_t *d_spmv, *d_elem, *d_sysl;  // device work lists
  uint4 *d_desc;  // per SpMV list entry: a row block's r0, r1, first entry, entries
  uint32_t cap_spmv, cap_elem;
  uint32_t nseg;                       // SpMV launches per step
  bool banded;                         // the SpMV build for banded batches
  uint32_t cap_seg[GB_SPLIT_MAX];      // their grids
  uint32_t *h_slice_nwork;             // host: max_slots x GB_SPLIT_MAX
  uint32_t *h_xt;                      // host: the active x-tile systems' slots
  uint32_t *h_xt_grid;                 // and their launch grids
  uint32_t n_xt;
  gb_bplan *d_slices;                  // max_slots x GB_SPLIT_MAX slice plans
  uint32_t *h_spmv, *h_elem, *h_sysl;  // pinned staging
  double delta;
  hipGraph_t graph[GB_B_NGRAPH];
  hipGraphExec_t gexec[GB_B_NGRAPH];
};

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double gb_b_sum_chunks(const double *part, uint32_t n,
                                                  double *red) {
  // same code and order in every workgroup of K4 and in the init reduction
  double s = 0.0;
  for (uint32_t base = 0; base < n; base += GB_BLOCK * 8) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      uint32_t i = base + j * GB_BLOCK + threadIdx.x;
      v[j] = i < n ? part[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  s = gb_wave_sum(s);
  const int lane = threadIdx.x & (GB_WAVE - 1), wave = threadIdx.x / GB_WAVE;
  if (lane == 0) red[wave] = s;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < GB_WAVES_PER_BLOCK; ++w) t += red[w];
  __syncthreads();
  return t;
}

// One SpMV work item of system s (same bins and orders as gb_spmv_kernel).
// EPI 0: y = A x, dot += x[row] * y   (CG step; T = value/vector type)
// EPI 1: out = b - A x in FP64, dot += out^2   (refresh step; T = double)
// A row block's rows and entries come from its descriptor (d), filled when
// the work list is built, so the block's loads start without waiting on the
// plan's block table and row pointers.
template <typename T, int EPI>
__device__ __forceinline__ double gb_b_spmv_item(const gb_bplan &S, uint32_t w,
                                                 const T *__restrict__ val,
                                                 const T *__restrict__ x,
                                                 T *__restrict__ y,
                                                 const double *__restrict__ b,
                                                 T *prod, T *tred, const uint4 d) {
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (GB_WAVE - 1), wave = tid / GB_WAVE;
  const uint32_t *__restrict__ rowptr = S.rowptr;
  const uint32_t *__restrict__ rowind = S.rowind;
  double dotp = 0.0;
  if (w < S.nstream) {
    const uint32_t r0 = d.x, r1 = d.y;
    const uint32_t k0 = d.z;
    const uint32_t cnt = d.w;
    const uint32_t ra = r0 + tid, rb = r0 + tid + GB_BLOCK;
    uint32_t pa0 = 0, pa1 = 0, pb0 = 0, pb1 = 0;
    T xa = 0, xb = 0;
    double ba = 0, bb = 0;
    if (ra < r1) {
      pa0 = rowptr[ra]; pa1 = rowptr[ra + 1];
      if (EPI == 0) xa = x[ra]; else ba = b[ra];
    }
    if (rb < r1) {
      pb0 = rowptr[rb]; pb1 = rowptr[rb + 1];
      if (EPI == 0) xb = x[rb]; else bb = b[rb];
    }
    if (cnt > 0) {
      const uint32_t klast = k0 + cnt - 1;
      // the matrix's plan decides between 16-bit offsets from r0 (plain
      // loads; FP32 values plain too) and 32-bit indices, as in
      // gb_spmv_kernel; the offsets are added at the gathers
      const int16_t *__restrict__ dcol = S.dcol;
      uint32_t c[GB_STREAM_ITERS];
      T v[GB_STREAM_ITERS], xv[GB_STREAM_ITERS];
      if (dcol) {
#pragma unroll
        for (int i = 0; i < GB_STREAM_ITERS; ++i) {
          if (i * GB_BLOCK < cnt) {
            uint32_t k = min(k0 + i * GB_BLOCK + tid, klast);
            c[i] = (uint32_t)(int32_t)dcol[k];
            if (sizeof(T) == 4) v[i] = val[k];
            else v[i] = __builtin_nontemporal_load(&val[k]);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < GB_STREAM_ITERS; ++i) {
          if (i * GB_BLOCK < cnt) {
            uint32_t k = min(k0 + i * GB_BLOCK + tid, klast);
            c[i] = __builtin_nontemporal_load(&rowind[k]);
            v[i] = __builtin_nontemporal_load(&val[k]);
          }
        }
      }
      const T *xg = dcol ? x + r0 : x;
#pragma unroll
      for (int i = 0; i < GB_STREAM_ITERS; ++i)
        if (i * GB_BLOCK < cnt)
          xv[i] = xg[dcol ? (int64_t)(int32_t)c[i] : (int64_t)c[i]];
#pragma unroll
      for (int i = 0; i < GB_STREAM_ITERS; ++i) {
        uint32_t e = i * GB_BLOCK + tid;
        if (i * GB_BLOCK < cnt && e < cnt) prod[e] = v[i] * xv[i];
      }
    }
    __syncthreads();
    if (ra < r1) {
      T s = 0;
      for (uint32_t k = pa0 - k0; k < pa1 - k0; ++k) s += prod[k];
      if (EPI == 0) { y[ra] = s; dotp += (double)xa * (double)s; }
      else { double o = ba - (double)s; ((double *)y)[ra] = o; dotp += o * o; }
    }
    if (rb < r1) {
      T s = 0;
      for (uint32_t k = pb0 - k0; k < pb1 - k0; ++k) s += prod[k];
      if (EPI == 0) { y[rb] = s; dotp += (double)xb * (double)s; }
      else { double o = bb - (double)s; ((double *)y)[rb] = o; dotp += o * o; }
    }
    __syncthreads();
  } else if (w < S.nstream + S.nwaveblocks) {
    const uint32_t wid = (w - S.nstream) * GB_WAVES_PER_BLOCK + wave;
    if (wid < S.nwave) {
      const uint32_t row = S.wave_rows[wid];
      const uint32_t k0 = rowptr[row], k1 = rowptr[row + 1];
      T s = 0;
      if (S.exact) {
        for (uint32_t base = k0; base < k1; base += GB_WAVE) {
          const uint32_t k = base + lane;
          T pr = 0;
          if (k < k1) pr = val[k] * x[rowind[k]];
          const uint32_t cn = min((uint32_t)GB_WAVE, k1 - base);
          for (uint32_t j = 0; j < cn; ++j) s += gb_readlane(pr, j);
        }
      } else {
        s = gb_wave_sum(gb_row_lane_sum<T, GB_WAVE>(rowind, val, x, k0, k1, lane));
      }
      if (lane == 0) {
        if (EPI == 0) { y[row] = s; dotp += (double)x[row] * (double)s; }
        else { double o = b[row] - (double)s; ((double *)y)[row] = o; dotp += o * o; }
      }
    }
  } else {
    const uint32_t row = S.block_rows[w - S.nstream - S.nwaveblocks];
    const uint32_t k0 = rowptr[row], k1 = rowptr[row + 1];
    T s = 0;
    if (S.exact) {
      for (uint32_t base = k0; base < k1; base += GB_STREAM_LDS) {
        const uint32_t cn = min((uint32_t)GB_STREAM_LDS, k1 - base);
        for (uint32_t e = tid; e < cn; e += GB_BLOCK)
          prod[e] = val[base + e] * x[rowind[base + e]];
        __syncthreads();
        if (tid == 0)
          for (uint32_t e = 0; e < cn; ++e) s += prod[e];
        __syncthreads();
      }
    } else {
      s = gb_wave_sum(gb_row_lane_sum<T, GB_BLOCK>(rowind, val, x, k0, k1, tid));
      if (lane == 0) tred[wave] = s;
      __syncthreads();
      if (tid == 0) {
        T t = tred[0];
#pragma unroll
        for (int q = 1; q < GB_WAVES_PER_BLOCK; ++q) t += tred[q];
        s = t;
      }
      __syncthreads();
    }
    if (tid == 0) {
      if (EPI == 0) { y[row] = s; dotp += (double)x[row] * (double)s; }
      else { double o = b[row] - (double)s; ((double *)y)[row] = o; dotp += o * o; }
    }
  }
  return dotp;
}

// ----------------------------------------------------------------- kernels

// The descriptors of the SpMV list entries [a, b) of segment seg: for a
// row block (item below its plan's nstream), its first and last-plus-one
// rows, its first entry and its entry count.
__global__ void __launch_bounds__(GB_BLOCK)
gb_b_describe(const gb_bsys *__restrict__ sys, const uint32_t *__restrict__ items,
              uint32_t a, uint32_t b, uint32_t seg, const gb_bplan *__restrict__ slice_tab,
              uint4 *__restrict__ desc) {
  for (uint32_t it = a + blockIdx.x * GB_BLOCK + threadIdx.x; it < b;
       it += gridDim.x * GB_BLOCK) {
    const uint32_t s = items[2 * it], lw = items[2 * it + 1];
    const gb_bsys &S = sys[s];
    const gb_bplan &L = S.nsplit ? slice_tab[(size_t)s * GB_SPLIT_MAX + seg] : S.pl;
    uint4 d = make_uint4(0u, 0u, 0u, 0u);
    if (lw < L.nstream) {
      d.x = L.blk[2 * lw];
      d.y = L.blk[2 * lw + 1];
      d.z = L.rowptr[d.x];
      d.w = L.rowptr[d.y] - d.z;
    }
    desc[it] = d;
  }
}

// Wavefronts per SIMD the batched SpMV's register budget must allow, by
// batch.  Measured: 512 Laplacians (mixed precision) 1645 solves/s at 5
// (87 VGPRs, the compiler's own choice), 1755 at 7; the served stream of
// random matrices 355-363 solves/s at 5, 344-346 at 7.  A step launches the
// seven-wavefront build when every admitted system's matrix reads 16-bit
// column offsets (banded rows, gathers that stay in L2), else the five.
#define GB_B_SPMV_WAVES_BANDED 7
#define GB_B_SPMV_WAVES_OTHER 5
template <int WAVES>
__global__ void __launch_bounds__(GB_BLOCK)
__attribute__((amdgpu_waves_per_eu(WAVES)))
gb_b_spmv(const gb_bsys *__restrict__ sys, const uint32_t *__restrict__ items,
          const uint4 *__restrict__ desc,
          const gb_blists *__restrict__ lists, uint32_t seg,
          const gbd_bstate *__restrict__ states, const gb_bplan *__restrict__ slice_tab) {
  __shared__ double prodd[GB_STREAM_LDS];
  __shared__ double tredd[GB_WAVES_PER_BLOCK];
  __shared__ double red[GB_WAVES_PER_BLOCK];
  const uint32_t end = lists->seg[seg + 1];
  for (uint32_t it = lists->seg[seg] + blockIdx.x; it < end; it += gridDim.x) {
    const uint32_t s = items[2 * it], lw = items[2 * it + 1];
    const uint4 d = desc[it];
    // the system, its state and its slice plan are addressed from the slot
    // alone, so their loads go out together
    const gb_bsys &S = sys[s];
    const gbd_bstate *st = states + s;
    if (st->done) continue;  // not written during this kernel: uniform
    const int mode = st->mode_next;
    double dotp;
    uint32_t w = lw;  // the item's partial slot
    if (S.nsplit) {
      // a column-split matrix runs every step over its slices, slice r in
      // launch r, each into its partial vector: CG steps in the system's
      // precision, refresh steps A x in FP64 (the update kernel adds them)
      const uint32_t r = seg;
      const gb_bplan &L = slice_tab[(size_t)s * GB_SPLIT_MAX + r];
      w = L.base + lw;
      if (mode == GB_MODE_REFRESH) {
        double *yp = (double *)S.ypart + (size_t)r * S.n;
        dotp = gb_b_spmv_item<double, 0>(L, lw, L.val, S.x, yp, nullptr, prodd, tredd, d);
      } else if (S.precision != GOOSEBERRY_PRECISION_FP64) {
        float *yp = (float *)S.ypart + (size_t)r * S.n;
        dotp = gb_b_spmv_item<float, 0>(L, lw, L.val32, S.p32, yp, nullptr,
                                        (float *)prodd, (float *)tredd, d);
      } else {
        double *yp = (double *)S.ypart + (size_t)r * S.n;
        dotp = gb_b_spmv_item<double, 0>(L, lw, L.val, S.p, yp, nullptr, prodd, tredd, d);
      }
    } else if (mode == GB_MODE_REFRESH) {
      dotp = gb_b_spmv_item<double, 1>(S.pl, w, S.pl.val, S.x, S.r, S.b, prodd, tredd, d);
    } else if (S.precision != GOOSEBERRY_PRECISION_FP64) {
      dotp = gb_b_spmv_item<float, 0>(S.pl, w, S.pl.val32, S.p32, S.ap32, nullptr,
                                      (float *)prodd, (float *)tredd, d);
    } else {
      dotp = gb_b_spmv_item<double, 0>(S.pl, w, S.pl.val, S.p, S.ap, nullptr, prodd, tredd, d);
    }
    // one partial per work item
    const double bs = gb_block_sum(dotp, red);
    if (threadIdx.x == 0) S.part1[w] = bs;
    __syncthreads();
  }
}

// The SpMV of a system on an x-tile plan (grid = its row blocks): Ap and
// the p.Ap partials of a CG step in the system's precision, or in a
// refresh step the FP64 residual b - A x and its squared-norm partials;
// one partial per row block
__global__ void __launch_bounds__(GB_XT_THREADS)
gb_b_spmv_xt(const gb_bsys *__restrict__ sys, uint32_t slot,
             const gbd_bstate *__restrict__ states) {
  __shared__ gb_xt_lds L;
  const gb_bsys &S = sys[slot];
  const gbd_bstate *st = states + slot;
  if (st->done) return;  // not written during this kernel: uniform
  if (st->mode_next == GB_MODE_REFRESH)
    gb_xt_spmv<double, GB_XT_RESID>(S.xt_ptr, S.xt_idx, S.xt_val, S.x, S.r, 1, S.b, S.part1,
                                    S.n, S.n, S.xt_R, S.xt_T, S.xt_lc, L);
  else if (S.precision != GOOSEBERRY_PRECISION_FP64)
    gb_xt_spmv<float, GB_XT_DOT>(S.xt_ptr, S.xt_idx, S.xt_val32, S.p32, S.ap32, 1, nullptr,
                                 S.part1, S.n, S.n, S.xt_R, S.xt_T, S.xt_lc, L);
  else
    gb_xt_spmv<double, GB_XT_DOT>(S.xt_ptr, S.xt_idx, S.xt_val, S.p, S.ap, 1, nullptr,
                                  S.part1, S.n, S.n, S.xt_R, S.xt_T, S.xt_lc, L);
}

__global__ void __launch_bounds__(GB_BRED)
gb_b_reduce(const gb_bsys *__restrict__ sys, const uint32_t *__restrict__ syslist,
            const gb_blists *__restrict__ lists) {
  __shared__ double red[GB_BRED / GB_WAVE];
  const uint32_t count = lists->n_sys;
  for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
    const gb_bsys &S = sys[syslist[it]];
    gbd_bstate *st = S.st;
    if (st->done) continue;
    const int mode = st->mode_next;
    if (S.nsplit && mode == GB_MODE_REFRESH) {
      // A x is still in slices: the update kernel forms b - A x and the
      // p-update kernel tests it
      if (threadIdx.x == 0) st->mode = mode;
      continue;
    }
    // one partial per SpMV work item (per row block on an x-tile plan)
    const uint32_t np = S.xt_ptr ? S.xt_W : S.nsplit ? S.nwork_split : S.pl.nwork;
    const double *part = S.part1;
    double s = 0.0;
    for (uint32_t base = 0; base < np; base += GB_BRED * 8) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        uint32_t i = base + j * GB_BRED + threadIdx.x;
        v[j] = i < np ? part[i] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) s += v[j];
    }
    s = gb_wave_sum(s);
    if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = red[0];
      for (int q = 1; q < GB_BRED / GB_WAVE; ++q) t += red[q];
      st->mode = mode;
      if (mode == GB_MODE_REFRESH) {
        // t = ||b - A x||^2 in FP64
        const double rn = sqrt(t);
        st->nrefine += 1;
        st->rnorm64 = rn;
        st->rnorm_ref = rn;
        if (rn <= st->tol) {
          st->done = 1;
          st->status = 1;
        } else if (!isfinite(rn)) {
          st->done = 1;
          st->status = 3;
        } else if (st->final_refresh) {
          st->done = 1;
          st->status = 2;
        }
      } else {
        const double dr = st->dr;
        const double alpha = dr / t;
        st->pap = t;
        st->alpha = alpha;
        if (!(t != 0.0) || !isfinite(alpha)) {
          st->done = 1;
          st->status = 3;
        } else {
          st->dr_prev = dr;
          st->iter_next = st->iter + 1;
        }
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_b_update(const gb_bsys *__restrict__ sys, const uint32_t *__restrict__ items,
            const gb_blists *__restrict__ lists) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  const uint32_t count = lists->n_elem;
  for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
    const uint32_t sidx = items[2 * it], c = items[2 * it + 1];
    const gb_bsys &S = sys[sidx];
    const gbd_bstate *st = S.st;
    if (st->done) continue;
    const int mode = st->mode;
    const uint32_t i0 = c * GB_BCHUNK;
    double d = 0.0;
    if (S.precision != GOOSEBERRY_PRECISION_FP64) {
      if (mode == GB_MODE_REFRESH) {
        // r32 = (float) r64 from the refresh SpMV (for a split matrix r64 =
        // b - A x from the slices' FP64 partials, and its squared norm goes
        // to part3); d = 0 was set when x was updated
        const double *yp = (const double *)S.ypart;
        double d64 = 0.0;
        for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
          uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
          if (i < S.n) {
            double rr;
            if (S.nsplit) {
              double ax = yp[i];
              for (uint32_t r = 1; r < S.nsplit; ++r) ax += yp[(size_t)r * S.n + i];
              rr = S.b[i] - ax;
              S.r[i] = rr;
              d64 += rr * rr;
            } else {
              rr = S.r[i];
            }
            float rf = S.jscale ? (float)(S.jscale[i] * rr) : (float)rr;
            S.r32[i] = rf;
            d += (double)rf * (double)rf;
          }
        }
        if (S.nsplit) {
          const double b64 = gb_block_sum(d64, red);
          if (threadIdx.x == 0) S.part3[c] = b64;
          __syncthreads();
        }
      } else {
        const float a = (float)st->alpha;
        const float *yp = (const float *)S.ypart;
        for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
          uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
          if (i < S.n) {
            float ap;
            if (S.nsplit) {  // slice partials added in slice order
              ap = yp[i];
              for (uint32_t r = 1; r < S.nsplit; ++r) ap += yp[(size_t)r * S.n + i];
            } else {
              ap = S.ap32[i];
            }
            // d += a p is done by gb_b_pupdate, which reads p anyway
            float rn = S.r32[i] - a * ap;
            S.r32[i] = rn;
            d += (double)rn * (double)rn;
          }
        }
      }
    } else {
      // src/cgd.c:92-100
      const double alpha = st->alpha;
      const double *yp = (const double *)S.ypart;
      for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
        uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
        if (i < S.n) {
          double ap;
          if (S.nsplit) {
            ap = yp[i];
            for (uint32_t r = 1; r < S.nsplit; ++r) ap += yp[(size_t)r * S.n + i];
          } else {
            ap = S.ap[i];
          }
          double ps = alpha * S.p[i];
          S.x[i] = S.x[i] + ps;
          double aps = alpha * ap;
          double rn = S.r[i] - aps;
          S.r[i] = rn;
          d += rn * rn;
        }
      }
    }
    double bs = gb_block_sum(d, red);
    if (threadIdx.x == 0) S.part2[c] = bs;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_b_pupdate(const gb_bsys *__restrict__ sys, const uint32_t *__restrict__ items,
             const gb_blists *__restrict__ lists, double delta) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  __shared__ int32_t s_done;
  const uint32_t count = lists->n_elem;
  for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
    const uint32_t sidx = items[2 * it], c = items[2 * it + 1];
    const gb_bsys &S = sys[sidx];
    gbd_bstate *st = S.st;
    // workgroup 0 of a system may set done during this kernel: decide once
    if (threadIdx.x == 0) s_done = *(volatile int32_t *)&st->done;
    __syncthreads();
    const int32_t was_done = s_done;
    __syncthreads();
    if (was_done) continue;
    const int mode = st->mode;
    const double dnr = gb_b_sum_chunks(S.part2, S.nchunks, red);
    const uint32_t i0 = c * GB_BCHUNK;
    if (mode == GB_MODE_REFRESH) {
      // MIXED only: restart the FP32 recurrence from the FP64 residual,
      // keep the direction p.  A split matrix's FP64 residual norm is
      // tested here (the reduction kernel tests the others').
      if (S.nsplit) {
        const double t = gb_b_sum_chunks(S.part3, S.nchunks, red);
        if (c == 0 && threadIdx.x == 0) {
          const double rn = sqrt(t);
          st->nrefine += 1;
          st->rnorm64 = rn;
          st->rnorm_ref = rn;
          if (rn <= st->tol) {
            st->done = 1;
            st->status = 1;
          } else if (!isfinite(rn)) {
            st->done = 1;
            st->status = 3;
          } else if (st->final_refresh) {
            st->done = 1;
            st->status = 2;
          }
        }
      }
      if (c == 0 && threadIdx.x == 0) {
        st->dr = dnr;
        st->rnorm = sqrt(dnr);
        st->inner = 0;
        st->mode_next = GB_MODE_CG;
        // the refresh test of a scaled system compares scaled residuals
        if (S.jscale) st->rnorm_ref = st->rnorm;
      }
      continue;
    }
    // src/cgd.c:100-119 (and its FP32-inner counterpart)
    const double alpha = st->alpha;
    const double beta = dnr / st->dr_prev;
    const double rnorm = sqrt(dnr);
    const unsigned long long it1 = st->iter_next;
    const bool cap = st->maxiter && it1 >= st->maxiter;
    int status = 0;        // FP64: stop here
    bool refresh = false;  // MIXED: fold d into x, FP64 residual next step
    if (S.precision != GOOSEBERRY_PRECISION_FP64) {
      if (!isfinite(dnr)) status = 3;
      // refresh after a drop by delta, or once the recurrence residual
      // predicts convergence (for a scaled system: it has dropped by
      // tol / ||b - A x|| since the last refresh)
      else if (rnorm <= delta * st->rnorm_ref || cap ||
               (S.jscale ? rnorm * st->rnorm64 <= st->tol * st->rnorm_ref
                         : rnorm <= st->tol))
        refresh = true;
    } else {
      if (rnorm <= st->tol) status = 1;
      else if (!isfinite(dnr)) status = 3;
      else if (cap) status = 2;
    }
    if (c == 0 && threadIdx.x == 0) {
      st->beta = beta;
      st->dr = dnr;
      st->rnorm = rnorm;
      st->iter = it1;
      st->inner += 1;
      if (status) {
        st->status = status;
        st->done = 1;
      }
      if (refresh) {
        st->mode_next = GB_MODE_REFRESH;
        if (cap) st->final_refresh = 1;
      }
    }
    if (status) continue;
    if (S.precision != GOOSEBERRY_PRECISION_FP64) {
      // this step's d += a p (the same float products and sums as beside
      // the r update, where it was one more read of p), then p = r + beta p.
      // A mixed system's CG step sets `done` above only on a non-finite
      // residual, after which x never takes d: a workgroup that sees that
      // flag early and skips this loop changes no result.
      const float a = (float)alpha;
      const float boa = (float)beta;
      for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
        uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
        if (i < S.n) {
          const float pv = S.p32[i];
          const float dn = S.d32[i] + a * pv;
          S.p32[i] = S.r32[i] + boa * pv;
          if (refresh) {
            // the FP64 solution update
            S.x[i] = S.x[i] + (S.jscale ? S.jscale[i] * (double)dn : (double)dn);
            S.d32[i] = 0.0f;
          } else {
            S.d32[i] = dn;
          }
        }
      }
    } else {
      const double boa = beta / alpha;
      for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
        uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
        if (i < S.n) {
          double ps = alpha * S.p[i];
          double t = ps * boa;
          S.p[i] = S.r[i] + t;
        }
      }
    }
  }
}

// Start of one system (not part of the step graph): x = 0, r = b, p = r
// (FP64) or r32 = p32 = (float) b, d = 0 (MIXED); chunk partials of b.b
// (part2) and of r32.r32 (part1, MIXED).
__global__ void __launch_bounds__(GB_BLOCK)
gb_b_init(const gb_bsys *__restrict__ sys, uint32_t slot) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  const gb_bsys &S = sys[slot];
  for (uint32_t c = blockIdx.x; c < S.nchunks; c += gridDim.x) {
    const uint32_t i0 = c * GB_BCHUNK;
    double d = 0.0, f = 0.0;
    for (int j = 0; j < GB_BCHUNK_ITEMS; ++j) {
      uint32_t i = i0 + j * GB_BLOCK + threadIdx.x;
      if (i < S.n) {
        double v = S.b[i];
        S.x[i] = 0.0;
        d += v * v;
        if (S.precision != GOOSEBERRY_PRECISION_FP64) {
          float vf = S.jscale ? (float)(S.jscale[i] * v) : (float)v;
          S.r32[i] = vf;
          S.p32[i] = vf;
          S.d32[i] = 0.0f;
          f += (double)vf * (double)vf;
        } else {
          S.r[i] = v;
          S.p[i] = v;
        }
      }
    }
    double bs = gb_block_sum(d, red);
    __syncthreads();
    double fs = gb_block_sum(f, red);
    if (threadIdx.x == 0) {
      S.part2[c] = bs;
      S.part1[c] = fs;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_b_init_state(const gb_bsys *__restrict__ sys, uint32_t slot, double tol,
                int rel, unsigned long long maxiter) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  const gb_bsys &S = sys[slot];
  const double bb = gb_b_sum_chunks(S.part2, S.nchunks, red);
  const double ff = gb_b_sum_chunks(S.part1, S.nchunks, red);
  if (threadIdx.x == 0) {
    gbd_bstate *st = S.st;
    uint32_t *words = (uint32_t *)st;
    for (uint32_t i = 0; i < sizeof(gbd_bstate) / 4; ++i) words[i] = 0u;
    st->rnorm0 = sqrt(bb);
    st->tol = rel ? tol * st->rnorm0 : tol;
    st->maxiter = maxiter;
    st->precision = S.precision;
    st->mode = GB_MODE_CG;
    st->mode_next = GB_MODE_CG;
    if (S.precision != GOOSEBERRY_PRECISION_FP64) {
      st->dr = ff;
      st->rnorm = sqrt(ff);
      st->rnorm64 = st->rnorm0;
      st->rnorm_ref = S.jscale ? st->rnorm : st->rnorm0;
    } else {
      st->dr = bb;
      st->rnorm = st->rnorm0;
    }
    if (bb == 0.0) {
      st->done = 1;
      st->status = 1;
    } else if (!isfinite(bb)) {
      st->done = 1;
      st->status = 3;
    }
  }
}

// ------------------------------------------------------------------ host

static void gb_b_drop_graphs(gbd_batch *B) {
  for (int i = 0; i < GB_B_NGRAPH; ++i) {
    if (B->gexec[i]) (void)hipGraphExecDestroy(B->gexec[i]);
    if (B->graph[i]) (void)hipGraphDestroy(B->graph[i]);
    B->gexec[i] = nullptr;
    B->graph[i] = nullptr;
  }
}

static int gb_b_enqueue_raw(gbd_batch *B, uint32_t k) {
  hipStream_t s = gb_stream();
  for (uint32_t i = 0; i < k; ++i) {
    for (uint32_t g = 0; g < B->nseg; ++g) {
      if (B->banded)
        gb_b_spmv<GB_B_SPMV_WAVES_BANDED><<<B->cap_seg[g], GB_BLOCK, 0, s>>>(
            B->d_sys, B->d_spmv, B->d_desc, B->d_lists, g, B->d_st, B->d_slices);
      else
        gb_b_spmv<GB_B_SPMV_WAVES_OTHER><<<B->cap_seg[g], GB_BLOCK, 0, s>>>(
            B->d_sys, B->d_spmv, B->d_desc, B->d_lists, g, B->d_st, B->d_slices);
      GB_LAUNCH_CHECK();
    }
    for (uint32_t j = 0; j < B->n_xt; ++j) {
      gb_b_spmv_xt<<<B->h_xt_grid[j], GB_XT_THREADS, 0, s>>>(B->d_sys, B->h_xt[j], B->d_st);
      GB_LAUNCH_CHECK();
    }
    gb_b_reduce<<<GB_BGRID_SYS, GB_BRED, 0, s>>>(B->d_sys, B->d_sysl, B->d_lists);
    GB_LAUNCH_CHECK();
    gb_b_update<<<B->cap_elem, GB_BLOCK, 0, s>>>(B->d_sys, B->d_elem, B->d_lists);
    GB_LAUNCH_CHECK();
    gb_b_pupdate<<<B->cap_elem, GB_BLOCK, 0, s>>>(B->d_sys, B->d_elem, B->d_lists,
                                                     B->delta);
    GB_LAUNCH_CHECK();
  }
  return GOOSEBERRY_SUCCESS;
}

static uint32_t gb_b_nchunks(uint32_t n) { return (n + GB_BCHUNK - 1) / GB_BCHUNK; }

// Workspace of one system: byte offsets of its buffers inside one
// allocation (each 256-byte aligned); the same function sizes and carves it.
struct gb_b_layout {
  size_t part1, part2, part3, r, p, ap, d32, r32, p32, ap32, ypart, slices, total;
};

// SpMV items of the matrix's column slices (0 when it is not split)
static uint32_t gb_b_nwork_split(const gbd_matrix *m) {
  uint32_t t = 0;
  for (uint32_t r = 0; r < m->nsplit; ++r) t += gb_spmv_nwork(m->split[r]);
  return t;
}

static gb_b_layout gb_b_plan_layout(const gbd_matrix *m, int precision) {
  const size_t n = m->nrows ? m->nrows : 1;
  size_t nwork = m->nsplit ? gb_b_nwork_split(m) : gb_spmv_nwork(m);
  if (m->xt_ptr && m->xt_W > nwork) nwork = m->xt_W;
  const size_t nch = gb_b_nchunks(m->nrows);
  const size_t np1 = (nwork > nch ? nwork : nch) + 1;
  const size_t np2 = nch + 1;
  gb_b_layout L;
  memset(&L, 0xff, sizeof(L));
  size_t off = 0;
  auto take = [&](size_t nb) {
    size_t o = off;
    off += (nb + 255) & ~(size_t)255;
    return o;
  };
  L.part1 = take(np1 * 8);
  L.part2 = take(np2 * 8);
  L.r = take(n * 8);
  if (precision != GOOSEBERRY_PRECISION_FP64) {
    L.d32 = take(n * 4);
    L.r32 = take(n * 4);
    L.p32 = take(n * 4);
    L.ap32 = take(n * 4);
  } else {
    L.p = take(n * 8);
    L.ap = take(n * 8);
  }
  if (m->nsplit) {
    // FP64-sized: the refresh steps of a mixed system use it in FP64
    L.ypart = take((size_t)m->nsplit * n * 8);
    L.slices = take((size_t)m->nsplit * sizeof(gb_bplan));
    L.part3 = take((nch + 1) * 8);
  }
  L.total = off;
  return L;
}

static gb_bplan gb_b_plan_of(const gbd_matrix *m, uint32_t base, bool jacobi) {
  gb_bplan P;
  memset(&P, 0, sizeof(P));
  P.rowptr = m->rowptr;
  P.rowind = m->rowind;
  P.dcol = m->dcol;
  P.blk = m->blk;
  P.wave_rows = m->rowbins + (m->nrows - m->nwave - m->nblockrows);
  P.block_rows = m->rowbins + (m->nrows - m->nblockrows);
  P.val = m->val;
  P.val32 = jacobi ? m->val32j : m->val32;
  P.nstream = m->nstream;
  P.nwave = m->nwave;
  P.nwaveblocks = gb_spmv_nwaveblocks(m);
  P.nblockrows = m->nblockrows;
  P.nwork = gb_spmv_nwork(m);
  P.exact = m->params.exact;
  P.base = base;
  return P;
}

extern "C" {

size_t gbd_batch_workspace_bytes(const gbd_matrix *m, int precision) {
  return gb_b_plan_layout(m, precision).total;
}

int gbd_batch_create(uint32_t max_slots, double delta, gbd_batch **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (max_slots == 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_batch *B = (gbd_batch *)calloc(1, sizeof(gbd_batch));
  if (!B) return GOOSEBERRY_ERROR_NOMEMORY;
  B->max_slots = max_slots;
  B->delta = delta > 0 && delta < 1 ? delta : 0.05;
  B->used = (uint8_t *)calloc(max_slots, 1);
  B->slot_mem = (void **)calloc(max_slots, sizeof(void *));
  B->slot_mat = (gbd_matrix **)calloc(max_slots, sizeof(gbd_matrix *));
  B->h_slice_nwork = (uint32_t *)calloc((size_t)max_slots * GB_SPLIT_MAX, sizeof(uint32_t));
  B->h_xt = (uint32_t *)calloc(max_slots, sizeof(uint32_t));
  B->h_xt_grid = (uint32_t *)calloc(max_slots, sizeof(uint32_t));
  B->h_sys = (gb_bsys *)calloc(max_slots, sizeof(gb_bsys));
  hipError_t e = gb_dev_alloc((void **)&B->d_sys, max_slots * sizeof(gb_bsys));
  if (e == hipSuccess) e = gb_dev_alloc((void **)&B->d_st, max_slots * sizeof(gbd_bstate));
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->h_st, max_slots * sizeof(gbd_bstate), hipHostMallocDefault);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&B->d_lists, sizeof(gb_blists));
  if (e == hipSuccess)
    e = gb_dev_alloc((void **)&B->d_slices, (size_t)max_slots * GB_SPLIT_MAX * sizeof(gb_bplan));
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->h_lists, sizeof(gb_blists), hipHostMallocDefault);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&B->d_sysl, max_slots * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->h_sysl, max_slots * sizeof(uint32_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(B->d_lists, 0, sizeof(gb_blists), gb_stream());
  if (e == hipSuccess)
    e = hipMemsetAsync(B->d_st, 0, max_slots * sizeof(gbd_bstate), gb_stream());
  if (e == hipSuccess)
    e = hipMemsetAsync(B->d_sys, 0, max_slots * sizeof(gb_bsys), gb_stream());
  if (e != hipSuccess || !B->used || !B->slot_mem || !B->slot_mat || !B->h_sys ||
      !B->h_slice_nwork || !B->h_xt || !B->h_xt_grid) {
    gbd_batch_destroy(B);
    return e != hipSuccess ? gb_hip_fail(e, "gbd_batch_create", __FILE__, __LINE__)
                           : GOOSEBERRY_ERROR_NOMEMORY;
  }
  *out = B;
  return GOOSEBERRY_SUCCESS;
}

void gbd_batch_destroy(gbd_batch *B) {
  if (!B) return;
  (void)hipStreamSynchronize(gb_stream());
  gb_b_drop_graphs(B);
  if (B->slot_mem)
    for (uint32_t i = 0; i < B->max_slots; ++i) (void)gb_dev_free(B->slot_mem[i]);
  if (B->slot_mat && B->used)
    for (uint32_t i = 0; i < B->max_slots; ++i)
      if (B->used[i] && B->slot_mat[i]) B->slot_mat[i]->holders--;
  (void)gb_dev_free(B->d_sys);
  (void)gb_dev_free(B->d_st);
  (void)gb_dev_free(B->d_lists);
  (void)gb_dev_free(B->d_slices);
  (void)gb_dev_free(B->d_spmv);
  (void)gb_dev_free(B->d_desc);
  (void)gb_dev_free(B->d_elem);
  (void)gb_dev_free(B->d_sysl);
  if (B->h_st) (void)hipHostFree(B->h_st);
  if (B->h_lists) (void)hipHostFree(B->h_lists);
  if (B->h_spmv) (void)hipHostFree(B->h_spmv);
  if (B->h_elem) (void)hipHostFree(B->h_elem);
  if (B->h_sysl) (void)hipHostFree(B->h_sysl);
  free(B->used);
  free(B->slot_mem);
  free(B->slot_mat);
  free(B->h_slice_nwork);
  free(B->h_xt);
  free(B->h_xt_grid);
  free(B->h_sys);
  free(B);
}

int gbd_batch_add(gbd_batch *B, const gbd_matrix *m, const double *d_b,
                  double *d_x, double tol, int rel, uint64_t maxiter,
                  int precision, uint32_t *slot_out) {
  if (m->nrows != m->ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (m->nparts) {
    gb_set_error("batched solves of a matrix held as row parts (64-bit row pointers) are "
                 "not supported");
    return GOOSEBERRY_ERROR_UNIMPLEMENTED;
  }
  if (precision != GOOSEBERRY_PRECISION_FP64 &&
      precision != GOOSEBERRY_PRECISION_MIXED &&
      precision != GOOSEBERRY_PRECISION_MIXED_JACOBI)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  const bool jac = precision == GOOSEBERRY_PRECISION_MIXED_JACOBI;
  if (precision == GOOSEBERRY_PRECISION_MIXED && !m->val32)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (jac && (!m->val32j || !m->jscale || (m->nsplit && !m->sval32j)))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  // an x-tile plan: the layout's copy of the inner solve's FP32 values
  if (m->xt_ptr && precision != GOOSEBERRY_PRECISION_FP64 &&
      !(jac ? m->xt_val32j : m->xt_val32))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t slot = 0;
  while (slot < B->max_slots && B->used[slot]) ++slot;
  if (slot == B->max_slots) return GOOSEBERRY_ERROR_NOMEMORY;
  if (precision == GOOSEBERRY_PRECISION_MIXED && m->nsplit && !m->sval32)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  const uint32_t nch = gb_b_nchunks(m->nrows);
  const gb_b_layout L = gb_b_plan_layout(m, precision);
  uint8_t *mem = nullptr;
  GB_HIP(gb_dev_alloc((void **)&mem, L.total));
  gb_bsys S;
  memset(&S, 0, sizeof(S));
  S.pl = gb_b_plan_of(m, 0, jac);
  S.jscale = jac ? m->jscale : nullptr;
  S.n = m->nrows;
  S.nchunks = nch;
  S.nsplit = m->nsplit;
  S.nwork_split = gb_b_nwork_split(m);
  S.b = d_b;
  S.x = d_x;
  S.part1 = (double *)(mem + L.part1);
  S.part2 = (double *)(mem + L.part2);
  S.r = (double *)(mem + L.r);
  if (precision != GOOSEBERRY_PRECISION_FP64) {
    S.d32 = (float *)(mem + L.d32);
    S.r32 = (float *)(mem + L.r32);
    S.p32 = (float *)(mem + L.p32);
    S.ap32 = (float *)(mem + L.ap32);
  } else {
    S.p = (double *)(mem + L.p);
    S.ap = (double *)(mem + L.ap);
  }
  S.st = B->d_st + slot;
  S.precision = precision;
  if (m->xt_ptr) {
    S.xt_ptr = m->xt_ptr;
    S.xt_idx = m->xt_idx;
    S.xt_val = m->xt_val;
    S.xt_val32 = jac ? m->xt_val32j : m->xt_val32;
    S.xt_R = m->xt_R;
    S.xt_W = m->xt_W;
    S.xt_T = m->xt_T;
    S.xt_lc = m->xt_lc;
  }
  hipStream_t s = gb_stream();
  gb_bplan *hsl = nullptr;
  if (m->nsplit) {
    // the slices' plans, their items numbered after the matrix's own
    S.ypart = mem + L.ypart;
    S.slices = (const gb_bplan *)(mem + L.slices);
    S.part3 = (double *)(mem + L.part3);
    hsl = (gb_bplan *)calloc(m->nsplit, sizeof(gb_bplan));
    if (!hsl) {
      (void)gb_dev_free(mem);
      return GOOSEBERRY_ERROR_NOMEMORY;
    }
    uint32_t base = 0;
    for (uint32_t r = 0; r < m->nsplit; ++r) {
      hsl[r] = gb_b_plan_of(m->split[r], base, jac);
      base += hsl[r].nwork;
      B->h_slice_nwork[(size_t)slot * GB_SPLIT_MAX + r] = hsl[r].nwork;
    }
    hipError_t e = hipMemcpyAsync(mem + L.slices, hsl, m->nsplit * sizeof(gb_bplan),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(B->d_slices + (size_t)slot * GB_SPLIT_MAX, hsl,
                         m->nsplit * sizeof(gb_bplan), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    free(hsl);
    if (e != hipSuccess) {
      (void)gb_dev_free(mem);
      return gb_hip_fail(e, "gbd_batch_add slices", __FILE__, __LINE__);
    }
  }
  B->h_sys[slot] = S;
  GB_HIP(hipMemcpyAsync(B->d_sys + slot, &B->h_sys[slot], sizeof(gb_bsys),
                        hipMemcpyHostToDevice, s));
  uint32_t g = nch > 1024 ? 1024 : (nch ? nch : 1);
  gb_b_init<<<g, GB_BLOCK, 0, s>>>(B->d_sys, slot);
  GB_LAUNCH_CHECK();
  gb_b_init_state<<<1, GB_BLOCK, 0, s>>>(B->d_sys, slot, tol, rel,
                                         (unsigned long long)maxiter);
  GB_LAUNCH_CHECK();
  // the staging copy of h_sys[slot] must not change before the copy runs
  GB_HIP(hipStreamSynchronize(s));
  B->used[slot] = 1;
  B->slot_mem[slot] = mem;
  // the slot's descriptor holds the plan's arrays: no replan or free of the
  // matrix until the slot is removed
  B->slot_mat[slot] = const_cast<gbd_matrix *>(m);
  B->slot_mat[slot]->holders++;
  *slot_out = slot;
  return GOOSEBERRY_SUCCESS;
}

int gbd_batch_remove(gbd_batch *B, uint32_t slot) {
  if (slot >= B->max_slots || !B->used[slot]) return GOOSEBERRY_ERROR_INVALIDINPUT;
  // ordered after the steps already queued that still read the workspace
  GB_HIP(gb_dev_free(B->slot_mem[slot]));
  B->slot_mem[slot] = nullptr;
  B->used[slot] = 0;
  if (B->slot_mat[slot]) B->slot_mat[slot]->holders--;
  B->slot_mat[slot] = nullptr;
  return GOOSEBERRY_SUCCESS;
}

// Rebuild the work lists from the active slots (active[i] != 0 and in use).
int gbd_batch_commit(gbd_batch *B, const uint8_t *active) {
  hipStream_t s = gb_stream();
  uint64_t ns = 0, ne = 0;
  uint32_t nsys = 0;
  for (uint32_t i = 0; i < B->max_slots; ++i) {
    if (!B->used[i] || (active && !active[i])) continue;
    if (!B->h_sys[i].xt_ptr)
      ns += B->h_sys[i].nsplit ? B->h_sys[i].nwork_split : B->h_sys[i].pl.nwork;
    ne += B->h_sys[i].nchunks;
    ++nsys;
  }
  if (ns > 0xffffffffull || ne > 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  GB_HIP(hipStreamSynchronize(s));  // staging buffers are reused
  auto resize = [](uint64_t need, uint32_t cap) {
    return need > cap || (cap > 4 * GB_BLIST_MIN && need < cap / 4);
  };
  if (resize(ns, B->cap_spmv) || !B->d_spmv) {
    uint32_t cap = (uint32_t)(ns + ns / 2 + GB_BLIST_MIN);
    if (B->d_spmv) GB_HIP(gb_dev_free(B->d_spmv));
    if (B->d_desc) GB_HIP(gb_dev_free(B->d_desc));
    if (B->h_spmv) GB_HIP(hipHostFree(B->h_spmv));
    GB_HIP(gb_dev_alloc((void **)&B->d_spmv, (size_t)cap * 8));
    GB_HIP(gb_dev_alloc((void **)&B->d_desc, (size_t)cap * sizeof(uint4)));
    GB_HIP(hipHostMalloc((void **)&B->h_spmv, (size_t)cap * 8, hipHostMallocDefault));
    B->cap_spmv = cap;
    gb_b_drop_graphs(B);  // the graphs hold the old list pointer
  }
  if (resize(ne, B->cap_elem) || !B->d_elem) {
    uint32_t cap = (uint32_t)(ne + ne / 2 + GB_BLIST_MIN);
    if (B->d_elem) GB_HIP(gb_dev_free(B->d_elem));
    if (B->h_elem) GB_HIP(hipHostFree(B->h_elem));
    GB_HIP(gb_dev_alloc((void **)&B->d_elem, (size_t)cap * 8));
    GB_HIP(hipHostMalloc((void **)&B->h_elem, (size_t)cap * 8, hipHostMallocDefault));
    B->cap_elem = cap;
    gb_b_drop_graphs(B);
  }
  // SpMV segments: slice k of every split matrix in segment k, the other
  // matrices in segment 0
  uint32_t nseg = 1;
  uint64_t cnt[GB_SPLIT_MAX] = {0};
  bool banded = true;  // every active system's matrix has 16-bit offsets
  for (uint32_t i = 0; i < B->max_slots; ++i) {
    if (!B->used[i] || (active && !active[i])) continue;
    const gb_bsys &S = B->h_sys[i];
    if (S.xt_ptr) continue;  // launched on its own
    banded = banded && !S.nsplit && S.pl.dcol;
    if (S.nsplit) {
      if (S.nsplit > nseg) nseg = S.nsplit;
      for (uint32_t r = 0; r < S.nsplit; ++r)
        cnt[r] += B->h_slice_nwork[(size_t)i * GB_SPLIT_MAX + r];
    } else {
      cnt[0] += S.pl.nwork;
    }
  }
  if (banded != B->banded) {
    B->banded = banded;
    gb_b_drop_graphs(B);  // the graphs hold the SpMV build
  }
  bool regrid = nseg != B->nseg;
  for (uint32_t g = 0; g < nseg; ++g) regrid = regrid || resize(cnt[g], B->cap_seg[g]);
  if (regrid) {
    for (uint32_t g = 0; g < GB_SPLIT_MAX; ++g)
      B->cap_seg[g] = g < nseg ? (uint32_t)(cnt[g] + cnt[g] / 2 + GB_BLIST_MIN) : 0;
    B->nseg = nseg;
    gb_b_drop_graphs(B);  // the graphs hold the launch grids
  }
  uint64_t a = 0, c = 0;
  uint32_t q = 0;
  for (uint32_t g = 0; g < nseg; ++g) {
    B->h_lists->seg[g] = (uint32_t)a;
    for (uint32_t i = 0; i < B->max_slots; ++i) {
      if (!B->used[i] || (active && !active[i])) continue;
      const gb_bsys &S = B->h_sys[i];
      uint32_t nw = 0;
      if (S.xt_ptr) nw = 0;
      else if (S.nsplit) nw = g < S.nsplit ? B->h_slice_nwork[(size_t)i * GB_SPLIT_MAX + g] : 0;
      else if (g == 0) nw = S.pl.nwork;
      for (uint32_t w = 0; w < nw; ++w) {
        B->h_spmv[2 * a] = i;
        B->h_spmv[2 * a + 1] = w;
        ++a;
      }
    }
  }
  B->h_lists->seg[nseg] = (uint32_t)a;
  B->h_lists->nseg = nseg;
  // the x-tile systems' launches (the graphs hold them)
  uint32_t nx = 0;
  bool xchanged = false;
  for (uint32_t i = 0; i < B->max_slots; ++i) {
    if (!B->used[i] || (active && !active[i]) || !B->h_sys[i].xt_ptr) continue;
    xchanged = xchanged || nx >= B->n_xt || B->h_xt[nx] != i ||
               B->h_xt_grid[nx] != B->h_sys[i].xt_W;
    B->h_xt[nx] = i;
    B->h_xt_grid[nx] = B->h_sys[i].xt_W;
    ++nx;
  }
  if (xchanged || nx != B->n_xt) {
    B->n_xt = nx;
    gb_b_drop_graphs(B);
  }
  for (uint32_t i = 0; i < B->max_slots; ++i) {
    if (!B->used[i] || (active && !active[i])) continue;
    for (uint32_t k = 0; k < B->h_sys[i].nchunks; ++k) {
      B->h_elem[2 * c] = i;
      B->h_elem[2 * c + 1] = k;
      ++c;
    }
    B->h_sysl[q++] = i;
  }
  B->h_lists->n_spmv = (uint32_t)ns;
  B->h_lists->n_elem = (uint32_t)ne;
  B->h_lists->n_sys = nsys;
  if (ns) {
    GB_HIP(hipMemcpyAsync(B->d_spmv, B->h_spmv, ns * 8, hipMemcpyHostToDevice, s));
    for (uint32_t g = 0; g < nseg; ++g) {
      const uint32_t a0 = B->h_lists->seg[g], a1 = B->h_lists->seg[g + 1];
      if (a1 > a0) {
        uint32_t gr = gb_div_up(a1 - a0, GB_BLOCK);
        if (gr > 4096u) gr = 4096u;
        gb_b_describe<<<gr, GB_BLOCK, 0, s>>>(B->d_sys, B->d_spmv, a0, a1, g, B->d_slices,
                                              B->d_desc);
        GB_LAUNCH_CHECK();
      }
    }
  }
  if (ne) GB_HIP(hipMemcpyAsync(B->d_elem, B->h_elem, ne * 8, hipMemcpyHostToDevice, s));
  if (nsys)
    GB_HIP(hipMemcpyAsync(B->d_sysl, B->h_sysl, (size_t)nsys * 4, hipMemcpyHostToDevice, s));
  GB_HIP(hipMemcpyAsync(B->d_lists, B->h_lists, sizeof(gb_blists), hipMemcpyHostToDevice, s));
  return GOOSEBERRY_SUCCESS;
}

// k steps, replayed from graphs of 2^i steps captured on first use.
int gbd_batch_step(gbd_batch *B, uint32_t k) {
  hipStream_t s = gb_stream();
  if (!B->d_spmv || !B->d_elem) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (!gb_graphs_enabled()) return gb_b_enqueue_raw(B, k);
  for (int i = GB_B_NGRAPH - 1; i >= 0 && k; --i) {
    const uint32_t len = 1u << i;
    while (k >= len) {
      if (!B->gexec[i]) {
        GB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rv = gb_b_enqueue_raw(B, len);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rv != GOOSEBERRY_SUCCESS) {
          if (g) (void)hipGraphDestroy(g);
          return rv;
        }
        GB_HIP(e);
        B->graph[i] = g;
        GB_HIP(hipGraphInstantiate(&B->gexec[i], g, nullptr, nullptr, 0));
      }
      GB_HIP(hipGraphLaunch(B->gexec[i], s));
      k -= len;
    }
  }
  return GOOSEBERRY_SUCCESS;
}

int gbd_batch_poll(gbd_batch *B, gbd_bstate *states) {
  hipStream_t s = gb_stream();
  GB_HIP(hipMemcpyAsync(B->h_st, B->d_st, (size_t)B->max_slots * sizeof(gbd_bstate),
                        hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  memcpy(states, B->h_st, (size_t)B->max_slots * sizeof(gbd_bstate));
  return GOOSEBERRY_SUCCESS;
}

int gbd_batch_slot_used(const gbd_batch *B, uint32_t slot) {
  return slot < B->max_slots && B->used[slot];
}

}  // extern "C"
