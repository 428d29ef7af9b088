// gb_spmv.hip -- CSR SpMV for gfx950 with row-length binning.
//
// Reference: src/blas.c:67-92 (__sparse_mult_dense_block) driven by
// src/blas.c:552-606 (blas_spmult).  Rows are binned by length on the GPU
// (stable counting sort of row ids by bin) into
//   * short rows (len <= short_max): LDS-staged row blocks.  A block is a
//     contiguous run of short rows whose nonzeros start inside one window
//     of `window` nonzeros (at most rowcap rows).  The block streams its
//     values/indices with fully coalesced loads, stages the products
//     val*x[col] in LDS, then one thread per row adds them left to right
//     from +0.0 -- the reference's exact operation order, so these rows are
//     bitwise identical to the reference (no FMA: this file is compiled
//     with -ffp-contract=off).
//   * medium rows (short_max < len <= wave_max): one wavefront per row,
//     lanes stride the row, 64-lane butterfly reduction.
//   * long rows (len > wave_max): one 256-thread workgroup per row.
// All three run in one launch: blockIdx ranges select the bin.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>

#include "gb_matrix.h"
#include "gb_xtile.h"

// ------------------------------------------------------------- plan build

__device__ __forceinline__ uint32_t gb_lower_bound(const uint32_t *a,
                                                   uint32_t n, uint32_t v) {
  // first i in [0, n) with a[i] >= v (n if none)
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t gb_row_class(uint32_t len,
                                                 const gbd_plan_params p) {
  return len <= p.short_max ? 0u : (len <= p.wave_max ? 1u : 2u);
}

// Per row: bin (0 short / 1 wave / 2 block), the "starts a row block" flag
// and the log2 row-length histogram (hist[0..32]); hist[33] counts the short
// rows whose first or last column rules out 16-bit offsets from any block
// they can fall in (column - row above 32767, or below -32768 - (rowcap - 1)),
// so that the plan skips building offsets that cannot fit.
__global__ void __launch_bounds__(GB_BLOCK)
gb_plan_classify(const uint32_t *__restrict__ rowptr,
                 const uint32_t *__restrict__ rowind, uint32_t nrows, uint32_t cshift,
                 gbd_plan_params p, uint32_t *__restrict__ cls,
                 uint32_t *__restrict__ flag,
                 unsigned long long *__restrict__ hist) {
  __shared__ uint32_t h[34];
  if (threadIdx.x < 34) h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t r = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (r < nrows) {
    uint32_t a = rowptr[r], b = rowptr[r + 1];
    uint32_t len = b - a;
    uint32_t c = gb_row_class(len, p);
    cls[r] = c;
    // floor(log2(len + 1)); len + 1 wraps to 0 only for len = 2^32 - 1
    uint32_t bin = (len == 0xffffffffu) ? 32u : 31u - __clz(len + 1);
    atomicAdd(&h[bin], 1u);
    if (c == 0 && len && !p.wide_index) {
      const int64_t lo = -32768 - (int64_t)(p.rowcap - 1);
      const int64_t rr = (int64_t)r + cshift;
      const int64_t d0 = (int64_t)rowind[a] - rr, d1 = (int64_t)rowind[b - 1] - rr;
      if (d0 > 32767 || d0 < lo || d1 > 32767 || d1 < lo) atomicAdd(&h[33], 1u);
    }
    uint32_t f = 0;
    if (c == 0) {
      if (r == 0) {
        f = 1;
      } else {
        uint32_t pa = rowptr[r - 1];
        uint32_t pc = gb_row_class(a - pa, p);
        uint32_t w = a / p.window;
        if (pc != 0 || pa / p.window != w) {
          f = 1;
        } else {
          uint32_t first = gb_lower_bound(rowptr, nrows + 1, w * p.window);
          if ((r - first) % p.rowcap == 0) f = 1;
        }
      }
    }
    flag[r] = f;
  }
  __syncthreads();
  if (threadIdx.x < 34 && h[threadIdx.x])
    atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// Structural check of a device CSR before any kernel indexes through it:
// rowptr[0] == 0, non-decreasing, and every column index < ncols.
__global__ void gb_validate_csr(const uint32_t *__restrict__ rowptr,
                                const uint32_t *__restrict__ rowind,
                                uint32_t nrows, uint32_t ncols, uint32_t nnz,
                                uint32_t *__restrict__ bad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  uint32_t b = 0;
  for (size_t r = i; r < nrows; r += stride) {
    uint32_t a = rowptr[r], e = rowptr[r + 1];
    if (e < a || e > nnz) b = 1;
    if (r == 0 && a != 0) b = 1;
  }
  for (size_t k = i; k < nnz; k += stride)
    if (rowind[k] >= ncols) b = 1;
  if (b) atomicOr(bad, 1u);
}

// The same check of a CSR with 64-bit row pointers.
__global__ void gb_validate_csr64(const uint64_t *__restrict__ rowptr,
                                  const uint32_t *__restrict__ rowind,
                                  uint32_t nrows, uint32_t ncols, uint64_t nnz,
                                  uint32_t *__restrict__ bad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  uint32_t b = 0;
  for (size_t r = i; r < nrows; r += stride) {
    uint64_t a = rowptr[r], e = rowptr[r + 1];
    if (e < a || e > nnz) b = 1;
    if (r == 0 && a != 0) b = 1;
  }
  for (size_t k = i; k < nnz; k += stride)
    if (rowind[k] >= ncols) b = 1;
  if (b) atomicOr(bad, 1u);
}

// out[i] = in[i] - base for i < n (64-bit row pointers to the 32-bit ones
// of a row part, or of a whole matrix with fewer than 2^32 entries)
__global__ void gb_rebase_rowptr(const uint64_t *__restrict__ in, uint64_t base, uint64_t n,
                                 uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)(in[i] - base);
}

// idx = exclusive scan of flag.  Block id of a short row r is
// idx[r] + flag[r] - 1.
__global__ void __launch_bounds__(GB_BLOCK)
gb_plan_blocks(const uint32_t *__restrict__ cls,
               const uint32_t *__restrict__ flag,
               const uint32_t *__restrict__ idx, uint32_t nrows,
               uint32_t *__restrict__ blk) {
  uint32_t r = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (r >= nrows || cls[r] != 0) return;
  uint32_t id = idx[r] + flag[r] - 1;
  if (flag[r]) blk[2 * id] = r;
  bool end = (r + 1 == nrows) || cls[r + 1] != 0 || flag[r + 1];
  if (end) blk[2 * id + 1] = r + 1;
}

// Each row block's first row, end row, first entry and entry count, so
// that the SpMV starts a block's data loads after one load instead of the
// block table and then the row pointers.
__global__ void __launch_bounds__(GB_BLOCK)
gb_plan_desc(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ blk,
             uint32_t nst, uint4 *__restrict__ desc) {
  const uint32_t b = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (b >= nst) return;
  const uint32_t r0 = blk[2 * b], r1 = blk[2 * b + 1];
  const uint32_t k0 = rowptr[r0];
  desc[b] = make_uint4(r0, r1, k0, rowptr[r1] - k0);
}

// dcol[k] = rowind[k] - r0 for the entries of each row block [r0, r1),
// one workgroup per block; *bad is set when an offset leaves int16.  On a
// matrix that does not fit, most blocks find offsets out of range: one
// atomic per workgroup, and workgroups that start after the flag is set
// (as their XCD sees it) return at once, so the atomics do not pile up on
// the one word.
__global__ void __launch_bounds__(GB_BLOCK)
gb_plan_narrow(const uint32_t *__restrict__ rowptr,
               const uint32_t *__restrict__ rowind,
               const uint32_t *__restrict__ blk, uint32_t cshift,
               int16_t *__restrict__ dcol, uint32_t *__restrict__ bad) {
  __shared__ uint32_t s_bad;
  if (threadIdx.x == 0)
    s_bad = __hip_atomic_load(bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (s_bad) return;
  const uint32_t r0 = blk[2 * blockIdx.x], r1 = blk[2 * blockIdx.x + 1];
  const uint32_t k1 = rowptr[r1];
  bool out = false;
  for (uint32_t k = rowptr[r0] + threadIdx.x; k < k1; k += GB_BLOCK) {
    const int64_t d = (int64_t)rowind[k] - ((int64_t)r0 + cshift);
    out |= d < -32768 || d > 32767;
    dcol[k] = (int16_t)d;
  }
  if (__syncthreads_or(out) && threadIdx.x == 0) atomicOr(bad, 1u);
}

// --------------------------------------------------------------- kernels

// One launch covers all bins.  Workgroups walk the work items
// w = blockIdx.x, blockIdx.x + gridDim.x, ...: first the LDS-staged row
// blocks, then groups of four wave rows, then long rows.  By default the
// grid has one workgroup per work item; a capped grid (plan.grid) makes it
// persistent.  With xcd_remap the row blocks are renumbered so that each
// XCD walks one contiguous range.
// With DOT the kernel also forms the block's share of p.Ap (x = p, y = Ap)
// and writes it to partials[blockIdx.x]; the next CG kernel adds the
// partials in a fixed order (no fences or atomics: deterministic).
// The pipelined row-block kernel's workgroup: 512 threads (one row of a
// block per thread), built for eight wavefronts per SIMD (at most 64
// VGPRs), four workgroups per CU.  Measured on the 160^3 Laplacian with
// value indices, per SpMV / CG iteration: 256 threads at four waves per SIMD
// 55.5 / 100.4 us, 512 threads at eight 54.1 / 95.5 us.
#define GB_SPMV_PIPE_THREADS 512
#define GB_SPMV_PIPE_WAVES 8
#ifndef GB_SPMV_WAVES
// wavefronts per SIMD the register budget must allow: 72 VGPRs (the FP64
// variants spill 2 to scratch).  Measured on the 160^3 Laplacian with the
// 16-bit offsets, per CG iteration: 5 waves 118.1 us, 6: 114.9, 7: 114.7.
#define GB_SPMV_WAVES 7
#endif
// With NARROW the row blocks read dcol, each entry's column minus the
// block's first row as a 16-bit signed offset, instead of rowind (the plan
// made dcol only when every block's offsets fit): 2 bytes an entry less.
// With VD the values are read as 8-bit indices into the table vtab of the
// matrix's distinct values (1 byte an entry instead of 8 or 4).
template <typename T, bool DOT, bool NARROW, bool VD>
__global__ void __launch_bounds__(GB_BLOCK) __attribute__((amdgpu_waves_per_eu(GB_SPMV_WAVES)))
gb_spmv_kernel(const uint32_t *__restrict__ rowptr,
               const uint32_t *__restrict__ rowind,
               const int16_t *__restrict__ dcol,
               const T *__restrict__ val, const uint8_t *__restrict__ vidx,
               const T *__restrict__ vtab, const T *__restrict__ x,
               T *__restrict__ y, uint32_t ystride, uint32_t cshift,
               const uint4 *__restrict__ bdesc, uint32_t nstream,
               const uint32_t *__restrict__ wave_rows, uint32_t nwave,
               uint32_t nwaveblocks,
               const uint32_t *__restrict__ block_rows, uint32_t nwork,
               uint32_t exact, uint32_t remap, double *__restrict__ partials,
               const gbd_cg_state *st) {
  __shared__ T prod[GB_STREAM_LDS];
  __shared__ double red[GB_WAVES_PER_BLOCK];
  __shared__ T tred[GB_WAVES_PER_BLOCK];
  // With DOT the launch writes only Ap and the partials, which nothing reads
  // once the solve has stopped, so it does not wait on the stop flag (that
  // read would add a dependent memory round trip to every workgroup).
  (void)st;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (GB_WAVE - 1);
  const uint32_t wave = tid / GB_WAVE;
  double dotp = 0.0;

  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    if (w < nstream) {
      // ---- LDS-staged row block of short rows
      const uint32_t lb = remap == 0 ? w
                          : remap == 1 ? gb_xcd_remap(w, nstream)
                                       : gb_xcd_chunk_remap(w, nstream, remap);
      const uint4 bd = bdesc[lb];
      const uint32_t r0 = bd.x, r1 = bd.y;
      const uint32_t k0 = bd.z;
      const uint32_t cnt = bd.w;
      // row bounds of this thread's (at most two) rows, loaded together
      // with the block's nonzeros so that all loads are in flight at once
      const uint32_t ra = r0 + tid, rb = r0 + tid + GB_BLOCK;
      uint32_t pa0 = 0, pa1 = 0, pb0 = 0, pb1 = 0;
      T xa = 0, xb = 0;  // p[row] for the dot product, fetched up front
      if (ra < r1) {
        pa0 = rowptr[ra];
        pa1 = rowptr[ra + 1];
        if (DOT) xa = x[ra];
      }
      if (rb < r1) {
        pb0 = rowptr[rb];
        pb1 = rowptr[rb + 1];
        if (DOT) xb = x[rb];
      }
      if (cnt > 0) {
        // stage val*x[col] for the block's nonzeros; out-of-range lanes
        // re-read the last entry (in bounds) and store nothing
        const uint32_t klast = k0 + cnt - 1;
        // with NARROW, c holds the offsets and the gathers index x + r0
        // (the add stays out of the load branches, which then need no
        // wait for the loaded value).  Measured on the 160^3 Laplacian:
        // offsets and FP32 values are faster as plain loads, FP64 values
        // as non-temporal ones.
        uint32_t c[GB_STREAM_ITERS];
        T v[GB_STREAM_ITERS], xv[GB_STREAM_ITERS];
#pragma unroll
        for (int i = 0; i < GB_STREAM_ITERS; ++i) {
          if (i * GB_BLOCK < cnt) {
            uint32_t k = min(k0 + i * GB_BLOCK + tid, klast);
            if (NARROW) {
              c[i] = (uint32_t)(int32_t)dcol[k];
              if (VD) v[i] = vtab[vidx[k]];
              else if (sizeof(T) == 4) v[i] = val[k];
              else v[i] = __builtin_nontemporal_load(&val[k]);
            } else {
              c[i] = __builtin_nontemporal_load(&rowind[k]);
              v[i] = gb_ld_val<T, VD>(val, vidx, vtab, k);
            }
          }
        }
        const T *xg = NARROW ? x + r0 + cshift : x;
#pragma unroll
        for (int i = 0; i < GB_STREAM_ITERS; ++i) {
          if (i * GB_BLOCK < cnt) {
            if (NARROW) xv[i] = xg[(int32_t)c[i]];
            else xv[i] = xg[c[i]];
          }
        }
#pragma unroll
        for (int i = 0; i < GB_STREAM_ITERS; ++i) {
          uint32_t e = i * GB_BLOCK + tid;
          if (i * GB_BLOCK < cnt && e < cnt) prod[e] = v[i] * xv[i];
        }
      }
      __syncthreads();
      // one thread per row: the reference's left-to-right sum from +0.0
      if (ra < r1) {
        T s = 0;
        for (uint32_t k = pa0 - k0; k < pa1 - k0; ++k) s += prod[k];
        y[(size_t)ra * ystride] = s;
        if (DOT) dotp += (double)xa * (double)s;
      }
      if (rb < r1) {
        T s = 0;
        for (uint32_t k = pb0 - k0; k < pb1 - k0; ++k) s += prod[k];
        y[(size_t)rb * ystride] = s;
        if (DOT) dotp += (double)xb * (double)s;
      }
    } else if (w < nstream + nwaveblocks) {
      // ---- one wavefront per medium row
      const uint32_t wid = (w - nstream) * GB_WAVES_PER_BLOCK + wave;
      if (wid < nwave) {
        const uint32_t row = wave_rows[wid];
        const uint32_t k0 = rowptr[row], k1 = rowptr[row + 1];
        T s = 0;
        if (exact) {
          // 64 products at a time, folded left to right across the lanes
          for (uint32_t base = k0; base < k1; base += GB_WAVE) {
            const uint32_t k = base + lane;
            T pr = 0;
            if (k < k1) {
              uint32_t c = __builtin_nontemporal_load(&rowind[k]);
              T v = gb_ld_val<T, VD>(val, vidx, vtab, k);
              pr = v * x[c];
            }
            const uint32_t cnt = min((uint32_t)GB_WAVE, k1 - base);
            for (uint32_t j = 0; j < cnt; ++j) s += gb_readlane(pr, j);
          }
        } else {
          s = gb_wave_sum(
              gb_row_lane_sum<T, GB_WAVE, VD>(rowind, val, x, k0, k1, lane, vidx, vtab));
        }
        if (lane == 0) {
          y[(size_t)row * ystride] = s;
          if (DOT) dotp += (double)x[row] * (double)s;
        }
      }
    } else {
      // ---- one workgroup per long row
      const uint32_t row = block_rows[w - nstream - nwaveblocks];
      const uint32_t k0 = rowptr[row], k1 = rowptr[row + 1];
      if (exact) {
        // products staged in LDS a chunk at a time, thread 0 adds them in
        // order
        T s = 0;
        for (uint32_t base = k0; base < k1; base += GB_STREAM_LDS) {
          const uint32_t cnt = min((uint32_t)GB_STREAM_LDS, k1 - base);
          for (uint32_t e = tid; e < cnt; e += GB_BLOCK) {
            uint32_t c = __builtin_nontemporal_load(&rowind[base + e]);
            T v = gb_ld_val<T, VD>(val, vidx, vtab, base + e);
            prod[e] = v * x[c];
          }
          __syncthreads();
          if (tid == 0)
            for (uint32_t e = 0; e < cnt; ++e) s += prod[e];
          __syncthreads();
        }
        if (tid == 0) {
          y[(size_t)row * ystride] = s;
          if (DOT) dotp += (double)x[row] * (double)s;
        }
      } else {
        T s = gb_wave_sum(
            gb_row_lane_sum<T, GB_BLOCK, VD>(rowind, val, x, k0, k1, tid, vidx, vtab));
        if (lane == 0) tred[wave] = s;
        __syncthreads();
        if (tid == 0) {
          T t = tred[0];
#pragma unroll
          for (int q = 1; q < GB_WAVES_PER_BLOCK; ++q) t += tred[q];
          y[(size_t)row * ystride] = t;
          if (DOT) dotp += (double)x[row] * (double)t;
        }
      }
    }
    // the LDS buffers are reused by the next work item
    if (w + gridDim.x < nwork) __syncthreads();
  }

  if (DOT) {
    // one partial per workgroup: the wavefront sums added in wave order
    // (`red` is used nowhere else, so no barrier is needed before it)
    const double ws = gb_wave_sum(dotp);
    if (lane == 0) red[wave] = ws;
    __syncthreads();
    if (tid == 0) {
      double t = red[0];
#pragma unroll
      for (int q = 1; q < GB_WAVES_PER_BLOCK; ++q) t += red[q];
      partials[blockIdx.x] = t;
    }
  }
}

// Persistent, software-pipelined form for matrices whose rows are all
// short (every work item an LDS-staged row block): workgroup w takes blocks
// w, w + G, w + 2G, ... and issues the next block's loads (its row bounds,
// column offsets or indices, and values or value indices) before it
// gathers, stages and adds the current block, so one block's memory
// latency overlaps the previous block's work.  Same products and sums as
// gb_spmv_kernel's row blocks; with DOT one p.Ap partial per workgroup.

// entries of one row block per thread of an NT-thread workgroup
template <uint32_t NT>
struct gb_pipe_iters {
  static constexpr uint32_t n = (GB_STREAM_LDS + NT - 1) / NT;
};

template <typename T, bool DOT, bool NARROW, bool VD, uint32_t NT>
struct gb_pipe_blk {
  static constexpr uint32_t ITERS = gb_pipe_iters<NT>::n;
  uint32_t r0, r1, k0, cnt;
  uint32_t pa0, pa1, pb0, pb1;
  T xa, xb;
  uint32_t c[ITERS];
  // values, or (VD) value indices
  typename std::conditional<VD, uint32_t, T>::type v[ITERS];
};

// rows per block are at most 2 GB_BLOCK (the plan's rowcap): one or two
// rows per thread
template <typename T, bool DOT, bool NARROW, bool VD, uint32_t NT>
__device__ __forceinline__ void gb_pipe_load(gb_pipe_blk<T, DOT, NARROW, VD, NT> &B,
                                             const uint4 d,
                                             const uint32_t *__restrict__ rowptr,
                                             const uint32_t *__restrict__ rowind,
                                             const int16_t *__restrict__ dcol,
                                             const T *__restrict__ val,
                                             const uint8_t *__restrict__ vidx,
                                             const T *__restrict__ x, uint32_t tid) {
  constexpr bool TWO = NT < 2 * GB_BLOCK;
  B.r0 = d.x;
  B.r1 = d.y;
  B.k0 = d.z;
  B.cnt = d.w;
  const uint32_t ra = d.x + tid, rb = d.x + tid + NT;
  B.pa0 = B.pa1 = B.pb0 = B.pb1 = 0;
  B.xa = B.xb = 0;
  if (ra < d.y) {
    B.pa0 = rowptr[ra];
    B.pa1 = rowptr[ra + 1];
    if (DOT) B.xa = x[ra];
  }
  if (TWO && rb < d.y) {
    B.pb0 = rowptr[rb];
    B.pb1 = rowptr[rb + 1];
    if (DOT) B.xb = x[rb];
  }
  if (d.w == 0) return;
  const uint32_t klast = d.z + d.w - 1;
#pragma unroll
  for (uint32_t i = 0; i < B.ITERS; ++i) {
    if (i * NT < d.w) {
      const uint32_t k = min(d.z + i * NT + tid, klast);
      if (NARROW) B.c[i] = (uint32_t)(int32_t)dcol[k];
      else B.c[i] = __builtin_nontemporal_load(&rowind[k]);
      if (VD) B.v[i] = vidx[k];
      else if (NARROW && sizeof(T) == 4) B.v[i] = val[k];
      else B.v[i] = __builtin_nontemporal_load(&val[k]);
    }
  }
}

template <typename T, bool DOT, bool NARROW, bool VD, uint32_t NT, int WAVES>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES)))
gb_spmv_pipe(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
             const int16_t *__restrict__ dcol, const T *__restrict__ val,
             const uint8_t *__restrict__ vidx, const T *__restrict__ vtab,
             const T *__restrict__ x, T *__restrict__ y, uint32_t ystride, uint32_t cshift,
             const uint4 *__restrict__ bdesc, uint32_t nstream, uint32_t nvals,
             double *__restrict__ partials) {
  constexpr bool TWO = NT < 2 * GB_BLOCK;
  using blk_t = gb_pipe_blk<T, DOT, NARROW, VD, NT>;
  __shared__ T prod[GB_STREAM_LDS];
  __shared__ T ltab[VD ? GB_VDICT_MAX : 1];
  __shared__ double red[NT / GB_WAVE];
  const uint32_t tid = threadIdx.x;
  const uint32_t G = gridDim.x;
  if (VD) {
    for (uint32_t i = tid; i < nvals; i += NT) ltab[i] = vtab[i];
  }
  double dotp = 0.0;
  blk_t cur, nxt;
  uint32_t b = blockIdx.x;
  uint4 dn = make_uint4(0u, 0u, 0u, 0u);
  if (b < nstream) {
    const uint4 d0 = bdesc[b];
    if (b + G < nstream) dn = bdesc[b + G];
    gb_pipe_load(cur, d0, rowptr, rowind, dcol, val, vidx, x, tid);
  }
  if (VD) __syncthreads();  // the value table
  for (; b < nstream; b += G) {
    // the next block's loads go out before this block's gathers
    const bool more = b + G < nstream;
    uint4 dnn = make_uint4(0u, 0u, 0u, 0u);
    if (more) {
      if (b + 2 * G < nstream) dnn = bdesc[b + 2 * G];
      gb_pipe_load(nxt, dn, rowptr, rowind, dcol, val, vidx, x, tid);
    }
    const uint32_t r0 = cur.r0, k0 = cur.k0, cnt = cur.cnt;
    if (cnt > 0) {
      const T *xg = NARROW ? x + r0 + cshift : x;
      T xv[blk_t::ITERS];
#pragma unroll
      for (uint32_t i = 0; i < blk_t::ITERS; ++i) {
        if (i * NT < cnt) {
          if (NARROW) xv[i] = xg[(int32_t)cur.c[i]];
          else xv[i] = xg[cur.c[i]];
        }
      }
#pragma unroll
      for (uint32_t i = 0; i < blk_t::ITERS; ++i) {
        const uint32_t e = i * NT + tid;
        if (i * NT < cnt && e < cnt) {
          T v;
          if constexpr (VD) v = ltab[cur.v[i]];
          else v = cur.v[i];
          prod[e] = v * xv[i];
        }
      }
    }
    __syncthreads();
    // one thread per row: the reference's left-to-right sum from +0.0
    const uint32_t ra = r0 + tid, rb = r0 + tid + NT;
    if (ra < cur.r1) {
      T s = 0;
      for (uint32_t k = cur.pa0 - k0; k < cur.pa1 - k0; ++k) s += prod[k];
      y[(size_t)ra * ystride] = s;
      if (DOT) dotp += (double)cur.xa * (double)s;
    }
    if (TWO && rb < cur.r1) {
      T s = 0;
      for (uint32_t k = cur.pb0 - k0; k < cur.pb1 - k0; ++k) s += prod[k];
      y[(size_t)rb * ystride] = s;
      if (DOT) dotp += (double)cur.xb * (double)s;
    }
    __syncthreads();  // prod is reused by the next block
    if (more) {
      cur = nxt;
      dn = dnn;
    }
  }
  if (DOT) {
    const double ws = gb_wave_sum(dotp);
    if ((tid & (GB_WAVE - 1)) == 0) red[tid / GB_WAVE] = ws;
    __syncthreads();
    if (tid == 0) {
      double t = red[0];
#pragma unroll
      for (uint32_t q = 1; q < NT / GB_WAVE; ++q) t += red[q];
      partials[blockIdx.x] = t;
    }
  }
}

// ------------------------------------------------------------- x tiles
// For matrices whose columns are scattered over an x larger than an XCD's
// L2, every gather is an L2 request; the x-tile layout (gb_xtile.h) moves
// the gathers into LDS.

// the layout's keys (group of each entry) and packed (row, column) words
__global__ void __launch_bounds__(GB_BLOCK)
gb_xt_keys(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
           uint32_t nrows, uint32_t R, uint32_t T, uint32_t lc, uint32_t *__restrict__ keys,
           uint32_t *__restrict__ word) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= nrows) return;
  const uint32_t base = (i / R) * T, lr = (i % R) << lc, cm = (1u << lc) - 1;
  const uint32_t k1 = rowptr[i + 1];
  for (uint32_t k = rowptr[i]; k < k1; ++k) {
    const uint32_t c = rowind[k];
    keys[k] = base + (c >> lc);
    word[k] = lr | (c & cm);
  }
}

__global__ void gb_xt_gather(const uint32_t *__restrict__ perm, uint64_t nnz,
                             const uint32_t *__restrict__ word, const double *__restrict__ val,
                             uint32_t *__restrict__ idx, double *__restrict__ oval) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = perm[j];
    idx[j] = word[k];
    oval[j] = val[k];
  }
}

// y = A x on the x-tile layout (V = double, or float for the FP32 SpMV);
// with `partials` (FP64, CG) the block's share of x.Ax as well
template <typename V, int EPI>
__global__ void __launch_bounds__(GB_XT_THREADS)
gb_spmv_xtile(const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
              const V *__restrict__ val, const V *__restrict__ x, V *__restrict__ y,
              uint32_t ystride, double *__restrict__ partials, uint32_t nrows,
              uint32_t ncols, uint32_t R, uint32_t T, uint32_t lc) {
  __shared__ gb_xt_lds L;
  gb_xt_spmv<V, EPI>(ptr, idx, val, x, y, ystride, nullptr, partials, nrows, ncols, R, T,
                     lc, L);
}

uint32_t gb_spmv_grid(const gbd_matrix *m) {
  const uint64_t nwork = gb_spmv_nwork(m);
  // one workgroup per work item, the pipelined kernel's persistent grid, or
  // the plan's cap
  const uint64_t gmax = m->pipe_grid ? m->pipe_grid
                        : m->params.grid ? m->params.grid : 0xffffffffull;
  return (uint32_t)(nwork < gmax ? nwork : gmax);
}

// The pipelined row-block kernel runs matrices whose rows are all short
// (every work item an LDS-staged row block): on the plan's grid when it
// sets one, and by default for value-indexed plans, four workgroups per CU
// (the most that are resident at eight waves per SIMD).  Measured on the
// 160^3 Laplacian with 8-bit value indices (256-thread workgroups): 56
// against 64 us per SpMV, 100 against 111 us per CG iteration;
// with FP64 values read as stored it is slower (78 against 68 us), the
// stream then being HBM-bound.
static void gb_pipe_choose(gbd_matrix *m) {
  m->pipe_grid = 0;
  if (m->nwave || m->nblockrows || !m->nstream) return;
  if (m->params.grid) m->pipe_grid = m->params.grid;
  else if (m->vidx) m->pipe_grid = 4 * gb_num_cus();
}

// ------------------------------------------------------- column split

// cnt[r * nrows + i] = entries of row i in column slice r
__global__ void __launch_bounds__(GB_BLOCK)
gb_split_count(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
               uint32_t nrows, uint32_t width, uint32_t P, uint32_t *__restrict__ cnt) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= nrows) return;
  const uint32_t k0 = rowptr[i], k1 = rowptr[i + 1];
  for (uint32_t r = 0; r < P; ++r) {
    uint32_t c = 0;
    for (uint32_t k = k0; k < k1; ++k) c += (rowind[k] / width == r);
    cnt[(size_t)r * nrows + i] = c;
  }
}

// the entries of row i in slice r, in row order, from srowptr[r nrows + i]
template <typename T>
__global__ void __launch_bounds__(GB_BLOCK)
gb_split_fill(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
              const T *__restrict__ val, uint32_t nrows, uint32_t width, uint32_t P,
              const uint32_t *__restrict__ srowptr, uint32_t *__restrict__ srowind,
              T *__restrict__ sval) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= nrows) return;
  const uint32_t k0 = rowptr[i], k1 = rowptr[i + 1];
  for (uint32_t r = 0; r < P; ++r) {
    uint32_t p = srowptr[(size_t)r * nrows + i];
    for (uint32_t k = k0; k < k1; ++k) {
      const uint32_t c = rowind[k];
      if (c / width == r) {
        if (srowind) srowind[p] = c;
        sval[p] = val[k];
        ++p;
      }
    }
  }
}

// y[i] = ((yp_0[i] + yp_1[i]) + ...) + yp_{P-1}[i]; with DOT also the
// per-workgroup partials of x.y
template <typename T, bool DOT>
__global__ void __launch_bounds__(GB_BLOCK)
gb_split_sum(const T *__restrict__ yp, uint32_t n, uint32_t P, T *__restrict__ y,
             uint32_t ystride, const T *__restrict__ x, double *__restrict__ partials) {
  double dotp = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n; i += gridDim.x * GB_BLOCK) {
    T s = yp[i];
    for (uint32_t r = 1; r < P; ++r) s += yp[(size_t)r * n + i];
    y[(size_t)i * ystride] = s;
    if (DOT) dotp += (double)x[i] * (double)s;
  }
  if (DOT) {
    __shared__ double red[GB_WAVES_PER_BLOCK];
    const double ws = gb_wave_sum(dotp);
    if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = red[0];
#pragma unroll
      for (int q = 1; q < GB_WAVES_PER_BLOCK; ++q) t += red[q];
      partials[blockIdx.x] = t;
    }
  }
}

template <typename T>
static int gb_spmv_launch_t(const gbd_matrix *m, const T *val, const uint8_t *vidx,
                            const T *vtab, const T *x, T *y, uint32_t ystride,
                            double *partials, const gbd_cg_state *cg);

// per-workgroup partials of x.y over n rows, grid-stride in a fixed order
// (the p.Ap of a row-parted matrix, whose part launches compute no dot)
__global__ void __launch_bounds__(GB_BLOCK)
gb_dot_partials(const double *__restrict__ x, const double *__restrict__ y, uint32_t n,
                double *__restrict__ partials) {
  double dotp = 0.0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n; i += gridDim.x * GB_BLOCK)
    dotp += x[i] * y[i];
  __shared__ double red[GB_WAVES_PER_BLOCK];
  const double ws = gb_wave_sum(dotp);
  if ((threadIdx.x & (GB_WAVE - 1)) == 0) red[threadIdx.x / GB_WAVE] = ws;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
#pragma unroll
    for (int q = 1; q < GB_WAVES_PER_BLOCK; ++q) t += red[q];
    partials[blockIdx.x] = t;
  }
}

// a row-parted matrix: one SpMV per part into its rows of y (FP64, or FP32
// over the parts' views of the FP32 copy), then for CG (FP64) the p.Ap
// partials over the whole vectors
template <typename T>
static int gb_spmv_launch_parts(const gbd_matrix *m, const T *x, T *y, uint32_t ystride,
                                double *partials, const gbd_cg_state *cg) {
  for (uint32_t q = 0; q < m->nparts; ++q) {
    const gbd_matrix *v = m->parts[q];
    const T *vv, *vt;
    if constexpr (sizeof(T) == 8) {
      vv = v->val;
      vt = v->vtab;
    } else {
      vv = v->val32;
      vt = v->vtab32;
    }
    if (!vv) return GOOSEBERRY_ERROR_INVALIDINPUT;
    int rv = gb_spmv_launch_t<T>(v, vv, v->vidx, vt, x, y + (size_t)m->prow[q] * ystride,
                                 ystride, nullptr, nullptr);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  if constexpr (sizeof(T) == 8) {
    if (cg) {
      gb_dot_partials<<<gb_split_sum_grid(m), GB_BLOCK, 0, gb_stream()>>>(x, y, m->nrows,
                                                                           partials);
      GB_LAUNCH_CHECK();
    }
  }
  return GOOSEBERRY_SUCCESS;
}

// a split matrix: one launch per slice into its partial vector, then the sum
template <typename T>
static int gb_spmv_launch_split(const gbd_matrix *m, const T *x, T *y, uint32_t ystride,
                                double *partials, const gbd_cg_state *cg) {
  T *yp = (T *)m->ypart;
  for (uint32_t r = 0; r < m->nsplit; ++r) {
    const gbd_matrix *v = m->split[r];
    const T *vval;
    if constexpr (sizeof(T) == 8) vval = (const T *)v->val;
    else vval = (const T *)v->val32;
    int rv = gb_spmv_launch_t<T>(v, vval, nullptr, nullptr, x, yp + (size_t)r * m->nrows, 1,
                                 nullptr, nullptr);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  hipStream_t s = gb_stream();
  if (cg) {
    gb_split_sum<T, true><<<gb_split_sum_grid(m), GB_BLOCK, 0, s>>>(
        yp, m->nrows, m->nsplit, y, ystride, x, partials);
  } else {
    gb_split_sum<T, false><<<gb_split_sum_grid(m), GB_BLOCK, 0, s>>>(
        yp, m->nrows, m->nsplit, y, ystride, nullptr, nullptr);
  }
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

// the persistent pipelined kernel serves plans with a grid cap on matrices
// whose rows are all short
static inline bool gb_spmv_piped(const gbd_matrix *m) { return m->pipe_grid != 0; }

template <typename T, bool DOT, bool NARROW, bool VD>
static void gb_spmv_go(const gbd_matrix *m, uint32_t grid, hipStream_t s, const T *val,
                       const uint8_t *vidx, const T *vtab, const T *x, T *y,
                       uint32_t ystride, double *partials, const gbd_cg_state *cg) {
  if (gb_spmv_piped(m)) {
    gb_spmv_pipe<T, DOT, NARROW, VD, GB_SPMV_PIPE_THREADS, GB_SPMV_PIPE_WAVES>
        <<<grid, GB_SPMV_PIPE_THREADS, 0, s>>>(
            m->rowptr, m->rowind, NARROW ? m->dcol : nullptr, val, vidx, vtab, x, y, ystride,
            m->cshift, m->bdesc, m->nstream, m->nvals, partials);
    return;
  }
  const uint32_t *wave_rows = m->rowbins + (m->nrows - m->nwave - m->nblockrows);
  const uint32_t *block_rows = m->rowbins + (m->nrows - m->nblockrows);
  gb_spmv_kernel<T, DOT, NARROW, VD><<<grid, GB_BLOCK, 0, s>>>(
      m->rowptr, m->rowind, NARROW ? m->dcol : nullptr, val, vidx, vtab, x, y, ystride,
      m->cshift, m->bdesc, m->nstream, wave_rows, m->nwave, gb_spmv_nwaveblocks(m), block_rows,
      gb_spmv_nwork(m), m->params.exact, m->params.xcd_remap, partials, cg);
}

template <typename T, bool DOT, bool NARROW>
static void gb_spmv_go_vd(const gbd_matrix *m, uint32_t grid, hipStream_t s, const T *val,
                          const uint8_t *vidx, const T *vtab, const T *x, T *y,
                          uint32_t ystride, double *partials, const gbd_cg_state *cg) {
  if (vidx)
    gb_spmv_go<T, DOT, NARROW, true>(m, grid, s, val, vidx, vtab, x, y, ystride, partials, cg);
  else
    gb_spmv_go<T, DOT, NARROW, false>(m, grid, s, val, vidx, vtab, x, y, ystride, partials, cg);
}

template <typename T>
static int gb_spmv_launch_t(const gbd_matrix *m, const T *val, const uint8_t *vidx,
                            const T *vtab, const T *x, T *y, uint32_t ystride,
                            double *partials, const gbd_cg_state *cg) {
  if (m->nsplit) return gb_spmv_launch_split<T>(m, x, y, ystride, partials, cg);
  if (m->nparts) return gb_spmv_launch_parts<T>(m, x, y, ystride, partials, cg);
  if (m->xt_ptr) {
    // x tiles: the FP64 values, or their FP32 copy in the layout's order
    // (made with the matrix's FP32 copy; CG runs in FP64 only)
    const T *xv;
    if constexpr (sizeof(T) == 8) xv = m->xt_val;
    else xv = cg ? nullptr : m->xt_val32;
    if (xv) {
      hipStream_t s = gb_stream();
      gb_spmv_xtile<T, GB_XT_STORE><<<m->xt_W, GB_XT_THREADS, 0, s>>>(
          m->xt_ptr, m->xt_idx, xv, x, y, ystride, nullptr, m->nrows, m->ncols, m->xt_R,
          m->xt_T, m->xt_lc);
      GB_LAUNCH_CHECK();
      if constexpr (sizeof(T) == 8) {
        if (cg) {
          gb_dot_partials<<<gb_split_sum_grid(m), GB_BLOCK, 0, s>>>(x, y, m->nrows, partials);
          GB_LAUNCH_CHECK();
        }
      }
      return GOOSEBERRY_SUCCESS;
    }
  }
  const uint32_t grid = gb_spmv_grid(m);
  if (grid == 0) return GOOSEBERRY_SUCCESS;
  hipStream_t s = gb_stream();
  if (!vtab) vidx = nullptr;
  if (cg && m->dcol)
    gb_spmv_go_vd<T, true, true>(m, grid, s, val, vidx, vtab, x, y, ystride, partials, cg);
  else if (cg)
    gb_spmv_go_vd<T, true, false>(m, grid, s, val, vidx, vtab, x, y, ystride, partials, cg);
  else if (m->dcol)
    gb_spmv_go_vd<T, false, true>(m, grid, s, val, vidx, vtab, x, y, ystride, nullptr, nullptr);
  else
    gb_spmv_go_vd<T, false, false>(m, grid, s, val, vidx, vtab, x, y, ystride, nullptr, nullptr);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gb_spmv_launch(const gbd_matrix *m, const double *x, double *y,
                   uint32_t ystride, double *partials, const gbd_cg_state *cg) {
  return gb_spmv_launch_t<double>(m, m->val, m->vidx, m->vtab, x, y, ystride, partials, cg);
}

int gb_spmv_launch_f32(const gbd_matrix *m, const float *x, float *y,
                       double *partials, const gbd_cg_state *cg) {
  if (!m->val32) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gb_spmv_launch_t<float>(m, m->val32, m->vidx, m->vtab32, x, y, 1, partials, cg);
}

__global__ void gb_to_f32(const double *__restrict__ in, float *__restrict__ out,
                          size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// Jacobi scaling, one wavefront per row: s_i = 1 / sqrt(a_ii), a_ii the sum
// of the row's entries in column i (added in a fixed order); *bad is set
// when a diagonal is not positive and finite.
__global__ void __launch_bounds__(GB_BLOCK)
gb_jacobi_scale(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
                const double *__restrict__ val, uint32_t n, double *__restrict__ scale,
                uint32_t *__restrict__ bad) {
  const uint32_t lane = threadIdx.x & (GB_WAVE - 1);
  const uint32_t nw = gridDim.x * GB_WAVES_PER_BLOCK;
  uint32_t b = 0;
  for (uint32_t i = blockIdx.x * GB_WAVES_PER_BLOCK + threadIdx.x / GB_WAVE; i < n; i += nw) {
    const uint32_t k0 = rowptr[i], k1 = rowptr[i + 1];
    double d = 0.0;
    for (uint32_t k = k0 + lane; k < k1; k += GB_WAVE)
      if (rowind[k] == i) d += val[k];
    d = gb_wave_sum(d);
    if (lane == 0) {
      if (d > 0.0 && isfinite(d)) scale[i] = 1.0 / sqrt(d);
      else { scale[i] = 1.0; b = 1; }
    }
  }
  if (b) atomicOr(bad, 1u);
}

// FP32 values of D^-1/2 A D^-1/2: (a_ij s_i) s_j rounded once to FP32, one
// wavefront per row
__global__ void __launch_bounds__(GB_BLOCK)
gb_jacobi_vals(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
               const double *__restrict__ val, uint32_t n, const double *__restrict__ scale,
               float *__restrict__ out) {
  const uint32_t lane = threadIdx.x & (GB_WAVE - 1);
  const uint32_t nw = gridDim.x * GB_WAVES_PER_BLOCK;
  for (uint32_t i = blockIdx.x * GB_WAVES_PER_BLOCK + threadIdx.x / GB_WAVE; i < n; i += nw) {
    const uint32_t k0 = rowptr[i], k1 = rowptr[i + 1];
    const double si = scale[i];
    for (uint32_t k = k0 + lane; k < k1; k += GB_WAVE)
      out[k] = (float)((val[k] * si) * scale[rowind[k]]);
  }
}

// ------------------------------------------------- value-indexed plans
// The distinct values of an array (64-bit patterns of FP64 values, or
// 32-bit patterns of FP32 values zero-extended) are collected in an
// open-addressing table: each workgroup first deduplicates its chunk in
// LDS and merges only its distinct keys into the global table; more than
// GB_VDICT_MAX distinct keys in a chunk, or the all-ones pattern (the empty
// mark), sets *over and the plan keeps plain values.
#define GB_VD_GTAB 1024
#define GB_VD_LTAB 512
#define GB_VD_EMPTY 0xffffffffffffffffull
#define GB_VD_CHUNK 65536

__device__ __forceinline__ uint32_t gb_vd_hash(unsigned long long k) {
  // the 64-bit finaliser of MurmurHash3
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

template <typename K>
__global__ void __launch_bounds__(GB_BLOCK)
gb_vd_collect(const K *__restrict__ vals, uint64_t nnz, unsigned long long *gkeys,
              uint32_t *over) {
  __shared__ unsigned long long lk[GB_VD_LTAB];
  __shared__ uint32_t lcount;
  __shared__ volatile uint32_t lover;
  for (uint32_t i = threadIdx.x; i < GB_VD_LTAB; i += GB_BLOCK) lk[i] = GB_VD_EMPTY;
  if (threadIdx.x == 0) {
    lcount = 0;
    lover = __hip_atomic_load(over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (lover) return;
  const uint64_t b = (uint64_t)blockIdx.x * GB_VD_CHUNK;
  const uint64_t e = nnz - b < GB_VD_CHUNK ? nnz : b + GB_VD_CHUNK;
  for (uint64_t k = b + threadIdx.x; k < e; k += GB_BLOCK) {
    const unsigned long long key = (unsigned long long)vals[k];
    if (key == GB_VD_EMPTY) {
      lover = 1;
      break;
    }
    uint32_t h = gb_vd_hash(key) & (GB_VD_LTAB - 1);
    bool placed = false;
    for (uint32_t p = 0; p < GB_VD_LTAB && !placed; ++p) {
      const unsigned long long cur = ((volatile unsigned long long *)lk)[h];
      if (cur == key) {
        placed = true;
      } else if (cur == GB_VD_EMPTY) {
        const unsigned long long prev = atomicCAS(&lk[h], GB_VD_EMPTY, key);
        if (prev == GB_VD_EMPTY) {
          if (atomicAdd(&lcount, 1u) >= GB_VDICT_MAX) lover = 1;
          placed = true;
        } else if (prev == key) {
          placed = true;
        }
      }
      if (!placed) h = (h + 1) & (GB_VD_LTAB - 1);
    }
    if (!placed) lover = 1;
    if (lover) break;
  }
  __syncthreads();
  if (lover) {
    if (threadIdx.x == 0) atomicOr(over, 1u);
    return;
  }
  for (uint32_t i = threadIdx.x; i < GB_VD_LTAB; i += GB_BLOCK) {
    const unsigned long long key = lk[i];
    if (key == GB_VD_EMPTY) continue;
    uint32_t h = gb_vd_hash(key) & (GB_VD_GTAB - 1);
    bool placed = false;
    for (uint32_t p = 0; p < GB_VD_GTAB && !placed; ++p) {
      const unsigned long long prev = atomicCAS(&gkeys[h], GB_VD_EMPTY, key);
      if (prev == GB_VD_EMPTY || prev == key) placed = true;
      else h = (h + 1) & (GB_VD_GTAB - 1);
    }
    if (!placed) atomicOr(over, 1u);
  }
}

// idx[k] = the table position of vals[k] (every value is in the table)
template <typename K>
__global__ void __launch_bounds__(GB_BLOCK)
gb_vd_map(const K *__restrict__ vals, uint64_t nnz, const unsigned long long *__restrict__ gkeys,
          const uint8_t *__restrict__ slot2id, uint8_t *__restrict__ idx) {
  __shared__ unsigned long long lk[GB_VD_GTAB];
  __shared__ uint8_t lid[GB_VD_GTAB];
  for (uint32_t i = threadIdx.x; i < GB_VD_GTAB; i += GB_BLOCK) {
    lk[i] = gkeys[i];
    lid[i] = slot2id[i];
  }
  __syncthreads();
  for (uint64_t k = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; k < nnz;
       k += (uint64_t)gridDim.x * GB_BLOCK) {
    const unsigned long long key = (unsigned long long)vals[k];
    uint32_t h = gb_vd_hash(key) & (GB_VD_GTAB - 1);
    for (uint32_t p = 0; p < GB_VD_GTAB && lk[h] != key; ++p) h = (h + 1) & (GB_VD_GTAB - 1);
    idx[k] = lid[h];
  }
}

// An 8-bit index per entry and the table of the distinct values of vals
// (width 8: FP64, 4: FP32; the table ordered by bit pattern, so the arrays
// are the same in every run).  *built = false, and nothing allocated, when
// there are more than GB_VDICT_MAX distinct values.
static int gb_vdict_build(const void *vals, int width, uint64_t nnz, uint8_t **r_idx,
                          void **r_tab, uint32_t *r_nvals, bool *built) {
  hipStream_t s = gb_stream();
  *built = false;
  *r_idx = nullptr;
  *r_tab = nullptr;
  if (nnz == 0) return GOOSEBERRY_SUCCESS;
  unsigned long long *gkeys = nullptr;
  uint32_t *over = nullptr;
  GB_HIP(gb_dev_alloc((void **)&gkeys, GB_VD_GTAB * 8 + 256));
  over = (uint32_t *)(gkeys + GB_VD_GTAB);
  GB_HIP(hipMemsetAsync(gkeys, 0xff, GB_VD_GTAB * 8, s));
  GB_HIP(hipMemsetAsync(over, 0, 4, s));
  const uint64_t nwg = (nnz + GB_VD_CHUNK - 1) / GB_VD_CHUNK;
  if (nwg > 0x7fffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (width == 8)
    gb_vd_collect<uint64_t><<<(uint32_t)nwg, GB_BLOCK, 0, s>>>((const uint64_t *)vals, nnz,
                                                               gkeys, over);
  else
    gb_vd_collect<uint32_t><<<(uint32_t)nwg, GB_BLOCK, 0, s>>>((const uint32_t *)vals, nnz,
                                                               gkeys, over);
  GB_LAUNCH_CHECK();
  unsigned long long hk[GB_VD_GTAB];
  uint32_t hover = 0;
  GB_HIP(hipMemcpyAsync(hk, gkeys, sizeof(hk), hipMemcpyDeviceToHost, s));
  GB_HIP(hipMemcpyAsync(&hover, over, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  unsigned long long keys[GB_VDICT_MAX];
  uint32_t nk = 0;
  for (uint32_t i = 0; i < GB_VD_GTAB && !hover; ++i) {
    if (hk[i] == GB_VD_EMPTY) continue;
    if (nk == GB_VDICT_MAX) hover = 1;
    else keys[nk++] = hk[i];
  }
  if (hover) {
    GB_HIP(gb_dev_free(gkeys));
    return GOOSEBERRY_SUCCESS;
  }
  // table order: ascending bit pattern
  for (uint32_t i = 1; i < nk; ++i)
    for (uint32_t j = i; j > 0 && keys[j - 1] > keys[j]; --j) {
      const unsigned long long t = keys[j];
      keys[j] = keys[j - 1];
      keys[j - 1] = t;
    }
  uint8_t hid[GB_VD_GTAB];
  memset(hid, 0, sizeof(hid));
  for (uint32_t i = 0; i < GB_VD_GTAB; ++i) {
    if (hk[i] == GB_VD_EMPTY) continue;
    uint32_t lo = 0, hi = nk;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (keys[mid] < hk[i]) lo = mid + 1; else hi = mid;
    }
    hid[i] = (uint8_t)lo;
  }
  unsigned char htab[GB_VDICT_MAX * 8];
  memset(htab, 0, sizeof(htab));
  for (uint32_t i = 0; i < nk; ++i) {
    if (width == 8) {
      const uint64_t b = keys[i];
      memcpy(htab + 8 * i, &b, 8);
    } else {
      const uint32_t b = (uint32_t)keys[i];
      memcpy(htab + 4 * i, &b, 4);
    }
  }
  uint8_t *slot2id = nullptr, *idx = nullptr;
  void *tab = nullptr;
  GB_HIP(gb_dev_alloc((void **)&slot2id, GB_VD_GTAB));
  GB_HIP(gb_dev_alloc(&tab, GB_VDICT_MAX * (size_t)width));
  GB_HIP(gb_dev_alloc((void **)&idx, nnz));
  GB_HIP(hipMemcpyAsync(slot2id, hid, GB_VD_GTAB, hipMemcpyHostToDevice, s));
  GB_HIP(hipMemcpyAsync(tab, htab, GB_VDICT_MAX * (size_t)width, hipMemcpyHostToDevice, s));
  const uint64_t g = (nnz + GB_BLOCK - 1) / GB_BLOCK;
  const uint32_t grid = g > 16384 ? 16384u : (uint32_t)g;
  if (width == 8)
    gb_vd_map<uint64_t><<<grid, GB_BLOCK, 0, s>>>((const uint64_t *)vals, nnz, gkeys, slot2id, idx);
  else
    gb_vd_map<uint32_t><<<grid, GB_BLOCK, 0, s>>>((const uint32_t *)vals, nnz, gkeys, slot2id, idx);
  GB_LAUNCH_CHECK();
  // the host staging (hid, htab) must outlive the copies
  GB_HIP(hipStreamSynchronize(s));
  GB_HIP(gb_dev_free(slot2id));
  GB_HIP(gb_dev_free(gkeys));
  *r_idx = idx;
  *r_tab = tab;
  *r_nvals = nk;
  *built = true;
  return GOOSEBERRY_SUCCESS;
}

// vtab32 = the FP32 roundings of vtab (the FP32 copy of a value-indexed
// matrix shares its indices)
static int gb_vdict_f32(gbd_matrix *m) {
  if (!m->vidx || m->vtab32) return GOOSEBERRY_SUCCESS;
  hipStream_t s = gb_stream();
  double h64[GB_VDICT_MAX];
  float h32[GB_VDICT_MAX];
  GB_HIP(hipMemcpyAsync(h64, m->vtab, sizeof(h64), hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < GB_VDICT_MAX; ++i) h32[i] = (float)h64[i];
  GB_HIP(gb_dev_alloc((void **)&m->vtab32, sizeof(h32)));
  GB_HIP(hipMemcpyAsync(m->vtab32, h32, sizeof(h32), hipMemcpyHostToDevice, s));
  GB_HIP(hipStreamSynchronize(s));
  return GOOSEBERRY_SUCCESS;
}

static void gb_vdict_free(gbd_matrix *m) {
  (void)gb_dev_free(m->vidx);
  (void)gb_dev_free(m->vtab);
  (void)gb_dev_free(m->vtab32);
  m->vidx = nullptr;
  m->vtab = nullptr;
  m->vtab32 = nullptr;
  m->nvals = 0;
}

// ------------------------------------------------------------- host side

static void gb_xtile_free(gbd_matrix *m) {
  (void)gb_dev_free(m->xt_ptr);
  (void)gb_dev_free(m->xt_idx);
  (void)gb_dev_free(m->xt_val);
  (void)gb_dev_free(m->xt_val32);
  (void)gb_dev_free(m->xt_val32j);
  m->xt_ptr = m->xt_idx = nullptr;
  m->xt_val = nullptr;
  m->xt_val32 = m->xt_val32j = nullptr;
  m->xt_R = m->xt_W = m->xt_T = m->xt_lc = 0;
}

// The Jacobi-scaled FP32 values in the layout's order: the same
// (a_ij s_i) s_j rounded once as the CSR-order copy, the row and column
// taken from the group and the packed word
__global__ void __launch_bounds__(GB_BLOCK)
gb_xt_jacobi_vals(const uint32_t *__restrict__ ptr, uint32_t ngroups, uint32_t T, uint32_t R,
                  uint32_t lc, const uint32_t *__restrict__ word,
                  const double *__restrict__ val, const double *__restrict__ scale,
                  float *__restrict__ out) {
  const uint32_t cm = (1u << lc) - 1;
  for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const uint32_t rb = (g / T) * R, c0 = (g % T) << lc, e = ptr[g + 1];
    for (uint32_t j = ptr[g] + threadIdx.x; j < e; j += GB_BLOCK) {
      const uint32_t w = word[j];
      out[j] = (float)((val[j] * scale[rb + (w >> lc)]) * scale[c0 | (w & cm)]);
    }
  }
}

// The layout's FP32 copies of the values the matrix has in FP32 (val32,
// val32j), made with them or with the layout, whichever comes second
static int gb_xtile_f32(gbd_matrix *m) {
  if (!m->xt_ptr) return GOOSEBERRY_SUCCESS;
  hipStream_t s = gb_stream();
  const size_t nz = m->nnz;
  if (m->val32 && !m->xt_val32) {
    GB_HIP(gb_dev_alloc((void **)&m->xt_val32, (nz + GB_XT_PAD) * sizeof(float)));
    GB_HIP(hipMemsetAsync(m->xt_val32 + nz, 0, GB_XT_PAD * sizeof(float), s));
    if (nz) {
      gb_to_f32<<<gb_div_up(nz, GB_BLOCK), GB_BLOCK, 0, s>>>(m->xt_val, m->xt_val32, nz);
      GB_LAUNCH_CHECK();
    }
  }
  if (m->val32j && m->jscale && !m->xt_val32j) {
    GB_HIP(gb_dev_alloc((void **)&m->xt_val32j, (nz + GB_XT_PAD) * sizeof(float)));
    GB_HIP(hipMemsetAsync(m->xt_val32j + nz, 0, GB_XT_PAD * sizeof(float), s));
    const uint32_t ng = m->xt_W * m->xt_T;
    if (nz && ng) {
      gb_xt_jacobi_vals<<<ng < 65536u ? ng : 65536u, GB_BLOCK, 0, s>>>(
          m->xt_ptr, ng, m->xt_T, m->xt_R, m->xt_lc, m->xt_idx, m->xt_val, m->jscale,
          m->xt_val32j);
      GB_LAUNCH_CHECK();
    }
  }
  return GOOSEBERRY_SUCCESS;
}

// The x-tile layout: keys = (row block, column tile) per entry, a stable
// sort of the entries by key (each group keeps CSR order), then the packed
// words and values gathered in that order.  Row blocks of
// ceil(nrows / CUs) rows (at most GB_XT_RMAX): one workgroup per CU.
static int gb_xtile_build(gbd_matrix *m) {
  hipStream_t s = gb_stream();
  const uint32_t n = m->nrows, C = m->params.xtile;
  uint32_t lc = 0;
  while ((1u << lc) < C) ++lc;
  uint32_t R = gb_div_up(n, gb_num_cus());
  if (R > GB_XT_RMAX) R = GB_XT_RMAX;
  if (R == 0) R = 1;
  const uint32_t W = gb_div_up(n, R);
  const uint32_t T = gb_div_up(m->ncols ? m->ncols : 1, C);
  const uint64_t groups = (uint64_t)W * T;
  if (groups >= 0xffffffffull || m->nnz + GB_XT_THREADS >= 0xffffffffull) {
    gb_set_error("x tiles of %u columns: too many groups (%llu) or entries", C,
                 (unsigned long long)groups);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  uint32_t *keys = nullptr, *word = nullptr, *perm = nullptr;
  const size_t nz = m->nnz;
  hipError_t e = gb_dev_alloc((void **)&keys, nz * 4);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&word, nz * 4);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&perm, nz * 4);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&m->xt_ptr, (groups + 1) * 4);
  // zero entries past the end: the SpMV reads the layout E entries a lane
  if (e == hipSuccess) e = gb_dev_alloc((void **)&m->xt_idx, (nz + GB_XT_PAD) * 4);
  if (e == hipSuccess) e = gb_dev_alloc((void **)&m->xt_val, (nz + GB_XT_PAD) * 8);
  if (e == hipSuccess) e = hipMemsetAsync(m->xt_idx + nz, 0, GB_XT_PAD * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(m->xt_val + nz, 0, GB_XT_PAD * 8, s);
  int rv = e == hipSuccess ? GOOSEBERRY_SUCCESS : GOOSEBERRY_ERROR_NOMEMORY;
  if (rv == GOOSEBERRY_SUCCESS) {
    gb_xt_keys<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, n, R, T, lc,
                                                           keys, word);
    if (gb_launch_status() != hipSuccess) rv = GOOSEBERRY_ERROR_DEVICE;
  }
  if (rv == GOOSEBERRY_SUCCESS)
    rv = gbd_countingsort_kv_unchecked(keys, nullptr, 0, (uint32_t)(groups - 1), nz, perm,
                                       m->xt_ptr);
  if (rv == GOOSEBERRY_SUCCESS) {
    const uint64_t g = gb_div_up(nz, GB_BLOCK);
    gb_xt_gather<<<g > 65536 ? 65536u : (uint32_t)g, GB_BLOCK, 0, s>>>(perm, nz, word, m->val,
                                                                      m->xt_idx, m->xt_val);
    if (gb_launch_status() != hipSuccess) rv = GOOSEBERRY_ERROR_DEVICE;
  }
  (void)gb_dev_free(keys);
  (void)gb_dev_free(word);
  (void)gb_dev_free(perm);
  if (rv != GOOSEBERRY_SUCCESS) {
    gb_xtile_free(m);
    return rv;
  }
  m->xt_R = R;
  m->xt_W = W;
  m->xt_T = T;
  m->xt_lc = lc;
  rv = gb_xtile_f32(m);
  if (rv != GOOSEBERRY_SUCCESS) gb_xtile_free(m);
  return rv;
}

static int gb_plan_check(const gbd_plan_params *p) {
  if (p->short_max == 0 || p->window == 0 || p->rowcap == 0) return 0;
  if ((uint64_t)p->window + p->short_max > GB_STREAM_LDS) return 0;
  if (p->wave_max < p->short_max) return 0;
  // the row-block kernel gives each thread at most two rows
  if (p->rowcap > 2 * GB_BLOCK) return 0;
  if (p->grid > 65536) return 0;
  if (p->csplit > GB_SPLIT_MAX) return 0;
  if (p->xcd_remap > 4096) return 0;
  // slices change the summation order: not with the bitwise plan
  if (p->csplit > 1 && p->exact) return 0;
  // x tiles: a power of two of 256..GB_XT_CMAX columns, alone
  if (p->xtile) {
    if (p->xtile < 256 || p->xtile > GB_XT_CMAX || (p->xtile & (p->xtile - 1))) return 0;
    if (p->csplit > 1 || p->exact) return 0;
  }
  return 1;
}

static void gb_split_free(gbd_matrix *m) {
  for (uint32_t r = 0; r < m->nsplit; ++r) gbd_matrix_destroy(m->split[r]);
  free(m->split);
  m->split = nullptr;
  m->nsplit = 0;
  (void)gb_dev_free(m->srowptr);
  (void)gb_dev_free(m->srowind);
  (void)gb_dev_free(m->sval);
  (void)gb_dev_free(m->sval32);
  (void)gb_dev_free(m->sval32j);
  (void)gb_dev_free(m->ypart);
  m->srowptr = m->srowind = nullptr;
  m->sval = m->ypart = nullptr;
  m->sval32 = m->sval32j = nullptr;
}

static int gb_plan_build(gbd_matrix *m);

// P column slices of width ceil(ncols / P) as views with their own plans
static int gb_split_build(gbd_matrix *m, uint32_t P) {
  hipStream_t s = gb_stream();
  const uint32_t n = m->nrows;
  const uint32_t width = (m->ncols + P - 1) / P;
  const uint64_t tot = (uint64_t)P * n;
  if (tot >= 0xffffffffull || width == 0) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t *cnt = nullptr;
  uint32_t bounds[GB_SPLIT_MAX + 1];
  GB_HIP(gb_dev_alloc((void **)&m->srowptr, (tot + 1) * 4));
  GB_HIP(gb_dev_alloc((void **)&m->srowind, (m->nnz ? m->nnz : 1) * 4));
  GB_HIP(gb_dev_alloc((void **)&m->sval, (m->nnz ? m->nnz : 1) * 8));
  GB_HIP(gb_dev_alloc((void **)&m->ypart, tot * 8));
  GB_HIP(gb_dev_alloc((void **)&cnt, tot * 4));
  const uint32_t g = gb_div_up(n, GB_BLOCK);
  gb_split_count<<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, n, width, P, cnt);
  GB_LAUNCH_CHECK();
  int rv = gbd_scan_u32(cnt, m->srowptr, tot, m->srowptr + tot);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  GB_HIP(gb_dev_free(cnt));
  gb_split_fill<double><<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->val, n, width, P,
                                               m->srowptr, m->srowind, m->sval);
  GB_LAUNCH_CHECK();
  if (m->val32) {
    GB_HIP(gb_dev_alloc((void **)&m->sval32, (m->nnz ? m->nnz : 1) * 4));
    gb_split_fill<float><<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->val32, n, width, P,
                                                m->srowptr, nullptr, m->sval32);
    GB_LAUNCH_CHECK();
  }
  if (m->val32j) {
    GB_HIP(gb_dev_alloc((void **)&m->sval32j, (m->nnz ? m->nnz : 1) * 4));
    gb_split_fill<float><<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->val32j, n, width, P,
                                                m->srowptr, nullptr, m->sval32j);
    GB_LAUNCH_CHECK();
  }
  // slice boundaries for the views' nonzero counts
  for (uint32_t r = 0; r <= P; ++r)
    GB_HIP(hipMemcpyAsync(&bounds[r], m->srowptr + (size_t)r * n, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  m->split = (gbd_matrix **)calloc(P, sizeof(gbd_matrix *));
  if (!m->split) return GOOSEBERRY_ERROR_NOMEMORY;
  m->nsplit = P;
  for (uint32_t r = 0; r < P; ++r) {
    gbd_matrix *v = (gbd_matrix *)calloc(1, sizeof(gbd_matrix));
    if (!v) return GOOSEBERRY_ERROR_NOMEMORY;
    m->split[r] = v;
    v->view = true;
    v->nrows = n;
    v->ncols = m->ncols;
    v->nnz = bounds[r + 1] - bounds[r];
    v->rowptr = m->srowptr + (size_t)r * n;
    v->rowind = m->srowind;
    v->val = m->sval;
    v->val32 = m->sval32;
    v->val32j = m->sval32j;
    v->params = m->params;
    v->params.csplit = 0;
    if ((rv = gb_plan_build(v)) != GOOSEBERRY_SUCCESS) return rv;
  }
  return GOOSEBERRY_SUCCESS;
}

// the plans of a row-parted matrix: one per part; the parent's histogram
// and bin counts are the parts' sums
static int gb_parts_plan(gbd_matrix *m) {
  if (m->params.csplit > 1 || m->params.xtile) {
    gb_set_error("a matrix held as row parts (64-bit row pointers) takes no column split "
                 "and no x tiles");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  for (uint32_t q = 0; q < m->nparts; ++q) {
    gbd_matrix *v = m->parts[q];
    v->params = m->params;
    int rv = gb_plan_build(v);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    for (int i = 0; i < 33; ++i) m->hist[i] += v->hist[i];
    m->nstream += v->nstream;
    m->nwave += v->nwave;
    m->nblockrows += v->nblockrows;
  }
  return GOOSEBERRY_SUCCESS;
}

static int gb_plan_build(gbd_matrix *m) {
  hipStream_t s = gb_stream();
  const uint32_t n = m->nrows;
  if (!m->view) gb_split_free(m);
  gb_vdict_free(m);
  if (!m->view) gb_xtile_free(m);
  m->pipe_grid = 0;
  memset(m->hist, 0, sizeof(m->hist));
  m->nstream = m->nwave = m->nblockrows = 0;
  if (m->nparts) return gb_parts_plan(m);
  if (m->blk) { GB_HIP(gb_dev_free(m->blk)); m->blk = nullptr; }
  if (m->bdesc) { GB_HIP(gb_dev_free(m->bdesc)); m->bdesc = nullptr; }
  if (m->rowbins) { GB_HIP(gb_dev_free(m->rowbins)); m->rowbins = nullptr; }
  if (m->dcol) { GB_HIP(gb_dev_free(m->dcol)); m->dcol = nullptr; }
  if (n == 0) return GOOSEBERRY_SUCCESS;

  uint32_t *cls = nullptr, *flag = nullptr, *idx = nullptr, *dist = nullptr;
  unsigned long long *hist = nullptr;
  GB_HIP(gb_dev_alloc((void **)&cls, (size_t)n * 4));
  GB_HIP(gb_dev_alloc((void **)&flag, (size_t)n * 4));
  GB_HIP(gb_dev_alloc((void **)&idx, ((size_t)n + 1) * 4));
  GB_HIP(gb_dev_alloc((void **)&dist, 4 * 4));
  GB_HIP(gb_dev_alloc((void **)&hist, 34 * 8));
  GB_HIP(hipMemsetAsync(hist, 0, 34 * 8, s));
  gb_plan_classify<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(
      m->rowptr, m->rowind, n, m->cshift, m->params, cls, flag, hist);
  GB_LAUNCH_CHECK();
  GB_HIP(gb_dev_alloc((void **)&m->rowbins, (size_t)n * 4));
  int rv = gbd_countingsort_kv_unchecked(cls, nullptr, 0, 2, n, m->rowbins, dist);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  rv = gbd_scan_u32(flag, idx, n, idx + n);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  uint32_t hd[4], nst = 0;
  unsigned long long hh[34];
  GB_HIP(hipMemcpyAsync(hd, dist, sizeof(hd), hipMemcpyDeviceToHost, s));
  GB_HIP(hipMemcpyAsync(&nst, idx + n, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 33; ++i) m->hist[i] = hh[i];
  m->nstream = nst;
  m->nwave = hd[2] - hd[1];
  m->nblockrows = hd[3] - hd[2];
  if (nst) {
    GB_HIP(gb_dev_alloc((void **)&m->blk, (size_t)nst * 2 * 4));
    gb_plan_blocks<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(cls, flag, idx,
                                                              n, m->blk);
    GB_LAUNCH_CHECK();
    GB_HIP(gb_dev_alloc((void **)&m->bdesc, (size_t)nst * sizeof(uint4)));
    gb_plan_desc<<<gb_div_up(nst, GB_BLOCK), GB_BLOCK, 0, s>>>(m->rowptr, m->blk, nst,
                                                              m->bdesc);
    GB_LAUNCH_CHECK();
  }
  GB_HIP(gb_dev_free(cls));
  GB_HIP(gb_dev_free(flag));
  GB_HIP(gb_dev_free(idx));
  GB_HIP(gb_dev_free(dist));
  GB_HIP(gb_dev_free(hist));
  // hh[33]: rows whose columns already rule the 16-bit offsets out
  // (a column slice's entries are addressed from the whole split array, so
  // slices take neither the offsets nor the value index; row parts do)
  const bool own = !m->view || m->part;
  if (own && !m->params.wide_index && nst && m->nnz && hh[33] == 0) {
    // 16-bit column offsets for the row blocks, kept only if all fit
    uint32_t *bad = nullptr, hbad = 0;
    GB_HIP(gb_dev_alloc((void **)&m->dcol, m->nnz * sizeof(int16_t)));
    GB_HIP(gb_dev_alloc((void **)&bad, 4));
    GB_HIP(hipMemsetAsync(bad, 0, 4, s));
    gb_plan_narrow<<<nst, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->blk, m->cshift, m->dcol,
                                            bad);
    GB_LAUNCH_CHECK();
    GB_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    GB_HIP(hipStreamSynchronize(s));
    GB_HIP(gb_dev_free(bad));
    if (hbad) {
      GB_HIP(gb_dev_free(m->dcol));
      m->dcol = nullptr;
    }
  }
  if (!m->view && m->params.csplit > 1) {
    int rv2 = gb_split_build(m, m->params.csplit);
    if (rv2 != GOOSEBERRY_SUCCESS) {
      gb_split_free(m);
      return rv2;
    }
  }
  if (own && !m->params.wide_values && !m->nsplit && m->nnz) {
    // 8-bit value indices when the matrix has few distinct values
    bool built = false;
    void *tab = nullptr;
    int rv3 = gb_vdict_build(m->val, 8, m->nnz, &m->vidx, &tab, &m->nvals, &built);
    m->vtab = (double *)tab;
    if (rv3 == GOOSEBERRY_SUCCESS && m->val32) rv3 = gb_vdict_f32(m);
    if (rv3 != GOOSEBERRY_SUCCESS) {
      gb_vdict_free(m);
      return rv3;
    }
  }
  if (!m->view && m->params.xtile && !m->nsplit && m->nnz) {
    int rv4 = gb_xtile_build(m);
    if (rv4 != GOOSEBERRY_SUCCESS) return rv4;
  }
  gb_pipe_choose(m);
  return GOOSEBERRY_SUCCESS;
}

extern "C" {

void gbd_plan_defaults(const uint64_t *hist, uint64_t nrows, uint64_t nnz,
                       gbd_plan_params *p) {
  (void)hist;
  (void)nrows;
  (void)nnz;
  p->short_max = 64;
  p->wave_max = 4096;
  p->window = 2048;
  p->rowcap = 512;
  p->exact = 0;
  p->grid = 0;
  p->xcd_remap = 0;
  p->csplit = 0;
  p->wide_index = 0;
  p->wide_values = 0;
  p->xtile = 0;
}

int gbd_csr_check(uint32_t nrows, uint32_t ncols, const uint32_t *d_rowptr,
                  const uint32_t *d_rowind, uint32_t *r_nnz) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipStream_t s = gb_stream();
  // nnz from the device, then every row pointer and column index checked
  uint32_t *bad = nullptr, hbad = 0, nnz = 0;
  GB_HIP(hipMemcpyAsync(&nnz, d_rowptr + nrows, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  GB_HIP(gb_dev_alloc((void **)&bad, 4));
  GB_HIP(hipMemsetAsync(bad, 0, 4, s));
  uint64_t work = (uint64_t)nrows > nnz ? nrows : nnz;
  uint32_t g = gb_div_up(work ? work : 1, GB_BLOCK);
  gb_validate_csr<<<g > 8192u ? 8192u : g, GB_BLOCK, 0, s>>>(
      d_rowptr, d_rowind, nrows, ncols, nnz, bad);
  GB_LAUNCH_CHECK();
  GB_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(gb_dev_free(bad));
  GB_HIP(hipStreamSynchronize(s));
  if (hbad) return GOOSEBERRY_ERROR_INVALIDINPUT;
  *r_nnz = nnz;
  return GOOSEBERRY_SUCCESS;
}

static int gb_matrix_create(uint32_t nrows, uint32_t ncols, uint32_t *d_rowptr,
                            uint32_t *d_rowind, double *d_rowval, bool trusted,
                            uint32_t nnz, const gbd_plan_params *params,
                            gbd_matrix **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipStream_t s = gb_stream();
  (void)s;
  if (!trusted) {
    rv = gbd_csr_check(nrows, ncols, d_rowptr, d_rowind, &nnz);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  gbd_plan_params p;
  if (params) p = *params;
  else gbd_plan_defaults(nullptr, nrows, nnz, &p);
  if (!gb_plan_check(&p)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gbd_matrix *m = (gbd_matrix *)calloc(1, sizeof(gbd_matrix));
  if (!m) return GOOSEBERRY_ERROR_NOMEMORY;
  m->nrows = nrows;
  m->ncols = ncols;
  m->rowptr = d_rowptr;
  m->rowind = d_rowind;
  m->val = d_rowval;
  m->nnz = nnz;
  m->params = p;
  rv = gb_plan_build(m);
  if (rv != GOOSEBERRY_SUCCESS) {
    // the caller keeps the arrays it passed in
    m->rowptr = nullptr; m->rowind = nullptr; m->val = nullptr;
    gbd_matrix_destroy(m);
    return rv;
  }
  *out = m;
  return GOOSEBERRY_SUCCESS;
}

int gbd_matrix_create(uint32_t nrows, uint32_t ncols, uint32_t *d_rowptr,
                      uint32_t *d_rowind, double *d_rowval,
                      const gbd_plan_params *params, gbd_matrix **out) {
  return gb_matrix_create(nrows, ncols, d_rowptr, d_rowind, d_rowval, false, 0,
                          params, out);
}

int gbd_matrix_create_trusted(uint32_t nrows, uint32_t ncols, uint32_t *d_rowptr,
                              uint32_t *d_rowind, double *d_rowval, uint32_t nnz,
                              const gbd_plan_params *params, gbd_matrix **out) {
  return gb_matrix_create(nrows, ncols, d_rowptr, d_rowind, d_rowval, true, nnz,
                          params, out);
}

// One 64-bit row pointer read back (the part boundaries' binary search).
static int gb_rowptr64_at(const uint64_t *d_rowptr64, uint32_t i, uint64_t *v) {
  hipStream_t s = gb_stream();
  GB_HIP(hipMemcpyAsync(v, d_rowptr64 + i, 8, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  return GOOSEBERRY_SUCCESS;
}

int gbd_matrix_create64(uint32_t nrows, uint32_t ncols, uint64_t *d_rowptr64,
                        uint32_t *d_rowind, double *d_rowval, uint64_t part_nnz,
                        const gbd_plan_params *params, gbd_matrix **out) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipStream_t s = gb_stream();
  if (part_nnz == 0 || part_nnz > GB_PART_NNZ) part_nnz = GB_PART_NNZ;
  gbd_plan_params p;
  if (params) p = *params;
  else gbd_plan_defaults(nullptr, nrows, 0, &p);
  if (!gb_plan_check(&p)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  // structure check, as for 32-bit row pointers
  uint64_t nnz = 0;
  if ((rv = gb_rowptr64_at(d_rowptr64, nrows, &nnz)) != GOOSEBERRY_SUCCESS) return rv;
  uint32_t *bad = nullptr, hbad = 0;
  GB_HIP(gb_dev_alloc((void **)&bad, 4));
  GB_HIP(hipMemsetAsync(bad, 0, 4, s));
  const uint64_t work = (uint64_t)nrows > nnz ? nrows : nnz;
  const uint64_t gw = gb_div_up(work ? work : 1, GB_BLOCK);
  gb_validate_csr64<<<gw > 8192u ? 8192u : (uint32_t)gw, GB_BLOCK, 0, s>>>(
      d_rowptr64, d_rowind, nrows, ncols, nnz, bad);
  GB_LAUNCH_CHECK();
  GB_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(gb_dev_free(bad));
  GB_HIP(hipStreamSynchronize(s));
  if (hbad) return GOOSEBERRY_ERROR_INVALIDINPUT;
  const uint64_t gr = gb_div_up((uint64_t)nrows + 1, GB_BLOCK);
  const uint32_t grid_rows = gr > 8192u ? 8192u : (uint32_t)gr;
  if (nnz <= part_nnz) {
    // an ordinary matrix: 32-bit row pointers
    uint32_t *rp32 = nullptr;
    GB_HIP(gb_dev_alloc((void **)&rp32, ((size_t)nrows + 1) * 4));
    gb_rebase_rowptr<<<grid_rows, GB_BLOCK, 0, s>>>(d_rowptr64, 0, (uint64_t)nrows + 1, rp32);
    GB_LAUNCH_CHECK();
    rv = gb_matrix_create(nrows, ncols, rp32, d_rowind, d_rowval, true, (uint32_t)nnz, &p,
                          out);
    if (rv != GOOSEBERRY_SUCCESS) {
      (void)gb_dev_free(rp32);
      return rv;
    }
    (void)gb_dev_free(d_rowptr64);
    return GOOSEBERRY_SUCCESS;
  }
  if (p.csplit > 1) {
    gb_set_error("a matrix held as row parts (64-bit row pointers) takes no column split");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  // part boundaries: part q ends at the last row r1 with
  // rowptr[r1] - rowptr[r0] <= part_nnz; two consecutive parts hold more
  // than part_nnz entries, so there are at most 2 ceil(nnz / part_nnz)
  const uint64_t cap = 2 * gb_div_up(nnz, part_nnz) + 2;
  if (cap > 65536) {
    gb_set_error("%llu entries in parts of at most %llu make too many row parts",
                 (unsigned long long)nnz, (unsigned long long)part_nnz);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  uint32_t *prow = (uint32_t *)malloc(cap * sizeof(uint32_t));
  uint64_t *pbase = (uint64_t *)malloc(cap * sizeof(uint64_t));
  if (!prow || !pbase) {
    free(prow);
    free(pbase);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  uint32_t np = 0;
  prow[0] = 0;
  pbase[0] = 0;
  while (prow[np] < nrows) {
    const uint32_t r0 = prow[np];
    const uint64_t b0 = pbase[np];
    uint32_t lo = r0, hi = nrows;  // rowptr[lo] - b0 <= part_nnz holds for lo
    uint64_t vlo = b0;
    while (lo < hi) {
      const uint32_t mid = lo + (uint32_t)(((uint64_t)hi - lo + 1) / 2);
      uint64_t v = 0;
      if ((rv = gb_rowptr64_at(d_rowptr64, mid, &v)) != GOOSEBERRY_SUCCESS) break;
      if (v - b0 <= part_nnz) {
        lo = mid;
        vlo = v;
      } else {
        hi = mid - 1;
      }
    }
    if (rv == GOOSEBERRY_SUCCESS && lo == r0) {
      gb_set_error("row %u has more than %llu entries, the most one row part holds", r0,
                   (unsigned long long)part_nnz);
      rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    }
    if (rv != GOOSEBERRY_SUCCESS) {
      free(prow);
      free(pbase);
      return rv;
    }
    ++np;
    prow[np] = lo;
    pbase[np] = vlo;
  }
  gbd_matrix *m = (gbd_matrix *)calloc(1, sizeof(gbd_matrix));
  if (!m) {
    free(prow);
    free(pbase);
    return GOOSEBERRY_ERROR_NOMEMORY;
  }
  m->nrows = nrows;
  m->ncols = ncols;
  m->nnz = nnz;
  m->params = p;
  m->prow = prow;
  m->pbase = pbase;
  m->parts = (gbd_matrix **)calloc(np, sizeof(gbd_matrix *));
  rv = m->parts ? GOOSEBERRY_SUCCESS : GOOSEBERRY_ERROR_NOMEMORY;
  if (rv == GOOSEBERRY_SUCCESS &&
      gb_dev_alloc((void **)&m->prowptr, ((size_t)nrows + np) * 4) != hipSuccess)
    rv = GOOSEBERRY_ERROR_NOMEMORY;
  for (uint32_t q = 0; rv == GOOSEBERRY_SUCCESS && q < np; ++q) {
    const uint64_t rows = prow[q + 1] - prow[q];
    const uint64_t g = gb_div_up(rows + 1, GB_BLOCK);
    gb_rebase_rowptr<<<g > 8192u ? 8192u : (uint32_t)g, GB_BLOCK, 0, s>>>(
        d_rowptr64 + prow[q], pbase[q], rows + 1, m->prowptr + prow[q] + q);
    if (gb_launch_status() != hipSuccess) {
      rv = GOOSEBERRY_ERROR_DEVICE;
      break;
    }
    gbd_matrix *v = (gbd_matrix *)calloc(1, sizeof(gbd_matrix));
    if (!v) {
      rv = GOOSEBERRY_ERROR_NOMEMORY;
      break;
    }
    m->parts[q] = v;
    m->nparts = q + 1;
    v->view = true;
    v->part = true;
    v->nrows = (uint32_t)rows;
    v->ncols = ncols;
    v->nnz = pbase[q + 1] - pbase[q];
    v->rowptr = m->prowptr + prow[q] + q;
    v->cshift = prow[q];
    v->rowind = d_rowind + pbase[q];
    v->val = d_rowval + pbase[q];
  }
  if (rv == GOOSEBERRY_SUCCESS) {
    m->rowind = d_rowind;
    m->val = d_rowval;
    m->rowptr64 = d_rowptr64;
    rv = gb_plan_build(m);
  }
  if (rv != GOOSEBERRY_SUCCESS) {
    // the caller keeps the arrays it passed in
    m->rowind = nullptr;
    m->val = nullptr;
    m->rowptr64 = nullptr;
    gbd_matrix_destroy(m);
    return rv;
  }
  *out = m;
  return GOOSEBERRY_SUCCESS;
}

uint32_t gbd_matrix_nparts(const gbd_matrix *m) { return m->nparts; }

int gbd_matrix_replan(gbd_matrix *m, const gbd_plan_params *params) {
  // a batch slot copied the plan's arrays into its descriptor
  if (m->holders) {
    gb_set_error("the matrix is held by %u batch or server slot(s): its plan cannot be "
                 "rebuilt until they are removed", m->holders);
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  gbd_plan_params p;
  if (params) p = *params;
  else gbd_plan_defaults(m->hist, m->nrows, m->nnz, &p);
  if (!gb_plan_check(&p)) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (m->nparts && (p.csplit > 1 || p.xtile)) {
    gb_set_error("a matrix held as row parts (64-bit row pointers) takes no column split "
                 "and no x tiles");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  m->params = p;
  return gb_plan_build(m);
}

void gbd_matrix_destroy(gbd_matrix *m) {
  if (!m) return;
  // pool frees are ordered after the queued work that reads the arrays
  if (m->view) {
    (void)gb_dev_free(m->blk);
    (void)gb_dev_free(m->bdesc);
    (void)gb_dev_free(m->rowbins);
    (void)gb_dev_free(m->dcol);
    gb_vdict_free(m);
    free(m);
    return;
  }
  gb_split_free(m);
  gb_vdict_free(m);
  gb_xtile_free(m);
  for (uint32_t q = 0; q < m->nparts; ++q) gbd_matrix_destroy(m->parts[q]);
  free(m->parts);
  free(m->prow);
  free(m->pbase);
  if (m->prowptr) (void)gb_dev_free(m->prowptr);
  if (m->rowptr64) (void)gb_dev_free(m->rowptr64);
  if (m->rowptr) (void)gb_dev_free(m->rowptr);
  if (m->rowind) (void)gb_dev_free(m->rowind);
  if (m->val) (void)gb_dev_free(m->val);
  if (m->val32) (void)gb_dev_free(m->val32);
  if (m->val32j) (void)gb_dev_free(m->val32j);
  if (m->jscale) (void)gb_dev_free(m->jscale);
  if (m->blk) (void)gb_dev_free(m->blk);
  if (m->bdesc) (void)gb_dev_free(m->bdesc);
  if (m->rowbins) (void)gb_dev_free(m->rowbins);
  if (m->dcol) (void)gb_dev_free(m->dcol);
  free(m);
}

int gbd_matrix_info(const gbd_matrix *m, uint32_t *nrows, uint32_t *ncols,
                    uint64_t *nnz, gbd_plan_params *params,
                    uint32_t *nblocks_stream, uint32_t *nrows_wave,
                    uint32_t *nrows_block, uint64_t *hist33) {
  if (nrows) *nrows = m->nrows;
  if (ncols) *ncols = m->ncols;
  if (nnz) *nnz = m->nnz;
  if (params) *params = m->params;
  if (nblocks_stream) *nblocks_stream = m->nstream;
  if (nrows_wave) *nrows_wave = m->nwave;
  if (nrows_block) *nrows_block = m->nblockrows;
  if (hist33) memcpy(hist33, m->hist, sizeof(m->hist));
  return GOOSEBERRY_SUCCESS;
}

// row parts: 1 when every part reads 16-bit offsets; the largest value
// table of the parts when every part is value-indexed
uint32_t gbd_matrix_index16(const gbd_matrix *m) {
  if (!m->nparts) return m->dcol ? 1u : 0u;
  for (uint32_t q = 0; q < m->nparts; ++q)
    if (!m->parts[q]->dcol) return 0u;
  return 1u;
}

uint32_t gbd_matrix_values8(const gbd_matrix *m) {
  if (!m->nparts) return m->vidx ? m->nvals : 0u;
  uint32_t nv = 0;
  for (uint32_t q = 0; q < m->nparts; ++q) {
    if (!m->parts[q]->vidx) return 0u;
    if (m->parts[q]->nvals > nv) nv = m->parts[q]->nvals;
  }
  return nv;
}

uint32_t gbd_matrix_holders(const gbd_matrix *m) { return m->holders; }

const uint32_t *gbd_matrix_rowptr(const gbd_matrix *m) { return m->rowptr; }
const uint32_t *gbd_matrix_rowind(const gbd_matrix *m) { return m->rowind; }
const double *gbd_matrix_rowval(const gbd_matrix *m) { return m->val; }

// the FP32 paths (mixed precision, batches, servers) take no row-parted
// matrix
static int gb_parts_refuse(const gbd_matrix *m, const char *what) {
  if (!m->nparts) return GOOSEBERRY_SUCCESS;
  gb_set_error("%s of a matrix held as row parts (64-bit row pointers, %u parts) is not "
               "supported", what, m->nparts);
  return GOOSEBERRY_ERROR_UNIMPLEMENTED;
}

int gbd_matrix_make_f32(gbd_matrix *m) {
  if (m->val32) return GOOSEBERRY_SUCCESS;
  GB_HIP(gb_dev_alloc((void **)&m->val32, (m->nnz ? m->nnz : 1) * sizeof(float)));
  if (m->nnz) {
    gb_to_f32<<<gb_div_up(m->nnz, GB_BLOCK), GB_BLOCK, 0, gb_stream()>>>(
        m->val, m->val32, m->nnz);
    GB_LAUNCH_CHECK();
  }
  if (m->nparts) {
    // each part views its share of the copy (as its values view the
    // matrix's) and rounds its own value table
    for (uint32_t q = 0; q < m->nparts; ++q) {
      gbd_matrix *v = m->parts[q];
      v->val32 = m->val32 + m->pbase[q];
      int rv = gb_vdict_f32(v);
      if (rv != GOOSEBERRY_SUCCESS) return rv;
    }
    return GOOSEBERRY_SUCCESS;
  }
  if (m->nsplit) {
    GB_HIP(gb_dev_alloc((void **)&m->sval32, (m->nnz ? m->nnz : 1) * sizeof(float)));
    if (m->nnz) {
      gb_to_f32<<<gb_div_up(m->nnz, GB_BLOCK), GB_BLOCK, 0, gb_stream()>>>(
          m->sval, m->sval32, m->nnz);
      GB_LAUNCH_CHECK();
    }
    for (uint32_t r = 0; r < m->nsplit; ++r) m->split[r]->val32 = m->sval32;
  }
  int rx = gb_xtile_f32(m);
  if (rx != GOOSEBERRY_SUCCESS) return rx;
  return gb_vdict_f32(m);
}

int gbd_matrix_make_f32_jacobi(gbd_matrix *m) {
  if (m->val32j) return GOOSEBERRY_SUCCESS;
  int rp = gb_parts_refuse(m, "Jacobi scaling");
  if (rp != GOOSEBERRY_SUCCESS) return rp;
  hipStream_t s = gb_stream();
  const uint32_t n = m->nrows;
  if (m->nrows != m->ncols) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t *bad = nullptr, hbad = 0;
  GB_HIP(gb_dev_alloc((void **)&m->jscale, (size_t)(n ? n : 1) * sizeof(double)));
  GB_HIP(gb_dev_alloc((void **)&bad, 4));
  GB_HIP(hipMemsetAsync(bad, 0, 4, s));
  const uint32_t g = n ? (gb_div_up(n, GB_WAVES_PER_BLOCK) > 16384u
                              ? 16384u : gb_div_up(n, GB_WAVES_PER_BLOCK)) : 0;
  if (n) {
    gb_jacobi_scale<<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->val, n, m->jscale, bad);
    GB_LAUNCH_CHECK();
  }
  GB_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  GB_HIP(gb_dev_free(bad));
  if (hbad) {
    GB_HIP(gb_dev_free(m->jscale));
    m->jscale = nullptr;
    gb_set_error("Jacobi scaling needs a positive, finite diagonal in every row");
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  GB_HIP(gb_dev_alloc((void **)&m->val32j, (m->nnz ? m->nnz : 1) * sizeof(float)));
  if (n && m->nnz) {
    gb_jacobi_vals<<<g, GB_BLOCK, 0, s>>>(m->rowptr, m->rowind, m->val, n, m->jscale,
                                          m->val32j);
    GB_LAUNCH_CHECK();
  }
  if (m->nsplit) {
    // the slices' copy in slice order, as the split build makes it
    const uint32_t P = m->nsplit;
    const uint32_t width = (m->ncols + P - 1) / P;
    GB_HIP(gb_dev_alloc((void **)&m->sval32j, (m->nnz ? m->nnz : 1) * sizeof(float)));
    gb_split_fill<float><<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(
        m->rowptr, m->rowind, m->val32j, n, width, P, m->srowptr, nullptr, m->sval32j);
    GB_LAUNCH_CHECK();
    for (uint32_t r = 0; r < P; ++r) m->split[r]->val32j = m->sval32j;
  }
  return gb_xtile_f32(m);
}

int gbd_spmv(const gbd_matrix *m, const double *d_x, double *d_y) {
  return gb_spmv_launch(m, d_x, d_y, 1, nullptr, nullptr);
}

int gbd_spmm(const gbd_matrix *m, uint32_t bncols, const double *d_b,
             double *d_c) {
  for (uint32_t j = 0; j < bncols; ++j) {
    int rv = gb_spmv_launch(m, d_b + (size_t)j * m->ncols, d_c + j, bncols,
                            nullptr, nullptr);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  return GOOSEBERRY_SUCCESS;
}

int gbd_spmv_f32(const gbd_matrix *m, const float *d_x, float *d_y) {
  return gb_spmv_launch_f32(m, d_x, d_y, nullptr, nullptr);
}

}  // extern "C"
