// gb_pagerank.hip -- PageRank power iteration over the resident SpMV
// (src/gooseberry_bin.c:2622-2737).
//
// Preparation, once per call:
//   column sums  indeg[c] = the values of column c added in row order (a
//                stable counting sort of the entries by column, then one
//                sequential sum per column: the reference's order)
//   normalise    nval[l] = val[l] / indeg[rowind[l]] (IEEE division), over
//                the copy of the values the SpMV plan reads (the CSR's, the
//                x-tile layout's or the column slices')
//   sinks        the rows without entries, ascending (scan + compaction)
// One iteration, three launches:
//   SpMV         out = N rank (the matrix's own plan over nval; with a
//                column-split plan, over the slices' copy of the values
//                normalised the same way, so each row is then summed per
//                slice and the slices added in order)
//   gb_pr_shift  c = (1 - d) / n + sum of rank[sinks] (the sink ranks are
//                added in sink order, as the reference does, up to
//                GB_PR_SEQ_SINKS sinks)
//   gb_pr_update out = out * d + c (scale, then add, separately rounded)
// and every tenth iteration the squared difference of rank and out
// (fixed-order two-level reduction) for the stopping test.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gb_matrix.h"
#include "gb_xtile.h"

// Sink ranks added one at a time in sink order up to this many sinks; above
// it, lane-strided sums and a butterfly (fixed order, not the reference's).
#define GB_PR_SEQ_SINKS 65536u
#define GB_PR_ERR_GRID 1024u
// niter == 0 iterates until the error test passes; the reference has no
// bound, this build stops after this many iterations.
#define GB_PR_MAX_ITER 1000000ull

// Column sums in row order, s = ((0 + v0) + v1) + ..., one dependent add
// after another as the reference's indeg[col] += val loop.  Columns of up to
// GB_PR_SHORT entries take one thread each, the loads issued eight at a time
// ahead of their adds; longer ones take a wavefront each (below), so a long
// column's loads do not serialise behind its adds.
#define GB_PR_SHORT 64

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_colsum(uint32_t ncols, const uint32_t *__restrict__ dist,
             const uint32_t *__restrict__ perm, const double *__restrict__ val,
             double *__restrict__ indeg) {
  const uint32_t c = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (c >= ncols) return;
  const uint32_t p0 = dist[c], p1 = dist[c + 1];
  if (p1 - p0 > GB_PR_SHORT) return;
  double s = 0.0;
  for (uint32_t p = p0; p < p1; p += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p + u < p1 ? val[perm[p + u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (p + u < p1) s += v[u];
  }
  indeg[c] = s;
}

// the columns longer than GB_PR_SHORT: one wavefront each, 64 values loaded
// at once (the next 64 while the current ones are added), added in order
__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_colsum_long(uint32_t ncols, const uint32_t *__restrict__ dist,
                  const uint32_t *__restrict__ perm, const double *__restrict__ val,
                  double *__restrict__ indeg) {
  const uint32_t c = blockIdx.x * GB_WAVES_PER_BLOCK + threadIdx.x / GB_WAVE;
  const uint32_t lane = threadIdx.x & (GB_WAVE - 1);
  if (c >= ncols) return;
  const uint32_t p0 = dist[c], p1 = dist[c + 1];
  if (p1 - p0 <= GB_PR_SHORT) return;
  double s = 0.0;
  double cur = p0 + lane < p1 ? val[perm[p0 + lane]] : 0.0;
  for (uint32_t p = p0; p < p1; p += GB_WAVE) {
    const uint32_t q = p + GB_WAVE + lane;
    const double next = q < p1 ? val[perm[q]] : 0.0;
    const uint32_t m = p1 - p < GB_WAVE ? p1 - p : GB_WAVE;
    // lane u's value read into scalar registers: the reads do not wait on
    // the adds, so only the adds form the chain
    const long long bits = __double_as_longlong(cur);
    const int lo = (int)(uint32_t)bits, hi = (int)(bits >> 32);
#pragma unroll
    for (int u = 0; u < GB_WAVE; ++u) {
      if ((uint32_t)u < m) {
        const uint32_t l = (uint32_t)__builtin_amdgcn_readlane(lo, u);
        const uint32_t h = (uint32_t)__builtin_amdgcn_readlane(hi, u);
        s += __longlong_as_double((long long)(((uint64_t)h << 32) | l));
      }
    }
    cur = next;
  }
  if (lane == 0) indeg[c] = s;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_normalise(uint64_t nnz, const uint32_t *__restrict__ rowind,
                const double *__restrict__ val, const double *__restrict__ indeg,
                double *__restrict__ nval) {
  for (uint64_t l = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; l < nnz;
       l += (uint64_t)gridDim.x * GB_BLOCK)
    nval[l] = val[l] / indeg[rowind[l]];
}

// The same quotients on an x-tile layout: entry j of group g = b T + t has
// column (t << lc) | (its word's low lc bits)
__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_normalise_xt(const uint32_t *__restrict__ ptr, uint32_t ngroups, uint32_t T,
                   uint32_t lc, const uint32_t *__restrict__ word,
                   const double *__restrict__ val, const double *__restrict__ indeg,
                   double *__restrict__ nval) {
  const uint32_t cm = (1u << lc) - 1;
  for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const uint32_t c0 = (g % T) << lc, e = ptr[g + 1];
    for (uint32_t j = ptr[g] + threadIdx.x; j < e; j += GB_BLOCK)
      nval[j] = val[j] / indeg[c0 | (word[j] & cm)];
  }
}

// Column sums in row order and the normalised values on an x-tile layout,
// one workgroup per column tile, with no sort.  The tile's entries in layout order (row blocks in
// order, each group in CSR order) are its columns' entries in row order.
// The workgroup walks them in chunks of GB_PRX_THREADS positions, one entry
// a thread (the next chunk's entries load while the current one is added),
// taking the groups of up to GB_PRX_THREADS row blocks at a time laid end to
// end.  A wavefront holds 64 consecutive positions: ballots of the column
// bits give every lane the lanes of its column, and the first of them adds
// their values to the column's sum in LDS in lane order.  A column met by
// one wavefront of the chunk is added at once; a column met by several is
// added by those wavefronts one after another, in wavefront order.  The
// sums are bitwise those of the sort-based path.  Then the workgroup divides
// every entry of the tile by its column's sum.
#define GB_PRX_THREADS 1024
#define GB_PRX_NORM 4

// A workgroup barrier that orders LDS only: global loads still in flight
// (the next chunk's entries) stay in flight across it, where __syncthreads
// would wait for them.  Every exchange across it here goes through LDS.
__device__ __forceinline__ void gb_prx_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// exclusive scan of one value per thread over GB_PRX_THREADS threads;
// *total gets the sum
__device__ __forceinline__ uint32_t gb_prx_scan_excl(uint32_t v, uint32_t *wsum,
                                                     uint32_t *total) {
  const uint32_t lane = threadIdx.x & (GB_WAVE - 1), wave = threadIdx.x / GB_WAVE;
  const uint32_t inc = gb_wave_scan_incl(v);
  if (lane == GB_WAVE - 1) wsum[wave] = inc;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < GB_PRX_THREADS / GB_WAVE; ++w) {
    const uint32_t x = wsum[w];
    if (w < wave) woff += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

// the layout position of window position q: the group k whose start
// goff[k] is the last one <= q (goff[0] = 0, non-decreasing over nb groups)
__device__ __forceinline__ uint32_t gb_prx_locate(const uint32_t *goff, const uint32_t *gst,
                                                  uint32_t nb, uint32_t q) {
  uint32_t lo = 0, hi = nb;  // goff[lo] <= q < goff[hi] (goff[nb] = the total)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (goff[mid] <= q) lo = mid;
    else hi = mid;
  }
  return gst[lo] + (q - goff[lo]);
}

// The tile's groups of row blocks b0 .. b0 + nb - 1 laid end to end: the
// first entry (gst) and first window position (goff) of each, goff[nb] = the
// window's entry count (returned).  Ends with a barrier.
__device__ __forceinline__ uint32_t gb_prx_window(const uint32_t *__restrict__ ptr,
                                                  uint32_t b0, uint32_t nb, uint32_t T,
                                                  uint32_t t, uint32_t *gst, uint32_t *goff,
                                                  uint32_t *wsum) {
  const uint32_t tid = threadIdx.x;
  uint32_t st = 0, sz = 0;
  if (tid < nb) {
    const uint32_t g = (b0 + tid) * T + t;
    st = ptr[g];
    sz = ptr[g + 1] - st;
  }
  uint32_t tot;
  const uint32_t off = gb_prx_scan_excl(sz, wsum, &tot);  // barriers inside
  gst[tid] = st;
  goff[tid] = off;
  if (tid == 0) goff[nb] = tot;
  __syncthreads();
  return tot;
}

// s + the values of the lanes in `peers` above this lane, in lane order
// (every lane of the wavefront takes part; `peers` is empty for most)
__device__ __forceinline__ double gb_prx_chain(double s, double v, uint64_t peers) {
  for (;;) {
    const bool more = peers != 0;
    if (!__ballot(more)) break;
    const int src = more ? __builtin_ctzll(peers) : (int)(threadIdx.x & (GB_WAVE - 1));
    const double w = __shfl(v, src, GB_WAVE);
    if (more) {
      s = s + w;
      peers &= peers - 1;
    }
  }
  return s;
}

__global__ void __launch_bounds__(GB_PRX_THREADS)
gb_pr_colsum_xt(const uint32_t *__restrict__ ptr, uint32_t W, uint32_t T, uint32_t lc,
                const uint32_t *__restrict__ word, const double *__restrict__ val,
                uint32_t ncols, double *__restrict__ indeg, double *__restrict__ nval) {
  __shared__ double cs[GB_XT_CMAX];           // the tile's column sums
  __shared__ uint32_t wset[GB_XT_CMAX];       // a column's wavefronts in the chunk
  __shared__ uint32_t gst[GB_PRX_THREADS];    // the window's groups: first entry
  __shared__ uint32_t goff[GB_PRX_THREADS + 1];  // and first window position
  __shared__ uint32_t wsum[GB_PRX_THREADS / GB_WAVE];
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  const uint32_t lane = tid & (GB_WAVE - 1), wave = tid / GB_WAVE;
  const uint64_t below = (1ull << lane) - 1;
  const uint32_t C = 1u << lc, cm = C - 1;
  for (uint32_t j = tid; j < C; j += GB_PRX_THREADS) {
    cs[j] = 0.0;
    wset[j] = 0;
  }
  for (uint32_t b0 = 0; b0 < W; b0 += GB_PRX_THREADS) {
    const uint32_t nb = W - b0 < GB_PRX_THREADS ? W - b0 : GB_PRX_THREADS;
    const uint32_t tot = gb_prx_window(ptr, b0, nb, T, t, gst, goff, wsum);
    // this thread's entry of the first chunk
    bool have = tid < tot;
    uint32_t col = 0;
    double v = 0.0;
    if (have) {
      const uint32_t j = gb_prx_locate(goff, gst, nb, tid);
      col = word[j] & cm;
      v = val[j];
    }
    // the first chunk's loads complete here: otherwise the compiler counts
    // them as pending at the top of the loop and waits for every load there,
    // the next chunk's included
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    for (uint32_t q0 = 0; q0 < tot; q0 += GB_PRX_THREADS) {
      // the next chunk's entry, loading while this chunk is added
      const uint32_t qn = q0 + GB_PRX_THREADS + tid;
      const bool nhave = qn < tot;
      uint32_t ncol = 0;
      double nv = 0.0;
      if (nhave) {
        const uint32_t j = gb_prx_locate(goff, gst, nb, qn);
        ncol = word[j];
        nv = val[j];
      }
      // the lanes of this wavefront with this lane's column
      uint64_t peers = __ballot(have);
      for (uint32_t bit = 0; bit < lc; ++bit) {
        const bool on = (col >> bit) & 1u;
        const uint64_t bb = __ballot(on);
        peers &= on ? bb : ~bb;
      }
      const bool first = have && (peers & below) == 0;
      if (first) atomicOr(&wset[col], 1u << wave);
      gb_prx_lds_barrier();
      const uint32_t ws = first ? wset[col] : 0u;
      const bool alone = first && ws == (1u << wave);
      const uint64_t rest = peers & ~(below | (1ull << lane));
      // columns of one wavefront: added now, every wavefront at once
      {
        double sum = alone ? cs[col] + v : 0.0;
        sum = gb_prx_chain(sum, v, alone ? rest : 0);
        if (alone) cs[col] = sum;
      }
      // columns of several wavefronts: wavefront by wavefront, in order
      const bool shared_col = first && !alone;
      for (uint32_t w = 0; w < GB_PRX_THREADS / GB_WAVE; ++w) {
        if (wave == w) {
          double sum = shared_col ? cs[col] + v : 0.0;
          sum = gb_prx_chain(sum, v, shared_col ? rest : 0);
          if (shared_col) cs[col] = sum;
        }
        gb_prx_lds_barrier();
      }
      if (first) wset[col] = 0;
      have = nhave;
      col = ncol & cm;
      v = nv;
      gb_prx_lds_barrier();
    }
  }
  for (uint32_t j = tid; j < C; j += GB_PRX_THREADS)
    if ((t << lc) + j < ncols) indeg[(t << lc) + j] = cs[j];
  // the quotients: every entry of the tile divided by its column's sum,
  // GB_PRX_NORM entries a thread loaded before any is divided
  for (uint32_t b0 = 0; b0 < W; b0 += GB_PRX_THREADS) {
    const uint32_t nb = W - b0 < GB_PRX_THREADS ? W - b0 : GB_PRX_THREADS;
    const uint32_t tot = gb_prx_window(ptr, b0, nb, T, t, gst, goff, wsum);
    for (uint32_t q0 = 0; q0 < tot; q0 += GB_PRX_NORM * GB_PRX_THREADS) {
      uint32_t jj[GB_PRX_NORM], ww[GB_PRX_NORM];
      double vv[GB_PRX_NORM];
#pragma unroll
      for (int u = 0; u < GB_PRX_NORM; ++u) {
        const uint32_t q = q0 + u * GB_PRX_THREADS + tid;
        if (q < tot) {
          jj[u] = gb_prx_locate(goff, gst, nb, q);
          ww[u] = word[jj[u]];
          vv[u] = val[jj[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < GB_PRX_NORM; ++u)
        if (q0 + u * GB_PRX_THREADS + tid < tot) nval[jj[u]] = vv[u] / cs[ww[u] & cm];
    }
    __syncthreads();  // gst and goff are rewritten by the next window
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_sink_flags(uint32_t n, const uint32_t *__restrict__ rowptr,
                 uint32_t *__restrict__ flag) {
  const uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (k < n) flag[k] = rowptr[k] == rowptr[k + 1] ? 1u : 0u;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_sink_compact(uint32_t n, const uint32_t *__restrict__ rowptr,
                   const uint32_t *__restrict__ idx, uint32_t *__restrict__ sinks) {
  const uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (k < n && rowptr[k] == rowptr[k + 1]) sinks[idx[k]] = k;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_fill(uint32_t n, double v, double *__restrict__ x) {
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < n; k += gridDim.x * GB_BLOCK)
    x[k] = v;
}

// one wavefront: c = ((1 - d) / n) + dist
__global__ void __launch_bounds__(GB_WAVE)
gb_pr_shift(const double *__restrict__ rank, const uint32_t *__restrict__ sinks,
            uint32_t nsinks, double damping, uint32_t n, double *__restrict__ c) {
  const uint32_t lane = threadIdx.x;
  double dist = 0;
  if (nsinks <= GB_PR_SEQ_SINKS) {
    // 64 ranks loaded at once, added in sink order by every lane
    for (uint32_t base = 0; base < nsinks; base += GB_WAVE) {
      const uint32_t k = base + lane;
      const double v = k < nsinks ? rank[sinks[k]] : 0.0;
      const uint32_t cnt = min((uint32_t)GB_WAVE, nsinks - base);
      for (uint32_t j = 0; j < cnt; ++j) dist += gb_readlane(v, j);
    }
  } else {
    for (uint32_t k = lane; k < nsinks; k += GB_WAVE) dist += rank[sinks[k]];
    dist = gb_wave_sum(dist);
  }
  if (lane == 0) c[0] = ((1.0 - damping) / n) + dist;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_update(uint32_t n, double *__restrict__ out, double damping,
             const double *__restrict__ c) {
  const double cc = c[0];
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < n; k += gridDim.x * GB_BLOCK) {
    const double t = out[k] * damping;  // gooseberry_scale
    out[k] = t + cc;                    // gooseberry_add_scalar
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_err_partials(uint32_t n, const double *__restrict__ a,
                   const double *__restrict__ b, double *__restrict__ part) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  double e = 0.0;
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < n; k += gridDim.x * GB_BLOCK) {
    const double d = a[k] - b[k];
    e += d * d;
  }
  e = gb_block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = e;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_pr_err_final(const double *__restrict__ part, uint32_t np, double *__restrict__ err) {
  __shared__ double red[GB_WAVES_PER_BLOCK];
  double e = 0.0;
  for (uint32_t i = threadIdx.x; i < np; i += GB_BLOCK) e += part[i];
  e = gb_block_sum(e, red);
  if (threadIdx.x == 0) err[0] = e;
}

static uint32_t gb_pr_grid(uint64_t n) {
  uint64_t g = (n + GB_BLOCK - 1) / GB_BLOCK;
  if (g > 65536) g = 65536;
  return g ? (uint32_t)g : 1u;
}

extern "C" {

int gbd_pagerank(const gbd_matrix *m, double damping, uint64_t niter,
                 double minerror, double *d_out, uint64_t *r_iters,
                 double *r_rmse, uint32_t *r_nsinks) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!m || !d_out || m->nrows != m->ncols || m->nrows == 0)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (m->nparts) {
    gb_set_error("PageRank of a matrix held as row parts (64-bit row pointers) is not "
                 "supported");
    return GOOSEBERRY_ERROR_UNIMPLEMENTED;
  }
  hipStream_t s = gb_stream();
  const uint32_t n = m->nrows;
  const uint64_t nnz = m->nnz;
  const uint32_t g = gb_pr_grid(n);
  double *indeg = nullptr, *nval = nullptr, *rank = nullptr, *out = nullptr,
         *c = nullptr, *part = nullptr, *err = nullptr;
  uint32_t *perm = nullptr, *dist = nullptr, *flag = nullptr, *idx = nullptr,
           *sinks = nullptr;
  uint32_t nsinks = 0;
  uint64_t i = 0;
  double e = 0.0;
  bool finished = false;
  gbd_matrix mm;
  gbd_matrix sv[GB_SPLIT_MAX];
  gbd_matrix *svp[GB_SPLIT_MAX];
  double *snval = nullptr;
  hipError_t he = hipSuccess;
  // An x-tile or column-split plan reads its own copy of the values, which
  // is normalised below; the CSR's order is then never read.
  const bool own_copy = m->xt_ptr || m->nsplit;
  // An x-tile plan sums the columns on its own layout, without the sort;
  // GOOSEBERRY_PAGERANK_SORTED_SUMS (set, not "0") takes the sort instead
  // (the same bits: a check of one path against the other).
  const char *sorted_env = getenv("GOOSEBERRY_PAGERANK_SORTED_SUMS");
  const bool tile_sums = m->xt_ptr && nnz &&
                         !(sorted_env && *sorted_env && strcmp(sorted_env, "0") != 0);
#define GB_PR_TRY(call)                                         \
  do {                                                          \
    he = (call);                                                \
    if (he != hipSuccess) {                                     \
      rv = gb_hip_fail(he, #call, __FILE__, __LINE__);          \
      goto done;                                                \
    }                                                           \
  } while (0)
#define GB_PR_RV(call)                                          \
  do {                                                          \
    rv = (call);                                                \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                    \
  } while (0)

  GB_PR_TRY(gb_dev_alloc((void **)&indeg, (size_t)n * 8));
  if (!own_copy) GB_PR_TRY(gb_dev_alloc((void **)&nval, (size_t)(nnz ? nnz : 1) * 8));
  if (!tile_sums) {
    GB_PR_TRY(gb_dev_alloc((void **)&perm, (size_t)(nnz ? nnz : 1) * 4));
    GB_PR_TRY(gb_dev_alloc((void **)&dist, ((size_t)n + 1) * 4));
  }
  GB_PR_TRY(gb_dev_alloc((void **)&flag, (size_t)n * 4));
  GB_PR_TRY(gb_dev_alloc((void **)&idx, ((size_t)n + 1) * 4));
  GB_PR_TRY(gb_dev_alloc((void **)&sinks, (size_t)n * 4));
  GB_PR_TRY(gb_dev_alloc((void **)&rank, (size_t)n * 8));
  GB_PR_TRY(gb_dev_alloc((void **)&out, (size_t)n * 8));
  GB_PR_TRY(gb_dev_alloc((void **)&c, 8));
  GB_PR_TRY(gb_dev_alloc((void **)&part, GB_PR_ERR_GRID * 8));
  GB_PR_TRY(gb_dev_alloc((void **)&err, 8));

  if (tile_sums) {
    // column sums and the layout's quotients, one workgroup per column tile
    GB_PR_TRY(gb_dev_alloc((void **)&snval, ((size_t)nnz + GB_XT_PAD) * 8));
    GB_PR_TRY(hipMemsetAsync(snval + nnz, 0, GB_XT_PAD * 8, s));
    gb_pr_colsum_xt<<<m->xt_T, GB_PRX_THREADS, 0, s>>>(m->xt_ptr, m->xt_W, m->xt_T, m->xt_lc,
                                                     m->xt_idx, m->xt_val, n, indeg, snval);
    GB_PR_TRY(hipGetLastError());
  } else {
    // column sums in row order (column indices were validated with the matrix)
    GB_PR_RV(gbd_countingsort_kv_unchecked(m->rowind, nullptr, 0, n - 1, nnz, perm, dist));
    gb_pr_colsum<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(n, dist, perm, m->val, indeg);
    GB_PR_TRY(hipGetLastError());
    gb_pr_colsum_long<<<gb_div_up(n, GB_WAVES_PER_BLOCK), GB_BLOCK, 0, s>>>(n, dist, perm,
                                                                           m->val, indeg);
    GB_PR_TRY(hipGetLastError());
  }
  if (nnz && !own_copy) {
    gb_pr_normalise<<<gb_pr_grid(nnz), GB_BLOCK, 0, s>>>(nnz, m->rowind, m->val, indeg,
                                                         nval);
    GB_PR_TRY(hipGetLastError());
  }
  // sinks
  gb_pr_sink_flags<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(n, m->rowptr, flag);
  GB_PR_TRY(hipGetLastError());
  GB_PR_RV(gbd_scan_u32(flag, idx, n, idx + n));
  gb_pr_sink_compact<<<gb_div_up(n, GB_BLOCK), GB_BLOCK, 0, s>>>(n, m->rowptr, idx, sinks);
  GB_PR_TRY(hipGetLastError());
  GB_PR_TRY(hipMemcpyAsync(&nsinks, idx + n, 4, hipMemcpyDeviceToHost, s));
  GB_PR_TRY(hipStreamSynchronize(s));

  // the normalised matrix: the same structure and plan over nval (null with
  // a plan of its own copy: its launches read only that copy)
  mm = *m;
  mm.val = nval;
  mm.val32 = nullptr;
  // the normalised values are read as stored (no value index)
  mm.vidx = nullptr;
  mm.vtab = nullptr;
  mm.vtab32 = nullptr;
  mm.nvals = 0;
  mm.nsplit = 0;
  mm.split = nullptr;
  if (m->xt_ptr && nnz) {
    // an x-tile plan: the layout's copy of the values divided by the same
    // column sums (the same quotient per entry), with the layout's zero
    // entries past the end (its SpMV reads vectors); made above with the
    // sums unless they came from the sort
    if (!tile_sums) {
      GB_PR_TRY(gb_dev_alloc((void **)&snval, ((size_t)nnz + GB_XT_PAD) * 8));
      GB_PR_TRY(hipMemsetAsync(snval + nnz, 0, GB_XT_PAD * 8, s));
      const uint32_t ng = m->xt_W * m->xt_T;
      gb_pr_normalise_xt<<<ng < 65536u ? ng : 65536u, GB_BLOCK, 0, s>>>(
          m->xt_ptr, ng, m->xt_T, m->xt_lc, m->xt_idx, m->xt_val, indeg, snval);
      GB_PR_TRY(hipGetLastError());
    }
    mm.xt_val = snval;
  }
  if (m->nsplit && nnz) {
    // a column-split plan: the slices' copy of the values divided by the
    // same column sums (the same quotient per entry), and views over it
    GB_PR_TRY(gb_dev_alloc((void **)&snval, (size_t)nnz * 8));
    gb_pr_normalise<<<gb_pr_grid(nnz), GB_BLOCK, 0, s>>>(nnz, m->srowind, m->sval, indeg,
                                                         snval);
    GB_PR_TRY(hipGetLastError());
    for (uint32_t r = 0; r < m->nsplit; ++r) {
      sv[r] = *m->split[r];
      sv[r].val = snval;
      sv[r].val32 = nullptr;
      svp[r] = &sv[r];
    }
    mm.nsplit = m->nsplit;
    mm.split = svp;
    mm.sval = snval;
    mm.sval32 = nullptr;
  }
  gb_pr_fill<<<g, GB_BLOCK, 0, s>>>(n, 1.0 / n, rank);
  GB_PR_TRY(hipGetLastError());
  for (i = 0; (niter == 0 && i < GB_PR_MAX_ITER) || i < niter; ++i) {
    GB_PR_RV(gbd_spmv(&mm, rank, out));
    gb_pr_shift<<<1, GB_WAVE, 0, s>>>(rank, sinks, nsinks, damping, n, c);
    GB_PR_TRY(hipGetLastError());
    gb_pr_update<<<g, GB_BLOCK, 0, s>>>(n, out, damping, c);
    GB_PR_TRY(hipGetLastError());
    if (i % 10 == 0) {
      gb_pr_err_partials<<<GB_PR_ERR_GRID, GB_BLOCK, 0, s>>>(n, rank, out, part);
      GB_PR_TRY(hipGetLastError());
      gb_pr_err_final<<<1, GB_BLOCK, 0, s>>>(part, GB_PR_ERR_GRID, err);
      GB_PR_TRY(hipGetLastError());
      GB_PR_TRY(hipMemcpyAsync(&e, err, 8, hipMemcpyDeviceToHost, s));
      GB_PR_TRY(hipStreamSynchronize(s));
      if (e <= minerror * minerror) {
        ++i;
        finished = true;
        break;
      }
    }
    double *t = rank;
    rank = out;
    out = t;
  }
  if (!finished) {
    gb_pr_err_partials<<<GB_PR_ERR_GRID, GB_BLOCK, 0, s>>>(n, rank, out, part);
    GB_PR_TRY(hipGetLastError());
    gb_pr_err_final<<<1, GB_BLOCK, 0, s>>>(part, GB_PR_ERR_GRID, err);
    GB_PR_TRY(hipGetLastError());
    GB_PR_TRY(hipMemcpyAsync(&e, err, 8, hipMemcpyDeviceToHost, s));
  }
  // the reference writes its `out` buffer: after a completed loop that is
  // the iterate before the last one
  GB_PR_TRY(hipMemcpyAsync(d_out, out, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  GB_PR_TRY(hipStreamSynchronize(s));
  if (r_iters) *r_iters = i;
  if (r_rmse) *r_rmse = sqrt(e);
  if (r_nsinks) *r_nsinks = nsinks;
  rv = GOOSEBERRY_SUCCESS;
done:
  (void)gb_dev_free(indeg);
  (void)gb_dev_free(nval);
  (void)gb_dev_free(snval);
  (void)gb_dev_free(perm);
  (void)gb_dev_free(dist);
  (void)gb_dev_free(flag);
  (void)gb_dev_free(idx);
  (void)gb_dev_free(sinks);
  (void)gb_dev_free(rank);
  (void)gb_dev_free(out);
  (void)gb_dev_free(c);
  (void)gb_dev_free(part);
  (void)gb_dev_free(err);
  return rv;
#undef GB_PR_TRY
#undef GB_PR_RV
}

}  // extern "C"
