// gb_spgemm.hip -- C = A B for sparse A and B (src/blas.c:94-234,609-791),
// bit for bit the reference's row-by-row scan.
//
// The reference walks row i of A entry by entry (j), and for each the row
// arowind[j] of B entry by entry (k): the product x_j * b_k goes to column
// browind[k] of C.  Each C element is the sum of its products in that scan
// order; in the sparse result the columns of a row appear in the order of
// their first product.  A row's products are numbered t = 0, 1, ... in that
// scan order.  Rows take one of two paths:
//
// Short rows (at most 64 products from at most 64 entries of A): one
// wavefront per row, lane t holding product t.  The lanes enter their
// columns in a 128-slot LDS hash table, each slot keeping the mask of lanes
// with its column: a lane with no lower lane in its mask is the head of its
// column, and its sum is its own product followed by the other lanes' in
// the mask, added in lane (= t) order.  The heads in lane order are the
// columns in first-appearance order (a ballot gives each head its rank).
//
// Other rows: every product is materialised at its position t (numbered
// over these rows, row-major), then
//   1. stable radix sort of t by column, then by row: (row, col, t) order
//   2. one sequential sum per (row, col) run, in t order
//   3. the runs sorted by their first t (unique, row-major), so each row
//      lists its columns by first appearance.
//
// Both paths add a column's products one after another in t order: the
// first product, then += the others (the dense variant starts from +0.0 as
// the reference's zeroed row does, so a lone -0.0 becomes +0.0).  The sparse
// result needs each row's entry count before the rows can be placed: the
// short rows write their entries in first-appearance order to a staging
// area at the row's first product position, with their counts; the scan of
// the counts gives the row pointers and a copy puts the staged rows in
// place.  (When the staging area would pass its limit, 8 GiB unless
// GOOSEBERRY_SPGEMM_STAGE_MAX says otherwise, the short rows are counted in
// one pass and written in another instead.)
#include <stdlib.h>

#include "gb_common.h"

#define GB_SG_SHORT 64  // products (and A entries) of a wavefront row

enum { GB_SG_COUNT = 0, GB_SG_FILL = 1, GB_SG_DENSE = 2, GB_SG_STAGE = 3 };
#define GB_SG_STAGE_MAX (8ull << 30)  // bytes of staged short-row entries

// the staging limit: GOOSEBERRY_SPGEMM_STAGE_MAX (bytes; 0 turns staging
// off) when set, read at every product
static uint64_t gb_sg_stage_max() {
  const char *e = getenv("GOOSEBERRY_SPGEMM_STAGE_MAX");
  if (e && *e) {
    char *end = nullptr;
    const unsigned long long v = strtoull(e, &end, 10);
    if (end != e) return v;
  }
  return GB_SG_STAGE_MAX;
}

// products per A entry: the length of the B row it selects
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_len(uint64_t annz, const uint32_t *__restrict__ arowind,
          const uint32_t *__restrict__ browptr, uint32_t *__restrict__ len) {
  for (uint64_t j = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; j < annz;
       j += (uint64_t)gridDim.x * GB_BLOCK) {
    const uint32_t l = arowind[j];
    len[j] = browptr[l + 1] - browptr[l];
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_sum64(const uint32_t *__restrict__ a, uint64_t n, unsigned long long *__restrict__ total) {
  __shared__ unsigned long long lds[GB_WAVES_PER_BLOCK];
  unsigned long long v = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * GB_BLOCK)
    v += a[i];
  v = gb_block_sum(v, lds);
  if (threadIdx.x == 0) atomicAdd(total, v);
}

// row classes: isshort[i] = 1 for a wavefront row, islong[i] = 1 for a row
// of the sort path (off = exclusive scan of the products per A entry)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_classify(uint32_t anrows, const uint32_t *__restrict__ arowptr,
               const uint32_t *__restrict__ off, uint32_t *__restrict__ isshort,
               uint32_t *__restrict__ islong) {
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < anrows;
       i += gridDim.x * GB_BLOCK) {
    const uint32_t j0 = arowptr[i], j1 = arowptr[i + 1];
    const uint32_t p = off[j1] - off[j0];
    const bool s = p <= GB_SG_SHORT && j1 - j0 <= GB_SG_SHORT;
    isshort[i] = s ? 1u : 0u;
    islong[i] = s ? 0u : 1u;
  }
}

// list[pos[i]] = i for flagged rows (pos = exclusive scan of the flags)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_compact(uint32_t n, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
              uint32_t *__restrict__ list) {
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n; i += gridDim.x * GB_BLOCK)
    if (flag[i]) list[pos[i]] = i;
}

#define GB_SG_HASH_BITS 7
#define GB_SG_HASH (1u << GB_SG_HASH_BITS)  // slots per wavefront row (>= 2 x 64)
#define GB_SG_EMPTY 0xffffffffu            // no column reaches 2^32 - 1

// LDS operations of one wavefront complete in order; this keeps the
// compiler from moving them across the step and waits for them
__device__ __forceinline__ void gb_sg_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One wavefront per short row.  COUNT: cnt[i] = distinct columns; FILL:
// the row's entries at crowptr[i]; DENSE: the sums into cdense row i.
template <int MODE>
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_wave(uint32_t nlist, const uint32_t *__restrict__ list,
           const uint32_t *__restrict__ arowptr, const uint32_t *__restrict__ arowind,
           const double *__restrict__ arowval, const uint32_t *__restrict__ browptr,
           const uint32_t *__restrict__ browind, const double *__restrict__ browval,
           const uint32_t *__restrict__ crowptr, uint32_t *__restrict__ cnt,
           uint32_t *__restrict__ crowind, double *__restrict__ crowval,
           double *__restrict__ cdense, uint32_t bncols,
           const uint32_t *__restrict__ pbase) {
  __shared__ uint32_t hkey[GB_WAVES_PER_BLOCK][GB_SG_HASH];
  __shared__ unsigned long long hmask[GB_WAVES_PER_BLOCK][GB_SG_HASH];
  __shared__ double hval[GB_WAVES_PER_BLOCK][GB_WAVE];
  const uint32_t w = blockIdx.x * GB_WAVES_PER_BLOCK + threadIdx.x / GB_WAVE;
  const uint32_t lane = threadIdx.x & (GB_WAVE - 1);
  if (w >= nlist) return;  // the whole wavefront
  // one row per wavefront: its scalars live in scalar registers
  const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)list[w]);
  const uint32_t j0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)arowptr[i]);
  const uint32_t m = (uint32_t)__builtin_amdgcn_readfirstlane((int)arowptr[i + 1]) - j0;
  // lane a < m: A entry j0 + a and the B row it selects
  uint32_t bstart = 0, blen = 0;
  double x = 0.0;
  if (lane < m) {
    const uint32_t l = arowind[j0 + lane];
    bstart = browptr[l];
    blen = browptr[l + 1] - bstart;
    x = arowval[j0 + lane];
  }
  const uint32_t incl = gb_wave_scan_incl(blen);
  const uint32_t excl = incl - blen;
  const uint32_t P = (uint32_t)__builtin_amdgcn_readlane((int)incl, GB_WAVE - 1);  // <= 64
  // product t = lane comes from the last A entry a whose products start at
  // or before t (entries selecting empty rows start where the next starts)
  uint32_t a = 0;
  for (uint32_t u = 0; u < m; ++u)
    if ((uint32_t)__builtin_amdgcn_readlane((int)excl, (int)u) <= lane) a = u;
  const uint32_t ex_a = (uint32_t)__shfl((int)excl, (int)a, GB_WAVE);
  const uint32_t bs_a = (uint32_t)__shfl((int)bstart, (int)a, GB_WAVE);
  const double x_a = __shfl(x, (int)a, GB_WAVE);
  const bool live = lane < P;
  uint32_t col = 0xffffffffu;
  double val = 0.0;
  if (live) {
    const uint32_t k = bs_a + (lane - ex_a);
    col = browind[k];
    val = x_a * browval[k];
  }
  // the row's columns in a small LDS hash table: slot -> column and the
  // mask of lanes holding it; a lane is its column's head when no lower
  // lane is in the mask, and the head adds the other lanes' products in
  // lane (= scan) order
  uint32_t *key = hkey[threadIdx.x / GB_WAVE];
  unsigned long long *msk = hmask[threadIdx.x / GB_WAVE];
  double *hv = hval[threadIdx.x / GB_WAVE];
  key[lane] = GB_SG_EMPTY;
  key[lane + GB_WAVE] = GB_SG_EMPTY;
  msk[lane] = 0ull;
  msk[lane + GB_WAVE] = 0ull;
  hv[lane] = val;
  gb_sg_wave_sync();
  uint32_t slot = 0;
  if (live) {
    uint32_t h = (col * 2654435761u) >> (32 - GB_SG_HASH_BITS);
    for (;;) {  // at most 64 columns in 128 slots: a free or equal slot exists
      const uint32_t prev = atomicCAS(&key[h], GB_SG_EMPTY, col);
      if (prev == GB_SG_EMPTY || prev == col) break;
      h = (h + 1) & (GB_SG_HASH - 1);
    }
    slot = h;
    atomicOr(&msk[slot], 1ull << lane);
  }
  gb_sg_wave_sync();
  bool head = false;
  double sum = 0.0;
  if (live) {
    const unsigned long long mk = msk[slot];
    head = (mk & ((1ull << lane) - 1ull)) == 0;
    if (head && MODE != GB_SG_COUNT) {
      sum = MODE == GB_SG_DENSE ? 0.0 + val : val;
      unsigned long long rest = mk & ~((2ull << lane) - 1ull);  // lanes above this one
      while (rest) {
        sum += hv[__ffsll(rest) - 1];
        rest &= rest - 1ull;
      }
    }
  }
  const uint64_t heads = __ballot(head);
  if (MODE == GB_SG_COUNT || MODE == GB_SG_STAGE) {
    if (lane == 0) cnt[i] = (uint32_t)__popcll(heads);
  }
  if (MODE != GB_SG_COUNT && head) {
    if (MODE == GB_SG_FILL || MODE == GB_SG_STAGE) {
      // STAGE: the row's staging slots start at its first product position
      // (pbase[j0]) and it has at most P <= its product count entries
      const uint32_t rank = (uint32_t)__popcll(heads & ((1ull << lane) - 1ull));
      const uint32_t o = (MODE == GB_SG_FILL ? crowptr[i] : pbase[j0]) + rank;
      crowind[o] = col;
      crowval[o] = sum;
    } else {
      cdense[(size_t)i * bncols + col] = sum;
    }
  }
}

// the staged short rows to their places, GB_BLOCK list rows per workgroup:
// row i's cnt[i] entries from its staging slots (pbase[arowptr[i]] on) to
// crowptr[i] on.  The workgroup numbers its rows' entries k = 0, 1, ...
// (a scan of the counts); thread t copies entries t, t + GB_BLOCK, ...,
// finding each one's row by binary search over the scan, so consecutive
// entries go to consecutive places.
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_unstage(uint32_t nlist, const uint32_t *__restrict__ list,
              const uint32_t *__restrict__ arowptr, const uint32_t *__restrict__ pbase,
              const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ sind,
              const double *__restrict__ sval, const uint32_t *__restrict__ crowptr,
              uint32_t *__restrict__ crowind, double *__restrict__ crowval) {
  __shared__ uint32_t lds4[GB_WAVES_PER_BLOCK];
  __shared__ uint32_t first[GB_BLOCK + 1];  // block entry number of each row's first
  __shared__ uint32_t src[GB_BLOCK], dst[GB_BLOCK];
  const uint32_t w = blockIdx.x * GB_BLOCK + threadIdx.x;
  uint32_t c = 0;
  if (w < nlist) {
    const uint32_t i = list[w];
    c = cnt[i];
    src[threadIdx.x] = pbase[arowptr[i]];
    dst[threadIdx.x] = crowptr[i];
  }
  uint32_t total = 0;
  const uint32_t e = gb_block_scan_excl(c, lds4, &total);
  first[threadIdx.x] = e;
  if (threadIdx.x == 0) first[GB_BLOCK] = total;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < total; k += GB_BLOCK) {
    // the last row r with first[r] <= k (rows without entries share their
    // first with the next row, and the search skips them)
    uint32_t lo = 0, hi = GB_BLOCK - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (first[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t d = k - first[lo];
    crowind[dst[lo] + d] = sind[src[lo] + d];
    crowval[dst[lo] + d] = sval[src[lo] + d];
  }
}

// the row of A entry j: the last row i with arowptr[i] <= j
__device__ __forceinline__ uint32_t gb_sg_row_of(const uint32_t *arowptr, uint32_t nrows,
                                                 uint32_t j) {
  uint32_t lo = 0, hi = nrows - 1;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    if (arowptr[mid] <= j) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// products per A entry of the sort path's rows (0 for short rows)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_len_long(uint32_t anrows, uint64_t annz, const uint32_t *__restrict__ arowptr,
               const uint32_t *__restrict__ islong, const uint32_t *__restrict__ len,
               uint32_t *__restrict__ len2) {
  for (uint64_t j = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; j < annz;
       j += (uint64_t)gridDim.x * GB_BLOCK) {
    const uint32_t i = gb_sg_row_of(arowptr, anrows, (uint32_t)j);
    len2[j] = islong[i] ? len[j] : 0u;
  }
}

// the products of A entry j (row i, column l) with B row l, at positions
// off[j] .. off[j] + len[j] - 1: their scan order
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_expand(uint32_t anrows, uint64_t annz, const uint32_t *__restrict__ arowptr,
             const uint32_t *__restrict__ arowind, const double *__restrict__ arowval,
             const uint32_t *__restrict__ browptr, const uint32_t *__restrict__ browind,
             const double *__restrict__ browval, const uint32_t *__restrict__ off,
             const uint32_t *__restrict__ len, uint32_t *__restrict__ erow,
             uint32_t *__restrict__ ecol, double *__restrict__ eval) {
  for (uint64_t j = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; j < annz;
       j += (uint64_t)gridDim.x * GB_BLOCK) {
    const uint32_t nk = len[j];
    if (!nk) continue;
    const uint32_t i = gb_sg_row_of(arowptr, anrows, (uint32_t)j);
    const uint32_t l = arowind[j];
    const double x = arowval[j];
    const uint32_t k0 = browptr[l];
    uint32_t t = off[j];
    for (uint32_t k = k0; k < k0 + nk; ++k, ++t) {
      erow[t] = i;
      ecol[t] = browind[k];
      eval[t] = x * browval[k];
    }
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_gather_u32(uint32_t n, const uint32_t *__restrict__ idx,
                 const uint32_t *__restrict__ in, uint32_t *__restrict__ out) {
  for (uint32_t p = blockIdx.x * GB_BLOCK + threadIdx.x; p < n; p += gridDim.x * GB_BLOCK)
    out[p] = in[idx[p]];
}

// run starts of equal (row, col) in sorted order (ord = t sorted)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_runflag(uint32_t n, const uint32_t *__restrict__ ord, const uint32_t *__restrict__ erow,
              const uint32_t *__restrict__ ecol, uint32_t *__restrict__ flag) {
  for (uint32_t p = blockIdx.x * GB_BLOCK + threadIdx.x; p < n; p += gridDim.x * GB_BLOCK) {
    uint32_t f = 1;
    if (p > 0) {
      const uint32_t a = ord[p - 1], b = ord[p];
      f = (erow[a] != erow[b] || ecol[a] != ecol[b]) ? 1u : 0u;
    }
    flag[p] = f;
  }
}

// start position of each run (runstart[run] = p)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_runstart(uint32_t n, const uint32_t *__restrict__ flag,
               const uint32_t *__restrict__ rid, uint32_t *__restrict__ runstart) {
  for (uint32_t p = blockIdx.x * GB_BLOCK + threadIdx.x; p < n; p += gridDim.x * GB_BLOCK)
    if (flag[p]) runstart[rid[p]] = p;
}

// one sequential sum per run, in t order; the run's row, column, first t
template <bool DENSE>
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_runsum(uint32_t nruns, uint32_t n, const uint32_t *__restrict__ runstart,
             const uint32_t *__restrict__ ord, const uint32_t *__restrict__ erow,
             const uint32_t *__restrict__ ecol, const double *__restrict__ eval,
             uint32_t bncols, double *__restrict__ cdense, uint32_t *__restrict__ rrow,
             uint32_t *__restrict__ rcol, uint32_t *__restrict__ rfirst,
             double *__restrict__ rval) {
  for (uint32_t q = blockIdx.x * GB_BLOCK + threadIdx.x; q < nruns; q += gridDim.x * GB_BLOCK) {
    const uint32_t p0 = runstart[q], p1 = q + 1 < nruns ? runstart[q + 1] : n;
    const uint32_t t0 = ord[p0];
    double v;
    if (DENSE) {
      v = 0.0;
      for (uint32_t p = p0; p < p1; ++p) v += eval[ord[p]];
      cdense[(size_t)erow[t0] * bncols + ecol[t0]] = v;
    } else {
      v = eval[t0];
      for (uint32_t p = p0 + 1; p < p1; ++p) v += eval[ord[p]];
      rrow[q] = erow[t0];
      rcol[q] = ecol[t0];
      rfirst[q] = t0;
      rval[q] = v;
    }
  }
}

// runs are in (row, col) order: rbeg[r] = the first run of row r and
// cnt[r] = its number of runs, for every row that has runs (a row's runs
// are contiguous; its last one is found by binary search)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_rowruns(uint32_t nruns, const uint32_t *__restrict__ rrow, uint32_t *__restrict__ rbeg,
              uint32_t *__restrict__ cnt) {
  for (uint32_t q = blockIdx.x * GB_BLOCK + threadIdx.x; q < nruns;
       q += gridDim.x * GB_BLOCK) {
    const uint32_t r = rrow[q];
    if (q == 0 || rrow[q - 1] != r) {
      uint32_t lo = q, hi = nruns - 1;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (rrow[mid] == r) lo = mid; else hi = mid - 1;
      }
      rbeg[r] = q;
      cnt[r] = lo + 1 - q;
    }
  }
}

// runs in first-appearance order (ordr): run q is entry q - rbeg[r] of its
// row r (the order is row-major, so a row's runs are q = rbeg[r] ..)
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_out(uint32_t nruns, const uint32_t *__restrict__ ordr, const uint32_t *__restrict__ rrow,
          const uint32_t *__restrict__ rcol, const double *__restrict__ rval,
          const uint32_t *__restrict__ rbeg, const uint32_t *__restrict__ crowptr,
          uint32_t *__restrict__ crowind, double *__restrict__ crowval) {
  for (uint32_t q = blockIdx.x * GB_BLOCK + threadIdx.x; q < nruns; q += gridDim.x * GB_BLOCK) {
    const uint32_t sidx = ordr[q];
    const uint32_t r = rrow[sidx];
    const uint32_t o = crowptr[r] + (q - rbeg[r]);
    crowind[o] = rcol[sidx];
    crowval[o] = rval[sidx];
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_zero(size_t n, double *__restrict__ c) {
  for (size_t i = (size_t)blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += (size_t)gridDim.x * GB_BLOCK)
    c[i] = 0.0;
}

static uint32_t gb_sg_grid(uint64_t n) {
  uint64_t g = (n + GB_BLOCK - 1) / GB_BLOCK;
  if (g > 65536) g = 65536;
  return g ? (uint32_t)g : 1u;
}

extern "C" {

int gbd_spgemm(uint32_t anrows, uint32_t ancols, uint32_t bncols,
               const uint32_t *d_arowptr, const uint32_t *d_arowind,
               const double *d_arowval, const uint32_t *d_browptr,
               const uint32_t *d_browind, const double *d_browval,
               double *d_cdense, uint32_t **d_crowptr, uint32_t **d_crowind,
               double **d_crowval, uint32_t *r_cnnz) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  const bool dense = d_cdense != nullptr;
  if (!dense && (!d_crowptr || !d_crowind || !d_crowval || !r_cnnz))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  hipStream_t s = gb_stream();
  uint32_t annz = 0, E = 0, nruns = 0, nshort = 0, nlong = 0, cnnz = 0;
  uint32_t *len = nullptr, *off = nullptr, *isshort = nullptr, *islong = nullptr,
           *slist = nullptr, *len2 = nullptr, *off2 = nullptr, *erow = nullptr,
           *ecol = nullptr, *perm1 = nullptr, *row1 = nullptr, *ord = nullptr,
           *flag = nullptr, *rid = nullptr, *runstart = nullptr, *rrow = nullptr,
           *rcol = nullptr, *rfirst = nullptr, *ordr = nullptr, *rbeg = nullptr,
           *cnt = nullptr, *crowptr = nullptr, *crowind = nullptr, *sind = nullptr;
  double *eval = nullptr, *rval = nullptr, *crowval = nullptr, *sval = nullptr;
  unsigned long long e64 = 0, *tot = nullptr;
  bool staged = false;
  hipError_t he = hipSuccess;
#define GB_SG_TRY(call)                                         \
  do {                                                          \
    he = (call);                                                \
    if (he != hipSuccess) {                                     \
      rv = gb_hip_fail(he, #call, __FILE__, __LINE__);          \
      goto done;                                                \
    }                                                           \
  } while (0)
#define GB_SG_RV(call)                                          \
  do {                                                          \
    rv = (call);                                                \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                    \
  } while (0)

  (void)ancols;
  if (dense && (uint64_t)anrows * bncols) {
    gb_sg_zero<<<gb_sg_grid((uint64_t)anrows * bncols), GB_BLOCK, 0, s>>>(
        (size_t)anrows * bncols, d_cdense);
    GB_SG_TRY(hipGetLastError());
  }
  GB_SG_TRY(hipMemcpyAsync(&annz, d_arowptr + anrows, 4, hipMemcpyDeviceToHost, s));
  GB_SG_TRY(hipStreamSynchronize(s));
  GB_SG_TRY(gb_dev_alloc((void **)&len, ((size_t)annz + 1) * 4));
  GB_SG_TRY(gb_dev_alloc((void **)&off, ((size_t)annz + 1) * 4));
  if (annz) {
    gb_sg_len<<<gb_sg_grid(annz), GB_BLOCK, 0, s>>>(annz, d_arowind, d_browptr, len);
    GB_SG_TRY(hipGetLastError());
  }
  // every product position must fit 32 bits: the total in 64 bits first
  GB_SG_TRY(gb_dev_alloc((void **)&tot, 8));
  GB_SG_RV(gbd_zero(tot, 8));
  if (annz) {
    gb_sg_sum64<<<gb_sg_grid(annz) > 1024u ? 1024u : gb_sg_grid(annz), GB_BLOCK, 0, s>>>(
        len, annz, tot);
    GB_SG_TRY(hipGetLastError());
  }
  GB_SG_TRY(hipMemcpyAsync(&e64, tot, 8, hipMemcpyDeviceToHost, s));
  GB_SG_TRY(hipStreamSynchronize(s));
  if (e64 > 0xfffffffeull) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  GB_SG_RV(gbd_scan_u32(len, off, annz, off + annz));
  // row classes and the list of wavefront rows
  GB_SG_TRY(gb_dev_alloc((void **)&isshort, ((size_t)anrows + 1) * 4));
  GB_SG_TRY(gb_dev_alloc((void **)&islong, ((size_t)anrows + 1) * 4));
  GB_SG_TRY(gb_dev_alloc((void **)&slist, ((size_t)anrows + 1) * 4));
  GB_SG_TRY(gb_dev_alloc((void **)&cnt, ((size_t)anrows + 1) * 4));
  if (anrows) {
    gb_sg_classify<<<gb_sg_grid(anrows), GB_BLOCK, 0, s>>>(anrows, d_arowptr, off, isshort,
                                                          islong);
    GB_SG_TRY(hipGetLastError());
  }
  // cnt holds the list positions until the counts replace them
  GB_SG_RV(gbd_scan_u32(isshort, cnt, anrows, cnt + anrows));
  GB_SG_TRY(hipMemcpyAsync(&nshort, cnt + anrows, 4, hipMemcpyDeviceToHost, s));
  GB_SG_TRY(hipStreamSynchronize(s));
  nlong = anrows - nshort;
  if (nshort) {
    gb_sg_compact<<<gb_sg_grid(anrows), GB_BLOCK, 0, s>>>(anrows, isshort, cnt, slist);
    GB_SG_TRY(hipGetLastError());
  }
  if (!dense) {
    GB_SG_RV(gbd_zero(cnt, ((size_t)anrows + 1) * 4));
    staged = e64 * 12 <= gb_sg_stage_max();
    if (nshort && staged) {
      GB_SG_TRY(gb_dev_alloc((void **)&sind, (size_t)e64 * 4));
      GB_SG_TRY(gb_dev_alloc((void **)&sval, (size_t)e64 * 8));
      gb_sg_wave<GB_SG_STAGE><<<gb_div_up(nshort, GB_WAVES_PER_BLOCK), GB_BLOCK, 0, s>>>(
          nshort, slist, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval,
          nullptr, cnt, sind, sval, nullptr, 0, off);
      GB_SG_TRY(hipGetLastError());
    } else if (nshort) {
      gb_sg_wave<GB_SG_COUNT><<<gb_div_up(nshort, GB_WAVES_PER_BLOCK), GB_BLOCK, 0, s>>>(
          nshort, slist, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval,
          nullptr, cnt, nullptr, nullptr, nullptr, 0, nullptr);
      GB_SG_TRY(hipGetLastError());
    }
  } else if (nshort) {
    gb_sg_wave<GB_SG_DENSE><<<gb_div_up(nshort, GB_WAVES_PER_BLOCK), GB_BLOCK, 0, s>>>(
        nshort, slist, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval,
        nullptr, nullptr, nullptr, nullptr, d_cdense, bncols, nullptr);
    GB_SG_TRY(hipGetLastError());
  }
  // the sort path's products, numbered over its rows
  if (nlong && annz) {
    GB_SG_TRY(gb_dev_alloc((void **)&len2, ((size_t)annz + 1) * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&off2, ((size_t)annz + 1) * 4));
    gb_sg_len_long<<<gb_sg_grid(annz), GB_BLOCK, 0, s>>>(anrows, annz, d_arowptr, islong, len,
                                                        len2);
    GB_SG_TRY(hipGetLastError());
    GB_SG_RV(gbd_scan_u32(len2, off2, annz, off2 + annz));
    GB_SG_TRY(hipMemcpyAsync(&E, off2 + annz, 4, hipMemcpyDeviceToHost, s));
    GB_SG_TRY(hipStreamSynchronize(s));
  }
  if (E) {
    GB_SG_TRY(gb_dev_alloc((void **)&erow, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&ecol, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&eval, (size_t)E * 8));
    GB_SG_TRY(gb_dev_alloc((void **)&perm1, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&row1, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&ord, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&flag, (size_t)E * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&rid, ((size_t)E + 1) * 4));
    gb_sg_expand<<<gb_sg_grid(annz), GB_BLOCK, 0, s>>>(
        anrows, annz, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval, off2,
        len2, erow, ecol, eval);
    GB_SG_TRY(hipGetLastError());
    // 1. (row, col, t) order: stable by column, then stable by row
    GB_SG_RV(gbd_countingsort_kv_unchecked(ecol, nullptr, 0, bncols ? bncols - 1 : 0, E,
                                           perm1, nullptr));
    gb_sg_gather_u32<<<gb_sg_grid(E), GB_BLOCK, 0, s>>>(E, perm1, erow, row1);
    GB_SG_TRY(hipGetLastError());
    GB_SG_RV(gbd_countingsort_kv_unchecked(row1, perm1, 0, anrows ? anrows - 1 : 0, E, ord,
                                           nullptr));
    // 2. runs of equal (row, col) and their sums
    gb_sg_runflag<<<gb_sg_grid(E), GB_BLOCK, 0, s>>>(E, ord, erow, ecol, flag);
    GB_SG_TRY(hipGetLastError());
    GB_SG_RV(gbd_scan_u32(flag, rid, E, rid + E));
    GB_SG_TRY(hipMemcpyAsync(&nruns, rid + E, 4, hipMemcpyDeviceToHost, s));
    GB_SG_TRY(hipStreamSynchronize(s));
    GB_SG_TRY(gb_dev_alloc((void **)&runstart, ((size_t)nruns + 1) * 4));
    gb_sg_runstart<<<gb_sg_grid(E), GB_BLOCK, 0, s>>>(E, flag, rid, runstart);
    GB_SG_TRY(hipGetLastError());
    if (dense) {
      gb_sg_runsum<true><<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(
          nruns, E, runstart, ord, erow, ecol, eval, bncols, d_cdense, nullptr, nullptr,
          nullptr, nullptr);
      GB_SG_TRY(hipGetLastError());
    } else {
      GB_SG_TRY(gb_dev_alloc((void **)&rrow, ((size_t)nruns + 1) * 4));
      GB_SG_TRY(gb_dev_alloc((void **)&rcol, ((size_t)nruns + 1) * 4));
      GB_SG_TRY(gb_dev_alloc((void **)&rfirst, ((size_t)nruns + 1) * 4));
      GB_SG_TRY(gb_dev_alloc((void **)&rval, ((size_t)nruns + 1) * 8));
      GB_SG_TRY(gb_dev_alloc((void **)&ordr, ((size_t)nruns + 1) * 4));
      GB_SG_TRY(gb_dev_alloc((void **)&rbeg, ((size_t)anrows + 1) * 4));
      gb_sg_runsum<false><<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(
          nruns, E, runstart, ord, erow, ecol, eval, bncols, nullptr, rrow, rcol, rfirst,
          rval);
      GB_SG_TRY(hipGetLastError());
      gb_sg_rowruns<<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(nruns, rrow, rbeg, cnt);
      GB_SG_TRY(hipGetLastError());
      // 3. first-appearance order: the first positions are unique and
      // row-major, so sorting by them alone orders the rows too
      GB_SG_RV(gbd_countingsort_kv_unchecked(rfirst, nullptr, 0, E - 1, nruns, ordr,
                                             nullptr));
    }
  }
  if (!dense) {
    GB_SG_TRY(gb_dev_alloc((void **)&crowptr, ((size_t)anrows + 1) * 4));
    GB_SG_RV(gbd_scan_u32(cnt, crowptr, anrows, crowptr + anrows));
    GB_SG_TRY(hipMemcpyAsync(&cnnz, crowptr + anrows, 4, hipMemcpyDeviceToHost, s));
    GB_SG_TRY(hipStreamSynchronize(s));
    GB_SG_TRY(gb_dev_alloc((void **)&crowind, ((size_t)cnnz + 1) * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&crowval, ((size_t)cnnz + 1) * 8));
    if (nshort && staged) {
      gb_sg_unstage<<<gb_div_up(nshort, GB_BLOCK), GB_BLOCK, 0, s>>>(
          nshort, slist, d_arowptr, off, cnt, sind, sval, crowptr, crowind, crowval);
      GB_SG_TRY(hipGetLastError());
    } else if (nshort) {
      gb_sg_wave<GB_SG_FILL><<<gb_div_up(nshort, GB_WAVES_PER_BLOCK), GB_BLOCK, 0, s>>>(
          nshort, slist, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval,
          crowptr, nullptr, crowind, crowval, nullptr, 0, nullptr);
      GB_SG_TRY(hipGetLastError());
    }
    if (nruns) {
      gb_sg_out<<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(nruns, ordr, rrow, rcol, rval, rbeg,
                                                       crowptr, crowind, crowval);
      GB_SG_TRY(hipGetLastError());
    }
    *d_crowptr = crowptr;
    *d_crowind = crowind;
    *d_crowval = crowval;
    *r_cnnz = cnnz;
    crowptr = crowind = nullptr;
    crowval = nullptr;
  }
  GB_SG_TRY(hipStreamSynchronize(s));
  rv = GOOSEBERRY_SUCCESS;
done:
  {
    void *tmp[] = {len, off, isshort, islong, slist, len2, off2, erow, ecol, eval,
                   perm1, row1, ord, flag, rid, runstart, rrow, rcol, rfirst, rval,
                   ordr, rbeg, cnt, crowptr, crowind, crowval, tot, sind, sval};
    for (void *q : tmp) (void)gb_dev_free(q);
  }
  return rv;
#undef GB_SG_TRY
#undef GB_SG_RV
}

}  // extern "C"
