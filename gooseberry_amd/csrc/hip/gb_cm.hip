// gb_cm.hip -- Cuthill-McKee ordering on the GPU (src/permute.c:49-110,
// 288-332): perm[p] = the row placed at position p, per diagonal block.
//
// The reference, per block [start, end): deg[v] = v's entries with a column
// in the block; the block's vertices wait in a min-priority queue by
// degree; each component starts at the least-degree unplaced vertex, and
// the vertices are then processed in placement order, each placing its
// unplaced in-block neighbours by ascending degree.  Equal degrees are
// taken in ascending vertex order here (the reference's order among them
// is its priority queue's, which the reference tree does not hold).
//
// Processing a contiguous range of placed-but-unprocessed positions
// [sr, nordered) at once is exact: each newly reached vertex belongs to
// the lowest position in the range that reaches it, and the new vertices
// follow in (that position, degree, vertex) order -- what processing the
// positions one by one appends.  Large ranges take that parallel step
// (claim by atomicMin, two stable radix sorts, placement); small ones are
// advanced vertex by vertex by one workgroup until the range grows, which
// keeps long thin graphs from paying one launch per vertex.  Component
// starts come from the vertices sorted once by (degree, vertex): the next
// start is the first unplaced one; a run of vertices with no in-block
// neighbour but themselves is placed in one step.
#include <stdlib.h>

#include "gb_common.h"

#define GB_CM_NONE 0xffffffffu
#define GB_CM_BATCH 64       // ranges at least this long take the parallel step
#define GB_CM_CAND 1024      // neighbours a vertex may place in the serial kernel
#define GB_CM_SERIAL_CAP 65536  // vertices per serial launch

// deg[v]: in-block entries of row start + v; iso[v]: none of them is
// another vertex
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_deg(uint32_t start, uint32_t end, const uint32_t *__restrict__ rowptr,
          const uint32_t *__restrict__ rowind, uint32_t *__restrict__ deg,
          uint32_t *__restrict__ iso) {
  const uint32_t nb = end - start;
  for (uint32_t v = blockIdx.x * GB_BLOCK + threadIdx.x; v < nb; v += gridDim.x * GB_BLOCK) {
    const uint32_t r = start + v;
    uint32_t d = 0, other = 0;
    for (uint32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const uint32_t k = rowind[j];
      if (k >= start && k < end) {
        ++d;
        if (k != r) other = 1;
      }
    }
    deg[v] = d;
    iso[v] = other ? 0u : 1u;
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_fill(uint32_t n, uint32_t *__restrict__ a, uint32_t value) {
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n; i += gridDim.x * GB_BLOCK)
    a[i] = value;
}

// srank[S[r]] = r
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_rank(uint32_t n, const uint32_t *__restrict__ S, uint32_t *__restrict__ srank) {
  for (uint32_t r = blockIdx.x * GB_BLOCK + threadIdx.x; r < n; r += gridDim.x * GB_BLOCK)
    srank[S[r]] = r;
}

// *res = the least r in [c, n) with S[r] unplaced (and, with NONISO, not
// isolated); res starts at n
template <bool NONISO>
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_find(uint32_t c, uint32_t n, const uint32_t *__restrict__ S,
           const uint32_t *__restrict__ posof, const uint32_t *__restrict__ iso,
           uint32_t *__restrict__ res) {
  __shared__ uint32_t best;
  if (threadIdx.x == 0) best = GB_CM_NONE;
  __syncthreads();
  for (uint32_t r = c + blockIdx.x * GB_BLOCK + threadIdx.x; r < n; r += gridDim.x * GB_BLOCK) {
    const uint32_t v = S[r];
    if (posof[v] == GB_CM_NONE && (!NONISO || !iso[v])) {
      atomicMin(&best, r);
      break;  // later r of this thread are larger
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && best != GB_CM_NONE) atomicMin(res, best);
}

// flag[r - c] = S[r] unplaced, for r in [c, f)
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_isoflag(uint32_t c, uint32_t f, const uint32_t *__restrict__ S,
              const uint32_t *__restrict__ posof, uint32_t *__restrict__ flag) {
  for (uint32_t r = c + blockIdx.x * GB_BLOCK + threadIdx.x; r < f; r += gridDim.x * GB_BLOCK)
    flag[r - c] = posof[S[r]] == GB_CM_NONE ? 1u : 0u;
}

// the flagged vertices of S[c, f) placed in S order from position p0
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_isoplace(uint32_t c, uint32_t f, uint32_t start, uint32_t p0,
               const uint32_t *__restrict__ S, const uint32_t *__restrict__ flag,
               const uint32_t *__restrict__ rank, uint32_t *__restrict__ posof,
               uint32_t *__restrict__ perm) {
  for (uint32_t r = c + blockIdx.x * GB_BLOCK + threadIdx.x; r < f; r += gridDim.x * GB_BLOCK) {
    if (flag[r - c]) {
      const uint32_t p = p0 + rank[r - c];
      const uint32_t v = S[r];
      perm[p] = start + v;
      posof[v] = p;
    }
  }
}

__global__ void gb_cm_set(uint32_t *__restrict__ d, uint32_t a, uint32_t b, uint32_t c) {
  d[0] = a;
  d[1] = b;
  d[2] = c;
}

// one vertex (local v) placed at position p
__global__ void gb_cm_place1(uint32_t v, uint32_t p, uint32_t start,
                             uint32_t *__restrict__ posof, uint32_t *__restrict__ perm) {
  posof[v] = p;
  perm[p] = start + v;
}

// parallel step over positions [b0, b1): every unplaced in-block neighbour
// keeps the least position reaching it; the first claim lists it
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_claim(uint32_t b0, uint32_t b1, uint32_t start, uint32_t end,
            const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ rowind,
            const uint32_t *__restrict__ perm, const uint32_t *__restrict__ posof,
            uint32_t *__restrict__ claim, uint32_t *__restrict__ list,
            uint32_t *__restrict__ nlist) {
  for (uint32_t p = b0 + blockIdx.x * GB_BLOCK + threadIdx.x; p < b1;
       p += gridDim.x * GB_BLOCK) {
    const uint32_t r = perm[p];
    for (uint32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const uint32_t k = rowind[j];
      if (k < start || k >= end) continue;
      const uint32_t kl = k - start;
      if (posof[kl] != GB_CM_NONE) continue;
      if (atomicMin(&claim[kl], p) == GB_CM_NONE) list[atomicAdd(nlist, 1u)] = kl;
    }
  }
}

// keys for the sorts: out[t] = srank[list[t]] (KIND 0) or
// claim[list[t]] - b0 (KIND 1)
template <int KIND>
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_keys(uint32_t L, const uint32_t *__restrict__ list, const uint32_t *__restrict__ src,
           uint32_t b0, uint32_t *__restrict__ out) {
  for (uint32_t t = blockIdx.x * GB_BLOCK + threadIdx.x; t < L; t += gridDim.x * GB_BLOCK)
    out[t] = KIND == 0 ? src[list[t]] : src[list[t]] - b0;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_placelist(uint32_t L, uint32_t p0, uint32_t start, const uint32_t *__restrict__ list,
                uint32_t *__restrict__ posof, uint32_t *__restrict__ perm) {
  for (uint32_t t = blockIdx.x * GB_BLOCK + threadIdx.x; t < L; t += gridDim.x * GB_BLOCK) {
    const uint32_t v = list[t];
    posof[v] = p0 + t;
    perm[p0 + t] = start + v;
  }
}

// A step's new vertices (at most GB_CM_LDS_SORT) sorted in one workgroup:
// keys (claiming position - b0, srank) are distinct, so a bitonic sort in
// LDS gives the exact order; the vertex is S[srank]
#define GB_CM_LDS_SORT 8192
#define GB_CM_SORT_THREADS 1024

__device__ void gb_cm_bitonic_place(uint32_t L, uint32_t b0, uint32_t p0, uint32_t start,
                                    const uint32_t *__restrict__ list,
                                    const uint32_t *__restrict__ claim,
                                    const uint32_t *__restrict__ srank,
                                    const uint32_t *__restrict__ S,
                                    uint32_t *__restrict__ posof, uint32_t *__restrict__ perm,
                                    uint64_t *key) {
  uint32_t P2 = 1;
  while (P2 < L) P2 <<= 1;
  for (uint32_t t = threadIdx.x; t < P2; t += GB_CM_SORT_THREADS) {
    uint64_t k = ~0ull;
    if (t < L) {
      const uint32_t v = list[t];
      k = ((uint64_t)(claim[v] - b0) << 32) | srank[v];
    }
    key[t] = k;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= P2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P2; i += GB_CM_SORT_THREADS) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint64_t a = key[i], b = key[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            key[i] = b;
            key[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t t = threadIdx.x; t < L; t += GB_CM_SORT_THREADS) {
    const uint32_t v = S[(uint32_t)key[t]];
    posof[v] = p0 + t;
    perm[p0 + t] = start + v;
  }
}

__global__ void __launch_bounds__(GB_CM_SORT_THREADS)
gb_cm_sortplace(uint32_t L, uint32_t b0, uint32_t p0, uint32_t start,
                const uint32_t *__restrict__ list, const uint32_t *__restrict__ claim,
                const uint32_t *__restrict__ srank, const uint32_t *__restrict__ S,
                uint32_t *__restrict__ posof, uint32_t *__restrict__ perm) {
  __shared__ uint64_t key[GB_CM_LDS_SORT];
  gb_cm_bitonic_place(L, b0, p0, start, list, claim, srank, S, posof, perm, key);
}

// A step's placement by buckets, for at most GB_CM_LDS_SORT parents
// [b0, b1) and new vertices, all in LDS: each new vertex goes to the bucket
// of its claiming position (counts by LDS atomics, their scan, a scatter),
// and each bucket is put in srank order by one thread (insertion sort).
// Parents keep their order and siblings go by srank: the (claiming
// position, srank) order of the sort.  A bucket of more than
// GB_CM_SEG_BUCKET vertices sends the step to the sort instead.
#define GB_CM_SEG_BUCKET 32
__global__ void __launch_bounds__(GB_CM_SORT_THREADS)
gb_cm_segplace(uint32_t L, uint32_t b0, uint32_t b1, uint32_t p0, uint32_t start,
               const uint32_t *__restrict__ list, const uint32_t *__restrict__ claim,
               const uint32_t *__restrict__ srank, const uint32_t *__restrict__ S,
               uint32_t *__restrict__ posof, uint32_t *__restrict__ perm) {
  __shared__ uint64_t key[GB_CM_LDS_SORT];  // the sort's keys, if it is needed
  // per bucket: its count, then its first slot, then (slots taken) its end
  __shared__ uint32_t base[GB_CM_LDS_SORT];
  __shared__ uint32_t out[GB_CM_LDS_SORT];  // sranks in bucket order
  __shared__ uint32_t big;
  const uint32_t np = b1 - b0;
  for (uint32_t q = threadIdx.x; q < np; q += GB_CM_SORT_THREADS) base[q] = 0;
  if (threadIdx.x == 0) big = 0;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < L; t += GB_CM_SORT_THREADS) {
    const uint32_t v = list[t];
    const uint32_t c = claim[v] - b0;  // < np: v was claimed in this step
    const uint32_t r = srank[v];
    key[t] = ((uint64_t)c << 32) | r;
    const uint32_t had = atomicAdd(&base[c], 1u);
    if (had == GB_CM_SEG_BUCKET) big = 1;
  }
  __syncthreads();
  if (big) {  // uniform: read after the barrier
    gb_cm_bitonic_place(L, b0, p0, start, list, claim, srank, S, posof, perm, key);
    return;
  }
  // exclusive scan of the bucket counts by the first wavefront
  if (threadIdx.x < GB_WAVE) {
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < np; c0 += GB_WAVE) {
      const uint32_t q = c0 + threadIdx.x;
      const uint32_t n = q < np ? base[q] : 0u;
      const uint32_t inc = gb_wave_scan_incl(n);
      if (q < np) base[q] = carry + inc - n;
      carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, GB_WAVE - 1);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < L; t += GB_CM_SORT_THREADS) {
    const uint64_t k = key[t];
    const uint32_t c = (uint32_t)(k >> 32);
    out[atomicAdd(&base[c], 1u)] = (uint32_t)k;
  }
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < np; q += GB_CM_SORT_THREADS) {
    const uint32_t a = q ? base[q - 1] : 0u, n = base[q] - a;
    for (uint32_t x = 1; x < n; ++x) {  // at most GB_CM_SEG_BUCKET
      const uint32_t r = out[a + x];
      uint32_t y = x;
      while (y > 0 && out[a + y - 1] > r) {
        out[a + y] = out[a + y - 1];
        --y;
      }
      out[a + y] = r;
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < L; t += GB_CM_SORT_THREADS) {
    const uint32_t v = S[out[t]];
    posof[v] = p0 + t;
    perm[p0 + t] = start + v;
  }
}

// One workgroup processes positions one by one from st[0] (sr) while
// st[1] (nordered) - sr < GB_CM_BATCH.  Stops with st[2] = 0 when the
// component is done, 1 when the range has grown, 2 before a vertex whose
// row is longer than GB_CM_CAND (the parallel step takes it), 3 at the
// launch's vertex cap.
__global__ void __launch_bounds__(GB_BLOCK)
gb_cm_serial(uint32_t start, uint32_t end, const uint32_t *__restrict__ rowptr,
             const uint32_t *__restrict__ rowind, const uint32_t *__restrict__ srank,
             uint32_t *__restrict__ posof, uint32_t *__restrict__ claim,
             uint32_t *__restrict__ perm, uint32_t *__restrict__ st) {
  __shared__ uint32_t s_sr, s_no, s_m, s_why;
  __shared__ uint32_t cand[GB_CM_CAND], ckey[GB_CM_CAND];
  if (threadIdx.x == 0) {
    s_sr = st[0];
    s_no = st[1];
    s_why = 3;
  }
  __syncthreads();
  for (uint32_t it = 0; it < GB_CM_SERIAL_CAP; ++it) {
    const uint32_t sr = s_sr, no = s_no;
    if (sr >= no || no - sr >= GB_CM_BATCH) {
      if (threadIdx.x == 0) s_why = sr >= no ? 0u : 1u;
      break;
    }
    const uint32_t r = perm[sr];
    const uint32_t j0 = rowptr[r], j1 = rowptr[r + 1];
    if (j1 - j0 > GB_CM_CAND) {
      if (threadIdx.x == 0) s_why = 2;
      break;
    }
    if (threadIdx.x == 0) s_m = 0;
    __syncthreads();
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += GB_BLOCK) {
      const uint32_t k = rowind[j];
      if (k >= start && k < end) {
        const uint32_t kl = k - start;
        if (posof[kl] == GB_CM_NONE && atomicCAS(&claim[kl], GB_CM_NONE, sr) == GB_CM_NONE) {
          const uint32_t slot = atomicAdd(&s_m, 1u);
          cand[slot] = kl;
          ckey[slot] = srank[kl];
        }
      }
    }
    __syncthreads();
    const uint32_t m = s_m;
    // the new vertices in (degree, vertex) order: their srank order
    for (uint32_t t = threadIdx.x; t < m; t += GB_BLOCK) {
      const uint32_t key = ckey[t];
      uint32_t rank = 0;
      for (uint32_t u = 0; u < m; ++u) rank += ckey[u] < key ? 1u : 0u;
      posof[cand[t]] = no + rank;
      perm[no + rank] = start + cand[t];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_sr = sr + 1;
      s_no = no + m;
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st[0] = s_sr;
    st[1] = s_no;
    st[2] = s_why;
  }
}

static uint32_t gb_cm_grid(uint64_t n) {
  uint64_t g = (n + GB_BLOCK - 1) / GB_BLOCK;
  if (g > 8192) g = 8192;
  return g ? (uint32_t)g : 1u;
}

// one diagonal block [start, end) into perm[start, end)
static int gb_cm_block(uint32_t start, uint32_t end, const uint32_t *rowptr,
                       const uint32_t *rowind, uint32_t *perm) {
  const uint32_t nb = end - start;
  if (nb == 0) return GOOSEBERRY_SUCCESS;
  hipStream_t s = gb_stream();
  int rv = GOOSEBERRY_SUCCESS;
  hipError_t he = hipSuccess;
  uint32_t *deg = nullptr, *iso = nullptr, *S = nullptr, *srank = nullptr, *posof = nullptr,
           *claim = nullptr, *list = nullptr, *list2 = nullptr, *keys = nullptr,
           *dv = nullptr, *flag = nullptr, *frank = nullptr, maxdeg = 0;
  uint32_t h[4];
  uint32_t placed = 0, cursor = 0;
  // device scalars dv: [0] find result; [1] list length, or [1..3] the
  // serial kernel's state
#define GB_CM_TRY(call)                                         \
  do {                                                          \
    he = (call);                                                \
    if (he != hipSuccess) {                                     \
      rv = gb_hip_fail(he, #call, __FILE__, __LINE__);          \
      goto done;                                                \
    }                                                           \
  } while (0)
#define GB_CM_RV(call)                                          \
  do {                                                          \
    rv = (call);                                                \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                    \
  } while (0)
  GB_CM_TRY(gb_dev_alloc((void **)&deg, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&iso, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&S, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&srank, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&posof, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&claim, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&list, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&list2, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&keys, (size_t)nb * 4));
  GB_CM_TRY(gb_dev_alloc((void **)&flag, (size_t)nb * 4 + 4));
  GB_CM_TRY(gb_dev_alloc((void **)&frank, (size_t)nb * 4 + 4));
  GB_CM_TRY(gb_dev_alloc((void **)&dv, 16));
  gb_cm_deg<<<gb_cm_grid(nb), GB_BLOCK, 0, s>>>(start, end, rowptr, rowind, deg, iso);
  GB_CM_TRY(hipGetLastError());
  // the least-degree order of the block's vertices, equal degrees by
  // vertex; no degree exceeds the block rows' entry count
  GB_CM_TRY(hipMemcpyAsync(&h[0], rowptr + start, 4, hipMemcpyDeviceToHost, s));
  GB_CM_TRY(hipMemcpyAsync(&h[1], rowptr + end, 4, hipMemcpyDeviceToHost, s));
  GB_CM_TRY(hipStreamSynchronize(s));
  maxdeg = h[1] - h[0];
  GB_CM_RV(gbd_countingsort_kv_unchecked(deg, nullptr, 0, maxdeg, nb, S, nullptr));
  gb_cm_rank<<<gb_cm_grid(nb), GB_BLOCK, 0, s>>>(nb, S, srank);
  GB_CM_TRY(hipGetLastError());
  gb_cm_fill<<<gb_cm_grid(nb), GB_BLOCK, 0, s>>>(nb, posof, GB_CM_NONE);
  GB_CM_TRY(hipGetLastError());
  gb_cm_fill<<<gb_cm_grid(nb), GB_BLOCK, 0, s>>>(nb, claim, GB_CM_NONE);
  GB_CM_TRY(hipGetLastError());

  while (placed < nb) {
    // the next start: the first unplaced vertex in (degree, vertex) order
    gb_cm_set<<<1, 1, 0, s>>>(dv, nb, 0, 0);
    GB_CM_TRY(hipGetLastError());
    gb_cm_find<false><<<gb_cm_grid(nb - cursor), GB_BLOCK, 0, s>>>(cursor, nb, S, posof, iso,
                                                                   dv);
    GB_CM_TRY(hipGetLastError());
    GB_CM_TRY(hipMemcpyAsync(&h[1], dv, 4, hipMemcpyDeviceToHost, s));
    GB_CM_TRY(hipStreamSynchronize(s));
    const uint32_t idx = h[1];
    if (idx >= nb) {  // cannot happen while vertices remain
      rv = GOOSEBERRY_ERROR_UNKNOWN;
      goto done;
    }
    uint32_t sv = 0, siso = 0;
    GB_CM_TRY(hipMemcpyAsync(&sv, S + idx, 4, hipMemcpyDeviceToHost, s));
    GB_CM_TRY(hipStreamSynchronize(s));
    GB_CM_TRY(hipMemcpyAsync(&siso, iso + sv, 4, hipMemcpyDeviceToHost, s));
    GB_CM_TRY(hipStreamSynchronize(s));
    if (siso) {
      // a start with no neighbour is a component of its own, and so is
      // every following unplaced one up to the first that has neighbours
      gb_cm_set<<<1, 1, 0, s>>>(dv, nb, 0, 0);
      GB_CM_TRY(hipGetLastError());
      gb_cm_find<true><<<gb_cm_grid(nb - idx), GB_BLOCK, 0, s>>>(idx, nb, S, posof, iso, dv);
      GB_CM_TRY(hipGetLastError());
      GB_CM_TRY(hipMemcpyAsync(&h[1], dv, 4, hipMemcpyDeviceToHost, s));
      GB_CM_TRY(hipStreamSynchronize(s));
      const uint32_t f = h[1];
      gb_cm_isoflag<<<gb_cm_grid(f - idx), GB_BLOCK, 0, s>>>(idx, f, S, posof, flag);
      GB_CM_TRY(hipGetLastError());
      GB_CM_RV(gbd_scan_u32(flag, frank, f - idx, frank + (f - idx)));
      uint32_t cnt = 0;
      GB_CM_TRY(hipMemcpyAsync(&cnt, frank + (f - idx), 4, hipMemcpyDeviceToHost, s));
      gb_cm_isoplace<<<gb_cm_grid(f - idx), GB_BLOCK, 0, s>>>(idx, f, start, start + placed, S,
                                                             flag, frank, posof, perm);
      GB_CM_TRY(hipGetLastError());
      GB_CM_TRY(hipStreamSynchronize(s));
      placed += cnt;
      cursor = f;
      continue;
    }
    gb_cm_place1<<<1, 1, 0, s>>>(sv, start + placed, start, posof, perm);
    GB_CM_TRY(hipGetLastError());
    cursor = idx + 1;
    uint32_t sr = start + placed, no = start + placed + 1;
    for (;;) {
      if (no - sr < GB_CM_BATCH) {
        gb_cm_set<<<1, 1, 0, s>>>(dv + 1, sr, no, 0);
        GB_CM_TRY(hipGetLastError());
        gb_cm_serial<<<1, GB_BLOCK, 0, s>>>(start, end, rowptr, rowind, srank, posof, claim,
                                            perm, dv + 1);
        GB_CM_TRY(hipGetLastError());
        GB_CM_TRY(hipMemcpyAsync(h, dv + 1, 12, hipMemcpyDeviceToHost, s));
        GB_CM_TRY(hipStreamSynchronize(s));
        sr = h[0];
        no = h[1];
        if (h[2] == 0) break;     // component done
        if (h[2] == 3) continue;  // vertex cap: go on serially
        // 1: the range has grown; 2: a long row: the parallel step next
      }
      // the parallel step over [sr, no)
      gb_cm_set<<<1, 1, 0, s>>>(dv + 1, 0, 0, 0);
      GB_CM_TRY(hipGetLastError());
      gb_cm_claim<<<gb_cm_grid(no - sr), GB_BLOCK, 0, s>>>(sr, no, start, end, rowptr, rowind,
                                                          perm, posof, claim, list, dv + 1);
      GB_CM_TRY(hipGetLastError());
      uint32_t L = 0;
      GB_CM_TRY(hipMemcpyAsync(&L, dv + 1, 4, hipMemcpyDeviceToHost, s));
      GB_CM_TRY(hipStreamSynchronize(s));
      if (L && L <= GB_CM_LDS_SORT && no - sr <= GB_CM_LDS_SORT) {
        gb_cm_segplace<<<1, GB_CM_SORT_THREADS, 0, s>>>(L, sr, no, no, start, list, claim,
                                                       srank, S, posof, perm);
        GB_CM_TRY(hipGetLastError());
      } else if (L && L <= GB_CM_LDS_SORT) {
        gb_cm_sortplace<<<1, GB_CM_SORT_THREADS, 0, s>>>(L, sr, no, start, list, claim, srank,
                                                        S, posof, perm);
        GB_CM_TRY(hipGetLastError());
      } else if (L) {
        // (position, degree, vertex) order: by srank, then stably by position
        gb_cm_keys<0><<<gb_cm_grid(L), GB_BLOCK, 0, s>>>(L, list, srank, 0, keys);
        GB_CM_TRY(hipGetLastError());
        GB_CM_RV(gbd_countingsort_kv_unchecked(keys, list, 0, nb - 1, L, list2, nullptr));
        gb_cm_keys<1><<<gb_cm_grid(L), GB_BLOCK, 0, s>>>(L, list2, claim, sr, keys);
        GB_CM_TRY(hipGetLastError());
        GB_CM_RV(gbd_countingsort_kv_unchecked(keys, list2, 0, no - sr - 1, L, list, nullptr));
        gb_cm_placelist<<<gb_cm_grid(L), GB_BLOCK, 0, s>>>(L, no, start, list, posof, perm);
        GB_CM_TRY(hipGetLastError());
      }
      sr = no;
      no += L;
      if (sr == no) break;
    }
    placed = no - start;
  }
  GB_CM_TRY(hipStreamSynchronize(s));
  rv = GOOSEBERRY_SUCCESS;
done:
  (void)gb_dev_free(deg); (void)gb_dev_free(iso); (void)gb_dev_free(S); (void)gb_dev_free(srank);
  (void)gb_dev_free(posof); (void)gb_dev_free(claim); (void)gb_dev_free(list); (void)gb_dev_free(list2);
  (void)gb_dev_free(keys); (void)gb_dev_free(flag); (void)gb_dev_free(frank); (void)gb_dev_free(dv);
  return rv;
#undef GB_CM_TRY
#undef GB_CM_RV
}

extern "C" {

int gbd_cuthillmckee(uint32_t nrows, const uint32_t *d_rowptr, const uint32_t *d_rowind,
                     const uint32_t *blocks, uint32_t nblocks, uint32_t *d_perm) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (blocks) {
    for (uint32_t c = 0; c < nblocks; ++c) {
      rv = gb_cm_block(blocks[c], blocks[c + 1], d_rowptr, d_rowind, d_perm);
      if (rv != GOOSEBERRY_SUCCESS) return rv;
    }
    return GOOSEBERRY_SUCCESS;
  }
  return gb_cm_block(0, nrows, d_rowptr, d_rowind, d_perm);
}

}  // extern "C"
