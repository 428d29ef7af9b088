// gb_index.hip -- integer index work on the GPU: exclusive scan, stable
// counting sort (LSD radix over 8-bit digits for wide key ranges) and the
// CSR / dense permutations of src/permute.c.
#include <stdio.h>

#include "gb_common.h"

#define SCAN_ITEMS 8
#define SCAN_TILE (GB_BLOCK * SCAN_ITEMS)

// ----------------------------------------------------------------- scan

__global__ void __launch_bounds__(GB_BLOCK)
gb_scan_tile_sums(const uint32_t *__restrict__ in, size_t n,
                  uint32_t *__restrict__ sums) {
  __shared__ uint32_t lds4[GB_WAVES_PER_BLOCK];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    size_t k = base + (size_t)i * GB_BLOCK + threadIdx.x;
    if (k < n) s += in[k];
  }
  s = gb_block_sum(s, lds4);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

// One block: exclusive scan of sums[0..nb) in place; *total = grand total.
__global__ void __launch_bounds__(GB_BLOCK)
gb_scan_top(uint32_t *__restrict__ sums, size_t nb, uint32_t *total) {
  __shared__ uint32_t lds4[GB_WAVES_PER_BLOCK];
  uint32_t carry = 0;
  for (size_t base = 0; base < nb; base += GB_BLOCK) {
    size_t k = base + threadIdx.x;
    uint32_t v = k < nb ? sums[k] : 0u;
    uint32_t tot;
    uint32_t ex = gb_block_scan_excl(v, lds4, &tot);
    if (k < nb) sums[k] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// Each thread scans SCAN_ITEMS consecutive entries of the tile.
__global__ void __launch_bounds__(GB_BLOCK)
gb_scan_tile_apply(const uint32_t *in, uint32_t *out, size_t n,
                   const uint32_t *__restrict__ sums) {
  __shared__ uint32_t lds4[GB_WAVES_PER_BLOCK];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE +
                      (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    size_t k = base + i;
    v[i] = k < n ? in[k] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t run = sums[blockIdx.x] + gb_block_scan_excl(s, lds4, &tot);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    size_t k = base + i;
    if (k < n) out[k] = run;
    run += v[i];
  }
}

static int gb_scan_u32(const uint32_t *d_in, uint32_t *d_out, size_t n,
                       uint32_t *d_total) {
  hipStream_t s = gb_stream();
  if (n == 0) {
    if (d_total) GB_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
    return GOOSEBERRY_SUCCESS;
  }
  size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint32_t *sums = nullptr;
  GB_HIP(gb_dev_alloc((void **)&sums, nb * sizeof(uint32_t)));
  gb_scan_tile_sums<<<(uint32_t)nb, GB_BLOCK, 0, s>>>(d_in, n, sums);
  GB_LAUNCH_CHECK();
  gb_scan_top<<<1, GB_BLOCK, 0, s>>>(sums, nb, d_total);
  GB_LAUNCH_CHECK();
  gb_scan_tile_apply<<<(uint32_t)nb, GB_BLOCK, 0, s>>>(d_in, d_out, n, sums);
  GB_LAUNCH_CHECK();
  GB_HIP(gb_dev_free(sums));
  return GOOSEBERRY_SUCCESS;
}

// ------------------------------------------------------ counting sort

#define CS_RADIX 256                  // 8-bit digits
#define CS_ITEMS 16                   // keys per thread
#define CS_TILE (GB_BLOCK * CS_ITEMS)  // keys per workgroup tile (4096)
static_assert(GB_BLOCK == CS_RADIX, "one thread per digit in the sort's kernels");

__device__ __forceinline__ uint32_t cs_digit(uint32_t k, uint32_t kmin,
                                             int shift) {
  return ((k - kmin) >> shift) & (CS_RADIX - 1);
}

// Digit counts of each tile of CS_TILE consecutive keys, stored digit-major
// (counts[d * ntiles + tile]) so that one exclusive scan gives every
// (digit, tile) its first output position in stable order.
__global__ void __launch_bounds__(GB_BLOCK)
gb_cs_count(const uint32_t *__restrict__ keys, size_t n, uint32_t kmin,
            int shift, uint32_t *__restrict__ counts, uint32_t ntiles) {
  __shared__ uint32_t cnt[CS_RADIX];
  cnt[threadIdx.x] = 0;  // GB_BLOCK == CS_RADIX
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * CS_TILE;
#pragma unroll
  for (int i = 0; i < CS_ITEMS; ++i) {
    const size_t k = base + (size_t)i * GB_BLOCK + threadIdx.x;
    if (k < n) atomicAdd(&cnt[cs_digit(keys[k], kmin, shift)], 1u);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// One tile per workgroup.  Key i of thread t is key base + i * GB_BLOCK + t,
// so the keys in (i, t) order are the tile's keys in order.  Step i ranks
// its GB_BLOCK keys: within a wavefront by the mask of lanes with the same
// digit (eight ballots), across wavefronts by the per-wave digit counts in
// LDS, across steps by each digit's running count.  A scan of the tile's
// digit counts then places the keys in LDS in (digit, order) order, and
// thread t writes the LDS keys t, t + GB_BLOCK, ... to their places: a
// digit's keys of the tile go to consecutive places, so the writes come in
// runs (CS_TILE / CS_RADIX = 16 keys on average) instead of one place per
// key.
__global__ void __launch_bounds__(GB_BLOCK)
gb_cs_scatter(const uint32_t *__restrict__ keys,
              const uint32_t *__restrict__ vals, size_t n, uint32_t kmin,
              int shift, const uint32_t *__restrict__ offs, uint32_t ntiles,
              uint32_t *__restrict__ okeys, uint32_t *__restrict__ ovals) {
  __shared__ uint32_t wcnt[GB_WAVES_PER_BLOCK][CS_RADIX];
  __shared__ uint32_t run[CS_RADIX];     // keys of each digit ranked so far
  __shared__ uint32_t dstart[CS_RADIX];  // first LDS place of each digit
  __shared__ uint32_t gbase[CS_RADIX];   // first output place of each digit
  __shared__ uint32_t skey[CS_TILE], sval[CS_TILE];
  __shared__ uint32_t lds4[GB_WAVES_PER_BLOCK];
  const uint32_t tid = threadIdx.x;
  const int lane = tid & (GB_WAVE - 1);
  const int wave = tid / GB_WAVE;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  gbase[tid] = offs[(size_t)tid * ntiles + blockIdx.x];
  run[tid] = 0;
  const size_t base = (size_t)blockIdx.x * CS_TILE;
  uint32_t key[CS_ITEMS], val[CS_ITEMS], pos[CS_ITEMS];
#pragma unroll
  for (int i = 0; i < CS_ITEMS; ++i) {
    const size_t k = base + (size_t)i * GB_BLOCK + tid;
    key[i] = k < n ? keys[k] : kmin;
    val[i] = k < n ? (vals ? vals[k] : (uint32_t)k) : 0u;
  }
#pragma unroll
  for (int i = 0; i < CS_ITEMS; ++i) {
#pragma unroll
    for (int w = 0; w < GB_WAVES_PER_BLOCK; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const bool valid = base + (size_t)i * GB_BLOCK + tid < n;
    const uint32_t d = cs_digit(key[i], kmin, shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bb = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    if (valid && (63 - __clzll(peers)) == lane) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t o = run[d];
      for (int w = 0; w < wave; ++w) o += wcnt[w][d];
      pos[i] = o + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < GB_WAVES_PER_BLOCK; ++w) add += wcnt[w][tid];
    run[tid] += add;
  }
  __syncthreads();
  uint32_t total = 0;
  dstart[tid] = gb_block_scan_excl(run[tid], lds4, &total);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CS_ITEMS; ++i) {
    if (base + (size_t)i * GB_BLOCK + tid < n) {
      const uint32_t p = dstart[cs_digit(key[i], kmin, shift)] + pos[i];
      skey[p] = key[i];
      sval[p] = val[i];
    }
  }
  __syncthreads();
  const uint32_t tn = n - base < (size_t)CS_TILE ? (uint32_t)(n - base) : (uint32_t)CS_TILE;
  for (uint32_t j = tid; j < tn; j += GB_BLOCK) {
    const uint32_t k = skey[j];
    const uint32_t d = cs_digit(k, kmin, shift);
    const uint32_t o = gbase[d] + (j - dstart[d]);
    okeys[o] = k;
    ovals[o] = sval[j];
  }
}

// dist[k] = the number of sorted keys below kmin + k, for k in [0, range]
// (range = kmax - kmin + 1): a lower bound in the sorted keys per k
__global__ void __launch_bounds__(GB_BLOCK)
gb_cs_dist(const uint32_t *__restrict__ skeys, size_t n, uint32_t kmin, uint64_t range,
           uint32_t *__restrict__ dist) {
  for (uint64_t k = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; k <= range;
       k += (uint64_t)gridDim.x * GB_BLOCK) {
    size_t lo = 0, hi = n;  // first position whose key is >= kmin + k
    while (lo < hi) {
      const size_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)(skeys[mid] - kmin) < k) lo = mid + 1;
      else hi = mid;
    }
    dist[k] = (uint32_t)lo;
  }
}

__global__ void gb_cs_check_keys(const uint32_t *__restrict__ keys, size_t n,
                                 uint32_t kmin, uint32_t kmax,
                                 uint32_t *__restrict__ bad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  uint32_t b = 0;
  for (; i < n; i += stride) {
    uint32_t k = keys[i];
    b |= (k < kmin || k > kmax) ? 1u : 0u;
  }
  if (b) atomicOr(bad, 1u);
}

static uint32_t gb_grid_stride_blocks(size_t n) {
  uint32_t g = gb_div_up(n ? n : 1, GB_BLOCK);
  return g > 8192u ? 8192u : g;
}

// --------------------------------------------------------- permutations

__global__ void gb_perm_rowlen(const uint32_t *__restrict__ rowptr,
                               const uint32_t *__restrict__ rowperm,
                               uint32_t nrows, uint32_t *__restrict__ len) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrows) {
    uint32_t v = rowperm[i];
    len[i] = rowptr[v + 1] - rowptr[v];
  }
}

__global__ void gb_perm_rename(const uint32_t *__restrict__ colperm,
                               uint32_t ncols, uint32_t *__restrict__ rename) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncols) rename[colperm[i]] = i;
}

__global__ void gb_perm_rename_apply(const uint32_t *__restrict__ in,
                                     uint32_t *__restrict__ out, size_t nnz,
                                     const uint32_t *__restrict__ rename) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) out[k] = rename[in[k]];
}

// row of output entry k: the last i with prowptr[i] <= k, searched in [lo, hi]
__device__ __forceinline__ uint32_t gb_row_of(const uint32_t *prowptr,
                                              uint32_t lo, uint32_t hi,
                                              uint32_t k) {
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo + 1) / 2;
    if (prowptr[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define PERM_TILE 1024
// Gathers the entries of the permuted CSR; each block covers PERM_TILE
// consecutive output entries and first narrows the row range to search.
__global__ void __launch_bounds__(GB_BLOCK)
gb_perm_gather(uint32_t nrows, const uint32_t *__restrict__ rowptr,
               const uint32_t *__restrict__ rowind,
               const double *__restrict__ rowval,
               const uint32_t *__restrict__ rowperm,
               const uint32_t *__restrict__ rename,
               const uint32_t *__restrict__ prowptr, size_t nnz,
               uint32_t *__restrict__ prowind, double *__restrict__ prowval) {
  __shared__ uint32_t range[2];
  const size_t k0 = (size_t)blockIdx.x * PERM_TILE;
  size_t k1 = k0 + PERM_TILE;
  if (k1 > nnz) k1 = nnz;
  if (threadIdx.x == 0)
    range[0] = gb_row_of(prowptr, 0, nrows - 1, (uint32_t)k0);
  if (threadIdx.x == 1)
    range[1] = gb_row_of(prowptr, 0, nrows - 1, (uint32_t)(k1 - 1));
  __syncthreads();
  const uint32_t lo = range[0], hi = range[1];
  for (size_t k = k0 + threadIdx.x; k < k1; k += GB_BLOCK) {
    uint32_t i = gb_row_of(prowptr, lo, hi, (uint32_t)k);
    uint32_t v = rowperm ? rowperm[i] : i;
    size_t src = (size_t)rowptr[v] + (k - prowptr[i]);
    uint32_t c = rowind[src];
    prowind[k] = rename ? rename[c] : c;
    prowval[k] = rowval[src];
  }
}

__global__ void gb_perm_dense(uint32_t nrows, uint32_t ncols,
                              const double *__restrict__ in,
                              double *__restrict__ out,
                              const uint32_t *__restrict__ rp,
                              const uint32_t *__restrict__ cp, int rev) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t tot = (size_t)nrows * ncols;
  if (e >= tot) return;
  uint32_t i = (uint32_t)(e / ncols), j = (uint32_t)(e % ncols);
  uint32_t v = rp ? rp[i] : i;
  uint32_t u = cp ? cp[j] : j;
  size_t o = (size_t)v * ncols + u;
  if (rev) out[o] = in[e];
  else out[e] = in[o];
}

extern "C" {

int gbd_scan_u32(const uint32_t *d_in, uint32_t *d_out, size_t n,
                 uint32_t *d_total) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (n > 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  return gb_scan_u32(d_in, d_out, n, d_total);
}

static int gb_countingsort(const uint32_t *d_keys, const uint32_t *d_vals,
                           uint32_t kmin, uint32_t kmax, size_t n,
                           uint32_t *d_out, uint32_t *d_dist, bool check) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (kmax < kmin || n > 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  hipStream_t s = gb_stream();
  const uint64_t range = (uint64_t)kmax - kmin + 1;
  // keys outside [kmin, kmax] would index out of bounds: reject them
  if (n && check) {
    uint32_t *bad = nullptr, hbad = 0;
    GB_HIP(gb_dev_alloc((void **)&bad, sizeof(uint32_t)));
    GB_HIP(hipMemsetAsync(bad, 0, sizeof(uint32_t), s));
    gb_cs_check_keys<<<gb_grid_stride_blocks(n), GB_BLOCK, 0, s>>>(
        d_keys, n, kmin, kmax, bad);
    GB_LAUNCH_CHECK();
    GB_HIP(hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GB_HIP(gb_dev_free(bad));
    GB_HIP(hipStreamSynchronize(s));
    if (hbad) return GOOSEBERRY_ERROR_INVALIDINPUT;
  }
  if (d_dist && range > 0xfffffffeull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (n == 0) {
    if (d_dist) GB_HIP(hipMemsetAsync(d_dist, 0, (range + 1) * sizeof(uint32_t), s));
    return GOOSEBERRY_SUCCESS;
  }
  int bits = 0;
  while (bits < 32 && (range - 1) >> bits) bits += 8;
  if (bits == 0) bits = 8;
  const int npass = bits / 8;
  const uint32_t ntiles = gb_div_up(n, CS_TILE);
  const size_t ncounts = (size_t)CS_RADIX * ntiles;
  uint32_t *counts = nullptr, *tk[2] = {nullptr, nullptr},
           *tv[2] = {nullptr, nullptr};
  GB_HIP(gb_dev_alloc((void **)&counts, ncounts * sizeof(uint32_t)));
  GB_HIP(gb_dev_alloc((void **)&tk[0], n * sizeof(uint32_t)));
  GB_HIP(gb_dev_alloc((void **)&tk[1], n * sizeof(uint32_t)));
  GB_HIP(gb_dev_alloc((void **)&tv[0], n * sizeof(uint32_t)));
  if (npass > 1) GB_HIP(gb_dev_alloc((void **)&tv[1], n * sizeof(uint32_t)));
  const uint32_t *ck = d_keys, *cv = d_vals;
  for (int p = 0; p < npass; ++p) {
    const int shift = 8 * p;
    const bool last = (p == npass - 1);
    uint32_t *ok = tk[p & 1];
    uint32_t *ov = last ? d_out : tv[p & 1];
    gb_cs_count<<<ntiles, GB_BLOCK, 0, s>>>(ck, n, kmin, shift, counts, ntiles);
    GB_LAUNCH_CHECK();
    rv = gb_scan_u32(counts, counts, ncounts, nullptr);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
    gb_cs_scatter<<<ntiles, GB_BLOCK, 0, s>>>(ck, cv, n, kmin, shift, counts, ntiles, ok, ov);
    GB_LAUNCH_CHECK();
    ck = ok;
    cv = ov;
  }
  // the distribution from the sorted keys
  if (d_dist) {
    uint64_t g = (range + GB_BLOCK) / GB_BLOCK;
    if (g > 65536) g = 65536;
    gb_cs_dist<<<(uint32_t)g, GB_BLOCK, 0, s>>>(ck, n, kmin, range, d_dist);
    GB_LAUNCH_CHECK();
  }
  GB_HIP(gb_dev_free(counts));
  GB_HIP(gb_dev_free(tk[0]));
  GB_HIP(gb_dev_free(tk[1]));
  GB_HIP(gb_dev_free(tv[0]));
  if (tv[1]) GB_HIP(gb_dev_free(tv[1]));
  return GOOSEBERRY_SUCCESS;
}

int gbd_countingsort_kv(const uint32_t *d_keys, const uint32_t *d_vals,
                        uint32_t kmin, uint32_t kmax, size_t n,
                        uint32_t *d_out, uint32_t *d_dist) {
  return gb_countingsort(d_keys, d_vals, kmin, kmax, n, d_out, d_dist, true);
}

int gbd_countingsort_kv_unchecked(const uint32_t *d_keys, const uint32_t *d_vals,
                                  uint32_t kmin, uint32_t kmax, size_t n,
                                  uint32_t *d_out, uint32_t *d_dist) {
  return gb_countingsort(d_keys, d_vals, kmin, kmax, n, d_out, d_dist, false);
}

int gbd_permute_sparse(uint32_t nrows, uint32_t ncols, const uint32_t *d_rowptr,
                       const uint32_t *d_rowind, const double *d_rowval,
                       const uint32_t *d_rowperm, const uint32_t *d_colperm,
                       uint32_t *d_prowptr, uint32_t *d_prowind,
                       double *d_prowval) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipStream_t s = gb_stream();
  uint32_t nnz = 0;
  GB_HIP(hipMemcpyAsync(&nnz, d_rowptr + nrows, sizeof(uint32_t),
                        hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  uint32_t *rename = nullptr;
  if (d_colperm && ncols) {
    GB_HIP(gb_dev_alloc((void **)&rename, (size_t)ncols * sizeof(uint32_t)));
    gb_perm_rename<<<gb_div_up(ncols, GB_BLOCK), GB_BLOCK, 0, s>>>(d_colperm,
                                                                  ncols, rename);
    GB_LAUNCH_CHECK();
  }
  if (d_rowperm) {
    if (nrows) {
      uint32_t *len = nullptr;
      GB_HIP(gb_dev_alloc((void **)&len, (size_t)nrows * sizeof(uint32_t)));
      gb_perm_rowlen<<<gb_div_up(nrows, GB_BLOCK), GB_BLOCK, 0, s>>>(
          d_rowptr, d_rowperm, nrows, len);
      GB_LAUNCH_CHECK();
      rv = gb_scan_u32(len, d_prowptr, nrows, d_prowptr + nrows);
      if (rv != GOOSEBERRY_SUCCESS) return rv;
      GB_HIP(gb_dev_free(len));
    } else {
      GB_HIP(hipMemsetAsync(d_prowptr, 0, sizeof(uint32_t), s));
    }
    if (nnz) {
      gb_perm_gather<<<gb_div_up(nnz, PERM_TILE), GB_BLOCK, 0, s>>>(
          nrows, d_rowptr, d_rowind, d_rowval, d_rowperm, rename, d_prowptr,
          nnz, d_prowind, d_prowval);
      GB_LAUNCH_CHECK();
    }
  } else {
    GB_HIP(hipMemcpyAsync(d_prowptr, d_rowptr, ((size_t)nrows + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, s));
    if (nnz) {
      GB_HIP(hipMemcpyAsync(d_prowval, d_rowval, (size_t)nnz * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
      if (rename) {
        gb_perm_rename_apply<<<gb_div_up(nnz, GB_BLOCK), GB_BLOCK, 0, s>>>(
            d_rowind, d_prowind, nnz, rename);
        GB_LAUNCH_CHECK();
      } else {
        GB_HIP(hipMemcpyAsync(d_prowind, d_rowind, (size_t)nnz * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
      }
    }
  }
  if (rename) GB_HIP(gb_dev_free(rename));
  return GOOSEBERRY_SUCCESS;
}

int gbd_permute_dense(uint32_t nrows, uint32_t ncols, const double *d_in,
                      double *d_out, const uint32_t *d_rowperm,
                      const uint32_t *d_colperm, int rev) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  size_t tot = (size_t)nrows * ncols;
  if (!tot) return GOOSEBERRY_SUCCESS;
  gb_perm_dense<<<gb_div_up(tot, GB_BLOCK), GB_BLOCK, 0, gb_stream()>>>(
      nrows, ncols, d_in, d_out, d_rowperm, d_colperm, rev);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

}  // extern "C"
