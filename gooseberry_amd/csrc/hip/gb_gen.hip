// gb_gen.hip -- random symmetric, strictly diagonally dominant SPD matrices
// built directly in HBM (no transpose, no host round trip).
//
// Row i draws a target length t_i uniform in [dmin, dmax] (counting the
// diagonal).  K = ceil((dmax - 1) / 2) keyed bijections pi_k of [0, n)
// (4-round Feistel networks on the next even power of two, cycle-walked into
// [0, n)) pair row i with j = pi_k(i) and with j = pi_k^-1(i).  The pair
// {i, j} of bijection k is kept iff i != j and 2k + 1 < min(t_i, t_j), a
// rule both rows evaluate identically, so the pattern is symmetric and each
// column position is uniform over the rows.  The value of a kept pair is
// -u with u uniform in (0, 1] from a hash of (seed, k, min(i, j),
// max(i, j)), the same from both sides.  The diagonal is 1.01 times the
// row's absolute off-diagonal sum (1 for a row with no off-diagonals), so
// the matrix is symmetric positive definite.  A row lists its kept pairs in
// the order k = 0..K-1 (pi_k(i) then pi_k^-1(i)), then the diagonal.
#include <stdlib.h>

#include "gb_common.h"

__device__ __forceinline__ uint64_t gb_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t gb_hash3(uint64_t a, uint64_t b, uint64_t c) {
  return gb_mix64(a * 0x9e3779b97f4a7c15ull ^ gb_mix64(b + 0x632be59bd9b4e019ull) ^
                  gb_mix64(c * 0xd1342543de82ef95ull + 0x7f4a7c159e3779b9ull));
}

// uniform in (0, 1]
__device__ __forceinline__ double gb_u01(uint64_t h) {
  return (double)((h >> 11) + 1) * 0x1.0p-53;
}

struct gb_feistel {
  uint32_t half, mask;
  uint64_t key;
};

__device__ __forceinline__ uint32_t gb_fround(uint32_t r, uint64_t key, int round,
                                              uint32_t mask) {
  return (uint32_t)gb_hash3(key, (uint64_t)round, r) & mask;
}

__device__ __forceinline__ uint32_t gb_fperm(uint32_t x, const gb_feistel &F) {
  uint32_t L = x >> F.half, R = x & F.mask;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t t = R;
    R = L ^ gb_fround(R, F.key, k, F.mask);
    L = t;
  }
  return (L << F.half) | R;
}

__device__ __forceinline__ uint32_t gb_finv(uint32_t x, const gb_feistel &F) {
  uint32_t L = x >> F.half, R = x & F.mask;
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    uint32_t t = L;
    L = R ^ gb_fround(L, F.key, k, F.mask);
    R = t;
  }
  return (L << F.half) | R;
}

// cycle walking: a bijection of [0, n) (every cycle through a value < n
// returns below n, at the latest at x itself)
__device__ __forceinline__ uint32_t gb_pi(uint32_t x, uint32_t n, const gb_feistel &F) {
  uint32_t y = gb_fperm(x, F);
  while (y >= n) y = gb_fperm(y, F);
  return y;
}
__device__ __forceinline__ uint32_t gb_pi_inv(uint32_t x, uint32_t n, const gb_feistel &F) {
  uint32_t y = gb_finv(x, F);
  while (y >= n) y = gb_finv(y, F);
  return y;
}

struct gb_sdd_params {
  uint32_t n, dmin, dspan, K, half, mask;
  uint64_t seed;
};

__device__ __forceinline__ uint32_t gb_target(uint32_t i, const gb_sdd_params &P) {
  return P.dmin + (uint32_t)(gb_hash3(P.seed, 0xabcdefull, i) % P.dspan);
}

__device__ __forceinline__ gb_feistel gb_feistel_k(const gb_sdd_params &P, uint32_t k) {
  gb_feistel F;
  F.half = P.half;
  F.mask = P.mask;
  F.key = gb_hash3(P.seed, 0x5eedull, k);
  return F;
}

// MODE 0: row lengths; MODE 1: fill (rowptr given)
template <int MODE>
__global__ void __launch_bounds__(GB_BLOCK)
gb_sdd_rows(gb_sdd_params P, uint32_t *__restrict__ len,
            const uint32_t *__restrict__ rowptr, uint32_t *__restrict__ rowind,
            double *__restrict__ rowval) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= P.n) return;
  const uint32_t ti = gb_target(i, P);
  uint32_t cnt = 0, pos = MODE ? rowptr[i] : 0;
  double asum = 0.0;
  for (uint32_t k = 0; k < P.K; ++k) {
    if (2 * k + 1 >= ti) break;  // larger k are never kept for this row
    const gb_feistel F = gb_feistel_k(P, k);
    const uint32_t js[2] = {gb_pi(i, P.n, F), gb_pi_inv(i, P.n, F)};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const uint32_t j = js[d];
      if (j == i) continue;
      const uint32_t tj = gb_target(j, P);
      const uint32_t tmin = ti < tj ? ti : tj;
      if (2 * k + 1 >= tmin) continue;
      ++cnt;
      if (MODE) {
        const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
        const double v = -gb_u01(gb_hash3(P.seed ^ ((uint64_t)k << 40), lo, hi));
        rowind[pos] = j;
        rowval[pos] = v;
        ++pos;
        asum += -v;
      }
    }
  }
  if (MODE) {
    rowind[pos] = i;
    rowval[pos] = cnt ? 1.01 * asum : 1.0;
  } else {
    len[i] = cnt + 1;
  }
}

// out[i] = lo + (hi - lo) u, u = (hash(seed, stream, i) >> 11) 2^-53
__global__ void gb_fill_uniform(double *__restrict__ out, size_t n, uint64_t seed,
                                uint64_t stream, double lo, double hi) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  const double w = hi - lo;
  for (; i < n; i += stride) {
    double u = (double)(gb_hash3(seed, stream + 0x1234567ull, i) >> 11) * 0x1.0p-53;
    out[i] = lo + w * u;
  }
}

extern "C" {

int gbd_fill_uniform(double *d_out, size_t n, uint64_t seed, uint64_t stream,
                     double lo, double hi) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!n) return GOOSEBERRY_SUCCESS;
  uint32_t g = gb_div_up(n, GB_BLOCK);
  if (g > 16384u) g = 16384u;
  gb_fill_uniform<<<g, GB_BLOCK, 0, gb_stream()>>>(d_out, n, seed, stream, lo, hi);
  GB_LAUNCH_CHECK();
  return GOOSEBERRY_SUCCESS;
}

int gbd_generate_sdd(uint32_t n, uint64_t seed, uint32_t dmin, uint32_t dmax,
                     uint32_t **d_rowptr, uint32_t **d_rowind, double **d_rowval) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (n == 0 || dmin < 1 || dmax < dmin || dmax > 255) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if ((uint64_t)n * dmax >= 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gb_sdd_params P;
  P.n = n;
  P.dmin = dmin;
  P.dspan = dmax - dmin + 1;
  P.K = dmax / 2;  // 2k + 1 < t <= dmax  =>  k < dmax / 2
  P.seed = seed;
  uint32_t bits = 2;
  while ((1ull << bits) < n) bits += 2;
  P.half = bits / 2;
  P.mask = (1u << P.half) - 1;
  hipStream_t s = gb_stream();
  uint32_t *rowptr = nullptr, *rowind = nullptr;
  double *rowval = nullptr;
  GB_HIP(hipMalloc((void **)&rowptr, ((size_t)n + 1) * 4));
  const uint32_t grid = gb_div_up(n, GB_BLOCK);
  gb_sdd_rows<0><<<grid, GB_BLOCK, 0, s>>>(P, rowptr, nullptr, nullptr, nullptr);
  GB_LAUNCH_CHECK();
  rv = gbd_scan_u32(rowptr, rowptr, n, rowptr + n);
  if (rv != GOOSEBERRY_SUCCESS) { hipFree(rowptr); return rv; }
  uint32_t nnz = 0;
  GB_HIP(hipMemcpyAsync(&nnz, rowptr + n, 4, hipMemcpyDeviceToHost, s));
  GB_HIP(hipStreamSynchronize(s));
  hipError_t e = hipMalloc((void **)&rowind, (size_t)nnz * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&rowval, (size_t)nnz * 8);
  if (e != hipSuccess) {
    hipFree(rowptr); hipFree(rowind);
    return gb_hip_fail(e, "gbd_generate_sdd", __FILE__, __LINE__);
  }
  gb_sdd_rows<1><<<grid, GB_BLOCK, 0, s>>>(P, nullptr, rowptr, rowind, rowval);
  GB_LAUNCH_CHECK();
  *d_rowptr = rowptr;
  *d_rowind = rowind;
  *d_rowval = rowval;
  return GOOSEBERRY_SUCCESS;
}

}  // extern "C"
