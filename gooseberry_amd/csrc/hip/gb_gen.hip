// gb_gen.hip -- random symmetric, strictly diagonally dominant SPD matrices
// built directly in HBM (no host round trip).
//
// Row i draws a degree d_i -- uniform in [dmin, dmax], or Zipf(s) clipped
// to [dmin, dmax] -- and emits o_i = d_i - 1 directed edges e = off_i ..
// off_i + o_i - 1 (off = exclusive scan of o).  Edge e goes to a column j_e
// uniform over [0, n) \ {i} and carries the value -u_e, u_e uniform in
// (0, 1], both from counter-based hashes of (seed, e).  The matrix is the
// symmetrised pattern: edge e places -u_e at (i, j_e) and at (j_e, i);
// positions drawn twice stay separate entries (the matrix element is their
// sum), so A equals its transpose exactly.  Row i lists its out-edges in e
// order, then its in-edges in e order (a stable counting sort of the edges
// by head), then the diagonal: 1.01 times the row's absolute off-diagonal
// sum in that order (1 for a row without off-diagonals).  A symmetric
// matrix with positive, strictly dominant diagonal is positive definite.
//
// Zipf variates use Devroye's rejection method from the continuous Pareto
// tail (L. Devroye, Non-Uniform Random Variate Generation, Springer 1986,
// ch. X.6): X = floor(U^(-1/(s-1))), accepted iff
// V X (T - 1) / (b - 1) <= T / b with T = (1 + 1/X)^(s-1), b = 2^(s-1).
#include <math.h>
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

// hash streams
#define GB_H_DEGREE 0xd0ull
#define GB_H_ZIPF_U 0xd1ull
#define GB_H_ZIPF_V 0xd2ull
#define GB_H_COLUMN 0xe0ull
#define GB_H_VALUE 0xe1ull
// rejection rounds before a Zipf draw gives up (acceptance is ~0.9 per round
// at s = 1.5; 1024 rejections in a row do not happen in practice)
#define GB_ZIPF_TRIES 1024u

struct gb_gen_params {
  uint32_t n, dmin, dmax, law;
  double am1, b;  // s - 1 and 2^(s - 1) for the Zipf law
  uint64_t seed;
};

__device__ __forceinline__ uint32_t gb_gen_degree(uint32_t i, const gb_gen_params &P) {
  if (P.law == GOOSEBERRY_DEGREE_UNIFORM)
    return P.dmin + (uint32_t)(gb_hash3(P.seed, GB_H_DEGREE, i) % (P.dmax - P.dmin + 1));
  double x = 1.0;
  for (uint32_t t = 0; t < GB_ZIPF_TRIES; ++t) {
    const uint64_t r = (uint64_t)t << 8;
    const double U = gb_u01(gb_hash3(P.seed, GB_H_ZIPF_U + r, i));
    const double V = gb_u01(gb_hash3(P.seed, GB_H_ZIPF_V + r, i));
    x = floor(pow(U, -1.0 / P.am1));
    const double T = pow(1.0 + 1.0 / x, P.am1);
    if (V * x * (T - 1.0) / (P.b - 1.0) <= T / P.b) break;
  }
  if (!(x <= (double)P.dmax)) return P.dmax;  // also a NaN from an overflow
  if (x < (double)P.dmin) return P.dmin;
  return (uint32_t)x;
}

__device__ __forceinline__ double gb_gen_value(const gb_gen_params &P, uint32_t e) {
  return -gb_u01(gb_hash3(P.seed, GB_H_VALUE, e));
}

// o_i = d_i - 1 into outdeg[i]
__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_outdeg(gb_gen_params P, uint32_t *__restrict__ outdeg) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i < P.n) outdeg[i] = gb_gen_degree(i, P) - 1;
}

// 64-bit total of a uint32 array (integer sum, order does not matter)
__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_sum64(const uint32_t *__restrict__ a, uint32_t n,
             unsigned long long *__restrict__ total) {
  __shared__ unsigned long long lds[GB_WAVES_PER_BLOCK];
  unsigned long long v = 0;
  for (uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += gridDim.x * GB_BLOCK)
    v += a[i];
  v = gb_block_sum(v, lds);
  if (threadIdx.x == 0) atomicAdd(total, v);
}

// edges of row i: src[e] = i, dst[e] = j_e
__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_edges(gb_gen_params P, const uint32_t *__restrict__ off,
             uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= P.n) return;
  const uint32_t e1 = off[i + 1];
  for (uint32_t e = off[i]; e < e1; ++e) {
    uint32_t j = (uint32_t)(gb_hash3(P.seed, GB_H_COLUMN, e) % (P.n - 1));
    j += (j >= i);
    src[e] = i;
    dst[e] = j;
  }
}

// row lengths of the symmetrised matrix: out-edges, in-edges, diagonal
__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_rowlen(uint32_t n, const uint32_t *__restrict__ off,
              const uint32_t *__restrict__ inptr, uint32_t *__restrict__ len) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i < n) len[i] = (off[i + 1] - off[i]) + (inptr[i + 1] - inptr[i]) + 1;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_fill_out(gb_gen_params P, const uint32_t *__restrict__ off,
                const uint32_t *__restrict__ dst, const uint32_t *__restrict__ rowptr,
                uint32_t *__restrict__ rowind, double *__restrict__ rowval) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= P.n) return;
  const uint32_t e0 = off[i], e1 = off[i + 1];
  uint32_t pos = rowptr[i];
  for (uint32_t e = e0; e < e1; ++e, ++pos) {
    rowind[pos] = dst[e];
    rowval[pos] = gb_gen_value(P, e);
  }
}

// ine[p], p in [inptr[j], inptr[j+1]), are the edges into row j in e order
__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_fill_in(gb_gen_params P, uint32_t m, const uint32_t *__restrict__ ine,
               const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
               const uint32_t *__restrict__ off, const uint32_t *__restrict__ inptr,
               const uint32_t *__restrict__ rowptr, uint32_t *__restrict__ rowind,
               double *__restrict__ rowval) {
  for (uint32_t p = blockIdx.x * GB_BLOCK + threadIdx.x; p < m;
       p += gridDim.x * GB_BLOCK) {
    const uint32_t e = ine[p];
    const uint32_t j = dst[e];
    const uint32_t slot = rowptr[j] + (off[j + 1] - off[j]) + (p - inptr[j]);
    rowind[slot] = src[e];
    rowval[slot] = gb_gen_value(P, e);
  }
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_gen_diag(uint32_t n, const uint32_t *__restrict__ rowptr,
            uint32_t *__restrict__ rowind, double *__restrict__ rowval) {
  const uint32_t i = blockIdx.x * GB_BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = rowptr[i], last = rowptr[i + 1] - 1;
  double s = 0.0;
  for (uint32_t k = a; k < last; ++k) s = s + fabs(rowval[k]);
  rowind[last] = i;
  rowval[last] = last > a ? 1.01 * s : 1.0;
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

int gbd_generate_spd(uint32_t n, uint64_t seed, int law, double zipf_s,
                     uint32_t dmin, uint32_t dmax, uint32_t **d_rowptr,
                     uint32_t **d_rowind, double **d_rowval, uint32_t *r_nnz) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!d_rowptr || !d_rowind || !d_rowval || !r_nnz) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (n < 2 || n == 0xffffffffu || dmin < 1 || dmax < dmin || dmax > (1u << 20))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (law != GOOSEBERRY_DEGREE_UNIFORM && law != GOOSEBERRY_DEGREE_ZIPF)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (law == GOOSEBERRY_DEGREE_ZIPF && !(zipf_s > 1.0 && zipf_s <= 16.0))
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  gb_gen_params P;
  P.n = n;
  P.dmin = dmin;
  P.dmax = dmax;
  P.law = (uint32_t)law;
  P.am1 = law == GOOSEBERRY_DEGREE_ZIPF ? zipf_s - 1.0 : 1.0;
  P.b = exp2(P.am1);
  P.seed = seed;
  hipStream_t s = gb_stream();
  const uint32_t grid = gb_div_up(n, GB_BLOCK);
  uint32_t *off = nullptr, *src = nullptr, *dst = nullptr, *ine = nullptr,
           *inptr = nullptr, *rowptr = nullptr, *rowind = nullptr;
  double *rowval = nullptr;
  unsigned long long *tot = nullptr, m64 = 0;
  uint32_t m = 0, nnz = 0;
  hipError_t e = hipSuccess;
#define GB_GEN_TRY(call)                                       \
  do {                                                         \
    e = (call);                                                \
    if (e != hipSuccess) {                                     \
      rv = gb_hip_fail(e, #call, __FILE__, __LINE__);          \
      goto done;                                               \
    }                                                          \
  } while (0)
#define GB_GEN_RV(call)                                        \
  do {                                                         \
    rv = (call);                                               \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                   \
  } while (0)

  GB_GEN_TRY(gb_dev_alloc((void **)&off, ((size_t)n + 1) * 4));
  GB_GEN_TRY(gb_dev_alloc((void **)&tot, sizeof(*tot)));
  GB_GEN_TRY(hipMemsetAsync(tot, 0, sizeof(*tot), s));
  gb_gen_outdeg<<<grid, GB_BLOCK, 0, s>>>(P, off);
  GB_GEN_TRY(hipGetLastError());
  gb_gen_sum64<<<grid < 1024u ? grid : 1024u, GB_BLOCK, 0, s>>>(off, n, tot);
  GB_GEN_TRY(hipGetLastError());
  GB_GEN_TRY(hipMemcpyAsync(&m64, tot, sizeof(m64), hipMemcpyDeviceToHost, s));
  GB_GEN_TRY(hipStreamSynchronize(s));
  // every index and row pointer of the result must fit 32 bits
  if (2 * m64 + n > 0xffffffffull) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  m = (uint32_t)m64;
  nnz = 2 * m + n;
  GB_GEN_RV(gbd_scan_u32(off, off, n, off + n));
  GB_GEN_TRY(gb_dev_alloc((void **)&src, (size_t)m * 4 + 4));
  GB_GEN_TRY(gb_dev_alloc((void **)&dst, (size_t)m * 4 + 4));
  GB_GEN_TRY(gb_dev_alloc((void **)&ine, (size_t)m * 4 + 4));
  GB_GEN_TRY(gb_dev_alloc((void **)&inptr, ((size_t)n + 1) * 4));
  gb_gen_edges<<<grid, GB_BLOCK, 0, s>>>(P, off, src, dst);
  GB_GEN_TRY(hipGetLastError());
  // every dst is in [0, n) by construction
  GB_GEN_RV(gbd_countingsort_kv_unchecked(dst, nullptr, 0, n - 1, m, ine, inptr));
  GB_GEN_TRY(gb_dev_alloc((void **)&rowptr, ((size_t)n + 1) * 4));
  gb_gen_rowlen<<<grid, GB_BLOCK, 0, s>>>(n, off, inptr, rowptr);
  GB_GEN_TRY(hipGetLastError());
  GB_GEN_RV(gbd_scan_u32(rowptr, rowptr, n, rowptr + n));
  GB_GEN_TRY(gb_dev_alloc((void **)&rowind, (size_t)nnz * 4));
  GB_GEN_TRY(gb_dev_alloc((void **)&rowval, (size_t)nnz * 8));
  gb_gen_fill_out<<<grid, GB_BLOCK, 0, s>>>(P, off, dst, rowptr, rowind, rowval);
  GB_GEN_TRY(hipGetLastError());
  if (m) {
    uint32_t g = gb_div_up(m, GB_BLOCK);
    if (g > 65536u) g = 65536u;
    gb_gen_fill_in<<<g, GB_BLOCK, 0, s>>>(P, m, ine, src, dst, off, inptr, rowptr,
                                          rowind, rowval);
    GB_GEN_TRY(hipGetLastError());
  }
  gb_gen_diag<<<grid, GB_BLOCK, 0, s>>>(n, rowptr, rowind, rowval);
  GB_GEN_TRY(hipGetLastError());
  *d_rowptr = rowptr;
  *d_rowind = rowind;
  *d_rowval = rowval;
  *r_nnz = nnz;
  rowptr = rowind = nullptr;
  rowval = nullptr;
  rv = GOOSEBERRY_SUCCESS;
done:
  (void)gb_dev_free(off);
  (void)gb_dev_free(tot);
  (void)gb_dev_free(src);
  (void)gb_dev_free(dst);
  (void)gb_dev_free(ine);
  (void)gb_dev_free(inptr);
  (void)gb_dev_free(rowptr);
  (void)gb_dev_free(rowind);
  (void)gb_dev_free(rowval);
  return rv;
#undef GB_GEN_TRY
#undef GB_GEN_RV
}

}  // extern "C"
