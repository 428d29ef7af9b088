/*
 * gb_gen.c -- synthetic inputs: Dirichlet Laplacians and seeded uniform
 * vectors.  The random stream is xoshiro256** seeded through SplitMix64
 * (Blackman & Vigna), u = (x >> 11) * 2^-53 in [0, 1).
 */
#include <stdlib.h>
#include <string.h>

#include "gooseberry.h"
#include "gb_host.h"

static uint64_t splitmix64(uint64_t *x)
{
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k)
{
  return (x << k) | (x >> (64 - k));
}

typedef struct { uint64_t s[4]; } gb_rng_t;

/* stream k of a seed: SplitMix64 started at seed + k * 0xd1342543de82ef95 */
void gb_rng_init(gb_rng_t *r, uint64_t seed, uint64_t stream)
{
  uint64_t x = seed + stream * 0xd1342543de82ef95ull;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

uint64_t gb_rng_next(gb_rng_t *r)
{
  uint64_t *s = r->s;
  uint64_t result = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

double gb_rng_uniform(gb_rng_t *r)
{
  return (double)(gb_rng_next(r) >> 11) * 0x1.0p-53;
}

int gooseberry_generate_uniform(uint64_t seed, uint64_t stream, size_t n,
    real_t lo, real_t hi, real_t *out)
{
  if (n && !out) return GOOSEBERRY_ERROR_INVALIDINPUT;
  gb_rng_t r;
  gb_rng_init(&r, seed, stream);
  const double w = hi - lo;
  for (size_t i = 0; i < n; ++i) out[i] = lo + w * gb_rng_uniform(&r);
  return GOOSEBERRY_SUCCESS;
}

int gooseberry_generate_laplacian(int dims, dim_t nx, dim_t ny, dim_t nz,
    dim_t *nrows, ind_t **r_rowptr, dim_t **r_rowind, real_t **r_rowval)
{
  if (dims != 2 && dims != 3) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (dims == 2) nz = 1;
  if (!nx || !ny || !nz || !nrows || !r_rowptr || !r_rowind || !r_rowval)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint64_t n = (uint64_t)nx * ny * nz;
  uint64_t nnz = (dims == 2)
      ? 5 * n - 2 * (uint64_t)nx - 2 * (uint64_t)ny
      : 7 * n - 2 * ((uint64_t)nx * ny + (uint64_t)ny * nz + (uint64_t)nx * nz);
  if (n >= 0xffffffffull || nnz >= 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  ind_t *rowptr = malloc(sizeof(ind_t) * (n + 1));
  dim_t *rowind = malloc(sizeof(dim_t) * nnz);
  real_t *rowval = malloc(sizeof(real_t) * nnz);
  if (!rowptr || !rowind || !rowval) {
    free(rowptr); free(rowind); free(rowval);
    return GOOSEBERRY_ERROR_UNKNOWN;
  }
  const real_t diag = dims == 2 ? 4.0 : 6.0;
  const uint64_t sx = 1, sy = nx, sz = (uint64_t)nx * ny;
  uint64_t k = 0, i = 0;
  rowptr[0] = 0;
  for (dim_t z = 0; z < nz; ++z) {
    for (dim_t y = 0; y < ny; ++y) {
      for (dim_t x = 0; x < nx; ++x, ++i) {
        /* neighbours in ascending column order */
        if (dims == 3 && z > 0) { rowind[k] = (dim_t)(i - sz); rowval[k++] = -1.0; }
        if (y > 0) { rowind[k] = (dim_t)(i - sy); rowval[k++] = -1.0; }
        if (x > 0) { rowind[k] = (dim_t)(i - sx); rowval[k++] = -1.0; }
        rowind[k] = (dim_t)i; rowval[k++] = diag;
        if (x + 1 < nx) { rowind[k] = (dim_t)(i + sx); rowval[k++] = -1.0; }
        if (y + 1 < ny) { rowind[k] = (dim_t)(i + sy); rowval[k++] = -1.0; }
        if (dims == 3 && z + 1 < nz) { rowind[k] = (dim_t)(i + sz); rowval[k++] = -1.0; }
        rowptr[i + 1] = (ind_t)k;
      }
    }
  }
  *nrows = (dim_t)n;
  *r_rowptr = rowptr;
  *r_rowind = rowind;
  *r_rowval = rowval;
  return GOOSEBERRY_SUCCESS;
}
