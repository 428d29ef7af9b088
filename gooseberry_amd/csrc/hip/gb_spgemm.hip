// gb_spgemm.hip -- C = A B for sparse A and B (src/blas.c:94-234,609-791),
// bit for bit the reference's row-by-row scan.
//
// The reference walks row i of A entry by entry (j), and for each the row
// arowind[j] of B entry by entry (k): the product x_j * b_k goes to column
// browind[k] of C.  Each C element is the sum of its products in that scan
// order; in the sparse result the columns of a row appear in the order of
// their first product.  Here every product is materialised with its scan
// position t (row-major, so t orders rows too), then
//   1. stable counting sort of t by column, then by row: (row, col, t) order
//   2. one sequential sum per (row, col) run, in t order (the first product,
//      then += the others; the dense variant adds them to +0.0 as the
//      reference's zeroed row does)
//   3. dense result: scatter the sums; sparse result: the runs sorted by
//      their first t (unique, row-major), so each row lists its columns by
//      first appearance.
#include <stdlib.h>

#include "gb_common.h"

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

// the products of A entry j (row i, column l) with B row l, at positions
// off[j] .. off[j] + len[j] - 1: their scan order
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_expand(uint32_t anrows, uint64_t annz, const uint32_t *__restrict__ arowptr,
             const uint32_t *__restrict__ arowind, const double *__restrict__ arowval,
             const uint32_t *__restrict__ browptr, const uint32_t *__restrict__ browind,
             const double *__restrict__ browval, const uint32_t *__restrict__ off,
             uint32_t *__restrict__ erow, uint32_t *__restrict__ ecol,
             double *__restrict__ eval) {
  for (uint64_t j = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; j < annz;
       j += (uint64_t)gridDim.x * GB_BLOCK) {
    const uint32_t i = gb_sg_row_of(arowptr, anrows, (uint32_t)j);
    const uint32_t l = arowind[j];
    const double x = arowval[j];
    uint32_t t = off[j];
    for (uint32_t k = browptr[l]; k < browptr[l + 1]; ++k, ++t) {
      erow[t] = i;
      ecol[t] = browind[k];
      eval[t] = x * browval[k];
    }
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

// row pointers of C from the runs' rows (ascending in run order): thread q
// writes the pointers of the rows between run q-1's row and run q's row
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_rowptr(uint32_t nruns, uint32_t anrows, const uint32_t *__restrict__ rrow,
             uint32_t *__restrict__ crowptr) {
  for (uint32_t q = blockIdx.x * GB_BLOCK + threadIdx.x; q <= nruns;
       q += gridDim.x * GB_BLOCK) {
    const uint32_t lo = q == 0 ? 0u : rrow[q - 1] + 1u;
    const uint32_t hi = q == nruns ? anrows : rrow[q];
    for (uint32_t i = lo; i <= hi; ++i) crowptr[i] = q;
  }
}

// runs in first-appearance order (ordr) written out as the C arrays
__global__ void __launch_bounds__(GB_BLOCK)
gb_sg_out(uint32_t nruns, const uint32_t *__restrict__ ordr, const uint32_t *__restrict__ rcol,
          const double *__restrict__ rval, uint32_t *__restrict__ crowind,
          double *__restrict__ crowval) {
  for (uint32_t q = blockIdx.x * GB_BLOCK + threadIdx.x; q < nruns; q += gridDim.x * GB_BLOCK) {
    const uint32_t r = ordr[q];
    crowind[q] = rcol[r];
    crowval[q] = rval[r];
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
  uint32_t annz = 0, E = 0, nruns = 0;
  uint32_t *len = nullptr, *off = nullptr, *erow = nullptr, *ecol = nullptr,
           *perm1 = nullptr, *row1 = nullptr, *ord = nullptr, *flag = nullptr,
           *rid = nullptr, *runstart = nullptr, *rrow = nullptr, *rcol = nullptr,
           *rfirst = nullptr, *ordr = nullptr, *crowptr = nullptr,
           *crowind = nullptr;
  double *eval = nullptr, *rval = nullptr, *crowval = nullptr;
  unsigned long long e64 = 0, *tot = nullptr;
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
  E = (uint32_t)e64;
  GB_SG_RV(gbd_scan_u32(len, off, annz, off + annz));
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
        anrows, annz, d_arowptr, d_arowind, d_arowval, d_browptr, d_browind, d_browval, off,
        erow, ecol, eval);
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
      gb_sg_runsum<false><<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(
          nruns, E, runstart, ord, erow, ecol, eval, bncols, nullptr, rrow, rcol, rfirst,
          rval);
      GB_SG_TRY(hipGetLastError());
      // 3. first-appearance order: the first positions are unique and
      // row-major, so sorting by them alone orders the rows too
      GB_SG_RV(gbd_countingsort_kv_unchecked(rfirst, nullptr, 0, E - 1, nruns, ordr,
                                             nullptr));
    }
  }
  if (!dense) {
    GB_SG_TRY(gb_dev_alloc((void **)&crowptr, ((size_t)anrows + 1) * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&crowind, ((size_t)nruns + 1) * 4));
    GB_SG_TRY(gb_dev_alloc((void **)&crowval, ((size_t)nruns + 1) * 8));
    gb_sg_rowptr<<<gb_sg_grid((uint64_t)nruns + 1), GB_BLOCK, 0, s>>>(nruns, anrows, rrow,
                                                                      crowptr);
    GB_SG_TRY(hipGetLastError());
    if (nruns) {
      gb_sg_out<<<gb_sg_grid(nruns), GB_BLOCK, 0, s>>>(nruns, ordr, rcol, rval, crowind,
                                                       crowval);
      GB_SG_TRY(hipGetLastError());
    }
    *d_crowptr = crowptr;
    *d_crowind = crowind;
    *d_crowval = crowval;
    *r_cnnz = nruns;
    crowptr = crowind = nullptr;
    crowval = nullptr;
  }
  GB_SG_TRY(hipStreamSynchronize(s));
  rv = GOOSEBERRY_SUCCESS;
done:
  gb_dev_free(len); gb_dev_free(off); gb_dev_free(erow); gb_dev_free(ecol);
  gb_dev_free(eval); gb_dev_free(perm1); gb_dev_free(row1); gb_dev_free(ord);
  gb_dev_free(flag); gb_dev_free(rid); gb_dev_free(runstart); gb_dev_free(rrow);
  gb_dev_free(rcol); gb_dev_free(rfirst); gb_dev_free(rval); gb_dev_free(ordr);
  gb_dev_free(crowptr); gb_dev_free(crowind); gb_dev_free(crowval);
  gb_dev_free(tot);
  return rv;
#undef GB_SG_TRY
#undef GB_SG_RV
}

}  // extern "C"
