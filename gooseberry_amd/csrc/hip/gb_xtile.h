// gb_xtile.h -- the SpMV over the x-tile layout, shared by the matrix path
// (gb_spmv.hip) and the batched engine (gb_batch.hip).
//
// The layout groups a matrix's entries by (block of R rows, tile of
// C = 2^lc columns), each group in CSR order: ptr[g] .. ptr[g + 1] for group
// g = block * T + tile, each entry a packed word (row within the block << lc
// | column within the tile) and a value.  One workgroup owns a row block and
// keeps its y in LDS; tile by tile it copies x's tile into LDS and adds its
// group's products, so the gathers read LDS and only the matrix and x
// stream.  Each row gets one sum per tile, added to its y in tile order: a
// fixed regrouping of the row's sum (by tile, as with a column split),
// bitwise reproducible.  V is the value, vector and accumulation type
// (double, or float for the FP32 SpMV of the mixed-precision solves).  In
// FP32 a step covers two adjacent tiles (their groups are adjacent in the
// layout and their x fits the same LDS), whose runs add into two
// accumulators, summed when the block is written.
#pragma once

#include "gb_matrix.h"

#define GB_XT_THREADS 1024
#define GB_XT_RMAX 8192  // rows of y a workgroup keeps in LDS
#define GB_XT_DEPTH 1    // chunks a wavefront has in flight (on the Zipf matrix:
                         // FP64 1.56-1.59 ms at 1, 1.60 at 2, 1.72 at 3;
                         // FP32 1.08 ms at 1, 1.13 at 2)
#define GB_XT_PAD 4      // zero entries past the end of the layout's arrays

// what the workgroup does with its block of y = A x
enum {
  GB_XT_STORE = 0,  // y[row * ystride] = (A x)[row]
  GB_XT_DOT = 1,    // y[row] = (A x)[row]; partials[block] = sum of x[row] y[row]
                    // (x = p in a CG step: the block's part of p.Ap)
  GB_XT_RESID = 2,  // V = double: y[row] = b[row] - (A x)[row];
                    // partials[block] = sum of y[row]^2
};

// Cross-lane moves for the segmented scan, as DPP (data-parallel
// primitives: a lane reads another lane's register within the instruction,
// with no LDS round trip), with bound_ctrl set: a lane whose source lies
// outside the pattern reads 0, so no register needs a fallback value first.
template <int CTRL>
__device__ __forceinline__ uint32_t gb_dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double gb_dpp_val(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = gb_dpp_u32<CTRL>((uint32_t)b);
  const uint32_t hi = gb_dpp_u32<CTRL>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ float gb_dpp_val(float v) {
  return __uint_as_float(gb_dpp_u32<CTRL>(__float_as_uint(v)));
}
// one step of the segmented inclusive scan: a lane taking part in the step
// (`on`) adds the partial sum of its source lane when that lane holds the
// same row (rows of a group are contiguous, so the lanes between hold it
// too); the earlier part is the left operand.  A lane whose source lies
// outside the pattern reads row 0 and +0.0: adding +0.0 changes no partial
// sum except -0.0, which becomes +0.0, and y starts from +0.0, so every
// result is the same either way.
template <int CTRL, typename V>
__device__ __forceinline__ void gb_seg_step(V &sc, uint32_t r, bool on) {
  const V ps = gb_dpp_val<CTRL>(sc);
  const uint32_t pr = gb_dpp_u32<CTRL>(r);
  if (on && pr == r) sc = ps + sc;
}
// the scan over 64 lanes: within each row of 16 lanes by shifts of 1, 2,
// 4 and 8 lanes, then the last lane of row 0 into row 1 and of row 2 into
// row 3, then the last lane of row 1 into rows 2 and 3
template <typename V>
__device__ __forceinline__ V gb_seg_scan64(V sc, uint32_t r, uint32_t lane) {
  gb_seg_step<0x111>(sc, r, true);               // row_shr:1
  gb_seg_step<0x112>(sc, r, true);               // row_shr:2
  gb_seg_step<0x114>(sc, r, true);               // row_shr:4
  gb_seg_step<0x118>(sc, r, true);               // row_shr:8
  gb_seg_step<0x142>(sc, r, (lane & 16u) != 0);  // row_bcast:15, rows 1 and 3
  gb_seg_step<0x143>(sc, r, lane >= 32);         // row_bcast:31, rows 2 and 3
  return sc;
}
__device__ __forceinline__ double gb_readlane_val(double v, uint32_t l) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ float gb_readlane_val(float v, uint32_t l) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), l));
}

// The LDS a workgroup of the x-tile SpMV uses, declared once by each
// kernel and viewed as V (the FP32 SpMV uses the first halves)
struct gb_xt_lds {
  double xs[GB_XT_CMAX];  // the current tile of x
  double ys[GB_XT_RMAX];  // the block's y
  double wcarry[GB_XT_THREADS / GB_WAVE];  // a range's first run, cut off
                                           // from the one before
  double whead[GB_XT_THREADS / GB_WAVE];   // a range's last run, started in it
  uint32_t wfrow[GB_XT_THREADS / GB_WAVE]; // first run's row | cut (bit 31) |
                                           // the only run (bit 30)
  uint32_t wlrow[GB_XT_THREADS / GB_WAVE]; // last run's row | started here (bit 31)
  double red[GB_XT_THREADS / GB_WAVE];     // the block's partial
};

// y = A x on the x-tile layout for the block blockIdx.x (grid = the row
// blocks).  A group's entries are cut into one contiguous range per
// wavefront; a wavefront reads its range 64 E entries at a time, E
// consecutive entries per lane (vector loads, coalesced; GB_XT_DEPTH
// chunks in flight; the next group's tile of x and first chunks are loaded
// when this group starts).  A lane adds its products left to right while
// they share a row; a segmented scan across the lanes, keyed by each lane's
// last row, sums the runs (rows are contiguous within a group, so equal
// rows are one run), and a run that ends inside a lane after beginning at
// its first entry takes the previous lane's scan.  The run that reaches a chunk's end is carried into
// the next chunk; a run cut by the range's end is finished by the wavefront
// where it starts, which adds the following ranges' parts in order.  The
// layout's arrays hold GB_XT_PAD zero entries past the end, so a lane's
// vector never reads past them.
template <typename V, int EPI>
__device__ __forceinline__ void gb_xt_spmv(const uint32_t *__restrict__ ptr,
                                           const uint32_t *__restrict__ idx,
                                           const V *__restrict__ val, const V *__restrict__ x,
                                           V *__restrict__ y, uint32_t ystride,
                                           const double *__restrict__ b,
                                           double *__restrict__ partials, uint32_t nrows,
                                           uint32_t ncols, uint32_t R, uint32_t T, uint32_t lc,
                                           gb_xt_lds &L) {
  constexpr uint32_t NW = GB_XT_THREADS / GB_WAVE;
  // tiles a step covers: two adjacent tiles in FP32 (their x fits the same
  // 64 KB of LDS; on the Zipf matrix 1.01 ms against 1.07 ms a tile a
  // step), one in FP64; the second tile's rows carry bit 15 in their keys,
  // so that no run spans the two tiles
  constexpr uint32_t TP = sizeof(V) == 4 ? 2 : 1;
  constexpr uint32_t XPT = TP * GB_XT_CMAX / GB_XT_THREADS;  // x values a thread stages
  constexpr uint32_t NONE = 0xffffffffu;
  V *xs = (V *)L.xs;
  V *ys = (V *)L.ys;
  V *wcarry = (V *)L.wcarry;
  V *whead = (V *)L.whead;
  uint32_t *wfrow = L.wfrow, *wlrow = L.wlrow;
  const uint32_t tid = threadIdx.x, lane = tid & (GB_WAVE - 1), wv = tid / GB_WAVE;
  const uint32_t r0 = blockIdx.x * R;
  const uint32_t nr = min(R, nrows - r0);
  const uint32_t C = 1u << lc;
  // with two tiles a step, the second tile's runs add into their own
  // accumulator (the y region's unused half in FP32): a row can have a run
  // in each tile, and two wavefronts must not update one word at once
  V *ys2 = ys + GB_XT_RMAX;
  for (uint32_t i = tid; i < nr; i += GB_XT_THREADS) {
    ys[i] = (V)0;
    if (TP > 1) ys2[i] = (V)0;
  }
  const uint32_t *gp = ptr + (size_t)blockIdx.x * T;
  // a lane holds E consecutive entries, read as one vector: 2 in FP64, 4 in
  // FP32 (measured on the Zipf matrix: FP64 1.61 ms at 2, 1.66 ms at 4;
  // FP32 1.23 ms at 2, 1.13 ms at 4)
  constexpr uint32_t E = sizeof(V) == 8 ? 2 : 4;
  typedef uint32_t uE __attribute__((ext_vector_type(E)));
  typedef V vE __attribute__((ext_vector_type(E)));
  constexpr uint32_t CH = E * GB_WAVE;  // entries per chunk
  const uE *__restrict__ idxE = (const uE *)idx;
  const vE *__restrict__ valE = (const vE *)val;
  V xr[XPT];                // the next group's tile of x
  uE rw[GB_XT_DEPTH];       // chunks in flight: packed words and values
  vE rv[GB_XT_DEPTH];
  uE rw2[GB_XT_DEPTH];      // the next group's first chunks
  vE rv2[GB_XT_DEPTH];
  uint32_t wprev = 0, wprev2 = 0;  // the entry before the range (is its first
                                   // run cut?)
  // the group pointers are read through the constant address space: the
  // same words for every lane, loaded by the scalar unit, so no vector load
  // (and no wait on the vector loads in flight) comes between groups
  typedef const __attribute__((address_space(4))) uint32_t gb_const_u32;
  const gb_const_u32 *gpc = (const gb_const_u32 *)(uintptr_t)gp;
  const uint32_t U = (T + TP - 1) / TP;  // steps
  // this wavefront's range of step tt
  auto range = [&](uint32_t tt, uint32_t &ra, uint32_t &rz) {
    const uint32_t gs = gpc[tt * TP], ge = gpc[min(T, tt * TP + TP)];
    const uint32_t cnt = ge - gs;
    const uint32_t ew = (cnt + NW - 1) / NW;
    ra = gs + min(cnt, wv * ew);
    rz = min(ra + ew, ge);
  };
  // group tt's tile of x into registers, and its first chunks (from the
  // multiple of E at or before the range's start).  Every load is
  // unconditional, clamped to the range's last E-group (the extra values
  // are never used), so no load sits under a branch and the waits for them
  // stay exact.
  auto prefetch = [&](uint32_t tt, uE (&pw)[GB_XT_DEPTH], vE (&pv)[GB_XT_DEPTH],
                      uint32_t &pprev) {
    const uint32_t c0 = (tt * TP) << lc;
    const uint32_t cl = min(TP * C, ncols - c0) - 1;
#pragma unroll
    for (uint32_t i = 0; i < XPT; ++i) xr[i] = x[c0 + min(i * GB_XT_THREADS + tid, cl)];
    uint32_t ra, rz;
    range(tt, ra, rz);
    const uint32_t pl = (rz > ra ? rz - 1 : ra) / E;
    pprev = idx[ra > 0 ? ra - 1 : 0];
#pragma unroll
    for (uint32_t q = 0; q < GB_XT_DEPTH; ++q) {
      const uint32_t p = min(ra / E + q * GB_WAVE + lane, pl);
      pw[q] = __builtin_nontemporal_load(&idxE[p]);
      pv[q] = __builtin_nontemporal_load(&valE[p]);
    }
  };
  // the next group with entries after group tt (or T); the pointer words
  // are read as the same for every lane, so the loop stays scalar
  auto next_group = [&](uint32_t tt) {
    while (tt < U && gpc[tt * TP] == gpc[min(T, tt * TP + TP)]) ++tt;
    return tt;
  };
  uint32_t t = next_group(0);
  if (t < U) prefetch(t, rw, rv, wprev);
  while (t < U) {
    {
      const uint32_t cw = min(TP * C, ncols - ((t * TP) << lc));
#pragma unroll
      for (uint32_t i = 0; i < XPT; ++i) {
        const uint32_t c = i * GB_XT_THREADS + tid;
        if (c < cw) xs[c] = xr[i];
      }
    }
    __syncthreads();
    // the next group's x and first chunks go out now: they have this
    // group's work to arrive
    const uint32_t tn = next_group(t + 1);
    if (tn < U) prefetch(tn, rw2, rv2, wprev2);
    const uint32_t s = gpc[t * TP];
    // entries from mid on belong to the step's second tile
    const uint32_t mid = gpc[min(T, t * TP + 1)];
    auto key = [&](uint32_t w, uint32_t e) {
      if constexpr (TP == 1) return w >> lc;
      else return (w >> lc) | (e >= mid ? 0x8000u : 0u);
    };
    auto xcol = [&](uint32_t w, uint32_t e) {
      if constexpr (TP == 1) return w & (C - 1);
      else return (e >= mid ? C : 0u) | (w & (C - 1));
    };
    auto yat = [&](uint32_t k) -> V & {
      if constexpr (TP == 1) return ys[k];
      else return (k & 0x8000u) ? ys2[k & 0x7fffu] : ys[k];
    };
    uint32_t a, z;
    range(t, a, z);
    uint32_t fw = 0, lw = 0;
    if (a < z) {  // the same for the whole wavefront
      uint32_t w0 = rw[0][0];
#pragma unroll
      for (uint32_t i = 1; i < E; ++i)
        if (a % E == i) w0 = rw[0][i];
      const uint32_t rfirst = key((uint32_t)__builtin_amdgcn_readfirstlane((int)w0), a);
      const bool cut = a > s && key(wprev, a - 1) == rfirst;
      uint32_t crow = NONE;  // the run carried from the previous chunk
      V csum = (V)0;
      // a run's sum into y, or, for the range's first run when it continues
      // a run of the range before, into this wavefront's carry word
      auto put = [&](uint32_t row, V t) {
        if (row == rfirst && cut) wcarry[wv] = t;
        else yat(row) += t;
      };
      const uint32_t kb0 = a - a % E;
      for (uint32_t kb = kb0; kb < z; kb += GB_XT_DEPTH * CH) {
#pragma unroll
        for (uint32_t q = 0; q < GB_XT_DEPTH; ++q) {
          const uint32_t k0 = kb + q * CH;
          if (k0 < z) {
            const uE w = rw[q];
            const vE v = rv[q];
            const uint32_t e0 = k0 + E * lane;
            // entries outside the range get rows no entry has (valid rows
            // are < 2^13); the scan moves sums only to higher lanes and an
            // invalid entry is never joined to a valid one, so theirs
            // (products of other ranges' entries) reach no valid run
            uint32_t r[E];
            V p[E];
#pragma unroll
            for (uint32_t i = 0; i < E; ++i) {
              const bool ok = e0 + i >= a && e0 + i < z;
              r[i] = ok ? key(w[i], e0 + i) : (0x80000000u | (E * lane + i));
              p[i] = v[i] * xs[xcol(w[i], e0 + i)];
            }
            // the run carried from the previous chunk (whose entries end
            // just before this chunk's first, lane 0's first entry)
            const uint32_t rl0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)r[0]);
            if (crow != NONE) {
              if (crow == rl0) {
                if (lane == 0) p[0] = csum + p[0];
              } else if (lane == 0) {
                put(crow, csum);  // it ended with the previous chunk
              }
              crow = NONE;
            }
            // the lane's runs left to right: acc[i] = the lane's part of the
            // run ending at entry i
            V acc[E];
            acc[0] = p[0];
#pragma unroll
            for (uint32_t i = 1; i < E; ++i) acc[i] = r[i] == r[i - 1] ? acc[i - 1] + p[i] : p[i];
            // the runs across lanes, keyed by each lane's last row
            const V sc = gb_seg_scan64<V>(acc[E - 1], r[E - 1], lane);
            // a run ending inside the lane that began at its first entry
            // takes the previous lane's scan when it continues it
            const V sp = gb_dpp_val<0x138>(sc);      // wave_shr:1
            const uint32_t rp = gb_dpp_u32<0x138>(r[E - 1]);
            const bool cont = rp == r[0];
            V tot[E];
#pragma unroll
            for (uint32_t i = 0; i + 1 < E; ++i)
              tot[i] = cont && r[i] == r[0] ? sp + acc[i] : acc[i];
            tot[E - 1] = sc;
            const uint32_t rn = gb_dpp_u32<0x130>(r[0]);  // wave_shl:1
            // the chunk's last entry in the range: its run goes on
            const uint32_t el = min(k0 + CH, z) - 1 - k0;
            const uint32_t ll = el / E, jl = el % E;
#pragma unroll
            for (uint32_t i = 0; i < E; ++i) {
              const bool ok = e0 + i >= a && e0 + i < z;
              const bool ends = i + 1 < E ? r[i] != r[i + 1] : r[i] != rn;
              if (ok && ends && !(lane == ll && jl == i)) put(r[i], tot[i]);
            }
            uint32_t rlast = r[0];
            V slast = tot[0];
#pragma unroll
            for (uint32_t i = 1; i < E; ++i)
              if (jl == i) {
                rlast = r[i];
                slast = tot[i];
              }
            crow = (uint32_t)__builtin_amdgcn_readlane((int)rlast, ll);
            csum = gb_readlane_val(slast, ll);
          }
          // the chunk GB_XT_DEPTH further on takes this slot, once this
          // chunk's values are used (the same registers, no copies)
          const uint32_t pf = min(k0 / E + GB_XT_DEPTH * GB_WAVE + lane, (z - 1) / E);
          rw[q] = __builtin_nontemporal_load(&idxE[pf]);
          rv[q] = __builtin_nontemporal_load(&valE[pf]);
        }
      }
      const bool single = crow == rfirst && cut;
      if (single && lane == 0) wcarry[wv] = csum;
      if (!single && lane == 0) whead[wv] = csum;
      fw = rfirst | (cut ? 0x80000000u : 0u) | (single ? 0x40000000u : 0u);
      lw = single ? 0u : (crow | 0x80000000u);
    }
    if (lane == 0) {
      wfrow[wv] = fw;
      wlrow[wv] = lw;
    }
    __syncthreads();
    // a run started in this range and cut by its end takes the next
    // ranges' parts, in order
    if (lane == 0 && (lw & 0x80000000u)) {
      const uint32_t r = lw & 0x7fffffffu;
      V tot = whead[wv];
      for (uint32_t q = wv + 1; q < NW; ++q) {
        const uint32_t f = wfrow[q];
        if (!(f & 0x80000000u) || (f & 0x3fffffffu) != r) break;
        tot += wcarry[q];
        if (!(f & 0x40000000u)) break;
      }
      yat(r) += tot;
    }
    // no barrier here: every read of xs ended before the barrier above, and
    // the next group's range words and ys updates come after the next
    // group's first barrier, which every wavefront reaches after this step
#pragma unroll
    for (uint32_t q = 0; q < GB_XT_DEPTH; ++q) {
      rw[q] = rw2[q];
      rv[q] = rv2[q];
    }
    wprev = wprev2;
    t = tn;
  }
  __syncthreads();  // the last group's chained runs are in ys
  if constexpr (EPI == GB_XT_STORE) {
    for (uint32_t i = tid; i < nr; i += GB_XT_THREADS)
      y[(size_t)(r0 + i) * ystride] = TP > 1 ? ys[i] + ys2[i] : ys[i];
  } else {
    // the block's partial: each thread's rows in order, then the
    // wavefronts' sums in order
    double d = 0.0;
    for (uint32_t i = tid; i < nr; i += GB_XT_THREADS) {
      const V yi = TP > 1 ? ys[i] + ys2[i] : ys[i];
      if constexpr (EPI == GB_XT_DOT) {
        y[r0 + i] = yi;
        d += (double)x[r0 + i] * (double)yi;
      } else {
        const double o = b[r0 + i] - (double)yi;
        ((double *)y)[r0 + i] = o;
        d += o * o;
      }
    }
    d = gb_wave_sum(d);
    if (lane == 0) L.red[wv] = d;
    __syncthreads();
    if (tid == 0) {
      double s = L.red[0];
#pragma unroll
      for (uint32_t q = 1; q < NW; ++q) s += L.red[q];
      partials[blockIdx.x] = s;
    }
  }
}
