/*
 * gb_tune.c -- SpMV plan autotuner.
 *
 * The candidates come from the matrix's row-length histogram (rows with
 * floor(log2(len + 1)) == b): only thresholds that move rows between the
 * LDS row-block, wavefront-per-row and workgroup-per-row bins are tried.
 * Each candidate is built on the device (gbd_matrix_replan) and timed with
 * HIP events over `reps` SpMVs after one warm-up launch; the search is a
 * coordinate descent (short_max, then window x rowcap, then wave_max, then
 * the number of column slices for a gathered vector larger than half an
 * XCD's L2) from
 * the current plan, and a candidate is kept when it is at least 3% faster
 * than the best so far.  Every plan gives the
 * same results to the SpMV's stated tolerance (exact plans: bitwise), so
 * tuning changes speed only.
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "gb_host.h"

#define GB_TUNE_LDS 2304   /* window + short_max must fit the LDS stage */
/* a candidate replaces the best plan only when it is this much faster
 * (timing noise of short launches must not pick a plan) */
#define GB_TUNE_GAIN 0.97

typedef struct {
  gbd_matrix *m;
  int precision;
  void *x, *y;
  void *e0, *e1;
  uint32_t reps;
} tune_ctx;

static int time_plan(tune_ctx *c, const gbd_plan_params *p, double *ms)
{
  int rv = gbd_matrix_replan(c->m, p);
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  for (uint32_t k = 0; k <= c->reps; ++k) {
    if (k == 1 && (rv = gbd_event_record(c->e0)) != GOOSEBERRY_SUCCESS) return rv;
    rv = c->precision == GOOSEBERRY_PRECISION_MIXED
        ? gbd_spmv_f32(c->m, (const float *)c->x, (float *)c->y)
        : gbd_spmv(c->m, (const double *)c->x, (double *)c->y);
    if (rv != GOOSEBERRY_SUCCESS) return rv;
  }
  if ((rv = gbd_event_record(c->e1)) != GOOSEBERRY_SUCCESS) return rv;
  float t = 0.f;
  if ((rv = gbd_event_elapsed_ms(c->e0, c->e1, &t)) != GOOSEBERRY_SUCCESS) return rv;
  *ms = (double)t / c->reps;
  return GOOSEBERRY_SUCCESS;
}

/* some bucket of the log2 histogram meets the lengths (lo, hi]; bucket b
 * holds the lengths [2^b - 1, 2^(b+1) - 2] */
static int rows_between(const uint64_t *hist, uint64_t lo, uint64_t hi)
{
  for (int b = 0; b < 33; ++b) {
    uint64_t blo = (1ull << b) - 1;
    uint64_t bhi = b >= 32 ? 0xffffffffull : (1ull << (b + 1)) - 2;
    if (hist[b] && bhi > lo && blo <= hi) return 1;
  }
  return 0;
}

static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

int gooseberry_matrix_autotune(gooseberry_matrix_t mat, int precision,
    uint32_t reps, gooseberry_plan_t *r_plan, double *r_ms)
{
  GB_API_GUARD();
  if (!mat) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (precision != GOOSEBERRY_PRECISION_FP64 && precision != GOOSEBERRY_PRECISION_MIXED &&
      precision != GOOSEBERRY_PRECISION_MIXED_JACOBI)
    return GOOSEBERRY_ERROR_INVALIDINPUT;
  /* the scaled FP32 values stream like the plain ones: time those */
  if (precision == GOOSEBERRY_PRECISION_MIXED_JACOBI) precision = GOOSEBERRY_PRECISION_MIXED;
  /* a batch slot holds the current plan's arrays */
  if (gbd_matrix_holders(gb_matrix_dev(mat))) return GOOSEBERRY_ERROR_INVALIDINPUT;
  if (gbd_matrix_nparts(gb_matrix_dev(mat))) return gb_parts_unsupported();
  tune_ctx c;
  memset(&c, 0, sizeof(c));
  c.m = gb_matrix_dev(mat);
  c.precision = precision;
  c.reps = reps ? reps : 5;
  uint32_t nrows, ncols;
  uint64_t nnz, hist[33];
  gbd_plan_params cur;
  gbd_matrix_info(c.m, &nrows, &ncols, &nnz, &cur, NULL, NULL, NULL, hist);
  int rv;
  if (precision == GOOSEBERRY_PRECISION_MIXED &&
      (rv = gbd_matrix_make_f32(c.m)) != GOOSEBERRY_SUCCESS)
    return rv;
  const size_t es = precision == GOOSEBERRY_PRECISION_MIXED ? 4 : 8;
  if ((rv = gbd_malloc(&c.x, (size_t)(ncols ? ncols : 1) * es)) != GOOSEBERRY_SUCCESS ||
      (rv = gbd_malloc(&c.y, (size_t)(nrows ? nrows : 1) * es)) != GOOSEBERRY_SUCCESS ||
      (rv = gbd_zero(c.x, (size_t)(ncols ? ncols : 1) * es)) != GOOSEBERRY_SUCCESS ||
      (rv = gbd_event_create(&c.e0)) != GOOSEBERRY_SUCCESS ||
      (rv = gbd_event_create(&c.e1)) != GOOSEBERRY_SUCCESS)
    goto done;
  /* candidates start from the current plan without x tiles (a column
   * split and x tiles exclude each other) */
  cur.xtile = 0;
  gbd_plan_params best = cur;
  double best_ms = DBL_MAX, ms;
  if ((rv = time_plan(&c, &best, &best_ms)) != GOOSEBERRY_SUCCESS) goto done;

  /* 1. the LDS row-block threshold */
  static const uint32_t shorts[] = {8, 16, 32, 64, 128, 256};
  for (size_t i = 0; i < sizeof(shorts) / sizeof(shorts[0]); ++i) {
    gbd_plan_params p = best;
    p.short_max = shorts[i];
    if (p.short_max == best.short_max) continue;
    /* only thresholds that move some rows */
    if (!rows_between(hist, umin(p.short_max, best.short_max),
                      umax(p.short_max, best.short_max)))
      continue;
    if (p.window + p.short_max > GB_TUNE_LDS) p.window = GB_TUNE_LDS - p.short_max;
    if (p.wave_max < p.short_max) p.wave_max = p.short_max;
    if ((rv = time_plan(&c, &p, &ms)) != GOOSEBERRY_SUCCESS) goto done;
    if (ms < GB_TUNE_GAIN * best_ms) { best_ms = ms; best = p; }
  }
  /* 2. the row-block window and row cap */
  static const uint32_t windows[] = {512, 1024, 2048};
  static const uint32_t rowcaps[] = {128, 256, 512};
  if (rows_between(hist, 0, best.short_max)) {
    for (size_t i = 0; i < 3; ++i)
      for (size_t j = 0; j < 3; ++j) {
        gbd_plan_params p = best;
        p.window = windows[i];
        p.rowcap = rowcaps[j];
        if (p.window + p.short_max > GB_TUNE_LDS) continue;
        if (p.window == best.window && p.rowcap == best.rowcap) continue;
        if ((rv = time_plan(&c, &p, &ms)) != GOOSEBERRY_SUCCESS) goto done;
        if (ms < GB_TUNE_GAIN * best_ms) { best_ms = ms; best = p; }
      }
  }
  /* 3. wavefront-per-row against workgroup-per-row for the long rows */
  static const uint32_t waves[] = {512, 1024, 4096, 16384};
  if (rows_between(hist, best.short_max, 0xffffffffull)) {
    for (size_t i = 0; i < 4; ++i) {
      gbd_plan_params p = best;
      p.wave_max = waves[i];
      if (p.wave_max < p.short_max || p.wave_max == best.wave_max) continue;
      if (!rows_between(hist, umin(p.wave_max, best.wave_max),
                        umax(p.wave_max, best.wave_max)))
        continue;
      if ((rv = time_plan(&c, &p, &ms)) != GOOSEBERRY_SUCCESS) goto done;
      if (ms < GB_TUNE_GAIN * best_ms) { best_ms = ms; best = p; }
    }
  }
  /* 4. column slices, for a gathered vector larger than half an XCD's L2 */
  static const uint32_t splits[] = {4, 8, 16, 32};
  if (!best.exact && (uint64_t)ncols * es > (2u << 20)) {
    gbd_plan_params base = best;
    for (size_t i = 0; i < 4; ++i) {
      gbd_plan_params p = base;
      p.csplit = splits[i];
      if ((uint64_t)ncols * es / splits[i] < (256u << 10)) continue;
      if ((rv = time_plan(&c, &p, &ms)) != GOOSEBERRY_SUCCESS) goto done;
      if (ms < GB_TUNE_GAIN * best_ms) { best_ms = ms; best = p; }
    }
    /* 5. x through LDS in column tiles */
    static const uint32_t xtiles[] = {4096, 8192};
    for (size_t i = 0; i < 2; ++i) {
      gbd_plan_params p = base;
      p.csplit = 0;
      p.xtile = xtiles[i];
      if ((rv = time_plan(&c, &p, &ms)) != GOOSEBERRY_SUCCESS) goto done;
      if (ms < GB_TUNE_GAIN * best_ms) { best_ms = ms; best = p; }
    }
  }
  gooseberry_plan_t pub;
  pub.short_max = best.short_max;
  pub.wave_max = best.wave_max;
  pub.window = best.window;
  pub.rowcap = best.rowcap;
  pub.exact = best.exact;
  pub.grid = best.grid;
  pub.xcd_remap = best.xcd_remap;
  pub.csplit = best.csplit;
  pub.wide_index = best.wide_index;
  pub.wide_values = best.wide_values;
  pub.xtile = best.xtile;
  /* through the public replan: it also drops the CG workspace, which is
   * sized by the plan */
  if ((rv = gooseberry_matrix_replan(mat, &pub)) != GOOSEBERRY_SUCCESS) goto done;
  if (r_plan) *r_plan = pub;
  if (r_ms) *r_ms = best_ms;
  rv = GOOSEBERRY_SUCCESS;
done:
  gbd_sync();
  if (c.e0) gbd_event_destroy(c.e0);
  if (c.e1) gbd_event_destroy(c.e1);
  gbd_free(c.x);
  gbd_free(c.y);
  return rv;
}
