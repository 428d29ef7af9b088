// Host check of the text conversions in gb_strtod.h against the C
// library: for every string, glibc's strtod and strtoull(.., 10) and the
// restated conversions must give the same bits and the same end.
#define GB_HD
#include "gb_strtod.h"

#include <stdlib.h>
#include <string.h>

extern "C" {

// buf holds count NUL-terminated strings back to back.  Returns the
// number of mismatches; the indices of the first maxbad are stored in bad;
// *nhard counts the strings that needed the exact fallback.
long gbt_check(const char *buf, long count, long *bad, long maxbad, long *nhard) {
  static gbt_big L, R;
  long nbad = 0;
  *nhard = 0;
  const char *s = buf;
  for (long i = 0; i < count; ++i) {
    const uint64_t n = strlen(s);
    char *e = nullptr;
    const double v = strtod(s, &e);
    bool hard = false;
    double g = 0;
    const uint64_t end = gbt_strtod(s, n, 0, &g, &hard);
    if (hard) {
      g = gbt_strtod_exact(s, n, 0, &L, &R);
      ++*nhard;
    }
    char *e2 = nullptr;
    const unsigned long long u = strtoull(s, &e2, 10);
    uint64_t gu = 0;
    const uint64_t end2 = gbt_strtoull(s, n, 0, 10, &gu);
    const bool ok = gbt_u64(v) == gbt_u64(g) && (uint64_t)(e - s) == end &&
                    u == gu && (uint64_t)(e2 - s) == end2;
    if (!ok) {
      if (nbad < maxbad) bad[nbad] = i;
      ++nbad;
    }
    s += n + 1;
  }
  return nbad;
}

// one string: the restated conversions' results, for diagnostics
void gbt_one(const char *s, double *val, uint64_t *end, int *hard, uint64_t *uval,
             uint64_t *uend) {
  static gbt_big L, R;
  const uint64_t n = strlen(s);
  bool h = false;
  *end = gbt_strtod(s, n, 0, val, &h);
  if (h) *val = gbt_strtod_exact(s, n, 0, &L, &R);
  *hard = h;
  *uend = gbt_strtoull(s, n, 0, 10, uval);
}
}
