// gb_strtod.h -- the text-to-number conversions the reference's CSR reader
// uses (src/io.c:415-446: strtoull(.., 10) for a column, strtod for a
// value), restated as plain functions on a byte range so that the same
// code runs in a GPU kernel and on the host.  The results are those of the
// C library in the "C" locale: the same value, bit for bit, and the same
// end position.
//
// The including file defines GB_HD (the function qualifiers: "__host__
// __device__" in HIP sources, empty in a host-only build).
//
// Text is addressed as s[0 .. n); bytes at or past n read as NUL, and a
// NUL byte ends every conversion (as it ends a C string).
//
// strtod, decimal form: digits are gathered into a 64-bit integer w (19
// significant digits at most) and a power of ten q; w * 10^q is then
// rounded by the Eisel-Lemire method (a 64 x 128-bit product with a
// table of 5^q, gb_pow5.h), which decides every case with at most 19
// digits.  With more digits the method is applied to w and w + 1; when the
// two disagree (the dropped digits decide the rounding), or in the one
// case the method cannot settle, the conversion is marked hard and
// gb_strtod_exact settles it with exact big-integer arithmetic against the
// halfway point between the two neighbouring doubles.  Hexadecimal forms
// (0x...p...), infinities and NaNs (with glibc's payload rule for
// "nan(chars)") are converted exactly with integer arithmetic.
#pragma once
#include <stdint.h>

#include "gb_pow5.h"

#ifndef GB_HD
#error "define GB_HD before including gb_strtod.h"
#endif

GB_HD inline unsigned gbt_at(const char *s, uint64_t n, uint64_t p) {
  return p < n ? (unsigned char)s[p] : 0u;
}
GB_HD inline bool gbt_space(unsigned c) { return c == ' ' || (c >= 9u && c <= 13u); }
GB_HD inline bool gbt_digit(unsigned c) { return c >= '0' && c <= '9'; }
GB_HD inline int gbt_hexval(unsigned c) {
  if (c >= '0' && c <= '9') return (int)(c - '0');
  if (c >= 'a' && c <= 'f') return (int)(c - 'a' + 10);
  if (c >= 'A' && c <= 'F') return (int)(c - 'A' + 10);
  return -1;
}
GB_HD inline unsigned gbt_lower(unsigned c) { return (c >= 'A' && c <= 'Z') ? c + 32u : c; }

GB_HD inline double gbt_bits(uint64_t b) {
  union { uint64_t u; double d; } x;
  x.u = b;
  return x.d;
}
GB_HD inline uint64_t gbt_u64(double d) {
  union { uint64_t u; double d; } x;
  x.d = d;
  return x.u;
}

// ---------------------------------------------------------------- strtoull
// strtoull(s + p, &end, base) for base 10 or 16 (the two the conversions
// need): returns end (p when nothing converts) and the value in *out.
GB_HD inline uint64_t gbt_strtoull(const char *s, uint64_t n, uint64_t p, int base,
                                   uint64_t *out) {
  uint64_t q = p;
  while (gbt_space(gbt_at(s, n, q))) ++q;
  bool neg = false;
  unsigned c = gbt_at(s, n, q);
  if (c == '+' || c == '-') {
    neg = c == '-';
    ++q;
  }
  uint64_t v = 0;
  bool ovf = false;
  const uint64_t d0 = q;
  for (;;) {
    c = gbt_at(s, n, q);
    int d = base == 16 ? gbt_hexval(c) : (gbt_digit(c) ? (int)(c - '0') : -1);
    if (d < 0) break;
    if (!ovf) {
      if (v > (~0ull - (uint64_t)d) / (uint64_t)base) ovf = true;
      else v = v * (uint64_t)base + (uint64_t)d;
    }
    ++q;
  }
  if (q == d0) {
    *out = 0;
    return p;
  }
  *out = ovf ? ~0ull : (neg ? (uint64_t)0 - v : v);
  return q;
}

// strtoull(.., 0) on the n-char-sequence of "nan(...)": "0x"/"0X" hex,
// a leading 0 octal, else decimal; returns end within [p, e)
GB_HD inline uint64_t gbt_strtoull0(const char *s, uint64_t e, uint64_t p, uint64_t *out) {
  uint64_t q = p;
  unsigned base = 10;
  if (gbt_at(s, e, q) == '0') {
    if (gbt_lower(gbt_at(s, e, q + 1)) == 'x' && gbt_hexval(gbt_at(s, e, q + 2)) >= 0) {
      base = 16;
      q += 2;
    } else {
      base = 8;
    }
  }
  uint64_t v = 0;
  bool ovf = false;
  const uint64_t d0 = q;
  for (;;) {
    int d = gbt_hexval(gbt_at(s, e, q));
    if (d < 0 || (unsigned)d >= base) break;
    if (!ovf) {
      if (v > (~0ull - (uint64_t)d) / base) ovf = true;
      else v = v * base + (uint64_t)d;
    }
    ++q;
  }
  *out = ovf ? ~0ull : v;
  return q == d0 && base != 8 ? p : q;
}

// ------------------------------------------------------------ rounding
// The double nearest to (m + sticky) * 2^e2, m > 0, sticky meaning "a
// little more than m" (ties to even; sticky breaks a tie upward).
GB_HD inline double gbt_round_binary(uint64_t m, int64_t e2, bool sticky, bool neg) {
  const uint64_t sgn = neg ? (1ull << 63) : 0ull;
  int lz = __builtin_clzll(m);
  m <<= lz;
  e2 -= lz;                      // value = m 2^e2, m in [2^63, 2^64)
  const int64_t lead = e2 + 63;  // exponent of the leading bit
  if (lead > 1023) return gbt_bits(sgn | 0x7ff0000000000000ull);
  int64_t shift = lead >= -1022 ? 11 : 11 + (-1022 - lead);
  if (shift > 64) return gbt_bits(sgn);
  uint64_t kept, rem, half;
  if (shift == 64) {
    kept = 0;
    rem = m;
    half = 1ull << 63;
  } else {
    kept = m >> shift;
    rem = m & ((1ull << shift) - 1);
    half = 1ull << (shift - 1);
  }
  bool up = rem > half || (rem == half && (sticky || (kept & 1)));
  kept += up ? 1 : 0;
  uint64_t bits;
  if (lead >= -1022) bits = ((uint64_t)(lead + 1023) << 52) + (kept - (1ull << 52));
  else bits = kept;  // subnormal (kept == 2^52 encodes the least normal)
  return gbt_bits(sgn | bits);
}

// ------------------------------------------------------- Eisel-Lemire
struct gbt_u128 {
  uint64_t hi, lo;
};
GB_HD inline gbt_u128 gbt_mul64(uint64_t a, uint64_t b) {
  unsigned __int128 r = (unsigned __int128)a * b;
  gbt_u128 x;
  x.hi = (uint64_t)(r >> 64);
  x.lo = (uint64_t)r;
  return x;
}

// w 10^q rounded to nearest-even as raw bits (positive), for w >= 1 with
// at most 64 bits.  Returns false when the 128-bit product cannot decide;
// *bits is then within one unit of the result.
GB_HD inline bool gbt_eisel_lemire(uint64_t w, int64_t q, uint64_t *bits) {
  if (w == 0 || q < GB_POW5_QMIN) {
    *bits = 0;
    return true;
  }
  if (q > GB_POW5_QMAX) {
    *bits = 0x7ff0000000000000ull;
    return true;
  }
  const int lz = __builtin_clzll(w);
  w <<= lz;
  const int idx = 2 * (int)(q - GB_POW5_QMIN);
  gbt_u128 p = gbt_mul64(w, gb_pow5_128[idx]);
  bool sure = true;
  // 52 explicit bits + hidden + rounding + possible shift = 55 bits: the
  // low 9 bits of the high word all ones means a carry may still come
  if ((p.hi & 0x1ffull) == 0x1ffull) {
    gbt_u128 p2 = gbt_mul64(w, gb_pow5_128[idx + 1]);
    p.lo += p2.hi;
    if (p2.hi > p.lo) p.hi++;
    if (p.lo == ~0ull && !(q >= -27 && q <= 55)) sure = false;
  }
  const int upper = (int)(p.hi >> 63);
  const int shift = upper + 9;
  uint64_t mant = p.hi >> shift;
  // floor(q log2 10) + 63 via 217706 / 2^16 ~ log2 10
  int64_t pow2 = (((217706ll) * q) >> 16) + 63 + upper - lz + 1023;
  if (pow2 <= 0) {
    if (-pow2 + 1 >= 64) {
      *bits = 0;
      return true;
    }
    mant >>= -pow2 + 1;
    mant += mant & 1;
    mant >>= 1;
    pow2 = mant < (1ull << 52) ? 0 : 1;
    *bits = ((uint64_t)pow2 << 52) | (mant & ((1ull << 52) - 1));
    return sure;
  }
  // an exact halfway product rounds to even instead of up
  if (p.lo <= 1 && q >= -4 && q <= 23 && (mant & 3) == 1 && (mant << shift) == p.hi)
    mant &= ~1ull;
  mant += mant & 1;
  mant >>= 1;
  if (mant >= (2ull << 52)) {
    mant = 1ull << 52;
    pow2++;
  }
  mant &= ~(1ull << 52);
  if (pow2 >= 0x7ff) {
    *bits = 0x7ff0000000000000ull;
    return sure;
  }
  *bits = ((uint64_t)pow2 << 52) | mant;
  return sure;
}

// ------------------------------------------------------------- strtod
// Result of a decimal scan: value = (digits) * 10^(q) with the leading 19
// significant digits in w.
struct gbt_dec {
  uint64_t w;
  int64_t q;
  bool truncated;  // a nonzero significant digit beyond the 19th
  uint64_t mstart, mend;  // the mantissa text (digits and point)
  int64_t exp10;          // the exponent part's value (saturated)
};

// Scans [digits][.digits][(e|E)[sign]digits] at p (no sign, no space);
// returns the end, or p when no digit is present.
GB_HD inline uint64_t gbt_scan_decimal(const char *s, uint64_t n, uint64_t p, gbt_dec *d) {
  uint64_t q = p;
  uint64_t w = 0;
  int nsig = 0;
  int64_t dropped = 0, nfrac = 0;
  bool trunc = false, any = false, frac = false;
  for (;;) {
    unsigned c = gbt_at(s, n, q);
    if (gbt_digit(c)) {
      any = true;
      unsigned dig = c - '0';
      if (nsig == 0 && dig == 0) {
        // leading zero
      } else if (nsig < 19) {
        w = w * 10 + dig;
        ++nsig;
      } else {
        ++dropped;
        if (dig) trunc = true;
      }
      if (frac) ++nfrac;
      ++q;
    } else if (c == '.' && !frac) {
      frac = true;
      ++q;
    } else {
      break;
    }
  }
  if (!any) return p;
  d->mstart = p;
  d->mend = q;
  int64_t ex = 0;
  if (gbt_lower(gbt_at(s, n, q)) == 'e') {
    uint64_t r = q + 1;
    bool eneg = false;
    unsigned c = gbt_at(s, n, r);
    if (c == '+' || c == '-') {
      eneg = c == '-';
      ++r;
    }
    if (gbt_digit(gbt_at(s, n, r))) {
      while (gbt_digit(c = gbt_at(s, n, r))) {
        if (ex < 100000000) ex = ex * 10 + (c - '0');
        ++r;
      }
      if (eneg) ex = -ex;
      q = r;
    }
  }
  // dropped and nfrac are bounded by the text length; keep q in range
  int64_t qq = dropped - nfrac + ex;
  if (qq > 100000000) qq = 100000000;
  if (qq < -100000000) qq = -100000000;
  d->w = w;
  d->q = qq;
  d->truncated = trunc;
  d->exp10 = ex;
  return q;
}

// Hexadecimal after "0x": hex digits [. hex digits] [(p|P)[sign]digits];
// returns the end and the value (exact rounding).
GB_HD inline uint64_t gbt_scan_hex(const char *s, uint64_t n, uint64_t p, bool neg,
                                   double *out) {
  uint64_t q = p;
  uint64_t m = 0;
  int nsig = 0;
  int64_t e2 = 0;
  bool sticky = false, frac = false;
  for (;;) {
    unsigned c = gbt_at(s, n, q);
    int h = gbt_hexval(c);
    if (h >= 0) {
      if (nsig == 0 && h == 0) {
        if (frac) e2 -= 4;
      } else if (nsig < 16) {
        m = (m << 4) | (uint64_t)h;
        ++nsig;
        if (frac) e2 -= 4;
      } else {
        if (h) sticky = true;
        if (!frac) e2 += 4;
      }
      ++q;
    } else if (c == '.' && !frac) {
      frac = true;
      ++q;
    } else {
      break;
    }
  }
  if (gbt_lower(gbt_at(s, n, q)) == 'p') {
    uint64_t r = q + 1;
    bool eneg = false;
    unsigned c = gbt_at(s, n, r);
    if (c == '+' || c == '-') {
      eneg = c == '-';
      ++r;
    }
    if (gbt_digit(gbt_at(s, n, r))) {
      int64_t ex = 0;
      while (gbt_digit(c = gbt_at(s, n, r))) {
        if (ex < 100000000) ex = ex * 10 + (c - '0');
        ++r;
      }
      e2 += eneg ? -ex : ex;
      q = r;
    }
  }
  if (m == 0) *out = gbt_bits(neg ? (1ull << 63) : 0ull);
  else *out = gbt_round_binary(m, e2, sticky, neg);
  return q;
}

// strtod(s + p, &end): returns end (p when nothing converts).  *hard set:
// the dropped digits decide the rounding; *out is then a neighbour of the
// result and gb_strtod_exact gives the result.  want_value false: only the
// end is needed (a decimal value is not rounded, *out is 0).
GB_HD inline uint64_t gbt_strtod(const char *s, uint64_t n, uint64_t p, double *out,
                                 bool *hard, bool want_value = true) {
  *hard = false;
  uint64_t q = p;
  while (gbt_space(gbt_at(s, n, q))) ++q;
  bool neg = false;
  unsigned c = gbt_at(s, n, q);
  if (c == '+' || c == '-') {
    neg = c == '-';
    ++q;
    c = gbt_at(s, n, q);
  }
  const uint64_t sgn = neg ? (1ull << 63) : 0ull;
  const unsigned lc = gbt_lower(c);
  if (lc == 'i') {
    if (gbt_lower(gbt_at(s, n, q + 1)) == 'n' && gbt_lower(gbt_at(s, n, q + 2)) == 'f') {
      q += 3;
      const char *rest = "inity";
      uint64_t r = q;
      int k = 0;
      while (k < 5 && gbt_lower(gbt_at(s, n, r)) == (unsigned)rest[k]) {
        ++r;
        ++k;
      }
      if (k == 5) q = r;
      *out = gbt_bits(sgn | 0x7ff0000000000000ull);
      return q;
    }
    *out = 0.0;
    return p;
  }
  if (lc == 'n') {
    if (gbt_lower(gbt_at(s, n, q + 1)) == 'a' && gbt_lower(gbt_at(s, n, q + 2)) == 'n') {
      q += 3;
      uint64_t bits = 0x7ff8000000000000ull;
      if (gbt_at(s, n, q) == '(') {
        uint64_t r = q + 1;
        for (;;) {
          unsigned k = gbt_at(s, n, r);
          if (gbt_digit(k) || (k >= 'a' && k <= 'z') || (k >= 'A' && k <= 'Z') || k == '_')
            ++r;
          else
            break;
        }
        if (gbt_at(s, n, r) == ')') {
          uint64_t mant = 0;
          uint64_t e = gbt_strtoull0(s, r, q + 1, &mant);
          if (e == r) {
            const uint64_t pay = mant & ((1ull << 51) - 1);
            if (pay) bits = 0x7ff8000000000000ull | pay;
          }
          q = r + 1;
        }
      }
      *out = gbt_bits(sgn | bits);
      return q;
    }
    *out = 0.0;
    return p;
  }
  if (c == '0' && gbt_lower(gbt_at(s, n, q + 1)) == 'x') {
    const unsigned c2 = gbt_at(s, n, q + 2);
    if (gbt_hexval(c2) >= 0 || (c2 == '.' && gbt_hexval(gbt_at(s, n, q + 3)) >= 0))
      return gbt_scan_hex(s, n, q + 2, neg, out);
    // "0x" without hex digits: the "0" alone
    *out = gbt_bits(sgn);
    return q + 1;
  }
  gbt_dec d;
  const uint64_t e = gbt_scan_decimal(s, n, q, &d);
  if (e == q) {
    *out = 0.0;
    return p;
  }
  if (d.w == 0 || !want_value) {
    *out = gbt_bits(sgn);
    return e;
  }
  uint64_t b0;
  if (!d.truncated && d.q == 0 && d.w <= (1ull << 53)) {
    *out = neg ? -(double)d.w : (double)d.w;
    return e;
  }
  bool ok = gbt_eisel_lemire(d.w, d.q, &b0);
  if (ok && d.truncated) {
    uint64_t b1;
    ok = gbt_eisel_lemire(d.w + 1, d.q, &b1) && b1 == b0;
  }
  if (!ok) *hard = true;
  *out = gbt_bits(sgn | b0);
  return e;
}

// ------------------------------------------------ exact decimal fallback
// Little-endian base-2^32 integers large enough for every comparison the
// fallback makes (digits of a double's halfway point scaled to integers:
// under 3900 bits).
#define GBT_LIMBS 136
struct gbt_big {
  uint32_t v[GBT_LIMBS];
  int n;  // limbs in use
  bool ovf;
};
GB_HD inline void gbt_big_set(gbt_big *a, uint64_t x) {
  a->v[0] = (uint32_t)x;
  a->v[1] = (uint32_t)(x >> 32);
  a->n = a->v[1] ? 2 : (a->v[0] ? 1 : 0);
  a->ovf = false;
}
GB_HD inline void gbt_big_muladd(gbt_big *a, uint32_t m, uint32_t add) {
  uint64_t carry = add;
  for (int i = 0; i < a->n; ++i) {
    uint64_t t = (uint64_t)a->v[i] * m + carry;
    a->v[i] = (uint32_t)t;
    carry = t >> 32;
  }
  if (carry) {
    if (a->n < GBT_LIMBS) a->v[a->n++] = (uint32_t)carry;
    else a->ovf = true;
  }
}
GB_HD inline void gbt_big_pow10(gbt_big *a, int64_t k) {
  while (k >= 9) {
    gbt_big_muladd(a, 1000000000u, 0);
    k -= 9;
  }
  uint32_t m = 1;
  while (k-- > 0) m *= 10;
  if (m > 1) gbt_big_muladd(a, m, 0);
}
GB_HD inline void gbt_big_shl(gbt_big *a, int64_t k) {
  if (a->n == 0 || k <= 0) return;
  const int64_t limbs = k / 32;
  const int bits = (int)(k % 32);
  if (a->n + limbs + 1 > GBT_LIMBS) {
    a->ovf = true;
    return;
  }
  int nn = a->n + (int)limbs;
  a->v[nn] = 0;
  for (int i = a->n - 1; i >= 0; --i) {
    const uint64_t t = (uint64_t)a->v[i] << bits;
    a->v[i + limbs + 1] |= (uint32_t)(t >> 32);
    a->v[i + limbs] = (uint32_t)t;
  }
  for (int64_t i = 0; i < limbs; ++i) a->v[i] = 0;
  a->n = a->v[nn] ? nn + 1 : nn;
}
GB_HD inline int gbt_big_cmp(const gbt_big *a, const gbt_big *b) {
  if (a->n != b->n) return a->n < b->n ? -1 : 1;
  for (int i = a->n - 1; i >= 0; --i)
    if (a->v[i] != b->v[i]) return a->v[i] < b->v[i] ? -1 : 1;
  return 0;
}

// sign of (D 10^e10) - (2 m + 1) 2^(e2 - 1): the decimal against the
// halfway point above m 2^e2
GB_HD inline int gbt_cmp_halfway(const char *s, uint64_t mstart, uint64_t mend, int64_t ex,
                                 uint64_t m, int64_t e2, bool *sticky, gbt_big *L,
                                 gbt_big *R) {
  // D: up to 800 significant digits; later nonzero digits are a sticky
  // amount below one unit of the last kept digit
  gbt_big_set(L, 0);
  int nsig = 0;
  int64_t dropped = 0, nfrac = 0;
  bool frac = false;
  uint32_t chunk = 0, cm = 1;
  *sticky = false;
  for (uint64_t q = mstart; q < mend; ++q) {
    const unsigned c = (unsigned char)s[q];
    if (c == '.') {
      frac = true;
      continue;
    }
    const unsigned dig = c - '0';
    if (frac) ++nfrac;
    if (nsig == 0 && dig == 0) continue;
    if (nsig < 800) {
      chunk = chunk * 10 + dig;
      cm *= 10;
      ++nsig;
      if (cm == 1000000000u) {
        gbt_big_muladd(L, cm, chunk);
        chunk = 0;
        cm = 1;
      }
    } else {
      ++dropped;
      if (dig) *sticky = true;
    }
  }
  if (cm > 1) gbt_big_muladd(L, cm, chunk);
  const int64_t e10 = dropped - nfrac + ex;
  gbt_big_set(R, 2 * m + 1);
  if (e10 >= 0) gbt_big_pow10(L, e10);
  else gbt_big_pow10(R, -e10);
  if (e2 - 1 >= 0) gbt_big_shl(R, e2 - 1);
  else gbt_big_shl(L, 1 - e2);
  return gbt_big_cmp(L, R);
}

// (m, e2) of a finite non-negative double's bits: value = m 2^e2
GB_HD inline void gbt_decompose(uint64_t bits, uint64_t *m, int64_t *e2) {
  const uint64_t f = bits & ((1ull << 52) - 1);
  const int64_t be = (int64_t)(bits >> 52);
  if (be == 0) {
    *m = f;
    *e2 = -1074;
  } else {
    *m = f | (1ull << 52);
    *e2 = be - 1075;
  }
}

// The decimal conversion of a hard case, exactly.  Re-scans the text at p
// (as gbt_strtod did, so the end is the same) and moves from the
// Eisel-Lemire neighbour to the correctly rounded double by halfway
// comparisons.  L and R are the caller's scratch integers.
GB_HD inline double gbt_strtod_exact(const char *s, uint64_t n, uint64_t p, gbt_big *L,
                                     gbt_big *R) {
  uint64_t q = p;
  while (gbt_space(gbt_at(s, n, q))) ++q;
  bool neg = false;
  unsigned c = gbt_at(s, n, q);
  if (c == '+' || c == '-') {
    neg = c == '-';
    ++q;
  }
  gbt_dec d;
  gbt_scan_decimal(s, n, q, &d);
  uint64_t b = 0;
  gbt_eisel_lemire(d.w, d.q, &b);  // within an ulp of the result
  const uint64_t INF = 0x7ff0000000000000ull;
  for (int guard = 0; guard < 8; ++guard) {
    bool sticky = false;
    uint64_t m;
    int64_t e2;
    // step up while the value is above the halfway point over b
    if (b < INF) {
      gbt_decompose(b, &m, &e2);
      int cmp = gbt_cmp_halfway(s, d.mstart, d.mend, d.exp10, m, e2, &sticky, L, R);
      if (cmp > 0 || (cmp == 0 && (sticky || (m & 1)))) {
        ++b;
        continue;
      }
    }
    // step down while the value is below the halfway point under b
    if (b > 0) {
      const uint64_t a = b - 1;
      gbt_decompose(a, &m, &e2);
      int cmp = gbt_cmp_halfway(s, d.mstart, d.mend, d.exp10, m, e2, &sticky, L, R);
      if (cmp < 0 || (cmp == 0 && !sticky && !(m & 1))) {
        b = a;
        continue;
      }
    }
    break;
  }
  return gbt_bits((neg ? (1ull << 63) : 0ull) | b);
}
