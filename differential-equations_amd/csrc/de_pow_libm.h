// pow(x, y) with the bits of the C library that the reference links.
//
// The reference's controller and initial-step heuristic call f64::powf:
//   err.powf(-1/order)            src/methods/erk/dormandprince/ordinary.rs:154
//                                 src/methods/erk/adaptive/ordinary.rs:172
//   (0.01/der12).powf(1/order)    src/methods/h_init.rs:124
// Rust lowers powf to the platform `pow` (glibc on Linux x86-64).  Since glibc
// 2.28 that pow is the table-driven method of ARM's optimized-routines
// (Szabolcs Nagy, 2018; also musl's pow), and on x86-64 CPUs with FMA glibc
// runs the copy of it compiled with -mfma, in which the compiler fuses the
// products it can.  This file restates that method with every fused operation
// written as an explicit fma in the same place and every other operation
// rounded on its own, so the result is the reference binary's pow bit for bit
// on CPU and on gfx950 alike (max error about 0.52 ulp; it is not correctly
// rounded, and parity needs exactly its roundings, not better ones).
//
// Method:
//   x = 2^k z, z in [0x1.69555p-1, 0x1.69555p0); i = top 7 bits of
//   bits(x) - 0x3fe6955500000000; 1/c_i has few bits, so r = z/c_i - 1 is
//   exact in one fma; log x = k ln2 + log c_i + log1p(r) as hi + lo
//   (degree-8 polynomial); e = y log x as ehi + elo; then
//   exp(e) = 2^(m/128) exp(s), e = m ln2/128 + s, with 2^(j/128) = H (1 + T)
//   from a 128-entry table and a degree-5 polynomial for exp(s) - 1.
// Tables: de_pow_libm_tables.h (gen_pow_libm_tables.py).
#pragma once

#include "de_common.h"
#include "de_pow_libm_tables.h"

namespace de {

DE_TABLE double kLibmPowLog[128 * 3] = DE_LIBM_POW_LOG_TABLE;
DE_TABLE uint64_t kLibmPowExp[128 * 2] = DE_LIBM_EXP_TABLE;

namespace libm_pow_detail {

// log: ln2 split so that k*Ln2hi + logc is exact; polynomial coefficients of
// log1p(r) - r (scaled as the evaluation uses them: A[j] multiplies r^(j+2)).
static constexpr double kLn2hi = 0x1.62e42fefa3800p-1;
static constexpr double kLn2lo = 0x1.ef35793c76730p-45;
static constexpr double kA0 = -0x1p-1;
static constexpr double kA1 = 0x1.555555555556p-2 * -2;
static constexpr double kA2 = -0x1.0000000000006p-2 * -2;
static constexpr double kA3 = 0x1.999999959554ep-3 * 4;
static constexpr double kA4 = -0x1.555555529a47ap-3 * 4;
static constexpr double kA5 = 0x1.2495b9b4845e9p-3 * -8;
static constexpr double kA6 = -0x1.0002b8b263fc3p-3 * -8;
static constexpr uint64_t kOff = 0x3fe6955500000000ULL;

// exp: N/ln2, -ln2/N in two parts, the rounding shift 1.5*2^52, and the
// coefficients of exp(s) - 1 - s for s^2 .. s^5.
static constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128;
static constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;
static constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;
static constexpr double kShift = 0x1.8p52;
static constexpr double kC2 = 0x1.ffffffffffdbdp-2;
static constexpr double kC3 = 0x1.555555555543cp-3;
static constexpr double kC4 = 0x1.55555cf172b91p-5;
static constexpr double kC5 = 0x1.1111167a4d017p-7;
static constexpr uint64_t kSignBias = 0x800ULL << 7;

DE_HD uint32_t top12(double x) { return (uint32_t)(as_u64(x) >> 52); }

// 0 if iy is not an integer, 1 if odd, 2 if even (iy: finite, non-zero).
DE_HD int checkint(uint64_t iy) {
  int e = (int)(iy >> 52 & 0x7ff);
  if (e < 0x3ff) return 0;
  if (e > 0x3ff + 52) return 2;
  if (iy & ((1ULL << (0x3ff + 52 - e)) - 1)) return 0;
  if (iy & (1ULL << (0x3ff + 52 - e))) return 1;
  return 2;
}

// Bits of +-0, +-inf or NaN.
DE_HD bool zeroinfnan(uint64_t i) {
  return 2 * i - 1 >= 2 * 0x7ff0000000000000ULL - 1;
}

DE_HD bool is_snan(double x) {
  uint64_t u = as_u64(x);
  return (u & 0x7ff0000000000000ULL) == 0x7ff0000000000000ULL &&
         (u & 0x000fffffffffffffULL) != 0 && !(u & 0x0008000000000000ULL);
}

DE_HD double oflow(uint64_t sign) { return sign ? -INFINITY : INFINITY; }
DE_HD double uflow(uint64_t sign) { return sign ? -0.0 : 0.0; }

// log(x) as hi + *tail for the bits ix of a positive normal (or rescaled
// subnormal) x.
DE_HD DE_FORCE_INLINE double log_inline(uint64_t ix, double* tail) {
  DE_NO_CONTRACT
  const uint64_t tmp = ix - kOff;
  // Integer steps on the high word (the low one of tmp & 0xfff << 52 is 0):
  // i = bits 45..51 of tmp, k = its top 12 bits signed, z = x / 2^k.
  const uint32_t thi = (uint32_t)(tmp >> 32);
  const int i = (int)((thi >> 13) & 127u);
  const int32_t k = (int32_t)thi >> 20;
  const uint32_t zhi = (uint32_t)(ix >> 32) - (thi & 0xfff00000u);
  const double z = as_f64(((uint64_t)zhi << 32) | (uint32_t)ix);
  const double kd = (double)k;
  const double invc = kLibmPowLog[3 * i + 0];
  const double logc = kLibmPowLog[3 * i + 1];
  const double logctail = kLibmPowLog[3 * i + 2];
  const double r = fma(z, invc, -1.0);
  const double t1 = fma(kd, kLn2hi, logc);
  const double t2 = r + t1;
  const double lo1 = fma(kd, kLn2lo, logctail);
  const double lo2 = r + (t1 - t2);
  const double ar = r * kA0;
  const double ar2 = r * ar;
  const double ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = fma(ar, r, -ar2);
  const double lo4 = ar2 + (t2 - hi);
  const double q1 = fma(r, kA2, kA1);
  const double q2 = fma(r, kA4, kA3);
  const double q3 = fma(r, kA6, kA5);
  const double q = fma(ar2, fma(ar2, q3, q2), q1);
  const double lo = fma(ar3, q, lo4 + (lo3 + (lo1 + lo2)));
  const double y = hi + lo;
  *tail = lo + (hi - y);
  return y;
}

// 2^(-1022) < |result| handling for exponents outside (-1022, 1024).
DE_HD double exp_specialcase(double tmp, uint64_t sbits, uint64_t ki) {
  DE_NO_CONTRACT
  if ((ki & 0x80000000ULL) == 0) {
    sbits -= 1009ULL << 52;
    const double scale = as_f64(sbits);
    return 0x1p1009 * fma(tmp, scale, scale);
  }
  sbits += 1022ULL << 52;
  const double scale = as_f64(sbits);
  const double st = tmp * scale;
  double y = scale + st;
  if (fabs(y) < 1.0) {
    const double one = y < 0.0 ? -1.0 : 1.0;
    double lo = st + (scale - y);
    const double hi = y + one;
    lo = lo + (y + (one - hi));
    y = (hi + lo) - one;
    if (y == 0) y = as_f64(sbits & 0x8000000000000000ULL);
  }
  return 0x1p-1022 * y;
}

// exp(x + xtail), negated when sign_bias is set.
DE_HD DE_FORCE_INLINE double exp_inline(double x, double xtail, uint64_t sign_bias) {
  DE_NO_CONTRACT
  uint32_t abstop = top12(x) & 0x7ff;
  if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
    if (abstop - top12(0x1p-54) >= 0x80000000u) {
      const double one = 1.0 + x;
      return sign_bias ? -one : one;
    }
    if (abstop >= top12(1024.0)) return (as_u64(x) >> 63) ? uflow(sign_bias) : oflow(sign_bias);
    abstop = 0;
  }
  const double kd0 = fma(x, kInvLn2N, kShift);
  const uint64_t ki = as_u64(kd0);
  const double kd = kd0 - kShift;
  double r = fma(kd, kNegLn2loN, fma(kd, kNegLn2hiN, x));
  r = xtail + r;
  const uint64_t idx = 2 * (ki % 128);
  const uint64_t top = (ki + sign_bias) << 45;
  const double tail = as_f64(kLibmPowExp[idx]);
  const uint64_t sbits = kLibmPowExp[idx + 1] + top;
  const double r2 = r * r;
  const double t = fma(r2, fma(r, kC3, kC2), r + tail);
  const double tmp = fma(r2 * r2, fma(r, kC5, kC4), t);
  if (abstop == 0) return exp_specialcase(tmp, sbits, ki);
  const double scale = as_f64(sbits);
  return fma(tmp, scale, scale);
}

}  // namespace libm_pow_detail

DE_HD double libm_pow(double x, double y) {
  DE_NO_CONTRACT
  using namespace libm_pow_detail;
  uint64_t sign_bias = 0;
  uint64_t ix = as_u64(x);
  const uint64_t iy = as_u64(y);
  uint32_t topx = top12(x);
  const uint32_t topy = top12(y);
  if (topx - 0x001u >= 0x7ffu - 0x001u || (topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu) {
    // x subnormal, zero, negative, inf or NaN; or |y| < 2^-65, |y| >= 2^63,
    // y zero, inf or NaN.
    if (zeroinfnan(iy)) {
      if (2 * iy == 0) return is_snan(x) ? x + y : 1.0;
      if (ix == 0x3ff0000000000000ULL) return is_snan(y) ? x + y : 1.0;
      if (2 * ix > 2 * 0x7ff0000000000000ULL || 2 * iy > 2 * 0x7ff0000000000000ULL) return x + y;
      if (2 * ix == 2 * 0x3ff0000000000000ULL) return 1.0;
      if ((2 * ix < 2 * 0x3ff0000000000000ULL) == !(iy >> 63)) return 0.0;
      return y * y;
    }
    if (zeroinfnan(ix)) {
      double x2 = x * x;
      if ((ix >> 63) && checkint(iy) == 1) x2 = -x2;
      return (iy >> 63) ? 1.0 / x2 : x2;
    }
    // x and y are non-zero and finite.
    if (ix >> 63) {
      const int yint = checkint(iy);
      if (yint == 0) return (x - x) / (x - x);
      if (yint == 1) sign_bias = kSignBias;
      ix &= 0x7fffffffffffffffULL;
      topx &= 0x7ff;
    }
    if ((topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu) {
      if (ix == 0x3ff0000000000000ULL) return 1.0;
      if ((topy & 0x7ff) < 0x3beu) return ix > 0x3ff0000000000000ULL ? 1.0 + y : 1.0 - y;
      return (ix > 0x3ff0000000000000ULL) == (topy < 0x800) ? oflow(0) : uflow(0);
    }
    if (topx == 0) {
      // Subnormal x: rescale so that the exponent field becomes negative.
      ix = as_u64(x * 0x1p52);
      ix &= 0x7fffffffffffffffULL;
      ix -= 52ULL << 52;
    }
  }
  double lo;
  const double hi = log_inline(ix, &lo);
  const double ehi = y * hi;
  const double elo = fma(y, lo, fma(y, hi, -ehi));
  return exp_inline(ehi, elo, sign_bias);
}

}  // namespace de
