// Project pow(x, y) for the step-size controller, identical bits on host and
// device.
//
// The reference calls f64::powf in two places on the hot path:
//   err.powf(-1/order)            src/methods/erk/dormandprince/ordinary.rs:154
//   (0.01/der12).powf(1/order)    src/methods/h_init.rs:124
// Rust lowers powf to the platform libm `pow`, whose last-bit rounding is not
// specified and differs between libms.  The integrator instead evaluates pow in
// double-double arithmetic (relative error below about 2^-75 before the final
// rounding) and rounds once, so the result is the correctly rounded x^y except
// when x^y lies within about 2^-75 (relative) of a rounding boundary.  Because
// the same source runs on the CPU and on gfx950 with explicit, exactly rounded
// fma and no contraction, CPU and GPU produce the same bits.
//
// Method (textbook; Tang's table-driven log and exp):
//   x = 2^k z, z in [0.70703125, 1.4140625); i = top 7 bits of bits(x) - OFF
//   log x = k ln2 - log(invc_i) + log1p(r),   r = z*invc_i - 1 (exact, as dd)
//   Z = y log x (dd);  Z = m ln2/128 + s;  x^y = 2^(m>>7) 2^((m&127)/128) e^s
// Domain: finite y with |y| <= 0.5 (the controller uses +-1/order <= 0.2), any x.
// For those arguments the result is always a normal double.
#pragma once

#include "de_common.h"
#include "de_pow_tables.h"

namespace de {

DE_TABLE double kPowLogTable[128 * 3] = DE_POW_LOG_TABLE;
DE_TABLE double kPowExpTable[128 * 2] = DE_POW_EXP_TABLE;

struct dd {
  double hi, lo;
};

// Error-free transformations.
DE_HD dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

// Requires exponent(a) >= exponent(b) or a == 0.
DE_HD dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

DE_HD dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}

// Round-to-nearest-even integer of |v| < 2^51 using plain IEEE adds.
DE_HD double round_to_int(double v) {
  const double shift = 0x1.8p52;
  return (v + shift) - shift;
}

// log(x) as a double-double for finite x > 0.  Absolute error < 2^-74 for
// |log x| < 745.
DE_HD dd pow_log_dd(double x) {
  uint64_t ix = as_u64(x);
  int64_t kadj = 0;
  if ((ix >> 52) == 0) {  // subnormal: scale into the normal range (exact)
    ix = as_u64(x * 0x1p54);
    kadj = -54;
  }
  uint64_t tmp = ix - DE_POW_OFF;
  int i = (int)((tmp >> 45) & 127);
  int64_t k = ((int64_t)tmp >> 52) + kadj;
  uint64_t iz = ix - (tmp & (0xfffULL << 52));
  double z = as_f64(iz);
  double invc = kPowLogTable[3 * i + 0];
  double lc_hi = kPowLogTable[3 * i + 1];
  double lc_lo = kPowLogTable[3 * i + 2];

  // r = z*invc - 1, exactly, as (rh, rl).
  dd p = two_prod(z, invc);
  dd r = fast_two_sum(p.hi - 1.0, p.lo);
  double rh = r.hi, rl = r.lo;

  // log1p(r) = r - r^2/2 + r^3 Q(r), Q(r) = 1/3 - r/4 + ... - r^7/10.
  const double c[8] = DE_POW_LOG_POLY;
  double q = c[7];
  q = fma(rh, q, c[6]);
  q = fma(rh, q, c[5]);
  q = fma(rh, q, c[4]);
  q = fma(rh, q, c[3]);
  q = fma(rh, q, c[2]);
  q = fma(rh, q, c[1]);
  q = fma(rh, q, c[0]);
  dd s = two_prod(rh, rh);
  double r3 = s.hi * rh;
  dd a = fast_two_sum(rh, -0.5 * s.hi);
  double tail = a.lo + rl - 0.5 * s.lo - rh * rl + r3 * q;

  // k ln2 + log(c) + log1p(r)
  double kd = (double)k;
  dd b = two_sum(kd * DE_POW_LN2_HI, lc_hi);
  dd cc = two_sum(b.hi, a.hi);
  double lo = b.lo + cc.lo + lc_lo + kd * DE_POW_LN2_LO + tail;
  return fast_two_sum(cc.hi, lo);
}

// exp(Z) for a double-double Z with |Z| < 700, rounded once to double.
DE_HD double pow_exp_dd(dd Z) {
  double kz = round_to_int(Z.hi * DE_POW_EXP_INV_C);
  int64_t ki = (int64_t)kz;
  double sh = Z.hi - kz * DE_POW_EXP_C_HI;
  double sl = Z.lo - kz * DE_POW_EXP_C_LO;
  dd r = two_sum(sh, sl);

  // expm1(r) = r + r^2/2 + r^3 Q2(r), Q2(r) = 1/6 + r/24 + ... + r^5/8!.
  const double c[6] = DE_POW_EXP_POLY;
  double q = c[5];
  q = fma(r.hi, q, c[4]);
  q = fma(r.hi, q, c[3]);
  q = fma(r.hi, q, c[2]);
  q = fma(r.hi, q, c[1]);
  q = fma(r.hi, q, c[0]);
  dd s = two_prod(r.hi, r.hi);
  double r3 = s.hi * r.hi;
  dd e = fast_two_sum(r.hi, 0.5 * s.hi);
  double etail = e.lo + r.lo + 0.5 * s.lo + r.hi * r.lo + r3 * q;

  int j = (int)(ki & 127);
  int64_t m = ki >> 7;
  double th = kPowExpTable[2 * j + 0];
  double tl = kPowExpTable[2 * j + 1];
  // 2^(j/128) (1 + e) = th + th*e + tl + tl*e
  dd pe = two_prod(th, e.hi);
  double rlo = pe.lo + th * etail + tl + tl * e.hi;
  dd sum = fast_two_sum(th, pe.hi);
  double res = sum.hi + (sum.lo + rlo);
  // Scale by 2^m (exact for normal results, which the domain guarantees).
  if (m > 1000) {
    res *= 0x1p1000;
    m -= 1000;
  } else if (m < -1000) {
    res *= 0x1p-1000;
    m += 1000;
  }
  return res * as_f64((uint64_t)(m + 1023) << 52);
}

// pow(x, y) with C99 special cases for the arguments the integrator can see.
// The general path runs on a substitute argument for the special cases and
// their results are selected at the end (no divergent early returns).
DE_HD double pow(double x, double y) {
  const double inf = as_f64(0x7ff0000000000000ULL);
  const bool special = !(x > 0.0 && x < inf) || x == 1.0 || y == 0.0 || y != y;
  dd L = pow_log_dd(special ? 2.0 : x);
  dd zp = two_prod(y, L.hi);
  dd Z = fast_two_sum(zp.hi, zp.lo + y * L.lo);
  double r = pow_exp_dd(special ? dd{0.0, 0.0} : Z);
  if (special) {
    if (y == 0.0 || x == 1.0) r = 1.0;
    else if (x != x || y != y) r = x + y;
    else if (x == 0.0) r = y < 0.0 ? inf : 0.0;
    else if (x < 0.0) r = as_f64(0x7ff8000000000000ULL);  // non-integer y only
    else r = y < 0.0 ? 0.0 : inf;                          // x = +inf
  }
  return r;
}

}  // namespace de
