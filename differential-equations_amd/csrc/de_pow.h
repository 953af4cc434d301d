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
//
// pow(x, -0.2), the exponent of every DOPRI5 step, has a cheaper first try:
// a table-driven fifth root whose result is checked by Ziv's rounding test
// (in the form of the CRlibm documentation) and returned only when it is
// certainly the correctly rounded value; otherwise, and for non-normal x, the
// double-double evaluation above runs.  Both parts use only exactly rounded
// operations, so host and device still agree bit for bit.
#pragma once

#include "de_common.h"
#include "de_pow_tables.h"

// The error-free transformations below need every product and sum rounded
// on its own, also in translation units compiled with FMA contraction.
#if defined(__clang__)
#define DE_NO_CONTRACT _Pragma("clang fp contract(off)")
#define DE_FORCE_INLINE __attribute__((always_inline))
#else
#define DE_NO_CONTRACT
#define DE_FORCE_INLINE
#endif

namespace de {

DE_TABLE double kPowLogTable[128 * 3] = DE_POW_LOG_TABLE;
DE_TABLE double kPowExpTable[128 * 2] = DE_POW_EXP_TABLE;

struct dd {
  double hi, lo;
};

// Error-free transformations.
DE_HD dd two_sum(double a, double b) {
  DE_NO_CONTRACT
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

// Requires exponent(a) >= exponent(b) or a == 0.
DE_HD dd fast_two_sum(double a, double b) {
  DE_NO_CONTRACT
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

DE_HD dd two_prod(double a, double b) {
  DE_NO_CONTRACT
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}

// Round-to-nearest-even integer of |v| < 2^51 using plain IEEE adds.
DE_HD double round_to_int(double v) {
  DE_NO_CONTRACT
  const double shift = 0x1.8p52;
  return (v + shift) - shift;
}

// log(x) as a double-double for finite x > 0.  Absolute error < 2^-74 for
// |log x| < 745.
DE_HD DE_FORCE_INLINE dd pow_log_dd(double x) {
  DE_NO_CONTRACT
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
DE_HD DE_FORCE_INLINE double pow_exp_dd(dd Z) {
  DE_NO_CONTRACT
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
DE_HD DE_FORCE_INLINE double pow_dd(double x, double y) {
  DE_NO_CONTRACT
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

DE_TABLE double kPow5InvTable[128 * 2] = DE_POW5_INV_TABLE;
DE_TABLE double kPow5ATable[640 * 2] = DE_POW5_A_TABLE;
DE_TABLE double kPowR6InvTable[128 * 2] = DE_POWR6_INV_TABLE;
DE_TABLE double kPowR6ATable[6 * 128 * 2] = DE_POWR6_A_TABLE;
DE_TABLE double kPowR7InvTable[128 * 2] = DE_POWR7_INV_TABLE;
DE_TABLE double kPowR7ATable[7 * 128 * 2] = DE_POWR7_A_TABLE;
DE_TABLE double kPowR8InvTable[128 * 2] = DE_POWR8_INV_TABLE;
DE_TABLE double kPowR8ATable[8 * 128 * 2] = DE_POWR8_A_TABLE;
DE_TABLE double kPowR9InvTable[128 * 2] = DE_POWR9_INV_TABLE;
DE_TABLE double kPowR9ATable[9 * 128 * 2] = DE_POWR9_A_TABLE;

// The constants of the k-th-root path, as compile-time values (a static
// member array would be a device global whose loads are not folded).
#define DE_POW_ROOT(K, P, INV, A)                                                \
  template <>                                                                    \
  struct PowRoot<K> {                                                            \
    static constexpr double c1_hi = P##_C1_HI, c1_lo = P##_C1_LO;                \
    static constexpr double neg_delta = P##_NEG_DELTA;                           \
    static constexpr double neg_delta_ln2 = P##_NEG_DELTA_LN2;                   \
    DE_HD static constexpr double poly(int j) {                                  \
      constexpr double c[8] = P##_POLY;                                          \
      return c[j];                                                               \
    }                                                                            \
    DE_HD static const double* inv() { return INV; }                             \
    DE_HD static const double* a() { return A; }                                 \
  };
template <int K>
struct PowRoot;
DE_POW_ROOT(5, DE_POW5, kPow5InvTable, kPow5ATable)
DE_POW_ROOT(6, DE_POWR6, kPowR6InvTable, kPowR6ATable)
DE_POW_ROOT(7, DE_POWR7, kPowR7InvTable, kPowR7ATable)
DE_POW_ROOT(8, DE_POWR8, kPowR8InvTable, kPowR8ATable)
DE_POW_ROOT(9, DE_POWR9, kPowR9InvTable, kPowR9ATable)
#undef DE_POW_ROOT

// x^D with D = -fl(1/K), K = 5 .. 9, for positive normal x.  Returns false
// when the result could not be certified (then *out is meaningless); for
// other x the return is false too.
//
// Let D = -(1/K + delta), delta = fl(1/K) - 1/K (|delta| <= 2^-55; 0 for
// K = 8).  With x = 2^E m, m in [1,2), E = Kq + s (0 <= s < K),
// r = m*invc_i - 1 (exact, |r| < 2^-7.48):
//   x^D = 2^-q * [2^(-s/K) invc_i^(1/K)] * (1+r)^(-1/K) * exp(-delta log x)
// The bracket is the table entry A = ah + al.  (1+r)^(-1/K) is its Taylor
// series to r^9 (truncation < 2^-78), kept as 1 + p1 + ulo with p1 + e1 =
// -r/K exact to 2^-118.  corr = -delta log x = -delta (E ln2 - log invc_i + r)
// is at most 2^-46.5 and is needed to 2^-70 only (the dropped -delta r^2/2 is
// below 2^-71).  Every rounding after the exact steps is at most 2^-70
// relative to the result and there are fewer than six of them, so s + w
// below is within eps < 2^-67.4 (relative) of x^D.  Rounding test:
// res = RN(s + w) with residual d = s + w - res (exact, Fast2Sum); if
// RN(res + d*(1 + 2^-12)) == res then |x^D - res| < ulp(res)/2 (ulp/4 below a
// power of two), because 2^54 eps < 1 - 1/(1 + 2^-12).  The test fails for
// about 2^-12 of all x.
template <int K>
DE_HD bool pow_root_fast(double x, double* out) {
  DE_NO_CONTRACT
  using T = PowRoot<K>;
  // E + K*Q0 = u with u >= 0 for every normal x; the result's scale 2^-q is
  // 2^(Q0 - floor(u/K)).
  constexpr uint32_t kQ0 = (1025u + K - 1u) / K;
  constexpr uint32_t kOff = K * kQ0 - 1023u;
  const uint64_t ix = as_u64(x);
  const uint32_t hx = (uint32_t)(ix >> 32);
  const uint32_t i = (hx >> 13) & 127u;
  const uint32_t u = ((hx >> 20) & 0x7ffu) + kOff;
  const uint32_t fq = u / (uint32_t)K;
  const uint32_t s = u - (uint32_t)K * fq;
  const uint32_t row = s * 128u + i;  // < 128 K for every x
  const double m = as_f64((ix & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  const double invc = T::inv()[2 * i];
  const double dlinv = T::inv()[2 * i + 1];  // delta log(invc_i)
  const double ah = T::a()[2 * row];
  const double al = T::a()[2 * row + 1];

  const double r = fma(m, invc, -1.0);
  const double e = (double)((int32_t)u - (int32_t)(K * kQ0));
  const double corr = fma(T::neg_delta, r, fma(e, T::neg_delta_ln2, dlinv));

  double t = T::poly(0);  // binom(-1/K, j), j = 9 .. 2
  t = fma(r, t, T::poly(1));
  t = fma(r, t, T::poly(2));
  t = fma(r, t, T::poly(3));
  t = fma(r, t, T::poly(4));
  t = fma(r, t, T::poly(5));
  t = fma(r, t, T::poly(6));
  t = fma(r, t, T::poly(7));
  const double p1 = r * T::c1_hi;
  const double e1 = fma(r, T::c1_hi, -p1);
  const double v = fma(r, t, T::c1_lo);
  const double ulo = fma(r, v, e1);                    // (1+r)^(-1/K) = 1 + p1 + ulo
  const double lt = ulo + fma(corr, p1 + ulo, corr);  // (..)(1+corr) = 1 + p1 + lt
  const double alt = fma(al, p1 + lt, al);
  const double m1 = ah * p1;
  const double m1e = fma(ah, p1, -m1);
  const double lo = fma(ah, lt, m1e + alt);
  const double sh = ah + m1;  // |ah| > |m1|
  const double st = m1 - (sh - ah);
  const double w = st + lo;
  const double res = sh + w;
  const double d = w - (res - sh);
  const bool normal = (hx - 0x00100000u) < 0x7fe00000u;  // 0 < x < inf, not subnormal
  *out = res * as_f64((uint64_t)(1023u + kQ0 - fq) << 52);
  return normal && fma(d, 0x1.001p+0, res) == res;
}

// x^(-0.2) (the fifth-root path).
DE_HD bool pow_m02_fast(double x, double* out) { return pow_root_fast<5>(x, out); }

// The controllers' exponents -fl(1/k), k = 5 .. 9, try the k-th-root path
// first.  The double-double path is forced inline (above): as a call inside
// the step loop it would save and restore live registers through scratch
// memory.
DE_HD double pow(double x, double y) {
  double r;
  if (y == -0.2) {
    if (__builtin_expect(pow_root_fast<5>(x, &r), 1)) return r;
  } else if (y == -(1.0 / 6.0)) {
    if (__builtin_expect(pow_root_fast<6>(x, &r), 1)) return r;
  } else if (y == -(1.0 / 7.0)) {
    if (__builtin_expect(pow_root_fast<7>(x, &r), 1)) return r;
  } else if (y == -0.125) {
    if (__builtin_expect(pow_root_fast<8>(x, &r), 1)) return r;
  } else if (y == -(1.0 / 9.0)) {
    if (__builtin_expect(pow_root_fast<9>(x, &r), 1)) return r;
  }
  return pow_dd(x, y);
}

}  // namespace de
