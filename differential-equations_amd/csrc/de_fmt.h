// Rust's `{:?}` text of an f64, the number format of Solution::to_csv
// (src/solution.rs:190-230 writes `format!("{:?}", t)` and `{:?}` of every
// state component), on the host and the device.
//
// The digits are the shortest decimal that reads back as the same double,
// and of those the closest to it (the round-trip interval includes its ends
// when the mantissa is even, as round-half-even parsing does).  They come
// from the Ryu algorithm (U. Adams, "Ryu: fast float-to-string conversion",
// PLDI 2018): the value's rounding interval is scaled by a 128-bit power of
// five (tables in de_ryu_tables.h) so that its ends become integers
// vm < vr < vp, and decimal digits are removed while vp / 10 > vm / 10;
// exact trailing-zero tracking decides the ties.
//
// Layout (core's float_to_general_debug): NaN -> "NaN" (any sign), +-inf ->
// "inf" / "-inf", +-0 -> "0.0" / "-0.0"; 1e-4 <= |x| < 1e16 in positional
// notation with at least one fractional digit ("1.0", "0.0001",
// "1234500.0"), otherwise d[.ddd]e<exp> with no '+' and no leading zeros in
// the exponent ("1e-7", "1.5e16", "-2.2250738585072014e-308").
#pragma once

#include "de_common.h"
#include "de_ryu_tables.h"

namespace de {
namespace ryu {

DE_TABLE uint64_t kPow5Split[kPow5SplitSize][2] = DE_RYU_POW5_SPLIT;
DE_TABLE uint64_t kPow5InvSplit[kPow5InvSplitSize][2] = DE_RYU_POW5_INV_SPLIT;

constexpr int kMaxChars = 24;  // "-2.2250738585072014e-308"

// ceil(log2(5^e)) for 1 <= e <= 3528; 1 for e = 0.
DE_HD int32_t pow5bits(int32_t e) { return (int32_t)(((uint32_t)e * 1217359u) >> 19) + 1; }
// floor(log10(2^e)), floor(log10(5^e)) for 0 <= e <= 1650 / 2620.
DE_HD uint32_t log10_pow2(int32_t e) { return ((uint32_t)e * 78913u) >> 18; }
DE_HD uint32_t log10_pow5(int32_t e) { return ((uint32_t)e * 732923u) >> 20; }

// High 64 bits of a 64 x 64-bit product from four 32 x 32-bit products: the
// same integer operations on the host and the device, so the host build of
// this file exercises exactly the device's arithmetic.
DE_HD uint64_t umulh(uint64_t a, uint64_t b) {
  const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
  const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  const uint64_t mid = (p00 >> 32) + (uint32_t)p01 + (uint32_t)p10;
  return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// Every loop below has a fixed iteration cap: a 64-bit value has at most 27
// factors 5 and 20 decimal digits, so a cap is never reached by a valid
// input; reaching one marks the result bad (the writer reports it) instead
// of looping.
constexpr int kLoopCap = 32;

// Number of factors 5 of v > 0: v * 5^-1 mod 2^64 stays <= (2^64 - 1) / 5
// exactly while 5 divides v.  kLoopCap (and *bad set) for v = 0.
DE_HD uint32_t pow5_factor(uint64_t v, bool* bad) {
  uint32_t count = 0;
  for (int it = 0; it < kLoopCap; ++it) {
    v *= 14757395258967641293ull;
    if (v > 3689348814741910323ull) return count;
    ++count;
  }
  *bad = true;
  return count;
}
DE_HD bool multiple_of_pow5(uint64_t v, uint32_t p, bool* bad) { return pow5_factor(v, bad) >= p; }
DE_HD bool multiple_of_pow2(uint64_t v, uint32_t p) { return (v & ((1ull << p) - 1)) == 0; }

// floor(m * mul / 2^j) for a 128-bit mul, 64 < j < 128, m < 2^55.
DE_HD uint64_t mul_shift(uint64_t m, const uint64_t* mul, int32_t j) {
  const uint64_t lo1 = m * mul[1];
  uint64_t hi1 = umulh(m, mul[1]);
  const uint64_t hi0 = umulh(m, mul[0]);
  const uint64_t sum = hi0 + lo1;
  if (sum < hi0) ++hi1;
  const int32_t d = j - 64;  // 2 .. 59
  return (hi1 << (64 - d)) | (sum >> d);
}

struct Decimal {
  uint64_t digits;  // value = digits * 10^exp
  int32_t exp;
  bool bad;         // a loop cap was reached (never for a valid input)
};

// Shortest closest decimal of a finite nonzero double's magnitude.
DE_HD Decimal shortest(uint64_t ieee_mantissa, uint32_t ieee_exponent) {
  bool bad = false;
  int32_t e2;
  uint64_t m2;
  if (ieee_exponent == 0) {
    e2 = 1 - 1023 - 52 - 2;
    m2 = ieee_mantissa;
  } else {
    e2 = (int32_t)ieee_exponent - 1023 - 52 - 2;
    m2 = (1ull << 52) | ieee_mantissa;
  }
  const bool accept_bounds = (m2 & 1) == 0;
  // The interval (4 m2 - 1 - mm_shift, 4 m2 + 2) around mv = 4 m2, scaled by 2^e2.
  const uint64_t mv = 4 * m2;
  const uint32_t mm_shift = ieee_mantissa != 0 || ieee_exponent <= 1;
  uint64_t vr, vp, vm;
  int32_t e10;
  bool vm_tz = false, vr_tz = false;
  if (e2 >= 0) {
    const uint32_t q = log10_pow2(e2) - (e2 > 3);
    e10 = (int32_t)q;
    const int32_t k = kPow5InvBitcount + pow5bits((int32_t)q) - 1;
    const int32_t i = -e2 + (int32_t)q + k;
    const uint64_t* mul = kPow5InvSplit[q];
    vr = mul_shift(mv, mul, i);
    vp = mul_shift(mv + 2, mul, i);
    vm = mul_shift(mv - 1 - mm_shift, mul, i);
    if (q <= 21) {
      // Only one of mp, mv, mm can be a multiple of 5, if any.
      if (mv % 5 == 0)
        vr_tz = multiple_of_pow5(mv, q, &bad);
      else if (accept_bounds)
        vm_tz = multiple_of_pow5(mv - 1 - mm_shift, q, &bad);
      else
        vp -= multiple_of_pow5(mv + 2, q, &bad);
    }
  } else {
    const uint32_t q = log10_pow5(-e2) - (-e2 > 1);
    e10 = (int32_t)q + e2;
    const int32_t i = -e2 - (int32_t)q;
    const int32_t k = pow5bits(i) - kPow5Bitcount;
    const int32_t j = (int32_t)q - k;
    const uint64_t* mul = kPow5Split[i];
    vr = mul_shift(mv, mul, j);
    vp = mul_shift(mv + 2, mul, j);
    vm = mul_shift(mv - 1 - mm_shift, mul, j);
    if (q <= 1) {
      // mv = 4 m2 has at least two trailing zero bits.
      vr_tz = true;
      if (accept_bounds)
        vm_tz = mm_shift == 1;  // mm = mv - 1 - mm_shift
      else
        --vp;  // mp = mv + 2 has one
    } else if (q < 63) {
      vr_tz = multiple_of_pow2(mv, q);
    }
  }
  // Remove digits while the interval still holds a shorter decimal.
  int32_t removed = 0;
  uint32_t last = 0;
  uint64_t out;
  if (vm_tz || vr_tz) {
    for (int it = 0;; ++it) {
      const uint64_t vp10 = vp / 10, vm10 = vm / 10;
      if (vp10 <= vm10) break;
      if (it == kLoopCap) {
        bad = true;
        break;
      }
      const uint64_t vr10 = vr / 10;
      vm_tz &= (uint32_t)(vm - 10 * vm10) == 0;
      vr_tz &= last == 0;
      last = (uint32_t)(vr - 10 * vr10);
      vr = vr10;
      vp = vp10;
      vm = vm10;
      ++removed;
    }
    if (vm_tz) {
      for (int it = 0;; ++it) {
        const uint64_t vm10 = vm / 10;
        if ((uint32_t)(vm - 10 * vm10) != 0) break;
        if (it == kLoopCap) {  // vm = 0
          bad = true;
          break;
        }
        const uint64_t vp10 = vp / 10, vr10 = vr / 10;
        vr_tz &= last == 0;
        last = (uint32_t)(vr - 10 * vr10);
        vr = vr10;
        vp = vp10;
        vm = vm10;
        ++removed;
      }
    }
    // An exact tie rounds to even.
    if (vr_tz && last == 5 && (vr & 1) == 0) last = 4;
    out = vr + (((vr == vm) && (!accept_bounds || !vm_tz)) || last >= 5);
  } else {
    bool round_up = false;
    for (int it = 0;; ++it) {
      const uint64_t vp10 = vp / 10, vm10 = vm / 10;
      if (vp10 <= vm10) break;
      if (it == kLoopCap) {
        bad = true;
        break;
      }
      const uint64_t vr10 = vr / 10;
      round_up = (uint32_t)(vr - 10 * vr10) >= 5;
      vr = vr10;
      vp = vp10;
      vm = vm10;
      ++removed;
    }
    out = vr + (vr == vm || round_up);
  }
  return {out, e10 + removed, bad};
}

}  // namespace ryu

// Rust's `{:?}` text of one double, prepared: its length and what emit()
// needs to write it without a staging buffer.
struct FmtF64 {
  uint64_t digits;  // the shortest digits (finite nonzero values)
  int32_t point;    // value = 0.<digits> * 10^point
  int16_t olen;     // number of digits
  int8_t kind;      // 0 finite nonzero, 1 zero, 2 infinite, 3 NaN
  int8_t neg;
  int8_t bad;       // a loop cap of the digit search was reached
  int32_t len;      // bytes of the text
};

DE_HD int decimal_length(uint64_t v) {
  int n = 1;
  uint64_t p = 10;
  for (int i = 1; i < 20 && v >= p; ++i, p *= 10) ++n;
  return n;
}

DE_HD FmtF64 fmt_prepare(double x) {
  const uint64_t bits = as_u64(x);
  const uint64_t mant = bits & ((1ull << 52) - 1);
  const uint32_t ex = (uint32_t)(bits >> 52) & 0x7ffu;
  FmtF64 f;
  f.digits = 0;
  f.point = 0;
  f.olen = 0;
  f.bad = 0;
  f.neg = (int8_t)(bits >> 63);
  if (ex == 0x7ffu && mant != 0) {
    f.kind = 3;
    f.neg = 0;
    f.len = 3;  // NaN
    return f;
  }
  if (ex == 0x7ffu) {
    f.kind = 2;
    f.len = f.neg + 3;  // inf
    return f;
  }
  if (ex == 0 && mant == 0) {
    f.kind = 1;
    f.len = f.neg + 3;  // 0.0
    return f;
  }
  const ryu::Decimal d = ryu::shortest(mant, ex);
  f.kind = 0;
  f.digits = d.digits;
  f.olen = (int16_t)decimal_length(d.digits);
  f.point = d.exp + f.olen;
  if (d.bad || f.olen > 17 || f.point < -330 || f.point > 330) {
    // Never for a valid input (17 significant digits at most, exponents of a
    // double): flagged, and written as the 3 bytes "NaN" so that every text
    // keeps within its 24-byte bound.
    f.bad = 1;
    f.kind = 3;
    f.neg = 0;
    f.len = 3;
    return f;
  }
  const int olen = f.olen, point = f.point;
  if (point < -3 || point > 16) {
    const int e = point - 1 < 0 ? 1 - point : point - 1;
    f.len = f.neg + 1 + (olen > 1 ? olen : 0) + 1 + (point - 1 < 0) + (e >= 100 ? 3 : e >= 10 ? 2 : 1);
  } else if (point <= 0) {
    f.len = f.neg + 2 - point + olen;
  } else if (point < olen) {
    f.len = f.neg + olen + 1;
  } else {
    f.len = f.neg + point + 2;
  }
  return f;
}

// Writes the f.len bytes of the text to dst[0 ..).
DE_HD void fmt_emit(const FmtF64& f, char* dst) {
  int s = 0;
  if (f.kind == 3) {
    dst[0] = 'N';
    dst[1] = 'a';
    dst[2] = 'N';
    return;
  }
  if (f.neg) dst[s++] = '-';
  if (f.kind == 2) {
    dst[s] = 'i';
    dst[s + 1] = 'n';
    dst[s + 2] = 'f';
    return;
  }
  if (f.kind == 1) {
    dst[s] = '0';
    dst[s + 1] = '.';
    dst[s + 2] = '0';
    return;
  }
  const int olen = f.olen, point = f.point;
  // Position of digit k (0 = most significant) and the fixed characters.
  int first, gap;  // digit k goes to s + first + k (+ 1 from index `gap` on)
  if (point < -3 || point > 16) {
    first = 0;
    gap = 1;
    if (olen > 1) dst[s + 1] = '.';
    int n = s + (olen > 1 ? olen + 1 : 1);
    dst[n++] = 'e';
    int e = point - 1;
    if (e < 0) {
      dst[n++] = '-';
      e = -e;
    }
    if (e >= 100) dst[n++] = (char)('0' + e / 100);
    if (e >= 10) dst[n++] = (char)('0' + (e / 10) % 10);
    dst[n] = (char)('0' + e % 10);
  } else if (point <= 0) {
    dst[s] = '0';
    dst[s + 1] = '.';
    for (int i = 0; i < -point; ++i) dst[s + 2 + i] = '0';
    first = 2 - point;
    gap = 32;
  } else if (point < olen) {
    dst[s + point] = '.';
    first = 0;
    gap = point;
  } else {
    for (int i = olen; i < point; ++i) dst[s + i] = '0';
    dst[s + point] = '.';
    dst[s + point + 1] = '0';
    first = 0;
    gap = 32;
  }
  uint64_t v = f.digits;
  for (int k = olen - 1; k >= 0; --k) {
    const uint64_t v10 = v / 10;
    dst[s + first + k + (k >= gap ? 1 : 0)] = (char)('0' + (int)(v - 10 * v10));
    v = v10;
  }
}

// Writes Rust's `{:?}` text of x to buf (at most ryu::kMaxChars bytes, no
// terminator); returns the number of bytes.
DE_HD int fmt_f64_debug(double x, char* buf) {
  const FmtF64 f = fmt_prepare(x);
  fmt_emit(f, buf);
  return f.len;
}

}  // namespace de
