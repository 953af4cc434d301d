// Shared host/device helpers for the ensemble Runge-Kutta integrator.
//
// This header is compiled three ways: by hipcc for gfx950 device code, by hipcc
// for the host side of the library, and by g++ for CPU-side consumers.  Every
// floating-point expression in the integrator must produce the same bits in all
// three, so:
//   * all translation units are built with -ffp-contract=off (no implicit FMA);
//   * explicit fma() is used only where an exactly rounded fused operation is
//     intended (double-double arithmetic), never as an optimisation;
//   * max/min use fmax/fmin (IEEE maxNum/minNum: a NaN operand yields the other
//     operand), which is the semantics of Rust's f64::max/f64::min that the
//     reference solver uses (src/methods/erk/dormandprince/ordinary.rs:157).
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DE_HD __host__ __device__ inline
#define DE_DEV __device__ inline
// Run once per trajectory (init, result write-back), inside branches a wave
// takes only when one of its lanes starts or ends a trajectory.  Forced
// inline: as calls they would pass the lane state through scratch memory,
// and the partial-wave scratch writes would multiply the kernel's HBM traffic.
#define DE_COLD __host__ __device__ inline __attribute__((always_inline))
#else
#define DE_HD inline
#define DE_DEV inline
#define DE_COLD inline
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define DE_TABLE __device__ __constant__ static const
#else
#define DE_TABLE static const
#endif

// Code whose every product and sum must be rounded on its own (error-free
// transformations, operation sequences fixed for bit parity), also in
// translation units compiled with FMA contraction.
#if defined(__clang__)
#define DE_NO_CONTRACT _Pragma("clang fp contract(off)")
#define DE_FORCE_INLINE __attribute__((always_inline))
#else
#define DE_NO_CONTRACT
#define DE_FORCE_INLINE
#endif

namespace de {

// f64::EPSILON * 10, the end-of-interval threshold of the solve drivers
// (src/ode/solve_ivp.rs:198,263; src/sde/solve_ivp.rs:217,273).
static constexpr double kTenEps = 2.220446049250313080847e-16 * 10.0;

// Quiet NaN: dense-output points that a solve does not produce.
static constexpr uint64_t kNaNBits = 0x7ff8000000000000ULL;

DE_HD uint64_t as_u64(double x) {
  union { double d; uint64_t u; } v;
  v.d = x;
  return v.u;
}

DE_HD double as_f64(uint64_t u) {
  union { double d; uint64_t u; } v;
  v.u = u;
  return v.d;
}

// Rust f64::signum: 1.0 for +0 and positive values, -1.0 for -0 and negative
// values, NaN for NaN.
DE_HD double signum(double x) { return x != x ? x : copysign(1.0, x); }

// constrain_step_size (src/utils.rs:22-33): clamp |h| to [h_min, h_max] keeping
// the sign of h (written with selects; a NaN h is returned unchanged).
DE_HD double constrain_step_size(double h, double h_min, double h_max) {
  const double sign = signum(h);
  const double a = fabs(h);
  return a < h_min ? sign * h_min : (a > h_max ? sign * h_max : h);
}

// The stiffness test's decision h * sqrt(num / den) > 6.1, with den > 0, as
// the reference evaluates it (dormandprince/ordinary.rs:172-174: a division,
// a square root and a product, each rounded), but usually without them.  For
// h, num and den in [2^-300, 2^300] the products h^2 num and RN(6.1)^2 den
// are within a relative 3 * 2^-53 of their exact values, and the rounded
// expression within a relative 2.5 * 2^-53 of h sqrt(num / den); so when the
// two products differ by more than a relative 1e-12 their order is the
// decision.  Otherwise, and outside that range (h <= 0 included), the
// reference's own sequence decides.
DE_HD bool stiff_exceeds(double h, double num, double den) {
  constexpr double kLo = 0x1p-300, kHi = 0x1p300;
  constexpr double c = 6.1;
  constexpr double c2 = c * c;  // rounded once; inside the 3 * 2^-53 above
  const bool in_range = h > kLo && h < kHi && num > kLo && num < kHi && den > kLo && den < kHi;
  if (in_range) {
    const double lhs = (h * h) * num;
    const double rhs = c2 * den;
    if (lhs < rhs * (1.0 - 1e-12)) return false;
    if (lhs > rhs * (1.0 + 1e-12)) return true;
  }
  return h * sqrt(num / den) > c;
}

// Per-trajectory status codes.  0..4 mirror Status::Complete and the Error
// variants a solve can end with (src/error.rs:13-41), 5 Status::Interrupted
// (an event solout's Terminate), 6 a solout's unwrap of a failed
// interpolation (the reference panics); kRunning is internal.
enum Status : int32_t {
  kComplete = 0,
  kMaxSteps = 1,
  kStepSize = 2,
  kStiffness = 3,
  kBadInput = 4,
  kInterrupted = 5,
  kOutOfBounds = 6,
  kRunning = -1,
};

// validate_step_size_parameters (src/utils.rs:60-157): 0 when the parameters
// are valid, else the number of the first failed check in the reference's
// order (each an Err(Error::BadInput) with its own message):
//   1 tf == t0              2 sign(h0) != sign(tf - t0)   3 h_min < 0
//   4 h_max < 0             5 h_min > h_max                6 |h0| < h_min
//   7 |h0| > h_max          8 |h0| > |tf - t0|             9 h0 == 0
DE_HD int validate_step_size_check(double h0, double h_min, double h_max, double t0,
                                   double tf) {
  if (tf == t0) return 1;
  double sign = signum(tf - t0);
  if (signum(h0) != sign) return 2;
  if (h_min < 0.0) return 3;
  if (h_max < 0.0) return 4;
  if (h_min > h_max) return 5;
  if (fabs(h0) < h_min) return 6;
  if (fabs(h0) > h_max) return 7;
  if (fabs(h0) > fabs(tf - t0)) return 8;
  if (h0 == 0.0) return 9;
  return 0;
}

DE_HD bool validate_step_size(double h0, double h_min, double h_max, double t0,
                              double tf) {
  return validate_step_size_check(h0, h_min, h_max, t0, tf) == 0;
}

}  // namespace de
