// One-lane-per-trajectory solvers for fixed-dimension families (device code).
//
// Each lane runs the reference's solve loop for one trajectory with its whole
// state in registers:
//   solve_ode driver         src/ode/solve_ivp.rs:116-277
//   DOPRI5 init / step       src/methods/erk/dormandprince/ordinary.rs:16-270
//   DOPRI5 dense output      src/methods/erk/dormandprince/ordinary.rs:196-235,298-336
//   h_init                   src/methods/h_init.rs:45-135
//   fixed ERK (rk4, euler)   src/methods/erk/fixed/ordinary.rs:16-138, Hermite
//                            output src/interpolate.rs:40-60
//   TEvalSolout              src/solout/t_eval.rs:83-134
//   event solouts            de_event.h (src/solout/event.rs, crossing.rs,
//                            hyperplane.rs)
// Operation order is the reference's, with one difference: terms whose tableau
// coefficient is exactly zero (a[6][1], b[1], b[6], er[1], all of bi[4]) and
// the leading `0 +` of combinations that start from zeros_like() are omitted.
// For finite stages those terms add a signed zero, which can only change the
// sign of an exactly-zero component; a non-finite stage makes the attempt's
// error non-finite either way, so the attempt is rejected either way.  The
// values produced are therefore the reference's.
//
// The caller provides the RHS family F (de_families.h) and the launch
// arguments; the same code runs under the plain and the persistent schedules.
#pragma once

#include "de_common.h"
#include "de_dop853_table.h"
#include "de_event.h"
#include "de_families.h"
#include "de_pow_libm.h"

// Controller and initial-step pow: the C library's pow that f64::powf calls,
// restated with the same roundings (de_pow_libm.h), in both builds.
#define DE_CTRL_POW(x, y) ::de::libm_pow((x), (y))

namespace de {

struct OdeArgs {
  int64_t n;
  const double* params;
  const double* y0;
  const double* t0;
  const double* tf;
  double t0s, tfs;
  const double* rtol_traj;
  const double* atol_traj;
  double rtol, atol;
  double h0, h_min, h_max, safety, min_scale, max_scale;
  int64_t max_steps;
  int64_t max_rejects;  // embedded family: consecutive rejections that end a solve
  const double* t_eval;
  int32_t n_eval;
  double* y_final;
  double* t_final;
  int64_t* accepted;
  int64_t* rejected;
  int64_t* evals;
  int32_t* status;
  double* dense;
  int32_t* n_dense;
  unsigned long long* counter;  // persistent schedule: next trajectory index
  const uint32_t* order;        // persistent schedule: trajectory at each position, or null
  const double* rtol_comp;      // per-component tolerances [c*n + i] (or state layout), or null
  const double* atol_comp;
  EventArgs ev;                 // event solout (lanes built with EV != 0)
  const uint32_t* n_dev;        // persistent schedule: positions to take, read on the
                                // device (null: n)
  void* scratch;                // caller-provided launch workspace (no allocation), or null
  int64_t scratch_bytes;
  // Variable-count point output (kPoints*): slots pts_offset[i] .. pts_offset[i+1]
  // of pts_t / pts_y (state at [s * D + c]); n_pts[i] = points pushed.
  int32_t pts_mode;
  int32_t pts_n;     // DenseSolout n
  double pts_dt;     // EvenSolout dt
  const int64_t* pts_offset;
  double* pts_t;
  double* pts_y;
  int64_t* n_pts;
};

// Tolerances of one trajectory (Tolerance, src/tolerance.rs:11-68): a scalar
// pair, or one pair per component when V (per-component arrays were given).
template <int D, bool V>
struct LaneTol {
  double r, a;
  DE_HD void load(const OdeArgs& A, int64_t i) {
    r = A.rtol_traj ? A.rtol_traj[i] : A.rtol;
    a = A.atol_traj ? A.atol_traj[i] : A.atol;
  }
  DE_HD double rt(int) const { return r; }
  DE_HD double at(int) const { return a; }
};
template <int D>
struct LaneTol<D, true> {
  double r[D], a[D];
  DE_HD void load(const OdeArgs& A, int64_t i) {
    const double rs = A.rtol_traj ? A.rtol_traj[i] : A.rtol;
    const double as = A.atol_traj ? A.atol_traj[i] : A.atol;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      r[c] = A.rtol_comp ? A.rtol_comp[c * A.n + i] : rs;
      a[c] = A.atol_comp ? A.atol_comp[c * A.n + i] : as;
    }
  }
  DE_HD double rt(int c) const { return r[c]; }
  DE_HD double at(int c) const { return a[c]; }
};

// Bucket of a predicted cost c (attempts) for the longest-first order of the
// persistent schedule: 16 buckets per octave over [1, 2^16), 0 = longest; NaN,
// c < 1 and ended trajectories go to the last bucket (255).
DE_HD uint32_t cost_bucket(double c) {
  if (!(c >= 1.0)) return 255u;
  if (c >= 65536.0) return 0u;
  const uint64_t u = as_u64(c);
  const uint32_t e = (uint32_t)(u >> 52) - 1023u;  // 0 .. 15
  const uint32_t m = (uint32_t)(u >> 48) & 15u;    // top 4 fraction bits
  return 255u - (e * 16u + m);
}

// DOPRI5 coefficients (src/tableau/dorman_prince.rs:37-117).
namespace dp {
constexpr double c1 = 0.2, c2 = 0.3, c3 = 0.8, c4 = 8.0 / 9.0;
constexpr double a10 = 0.2;
constexpr double a20 = 3.0 / 40.0, a21 = 9.0 / 40.0;
constexpr double a30 = 44.0 / 45.0, a31 = -56.0 / 15.0, a32 = 32.0 / 9.0;
constexpr double a40 = 19372.0 / 6561.0, a41 = -25360.0 / 2187.0, a42 = 64448.0 / 6561.0,
                 a43 = -212.0 / 729.0;
constexpr double a50 = 9017.0 / 3168.0, a51 = -355.0 / 33.0, a52 = 46732.0 / 5247.0,
                 a53 = 49.0 / 176.0, a54 = -5103.0 / 18656.0;
constexpr double a60 = 35.0 / 384.0, a62 = 500.0 / 1113.0, a63 = 125.0 / 192.0,
                 a64 = -2187.0 / 6784.0, a65 = 11.0 / 84.0;
constexpr double b0 = 35.0 / 384.0, b2 = 500.0 / 1113.0, b3 = 125.0 / 192.0,
                 b4 = -2187.0 / 6784.0, b5 = 11.0 / 84.0;
constexpr double e0 = 71.0 / 57600.0, e2 = -71.0 / 16695.0, e3 = 71.0 / 1920.0,
                 e4 = -17253.0 / 339200.0, e5 = 22.0 / 525.0, e6 = -1.0 / 40.0;
}  // namespace dp

// TEvalSolout over [t_prev, t_curr]; interp(te, out) evaluates the step's
// interpolant and returns false for Err(OutOfBounds) (te outside the step,
// which is every interior point of a backward step), on which the
// reference's unwrap panics: teval_emit then returns false and the
// trajectory ends with kOutOfBounds.  Writes dense[(e * D + c) * n + i]; a
// point the solout passes over without producing it (the reference pushes
// nothing) gets NaN, as do the points after the last one reached
// (teval_tail_nan at the end).  `tail` is the time of the last point pushed.
template <int D, class Interp>
DE_HD bool teval_emit(const OdeArgs& a, int64_t i, double dir, int32_t& next, int32_t& pushed,
                      double t_curr, double t_prev, const double* y_curr, Interp&& interp,
                      double& tail) {
  int32_t idx = next;
  bool ok = true;
  while (idx < a.n_eval) {
    double te = a.t_eval[idx];
    bool in_range = dir > 0.0 ? ((te == t_prev && idx == 0) || (te > t_prev && te <= t_curr))
                              : ((te == t_prev && idx == 0) || (te < t_prev && te >= t_curr));
    if (in_range) {
      double out[D];
      if (te == t_curr) {
#pragma unroll
        for (int c = 0; c < D; ++c) out[c] = y_curr[c];
      } else if (!interp(te, out)) {
        ok = false;
        break;
      }
#pragma unroll
      for (int c = 0; c < D; ++c) a.dense[((int64_t)idx * D + c) * a.n + i] = out[c];
      pushed += 1;
      tail = te;
      idx += 1;
    } else {
      if ((dir > 0.0 && te > t_curr) || (dir < 0.0 && te < t_curr)) break;
#pragma unroll
      for (int c = 0; c < D; ++c) a.dense[((int64_t)idx * D + c) * a.n + i] = as_f64(kNaNBits);
      idx += 1;
    }
  }
  next = idx;
  return ok;
}

template <int D>
DE_HD void teval_tail_nan(const OdeArgs& a, int64_t i, int32_t from) {
  for (int32_t e = from; e < a.n_eval; ++e)
#pragma unroll
    for (int c = 0; c < D; ++c) a.dense[((int64_t)e * D + c) * a.n + i] = as_f64(kNaNBits);
}

// Variable-count point outputs: DefaultSolout (src/solout/default.rs:71-91),
// DenseSolout (src/solout/dense.rs:81-111) and EvenSolout (src/solout/
// even.rs:83-190).  Point k of trajectory i goes to slot pts_offset[i] + k
// while the trajectory's slots last; `pushed` counts every point.  `last` is
// EvenSolout's last_output_t, which the start call sets (that call has
// t_prev = t0 exactly, so the solout's other first-call branch never runs).
enum : int32_t { kPointsNone = 0, kPointsSteps = 1, kPointsDense = 2, kPointsEven = 3 };

template <int D>
DE_HD void points_push(const OdeArgs& a, int64_t i, int32_t& pushed, double t, const double* y) {
  const int64_t s = a.pts_offset[i] + pushed;
  if (s < a.pts_offset[i + 1]) {
    a.pts_t[s] = t;
#pragma unroll
    for (int c = 0; c < D; ++c) a.pts_y[s * D + c] = y[c];
  }
  pushed += 1;
}

// The solout call after init (t_curr = t_prev = t0): every strategy pushes (t0, y0).
template <int D>
DE_HD void points_start(const OdeArgs& a, int64_t i, int32_t& pushed, double& last, double t0,
                        const double* y0) {
  points_push<D>(a, i, pushed, t0, y0);
  last = t0;
}

// The solout call of an accepted step [t_prev, t_curr]; false when an
// interpolation failed (the solouts unwrap it: the reference panics).
template <int D, class Interp>
DE_HD bool points_emit(const OdeArgs& a, int64_t i, double dir, double tf, int32_t& pushed,
                       double& last, double t_curr, double t_prev, const double* y_curr,
                       Interp&& interp) {
  double yi[D];
  if (a.pts_mode == kPointsSteps) {
    points_push<D>(a, i, pushed, t_curr, y_curr);
    return true;
  }
  if (a.pts_mode == kPointsDense) {
    if (t_prev != t_curr) {
      for (int32_t k = 1; k < a.pts_n; ++k) {
        const double h_old = t_curr - t_prev;
        const double ti = t_prev + (double)k * h_old / (double)a.pts_n;
        if (!interp(ti, yi)) return false;
        points_push<D>(a, i, pushed, ti, yi);
      }
    }
    points_push<D>(a, i, pushed, t_curr, y_curr);
    return true;
  }
  const double dt = a.pts_dt;
  const double tol = fabs(dt) * 1e-12 + 2.220446049250313080847e-16 * 10.0;
  double ti = last + dt * dir;
  while ((dir > 0.0 && ti <= t_curr) || (dir < 0.0 && ti >= t_curr)) {
    if ((dir > 0.0 && ti >= t_prev && ti <= t_curr) ||
        (dir < 0.0 && ti <= t_prev && ti >= t_curr)) {
      if (!(fabs(ti - last) <= tol)) {
        if (!interp(ti, yi)) return false;
        points_push<D>(a, i, pushed, ti, yi);
        last = ti;
      }
    }
    ti += dt * dir;
  }
  if (t_curr == tf) {
    if (fabs(last - tf) <= tol) {
      pushed -= 1;  // Solution::pop, then the exact final point
      points_push<D>(a, i, pushed, tf, y_curr);
      last = tf;
    } else if (last != tf) {
      points_push<D>(a, i, pushed, tf, y_curr);
      last = tf;
    }
  }
  return true;
}

// InitialStepSize::<Ordinary>::compute (src/methods/h_init.rs:45-135).
template <class F, class Tol, int ORDER = 5>
DE_HD double lane_h_init(const F& f, double t0, double tf, const double* y0, const Tol& tol,
                         double h_min, double h_max) {
  constexpr int D = F::kDim;
  double posneg = signum(tf - t0);
  double f0[D], f1[D], sk[D], y1[D];
  f.diff(t0, y0, f0);
  double dnf = 0.0, dny = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    sk[c] = tol.at(c) + tol.rt(c) * fabs(y0[c]);
    double u = f0[c] / sk[c];
    dnf += u * u;
    double v = y0[c] / sk[c];
    dny += v * v;
  }
  double h = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
  h = fmin(h, h_max);
  h *= posneg;
#pragma unroll
  for (int c = 0; c < D; ++c) y1[c] = y0[c] + h * f0[c];
  f.diff(t0 + h, y1, f1);
  double der2 = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double d = (f1[c] - f0[c]) / sk[c];
    der2 += d * d;
  }
  der2 = sqrt(der2) / fabs(h);
  double der12 = fmax(sqrt(dnf), der2);
  double h1 = der12 <= 1.0e-15 ? fabs(h) * 1.0e-3
                               : DE_CTRL_POW(0.01 / der12, 1.0 / (double)ORDER);
  double interval = fabs(tf - t0);
  h = fmin(fmax(fmin(fmin(fabs(h) * 100.0, h1), h_max), h_min), interval);
  return h * posneg;
}

// ------------------------------------------------------------------ DOPRI5 --
// DENSE selects the t_eval output path, VTOL per-component tolerances and EV
// the event solout (de_event.h) at compile time.  The attempt counter
// of the reference (`steps`, which counts attempts) is acc + rej at every
// check, and evals = base + 6 rej + 7 acc on every exit (base = 3 with the
// automatic initial step, 1 without), so neither is carried in the loop.
template <class F, bool DENSE = false, bool VTOL = false, int EV = 0>
struct Dopri5Lane {
  static constexpr int D = F::kDim;
  static constexpr int kEv = EV;
  static constexpr int kWavesPerEu = 3;  // occupancy hint of the persistent kernel
  static constexpr bool kLiveStep = true;  // step(a, i, live_in): lanes without work ride along
  // Deferred event bracket: the meta fields, then y, y_new, f(t, y) and
  // f(t_new, y_new); the root pass forms the interpolant from them.
  static constexpr int kEntryWidth = kEntryMeta + 4 * D;
  F f;
  double y[D], k0[D];
  double t, h, h_prev, tf, dir;
  double tail;  // time of the Solution's last point (NaN: none), for the event solout
  LaneTol<D, VTOL> tol;
  LaneEvent<D, EV> ev;
  int32_t acc, rej, base_ev, stiff, nonstiff, countdown, next_eval, pushed;
  bool rejected_last;

  // Result write-back and trajectory start run once per trajectory; they take
  // and return the lane by value so that the step loop's copy of the lane
  // never has its address taken (it stays in registers).
  DE_COLD static void write_back(const OdeArgs& a, int64_t i, int32_t st, Dopri5Lane L) {
    if constexpr (EV != 0) {
      if (st == kDeferOverflow) {
        event_defer_finish(a.ev, i, 0, true);
        return;
      }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) a.y_final[c * a.n + i] = L.y[c];
    a.t_final[i] = L.t;
    a.accepted[i] = L.acc;
    a.rejected[i] = L.rej;
    a.evals[i] =
        st == kBadInput ? 0 : (int64_t)L.base_ev + 6 * (int64_t)L.rej + 7 * (int64_t)L.acc;
    a.status[i] = st;
    if (DENSE && a.n_dense) a.n_dense[i] = L.pushed;
    if (DENSE && EV == 0 && a.n_pts) a.n_pts[i] = L.pushed;
    if (DENSE && a.n_eval > 0) teval_tail_nan<D>(a, i, L.next_eval);
    if constexpr (EV != 0) {
      if (a.ev.brk)
        event_defer_finish(a.ev, i, L.ev.k, false);
      else
        L.ev.finish(a.ev, a.n, i);
    }
  }
  DE_HD void finish(const OdeArgs& a, int64_t i, int32_t st) const { write_back(a, i, st, *this); }

  // The interpolant of a deferred entry, as the step's (ordinary.rs:301-336):
  // the root pass reads the entry once and forms cont = (y, ydiff, bspl,
  // cont3) with the step's operations.
  DE_HD static void entry_bounds(const double* ent, double* t0, double* t1) {
    *t0 = ent[kEntryT0];
    *t1 = ent[kEntryT0] + ent[kEntryH];
  }
  DE_HD static bool teval_base(const OdeArgs& a) { return EV == 1 && DENSE && a.n_eval > 0; }
  DE_HD static void entry_root(const OdeArgs& a, int64_t i, const double* ent, double* res) {
    const double tp = ent[kEntryT0], hp = ent[kEntryH];
    const double t_new = tp + hp;
    const double* d = ent + kEntryMeta;
    double y0p[D], cont1[D], cont2[D], cont3[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double yo = d[c], yn = d[D + c], ko = d[2 * D + c], kn = d[3 * D + c];
      y0p[c] = yo;
      double ydiff = yn + (-1.0) * yo;
      cont1[c] = ydiff;
      double bspl = hp * ko + (-1.0) * ydiff;
      cont2[c] = bspl;
      cont3[c] = (ydiff + (-hp) * kn) + (-1.0) * bspl;
    }
    const double cont4 = 0.0 * hp;
    auto interp = [&](double te, double* out) -> bool {
      if (te < tp || te > t_new) return false;
      const double s = (te - tp) / hp;
      const double s1 = 1.0 - s;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double acc_ = cont4 * s1 + cont3[c];
        acc_ = acc_ * s + cont2[c];
        acc_ = acc_ * s1 + cont1[c];
        out[c] = y0p[c] + s * acc_;
      }
      return true;
    };
    LaneEvent<D, EV>::root(a.ev, i, ent, res, tp, t_new, signum(hp), teval_base(a), interp);
  }
  DE_HD static bool entry_book(const OdeArgs& a, int64_t i) {
    return LaneEvent<D, EV>::book(a.ev, a.n, i, teval_base(a));
  }

  DE_COLD static Dopri5Lane start(const OdeArgs& a, int64_t i, bool* ok) {
    Dopri5Lane L;
    *ok = L.init(a, i);
    return L;
  }
  DE_HD bool begin(const OdeArgs& a, int64_t i) {
    bool ok;
    *this = start(a, i, &ok);
    return ok;
  }

  // Predicted number of attempts, |tf - t0| / |h0|, with h0 the initial step
  // the solve will use (the automatic one when a.h0 == 0).  Only orders work.
  DE_HD static double predicted_cost(const OdeArgs& a, int64_t i) {
    F f;
    f.load(a.params, a.n, i);
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    const double tf = a.tf ? a.tf[i] : a.tfs;
    double h0 = a.h0;
    if (h0 == 0.0 && tf != t0) {
      double y[D];
#pragma unroll
      for (int c = 0; c < D; ++c) y[c] = a.y0[c * a.n + i];
      LaneTol<D, VTOL> tl;
      tl.load(a, i);
      h0 = lane_h_init(f, t0, tf, y, tl, a.h_min, a.h_max);
    }
    return fabs(tf - t0) / fabs(h0);
  }

  // Driver prologue + DOPRI5 init.  Returns false when the trajectory ended
  // (BadInput), after writing its results.
  DE_HD bool init(const OdeArgs& a, int64_t i) {
    f.load(a.params, a.n, i);
    double t0 = a.t0 ? a.t0[i] : a.t0s;
    tf = a.tf ? a.tf[i] : a.tfs;
    tol.load(a, i);
#pragma unroll
    for (int c = 0; c < D; ++c) y[c] = a.y0[c * a.n + i];
    t = t0;
    acc = rej = 0;
    pushed = 0;
    next_eval = 0;
    ev.reset();
    double s = signum(tf - t0);
    dir = s;
    bool ok = (s == 1.0 || s == -1.0);
    double h0 = a.h0;
    base_ev = 1;
    if (ok && h0 == 0.0) {
      h0 = lane_h_init(f, t0, tf, y, tol, a.h_min, a.h_max);
      base_ev = 3;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      finish(a, i, kBadInput);
      return false;
    }
    h = h0;
    f.diff(t, y, k0);
    h_prev = 0.0;
    stiff = 0;
    nonstiff = 0;
    countdown = 100;
    rejected_last = false;
    solout_start(a, i);
    return true;
  }

  // The solout call after init (solve_ivp.rs:152-184): no interpolation, no
  // event detection; the DefaultSolout pushes (t0, y0).
  DE_HD void solout_start(const OdeArgs& a, int64_t i) {
    tail = as_f64(kNaNBits);
    if (DENSE && EV == 0 && a.pts_mode != kPointsNone)
      points_start<D>(a, i, pushed, tail, t, y);
    else if (DENSE && a.n_eval > 0)
      teval_emit<D>(a, i, dir, next_eval, pushed, t, t, y,
                    [](double, double*) { return false; }, tail);
    else
      tail = t;
    if constexpr (EV != 0) ev.start(a.ev, i, t, y);
  }

  // One pass of the driver loop: overshoot clip, one DOPRI5 attempt, solout,
  // end test.  Returns -1 while running, else the final status (the caller
  // writes the results with finish()).  A lane with live_in false (no
  // trajectory) or that ends at the pre-step checks runs the attempt's
  // arithmetic with its results discarded: the body has no branch around it,
  // so the state updates stay selects on the lane's own registers.
  DE_HD int32_t step(const OdeArgs& a, int64_t i, bool live_in = true) {
    // Overshoot clip and the two pre-step error checks of the reference.
    const bool clip = live_in && (t + h - tf) * dir > 0.0;
    const double hn = tf - t;
    const bool done = clip && fabs(hn) < kTenEps;
    h = clip ? hn : h;
    const bool small = fabs(h) < fabs(h_prev) * 1e-14;
    const bool maxed = (int64_t)(acc + rej) >= a.max_steps;
    const int32_t early = done ? kComplete : (small ? kStepSize : kMaxSteps);
    const bool live = live_in && !(done | small | maxed);
    // The stiffness test runs on accepted attempts whose number is a multiple
    // of 100 (ordinary.rs:167); the countdown restarts on every such attempt.
    countdown -= live ? 1 : 0;
    const bool stiff_check = live && countdown == 0;
    if (stiff_check) countdown = 100;

    double k1[D], k2[D], k3[D], k4[D], k5[D], k6[D], ys[D];
    const double a10h = dp::a10 * h;
#pragma unroll
    for (int c = 0; c < D; ++c) ys[c] = y[c] + a10h * k0[c];
    f.diff(t + dp::c1 * h, ys, k1);
    const double a20h = dp::a20 * h, a21h = dp::a21 * h;
#pragma unroll
    for (int c = 0; c < D; ++c) ys[c] = (y[c] + a20h * k0[c]) + a21h * k1[c];
    f.diff(t + dp::c2 * h, ys, k2);
    const double a30h = dp::a30 * h, a31h = dp::a31 * h, a32h = dp::a32 * h;
#pragma unroll
    for (int c = 0; c < D; ++c) ys[c] = ((y[c] + a30h * k0[c]) + a31h * k1[c]) + a32h * k2[c];
    f.diff(t + dp::c3 * h, ys, k3);
    const double a40h = dp::a40 * h, a41h = dp::a41 * h, a42h = dp::a42 * h,
                 a43h = dp::a43 * h;
#pragma unroll
    for (int c = 0; c < D; ++c)
      ys[c] = (((y[c] + a40h * k0[c]) + a41h * k1[c]) + a42h * k2[c]) + a43h * k3[c];
    f.diff(t + dp::c4 * h, ys, k4);
    const double a50h = dp::a50 * h, a51h = dp::a51 * h, a52h = dp::a52 * h,
                 a53h = dp::a53 * h, a54h = dp::a54 * h;
#pragma unroll
    for (int c = 0; c < D; ++c)
      ys[c] = ((((y[c] + a50h * k0[c]) + a51h * k1[c]) + a52h * k2[c]) + a53h * k3[c]) +
              a54h * k4[c];
    f.diff(t + 1.0 * h, ys, k5);
    const double a60h = dp::a60 * h, a62h = dp::a62 * h, a63h = dp::a63 * h,
                 a64h = dp::a64 * h, a65h = dp::a65 * h;
#pragma unroll
    for (int c = 0; c < D; ++c)
      ys[c] = ((((y[c] + a60h * k0[c]) + a62h * k2[c]) + a63h * k3[c]) + a64h * k4[c]) +
              a65h * k5[c];
    f.diff(t + 1.0 * h, ys, k6);

    double yseg[D], ynew[D];
    double err = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yseg[c] = (((dp::b0 * k0[c] + dp::b2 * k2[c]) + dp::b3 * k3[c]) + dp::b4 * k4[c]) +
                dp::b5 * k5[c];
      ynew[c] = y[c] + h * yseg[c];
      double es = ((((dp::e0 * k0[c] + dp::e2 * k2[c]) + dp::e3 * k3[c]) + dp::e4 * k4[c]) +
                   dp::e5 * k5[c]) +
                  dp::e6 * k6[c];
      double sk = tol.at(c) + tol.rt(c) * fmax(fabs(y[c]), fabs(ynew[c]));
      double e = es / sk;
      err = c == 0 ? e * e : err + e * e;
    }
    double deno = err;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * (double)D));
    double scale = a.safety * DE_CTRL_POW(err, -0.2);
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);

    // Accept/reject.  f(t_new, y_new) is evaluated for every attempt (a wave
    // evaluates it whenever one lane accepts) and the state is updated by
    // selects, so the loop has no divergent update paths.
    const bool accept = err <= 1.0;
    const double t_new = t + h;
    double dn[D];
    f.diff(t_new, ynew, dn);
    if (accept && stiff_check) {
      double stdnum = 0.0, stden = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double u = yseg[c] - k6[c];
        stdnum = c == 0 ? u * u : stdnum + u * u;
        double v = dn[c] - ys[c];
        stden = c == 0 ? v * v : stden + v * v;
      }
      if (stden > 0.0) {
        if (stiff_exceeds(h, stdnum, stden)) {
          nonstiff = 0;
          stiff += 1;
        }
      } else {
        nonstiff += 1;
        if (nonstiff == 6) stiff = 0;
      }
    }
    // Error::Stiffness (stiff reaches 15 only on an accepted attempt) ends the
    // solve with (t, y) at the start of the step: the attempt is not taken.
    const bool stop = live && stiff == 15;
    const bool take = live && accept && !stop;
    int32_t sol = -1;  // solout outcome: -1 continue, else the final status
    if constexpr (DENSE || EV != 0) {
      // The event function at the attempt's end is evaluated branch-free; the
      // interpolant is built only when an accepted step has t_eval points or
      // brackets an event.
      bool need = false;
      if constexpr (EV != 0) need = take && ev.pre(a.ev, t_new, ynew);
      const bool tev = DENSE && EV != 2 && take && a.n_eval > 0;
      const bool pts = DENSE && EV == 0 && take && a.pts_mode != kPointsNone;
      // cont = (y, ydiff, bspl, cont3, 0*h) and the Horner form of
      // ordinary.rs:312-333, on [t, t_new] (ordinary.rs:303-309).
      double cont1[D], cont2[D], cont3[D];
      const double hp = h, tp = t;
      const double* y0p = y;
      if (tev || (pts && a.pts_mode != kPointsSteps) || (need && !a.ev.brk)) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double ydiff = ynew[c] + (-1.0) * y[c];
          cont1[c] = ydiff;
          double bspl = h * k0[c] + (-1.0) * ydiff;
          cont2[c] = bspl;
          cont3[c] = (ydiff + (-h) * dn[c]) + (-1.0) * bspl;
        }
      }
      const double cont4 = 0.0 * h;
      auto interp = [&](double te, double* out) -> bool {
        if (te < tp || te > t_new) return false;
        double s = (te - tp) / hp;
        double s1 = 1.0 - s;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double acc_ = cont4 * s1 + cont3[c];
          acc_ = acc_ * s + cont2[c];
          acc_ = acc_ * s1 + cont1[c];
          out[c] = y0p[c] + s * acc_;
        }
        return true;
      };
      // The solout chain of an accepted step (solve_ivp.rs:242-259): TEvalSolout
      // or DefaultSolout, then EventWrappedSolout's detection (event.rs:499-522);
      // or the crossing solout alone.
      if (tev) {
        if (!teval_emit<D>(a, i, dir, next_eval, pushed, t_new, tp, ynew, interp, tail))
          sol = kOutOfBounds;
      } else if (pts) {
        if (!points_emit<D>(a, i, dir, tf, pushed, tail, t_new, tp, ynew, interp))
          sol = kOutOfBounds;
      } else {
        tail = take ? t_new : tail;
      }
      if constexpr (EV != 0) {
        if (need && sol < 0) {
          if (a.ev.brk) {
            bool overflow = false;
            double* ent = ev.defer(a.ev, i, true, tp, hp, tail, pushed, &overflow);
            if (ent) {
              double* dst = ent + kEntryMeta;
#pragma unroll
              for (int c = 0; c < D; ++c) {
                dst[c] = y0p[c];
                dst[D + c] = ynew[c];
                dst[2 * D + c] = k0[c];
                dst[3 * D + c] = dn[c];
              }
            }
            if (overflow) sol = kDeferOverflow;
          } else {
            sol = ev.post(a.ev, a.n, i, dir, true, t_new, tp, tail, interp);
          }
        }
        ev.commit(take);
      }
    }
    h_prev = take ? h : h_prev;
    t = take ? t_new : t;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      y[c] = take ? ynew[c] : y[c];
      k0[c] = take ? dn[c] : k0[c];
    }
    // After a rejection the next accepted step may not grow h.
    scale = take && rejected_last ? fmin(scale, 1.0) : scale;
    rejected_last = live ? !accept : rejected_last;
    acc += take ? 1 : 0;
    rej += live && !accept ? 1 : 0;
    const int32_t result =
        !live ? (live_in ? early : -1)
              : (stop ? kStiffness
                      : (sol >= 0 ? sol : (take && fabs(tf - t) <= kTenEps ? kComplete : -1)));
    h = live ? constrain_step_size(h * scale, a.h_min, a.h_max) : h;
    return result;
  }

  // step() + finish(): returns true when the trajectory ended.
  DE_HD bool advance(const OdeArgs& a, int64_t i) {
    int32_t st = step(a, i);
    if (st < 0) return false;
    finish(a, i, st);
    return true;
  }
};

// ------------------------------------------------------------------ DOP853 --
// The reference's generic Dormand-Prince step (dormandprince/ordinary.rs:
// 63-270) with the DOP853 tableau (src/tableau/dorman_prince.rs:155-380): 12
// stages, the second error estimate bh (deno = err + 0.01 err2), order 8 in
// the controller and h_init, stiffness test on k[11].  On every accepted step
// the reference evaluates three more stages for dense output, so evals =
// base + 11 (acc + rej) + 4 acc; the lane evaluates them only when t_eval is
// requested (they do not enter the solution).  Zero coefficients are skipped
// as in the DOPRI5 lane: each of k1..k4 (whose b and er weights are zero)
// feeds, through nonzero a entries, a stage with nonzero b and er weights, so
// a non-finite stage still makes the attempt's error non-finite.
namespace d8 {
constexpr double C[16] = DE_DOP853_C;
constexpr double A[16][16] = DE_DOP853_A;
constexpr double B[12] = DE_DOP853_B;
constexpr double BH[12] = DE_DOP853_BH;
constexpr double ER[12] = DE_DOP853_ER;
constexpr double BI[16][16] = DE_DOP853_BI;
}  // namespace d8

template <class F, bool DENSE = false, bool VTOL = false, int EV = 0>
struct Dop853Lane {
  static constexpr int D = F::kDim;
  static constexpr int kEv = EV;
  // Occupancy hint: with an event solout the 16-stage lane does not fit the
  // 168 registers of three waves per SIMD without spilling in the step.
  static constexpr int kWavesPerEu = EV != 0 ? 2 : 3;
  static constexpr bool kLiveStep = true;  // step(a, i, live_in)
  // Deferred event bracket: the meta fields, then cont[0..7].
  static constexpr int kEntryWidth = kEntryMeta + 8 * D;
  F f;
  double y[D], k0[D];
  double t, h, h_prev, tf, dir;
  double tail;
  LaneTol<D, VTOL> tol;
  LaneEvent<D, EV> ev;
  int32_t acc, rej, base_ev, stiff, nonstiff, countdown, next_eval, pushed;
  bool rejected_last;

  DE_COLD static void write_back(const OdeArgs& a, int64_t i, int32_t st, Dop853Lane L) {
    if constexpr (EV != 0) {
      if (st == kDeferOverflow) {
        event_defer_finish(a.ev, i, 0, true);
        return;
      }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) a.y_final[c * a.n + i] = L.y[c];
    a.t_final[i] = L.t;
    a.accepted[i] = L.acc;
    a.rejected[i] = L.rej;
    a.evals[i] =
        st == kBadInput ? 0 : (int64_t)L.base_ev + 11 * (int64_t)(L.acc + L.rej) + 4 * (int64_t)L.acc;
    a.status[i] = st;
    if (DENSE && a.n_dense) a.n_dense[i] = L.pushed;
    if (DENSE && EV == 0 && a.n_pts) a.n_pts[i] = L.pushed;
    if (DENSE && a.n_eval > 0) teval_tail_nan<D>(a, i, L.next_eval);
    if constexpr (EV != 0) {
      if (a.ev.brk)
        event_defer_finish(a.ev, i, L.ev.k, false);
      else
        L.ev.finish(a.ev, a.n, i);
    }
  }
  DE_HD void finish(const OdeArgs& a, int64_t i, int32_t st) const { write_back(a, i, st, *this); }

  DE_HD static void entry_bounds(const double* ent, double* t0, double* t1) {
    *t0 = ent[kEntryT0];
    *t1 = ent[kEntryT0] + ent[kEntryH];
  }
  DE_HD static bool entry_interp(const double* ent, double te, double* out) {
    const double tp = ent[kEntryT0], hp = ent[kEntryH];
    const double t_new = tp + hp;
    if (te < tp || te > t_new) return false;
    const double* d = ent + kEntryMeta;
    const double s = (te - tp) / hp;
    const double s1 = 1.0 - s;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double q = d[7 * D + c];
      q = q * s1 + d[6 * D + c];
      q = q * s + d[5 * D + c];
      q = q * s1 + d[4 * D + c];
      q = q * s1 + d[3 * D + c];
      q = q * s + d[2 * D + c];
      q = q * s1 + d[D + c];
      out[c] = d[c] + s * q;
    }
    return true;
  }
  // The root pass reads the entry once into registers.
  DE_HD static bool teval_base(const OdeArgs& a) { return EV == 1 && DENSE && a.n_eval > 0; }
  DE_HD static void entry_root(const OdeArgs& a, int64_t i, const double* ent, double* res) {
    double loc[kEntryWidth];
#pragma unroll
    for (int f = 0; f < kEntryWidth; ++f) loc[f] = ent[f];
    auto interp = [&](double te, double* out) { return entry_interp(loc, te, out); };
    double t0, t1;
    entry_bounds(loc, &t0, &t1);
    LaneEvent<D, EV>::root(a.ev, i, loc, res, t0, t1, signum(loc[kEntryH]), teval_base(a),
                           interp);
  }
  DE_HD static bool entry_book(const OdeArgs& a, int64_t i) {
    return LaneEvent<D, EV>::book(a.ev, a.n, i, teval_base(a));
  }

  DE_COLD static Dop853Lane start(const OdeArgs& a, int64_t i, bool* ok) {
    Dop853Lane L;
    *ok = L.init(a, i);
    return L;
  }
  DE_HD bool begin(const OdeArgs& a, int64_t i) {
    bool ok;
    *this = start(a, i, &ok);
    return ok;
  }

  DE_HD static double predicted_cost(const OdeArgs& a, int64_t i) {
    F f;
    f.load(a.params, a.n, i);
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    const double tf = a.tf ? a.tf[i] : a.tfs;
    double h0 = a.h0;
    if (h0 == 0.0 && tf != t0) {
      double y[D];
#pragma unroll
      for (int c = 0; c < D; ++c) y[c] = a.y0[c * a.n + i];
      LaneTol<D, VTOL> tl;
      tl.load(a, i);
      h0 = lane_h_init<F, LaneTol<D, VTOL>, 8>(f, t0, tf, y, tl, a.h_min, a.h_max);
    }
    return fabs(tf - t0) / fabs(h0);
  }

  DE_HD bool init(const OdeArgs& a, int64_t i) {
    f.load(a.params, a.n, i);
    double t0 = a.t0 ? a.t0[i] : a.t0s;
    tf = a.tf ? a.tf[i] : a.tfs;
    tol.load(a, i);
#pragma unroll
    for (int c = 0; c < D; ++c) y[c] = a.y0[c * a.n + i];
    t = t0;
    acc = rej = 0;
    pushed = 0;
    next_eval = 0;
    ev.reset();
    double s = signum(tf - t0);
    dir = s;
    bool ok = (s == 1.0 || s == -1.0);
    double h0 = a.h0;
    base_ev = 1;
    if (ok && h0 == 0.0) {
      h0 = lane_h_init<F, LaneTol<D, VTOL>, 8>(f, t0, tf, y, tol, a.h_min, a.h_max);
      base_ev = 3;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      finish(a, i, kBadInput);
      return false;
    }
    h = h0;
    f.diff(t, y, k0);
    h_prev = 0.0;
    stiff = 0;
    nonstiff = 0;
    countdown = 100;
    rejected_last = false;
    tail = as_f64(kNaNBits);
    if (DENSE && EV == 0 && a.pts_mode != kPointsNone)
      points_start<D>(a, i, pushed, tail, t, y);
    else if (DENSE && a.n_eval > 0)
      teval_emit<D>(a, i, dir, next_eval, pushed, t, t, y,
                    [](double, double*) { return false; }, tail);
    else
      tail = t;
    if constexpr (EV != 0) ev.start(a.ev, i, t, y);
    return true;
  }

  // y + sum_{j<i} (a[i][j] h) k[j], in j order, zero coefficients skipped.
  template <int I>
  DE_HD static void stage_input(const double* y, const double (*k)[D], double h, double* ys) {
#pragma unroll
    for (int c = 0; c < D; ++c) ys[c] = y[c];
#pragma unroll
    for (int j = 0; j < I; ++j) {
      if (d8::A[I][j] != 0.0) {
        const double ah = d8::A[I][j] * h;
#pragma unroll
        for (int c = 0; c < D; ++c) ys[c] = ys[c] + ah * k[j][c];
      }
    }
  }

  template <int I>
  DE_HD void stages(double (*k)[D], double* ys) const {
    if constexpr (I < 12) {
      stage_input<I>(y, k, h, ys);
      f.diff(t + d8::C[I] * h, ys, k[I]);
      stages<I + 1>(k, ys);
    }
  }

  template <int I>
  DE_HD void dense_stages(double (*k)[D], const double* yy, double tt, double hh) const {
    if constexpr (I < 16) {
      double ys[D];
      stage_input<I>(yy, k, hh, ys);
      f.diff(tt + d8::C[I] * hh, ys, k[I]);
      dense_stages<I + 1>(k, yy, tt, hh);
    }
  }

  // sum_j w[j] k[j] starting from the first nonzero weight (0 + x == x).
  template <int N>
  DE_HD static double comb(const double* w, const double (*k)[D], int c) {
    double s = 0.0;
    bool first = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (w[j] != 0.0) {
        s = first ? w[j] * k[j][c] : s + w[j] * k[j][c];
        first = false;
      }
    }
    return s;
  }

  // As Dopri5Lane::step: lanes with live_in false or ending at the pre-step
  // checks run the attempt with its results discarded.
  DE_HD int32_t step(const OdeArgs& a, int64_t i, bool live_in = true) {
    const bool clip = live_in && (t + h - tf) * dir > 0.0;
    const double hn = tf - t;
    const bool done = clip && fabs(hn) < kTenEps;
    h = clip ? hn : h;
    const bool small = fabs(h) < fabs(h_prev) * 1e-14;
    const bool maxed = (int64_t)(acc + rej) >= a.max_steps;
    const int32_t early = done ? kComplete : (small ? kStepSize : kMaxSteps);
    const bool live = live_in && !(done | small | maxed);
    countdown -= live ? 1 : 0;
    const bool stiff_check = live && countdown == 0;
    if (stiff_check) countdown = 100;

    double k[(DENSE || EV != 0) ? 16 : 12][D];
#pragma unroll
    for (int c = 0; c < D; ++c) k[0][c] = k0[c];
    double ys[D];
    stages<1>(k, ys);  // ys ends as the stage-11 input (ysti)

    double yseg[D], ynew[D];
    double err = 0.0, err2 = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yseg[c] = comb<12>(d8::B, k, c);
      ynew[c] = y[c] + h * yseg[c];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double es = comb<12>(d8::ER, k, c);
      double e2s = yseg[c];
#pragma unroll
      for (int j = 0; j < 12; ++j)
        if (d8::BH[j] != 0.0) e2s = e2s + (-d8::BH[j]) * k[j][c];
      const double sk = tol.at(c) + tol.rt(c) * fmax(fabs(y[c]), fabs(ynew[c]));
      const double e = es / sk;
      const double e2 = e2s / sk;
      err = c == 0 ? e * e : err + e * e;
      err2 = c == 0 ? e2 * e2 : err2 + e2 * e2;
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * (double)D));
    double scale = a.safety * DE_CTRL_POW(err, -0.125);
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);

    const bool accept = err <= 1.0;
    const double t_new = t + h;
    double dn[D];
    f.diff(t_new, ynew, dn);
    if (accept && stiff_check) {
      double stdnum = 0.0, stden = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double u = yseg[c] - k[11][c];
        stdnum = c == 0 ? u * u : stdnum + u * u;
        double v = dn[c] - ys[c];
        stden = c == 0 ? v * v : stden + v * v;
      }
      if (stden > 0.0) {
        if (stiff_exceeds(h, stdnum, stden)) {
          nonstiff = 0;
          stiff += 1;
        }
      } else {
        nonstiff += 1;
        if (nonstiff == 6) stiff = 0;
      }
    }
    const bool stop = live && stiff == 15;
    const bool take = live && accept && !stop;
    int32_t sol = -1;
    if constexpr (DENSE || EV != 0) {
      // The continuous extension is built only when an accepted step has
      // t_eval points or brackets an event: cont[0..3] as DOPRI5; k[12] =
      // f(t_new, y_new); k[13..15] at c[13..15]; cont[i] = h sum_j bi[i][j] k[j]
      // for i = 4..7.
      bool need = false;
      if constexpr (EV != 0) need = take && ev.pre(a.ev, t_new, ynew);
      const bool tev = DENSE && EV != 2 && take && a.n_eval > 0;
      const bool pts = DENSE && EV == 0 && take && a.pts_mode != kPointsNone;
      double cont[8][D];
      if (tev || (pts && a.pts_mode != kPointsSteps) || need) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double ydiff = ynew[c] + (-1.0) * y[c];
          cont[0][c] = y[c];
          cont[1][c] = ydiff;
          double bspl = h * k[0][c] + (-1.0) * ydiff;
          cont[2][c] = bspl;
          cont[3][c] = (ydiff + (-h) * dn[c]) + (-1.0) * bspl;
          k[12][c] = dn[c];
        }
        dense_stages<13>(k, y, t, h);
#pragma unroll
        for (int r = 4; r < 8; ++r)
#pragma unroll
          for (int c = 0; c < D; ++c) cont[r][c] = comb<16>(d8::BI[r], k, c) * h;
      }
      const double hp = h, tp = t;
      auto interp = [&](double te, double* out) -> bool {
        // ordinary.rs:303-333 with 8 coefficients: factors s1, s, s1 for
        // cont[6], cont[5], cont[4], then s1, s, s1 for cont[3..1].
        if (te < tp || te > t_new) return false;
        const double s = (te - tp) / hp;
        const double s1 = 1.0 - s;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double q = cont[7][c];
          q = q * s1 + cont[6][c];
          q = q * s + cont[5][c];
          q = q * s1 + cont[4][c];
          q = q * s1 + cont[3][c];
          q = q * s + cont[2][c];
          q = q * s1 + cont[1][c];
          out[c] = cont[0][c] + s * q;
        }
        return true;
      };
      if (tev) {
        if (!teval_emit<D>(a, i, dir, next_eval, pushed, t_new, tp, ynew, interp, tail))
          sol = kOutOfBounds;
      } else if (pts) {
        if (!points_emit<D>(a, i, dir, tf, pushed, tail, t_new, tp, ynew, interp))
          sol = kOutOfBounds;
      } else {
        tail = take ? t_new : tail;
      }
      if constexpr (EV != 0) {
        if (need && sol < 0) {
          if (a.ev.brk) {
            bool overflow = false;
            double* ent = ev.defer(a.ev, i, true, tp, hp, tail, pushed, &overflow);
            if (ent) {
#pragma unroll
              for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) ent[kEntryMeta + r * D + c] = cont[r][c];
            }
            if (overflow) sol = kDeferOverflow;
          } else {
            sol = ev.post(a.ev, a.n, i, dir, true, t_new, tp, tail, interp);
          }
        }
        ev.commit(take);
      }
    }
    h_prev = take ? h : h_prev;
    t = take ? t_new : t;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      y[c] = take ? ynew[c] : y[c];
      k0[c] = take ? dn[c] : k0[c];
    }
    scale = take && rejected_last ? fmin(scale, 1.0) : scale;
    rejected_last = live ? !accept : rejected_last;
    acc += take ? 1 : 0;
    rej += live && !accept ? 1 : 0;
    const int32_t result =
        !live ? (live_in ? early : -1)
              : (stop ? kStiffness
                      : (sol >= 0 ? sol : (take && fabs(tf - t) <= kTenEps ? kComplete : -1)));
    h = live ? constrain_step_size(h * scale, a.h_min, a.h_max) : h;
    return result;
  }

  DE_HD bool advance(const OdeArgs& a, int64_t i) {
    int32_t st = step(a, i);
    if (st < 0) return false;
    finish(a, i, st);
    return true;
  }
};

// --------------------------------------------------- fixed-step rk4 / euler --
// RK4: c = (0, 1/2, 1/2, 1), a10 = a21 = 1/2, a32 = 1, b = (1/6, 1/3, 1/3, 1/6)
// (src/tableau/runge_kutta.rs:39-70); euler: b = (1) (:289-302).
template <class F, int STAGES, int EV = 0>
struct FixedLane {
  static constexpr int D = F::kDim;
  static constexpr int kEv = EV;
  static constexpr int kWavesPerEu = 3;  // occupancy hint of the persistent kernel
  // Deferred event bracket: the meta fields (kEntryH holds the step's end),
  // then y_prev, dydt_prev, y, dydt of the Hermite interpolant.
  static constexpr int kEntryWidth = kEntryMeta + 4 * D;
  F f;
  double y[D], k0[D];
  double t, h, tf, dir;
  double tail;
  int64_t steps, acc, ev;
  LaneEvent<D, EV> evs;
  int32_t next_eval, pushed;

  DE_HD void finish(const OdeArgs& a, int64_t i, int32_t st) const {
    if constexpr (EV != 0) {
      if (st == kDeferOverflow) {
        event_defer_finish(a.ev, i, 0, true);
        return;
      }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) a.y_final[c * a.n + i] = y[c];
    a.t_final[i] = t;
    a.accepted[i] = acc;
    a.rejected[i] = 0;
    a.evals[i] = ev;
    a.status[i] = st;
    if (a.n_dense) a.n_dense[i] = pushed;
    if (EV == 0 && a.n_pts) a.n_pts[i] = pushed;
    if (a.n_eval > 0) teval_tail_nan<D>(a, i, next_eval);
    if constexpr (EV != 0) {
      if (a.ev.brk)
        event_defer_finish(a.ev, i, evs.k, false);
      else
        evs.finish(a.ev, a.n, i);
    }
  }

  // cubic_hermite_interpolate of a deferred entry, as the step's.
  DE_HD static void entry_bounds(const double* ent, double* t0, double* t1) {
    *t0 = ent[kEntryT0];
    *t1 = ent[kEntryH];
  }
  DE_HD static bool entry_interp(const double* ent, double te, double* out) {
    const double tp = ent[kEntryT0], tc = ent[kEntryH];
    if (te < tp || te > tc) return false;
    const double* d = ent + kEntryMeta;
    double hh = tc - tp;
    double s = (te - tp) / hh;
    double s2 = s * s;
    double s3 = s2 * s;
    double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
    double h10 = s3 - 2.0 * s2 + s;
    double h01 = -2.0 * s3 + 3.0 * s2;
    double h11 = s3 - s2;
    double w10 = h10 * hh, w11 = h11 * hh;
#pragma unroll
    for (int c = 0; c < D; ++c)
      out[c] = ((h00 * d[c] + w10 * d[D + c]) + h01 * d[2 * D + c]) + w11 * d[3 * D + c];
    return true;
  }
  DE_HD static bool teval_base(const OdeArgs& a) { return EV == 1 && a.n_eval > 0; }
  DE_HD static void entry_root(const OdeArgs& a, int64_t i, const double* ent, double* res) {
    double loc[kEntryWidth];
#pragma unroll
    for (int f = 0; f < kEntryWidth; ++f) loc[f] = ent[f];
    auto interp = [&](double te, double* out) { return entry_interp(loc, te, out); };
    double t0, t1;
    entry_bounds(loc, &t0, &t1);
    LaneEvent<D, EV>::root(a.ev, i, loc, res, t0, t1, signum(t1 - t0), teval_base(a), interp);
  }
  DE_HD static bool entry_book(const OdeArgs& a, int64_t i) {
    return LaneEvent<D, EV>::book(a.ev, a.n, i, teval_base(a));
  }

  DE_HD bool begin(const OdeArgs& a, int64_t i) { return init(a, i); }

  // Number of steps |tf - t0| / |h0| (only orders work).
  DE_HD static double predicted_cost(const OdeArgs& a, int64_t i) {
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    const double tf = a.tf ? a.tf[i] : a.tfs;
    const double h0 = a.h0 == 0.0 ? fabs(tf - t0) / 100.0 : a.h0;
    return fabs(tf - t0) / fabs(h0);
  }

  DE_HD bool init(const OdeArgs& a, int64_t i) {
    f.load(a.params, a.n, i);
    double t0 = a.t0 ? a.t0[i] : a.t0s;
    tf = a.tf ? a.tf[i] : a.tfs;
#pragma unroll
    for (int c = 0; c < D; ++c) y[c] = a.y0[c * a.n + i];
    t = t0;
    acc = ev = 0;
    pushed = 0;
    next_eval = 0;
    evs.reset();
    double s = signum(tf - t0);
    dir = s;
    bool ok = (s == 1.0 || s == -1.0);
    double h0 = a.h0;
    if (h0 == 0.0) h0 = fabs(tf - t0) / 100.0;
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      finish(a, i, kBadInput);
      return false;
    }
    h = h0;
    f.diff(t, y, k0);
    ev = 1;
    steps = 0;
    tail = as_f64(kNaNBits);
    if (EV == 0 && a.pts_mode != kPointsNone)
      points_start<D>(a, i, pushed, tail, t, y);
    else if (EV != 2 && a.n_eval > 0)
      teval_emit<D>(a, i, dir, next_eval, pushed, t, t, y,
                    [](double, double*) { return false; }, tail);
    else
      tail = t;
    if constexpr (EV != 0) evs.start(a.ev, i, t, y);
    return true;
  }

  // One driver pass; -1 while running, else the final status.
  DE_HD int32_t step(const OdeArgs& a, int64_t i) {
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) return kComplete;
      h = hn;
    }
    if (steps >= a.max_steps) return kMaxSteps;
    steps += 1;
    double ynext[D];
    if (STAGES == 4) {
      double k1[D], k2[D], k3[D], ys[D];
      const double hh = 0.5 * h;
#pragma unroll
      for (int c = 0; c < D; ++c) ys[c] = y[c] + hh * k0[c];
      f.diff(t + 0.5 * h, ys, k1);
#pragma unroll
      for (int c = 0; c < D; ++c) ys[c] = y[c] + hh * k1[c];
      f.diff(t + 0.5 * h, ys, k2);
      const double h1 = 1.0 * h;
#pragma unroll
      for (int c = 0; c < D; ++c) ys[c] = y[c] + h1 * k2[c];
      f.diff(t + 1.0 * h, ys, k3);
      const double b0h = (1.0 / 6.0) * h, b1h = (1.0 / 3.0) * h;
#pragma unroll
      for (int c = 0; c < D; ++c)
        ynext[c] = (((y[c] + b0h * k0[c]) + b1h * k1[c]) + b1h * k2[c]) + b0h * k3[c];
      ev += 3;
    } else {
      const double b0h = 1.0 * h;
#pragma unroll
      for (int c = 0; c < D; ++c) ynext[c] = y[c] + b0h * k0[c];
    }
    double tp = t, hp = h;
    t += h;
    double dn[D];
    f.diff(t, ynext, dn);
    ev += 1;
    acc += 1;
    int32_t sol = -1;
    if (EV != 0 || a.n_eval > 0 || a.pts_mode != kPointsNone) {
      const double* yp = y;
      const double* kp = k0;
      const double tc = t;
      auto interp = [&](double te, double* out) -> bool {
        // cubic_hermite_interpolate(t_prev, t, y_prev, y, dydt_prev, dydt, te)
        // on [t_prev, t] (fixed/ordinary.rs:170-177).
        if (te < tp || te > tc) return false;
        double hh = tc - tp;
        double s = (te - tp) / hh;
        double s2 = s * s;
        double s3 = s2 * s;
        double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
        double h10 = s3 - 2.0 * s2 + s;
        double h01 = -2.0 * s3 + 3.0 * s2;
        double h11 = s3 - s2;
        double w10 = h10 * hh, w11 = h11 * hh;
#pragma unroll
        for (int c = 0; c < D; ++c)
          out[c] = ((h00 * yp[c] + w10 * kp[c]) + h01 * ynext[c]) + w11 * dn[c];
        return true;
      };
      if (EV != 2 && a.n_eval > 0) {
        if (!teval_emit<D>(a, i, dir, next_eval, pushed, t, tp, ynext, interp, tail))
          sol = kOutOfBounds;
      } else if (EV == 0 && a.pts_mode != kPointsNone) {
        if (!points_emit<D>(a, i, dir, tf, pushed, tail, t, tp, ynext, interp))
          sol = kOutOfBounds;
      } else {
        tail = t;
      }
      if constexpr (EV != 0) {
        const bool need = evs.pre(a.ev, t, ynext);
        if (need && sol < 0) {
          if (a.ev.brk) {
            bool overflow = false;
            double* ent = evs.defer(a.ev, i, true, tp, tc, tail, pushed, &overflow);
            if (ent) {
              double* dst = ent + kEntryMeta;
#pragma unroll
              for (int c = 0; c < D; ++c) {
                dst[c] = yp[c];
                dst[D + c] = kp[c];
                dst[2 * D + c] = ynext[c];
                dst[3 * D + c] = dn[c];
              }
            }
            if (overflow) sol = kDeferOverflow;
          } else {
            sol = evs.post(a.ev, a.n, i, dir, true, t, tp, tail, interp);
          }
        }
        evs.commit(true);
      }
    }
    (void)hp;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      y[c] = ynext[c];
      k0[c] = dn[c];
    }
    if (sol >= 0) return sol;
    if (fabs(tf - t) <= kTenEps) return kComplete;
    return -1;
  }

  DE_HD bool advance(const OdeArgs& a, int64_t i) {
    int32_t st = step(a, i);
    if (st < 0) return false;
    finish(a, i, st);
    return true;
  }
};

}  // namespace de
