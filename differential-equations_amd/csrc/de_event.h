// Event solouts of one lane (device code), restated from the reference:
//   EventWrappedSolout   src/solout/event.rs:320-523 (detect_event :335-393,
//                        brent_dekker :395-490, solout :499-522)
//   CrossingSolout       src/solout/crossing.rs:181-320
//   HyperplaneCrossing   src/solout/hyperplane.rs:126-340
// A lane calls pre() after every attempt with the attempt's end point (g is
// evaluated branch-free; only accepted steps count); when the accepted step
// brackets a root, the lane builds the interpolant and calls post() (inline
// root) or defer() (deferred root).  commit(take) then advances the solout's
// state to the step's end.
//
// Interpolants are passed as callables interp(t, out[D]) -> bool that return
// false for the reference's Err(OutOfBounds) (t outside [t_prev, t] of the
// step).  The reference's brent_dekker treats that as "no root" (`.ok()?`);
// the other interpolations unwrap it, i.e. the reference panics, which ends
// the trajectory here with kOutOfBounds.
//
// Deferred roots.  A root search runs for a few dozen iterations on one lane
// while the rest of its wave waits, and a Poincare-section event brackets
// every few dozen steps, so found inline the searches dominate the solve.
// Without a terminate count (EventConfig::terminate = None) nothing in the
// integration depends on a root, so the lane stores each bracketed step (its
// interpolant and bracket) in a per-trajectory region in HBM and moves on; a
// root pass then searches every bracket with one thread each, and a
// bookkeeping pass walks each trajectory's brackets in order, applying the
// solout's push rules.  A trajectory that outgrows its region, or whose
// search hits a failed interpolation (a reference panic, which must end the
// solve at that step), is solved again with inline roots.
#pragma once

#include "de_common.h"

namespace de {

enum : int32_t { kSoloutDefault = 0, kSoloutEvent = 1, kSoloutCrossing = 2, kSoloutHyperplane = 3 };
enum : int32_t { kEventComponent = 0, kEventTime = 1, kEventPlane = 2 };
enum : int32_t { kCrossBoth = 0, kCrossPositive = 1, kCrossNegative = 2 };

struct EventArgs {
  int32_t solout;      // kSolout*
  int32_t fn;          // kEvent*: kEventComponent for crossings, kEventPlane for hyperplanes
  int32_t comp;
  int32_t direction;   // kCross*
  int32_t terminate;   // EventConfig::terminate, 0 = None
  int32_t max_events;  // points kept per trajectory
  double value;
  const double* value_traj;
  const double* point;
  const double* normal;  // the hyperplane solout's is normalised before the launch
  double* t_out;         // [max_events * n]: t_out[k * n + i]
  double* y_out;         // [(k * D + c) * n + i]
  int32_t* n_out;        // [n] points pushed, or null
  // Deferred roots (null brk: inline), for the slice of trajectories that
  // starts at `base`.  Entry k of trajectory i is brk[((i - base) * bmax + k)
  // * width ...]: kEntryMeta fields, then the lane's interpolant data; its
  // result record is res[((i - base) * bmax + k) * kResWidth ...].
  double* brk;
  double* res;
  int32_t* bcount;  // [slice] brackets stored, -1: solve again inline
  int32_t bmax;
  int32_t width;
  int64_t base;
};

// Fields of a deferred bracket: the step's start and its length (or end, for
// the fixed-step lanes' Hermite interpolant), g at both ends, and the time and
// number of the t_eval points pushed before the bracket was seen.
enum : int { kEntryT0 = 0, kEntryH = 1, kEntryG0 = 2, kEntryG1 = 3, kEntryTail = 4,
             kEntryTevals = 5, kEntryMeta = 6 };
// Result record of a bracket, written by the root pass: flag, time, state,
// then the t_eval fields the push rule needs; one 64-byte aligned record for
// states of up to four components.
enum : int { kResFlag = 0, kResT = 1, kResY = 2 };
enum : int { kResNone = 0, kResPoint = 1, kResPanic = 2 };
constexpr int kResWidth = 8;
// Lane status of a trajectory whose deferred region is full: its results are
// produced by the inline solve that follows.
constexpr int32_t kDeferOverflow = 100;

// g(t, y) of the event function: y[comp] - value, t - value, or
// y.minus(&point).dot(&normal) (State::minus = plus_scaled(-1, ..), dot folds
// from zero; src/traits.rs:336-354,551-555).
template <int D>
DE_HD double event_g(const EventArgs& e, double value, double t, const double* y) {
  if (e.fn == kEventTime) return t - value;
  if (e.fn == kEventComponent) {
    // y[comp] as an OR of masked bit patterns: a select chain over y[c] is
    // turned back into an indexed load by the compiler, which moves y from
    // registers to scratch memory.
    uint64_t bits = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) bits |= (e.comp == c ? ~0ull : 0ull) & as_u64(y[c]);
    return as_f64(bits) - value;
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) s = s + (y[c] + (-1.0) * e.point[c]) * e.normal[c];
  return s;
}

// f64::signum(a) != f64::signum(b); a NaN signum differs from everything.
DE_HD bool sign_differs(double a, double b) { return !(signum(a) == signum(b)); }

// Writes point k of trajectory i when it fits.
template <int D>
DE_HD void event_store(const EventArgs& e, int64_t n, int64_t i, int32_t k, double t,
                       const double* y) {
  if (k < e.max_events) {
    e.t_out[(int64_t)k * n + i] = t;
#pragma unroll
    for (int c = 0; c < D; ++c) e.y_out[((int64_t)k * D + c) * n + i] = y[c];
  }
}

// Count and NaN for the slots past the last point.
template <int D>
DE_HD void event_finish(const EventArgs& e, int64_t n, int64_t i, int32_t pushed) {
  if (e.n_out) e.n_out[i] = pushed;
  for (int32_t k = pushed; k < e.max_events; ++k) {
    e.t_out[(int64_t)k * n + i] = as_f64(kNaNBits);
    for (int c = 0; c < D; ++c) e.y_out[((int64_t)k * D + c) * n + i] = as_f64(kNaNBits);
  }
}

// EventWrappedSolout::brent_dekker (src/solout/event.rs:395-490) as written.
// After every iteration it sets c = a (the previous iterate), so the bracket
// is not maintained and the iteration can stop short of the root; the values
// are the reference's.  interp(b, yb) false ends the search without a root.
template <int D, class Interp, class G>
DE_HD bool brent_dekker(double a, double b, double fa, double fb, Interp& interp, G& g,
                        double* root) {
  if (fabs(fa) < fabs(fb)) {
    double x = a;
    a = b;
    b = x;
    x = fa;
    fa = fb;
    fb = x;
  }
  double c = a, fc = fa;
  double d = b - a;
  double e = d;
  for (int it = 0; it < 50; ++it) {
    if (fb == 0.0) {
      *root = b;
      return true;
    }
    if (signum(fa) == signum(fb)) {
      a = c;
      fa = fc;
      c = b;
      fc = fb;
      d = b - a;
      e = d;
    }
    if (fabs(fa) < fabs(fb)) {
      c = b;
      b = a;
      a = c;
      fc = fb;
      fb = fa;
      fa = fc;
    }
    const double tol = fmax(1e-14, 1e-12 * fabs(b));
    const double m = 0.5 * (a - b);
    if (fabs(m) <= tol || fb == 0.0) {
      *root = b;
      return true;
    }
    bool bisect = true;
    if (fabs(e) > tol && fabs(fa) > fabs(fb)) {
      const double s = fb / fa;
      double p, q;
      if (a == c) {
        p = 2.0 * m * s;
        q = 1.0 - s;
      } else {
        const double q1 = fa / fc;
        const double r = fb / fc;
        p = s * (2.0 * m * q1 * (q1 - r) - (b - a) * (r - 1.0));
        q = (q1 - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (q > 0.0) p = -p;
      else q = -q;
      if (fabs(2.0 * p) < (3.0 * m * q - fabs(tol * q)) && p < fabs(e * 0.5 * q)) {
        e = d;
        d = p / q;
        bisect = false;
      }
    }
    if (bisect) {
      d = m;
      e = m;
    }
    a = b;
    fa = fb;
    b = fabs(d) > tol ? b + d : b + (m > 0.0 ? tol : -tol);
    double yb[D];
    if (!interp(b, yb)) return false;
    fb = g(b, yb);
    c = a;
    fc = fa;
  }
  return false;
}

// Event state of a lane; EV selects the solout at compile time:
// 0 none, 1 EventWrappedSolout, 2 CrossingSolout / HyperplaneCrossingSolout.
template <int D, int EV>
struct LaneEvent {
  DE_HD void reset() {}
  DE_HD void start(const EventArgs&, int64_t, double, const double*) {}
  DE_HD void finish(const EventArgs&, int64_t, int64_t) const {}
  DE_HD void commit(bool) {}
};

// Shared by the two solouts: the deferred entry of a bracketed step.
// Returns the entry to fill with the interpolant, or null (no bracket, or the
// region is full: *overflow).
DE_HD double* event_defer(const EventArgs& e, int64_t i, int32_t& k, bool need, double t0,
                          double h, double g0, double g1, double tail, int32_t tevals,
                          bool* overflow) {
  if (!need) return nullptr;
  if (k >= e.bmax) {
    *overflow = true;
    return nullptr;
  }
  double* ent = e.brk + ((i - e.base) * e.bmax + k) * e.width;
  ent[kEntryT0] = t0;
  ent[kEntryH] = h;
  ent[kEntryG0] = g0;
  ent[kEntryG1] = g1;
  ent[kEntryTail] = tail;
  ent[kEntryTevals] = (double)tevals;
  k += 1;
  return ent;
}

// Result record of a root search.
template <int D>
DE_HD void event_result(double* r, int flag, double t, const double* y, double tail,
                        double tevals) {
  static_assert(4 + D <= kResWidth, "result record");
  r[kResFlag] = (double)flag;
  r[kResT] = t;
#pragma unroll
  for (int c = 0; c < D; ++c) r[kResY + c] = flag == kResPoint ? y[c] : 0.0;
  r[kResY + D] = tail;
  r[kResY + D + 1] = tevals;
}

// Deferred-mode end of a trajectory: the bracket count (the bookkeeping pass
// writes the event outputs), or -1 after an overflow.
DE_HD void event_defer_finish(const EventArgs& e, int64_t i, int32_t k, bool overflow) {
  e.bcount[i - e.base] = overflow ? -1 : k;
}

// EventWrappedSolout around DefaultSolout or TEvalSolout.
template <int D>
struct LaneEvent<D, 1> {
  double value, last_g, g_new;
  int32_t count, pushed;
  int32_t k;  // deferred brackets stored

  DE_HD void reset() {
    pushed = 0;
    k = 0;
  }
  // The solout call after init: only records g(t0, y0) (event.rs:346-355).
  DE_HD void start(const EventArgs& e, int64_t i, double t0, const double* y0) {
    value = e.value_traj ? e.value_traj[i] : e.value;
    last_g = event_g<D>(e, value, t0, y0);
    count = 0;
  }
  // Sign change in the configured direction, with a bracket to search.
  DE_HD bool pre(const EventArgs& e, double t_curr, const double* y_curr) {
    g_new = event_g<D>(e, value, t_curr, y_curr);
    const double g_prev = last_g;
    const bool change = sign_differs(g_prev, g_new);
    const bool ok = e.direction == kCrossPositive   ? change && g_prev < 0.0 && g_new >= 0.0
                    : e.direction == kCrossNegative ? change && g_prev > 0.0 && g_new <= 0.0
                                                    : change;
    return ok && g_prev * g_new <= 0.0;
  }
  // detect_event after pre(): the root, the push unless within 1e-14 of the
  // Solution's last time `tail` (NaN: no point yet), the terminate count.
  // Returns -1 to continue, else the trajectory's final status.
  template <class Interp>
  DE_HD int32_t post(const EventArgs& e, int64_t n, int64_t i, double dir, bool need,
                     double t_curr, double t_prev, double& tail, Interp& interp) {
    int32_t st = -1;
    if (need) {
      double a = t_prev, b = t_curr, fa = last_g, fb = g_new;
      if ((dir > 0.0 && a > b) || (dir < 0.0 && a < b)) {
        double x = a;
        a = b;
        b = x;
        x = fa;
        fa = fb;
        fb = x;
      }
      const EventArgs& ee = e;
      const double v = value;
      auto g = [&](double t, const double* y) { return event_g<D>(ee, v, t, y); };
      double t_event;
      if (brent_dekker<D>(a, b, fa, fb, interp, g, &t_event)) {
        double ye[D];
        if (!interp(t_event, ye)) {
          st = kOutOfBounds;
        } else {
          if (tail != tail || fabs(t_event - tail) > 1e-14) {
            event_store<D>(e, n, i, pushed, t_event, ye);
            pushed += 1;
            tail = t_event;
          }
          count += 1;
          if (e.terminate > 0 && count >= e.terminate) st = kInterrupted;
        }
      }
    }
    return st;
  }
  DE_HD void commit(bool take) { last_g = take ? g_new : last_g; }
  DE_HD void finish(const EventArgs& e, int64_t n, int64_t i) const {
    event_finish<D>(e, n, i, pushed);
  }
  // Deferred mode after pre(): the bracket's entry (or null), and the state
  // advanced as post() does.  `tail` / `tevals` are the last t_eval point pushed
  // and the number pushed so far.
  DE_HD double* defer(const EventArgs& e, int64_t i, bool need, double t0, double h,
                      double tail, int32_t tevals, bool* overflow) {
    return event_defer(e, i, k, need, t0, h, last_g, g_new, tail, tevals, overflow);
  }
  // Root pass on one entry: brent_dekker on [t_prev, t_curr] as post() runs it.
  // Under DefaultSolout the Solution's last time is the step's end, so the
  // push rule is decided here; under TEvalSolout it depends on the earlier
  // brackets and is left to book().
  template <class Interp>
  DE_HD static void root(const EventArgs& e, int64_t i, const double* ent, double* res,
                         double t_prev, double t_curr, double dir, bool teval_base,
                         Interp& interp) {
    const double v = e.value_traj ? e.value_traj[i] : e.value;
    double a = t_prev, b = t_curr, fa = ent[kEntryG0], fb = ent[kEntryG1];
    if ((dir > 0.0 && a > b) || (dir < 0.0 && a < b)) {
      double x = a;
      a = b;
      b = x;
      x = fa;
      fa = fb;
      fb = x;
    }
    auto g = [&](double t, const double* y) { return event_g<D>(e, v, t, y); };
    double t_event = 0.0, ye[D];
    int flag = kResNone;
    if (brent_dekker<D>(a, b, fa, fb, interp, g, &t_event))
      flag = interp(t_event, ye) ? kResPoint : kResPanic;
    if (!teval_base && flag == kResPoint && !(fabs(t_event - t_curr) > 1e-14)) flag = kResNone;
    event_result<D>(res, flag, t_event, ye, ent[kEntryTail], ent[kEntryTevals]);
  }
  // Bookkeeping of one trajectory's entries, in order: the push rule of
  // detect_event (event.rs:379-385) against the Solution's last time, which is
  // the step's end under DefaultSolout, else the last t_eval point pushed
  // since the previous bracket or the previous event point.  Returns false when
  // the trajectory must be solved again inline.
  // One pass over the trajectory's result records; a failed interpolation
  // (a reference panic) sends the trajectory to the inline solve, which
  // rewrites all of its outputs.
  DE_HD static bool book(const EventArgs& e, int64_t n, int64_t i, bool teval_base) {
    const int32_t cnt = e.bcount[i - e.base];
    if (cnt < 0) return false;
    const double* base = e.res + (i - e.base) * e.bmax * kResWidth;
    int32_t pushed = 0;
    double tail = as_f64(kNaNBits), tevals = 0.0;
    for (int32_t k = 0; k < cnt; ++k) {
      const double* r = base + (int64_t)k * kResWidth;
      const double flag = r[kResFlag];
      if (flag == (double)kResPanic) return false;
      if (teval_base && r[kResY + D + 1] != tevals) {
        tail = r[kResY + D];
        tevals = r[kResY + D + 1];
      }
      if (flag == (double)kResPoint) {
        const double t = r[kResT];
        if (!teval_base || tail != tail || fabs(t - tail) > 1e-14) {
          event_store<D>(e, n, i, pushed, t, r + kResY);
          pushed += 1;
          tail = t;
        }
      }
    }
    event_finish<D>(e, n, i, pushed);
    return true;
  }
};

// CrossingSolout (offset y[comp] - threshold) and HyperplaneCrossingSolout
// (signed distance to the plane, normal normalised), with their Newton
// iterations (crossing.rs:253-320, hyperplane.rs:277-340).
template <int D>
struct LaneEvent<D, 2> {
  double value, last, off_new;
  int32_t pushed;
  int32_t k;  // deferred brackets stored

  DE_HD void reset() {
    pushed = 0;
    k = 0;
  }
  DE_HD void start(const EventArgs& e, int64_t i, double t0, const double* y0) {
    value = e.value_traj ? e.value_traj[i] : e.value;
    last = event_g<D>(e, value, t0, y0);
  }
  DE_HD bool pre(const EventArgs& e, double t_curr, const double* y_curr) {
    off_new = event_g<D>(e, value, t_curr, y_curr);
    const double lo = last, off = off_new;
    bool crossing = sign_differs(lo, off);
    if (e.solout == kSoloutHyperplane)
      crossing = crossing || (lo == 0.0 && off != 0.0) || (lo != 0.0 && off == 0.0);
    const bool rec = e.direction == kCrossPositive   ? lo < 0.0 && off >= 0.0
                     : e.direction == kCrossNegative ? lo > 0.0 && off <= 0.0
                                                     : true;
    return crossing && rec;
  }
  // find_crossing_newton; returns false when one of its unwrapped
  // interpolations fails (the reference panics).
  template <class Interp>
  DE_HD bool newton(const EventArgs& e, double tl, double tu, double ol, double ou,
                    Interp& interp, bool* found, double* t_out) const {
    const bool hyper = e.solout == kSoloutHyperplane;
    const double eps = 2.220446049250313080847e-16;
    const double tolerance = eps * 100.0;
    double t = tl - ol * (tu - tl) / (ou - ol);
    double yt[D];
    for (int it = 0; it < 10; ++it) {
      if (!interp(t, yt)) return false;
      const double offset = event_g<D>(e, value, t, yt);
      if (fabs(offset) < tolerance) {
        *found = true;
        *t_out = t;
        return true;
      }
      const double delta_t = (tu - tl) * 1e-6;
      const double t_plus = t + delta_t;
      double yp[D];
      if (!interp(t_plus, yp)) return false;
      const double derivative = (event_g<D>(e, value, t_plus, yp) - offset) / delta_t;
      if (fabs(derivative) < (hyper ? eps : eps * 10.0)) break;
      const double t_new = t - offset / derivative;
      if (t_new < tl || t_new > tu) {
        t = (tl + tu) / 2.0;
      } else {
        if (!hyper && fabs(t_new - t) < tolerance * 0.1) {
          t = t_new;
          break;
        }
        t = t_new;
      }
    }
    if (!interp(t, yt)) return false;
    *found = fabs(event_g<D>(e, value, t, yt)) < tolerance * 10.0;
    *t_out = t;
    return true;
  }
  template <class Interp>
  DE_HD int32_t post(const EventArgs& e, int64_t n, int64_t i, double, bool need,
                     double t_curr, double t_prev, double&, Interp& interp) {
    int32_t st = -1;
    if (need) {
      const double lo = last, off = off_new;
      bool found = false;
      double tc = 0.0;
      if (!newton(e, t_prev, t_curr, lo, off, interp, &found, &tc)) {
        st = kOutOfBounds;
      } else {
        if (!found) {
          const double frac = -lo / (off - lo);
          tc = t_prev + frac * (t_curr - t_prev);
        }
        double yc[D];
        if (!interp(tc, yc)) {
          st = kOutOfBounds;
        } else {
          event_store<D>(e, n, i, pushed, tc, yc);
          pushed += 1;
        }
      }
    }
    return st;
  }
  DE_HD void commit(bool take) { last = take ? off_new : last; }
  DE_HD void finish(const EventArgs& e, int64_t n, int64_t i) const {
    event_finish<D>(e, n, i, pushed);
  }
  DE_HD double* defer(const EventArgs& e, int64_t i, bool need, double t0, double h,
                      double tail, int32_t tevals, bool* overflow) {
    return event_defer(e, i, k, need, t0, h, last, off_new, tail, tevals, overflow);
  }
  template <class Interp>
  DE_HD static void root(const EventArgs& e, int64_t i, const double* ent, double* res,
                         double t_prev, double t_curr, double, bool, Interp& interp) {
    LaneEvent<D, 2> s;
    s.value = e.value_traj ? e.value_traj[i] : e.value;
    const double lo = ent[kEntryG0], off = ent[kEntryG1];
    bool found = false;
    double tc = 0.0, yc[D];
    int flag = kResPanic;
    if (s.newton(e, t_prev, t_curr, lo, off, interp, &found, &tc)) {
      if (!found) {
        const double frac = -lo / (off - lo);
        tc = t_prev + frac * (t_curr - t_prev);
      }
      if (interp(tc, yc)) flag = kResPoint;
    }
    event_result<D>(res, flag, tc, yc, 0.0, 0.0);
  }
  DE_HD static bool book(const EventArgs& e, int64_t n, int64_t i, bool) {
    const int32_t cnt = e.bcount[i - e.base];
    if (cnt < 0) return false;
    const double* base = e.res + (i - e.base) * e.bmax * kResWidth;
    int32_t pushed = 0;
    for (int32_t k = 0; k < cnt; ++k) {
      const double* r = base + (int64_t)k * kResWidth;
      if (r[kResFlag] == (double)kResPanic) return false;
      event_store<D>(e, n, i, pushed, r[kResT], r + kResY);
      pushed += 1;
    }
    event_finish<D>(e, n, i, pushed);
    return true;
  }
};

}  // namespace de
