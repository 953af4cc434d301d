// Lane-per-trajectory step of the embedded explicit Runge-Kutta family:
// ExplicitRungeKutta<Ordinary, Adaptive, ..> with the rkf45, cash_karp and
// Verner tableaux (src/methods/erk/adaptive/ordinary.rs:16-292), one
// trajectory per lane, the tableau a compile-time constant (de_erk_tables.h).
//
// Zero coefficients.  The error estimate is an infinity norm folded with
// max(), which ignores a NaN, so a stage that overflows is not necessarily
// rejected and every NaN the reference produces must be reproduced.  A term
// (0 h) k_j adds +-0 when k_j and h are finite (which changes nothing but,
// possibly, the sign of an exact zero) and NaN otherwise.  The lanes skip the
// zero terms and keep, per component, a bit mask of the stages whose value is
// not finite: a combination whose skipped terms include such a stage (or that
// has zero terms while h is not finite) is NaN, as in the reference.  Results
// therefore equal the reference's as IEEE values; only the sign of a zero
// result can differ.
//
// Registers: k[S][D] per attempt; with dense output and a tableau with
// interpolation rows, k[I][D] (the extra stages are evaluated only on
// accepted steps that reach a t_eval point; the evaluation count still adds
// them on every accepted step, as the reference evaluates them always).
#pragma once

#include <utility>

#include "de_erk_tables.h"
#include "de_lane.h"

namespace de {

// Bit j set where a row's coefficient j is zero (j < n).
DE_HD constexpr uint32_t erk_zero_mask(const double* row, int n) {
  uint32_t m = 0;
  for (int j = 0; j < n; ++j)
    if (row[j] == 0.0) m |= 1u << j;
  return m;
}

// Bit r set where interpolation row r is zero in every power.
template <class Tab>
DE_HD constexpr uint32_t erk_zero_bi_rows() {
  uint32_t m = 0;
  for (int r = 0; r < Tab::I; ++r) {
    bool z = true;
    for (int j = 0; j < Tab::O; ++j) z = z && Tab::BI[r][j] == 0.0;
    if (z) m |= 1u << r;
  }
  return m;
}

// Stages that some combination takes with a zero coefficient: rows 1 .. rows-1
// of a, then b and bh (and the interpolation rows that are zero throughout,
// when the dense output is built).  Only these need a non-finite record.
template <class Tab>
DE_HD constexpr uint32_t erk_zero_stages(int rows, bool dense) {
  uint32_t m = erk_zero_mask(Tab::B, Tab::S) | erk_zero_mask(Tab::BH, Tab::S);
  for (int r = 1; r < rows; ++r) m |= erk_zero_mask(Tab::A[r], r);
  if (dense) m |= erk_zero_bi_rows<Tab>();
  return m;
}

// EV: 0 none, 1 EventWrappedSolout, 2 crossing / hyperplane solouts (de_event.h).
template <class F, class Tab, bool DENSE, bool VTOL, int EV = 0>
struct ErkLane {
  static constexpr int D = F::kDim;
  static constexpr int S = Tab::S, I = Tab::I, O = Tab::O;
  static constexpr bool kBi = Tab::kBi, kFsal = Tab::kFsal;
  static constexpr int kEv = EV;
  // The interpolant is built when a t_eval point or an event bracket needs it.
  static constexpr bool kInterp = DENSE || EV != 0;
  static constexpr int KN = (kInterp && kBi) ? I : S;  // stages held in registers
  static constexpr uint32_t kZeroStages = erk_zero_stages<Tab>(KN, kInterp && kBi);
  // Deferred event bracket: the meta fields (kEntryH holds h), then y_prev and
  // the S stages of the step (the root pass builds the I - S dense stages from
  // them, with the step's own operations), or y_prev, dydt_prev, y, dydt
  // (cubic Hermite).
  static constexpr int kEntryWidth = EV == 0 ? 0 : kEntryMeta + (kBi ? (1 + S) * D : 4 * D);
  // Occupancy hint of the persistent kernel: three waves per SIMD (168 VGPRs);
  // two (256 VGPRs) for the lanes of more than 36 doubles of interpolation
  // stages, which spill at three.
  static constexpr int kWavesPerEu = (KN == S || KN * D <= 36) ? 3 : 2;
  static constexpr bool kLiveStep = true;  // step(a, i, live_in)
  // Evaluations per accepted step beyond the S - 1 of every attempt
  // (ordinary.rs:151-181): f(t_new, y_new) unless FSAL, the dense stages.
  static constexpr int kAccEvals = (kFsal ? 0 : 1) + (kBi ? I - S : 0);

  F f;
  double y[D], k0[D];
  double t, h, h_prev, tf, dir;
  double tail;
  LaneTol<D, VTOL> tol;
  LaneEvent<D, EV> ev;
  int32_t acc, rej, base_ev, stiff, next_eval, pushed;
  bool rejected_last;

  DE_COLD static void write_back(const OdeArgs& a, int64_t i, int32_t st, ErkLane L) {
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
    a.evals[i] = st == kBadInput ? 0
                                 : (int64_t)L.base_ev + (int64_t)(S - 1) * (L.acc + L.rej) +
                                       (int64_t)kAccEvals * L.acc;
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

  // The interpolant of a deferred entry, as the step's: the dense stages are
  // built here once per bracket (same inputs, same operations, same bits as
  // the integrator's), so the integrator does not build them on every step
  // where some lane of its wave brackets.
  DE_HD static bool teval_base(const OdeArgs& a) { return EV == 1 && DENSE && a.n_eval > 0; }
  DE_HD static void entry_root(const OdeArgs& a, int64_t i, const double* ent, double* res) {
    double meta[kEntryMeta];
#pragma unroll
    for (int m = 0; m < kEntryMeta; ++m) meta[m] = ent[m];
    const double tp = meta[kEntryT0], hp = meta[kEntryH];
    const double t_new = tp + hp;
    const double* d = ent + kEntryMeta;
    if constexpr (kBi) {
      double yp[D], k[I][D];
      uint32_t bad[D];
      const uint32_t hbad = isfinite(hp) ? 0u : ~0u;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        yp[c] = d[c];
        bad[c] = hbad;
      }
#pragma unroll
      for (int r = 0; r < S; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
          k[r][c] = d[D + r * D + c];
          if ((kZeroStages >> r) & 1u) bad[c] |= isfinite(k[r][c]) ? 0u : (1u << r);
        }
      ErkLane L;
      L.f.load(a.params, a.n, i);
      L.template stages<S, I>(k, yp, tp, hp, bad);
      const bool fin = stages_finite(k);
      auto interp = [&](double te, double* out) -> bool {
        if (te < tp || te > t_new) return false;
        horner_interp(yp, k, tp, hp, te, fin, out);
        return true;
      };
      LaneEvent<D, EV>::root(a.ev, i, meta, res, tp, t_new, signum(hp), teval_base(a), interp);
    } else {
      double loc[4 * D];
#pragma unroll
      for (int f = 0; f < 4 * D; ++f) loc[f] = d[f];
      auto interp = [&](double te, double* out) -> bool {
        if (te < tp || te > t_new) return false;
        hermite_interp(loc, loc + D, loc + 2 * D, loc + 3 * D, tp, t_new, te, out);
        return true;
      };
      LaneEvent<D, EV>::root(a.ev, i, meta, res, tp, t_new, signum(hp), teval_base(a), interp);
    }
  }
  DE_HD static bool entry_book(const OdeArgs& a, int64_t i) {
    return LaneEvent<D, EV>::book(a.ev, a.n, i, teval_base(a));
  }

  DE_COLD static ErkLane start(const OdeArgs& a, int64_t i, bool* ok) {
    ErkLane L;
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
      h0 = lane_h_init<F, LaneTol<D, VTOL>, O>(f, t0, tf, y, tl, a.h_min, a.h_max);
    }
    return fabs(tf - t0) / fabs(h0);
  }

  // ordinary.rs:16-61: the computed h0 costs InitialStepSize's two
  // evaluations plus two more, then f(t0, y0).
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
    double s = signum(tf - t0);
    dir = s;
    bool ok = (s == 1.0 || s == -1.0);
    double h0 = a.h0;
    base_ev = 1;
    if (ok && h0 == 0.0) {
      h0 = lane_h_init<F, LaneTol<D, VTOL>, O>(f, t0, tf, y, tol, a.h_min, a.h_max);
      base_ev = 5;
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
    rejected_last = false;
    tail = as_f64(kNaNBits);
    if (DENSE && EV == 0 && a.pts_mode != kPointsNone)
      points_start<D>(a, i, pushed, tail, t, y);
    else if (DENSE && EV != 2 && a.n_eval > 0)
      teval_emit<D>(a, i, dir, next_eval, pushed, t, t, y,
                    [](double, double*) { return false; }, tail);
    else
      tail = t;
    if constexpr (EV != 0) {
      ev.reset();
      ev.start(a.ev, i, t, y);
    }
    return true;
  }

  // Every coefficient is read into a constexpr local at a compile-time index
  // (index-sequence folds): the tableau's static arrays are device globals,
  // whose loads the compiler does not fold, and a coefficient loaded at run
  // time would hold scalar registers and turn the zero tests into selects.

  // ys += (a[R][J] h) k[J] for a nonzero a[R][J].
  template <int R, int J>
  DE_HD static void stage_term(const double (*k)[D], double hh, double* ys) {
    constexpr double a = Tab::A[R][J];
    if constexpr (a != 0.0) {
      const double ah = a * hh;
#pragma unroll
      for (int c = 0; c < D; ++c) ys[c] = ys[c] + ah * k[J][c];
    }
  }
  template <int R, int... J>
  DE_HD static void stage_terms(std::integer_sequence<int, J...>, const double (*k)[D], double hh,
                                double* ys) {
    (stage_term<R, J>(k, hh, ys), ...);
  }

  // y + sum_{j<R} (a[R][j] h) k[j] over the nonzero a[R][j], in j order
  // (ordinary.rs:98-106); NaN where a skipped term would be (bad: the
  // per-component masks of non-finite stages, all ones when h is not finite).
  template <int R>
  DE_HD static void stage_input(const double* yy, const double (*k)[D], double hh,
                                const uint32_t* bad, double* ys) {
#pragma unroll
    for (int c = 0; c < D; ++c) ys[c] = yy[c];
    stage_terms<R>(std::make_integer_sequence<int, R>{}, k, hh, ys);
    constexpr uint32_t z = erk_zero_mask(Tab::A[R], R);
    if constexpr (z != 0) {
      uint32_t any = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) any |= bad[c];
      if (__builtin_expect((any & z) != 0, 0)) {
#pragma unroll
        for (int c = 0; c < D; ++c) ys[c] = (bad[c] & z) ? as_f64(kNaNBits) : ys[c];
      }
    }
  }

  template <int R, int END>
  DE_HD void stages(double (*k)[D], const double* yy, double tt, double hh, uint32_t* bad) const {
    if constexpr (R < END) {
      constexpr double cr = Tab::C[R];
      double ys[D];
      stage_input<R>(yy, k, hh, bad, ys);
      f.diff(tt + cr * hh, ys, k[R]);
      if constexpr ((kZeroStages >> R) & 1u) {
#pragma unroll
        for (int c = 0; c < D; ++c) bad[c] |= isfinite(k[R][c]) ? 0u : (1u << R);
      }
      stages<R + 1, END>(k, yy, tt, hh, bad);
    }
  }

  // v += (w[J] h) k[J][c] for a nonzero weight (SEL 0: b, 1: bh).
  template <int SEL, int J>
  DE_HD static void weight_term(const double (*k)[D], double hh, int c, double& v) {
    constexpr double w = SEL ? Tab::BH[J] : Tab::B[J];
    if constexpr (w != 0.0) v = v + (w * hh) * k[J][c];
  }
  template <int SEL, int... J>
  DE_HD static void weight_terms(std::integer_sequence<int, J...>, const double (*k)[D], double hh,
                                 int c, double& v) {
    (weight_term<SEL, J>(k, hh, c, v), ...);
  }

  // y + sum_j (w[j] h) k[j] over the nonzero weights (y_high, y_low).
  template <int SEL>
  DE_HD static double weighted(double yc, const double (*k)[D], double hh, uint32_t bad, int c) {
    double v = yc;
    weight_terms<SEL>(std::make_integer_sequence<int, S>{}, k, hh, c, v);
    constexpr uint32_t z = erk_zero_mask(SEL ? Tab::BH : Tab::B, S);
    if constexpr (z != 0) v = (bad & z) ? as_f64(kNaNBits) : v;
    return v;
  }

  // Horner over bi[r][0 .. J], from the highest power down.
  template <int Rr, int J>
  DE_HD static double horner(double q, double s) {
    if constexpr (J < 0) {
      return q;
    } else {
      constexpr double b = Tab::BI[Rr][J];
      return horner<Rr, J - 1>(q * s + b, s);
    }
  }

  // v += (cont_r h_prev) k[r] for interpolation row r (ordinary.rs:256-275);
  // rows zero in every power only when `all`.
  template <int Rr>
  DE_HD static void interp_row(const double (*k)[D], double s, double hp, bool all, double* v) {
    constexpr bool zero = (erk_zero_bi_rows<Tab>() >> Rr) & 1u;
    if (!zero || all) {
      constexpr double top = Tab::BI[Rr][O - 1];
      const double q = horner<Rr, O - 2>(top, s) * s;
      const double w = q * hp;
#pragma unroll
      for (int c = 0; c < D; ++c) v[c] = v[c] + w * k[Rr][c];
    }
  }
  template <int... Rr>
  DE_HD static void interp_rows(std::integer_sequence<int, Rr...>, const double (*k)[D], double s,
                                double hp, bool all, double* v) {
    (interp_row<Rr>(k, s, hp, all, v), ...);
  }

  // Whether every stage that sits in an all-zero interpolation row is finite
  // (then those rows add +-0 and are skipped).
  DE_HD static bool stages_finite(const double (*k)[D]) {
    constexpr uint32_t zr = erk_zero_bi_rows<Tab>();
    bool fin = true;
#pragma unroll
    for (int r = 0; r < I; ++r)
      if ((zr >> r) & 1u) {
#pragma unroll
        for (int c = 0; c < D; ++c) fin = fin && isfinite(k[r][c]);
      }
    return fin;
  }

  // ordinary.rs:252-277: Horner over bi[r][0..O) times s, then
  // y_prev + sum_r (cont_r h_prev) k[r] in r order.  Rows that are zero in
  // every power add +-0 when their stage, s and h_prev are finite and are
  // skipped then (`fin`: those stages are finite).
  DE_HD static void horner_interp(const double* yp, const double (*k)[D], double tp, double hp,
                                  double te, bool fin, double* out) {
    const double s = (te - tp) / hp;
    const bool all = !fin || !isfinite(s) || !isfinite(hp);
    double v[D];
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = yp[c];
    interp_rows(std::make_integer_sequence<int, I>{}, k, s, hp, all, v);
#pragma unroll
    for (int c = 0; c < D; ++c) out[c] = v[c];
  }

  // cubic_hermite_interpolate(t_prev, t, y_prev, y, dydt_prev, dydt)
  // (src/interpolate.rs:40-60), a linear combination from zero.
  DE_HD static void hermite_interp(const double* yp, const double* dp, const double* yc,
                                   const double* dc, double tp, double tc, double te,
                                   double* out) {
    const double hh = tc - tp;
    const double s = (te - tp) / hh;
    const double s2 = s * s;
    const double s3 = s2 * s;
    const double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
    const double h10 = s3 - 2.0 * s2 + s;
    const double h01 = -2.0 * s3 + 3.0 * s2;
    const double h11 = s3 - s2;
    const double w10 = h10 * hh, w11 = h11 * hh;
#pragma unroll
    for (int c = 0; c < D; ++c)
      out[c] = (((0.0 + h00 * yp[c]) + w10 * dp[c]) + h01 * yc[c]) + w11 * dc[c];
  }

  // One pass of the driver loop (solve_ivp.rs:190-263 around ordinary.rs:63-210).
  // Lanes with live_in false (no trajectory) or ending at the pre-step
  // checks run the attempt with its results discarded, so the state updates
  // stay selects on the lane's own registers.
  DE_HD int32_t step(const OdeArgs& a, int64_t i, bool live_in = true) {
    const bool clip = live_in && (t + h - tf) * dir > 0.0;
    const double hn = tf - t;
    const bool done = clip && fabs(hn) < kTenEps;
    h = clip ? hn : h;
    const bool small = fabs(h) < fabs(h_prev) * 1e-14;
    const bool maxed = (int64_t)(acc + rej) >= a.max_steps;
    const int32_t early = done ? kComplete : (small ? kStepSize : kMaxSteps);
    const bool live = live_in && !(done | small | maxed);
    double k[KN][D];
    uint32_t bad[D];
    const uint32_t hbad = isfinite(h) ? 0u : ~0u;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      k[0][c] = k0[c];
      bad[c] = hbad | ((kZeroStages & 1u) && !isfinite(k0[c]) ? 1u : 0u);
    }
    stages<1, S>(k, y, t, h, bad);

    double yh[D], m = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double hi = weighted<0>(y[c], k, h, bad[c], c);
      const double lo = weighted<1>(y[c], k, h, bad[c], c);
      yh[c] = hi;
      const double e = hi + (-1.0) * lo;
      const double sk = tol.at(c) + tol.rt(c) * fmax(fabs(y[c]), fabs(hi));
      m = fmax(m, fabs(e / sk));
    }
    double scale = a.safety * DE_CTRL_POW(m, -(1.0 / (double)O));
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);

    const bool accept = live && m <= 1.0;
    const int32_t stiff_next = accept ? (rejected_last ? 0 : stiff) : stiff + 1;
    const bool stop = live && !accept && (int64_t)stiff_next >= a.max_rejects;
    const double t_new = t + h;
    double dn[D];
    if constexpr (kFsal) {
#pragma unroll
      for (int c = 0; c < D; ++c) dn[c] = k[S - 1][c];
    } else {
      f.diff(t_new, yh, dn);
    }
    int32_t sol = -1;
    if constexpr (kInterp) {
      // The interpolant is built when the next t_eval point is not past the
      // new time (a NaN point is handed to teval_emit as well) or an event
      // is bracketed.
      const bool teval_out = DENSE && EV != 2 && a.n_eval > 0;
      bool tev = false;
      if (teval_out && accept && next_eval < a.n_eval) {
        const double te = a.t_eval[next_eval];
        tev = !(dir > 0.0 ? te > t_new : te < t_new);
      }
      // Without a t_eval output the Solution's last point is the step's end
      // (DefaultSolout pushes it before an event solout looks at the step).
      const bool pts_out = DENSE && EV == 0 && a.pts_mode != kPointsNone;
      if (!teval_out && !pts_out) tail = accept ? t_new : tail;
      // A point output: DefaultSolout pushes the step's end; the dense and even
      // outputs interpolate.
      const bool pts_interp = pts_out && accept && a.pts_mode != kPointsSteps;
      if (pts_out && accept && a.pts_mode == kPointsSteps) points_push<D>(a, i, pushed, t_new, yh);
      bool need = false;
      if constexpr (EV != 0) need = accept && ev.pre(a.ev, t_new, yh);
      if (tev || need || pts_interp) {
        // The dense stages: for a t_eval point, a point output or an inline
        // root search (a deferred bracket stores the S stages and the root
        // pass builds them).
        if constexpr (kBi) {
          bool build = tev || pts_interp;
          if constexpr (EV != 0) build = build || !a.ev.brk;
          if (build) stages<S, I>(k, y, t, h, bad);
        }
        const double tp = t, hp = h;
        bool fin = true;
        if constexpr (kBi) {
          uint32_t anyb = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) anyb |= bad[c];
          fin = (anyb & erk_zero_bi_rows<Tab>()) == 0;
        }
        auto interp = [&](double te, double* out) -> bool {
          if (te < tp || te > t_new) return false;
          if constexpr (kBi)
            horner_interp(y, k, tp, hp, te, fin, out);
          else
            hermite_interp(y, k[0], yh, dn, tp, t_new, te, out);
          return true;
        };
        if (tev && !teval_emit<D>(a, i, dir, next_eval, pushed, t_new, tp, yh, interp, tail))
          sol = kOutOfBounds;
        if (pts_interp && !points_emit<D>(a, i, dir, tf, pushed, tail, t_new, tp, yh, interp))
          sol = kOutOfBounds;
        if constexpr (EV != 0) {
          if (need && sol < 0) {
            if (a.ev.brk) {
              bool overflow = false;
              double* ent = ev.defer(a.ev, i, true, tp, hp, tail, pushed, &overflow);
              if (ent) {
                double* dst = ent + kEntryMeta;
                if constexpr (kBi) {
#pragma unroll
                  for (int c = 0; c < D; ++c) dst[c] = y[c];
#pragma unroll
                  for (int r = 0; r < S; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) dst[D + r * D + c] = k[r][c];
                } else {
#pragma unroll
                  for (int c = 0; c < D; ++c) {
                    dst[c] = y[c];
                    dst[D + c] = k[0][c];
                    dst[2 * D + c] = yh[c];
                    dst[3 * D + c] = dn[c];
                  }
                }
              }
              if (overflow) sol = kDeferOverflow;
            } else {
              sol = ev.post(a.ev, a.n, i, dir, true, t_new, tp, tail, interp);
            }
          }
        }
      }
      if constexpr (EV != 0) ev.commit(accept);
    }
    h_prev = accept ? h : h_prev;
    t = accept ? t_new : t;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      y[c] = accept ? yh[c] : y[c];
      k0[c] = accept ? dn[c] : k0[c];
    }
    scale = accept && rejected_last ? fmin(scale, 1.0) : scale;
    rejected_last = live ? !accept : rejected_last;
    stiff = live ? stiff_next : stiff;
    acc += accept ? 1 : 0;
    rej += live && !accept && !stop ? 1 : 0;
    const int32_t result =
        !live ? (live_in ? early : -1)
              : (stop ? kStiffness
                      : (sol >= 0 ? sol : (accept && fabs(tf - t) <= kTenEps ? kComplete : -1)));
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

}  // namespace de
