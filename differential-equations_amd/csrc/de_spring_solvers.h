// Solvers for ragged damped spring-mass chains (device code).
//
// Same algorithm and operation order as the lane solvers of de_lane.h and
// de_erk_lane.h (solve_ode src/ode/solve_ivp.rs:116-277, DOPRI5
// ordinary.rs:16-336, h_init h_init.rs:45-135, the solouts of src/solout/),
// for a state of runtime length n = 2N.  Two mappings:
//
//   Vec*          one lane per system (VecDopri5, VecDop853, VecFixed for RK4
//                 and Euler, VecErk for the embedded family); vectors live in
//                 an HBM workspace.  The error norm is the reference's
//                 sequential sum, so results are bit-identical to the CPU
//                 restatement.  The t_eval output and the point outputs.
//   GroupDopri5   G lanes per system, C masses per lane held in registers.
//                 Neighbour displacements cross lanes by DPP wavefront shifts
//                 and the norms are fixed butterfly reductions (DPP within a
//                 row, permutes across rows), identical on every lane of the
//                 group.  The butterfly sums in a different order
//                 than the reference's sequential loop (traits.rs:398-409), so
//                 results agree to rounding rather than bit for bit.  DOPRI5
//                 with the t_eval output or a point output, and event solouts.
#pragma once

#include "de_erk_tables.h"
#include "de_lane.h"

namespace de {

struct SpringArgs {
  OdeArgs o;                   // common settings and per-trajectory outputs
  const int32_t* sizes;        // N per system
  const int64_t* state_offset; // n + 1
  const int64_t* param_offset; // n + 1
  double* work;                // one-lane solvers: kVectors vectors per system at
                               // kVectors * state_offset
  const int64_t* pts_ypos;     // point output: start of system i's states in pts_y
  double* ev_work;             // event solouts: one state per system at state_offset
  double* ystate;              // one-lane solvers' working state at state_offset, or
                               // null: y_final (which the result then needs no copy to)
  const int32_t* list;         // GroupDopri5: system indices of this size class
  int64_t list_len;
};

// Shared pieces of the one-lane-per-system solvers: TEvalSolout over a
// system's dense rows (t_eval.rs:83-134) and the variable-count point outputs
// (DefaultSolout default.rs:71-91, DenseSolout dense.rs:81-111, EvenSolout
// even.rs:83-190, the same steps as points_emit in de_lane.h) with an
// interpolation callback, the write-back, and the automatic initial step
// (h_init.rs:45-135) of order ORDER.
//
// interp(te, out) returns false when te is outside the step (every interior
// point of a backward step), where the reference's solouts unwrap the
// interpolation and panic: the solve then ends OutOfBounds.  Otherwise it
// writes components ic0 .. ic1 - 1 of the state at te to out (all of them but
// in an event's root search), unless out is null (a point past the system's
// slots, counted and not kept, or a time event's search).
//
// Event solouts (EventWrappedSolout around the t_eval or default output, or
// CrossingSolout; event.rs:320-523, crossing.rs:181-320) search their roots
// inline with the lane code of de_event.h, on the event component projected
// to a one-component state.  Event point k of system i: time
// t_out[k * n_systems + i], state at y_out[max_events * state_offset[i] +
// k * n].
struct VecCommon {
  const SpringArgs* A;
  int64_t i;
  int nm, n;  // masses, state length
  const double* p;
  double* y;  // working state: y_final's slot, or ystate's
  double t, h, tf, dir;
  double last;  // EvenSolout::last_output_t
  int32_t next_eval, pushed;
  int ic0, ic1;                  // components the interpolation callbacks write
  double ev_value, ev_last, ev_new;
  double tail;                   // time of the Solution's last point (NaN: none)
  int32_t ev_count, ev_pushed;

  DE_HD void diff(double, const double* yy, double* dy) const { spring_chain_diff(nm, p, yy, dy); }

  DE_HD void write(int32_t st, int64_t acc, int64_t rej, int64_t ev) {
    const OdeArgs& a = A->o;
    if (A->ystate) {
      double* yf = a.y_final + A->state_offset[i];
      for (int c = 0; c < n; ++c) yf[c] = y[c];
    }
    a.t_final[i] = t;
    a.accepted[i] = acc;
    a.rejected[i] = rej;
    a.evals[i] = ev;
    a.status[i] = st;
    if (a.n_dense) a.n_dense[i] = pushed;
    if (a.pts_mode != kPointsNone && a.n_pts) a.n_pts[i] = pushed;
    if (a.n_eval > 0) {  // points not reached (all of them on BadInput)
      double* dbase = a.dense + (int64_t)a.n_eval * A->state_offset[i];
      for (int64_t e = (int64_t)next_eval * n; e < (int64_t)a.n_eval * n; ++e)
        dbase[e] = as_f64(kNaNBits);
    }
    if (a.ev.solout != kSoloutDefault) {  // event count, NaN past the last point
      const EventArgs& e = a.ev;
      if (e.n_out) e.n_out[i] = ev_pushed;
      double* yb = e.y_out + (int64_t)e.max_events * A->state_offset[i];
      for (int32_t k = ev_pushed; k < e.max_events; ++k) {
        e.t_out[(int64_t)k * a.n + i] = as_f64(kNaNBits);
        for (int c = 0; c < n; ++c) yb[(int64_t)k * n + c] = as_f64(kNaNBits);
      }
    }
  }

  // ---- event solouts
  DE_HD bool events_on() const { return A->o.ev.solout != kSoloutDefault; }
  DE_HD double ev_g(double tt, const double* yy) const {
    const EventArgs& e = A->o.ev;
    return e.fn == kEventTime ? tt - ev_value : yy[e.comp] - ev_value;
  }
  // The solout call after init records g(t0, y0) (event.rs:346-355,
  // crossing.rs:240-246).
  DE_HD void ev_start() {
    if (!events_on()) return;
    const EventArgs& e = A->o.ev;
    ev_value = e.value_traj ? e.value_traj[i] : e.value;
    ev_last = ev_g(t, y);
  }
  // g at an accepted step's end; true when the step brackets a root to search
  // in the configured direction (LaneEvent::pre of both solouts).
  DE_HD bool ev_pre(double t_curr, const double* y_curr) {
    if (!events_on()) return false;
    const EventArgs& e = A->o.ev;
    ev_new = ev_g(t_curr, y_curr);
    const double lo = ev_last, g = ev_new;
    const bool change = sign_differs(lo, g);
    if (e.solout == kSoloutEvent) {
      const bool ok = e.direction == kCrossPositive   ? change && lo < 0.0 && g >= 0.0
                      : e.direction == kCrossNegative ? change && lo > 0.0 && g <= 0.0
                                                      : change;
      return ok && lo * g <= 0.0;
    }
    const bool rec = e.direction == kCrossPositive   ? lo < 0.0 && g >= 0.0
                     : e.direction == kCrossNegative ? lo > 0.0 && g <= 0.0
                                                     : true;
    return change && rec;
  }
  // The state slot of the next event point, or null past max_events.
  DE_HD double* ev_dst() const {
    const EventArgs& e = A->o.ev;
    return ev_pushed < e.max_events
               ? e.y_out + (int64_t)e.max_events * A->state_offset[i] + (int64_t)ev_pushed * n
               : nullptr;
  }
  DE_HD void ev_push(double tt) {
    const EventArgs& e = A->o.ev;
    if (ev_pushed < e.max_events) e.t_out[(int64_t)ev_pushed * A->o.n + i] = tt;
    ev_pushed += 1;
  }
  // The root search of a bracketed step [t_prev, t_curr] and its push
  // (detect_event event.rs:335-393, CrossingSolout::solout crossing.rs:
  // 186-248); -1 to continue, else the solve's final status.  The Solution's
  // last time is the step's end under the default output, else `tail`.
  template <class Interp>
  DE_HD int32_t ev_post(bool need, double t_curr, double t_prev, Interp& interp) {
    if (!need) return -1;
    const EventArgs& e = A->o.ev;
    EventArgs e1 = e;
    e1.comp = 0;  // g of the projected one-component state
    double* ew = A->ev_work + A->state_offset[i];
    auto interp1 = [&](double tt, double* o) -> bool {
      if (e.fn == kEventTime) {
        o[0] = 0.0;
        return interp(tt, nullptr);
      }
      ic0 = e.comp;
      ic1 = e.comp + 1;
      const bool ok = interp(tt, ew);
      ic0 = 0;
      ic1 = n;
      o[0] = ok ? ew[e.comp] : 0.0;
      return ok;
    };
    int32_t st = -1;
    if (e.solout == kSoloutEvent) {
      double a0 = t_prev, b0 = t_curr, fa = ev_last, fb = ev_new;
      if ((dir > 0.0 && a0 > b0) || (dir < 0.0 && a0 < b0)) {
        double x = a0;
        a0 = b0;
        b0 = x;
        x = fa;
        fa = fb;
        fb = x;
      }
      const double v = ev_value;
      auto g = [&](double tt, const double* yy) { return event_g<1>(e1, v, tt, yy); };
      double t_event;
      if (brent_dekker<1>(a0, b0, fa, fb, interp1, g, &t_event)) {
        if (!interp(t_event, ev_dst())) {
          st = kOutOfBounds;
        } else {
          const double tl = A->o.n_eval > 0 ? tail : t_curr;
          if (tl != tl || fabs(t_event - tl) > 1e-14) {
            ev_push(t_event);
            tail = t_event;
          }
          ev_count += 1;
          if (e.terminate > 0 && ev_count >= e.terminate) st = kInterrupted;
        }
      }
    } else {
      LaneEvent<1, 2> cs;
      cs.value = ev_value;
      bool found = false;
      double tc = 0.0;
      if (!cs.newton(e1, t_prev, t_curr, ev_last, ev_new, interp1, &found, &tc)) {
        st = kOutOfBounds;
      } else {
        if (!found) {
          const double frac = -ev_last / (ev_new - ev_last);
          tc = t_prev + frac * (t_curr - t_prev);
        }
        if (!interp(tc, ev_dst()))
          st = kOutOfBounds;
        else
          ev_push(tc);
      }
    }
    return st;
  }
  // After an accepted step's solout call: g at its end is the next last value.
  DE_HD void ev_commit() { ev_last = ev_new; }

  DE_HD bool points_on() const { return A->o.pts_mode != kPointsNone; }
  // An accepted step writes output: t_eval rows or a point output.
  DE_HD bool output_on() const { return A->o.n_eval > 0 || points_on(); }

  // The state slot of the next point, or null past this system's slots.
  DE_HD double* pts_dst() const {
    const OdeArgs& a = A->o;
    return a.pts_offset[i] + pushed < a.pts_offset[i + 1]
               ? a.pts_y + A->pts_ypos[i] + (int64_t)pushed * n
               : nullptr;
  }
  // Counts the next point and keeps its time while the slots last.
  DE_HD void pts_commit(double tt) {
    const OdeArgs& a = A->o;
    const int64_t s = a.pts_offset[i] + pushed;
    if (s < a.pts_offset[i + 1]) a.pts_t[s] = tt;
    pushed += 1;
  }
  DE_HD void pts_push(double tt, const double* yy) {
    double* d = pts_dst();
    if (d)
      for (int c = 0; c < n; ++c) d[c] = yy[c];
    pts_commit(tt);
  }
  template <class Interp>
  DE_HD bool pts_interp(double ti, Interp& interp) {
    if (!interp(ti, pts_dst())) return false;
    pts_commit(ti);
    return true;
  }

  // The solout call after init (t_curr = t_prev = t0): TEvalSolout writes the
  // points equal to t0, every point output pushes (t0, y0).
  DE_HD void output_start() {
    const OdeArgs& a = A->o;
    if (points_on()) {
      pts_push(t, y);
      last = t;
    } else if (a.n_eval > 0) {
      emit(t, t, y, false, [](double, double*) { return true; });
    }
  }

  // The solout call of an accepted step [t_prev, t_curr] ending in y_curr.
  template <class Interp>
  DE_HD bool output(double t_curr, double t_prev, const double* y_curr, Interp interp) {
    return points_on() ? points_step(t_curr, t_prev, y_curr, interp)
                       : emit(t_curr, t_prev, y_curr, true, interp);
  }

  template <class Interp>
  DE_HD bool points_step(double t_curr, double t_prev, const double* y_curr, Interp& interp) {
    const OdeArgs& a = A->o;
    if (a.pts_mode == kPointsSteps) {
      pts_push(t_curr, y_curr);
      return true;
    }
    if (a.pts_mode == kPointsDense) {
      if (t_prev != t_curr) {
        for (int32_t k = 1; k < a.pts_n; ++k) {
          const double h_old = t_curr - t_prev;
          const double ti = t_prev + (double)k * h_old / (double)a.pts_n;
          if (!pts_interp(ti, interp)) return false;
        }
      }
      pts_push(t_curr, y_curr);
      return true;
    }
    const double dt = a.pts_dt;
    const double tol = fabs(dt) * 1e-12 + 2.220446049250313080847e-16 * 10.0;
    double ti = last + dt * dir;
    while ((dir > 0.0 && ti <= t_curr) || (dir < 0.0 && ti >= t_curr)) {
      if ((dir > 0.0 && ti >= t_prev && ti <= t_curr) ||
          (dir < 0.0 && ti <= t_prev && ti >= t_curr)) {
        if (!(fabs(ti - last) <= tol)) {
          if (!pts_interp(ti, interp)) return false;
          last = ti;
        }
      }
      ti += dt * dir;
    }
    if (t_curr == tf) {
      if (fabs(last - tf) <= tol) {
        pushed -= 1;  // Solution::pop, then the exact final point
        pts_push(tf, y_curr);
        last = tf;
      } else if (last != tf) {
        pts_push(tf, y_curr);
        last = tf;
      }
    }
    return true;
  }

  // TEvalSolout: with have_cont false only the step's end is written (the
  // start of a solve).
  template <class Interp>
  DE_HD bool emit(double t_curr, double t_prev, const double* ycur, bool have_cont,
                  Interp interp) {
    const OdeArgs& a = A->o;
    int32_t idx = next_eval;
    double* dbase = a.dense + (int64_t)a.n_eval * A->state_offset[i];
    bool ok = true;
    while (idx < a.n_eval) {
      double te = a.t_eval[idx];
      bool in_range = dir > 0.0 ? ((te == t_prev && idx == 0) || (te > t_prev && te <= t_curr))
                                : ((te == t_prev && idx == 0) || (te < t_prev && te >= t_curr));
      if (in_range) {
        double* out = dbase + (int64_t)idx * n;
        if (te == t_curr || !have_cont) {
          for (int c = 0; c < n; ++c) out[c] = ycur[c];
        } else if (!interp(te, out)) {
          ok = false;
          break;
        }
        pushed += 1;
        tail = te;
        idx += 1;
      } else {
        if ((dir > 0.0 && te > t_curr) || (dir < 0.0 && te < t_curr)) break;
        double* out = dbase + (int64_t)idx * n;  // passed over: not produced
        for (int c = 0; c < n; ++c) out[c] = as_f64(kNaNBits);
        idx += 1;
      }
    }
    next_eval = idx;
    return ok;
  }

  // Common prologue: the system, its state and time span.  w: this system's
  // workspace.
  DE_HD double setup(const SpringArgs* args, int64_t sys) {
    A = args;
    i = sys;
    const OdeArgs& a = A->o;
    nm = A->sizes[i];
    n = 2 * nm;
    p = a.params + A->param_offset[i];
    const int64_t so = A->state_offset[i];
    y = (A->ystate ? A->ystate : a.y_final) + so;
    const double* y0 = a.y0 + so;
    for (int c = 0; c < n; ++c) y[c] = y0[c];
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    tf = a.tf ? a.tf[i] : a.tfs;
    t = t0;
    pushed = next_eval = 0;
    ic0 = 0;
    ic1 = n;
    ev_count = ev_pushed = 0;
    tail = as_f64(kNaNBits);
    dir = signum(tf - t0);
    return t0;
  }

  // h_init with f0, f1, sk, y1 in workspace vectors.
  template <int ORDER>
  DE_HD double h_init(double t0, double* f0, double* f1, double* sk, double* y1) const {
    const OdeArgs& a = A->o;
    const int64_t so = A->state_offset[i];
    const double rtol = a.rtol_traj ? a.rtol_traj[i] : a.rtol;
    const double atol = a.atol_traj ? a.atol_traj[i] : a.atol;
    const double* rtv = a.rtol_comp ? a.rtol_comp + so : nullptr;
    const double* atv = a.atol_comp ? a.atol_comp + so : nullptr;
    diff(t0, y, f0);
    double dnf = 0.0, dny = 0.0;
    for (int c = 0; c < n; ++c) {
      sk[c] = (atv ? atv[c] : atol) + (rtv ? rtv[c] : rtol) * fabs(y[c]);
      double u = f0[c] / sk[c];
      dnf += u * u;
      double w2 = y[c] / sk[c];
      dny += w2 * w2;
    }
    double hh = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    hh = fmin(hh, a.h_max);
    hh *= dir;
    for (int c = 0; c < n; ++c) y1[c] = y[c] + hh * f0[c];
    diff(t0 + hh, y1, f1);
    double der2 = 0.0;
    for (int c = 0; c < n; ++c) {
      double d = (f1[c] - f0[c]) / sk[c];
      der2 += d * d;
    }
    der2 = sqrt(der2) / fabs(hh);
    double der12 = fmax(sqrt(dnf), der2);
    double h1 = der12 <= 1.0e-15 ? fabs(hh) * 1.0e-3
                                 : DE_CTRL_POW(0.01 / der12, 1.0 / (double)ORDER);
    double interval = fabs(tf - t0);
    hh = fmin(fmax(fmin(fmin(fabs(hh) * 100.0, h1), a.h_max), a.h_min), interval);
    return hh * dir;
  }
};

// ------------------------------------------------------------- VecDopri5 --
// Workspace: k[0..7), ys, ynew.
struct VecDopri5 : VecCommon {
  static constexpr int kVectors = 9;
  double *k[7], *ys, *ynew;
  double h_prev, rtol, atol;
  const double *rtv, *atv;  // per-component tolerances of this system, or null
  int64_t acc, rej, ev, steps;
  int32_t stiff, nonstiff, countdown;
  bool rejected_last;

  DE_HD double rt(int c) const { return rtv ? rtv[c] : rtol; }
  DE_HD double at(int c) const { return atv ? atv[c] : atol; }
  DE_HD void finish(int32_t st) { write(st, acc, rej, ev); }

  DE_HD bool init(const SpringArgs* args, int64_t sys) {
    const double t0 = setup(args, sys);
    const OdeArgs& a = A->o;
    const int64_t so = A->state_offset[i];
    double* w = A->work + kVectors * so;
    for (int v = 0; v < 7; ++v) k[v] = w + (int64_t)v * n;
    ys = w + 7 * (int64_t)n;
    ynew = w + 8 * (int64_t)n;
    rtol = a.rtol_traj ? a.rtol_traj[i] : a.rtol;
    atol = a.atol_traj ? a.atol_traj[i] : a.atol;
    rtv = a.rtol_comp ? a.rtol_comp + so : nullptr;
    atv = a.atol_comp ? a.atol_comp + so : nullptr;
    acc = rej = ev = steps = 0;
    bool ok = dir == 1.0 || dir == -1.0;
    double h0 = a.h0;
    if (ok && h0 == 0.0) {
      h0 = h_init<5>(t0, k[0], k[1], ys, ynew);
      ev += 2;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      ev = 0;
      finish(kBadInput);
      return false;
    }
    h = h0;
    diff(t, y, k[0]);
    ev += 1;
    h_prev = 0.0;
    stiff = nonstiff = 0;
    countdown = 100;
    rejected_last = false;
    output_start();
    ev_start();
    return true;
  }

  DE_HD bool advance() {
    const OdeArgs& a = A->o;
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) {
        finish(kComplete);
        return true;
      }
      h = hn;
    }
    if (fabs(h) < fabs(h_prev) * 1e-14) {
      finish(kStepSize);
      return true;
    }
    if (steps >= a.max_steps) {
      finish(kMaxSteps);
      return true;
    }
    steps += 1;
    // The stiffness test runs on accepted attempts whose number is a multiple
    // of 100 (ordinary.rs:167); the countdown restarts on every such attempt.
    countdown -= 1;
    const bool stiff_check = countdown == 0;
    if (stiff_check) countdown = 100;
    const double a10h = dp::a10 * h;
    for (int c = 0; c < n; ++c) ys[c] = y[c] + a10h * k[0][c];
    diff(t + dp::c1 * h, ys, k[1]);
    const double a20h = dp::a20 * h, a21h = dp::a21 * h;
    for (int c = 0; c < n; ++c) ys[c] = (y[c] + a20h * k[0][c]) + a21h * k[1][c];
    diff(t + dp::c2 * h, ys, k[2]);
    const double a30h = dp::a30 * h, a31h = dp::a31 * h, a32h = dp::a32 * h;
    for (int c = 0; c < n; ++c)
      ys[c] = ((y[c] + a30h * k[0][c]) + a31h * k[1][c]) + a32h * k[2][c];
    diff(t + dp::c3 * h, ys, k[3]);
    const double a40h = dp::a40 * h, a41h = dp::a41 * h, a42h = dp::a42 * h,
                 a43h = dp::a43 * h;
    for (int c = 0; c < n; ++c)
      ys[c] = (((y[c] + a40h * k[0][c]) + a41h * k[1][c]) + a42h * k[2][c]) + a43h * k[3][c];
    diff(t + dp::c4 * h, ys, k[4]);
    const double a50h = dp::a50 * h, a51h = dp::a51 * h, a52h = dp::a52 * h,
                 a53h = dp::a53 * h, a54h = dp::a54 * h;
    for (int c = 0; c < n; ++c)
      ys[c] = ((((y[c] + a50h * k[0][c]) + a51h * k[1][c]) + a52h * k[2][c]) + a53h * k[3][c]) +
              a54h * k[4][c];
    diff(t + 1.0 * h, ys, k[5]);
    const double a60h = dp::a60 * h, a62h = dp::a62 * h, a63h = dp::a63 * h,
                 a64h = dp::a64 * h, a65h = dp::a65 * h;
    for (int c = 0; c < n; ++c)
      ys[c] = ((((y[c] + a60h * k[0][c]) + a62h * k[2][c]) + a63h * k[3][c]) + a64h * k[4][c]) +
              a65h * k[5][c];
    diff(t + 1.0 * h, ys, k[6]);
    // yseg is kept in k[1] (k1 has a zero weight in b and er and is not used
    // again this attempt); ynew in ynew.
    double err = 0.0;
    for (int c = 0; c < n; ++c) {
      double sg = (((dp::b0 * k[0][c] + dp::b2 * k[2][c]) + dp::b3 * k[3][c]) +
                   dp::b4 * k[4][c]) +
                  dp::b5 * k[5][c];
      double yn = y[c] + h * sg;
      double es = ((((dp::e0 * k[0][c] + dp::e2 * k[2][c]) + dp::e3 * k[3][c]) +
                    dp::e4 * k[4][c]) +
                   dp::e5 * k[5][c]) +
                  dp::e6 * k[6][c];
      double sk = at(c) + rt(c) * fmax(fabs(y[c]), fabs(yn));
      double e = es / sk;
      err = c == 0 ? e * e : err + e * e;
      k[1][c] = sg;
      ynew[c] = yn;
    }
    double deno = err;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * (double)n));
    double scale = a.safety * DE_CTRL_POW(err, -0.2);
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);
    if (err <= 1.0) {
      double t_new = t + h;
      // f(t_new, y_new) goes to k[3] (no longer needed); ys still holds the
      // stage-6 state for the stiffness test.
      double* dn = k[3];
      diff(t_new, ynew, dn);
      if (stiff_check) {
        double stdnum = 0.0, stden = 0.0;
        for (int c = 0; c < n; ++c) {
          double u = k[1][c] - k[6][c];
          stdnum = c == 0 ? u * u : stdnum + u * u;
          double w2 = dn[c] - ys[c];
          stden = c == 0 ? w2 * w2 : stden + w2 * w2;
        }
        if (stden > 0.0) {
          if (stiff_exceeds(h, stdnum, stden)) {
            nonstiff = 0;
            stiff += 1;
            if (stiff == 15) {
              finish(kStiffness);
              return true;
            }
          }
        } else {
          nonstiff += 1;
          if (nonstiff == 6) stiff = 0;
        }
      }
      const bool need = ev_pre(t_new, ynew);
      int32_t sol = -1;
      if (a.pts_mode == kPointsSteps) {
        if (!output(t_new, t, ynew, [](double, double*) { return true; })) sol = kOutOfBounds;
      } else if (output_on() || need) {
        // cont1 = ydiff (k[4]), cont2 = bspl (k[5]), cont3 (k[6]).
        for (int c = 0; c < n; ++c) {
          double ydiff = ynew[c] + (-1.0) * y[c];
          double bspl = h * k[0][c] + (-1.0) * ydiff;
          double c3 = (ydiff + (-h) * dn[c]) + (-1.0) * bspl;
          k[4][c] = ydiff;
          k[5][c] = bspl;
          k[6][c] = c3;
        }
        const double tp = t, hp = h;
        const double *c1 = k[4], *c2 = k[5], *c3 = k[6];
        auto interp = [&](double te, double* out) {
          // ordinary.rs:303-333: factors s1, s, s1 over cont[4] = 0 .. cont[1].
          if (te < tp || te > t_new) return false;
          if (!out) return true;
          const double s = (te - tp) / hp;
          const double s1 = 1.0 - s;
          const double cont4 = 0.0 * hp;
          for (int c = ic0; c < ic1; ++c) {
            double acc_ = cont4 * s1 + c3[c];
            acc_ = acc_ * s + c2[c];
            acc_ = acc_ * s1 + c1[c];
            out[c] = y[c] + s * acc_;
          }
          return true;
        };
        if (output_on() && !output(t_new, tp, ynew, interp)) sol = kOutOfBounds;
        if (sol < 0) sol = ev_post(need, t_new, tp, interp);
      }
      if (events_on()) ev_commit();
      h_prev = h;
      t = t_new;
      for (int c = 0; c < n; ++c) {
        y[c] = ynew[c];
        k[0][c] = dn[c];
      }
      if (rejected_last) {
        rejected_last = false;
        scale = fmin(scale, 1.0);
      }
      acc += 1;
      ev += 7;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
      if (sol >= 0) {
        finish(sol);
        return true;
      }
      if (fabs(tf - t) <= kTenEps) {
        finish(kComplete);
        return true;
      }
    } else {
      rejected_last = true;
      rej += 1;
      ev += 6;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
    }
    return false;
  }
};

// ---------------------------------------------------------------- VecFixed --
// RK4 / Euler (fixed/ordinary.rs, the same steps as FixedLane) on one lane per
// system, vectors in the workspace: k0..k3, ys, ynext, dn.
template <int STAGES>
struct VecFixed : VecCommon {
  static constexpr int kVectors = 7;
  double *k0, *k1, *k2, *k3, *ys, *ynext, *dn;
  int64_t steps, acc, ev;

  DE_HD void finish(int32_t st) { write(st, acc, 0, ev); }

  DE_HD bool init(const SpringArgs* args, int64_t sys) {
    const double t0 = setup(args, sys);
    const OdeArgs& a = A->o;
    double* w = A->work + kVectors * A->state_offset[i];
    k0 = w;
    k1 = w + n;
    k2 = w + 2 * (int64_t)n;
    k3 = w + 3 * (int64_t)n;
    ys = w + 4 * (int64_t)n;
    ynext = w + 5 * (int64_t)n;
    dn = w + 6 * (int64_t)n;
    acc = ev = steps = 0;
    bool ok = dir == 1.0 || dir == -1.0;
    double h0 = a.h0;
    if (h0 == 0.0) h0 = fabs(tf - t0) / 100.0;
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      finish(kBadInput);
      return false;
    }
    h = h0;
    diff(t, y, k0);
    ev = 1;
    output_start();
    ev_start();
    return true;
  }

  DE_HD bool advance() {
    const OdeArgs& a = A->o;
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) {
        finish(kComplete);
        return true;
      }
      h = hn;
    }
    if (steps >= a.max_steps) {
      finish(kMaxSteps);
      return true;
    }
    steps += 1;
    if (STAGES == 4) {
      const double hh = 0.5 * h;
      for (int c = 0; c < n; ++c) ys[c] = y[c] + hh * k0[c];
      diff(t + 0.5 * h, ys, k1);
      for (int c = 0; c < n; ++c) ys[c] = y[c] + hh * k1[c];
      diff(t + 0.5 * h, ys, k2);
      const double h1 = 1.0 * h;
      for (int c = 0; c < n; ++c) ys[c] = y[c] + h1 * k2[c];
      diff(t + 1.0 * h, ys, k3);
      const double b0h = (1.0 / 6.0) * h, b1h = (1.0 / 3.0) * h;
      for (int c = 0; c < n; ++c)
        ynext[c] = (((y[c] + b0h * k0[c]) + b1h * k1[c]) + b1h * k2[c]) + b0h * k3[c];
      ev += 3;
    } else {
      const double b0h = 1.0 * h;
      for (int c = 0; c < n; ++c) ynext[c] = y[c] + b0h * k0[c];
    }
    const double tp = t;
    t += h;
    diff(t, ynext, dn);
    ev += 1;
    acc += 1;
    const bool need = ev_pre(t, ynext);
    int32_t sol = -1;
    if (output_on() || need) {
      const double tc = t;
      // cubic_hermite_interpolate on [t_prev, t] (fixed/ordinary.rs:170-177).
      auto interp = [&](double te, double* out) {
        if (te < tp || te > tc) return false;
        if (!out) return true;
        double hh = tc - tp;
        double s = (te - tp) / hh;
        double s2 = s * s;
        double s3 = s2 * s;
        double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
        double h10 = s3 - 2.0 * s2 + s;
        double h01 = -2.0 * s3 + 3.0 * s2;
        double h11 = s3 - s2;
        double w10 = h10 * hh, w11 = h11 * hh;
        for (int c = ic0; c < ic1; ++c)
          out[c] = ((h00 * y[c] + w10 * k0[c]) + h01 * ynext[c]) + w11 * dn[c];
        return true;
      };
      if (output_on() && !output(t, tp, ynext, interp)) sol = kOutOfBounds;
      if (sol < 0) sol = ev_post(need, t, tp, interp);
    }
    if (events_on()) ev_commit();
    for (int c = 0; c < n; ++c) {
      y[c] = ynext[c];
      k0[c] = dn[c];
    }
    if (sol >= 0) {
      finish(sol);
      return true;
    }
    if (fabs(tf - t) <= kTenEps) {
      finish(kComplete);
      return true;
    }
    return false;
  }
};

// --------------------------------------------------------------- VecDop853 --
// DOP853 (the generic Dormand-Prince step of dormandprince/ordinary.rs:63-270
// with the tableau of src/tableau/dorman_prince.rs:155-380; the same operations
// as Dop853Lane) on one lane per system.  Workspace: k[0..15], ys, ynew, yseg
// and the dense-output coefficients cont[1..7] (cont[0] is y).
struct VecDop853 : VecCommon {
  static constexpr int kVectors = 26;
  double *k[16], *ys, *ynew, *yseg, *cont[8];
  double h_prev, rtol, atol;
  const double *rtv, *atv;
  int64_t acc, rej, ev, steps;
  int32_t stiff, nonstiff, countdown;
  bool rejected_last;

  DE_HD double rt(int c) const { return rtv ? rtv[c] : rtol; }
  DE_HD double at(int c) const { return atv ? atv[c] : atol; }
  DE_HD void finish(int32_t st) { write(st, acc, rej, ev); }

  // ys = y + sum_{j<I} (a[I][j] h) k[j], zero coefficients skipped.
  DE_HD void stage_input(int I, const double* yy, double hh, double* out) const {
    for (int c = 0; c < n; ++c) out[c] = yy[c];
    for (int j = 0; j < I; ++j) {
      if (d8::A[I][j] != 0.0) {
        const double ah = d8::A[I][j] * hh;
        for (int c = 0; c < n; ++c) out[c] = out[c] + ah * k[j][c];
      }
    }
  }

  // sum_j w[j] k[j][c] from the first nonzero weight.
  DE_HD double comb(const double* wts, int nw, int c) const {
    double s = 0.0;
    bool first = true;
    for (int j = 0; j < nw; ++j) {
      if (wts[j] != 0.0) {
        s = first ? wts[j] * k[j][c] : s + wts[j] * k[j][c];
        first = false;
      }
    }
    return s;
  }

  DE_HD bool init(const SpringArgs* args, int64_t sys) {
    const double t0 = setup(args, sys);
    const OdeArgs& a = A->o;
    const int64_t so = A->state_offset[i];
    double* w = A->work + kVectors * so;
    for (int v = 0; v < 16; ++v) k[v] = w + (int64_t)v * n;
    ys = w + 16 * (int64_t)n;
    ynew = w + 17 * (int64_t)n;
    yseg = w + 18 * (int64_t)n;
    cont[0] = y;
    for (int r = 1; r < 8; ++r) cont[r] = w + (int64_t)(18 + r) * n;
    rtol = a.rtol_traj ? a.rtol_traj[i] : a.rtol;
    atol = a.atol_traj ? a.atol_traj[i] : a.atol;
    rtv = a.rtol_comp ? a.rtol_comp + so : nullptr;
    atv = a.atol_comp ? a.atol_comp + so : nullptr;
    acc = rej = ev = steps = 0;
    bool ok = dir == 1.0 || dir == -1.0;
    double h0 = a.h0;
    ev = 1;
    if (ok && h0 == 0.0) {
      h0 = h_init<8>(t0, k[0], k[1], ys, ynew);
      ev = 3;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      ev = 0;
      finish(kBadInput);
      return false;
    }
    h = h0;
    diff(t, y, k[0]);
    h_prev = 0.0;
    stiff = nonstiff = 0;
    countdown = 100;
    rejected_last = false;
    output_start();
    ev_start();
    return true;
  }

  DE_HD bool advance() {
    const OdeArgs& a = A->o;
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) {
        finish(kComplete);
        return true;
      }
      h = hn;
    }
    if (fabs(h) < fabs(h_prev) * 1e-14) {
      finish(kStepSize);
      return true;
    }
    if (steps >= a.max_steps) {
      finish(kMaxSteps);
      return true;
    }
    steps += 1;
    countdown -= 1;
    const bool stiff_check = countdown == 0;
    if (stiff_check) countdown = 100;
    for (int I = 1; I < 12; ++I) {  // ys ends as the stage-11 input
      stage_input(I, y, h, ys);
      diff(t + d8::C[I] * h, ys, k[I]);
    }
    double err = 0.0, err2 = 0.0;
    for (int c = 0; c < n; ++c) {
      yseg[c] = comb(d8::B, 12, c);
      ynew[c] = y[c] + h * yseg[c];
    }
    for (int c = 0; c < n; ++c) {
      const double es = comb(d8::ER, 12, c);
      double e2s = yseg[c];
      for (int j = 0; j < 12; ++j)
        if (d8::BH[j] != 0.0) e2s = e2s + (-d8::BH[j]) * k[j][c];
      const double sk = at(c) + rt(c) * fmax(fabs(y[c]), fabs(ynew[c]));
      const double e = es / sk;
      const double e2 = e2s / sk;
      err = c == 0 ? e * e : err + e * e;
      err2 = c == 0 ? e2 * e2 : err2 + e2 * e2;
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * (double)n));
    double scale = a.safety * DE_CTRL_POW(err, -0.125);
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);
    if (err <= 1.0) {
      const double t_new = t + h;
      double* dn = k[12];  // f(t_new, y_new), the 13th stage of the dense output
      diff(t_new, ynew, dn);
      if (stiff_check) {
        double stdnum = 0.0, stden = 0.0;
        for (int c = 0; c < n; ++c) {
          double u = yseg[c] - k[11][c];
          stdnum = c == 0 ? u * u : stdnum + u * u;
          double v = dn[c] - ys[c];
          stden = c == 0 ? v * v : stden + v * v;
        }
        if (stden > 0.0) {
          if (stiff_exceeds(h, stdnum, stden)) {
            nonstiff = 0;
            stiff += 1;
            if (stiff == 15) {
              finish(kStiffness);
              return true;
            }
          }
        } else {
          nonstiff += 1;
          if (nonstiff == 6) stiff = 0;
        }
      }
      const bool need = ev_pre(t_new, ynew);
      int32_t sol = -1;
      if (a.pts_mode == kPointsSteps) {
        if (!output(t_new, t, ynew, [](double, double*) { return true; })) sol = kOutOfBounds;
      } else if (output_on() || need) {
        // cont[1..3] as DOPRI5's, k[13..15] at c[13..15], cont[r] = h sum_j
        // bi[r][j] k[j] for r = 4..7.
        for (int c = 0; c < n; ++c) {
          double ydiff = ynew[c] + (-1.0) * y[c];
          cont[1][c] = ydiff;
          double bspl = h * k[0][c] + (-1.0) * ydiff;
          cont[2][c] = bspl;
          cont[3][c] = (ydiff + (-h) * dn[c]) + (-1.0) * bspl;
        }
        for (int I = 13; I < 16; ++I) {
          stage_input(I, y, h, ys);
          diff(t + d8::C[I] * h, ys, k[I]);
        }
        for (int r = 4; r < 8; ++r)
          for (int c = 0; c < n; ++c) cont[r][c] = comb(d8::BI[r], 16, c) * h;
        const double tp = t, hp = h;
        auto interp = [&](double te, double* out) {
          if (te < tp || te > t_new) return false;
          if (!out) return true;
          const double s = (te - tp) / hp;
          const double s1 = 1.0 - s;
          for (int c = ic0; c < ic1; ++c) {
            double q = cont[7][c];
            q = q * s1 + cont[6][c];
            q = q * s + cont[5][c];
            q = q * s1 + cont[4][c];
            q = q * s1 + cont[3][c];
            q = q * s + cont[2][c];
            q = q * s1 + cont[1][c];
            out[c] = y[c] + s * q;
          }
          return true;
        };
        if (output_on() && !output(t_new, tp, ynew, interp)) sol = kOutOfBounds;
        if (sol < 0) sol = ev_post(need, t_new, tp, interp);
      }
      if (events_on()) ev_commit();
      h_prev = h;
      t = t_new;
      for (int c = 0; c < n; ++c) {
        y[c] = ynew[c];
        k[0][c] = dn[c];
      }
      if (rejected_last) {
        rejected_last = false;
        scale = fmin(scale, 1.0);
      }
      acc += 1;
      ev += 15;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
      if (sol >= 0) {
        finish(sol);
        return true;
      }
      if (fabs(tf - t) <= kTenEps) {
        finish(kComplete);
        return true;
      }
    } else {
      rejected_last = true;
      rej += 1;
      ev += 11;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
    }
    return false;
  }
};

// ------------------------------------------------------------------ VecErk --
// The embedded family (adaptive/ordinary.rs:16-292, the same operations as
// ErkLane in de_erk_lane.h) on one lane per system: stages k[0..I), ys, yh,
// dn and one vector of per-component masks of the stages whose value is not
// finite (a skipped zero-coefficient term of such a stage is NaN in the
// reference; see de_erk_lane.h).  Every stage is tracked, which changes
// nothing where no zero coefficient meets it.
template <class Tab>
struct VecErk : VecCommon {
  static constexpr int S = Tab::S, I = Tab::I, O = Tab::O;
  static constexpr bool kBi = Tab::kBi, kFsal = Tab::kFsal;
  static constexpr int kVectors = I + 4;
  static constexpr int kAccEvals = (kFsal ? 0 : 1) + (kBi ? I - S : 0);
  double *k[I], *ys, *yh, *dn;
  uint32_t* bad;
  double h_prev, rtol, atol;
  const double *rtv, *atv;
  int64_t acc, rej, base_ev;
  int32_t stiff;
  bool rejected_last;

  DE_HD double rt(int c) const { return rtv ? rtv[c] : rtol; }
  DE_HD double at(int c) const { return atv ? atv[c] : atol; }
  DE_HD void finish(int32_t st) {
    write(st, acc, rej,
          st == kBadInput ? 0 : base_ev + (int64_t)(S - 1) * (acc + rej) + (int64_t)kAccEvals * acc);
  }

  DE_HD static uint32_t zero_mask(const double* row, int nrow) {
    uint32_t m = 0;
    for (int j = 0; j < nrow; ++j)
      if (row[j] == 0.0) m |= 1u << j;
    return m;
  }
  DE_HD static uint32_t zero_bi_rows() {
    uint32_t m = 0;
    for (int r = 0; r < I; ++r) {
      bool z = true;
      for (int j = 0; j < O; ++j) z = z && Tab::BI[r][j] == 0.0;
      if (z) m |= 1u << r;
    }
    return m;
  }

  // Stages R .. END-1 from y at (tt, hh): ys = y + sum_{j<R} (a[R][j] hh) k[j]
  // over the nonzero a in j order, NaN where a skipped term meets a
  // non-finite stage.
  DE_HD void stages(int R0, int END, double tt, double hh) {
    for (int R = R0; R < END; ++R) {
      for (int c = 0; c < n; ++c) ys[c] = y[c];
      for (int J = 0; J < R; ++J) {
        const double aj = Tab::A[R][J];
        if (aj != 0.0) {
          const double ah = aj * hh;
          for (int c = 0; c < n; ++c) ys[c] = ys[c] + ah * k[J][c];
        }
      }
      const uint32_t z = zero_mask(Tab::A[R], R);
      if (z != 0)
        for (int c = 0; c < n; ++c) ys[c] = (bad[c] & z) ? as_f64(kNaNBits) : ys[c];
      diff(tt + Tab::C[R] * hh, ys, k[R]);
      for (int c = 0; c < n; ++c) bad[c] |= isfinite(k[R][c]) ? 0u : (1u << R);
    }
  }

  // y[c] + sum_j (w[j] h) k[j][c] over the nonzero weights (SEL 0: b, 1: bh).
  DE_HD double weighted(const double* w, int c) const {
    double v = y[c];
    for (int J = 0; J < S; ++J)
      if (w[J] != 0.0) v = v + (w[J] * h) * k[J][c];
    const uint32_t z = zero_mask(w, S);
    return (z != 0 && (bad[c] & z)) ? as_f64(kNaNBits) : v;
  }

  DE_HD bool init(const SpringArgs* args, int64_t sys) {
    const double t0 = setup(args, sys);
    const OdeArgs& a = A->o;
    const int64_t so = A->state_offset[i];
    double* w = A->work + kVectors * so;
    for (int v = 0; v < I; ++v) k[v] = w + (int64_t)v * n;
    ys = w + (int64_t)I * n;
    yh = w + (int64_t)(I + 1) * n;
    dn = w + (int64_t)(I + 2) * n;
    bad = (uint32_t*)(w + (int64_t)(I + 3) * n);
    rtol = a.rtol_traj ? a.rtol_traj[i] : a.rtol;
    atol = a.atol_traj ? a.atol_traj[i] : a.atol;
    rtv = a.rtol_comp ? a.rtol_comp + so : nullptr;
    atv = a.atol_comp ? a.atol_comp + so : nullptr;
    acc = rej = 0;
    base_ev = 1;
    bool ok = dir == 1.0 || dir == -1.0;
    double h0 = a.h0;
    if (ok && h0 == 0.0) {
      // ordinary.rs:16-61: InitialStepSize's two evaluations plus two more.
      h0 = h_init<O>(t0, k[0], k[1], ys, yh);
      base_ev = 5;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      finish(kBadInput);
      return false;
    }
    h = h0;
    diff(t, y, k[0]);
    h_prev = 0.0;
    stiff = 0;
    rejected_last = false;
    output_start();
    ev_start();
    return true;
  }

  DE_HD bool advance() {
    const OdeArgs& a = A->o;
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) {
        finish(kComplete);
        return true;
      }
      h = hn;
    }
    if (fabs(h) < fabs(h_prev) * 1e-14) {
      finish(kStepSize);
      return true;
    }
    if (acc + rej >= a.max_steps) {
      finish(kMaxSteps);
      return true;
    }
    const uint32_t hbad = isfinite(h) ? 0u : ~0u;
    for (int c = 0; c < n; ++c) bad[c] = hbad | (isfinite(k[0][c]) ? 0u : 1u);
    stages(1, S, t, h);
    double m = 0.0;
    for (int c = 0; c < n; ++c) {
      const double hi = weighted(Tab::B, c);
      const double lo = weighted(Tab::BH, c);
      yh[c] = hi;
      const double e = hi + (-1.0) * lo;
      const double sk = at(c) + rt(c) * fmax(fabs(y[c]), fabs(hi));
      m = fmax(m, fabs(e / sk));
    }
    double scale = a.safety * DE_CTRL_POW(m, -(1.0 / (double)O));
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);
    const bool accept = m <= 1.0;
    const int32_t stiff_next = accept ? (rejected_last ? 0 : stiff) : stiff + 1;
    const bool stop = !accept && (int64_t)stiff_next >= a.max_rejects;
    const double t_new = t + h;
    if (kFsal) {
      for (int c = 0; c < n; ++c) dn[c] = k[S - 1][c];
    } else {
      diff(t_new, yh, dn);
    }
    int32_t sol = -1;
    // The interpolant is built for a dense or even point output, when the
    // next t_eval point is not past the new time or for an event's root
    // search; DefaultSolout pushes the step's end.
    const bool need = accept && ev_pre(t_new, yh);
    bool interp_out = false;
    if (accept && points_on()) {
      if (a.pts_mode == kPointsSteps)
        pts_push(t_new, yh);
      else
        interp_out = true;
    } else if (a.n_eval > 0 && accept && next_eval < a.n_eval) {
      const double te0 = a.t_eval[next_eval];
      interp_out = !(dir > 0.0 ? te0 > t_new : te0 < t_new);
    }
    if (interp_out || need) {
      // The dense stages for the interpolant (ordinary.rs:240-277), or the
      // cubic Hermite interpolant without them.
      if (kBi) stages(S, I, t, h);
      bool fin = true;
      if (kBi) {
        uint32_t anyb = 0;
        for (int c = 0; c < n; ++c) anyb |= bad[c];
        fin = (anyb & zero_bi_rows()) == 0;
      }
      const double tp = t, hp = h;
      auto interp = [&](double te, double* out) {
        if (te < tp || te > t_new) return false;
        if (!out) return true;
        if (kBi) {
          // Horner over bi[r][0..O) times s, then y_prev + sum_r
          // (cont_r h_prev) k[r] in r order; all-zero rows add +-0 and are
          // skipped when their stages, s and h_prev are finite.
          const double s = (te - tp) / hp;
          const bool all = !fin || !isfinite(s) || !isfinite(hp);
          const uint32_t zr = zero_bi_rows();
          for (int c = ic0; c < ic1; ++c) out[c] = y[c];
          for (int r = 0; r < I; ++r) {
            if (((zr >> r) & 1u) && !all) continue;
            double q = Tab::BI[r][O - 1];
            for (int j = O - 2; j >= 0; --j) q = q * s + Tab::BI[r][j];
            q = q * s;
            const double wq = q * hp;
            for (int c = ic0; c < ic1; ++c) out[c] = out[c] + wq * k[r][c];
          }
        } else {
          // cubic_hermite_interpolate (src/interpolate.rs:40-60).
          const double hh = t_new - tp;
          const double s = (te - tp) / hh;
          const double s2 = s * s;
          const double s3 = s2 * s;
          const double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
          const double h10 = s3 - 2.0 * s2 + s;
          const double h01 = -2.0 * s3 + 3.0 * s2;
          const double h11 = s3 - s2;
          const double w10 = h10 * hh, w11 = h11 * hh;
          for (int c = ic0; c < ic1; ++c)
            out[c] = (((0.0 + h00 * y[c]) + w10 * k[0][c]) + h01 * yh[c]) + w11 * dn[c];
        }
        return true;
      };
      if (interp_out && !output(t_new, tp, yh, interp)) sol = kOutOfBounds;
      if (sol < 0) sol = ev_post(need, t_new, tp, interp);
    }
    if (accept && events_on()) ev_commit();
    if (accept) {
      h_prev = h;
      t = t_new;
      for (int c = 0; c < n; ++c) {
        y[c] = yh[c];
        k[0][c] = dn[c];
      }
    }
    scale = accept && rejected_last ? fmin(scale, 1.0) : scale;
    rejected_last = !accept;
    stiff = stiff_next;
    acc += accept ? 1 : 0;
    rej += (accept || stop) ? 0 : 1;
    h = constrain_step_size(h * scale, a.h_min, a.h_max);
    if (stop) {
      finish(kStiffness);
      return true;
    }
    if (sol >= 0) {
      finish(sol);
      return true;
    }
    if (accept && fabs(tf - t) <= kTenEps) {
      finish(kComplete);
      return true;
    }
    return false;
  }
};

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// ----------------------------------------------------------- GroupDopri5 --
// G lanes per system (G a power of two <= 64), C masses per lane:
// lane l of the group owns masses j = l*C + q, q < C, and keeps x_j, v_j of
// every vector in registers.
// A double moved across lanes by one DPP operation on each half (dpp_ctrl
// CTRL; a lane whose source is out of range or inactive reads 0).  Lane
// exchanges within a row and wavefront shifts are VALU operations, without
// the LDS crossbar round trip of a permute.
template <int CTRL>
DE_DEV double dpp_double(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
constexpr int kDppQuadSwap1 = 0xB1;   // quad_perm [1,0,3,2]: lane i reads i ^ 1
constexpr int kDppQuadSwap2 = 0x4E;   // quad_perm [2,3,0,1]: lane i reads i ^ 2
constexpr int kDppHalfMirror = 0x141; // row_half_mirror: lane i reads 7 - i in its 8
constexpr int kDppMirror = 0x140;     // row_mirror: lane i reads 15 - i in its 16
constexpr int kDppWaveShr1 = 0x138;   // wave_shr:1: lane i reads lane i - 1
constexpr int kDppWaveShl1 = 0x130;   // wave_shl:1: lane i reads lane i + 1

// OUT: the point outputs and event solouts are compiled in (their paths
// would otherwise raise the register count of every group kernel).
template <int G, int C, bool OUT = false>
struct GroupDopri5 {
  double x[C], v[C];      // y
  double k0x[C], k0v[C];  // f(t, y)
  double kk[C + 1];       // springs k_{lC} .. k_{lC+C}
  double cd;              // damping
  int nm, l;
  int64_t sys, so;
  double t, h, h_prev, tf, dir, rtol, atol;
  double last;  // EvenSolout::last_output_t
  double ev_value, ev_last, ev_new, tail;  // event solouts (as VecCommon's)
  int32_t acc, rej, ev, steps, stiff, nonstiff, countdown, next_eval, pushed;
  int32_t ev_count, ev_pushed;
  bool rejected_last;

  // Tolerances of state component c (x_j: c = j, v_j: c = nm + j); the
  // per-component arrays are read where used (their pointer is uniform).
  DE_DEV double rt(const OdeArgs& a, int c) const { return a.rtol_comp ? a.rtol_comp[so + c] : rtol; }
  DE_DEV double at(const OdeArgs& a, int c) const { return a.atol_comp ? a.atol_comp[so + c] : atol; }

  // Sum over the group's G lanes, identical on every lane: each level adds
  // the value of a partner lane under an involution (xor 1, xor 2, the mirror
  // within 8 and within 16, then xor 16 and xor 32), so a lane and its partner
  // add the same two operands.
  DE_DEV double gsum(double s) const {
    if constexpr (G >= 2) s += dpp_double<kDppQuadSwap1>(s);
    if constexpr (G >= 4) s += dpp_double<kDppQuadSwap2>(s);
    if constexpr (G >= 8) s += dpp_double<kDppHalfMirror>(s);
    if constexpr (G >= 16) s += dpp_double<kDppMirror>(s);
    if constexpr (G >= 32) s += __shfl_xor(s, 16, G);
    if constexpr (G >= 64) s += __shfl_xor(s, 32, G);
    return s;
  }

  // (dx, dv) = f(x, v): dx_j = v_j, dv_j = spring_accel(...), zero beyond N.
  // The neighbours' displacements come by a wavefront shift; a group's first
  // and last lanes take the chain's fixed ends instead.
  DE_DEV void rhs(const double* xs, const double* vs, double* dx, double* dv) const {
    double left = dpp_double<kDppWaveShr1>(xs[C - 1]);
    double right = dpp_double<kDppWaveShl1>(xs[0]);
    if (l == 0) left = 0.0;
    if (l == G - 1) right = 0.0;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = l * C + q;
      double xm = q == 0 ? left : xs[q - 1];
      double xp = q == C - 1 ? right : xs[q + 1];
      if (j + 1 >= nm) xp = 0.0;
      bool real = j < nm;
      dx[q] = real ? vs[q] : 0.0;
      dv[q] = real ? spring_accel(xm, xs[q], xp, vs[q], kk[q], kk[q + 1], cd) : 0.0;
    }
  }

  DE_DEV void write_state(double* dst, const double* xs, const double* vs) const {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = l * C + q;
      if (j < nm) {
        dst[j] = xs[q];
        dst[nm + j] = vs[q];
      }
    }
  }

  DE_DEV void finish(const SpringArgs& A, int32_t st) const {
    const OdeArgs& a = A.o;
    write_state(a.y_final + so, x, v);
    if (l == 0) {
      a.t_final[sys] = t;
      a.accepted[sys] = acc;
      a.rejected[sys] = rej;
      a.evals[sys] = ev;
      a.status[sys] = st;
      if (a.n_dense) a.n_dense[sys] = pushed;
      if (OUT && a.pts_mode != kPointsNone && a.n_pts) a.n_pts[sys] = pushed;
    }
    if (a.n_eval > 0) {  // points not reached (all of them on BadInput)
      double nan_[C];
#pragma unroll
      for (int q = 0; q < C; ++q) nan_[q] = as_f64(kNaNBits);
      double* dbase = a.dense + (int64_t)a.n_eval * so;
      for (int32_t e = next_eval; e < a.n_eval; ++e)
        write_state(dbase + (int64_t)e * 2 * nm, nan_, nan_);
    }
    if (OUT && a.ev.solout != kSoloutDefault) {  // event count, NaN past the last point
      const EventArgs& e = a.ev;
      if (l == 0 && e.n_out) e.n_out[sys] = ev_pushed;
      double nan_[C];
#pragma unroll
      for (int q = 0; q < C; ++q) nan_[q] = as_f64(kNaNBits);
      for (int32_t k = ev_pushed; k < e.max_events; ++k) {
        if (l == 0) e.t_out[(int64_t)k * a.n + sys] = as_f64(kNaNBits);
        write_state(e.y_out + (int64_t)e.max_events * so + (int64_t)k * 2 * nm, nan_, nan_);
      }
    }
  }

  // ---- event solouts (inline roots, as VecCommon's): the event component
  // lives on one lane of the group and reaches the others by a permute, so
  // every lane runs the same root search on the same values.
  DE_DEV double component(const double* xs, const double* vs, int c) const {
    const int j = c < nm ? c : c - nm;
    const int q = j % C;
    double mine = c < nm ? xs[0] : vs[0];
#pragma unroll
    for (int r = 1; r < C; ++r)
      if (q == r) mine = c < nm ? xs[r] : vs[r];
    return __shfl(mine, j / C, G);
  }
  DE_DEV double ev_g(const EventArgs& e, double tt, const double* xs, const double* vs) const {
    return e.fn == kEventTime ? tt - ev_value : component(xs, vs, e.comp) - ev_value;
  }
  DE_DEV void ev_start(const OdeArgs& a) {
    if (a.ev.solout == kSoloutDefault) return;
    ev_value = a.ev.value_traj ? a.ev.value_traj[sys] : a.ev.value;
    ev_last = ev_g(a.ev, t, x, v);
  }
  DE_DEV bool ev_pre(const OdeArgs& a, double t_curr, const double* xs, const double* vs) {
    const EventArgs& e = a.ev;
    if (e.solout == kSoloutDefault) return false;
    ev_new = ev_g(e, t_curr, xs, vs);
    const double lo = ev_last, g = ev_new;
    const bool change = sign_differs(lo, g);
    if (e.solout == kSoloutEvent) {
      const bool ok = e.direction == kCrossPositive   ? change && lo < 0.0 && g >= 0.0
                      : e.direction == kCrossNegative ? change && lo > 0.0 && g <= 0.0
                                                      : change;
      return ok && lo * g <= 0.0;
    }
    const bool rec = e.direction == kCrossPositive   ? lo < 0.0 && g >= 0.0
                     : e.direction == kCrossNegative ? lo > 0.0 && g <= 0.0
                                                     : true;
    return change && rec;
  }
  // Interpolates the state at tt into the next event slot (while the slots
  // last) and counts the point.
  template <class Interp>
  DE_DEV bool ev_record(const OdeArgs& a, double tt, Interp& interp) {
    const EventArgs& e = a.ev;
    double ox[C], oy[C];
    if (!interp(tt, ox, oy)) return false;
    if (ev_pushed < e.max_events) {
      if (l == 0) e.t_out[(int64_t)ev_pushed * a.n + sys] = tt;
      write_state(e.y_out + (int64_t)e.max_events * so + (int64_t)ev_pushed * 2 * nm, ox, oy);
    }
    ev_pushed += 1;
    return true;
  }
  template <class Interp>
  DE_DEV int32_t ev_post(const OdeArgs& a, bool need, double t_curr, double t_prev,
                         Interp& interp) {
    if (!need) return -1;
    const EventArgs& e = a.ev;
    EventArgs e1 = e;
    e1.comp = 0;  // g of the projected one-component state
    auto interp1 = [&](double tt, double* o) -> bool {
      double ox[C], oy[C];
      if (!interp(tt, ox, oy)) return false;
      o[0] = e.fn == kEventTime ? 0.0 : component(ox, oy, e.comp);
      return true;
    };
    int32_t st = -1;
    if (e.solout == kSoloutEvent) {
      double a0 = t_prev, b0 = t_curr, fa = ev_last, fb = ev_new;
      if ((dir > 0.0 && a0 > b0) || (dir < 0.0 && a0 < b0)) {
        double w = a0;
        a0 = b0;
        b0 = w;
        w = fa;
        fa = fb;
        fb = w;
      }
      const double val = ev_value;
      auto g = [&](double tt, const double* yy) { return event_g<1>(e1, val, tt, yy); };
      double t_event;
      if (brent_dekker<1>(a0, b0, fa, fb, interp1, g, &t_event)) {
        const double tl = a.n_eval > 0 ? tail : t_curr;
        if (tl != tl || fabs(t_event - tl) > 1e-14) {
          if (!ev_record(a, t_event, interp)) return kOutOfBounds;
          tail = t_event;
        } else {
          double ox[C], oy[C];  // the reference interpolates before the push test
          if (!interp(t_event, ox, oy)) return kOutOfBounds;
        }
        ev_count += 1;
        if (e.terminate > 0 && ev_count >= e.terminate) st = kInterrupted;
      }
    } else {
      LaneEvent<1, 2> cs;
      cs.value = ev_value;
      bool found = false;
      double tc = 0.0;
      if (!cs.newton(e1, t_prev, t_curr, ev_last, ev_new, interp1, &found, &tc)) return kOutOfBounds;
      if (!found) {
        const double frac = -ev_last / (ev_new - ev_last);
        tc = t_prev + frac * (t_curr - t_prev);
      }
      if (!ev_record(a, tc, interp)) st = kOutOfBounds;
    }
    return st;
  }

  // As VecCommon::emit; interp returns false outside the step.  The time
  // tests are the same on every lane of the group, so the group agrees.
  template <class Interp>
  DE_DEV bool emit(const SpringArgs& A, double t_curr, double t_prev, const double* xc,
                   const double* vc, Interp&& interp) {
    const OdeArgs& a = A.o;
    int32_t idx = next_eval;
    double* dbase = a.dense + (int64_t)a.n_eval * so;
    bool ok = true;
    while (idx < a.n_eval) {
      double te = a.t_eval[idx];
      bool in_range = dir > 0.0 ? ((te == t_prev && idx == 0) || (te > t_prev && te <= t_curr))
                                : ((te == t_prev && idx == 0) || (te < t_prev && te >= t_curr));
      if (in_range) {
        double* out = dbase + (int64_t)idx * 2 * nm;
        if (te == t_curr) {
          write_state(out, xc, vc);
        } else {
          double ox[C], ov[C];
          if (!interp(te, ox, ov)) {
            ok = false;
            break;
          }
          write_state(out, ox, ov);
        }
        pushed += 1;
        tail = te;
        idx += 1;
      } else {
        if ((dir > 0.0 && te > t_curr) || (dir < 0.0 && te < t_curr)) break;
        double nan_[C];  // passed over: not produced
#pragma unroll
        for (int q = 0; q < C; ++q) nan_[q] = as_f64(kNaNBits);
        write_state(dbase + (int64_t)idx * 2 * nm, nan_, nan_);
        idx += 1;
      }
    }
    next_eval = idx;
    return ok;
  }

  // Point outputs, as VecCommon's: point k of the system has its time at
  // pts_t[pts_offset[sys] + k] (written by the group's first lane) and its
  // state at pts_y[pts_ypos[sys] + k * 2N] while the system's slots last.
  DE_DEV void pts_push(const SpringArgs& A, double tt, const double* xs, const double* vs) {
    const OdeArgs& a = A.o;
    const int64_t lo = a.pts_offset[sys];
    if (lo + pushed < a.pts_offset[sys + 1]) {
      if (l == 0) a.pts_t[lo + pushed] = tt;
      write_state(a.pts_y + A.pts_ypos[sys] + (int64_t)pushed * 2 * nm, xs, vs);
    }
    pushed += 1;
  }
  template <class Interp>
  DE_DEV bool pts_interp(const SpringArgs& A, double ti, Interp& interp) {
    double ox[C], ov[C];
    if (!interp(ti, ox, ov)) return false;
    pts_push(A, ti, ox, ov);
    return true;
  }
  template <class Interp>
  DE_DEV bool points_step(const SpringArgs& A, double t_curr, double t_prev, const double* xc,
                          const double* vc, Interp& interp) {
    const OdeArgs& a = A.o;
    if (a.pts_mode == kPointsSteps) {
      pts_push(A, t_curr, xc, vc);
      return true;
    }
    if (a.pts_mode == kPointsDense) {
      if (t_prev != t_curr) {
        for (int32_t k = 1; k < a.pts_n; ++k) {
          const double h_old = t_curr - t_prev;
          const double ti = t_prev + (double)k * h_old / (double)a.pts_n;
          if (!pts_interp(A, ti, interp)) return false;
        }
      }
      pts_push(A, t_curr, xc, vc);
      return true;
    }
    const double dt = a.pts_dt;
    const double tol = fabs(dt) * 1e-12 + 2.220446049250313080847e-16 * 10.0;
    double ti = last + dt * dir;
    while ((dir > 0.0 && ti <= t_curr) || (dir < 0.0 && ti >= t_curr)) {
      if ((dir > 0.0 && ti >= t_prev && ti <= t_curr) ||
          (dir < 0.0 && ti <= t_prev && ti >= t_curr)) {
        if (!(fabs(ti - last) <= tol)) {
          if (!pts_interp(A, ti, interp)) return false;
          last = ti;
        }
      }
      ti += dt * dir;
    }
    if (t_curr == tf) {
      if (fabs(last - tf) <= tol) {
        pushed -= 1;  // Solution::pop, then the exact final point
        pts_push(A, tf, xc, vc);
        last = tf;
      } else if (last != tf) {
        pts_push(A, tf, xc, vc);
        last = tf;
      }
    }
    return true;
  }

  // Load system `s`; returns false (after writing results) on BadInput.
  DE_DEV bool init(const SpringArgs& A, int64_t s, int lane_in_group) {
    const OdeArgs& a = A.o;
    l = lane_in_group;
    sys = s;
    nm = A.sizes[s];
    so = A.state_offset[s];
    const double* p = a.params + A.param_offset[s];
    cd = p[0];
#pragma unroll
    for (int q = 0; q <= C; ++q) {
      const int j = l * C + q;
      kk[q] = j <= nm ? p[1 + j] : 0.0;
    }
    const double* y0 = a.y0 + so;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = l * C + q;
      x[q] = j < nm ? y0[j] : 0.0;
      v[q] = j < nm ? y0[nm + j] : 0.0;
    }
    double t0 = a.t0 ? a.t0[s] : a.t0s;
    tf = a.tf ? a.tf[s] : a.tfs;
    rtol = a.rtol_traj ? a.rtol_traj[s] : a.rtol;
    atol = a.atol_traj ? a.atol_traj[s] : a.atol;
    t = t0;
    acc = rej = ev = steps = 0;
    pushed = next_eval = 0;
    ev_count = ev_pushed = 0;
    tail = as_f64(kNaNBits);
    dir = signum(tf - t0);
    bool ok = dir == 1.0 || dir == -1.0;
    double h0 = a.h0;
    if (ok && h0 == 0.0) {
      double f0x[C], f0v[C], skx[C], skv[C], y1x[C], y1v[C], f1x[C], f1v[C];
      rhs(x, v, f0x, f0v);
      double dnf = 0.0, dny = 0.0;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        if (l * C + q < nm) {
          const int j = l * C + q;
          skx[q] = at(a, j) + rt(a, j) * fabs(x[q]);
          skv[q] = at(a, nm + j) + rt(a, nm + j) * fabs(v[q]);
          double u = f0x[q] / skx[q], u2 = f0v[q] / skv[q];
          dnf += u * u + u2 * u2;
          double w = x[q] / skx[q], w2 = v[q] / skv[q];
          dny += w * w + w2 * w2;
        } else {
          skx[q] = skv[q] = 1.0;
        }
      }
      dnf = gsum(dnf);
      dny = gsum(dny);
      double hh = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
      hh = fmin(hh, a.h_max);
      hh *= dir;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        y1x[q] = x[q] + hh * f0x[q];
        y1v[q] = v[q] + hh * f0v[q];
      }
      rhs(y1x, y1v, f1x, f1v);
      double der2 = 0.0;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        if (l * C + q < nm) {
          double d = (f1x[q] - f0x[q]) / skx[q], d2 = (f1v[q] - f0v[q]) / skv[q];
          der2 += d * d + d2 * d2;
        }
      }
      der2 = sqrt(gsum(der2)) / fabs(hh);
      double der12 = fmax(sqrt(dnf), der2);
      double h1 = der12 <= 1.0e-15 ? fabs(hh) * 1.0e-3 : DE_CTRL_POW(0.01 / der12, 1.0 / 5.0);
      double interval = fabs(tf - t0);
      hh = fmin(fmax(fmin(fmin(fabs(hh) * 100.0, h1), a.h_max), a.h_min), interval);
      h0 = hh * dir;
      ev += 2;
    }
    ok = ok && validate_step_size(h0, a.h_min, a.h_max, t0, tf);
    if (!ok) {
      ev = 0;
      finish(A, kBadInput);
      return false;
    }
    h = h0;
    rhs(x, v, k0x, k0v);
    ev += 1;
    h_prev = 0.0;
    stiff = nonstiff = 0;
    countdown = 100;
    rejected_last = false;
    if (OUT && a.pts_mode != kPointsNone) {
      pts_push(A, t, x, v);
      last = t;
    } else if (a.n_eval > 0) {
      emit(A, t, t, x, v, [](double, double*, double*) { return true; });
    }
    if constexpr (OUT) ev_start(a);
    return true;
  }

  DE_DEV bool advance(const SpringArgs& A) {
    const OdeArgs& a = A.o;
    if ((t + h - tf) * dir > 0.0) {
      double hn = tf - t;
      if (fabs(hn) < kTenEps) {
        finish(A, kComplete);
        return true;
      }
      h = hn;
    }
    if (fabs(h) < fabs(h_prev) * 1e-14) {
      finish(A, kStepSize);
      return true;
    }
    if (steps >= a.max_steps) {
      finish(A, kMaxSteps);
      return true;
    }
    steps += 1;
    // The stiffness test runs on accepted attempts whose number is a multiple
    // of 100 (ordinary.rs:167); the countdown restarts on every such attempt.
    countdown -= 1;
    const bool stiff_check = countdown == 0;
    if (stiff_check) countdown = 100;
    double k1x[C], k1v[C], k2x[C], k2v[C], k3x[C], k3v[C], k4x[C], k4v[C], k5x[C], k5v[C],
        k6x[C], k6v[C], sx[C], sv[C];
    const double a10h = dp::a10 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = x[q] + a10h * k0x[q];
      sv[q] = v[q] + a10h * k0v[q];
    }
    rhs(sx, sv, k1x, k1v);
    const double a20h = dp::a20 * h, a21h = dp::a21 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = (x[q] + a20h * k0x[q]) + a21h * k1x[q];
      sv[q] = (v[q] + a20h * k0v[q]) + a21h * k1v[q];
    }
    rhs(sx, sv, k2x, k2v);
    const double a30h = dp::a30 * h, a31h = dp::a31 * h, a32h = dp::a32 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = ((x[q] + a30h * k0x[q]) + a31h * k1x[q]) + a32h * k2x[q];
      sv[q] = ((v[q] + a30h * k0v[q]) + a31h * k1v[q]) + a32h * k2v[q];
    }
    rhs(sx, sv, k3x, k3v);
    const double a40h = dp::a40 * h, a41h = dp::a41 * h, a42h = dp::a42 * h,
                 a43h = dp::a43 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = (((x[q] + a40h * k0x[q]) + a41h * k1x[q]) + a42h * k2x[q]) + a43h * k3x[q];
      sv[q] = (((v[q] + a40h * k0v[q]) + a41h * k1v[q]) + a42h * k2v[q]) + a43h * k3v[q];
    }
    rhs(sx, sv, k4x, k4v);
    const double a50h = dp::a50 * h, a51h = dp::a51 * h, a52h = dp::a52 * h,
                 a53h = dp::a53 * h, a54h = dp::a54 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = ((((x[q] + a50h * k0x[q]) + a51h * k1x[q]) + a52h * k2x[q]) + a53h * k3x[q]) +
              a54h * k4x[q];
      sv[q] = ((((v[q] + a50h * k0v[q]) + a51h * k1v[q]) + a52h * k2v[q]) + a53h * k3v[q]) +
              a54h * k4v[q];
    }
    rhs(sx, sv, k5x, k5v);
    const double a60h = dp::a60 * h, a62h = dp::a62 * h, a63h = dp::a63 * h,
                 a64h = dp::a64 * h, a65h = dp::a65 * h;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      sx[q] = ((((x[q] + a60h * k0x[q]) + a62h * k2x[q]) + a63h * k3x[q]) + a64h * k4x[q]) +
              a65h * k5x[q];
      sv[q] = ((((v[q] + a60h * k0v[q]) + a62h * k2v[q]) + a63h * k3v[q]) + a64h * k4v[q]) +
              a65h * k5v[q];
    }
    rhs(sx, sv, k6x, k6v);
    double segx[C], segv[C], nx[C], nv[C];
    double err = 0.0;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      segx[q] = (((dp::b0 * k0x[q] + dp::b2 * k2x[q]) + dp::b3 * k3x[q]) + dp::b4 * k4x[q]) +
                dp::b5 * k5x[q];
      segv[q] = (((dp::b0 * k0v[q] + dp::b2 * k2v[q]) + dp::b3 * k3v[q]) + dp::b4 * k4v[q]) +
                dp::b5 * k5v[q];
      nx[q] = x[q] + h * segx[q];
      nv[q] = v[q] + h * segv[q];
      if (l * C + q < nm) {
        double ex = ((((dp::e0 * k0x[q] + dp::e2 * k2x[q]) + dp::e3 * k3x[q]) +
                      dp::e4 * k4x[q]) +
                     dp::e5 * k5x[q]) +
                    dp::e6 * k6x[q];
        double evv = ((((dp::e0 * k0v[q] + dp::e2 * k2v[q]) + dp::e3 * k3v[q]) +
                       dp::e4 * k4v[q]) +
                      dp::e5 * k5v[q]) +
                     dp::e6 * k6v[q];
        const int j = l * C + q;
        double e1 = ex / (at(a, j) + rt(a, j) * fmax(fabs(x[q]), fabs(nx[q])));
        double e2 = evv / (at(a, nm + j) + rt(a, nm + j) * fmax(fabs(v[q]), fabs(nv[q])));
        err += e1 * e1 + e2 * e2;
      }
    }
    err = gsum(err);
    double deno = err;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * (double)(2 * nm)));
    double scale = a.safety * DE_CTRL_POW(err, -0.2);
    scale = fmin(fmax(scale, a.min_scale), a.max_scale);
    if (err <= 1.0) {
      double t_new = t + h;
      double dnx[C], dnv[C];
      rhs(nx, nv, dnx, dnv);
      if (stiff_check) {
        double stdnum = 0.0, stden = 0.0;
#pragma unroll
        for (int q = 0; q < C; ++q) {
          double u = segx[q] - k6x[q], u2 = segv[q] - k6v[q];
          stdnum += u * u + u2 * u2;
          double w = dnx[q] - sx[q], w2 = dnv[q] - sv[q];
          stden += w * w + w2 * w2;
        }
        stdnum = gsum(stdnum);
        stden = gsum(stden);
        if (stden > 0.0) {
          if (stiff_exceeds(h, stdnum, stden)) {
            nonstiff = 0;
            stiff += 1;
            if (stiff == 15) {
              finish(A, kStiffness);
              return true;
            }
          }
        } else {
          nonstiff += 1;
          if (nonstiff == 6) stiff = 0;
        }
      }
      bool emitted = true;
      int32_t sol = -1;
      const bool pts = OUT && a.pts_mode != kPointsNone;
      bool need = false;
      if constexpr (OUT) need = ev_pre(a, t_new, nx, nv);
      if (pts && a.pts_mode == kPointsSteps) {
        pts_push(A, t_new, nx, nv);
      } else if (a.n_eval > 0 || pts || need) {
        double c1x[C], c1v[C], c2x[C], c2v[C], c3x[C], c3v[C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
          double dx = nx[q] + (-1.0) * x[q], dv2 = nv[q] + (-1.0) * v[q];
          double bx = h * k0x[q] + (-1.0) * dx, bv = h * k0v[q] + (-1.0) * dv2;
          c1x[q] = dx;
          c1v[q] = dv2;
          c2x[q] = bx;
          c2v[q] = bv;
          c3x[q] = (dx + (-h) * dnx[q]) + (-1.0) * bx;
          c3v[q] = (dv2 + (-h) * dnv[q]) + (-1.0) * bv;
        }
        const double hp = h, tp = t, cont4 = 0.0 * h;
        auto interp = [&](double te, double* ox, double* ov) {
          if (te < tp || te > t_new) return false;
          double s = (te - tp) / hp;
          double s1 = 1.0 - s;
#pragma unroll
          for (int q = 0; q < C; ++q) {
            double ax = cont4 * s1 + c3x[q];
            ax = ax * s + c2x[q];
            ax = ax * s1 + c1x[q];
            ox[q] = x[q] + s * ax;
            double av = cont4 * s1 + c3v[q];
            av = av * s + c2v[q];
            av = av * s1 + c1v[q];
            ov[q] = v[q] + s * av;
          }
          return true;
        };
        if constexpr (OUT) {
          if (pts)
            emitted = points_step(A, t_new, t, nx, nv, interp);
          else if (a.n_eval > 0)
            emitted = emit(A, t_new, t, nx, nv, interp);
          sol = emitted ? ev_post(a, need, t_new, t, interp) : kOutOfBounds;
        } else {
          emitted = emit(A, t_new, t, nx, nv, interp);
          sol = emitted ? -1 : kOutOfBounds;
        }
      }
      if constexpr (OUT) {
        if (a.ev.solout != kSoloutDefault) ev_last = ev_new;
      }
      h_prev = h;
      t = t_new;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        x[q] = nx[q];
        v[q] = nv[q];
        k0x[q] = dnx[q];
        k0v[q] = dnv[q];
      }
      if (rejected_last) {
        rejected_last = false;
        scale = fmin(scale, 1.0);
      }
      acc += 1;
      ev += 7;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
      if (sol >= 0) {
        finish(A, sol);
        return true;
      }
      if (fabs(tf - t) <= kTenEps) {
        finish(A, kComplete);
        return true;
      }
    } else {
      rejected_last = true;
      rej += 1;
      ev += 6;
      h = constrain_step_size(h * scale, a.h_min, a.h_max);
    }
    return false;
  }
};
#endif

}  // namespace de
