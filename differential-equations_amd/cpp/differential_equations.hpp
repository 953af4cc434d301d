// C++ host mirror of the reference crate's builder API over the C-ABI
// (include/de_api.h), for ensembles.  Header-only; link libde_amd.so.
//
// The reference (Rust) reads
//     let sol = IVP::ode(&lorenz, t0, tf, y0)
//         .method(ExplicitRungeKutta::dopri5().rtol(1e-8).atol(1e-10))
//         .t_eval(points)
//         .solve()?;                                   // src/ivp.rs:277,632,656,781
// and this header reads
//     auto sol = IVP::ode(Lorenz{sigma, rho, beta}, t0, tf, y0s)
//         .method(ExplicitRungeKutta::dopri5().rtol(1e-8).atol(1e-10))
//         .t_eval(points)
//         .solve();                                    // one GPU solve of the ensemble
//     Result r = sol.at(i);    // r.ok ? r.value (Solution) : r.error (Error)
// with the same names, argument meanings and defaults
// (ExplicitRungeKutta::default, src/methods/erk/mod.rs:115-158).  Per-trajectory
// failures come back as Error values (src/error.rs:13-41); a failing call (bad
// descriptor, HIP error) throws std::runtime_error with de_last_error().
#pragma once

#include <stdint.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/de_api.h"

namespace differential_equations {

// ------------------------------------------------------------ results ----
// OutOfBounds: the reference solve panics here (a solout unwrapped an
// Error::OutOfBounds interpolation); t, y are the state the step reached.
enum class ErrorKind { BadInput, MaxSteps, StepSize, Stiffness, OutOfBounds };

struct Error {  // Error::{BadInput, MaxSteps{t,y}, StepSize{t,y}, Stiffness{t,y}}
  ErrorKind kind;
  double t;
  std::vector<double> y;
};

struct Steps {  // src/stats.rs:69
  int64_t accepted = 0, rejected = 0;
  int64_t total() const { return accepted + rejected; }
};

struct Evals {  // src/stats.rs:16
  int64_t function = 0;
};

enum class Status { Complete, Interrupted };  // src/status.rs (successful outcomes)

struct Solution {  // src/solution.rs:30-53 (points kept by the solout)
  std::vector<double> t;
  std::vector<std::vector<double>> y;
  // The points an event solout pushed (IVP::event); the crossing solouts'
  // points are t / y themselves, as they replace the output.
  std::vector<double> event_t;
  std::vector<std::vector<double>> event_y;
  Steps steps;
  Evals evals;
  Status status = Status::Complete;
};

struct Result {  // Result<Solution, Error>
  bool ok;
  Solution value;
  Error error;
};

// ------------------------------------------------------------- events -----
enum class CrossingDirection { Both, Positive, Negative };  // src/solout/mod.rs:38-45

struct EventConfig {  // src/solout/event.rs:8-50
  CrossingDirection dir = CrossingDirection::Both;
  int32_t terminate = 0;  // 0: None
  EventConfig direction(CrossingDirection d) const { EventConfig c = *this; c.dir = d; return c; }
  EventConfig terminate_after(int32_t k) const { EventConfig c = *this; c.terminate = k; return c; }
  EventConfig terminal() const { return terminate_after(1); }
};

// An event function built into the device kernels (the Event trait): y[c] - v,
// t - v or y.minus(&point).dot(&normal); v is one value or one per trajectory.
struct Event {
  int32_t fn = DE_EVENT_COMPONENT;
  int32_t component = 0;
  std::vector<double> value{0.0};
  std::vector<double> point, normal;
  EventConfig config;
  static Event component_at(int32_t c, std::vector<double> v, EventConfig cfg = {}) {
    Event e;
    e.component = c;
    e.value = std::move(v);
    e.config = cfg;
    return e;
  }
  static Event time(std::vector<double> v, EventConfig cfg = {}) {
    Event e;
    e.fn = DE_EVENT_TIME;
    e.value = std::move(v);
    e.config = cfg;
    return e;
  }
  static Event plane(std::vector<double> p, std::vector<double> nrm, EventConfig cfg = {}) {
    Event e;
    e.fn = DE_EVENT_PLANE;
    e.point = std::move(p);
    e.normal = std::move(nrm);
    e.config = cfg;
    return e;
  }
};

// ------------------------------------------------------------ methods ----
class ExplicitRungeKutta {
 public:
  static ExplicitRungeKutta dopri5() { return ExplicitRungeKutta(DE_METHOD_DOPRI5, 0.0); }
  static ExplicitRungeKutta dop853() { return ExplicitRungeKutta(DE_METHOD_DOP853, 0.0); }
  static ExplicitRungeKutta rk4(double h0) { return ExplicitRungeKutta(DE_METHOD_RK4, h0); }
  static ExplicitRungeKutta euler(double h0) { return ExplicitRungeKutta(DE_METHOD_EULER, h0); }
  // The embedded family (src/methods/erk/adaptive/mod.rs:43-122): h0 computed.
  static ExplicitRungeKutta rkf45() { return ExplicitRungeKutta(DE_METHOD_RKF45, 0.0); }
  static ExplicitRungeKutta cash_karp() { return ExplicitRungeKutta(DE_METHOD_CASH_KARP, 0.0); }
  static ExplicitRungeKutta rkv655e() { return ExplicitRungeKutta(DE_METHOD_RKV655E, 0.0); }
  static ExplicitRungeKutta rkv656e() { return ExplicitRungeKutta(DE_METHOD_RKV656E, 0.0); }
  static ExplicitRungeKutta rkv766e() { return ExplicitRungeKutta(DE_METHOD_RKV766E, 0.0); }
  static ExplicitRungeKutta rkv767e() { return ExplicitRungeKutta(DE_METHOD_RKV767E, 0.0); }
  static ExplicitRungeKutta rkv877e() { return ExplicitRungeKutta(DE_METHOD_RKV877E, 0.0); }
  static ExplicitRungeKutta rkv878e() { return ExplicitRungeKutta(DE_METHOD_RKV878E, 0.0); }
  static ExplicitRungeKutta rkv988e() { return ExplicitRungeKutta(DE_METHOD_RKV988E, 0.0); }
  static ExplicitRungeKutta rkv989e() { return ExplicitRungeKutta(DE_METHOD_RKV989E, 0.0); }
  ExplicitRungeKutta rtol(double v) const { auto m = *this; m.b_.rtol = v; m.rtol_v_.clear(); return m; }
  ExplicitRungeKutta atol(double v) const { auto m = *this; m.b_.atol = v; m.atol_v_.clear(); return m; }
  // Tolerance::Vector: one value per state component (src/tolerance.rs:11-68).
  ExplicitRungeKutta rtol(std::vector<double> v) const { auto m = *this; m.rtol_v_ = v; return m; }
  ExplicitRungeKutta atol(std::vector<double> v) const { auto m = *this; m.atol_v_ = v; return m; }
  ExplicitRungeKutta h0(double v) const { auto m = *this; m.b_.h0 = v; return m; }
  ExplicitRungeKutta h_min(double v) const { auto m = *this; m.b_.h_min = v; return m; }
  ExplicitRungeKutta h_max(double v) const { auto m = *this; m.b_.h_max = v; return m; }
  ExplicitRungeKutta max_steps(int64_t v) const { auto m = *this; m.b_.max_steps = v; return m; }
  // Consecutive rejections that end an embedded-family solve (erk/mod.rs:201).
  ExplicitRungeKutta max_rejects(int64_t v) const { auto m = *this; m.b_.max_rejects = v; return m; }
  ExplicitRungeKutta safety_factor(double v) const { auto m = *this; m.b_.safety_factor = v; return m; }
  ExplicitRungeKutta min_scale(double v) const { auto m = *this; m.b_.min_scale = v; return m; }
  ExplicitRungeKutta max_scale(double v) const { auto m = *this; m.b_.max_scale = v; return m; }
  const de_ode_batch& settings() const { return b_; }
  const std::vector<double>& rtol_vector() const { return rtol_v_; }
  const std::vector<double>& atol_vector() const { return atol_v_; }

 private:
  ExplicitRungeKutta(int32_t method, double h0) {
    de_ode_batch_init(&b_);
    b_.method = method;
    b_.h0 = h0;
  }
  de_ode_batch b_;
  std::vector<double> rtol_v_, atol_v_;
};

// ------------------------------------------------------------ systems ----
// Device-compiled RHS families; each parameter is given per trajectory (a
// vector of length n) or once for all (a vector of length 1).
struct Lorenz {  // tests/ode/systems.rs:85-101
  static constexpr int32_t family = DE_FAMILY_LORENZ;
  static constexpr int dim = 3;
  std::vector<double> sigma, rho, beta;
  std::vector<std::vector<double>> params() const { return {sigma, rho, beta}; }
};
struct VanDerPol {  // tests/ode/systems.rs:66-78
  static constexpr int32_t family = DE_FAMILY_VAN_DER_POL;
  static constexpr int dim = 2;
  std::vector<double> mu;
  std::vector<std::vector<double>> params() const { return {mu}; }
};
struct ExponentialGrowth {  // tests/ode/systems.rs:9-17
  static constexpr int32_t family = DE_FAMILY_EXPONENTIAL;
  static constexpr int dim = 1;
  std::vector<double> k;
  std::vector<std::vector<double>> params() const { return {k}; }
};
struct LinearEquation {  // tests/ode/systems.rs:22-32
  static constexpr int32_t family = DE_FAMILY_LINEAR;
  static constexpr int dim = 1;
  std::vector<double> a, b;
  std::vector<std::vector<double>> params() const { return {a, b}; }
};
struct HarmonicOscillator {  // tests/ode/systems.rs:37-48
  static constexpr int32_t family = DE_FAMILY_HARMONIC;
  static constexpr int dim = 2;
  std::vector<double> k;
  std::vector<std::vector<double>> params() const { return {k}; }
};
struct LogisticEquation {  // tests/ode/systems.rs:53-61
  static constexpr int32_t family = DE_FAMILY_LOGISTIC;
  static constexpr int dim = 1;
  std::vector<double> k, m;
  std::vector<std::vector<double>> params() const { return {k, m}; }
};
struct RobertsonProblem {  // tests/ode/systems.rs:155-173
  static constexpr int32_t family = DE_FAMILY_ROBERTSON;
  static constexpr int dim = 3;
  std::vector<std::vector<double>> params() const { return {}; }
};

// ------------------------------------------------------- ensemble result --
class EnsembleSolution {
 public:
  int64_t n = 0;
  int dim = 0;
  double t0 = 0.0;
  std::vector<double> t_eval;
  std::vector<double> y, t, dense;
  std::vector<int64_t> accepted, rejected, evals;
  std::vector<int32_t> status, n_dense;
  int32_t solout = DE_SOLOUT_DEFAULT;
  int32_t max_events = 0;
  std::vector<double> event_t, event_y;
  std::vector<int32_t> n_events;
  // A point output (all_steps / dense / even): trajectory i's points in slots
  // points_offset[i] .. points_offset[i] + min(n_points[i], its slots).
  std::vector<int64_t> points_offset, n_points;
  std::vector<double> points_t, points_y;

  // The reference's Result<Solution, Error> of trajectory i.
  Result at(int64_t i) const {
    Result r;
    std::vector<double> yi(dim);
    for (int c = 0; c < dim; ++c) yi[c] = y[(size_t)c * n + i];
    if (status[i] != DE_STATUS_COMPLETE && status[i] != DE_STATUS_INTERRUPTED) {
      r.ok = false;
      r.error.kind = status[i] == DE_STATUS_MAX_STEPS       ? ErrorKind::MaxSteps
                     : status[i] == DE_STATUS_STEP_SIZE     ? ErrorKind::StepSize
                     : status[i] == DE_STATUS_STIFFNESS     ? ErrorKind::Stiffness
                     : status[i] == DE_STATUS_OUT_OF_BOUNDS ? ErrorKind::OutOfBounds
                                                            : ErrorKind::BadInput;
      r.error.t = t[i];
      r.error.y = yi;
      return r;
    }
    r.ok = true;
    r.value.status = status[i] == DE_STATUS_INTERRUPTED ? Status::Interrupted : Status::Complete;
    r.value.steps.accepted = accepted[i];
    r.value.steps.rejected = rejected[i];
    r.value.evals.function = evals[i];
    const int32_t kept = n_events.empty() ? 0 : std::min(n_events[i], max_events);
    std::vector<double> et;
    std::vector<std::vector<double>> ey;
    for (int32_t k = 0; k < kept; ++k) {
      std::vector<double> yk(dim);
      for (int c = 0; c < dim; ++c) yk[c] = event_y[((size_t)k * dim + c) * n + i];
      et.push_back(event_t[(size_t)k * n + i]);
      ey.push_back(yk);
    }
    if (solout == DE_SOLOUT_CROSSING || solout == DE_SOLOUT_HYPERPLANE) {
      r.value.t = et;  // the crossing solouts replace the output
      r.value.y = ey;
      return r;
    }
    r.value.event_t = et;
    r.value.event_y = ey;
    if (!points_offset.empty()) {  // the Solution's points
      const int64_t lo = points_offset[i];
      const int64_t kept = std::min(n_points[i], points_offset[i + 1] - lo);
      for (int64_t k = 0; k < kept; ++k) {
        r.value.t.push_back(points_t[lo + k]);
        r.value.y.emplace_back(points_y.begin() + (lo + k) * dim,
                               points_y.begin() + (lo + k + 1) * dim);
      }
    } else if (t_eval.empty()) {  // final state (the last point DefaultSolout keeps)
      r.value.t.push_back(t[i]);
      r.value.y.push_back(yi);
    } else {               // TEvalSolout points that were produced
      for (size_t e = 0; e < t_eval.size(); ++e) {
        double v0 = dense[(e * dim) * n + i];
        if (v0 != v0) continue;
        std::vector<double> ye(dim);
        for (int c = 0; c < dim; ++c) ye[c] = dense[(e * dim + c) * n + i];
        r.value.t.push_back(t_eval[e]);
        r.value.y.push_back(ye);
      }
    }
    return r;
  }
};

// ------------------------------------------------------------------ IVP ---
template <class System>
class IVP {
 public:
  // IVP::ode(system, t0, tf, y0) with one initial state per trajectory.
  static IVP ode(const System& sys, double t0, double tf,
                 const std::vector<std::array<double, System::dim>>& y0) {
    return IVP(sys, t0, tf, y0);
  }
  IVP method(const ExplicitRungeKutta& m) const { IVP p = *this; p.m_ = m; return p; }
  IVP t_eval(std::vector<double> points) const {
    IVP p = *this;
    p.t_eval_ = points;
    p.solout_ = DE_SOLOUT_DEFAULT;  // with_solout replaces the whole solout
    p.points_ = DE_POINTS_NONE;
    return p;
  }
  // The default Solution of .solve() (src/solout/default.rs:71-91): (t0, y0)
  // and every accepted step.  Without it (or another output) the ensemble
  // keeps only each trajectory's final state.
  IVP all_steps() const { return with_points(DE_POINTS_STEPS, 0, 0.0); }
  // IVP::dense(n) and IVP::even(dt) (src/ivp.rs:643-655).
  IVP dense(int32_t n) const { return with_points(DE_POINTS_DENSE, n, 0.0); }
  IVP even(double dt) const { return with_points(DE_POINTS_EVEN, 0, dt); }
  // Launch schedule (DE_SCHED_*): persistent lane refill by default.
  IVP schedule(int32_t s) const { IVP p = *this; p.schedule_ = s; return p; }
  // IVP::event (src/ivp.rs:662-678): event detection around the current output.
  IVP event(const Event& e) const {
    IVP p = *this;
    p.solout_ = DE_SOLOUT_EVENT;
    p.event_ = e;
    return p;
  }
  // IVP::crossing (src/ivp.rs:682-692): replaces the output (and any t_eval).
  IVP crossing(int32_t component, double threshold,
               CrossingDirection d = CrossingDirection::Both) const {
    IVP p = *this;
    p.solout_ = DE_SOLOUT_CROSSING;
    p.event_ = Event::component_at(component, {threshold}, EventConfig{}.direction(d));
    p.t_eval_.clear();
    return p;
  }
  // IVP::hyperplane_crossing (src/ivp.rs:695-705) with the identity extractor.
  IVP hyperplane_crossing(std::vector<double> point, std::vector<double> normal,
                          CrossingDirection d = CrossingDirection::Both) const {
    IVP p = *this;
    p.solout_ = DE_SOLOUT_HYPERPLANE;
    p.event_ = Event::plane(std::move(point), std::move(normal), EventConfig{}.direction(d));
    p.t_eval_.clear();
    return p;
  }
  // Event points kept per trajectory.
  IVP max_events(int32_t k) const { IVP p = *this; p.max_events_ = k; return p; }

  // With a point output from all_steps() or dense(n), a first solve counts
  // each trajectory's accepted steps (results are deterministic) and a second
  // writes the points; even(dt) needs floor(|tf - t0| / dt) + 3 slots.
  EnsembleSolution solve() const {
    const int64_t n = (int64_t)y0_.size();
    if (points_ == DE_POINTS_NONE) return solve_once(nullptr);
    std::vector<int64_t> off((size_t)n + 1, 0);
    if (points_ == DE_POINTS_EVEN) {
      const int64_t cnt = (int64_t)std::floor(std::fabs(tf_ - t0_) / points_dt_) + 3;
      for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + cnt;
    } else {
      IVP first = *this;
      first.points_ = DE_POINTS_NONE;
      const EnsembleSolution s = first.solve_once(nullptr);
      for (int64_t i = 0; i < n; ++i)
        off[i + 1] = off[i] + (points_ == DE_POINTS_STEPS ? s.accepted[i] + 1
                                                          : points_n_ * s.accepted[i] + 1);
    }
    return solve_once(&off);
  }

 private:
  IVP with_points(int32_t mode, int32_t n, double dt) const {
    IVP p = *this;
    p.points_ = mode;
    p.points_n_ = n;
    p.points_dt_ = dt;
    p.t_eval_.clear();
    p.solout_ = DE_SOLOUT_DEFAULT;
    return p;
  }

  EnsembleSolution solve_once(const std::vector<int64_t>* points_offset) const {
    const int64_t n = (int64_t)y0_.size();
    de_ode_batch b = m_.settings();
    b.family = System::family;
    b.dim = System::dim;
    b.n = n;
    b.t0_scalar = t0_;
    b.tf_scalar = tf_;
    b.schedule = schedule_;
    b.flags = DE_FLAG_HOST_MEMORY;
    std::vector<double> params;
    for (const auto& col : sys_.params()) {
      if (col.size() != 1 && (int64_t)col.size() != n)
        throw std::runtime_error("parameter vector length must be 1 or n");
      for (int64_t i = 0; i < n; ++i) params.push_back(col.size() == 1 ? col[0] : col[i]);
    }
    std::vector<double> y0((size_t)System::dim * n);
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < System::dim; ++c) y0[(size_t)c * n + i] = y0_[i][c];
    // Tolerance vectors become per-component arrays in the layout of y0.
    auto expand = [&](const std::vector<double>& v, std::vector<double>& out) -> const double* {
      if (v.empty()) return nullptr;
      if ((int)v.size() != System::dim)
        throw std::runtime_error("tolerance vector length must equal the system dimension");
      out.resize((size_t)System::dim * n);
      for (int c = 0; c < System::dim; ++c)
        for (int64_t i = 0; i < n; ++i) out[(size_t)c * n + i] = v[c];
      return out.data();
    };
    std::vector<double> rtol_comp, atol_comp;
    b.rtol_comp = expand(m_.rtol_vector(), rtol_comp);
    b.atol_comp = expand(m_.atol_vector(), atol_comp);
    std::vector<double> te = t_eval_;
    if (!te.empty()) {
      std::sort(te.begin(), te.end());
      if (tf_ < t0_) std::reverse(te.begin(), te.end());
    }
    b.params = params.empty() ? nullptr : params.data();
    b.y0 = y0.data();
    b.n_eval = (int32_t)te.size();
    b.t_eval = te.empty() ? nullptr : te.data();
    EnsembleSolution s;
    s.n = n;
    s.dim = System::dim;
    s.t0 = t0_;
    s.t_eval = te;
    s.y.resize(y0.size());
    s.t.resize(n);
    s.accepted.resize(n);
    s.rejected.resize(n);
    s.evals.resize(n);
    s.status.resize(n);
    if (!te.empty()) {
      s.dense.resize(te.size() * y0.size());
      s.n_dense.resize(n);
    }
    std::vector<double> ev_value;
    if (solout_ != DE_SOLOUT_DEFAULT) {
      const Event& e = event_;
      if (e.value.size() != 1 && (int64_t)e.value.size() != n)
        throw std::runtime_error("event value vector length must be 1 or n");
      b.event_solout = solout_;
      b.event_fn = e.fn;
      b.event_component = e.component;
      b.event_direction = e.config.dir == CrossingDirection::Positive   ? DE_CROSS_POSITIVE
                          : e.config.dir == CrossingDirection::Negative ? DE_CROSS_NEGATIVE
                                                                        : DE_CROSS_BOTH;
      b.event_terminate = e.config.terminate;
      b.max_events = max_events_;
      b.event_value = e.value[0];
      if (e.value.size() != 1) {
        ev_value = e.value;
        b.event_value_traj = ev_value.data();
      }
      b.event_point = e.point.empty() ? nullptr : e.point.data();
      b.event_normal = e.normal.empty() ? nullptr : e.normal.data();
      s.solout = solout_;
      s.max_events = max_events_;
      s.event_t.resize((size_t)max_events_ * n);
      s.event_y.resize((size_t)max_events_ * y0.size());
      s.n_events.resize(n);
    }
    de_ode_result r{};
    r.event_t = s.event_t.empty() ? nullptr : s.event_t.data();
    r.event_y = s.event_y.empty() ? nullptr : s.event_y.data();
    r.n_events = s.n_events.empty() ? nullptr : s.n_events.data();
    r.y_final = s.y.data();
    r.t_final = s.t.data();
    r.accepted = s.accepted.data();
    r.rejected = s.rejected.data();
    r.evals = s.evals.data();
    r.status = s.status.data();
    r.dense = te.empty() ? nullptr : s.dense.data();
    r.n_dense = te.empty() ? nullptr : s.n_dense.data();
    if (points_offset) {
      b.points = points_;
      b.points_n = points_n_;
      b.points_dt = points_dt_;
      s.points_offset = *points_offset;
      s.n_points.resize(n);
      s.points_t.resize((size_t)points_offset->back());
      s.points_y.resize((size_t)points_offset->back() * System::dim);
      r.points_offset = s.points_offset.data();
      r.n_points = s.n_points.data();
      r.points_t = s.points_t.empty() ? nullptr : s.points_t.data();
      r.points_y = s.points_y.empty() ? nullptr : s.points_y.data();
    }
    int rc = de_ode_batch_solve(&b, &r, nullptr);
    if (rc != DE_OK) throw std::runtime_error(std::string("de_ode_batch_solve: ") + de_last_error());
    return s;
  }

  IVP(const System& sys, double t0, double tf,
      const std::vector<std::array<double, System::dim>>& y0)
      : sys_(sys), t0_(t0), tf_(tf), y0_(y0), m_(ExplicitRungeKutta::dopri5()) {}
  System sys_;
  double t0_, tf_;
  std::vector<std::array<double, System::dim>> y0_;
  ExplicitRungeKutta m_;
  std::vector<double> t_eval_;
  int32_t schedule_ = DE_SCHED_AUTO;
  int32_t solout_ = DE_SOLOUT_DEFAULT;
  Event event_;
  int32_t max_events_ = 16;
  int32_t points_ = DE_POINTS_NONE, points_n_ = 0;
  double points_dt_ = 0.0;
};

}  // namespace differential_equations
