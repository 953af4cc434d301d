// The reference's DOPRI5 / RK4 / Euler / embedded-family accuracy and error tests
// (tests/ode/accuracy.rs, tests/ode/interpolation.rs, tests/ode/errors.rs,
// tests/ode/from_fn.rs) written against the C++ builder mirror and run on the
// GPU through the C-ABI.  Each case solves a small ensemble (the same problem
// in every lane) and checks every lane like the reference checks its one solve.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../differential-equations_amd/cpp/differential_equations.hpp"

using namespace differential_equations;

static int failures = 0;

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++failures;                        \
    }                                    \
  } while (0)

template <class System>
static void test_ode(const char* name, const System& sys, double t0, double tf,
                     std::array<double, System::dim> y0,
                     std::array<double, System::dim> expected, const ExplicitRungeKutta& m,
                     double tolerance) {
  std::vector<std::array<double, System::dim>> y0s(8, y0);
  auto sol = IVP<System>::ode(sys, t0, tf, y0s).method(m).solve();
  for (int64_t i = 0; i < sol.n; ++i) {
    Result r = sol.at(i);
    CHECK(r.ok, "%s: trajectory %ld did not complete", name, (long)i);
    if (!r.ok) continue;
    const auto& yf = r.value.y.back();
    for (int c = 0; c < System::dim; ++c)
      CHECK(std::fabs(yf[c] - expected[c]) < tolerance, "%s: component %d: expected %.10g got %.10g",
            name, c, expected[c], yf[c]);
  }
  std::printf("%s passed\n", name);
}

int main() {
  // tests/ode/accuracy.rs:70-103
  test_ode("DOPRI5 exponential_growth_positive_direction", ExponentialGrowth{{1.0}}, 0.0, 10.0,
           {1.0}, {22026.46579479}, ExplicitRungeKutta::dopri5(), 1e1);
  test_ode("RK4 exponential_growth_positive_direction", ExponentialGrowth{{1.0}}, 0.0, 10.0,
           {1.0}, {22026.46579479}, ExplicitRungeKutta::rk4(0.01), 1e3);
  test_ode("Euler exponential_growth_positive_direction", ExponentialGrowth{{1.0}}, 0.0, 10.0,
           {1.0}, {22026.46579479}, ExplicitRungeKutta::euler(0.01), 2e3);
  // tests/ode/accuracy.rs:177-213
  test_ode("DOPRI5 exponential_growth_negative_direction", ExponentialGrowth{{1.0}}, 0.0, -10.0,
           {22026.46579479}, {1.0}, ExplicitRungeKutta::dopri5(), 1e-2);
  test_ode("RK4 exponential_growth_negative_direction", ExponentialGrowth{{1.0}}, 0.0, -10.0,
           {22026.46579479}, {1.0}, ExplicitRungeKutta::rk4(-0.01), 1e-2);
  // tests/ode/accuracy.rs:286-321
  test_ode("DOPRI5 linear_equation", LinearEquation{{1.0}, {1.0}}, 0.0, 10.0, {1.0},
           {44051.93158958}, ExplicitRungeKutta::dopri5(), 1e1);
  // tests/ode/accuracy.rs:395-430
  test_ode("DOPRI5 harmonic_oscillator", HarmonicOscillator{{1.0}}, 0.0, 10.0, {1.0, 0.0},
           {-0.83907153, 0.54402111}, ExplicitRungeKutta::dopri5(), 1e-2);
  test_ode("RK4 harmonic_oscillator", HarmonicOscillator{{1.0}}, 0.0, 10.0, {1.0, 0.0},
           {-0.83907153, 0.54402111}, ExplicitRungeKutta::rk4(0.01), 1e-1);
  // tests/ode/accuracy.rs:503-539
  test_ode("DOPRI5 logistic_equation", LogisticEquation{{1.0}, {10.0}}, 0.0, 10.0, {0.1},
           {9.95525518}, ExplicitRungeKutta::dopri5(), 1e-3);
  // tests/ode/accuracy.rs:622-671
  test_ode("DOPRI5 robertson_stiff_problem", RobertsonProblem{}, 0.0, 0.4, {1.0, 0.0, 0.0},
           {0.9851721139, 3.3863953790e-05, 0.0147940222},
           ExplicitRungeKutta::dopri5().rtol(1e-8).atol(1e-10), 5e-1);
  // tests/ode/from_fn.rs:3-18
  test_ode("Euler from_fn", ExponentialGrowth{{1.0}}, 0.0, 1.0, {1.0}, {2.5937},
           ExplicitRungeKutta::euler(0.1), 1e-3);

  // tests/ode/accuracy.rs rows of the embedded family (rkv656e has no row and
  // takes rkv655e's tolerances); each tableau, each system.
  {
    struct Named {
      const char* name;
      ExplicitRungeKutta m;
      bool verner;
    };
    const Named family[] = {
        {"RKF45", ExplicitRungeKutta::rkf45(), false},
        {"CashKarp", ExplicitRungeKutta::cash_karp(), false},
        {"RKV655e", ExplicitRungeKutta::rkv655e(), true},
        {"RKV656e", ExplicitRungeKutta::rkv656e(), true},
        {"RKV766e", ExplicitRungeKutta::rkv766e(), true},
        {"RKV767e", ExplicitRungeKutta::rkv767e(), true},
        {"RKV877e", ExplicitRungeKutta::rkv877e(), true},
        {"RKV878e", ExplicitRungeKutta::rkv878e(), true},
        {"RKV988e", ExplicitRungeKutta::rkv988e(), true},
        {"RKV989e", ExplicitRungeKutta::rkv989e(), true},
    };
    char label[128];
    for (const Named& f : family) {
      std::snprintf(label, sizeof label, "%s exponential_growth_positive_direction", f.name);
      test_ode(label, ExponentialGrowth{{1.0}}, 0.0, 10.0, {1.0}, {22026.46579479}, f.m,
               f.verner ? 1e-1 : 1e3);  // :70-176
      std::snprintf(label, sizeof label, "%s exponential_growth_negative_direction", f.name);
      test_ode(label, ExponentialGrowth{{1.0}}, 0.0, -10.0, {22026.46579479}, {1.0}, f.m,
               1e-2);  // :177-285
      std::snprintf(label, sizeof label, "%s linear_equation", f.name);
      test_ode(label, LinearEquation{{1.0}, {1.0}}, 0.0, 10.0, {1.0}, {44051.93158958}, f.m,
               f.verner ? 1e1 : 1e3);  // :286-394
      std::snprintf(label, sizeof label, "%s harmonic_oscillator", f.name);
      test_ode(label, HarmonicOscillator{{1.0}}, 0.0, 10.0, {1.0, 0.0},
               {-0.83907153, 0.54402111}, f.m, 1e-3);  // :395-503
      std::snprintf(label, sizeof label, "%s logistic_equation", f.name);
      test_ode(label, LogisticEquation{{1.0}, {10.0}}, 0.0, 10.0, {0.1}, {9.95525518}, f.m,
               f.verner ? 1e-3 : 1e-2);  // :504-612
    }
  }

  // tests/ode/interpolation.rs:22-72
  for (auto m : {ExplicitRungeKutta::dopri5(), ExplicitRungeKutta::rk4(0.01),
                 ExplicitRungeKutta::rkv655e(), ExplicitRungeKutta::rkv877e(),
                 ExplicitRungeKutta::rkv988e(), ExplicitRungeKutta::rkf45()}) {
    std::vector<std::array<double, 1>> y0s(4, {1.0});
    auto sol = IVP<ExponentialGrowth>::ode(ExponentialGrowth{{1.0}}, 0.0, 2.0, y0s)
                   .t_eval({0.5, 1.0, 1.69})
                   .method(m)
                   .solve();
    Result r = sol.at(0);
    CHECK(r.ok && r.value.y.size() == 3, "interpolation: expected 3 points");
    if (r.ok) CHECK(std::fabs(r.value.y[0][0] - std::exp(0.5)) < 1e-3, "interpolation value");
  }
  std::printf("interpolation passed\n");

  // tests/ode/errors.rs:87-130
  {
    std::vector<std::array<double, 1>> y0s(2, {1.0});
    auto a = IVP<ExponentialGrowth>::ode(ExponentialGrowth{{1.0}}, 0.0, 0.0, y0s).solve();
    CHECK(!a.at(0).ok && a.at(0).error.kind == ErrorKind::BadInput, "invalid_time_span");
    auto b = IVP<ExponentialGrowth>::ode(ExponentialGrowth{{1.0}}, 0.0, 1.0, y0s)
                 .method(ExplicitRungeKutta::dopri5().h0(10.0))
                 .solve();
    CHECK(!b.at(0).ok && b.at(0).error.kind == ErrorKind::BadInput, "initial_step_size_too_big");
    auto c = IVP<ExponentialGrowth>::ode(ExponentialGrowth{{1.0}}, 0.0, 1.0, y0s)
                 .method(ExplicitRungeKutta::rk4(10.0))
                 .solve();
    CHECK(!c.at(0).ok && c.at(0).error.kind == ErrorKind::BadInput, "rk4 h0 too big");
  }
  std::printf("errors passed\n");

  // tests/ode/errors.rs SimpleODE with its event g = t - 10,
  // EventConfig::default().terminal(): the solve stops Interrupted after the
  // step that crosses 10, with one event point at 10.
  for (auto m : {ExplicitRungeKutta::dopri5(), ExplicitRungeKutta::dop853(),
                 ExplicitRungeKutta::rk4(0.1), ExplicitRungeKutta::euler(0.1)}) {
    std::vector<std::array<double, 1>> y0s(4, {1.0});
    auto sol = IVP<ExponentialGrowth>::ode(ExponentialGrowth{{1.0}}, 0.0, 20.0, y0s)
                   .method(m)
                   .event(Event::time({10.0}, EventConfig{}.terminal()))
                   .solve();
    for (int64_t i = 0; i < sol.n; ++i) {
      Result r = sol.at(i);
      CHECK(r.ok && r.value.status == Status::Interrupted, "time event: not interrupted");
      CHECK(r.value.event_t.size() == 1 && std::fabs(r.value.event_t[0] - 10.0) < 1e-9,
            "time event: point");
      CHECK(r.value.t.size() == 1 && r.value.t[0] >= 10.0, "time event: final time");
    }
  }
  // examples/ode/03_logistic_growth: k = 1, m = 10, y0 = 1, g = y - 0.9 m,
  // terminal, DOP853 rtol = atol = 1e-12; y(t) = 10 / (1 + 9 e^-t) = 9 at ln 81.
  {
    std::vector<std::array<double, 1>> y0s(4, {1.0});
    auto sol = IVP<LogisticEquation>::ode(LogisticEquation{{1.0}, {10.0}}, 0.0, 10.0, y0s)
                   .method(ExplicitRungeKutta::dop853().rtol(1e-12).atol(1e-12))
                   .event(Event::component_at(0, {9.0}, EventConfig{}.terminal()))
                   .solve();
    Result r = sol.at(0);
    CHECK(r.ok && r.value.status == Status::Interrupted && r.value.event_t.size() == 1,
          "logistic event");
    if (r.ok && r.value.event_t.size() == 1)
      CHECK(std::fabs(r.value.event_t[0] - std::log(81.0)) < 1e-5, "logistic event time %.12g",
            r.value.event_t[0]);
  }
  // IVP::crossing(0, 0.0) on y'' = -y, y(0) = (1, 0): the zeros of cos t.
  {
    std::vector<std::array<double, 2>> y0s(4, {1.0, 0.0});
    auto sol = IVP<HarmonicOscillator>::ode(HarmonicOscillator{{1.0}}, 0.0, 20.0, y0s)
                   .method(ExplicitRungeKutta::dopri5().rtol(1e-10).atol(1e-12))
                   .crossing(0, 0.0)
                   .solve();
    Result r = sol.at(1);
    CHECK(r.ok && r.value.t.size() == 6, "crossing: expected 6 zeros");
    for (size_t k = 0; r.ok && k < r.value.t.size(); ++k)
      CHECK(std::fabs(r.value.t[k] - (M_PI / 2 + M_PI * k)) < 1e-8, "crossing zero %zu", k);
  }
  std::printf("events passed\n");

  // The default Solution (every accepted step), dense(3) and even(0.5) on
  // y'' = -y, y(0) = (1, 0): (t0, y0) first, the final state last, A + 1 and
  // 3A + 1 points, the grid 0, 0.5, ..., 20, and cos t along the way.
  {
    std::vector<std::array<double, 2>> y0s(3, {1.0, 0.0});
    auto ivp = IVP<HarmonicOscillator>::ode(HarmonicOscillator{{1.0}}, 0.0, 20.0, y0s)
                   .method(ExplicitRungeKutta::dopri5().rtol(1e-10).atol(1e-12));
    auto steps = ivp.all_steps().solve();
    auto dense = ivp.dense(3).solve();
    auto even = ivp.even(0.5).solve();
    for (int64_t i = 0; i < steps.n; ++i) {
      Result a = steps.at(i), d = dense.at(i), e = even.at(i);
      CHECK(a.ok && (int64_t)a.value.t.size() == a.value.steps.accepted + 1, "steps: A + 1");
      CHECK(d.ok && (int64_t)d.value.t.size() == 3 * d.value.steps.accepted + 1, "dense: 3A + 1");
      CHECK(e.ok && e.value.t.size() == 41, "even: 41 points");
      if (!a.ok || !d.ok || !e.ok) continue;
      CHECK(a.value.t.front() == 0.0 && a.value.y.front()[0] == 1.0, "steps: t0, y0 first");
      CHECK(a.value.t.back() == steps.t[i] && a.value.y.back()[0] == steps.y[i], "steps: last");
      for (size_t k = 0; k < e.value.t.size(); ++k) {
        CHECK(e.value.t[k] == 0.5 * (double)k, "even grid %zu", k);
        CHECK(std::fabs(e.value.y[k][0] - std::cos(e.value.t[k])) < 1e-8, "even: cos t");
      }
      for (size_t k = 0; k < d.value.t.size(); ++k)
        CHECK(std::fabs(d.value.y[k][0] - std::cos(d.value.t[k])) < 1e-8, "dense: cos t");
    }
  }
  std::printf("point outputs passed\n");

  // Ensemble: per-trajectory parameters (the builder's vectors) on Lorenz.
  {
    const int n = 256;
    std::vector<double> rho(n);
    std::vector<std::array<double, 3>> y0s(n);
    for (int i = 0; i < n; ++i) {
      rho[i] = 50.0 * i / n;
      y0s[i] = {1.0, 1.0, 1.0};
    }
    auto sol = IVP<Lorenz>::ode(Lorenz{{10.0}, rho, {8.0 / 3.0}}, 0.0, 2.0, y0s)
                   .method(ExplicitRungeKutta::dopri5().rtol(1e-8).atol(1e-10).max_steps(100000))
                   .solve();
    for (int i = 0; i < n; ++i) {
      Result r = sol.at(i);
      CHECK(r.ok, "lorenz trajectory %d", i);
      if (r.ok)
        CHECK(r.value.evals.function == 3 + 6 * r.value.steps.total() + r.value.steps.accepted,
              "lorenz evals invariant %d", i);
    }
    std::printf("lorenz ensemble passed\n");
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
  return failures ? 1 : 0;
}
