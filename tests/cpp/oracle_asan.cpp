// The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host
// code only).  Every output layout -- t_eval rows, event points, point
// outputs with short slots, the ragged spring-chain states of each, SDE
// paths, CSV documents -- is written into buffers of exactly the sizes the
// C-ABI documents, so a write or read past a layout's end aborts the run.
// Prints "ok" when every case ran.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/de_api.h"

extern "C" {
int de_oracle_ode_solve(const de_ode_batch* b, de_ode_result* out, int pow_mode, int threads);
int de_oracle_sde_solve(const de_sde_batch* b, de_sde_result* out, int threads);
int de_oracle_ode_csv(const de_ode_batch* b, const de_ode_result* r, int32_t source,
                      int64_t* offsets, char* out);
}

namespace {

void check(bool ok, const char* what) {
  if (!ok) {
    fprintf(stderr, "check failed: %s\n", what);
    exit(1);
  }
}

de_ode_batch defaults(int family, int method, int64_t n) {
  de_ode_batch b;
  memset(&b, 0, sizeof b);
  b.family = family;
  b.method = method;
  b.n = n;
  b.rtol = 1e-7;
  b.atol = 1e-9;
  b.h_max = INFINITY;
  b.max_steps = 200000;
  b.max_rejects = 100;
  b.safety_factor = 0.9;
  b.min_scale = 0.2;
  b.max_scale = 10.0;
  return b;
}

// Result buffers of exactly the documented lengths (vectors allocate exactly
// what they are sized to, so one element past the end is a redzone).
struct Result {
  std::vector<double> y, t, dense, et, ey, pt, py;
  std::vector<int64_t> acc, rej, evals, npts, poff;
  std::vector<int32_t> status, ndense, nev;
  de_ode_result r;

  // state_len: all states; slot_states: points_y length.
  Result(int64_t n, int64_t state_len, int n_eval, int max_events, bool events,
         const std::vector<int64_t>& offsets, int64_t slot_states)
      : y(state_len), t(n), acc(n), rej(n), evals(n), status(n) {
    memset(&r, 0, sizeof r);
    r.y_final = y.data();
    r.t_final = t.data();
    r.accepted = acc.data();
    r.rejected = rej.data();
    r.evals = evals.data();
    r.status = status.data();
    if (n_eval > 0) {
      dense.resize((size_t)n_eval * state_len);
      ndense.resize(n);
      r.dense = dense.data();
      r.n_dense = ndense.data();
    }
    if (events) {
      et.resize((size_t)max_events * n);
      ey.resize((size_t)max_events * state_len);
      nev.resize(n);
      r.event_t = et.empty() ? nullptr : et.data();
      r.event_y = ey.empty() ? nullptr : ey.data();
      r.n_events = nev.data();
    }
    if (!offsets.empty()) {
      poff = offsets;
      pt.resize(offsets.back());
      py.resize(slot_states);
      npts.resize(n);
      r.points_offset = poff.data();
      r.points_t = pt.empty() ? nullptr : pt.data();
      r.points_y = py.empty() ? nullptr : py.data();
      r.n_points = npts.data();
    }
  }
};

// CSV documents of a solved batch into a buffer of exactly their length.
void csv(const de_ode_batch& b, const Result& res, int source) {
  std::vector<int64_t> off(b.n + 1);
  check(de_oracle_ode_csv(&b, &res.r, source, off.data(), nullptr) == 0, "csv sizes");
  std::vector<char> text(off.back() > 0 ? off.back() : 1);
  check(de_oracle_ode_csv(&b, &res.r, source, off.data(), text.data()) == 0, "csv write");
}

// Lorenz: t_eval (points before t0 and past tf), events, crossings, point
// outputs with short slots.
void lorenz_cases() {
  const int64_t n = 48;
  std::vector<double> p(3 * n), y0(3 * n);
  for (int64_t i = 0; i < n; ++i) {
    p[i] = 10.0;
    p[n + i] = 5.0 + 40.0 * (double)i / (double)n;
    p[2 * n + i] = 8.0 / 3.0;
    for (int c = 0; c < 3; ++c) y0[c * n + i] = 1.0 + 0.01 * (double)(i + c);
  }
  std::vector<double> te;
  for (int k = 0; k < 25; ++k) te.push_back(-1.0 + 0.35 * k);
  const int methods[] = {DE_METHOD_DOPRI5, DE_METHOD_DOP853, DE_METHOD_RK4, DE_METHOD_RKV878E};
  for (int m : methods) {
    de_ode_batch b = defaults(DE_FAMILY_LORENZ, m, n);
    b.dim = 3;
    b.params = p.data();
    b.y0 = y0.data();
    b.tf_scalar = 6.0;
    if (m == DE_METHOD_RK4) b.h0 = 0.002;
    b.n_eval = (int32_t)te.size();
    b.t_eval = te.data();
    {
      Result res(n, 3 * n, b.n_eval, 0, false, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "lorenz t_eval");
      csv(b, res, DE_CSV_TEVAL);
      csv(b, res, DE_CSV_FINAL);
    }
    // EventWrappedSolout around t_eval, few slots, a terminate count.
    b.event_solout = DE_SOLOUT_EVENT;
    b.event_fn = DE_EVENT_COMPONENT;
    b.event_component = 2;
    b.event_value = 20.0;
    b.max_events = 3;
    b.event_terminate = 5;
    {
      Result res(n, 3 * n, b.n_eval, b.max_events, true, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "lorenz events");
      csv(b, res, DE_CSV_EVENTS);
    }
    // Crossing and hyperplane solouts (no t_eval), one slot.
    b.n_eval = 0;
    b.t_eval = nullptr;
    b.event_terminate = 0;
    b.max_events = 1;
    b.event_solout = DE_SOLOUT_CROSSING;
    b.event_component = 0;
    b.event_value = 0.0;
    {
      Result res(n, 3 * n, 0, b.max_events, true, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "lorenz crossing");
    }
    const double point[3] = {0.0, 0.0, 25.0}, normal[3] = {1.0, 1.0, 3.0};
    b.event_solout = DE_SOLOUT_HYPERPLANE;
    b.event_point = point;
    b.event_normal = normal;
    {
      Result res(n, 3 * n, 0, b.max_events, true, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "lorenz hyperplane");
    }
    b.event_solout = DE_SOLOUT_DEFAULT;
    b.event_point = b.event_normal = nullptr;
    b.max_events = 0;
    // Point outputs: 0 .. 6 slots per trajectory, far fewer than the points.
    const int modes[] = {DE_POINTS_STEPS, DE_POINTS_DENSE, DE_POINTS_EVEN};
    for (int mode : modes) {
      b.points = mode;
      b.points_n = 3;
      b.points_dt = 0.1;
      std::vector<int64_t> off(n + 1, 0);
      for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + i % 7;
      Result res(n, 3 * n, 0, 0, false, off, off.back() * 3);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "lorenz points");
      for (int64_t i = 0; i < n; ++i) check(res.npts[i] >= 1, "points counted");
      csv(b, res, DE_CSV_POINTS);
    }
  }
}

// Spring chains of ragged sizes: t_eval, events, crossings and point
// outputs, each with its ragged layout.
void spring_cases() {
  const int32_t sizes[] = {1, 7, 2, 40, 5, 3, 16, 1, 9, 4};
  const int64_t n = sizeof sizes / sizeof sizes[0];
  std::vector<int64_t> so(n + 1, 0), po(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    so[i + 1] = so[i] + 2 * sizes[i];
    po[i + 1] = po[i] + sizes[i] + 2;
  }
  std::vector<double> params(po[n]), y0(so[n], 0.0);
  for (int64_t i = 0; i < n; ++i) {
    params[po[i]] = 0.01 * (double)(i % 3);
    for (int j = 0; j <= sizes[i]; ++j) params[po[i] + 1 + j] = 0.5 + 0.1 * (double)((i + j) % 5);
    for (int j = 0; j < sizes[i]; ++j) y0[so[i] + j] = std::sin(0.7 * (double)(j + i));
  }
  std::vector<double> te;
  for (int k = 0; k < 13; ++k) te.push_back(-0.5 + 0.25 * k);
  const int methods[] = {DE_METHOD_DOPRI5, DE_METHOD_DOP853, DE_METHOD_EULER, DE_METHOD_RKV767E};
  for (int m : methods) {
    de_ode_batch b = defaults(DE_FAMILY_SPRING_CHAIN, m, n);
    b.sizes = sizes;
    b.state_offset = so.data();
    b.param_offset = po.data();
    b.params = params.data();
    b.y0 = y0.data();
    b.tf_scalar = 2.5;
    if (m == DE_METHOD_EULER) b.h0 = 0.001;
    b.n_eval = (int32_t)te.size();
    b.t_eval = te.data();
    {
      Result res(n, so[n], b.n_eval, 0, false, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "spring t_eval");
    }
    b.event_solout = DE_SOLOUT_EVENT;
    b.event_fn = DE_EVENT_COMPONENT;
    b.event_component = 1;
    b.max_events = 2;
    {
      Result res(n, so[n], b.n_eval, b.max_events, true, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "spring events");
    }
    b.n_eval = 0;
    b.t_eval = nullptr;
    b.event_solout = DE_SOLOUT_CROSSING;
    b.event_component = 0;
    b.max_events = 1;
    {
      Result res(n, so[n], 0, b.max_events, true, {}, 0);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "spring crossing");
    }
    b.event_solout = DE_SOLOUT_DEFAULT;
    b.max_events = 0;
    const int modes[] = {DE_POINTS_STEPS, DE_POINTS_DENSE, DE_POINTS_EVEN};
    for (int mode : modes) {
      b.points = mode;
      b.points_n = 2;
      b.points_dt = 0.2;
      std::vector<int64_t> off(n + 1, 0);
      int64_t states = 0;
      for (int64_t i = 0; i < n; ++i) {
        off[i + 1] = off[i] + (i * 3) % 5;
        states += (off[i + 1] - off[i]) * 2 * sizes[i];
      }
      Result res(n, so[n], 0, 0, false, off, states);
      check(de_oracle_ode_solve(&b, &res.r, 0, 2) == 0, "spring points");
    }
    b.points = DE_POINTS_NONE;
  }
}

// SDE paths with a recorded trajectory.
void sde_cases() {
  const int64_t n = 32;
  const int dim = 8;
  std::vector<double> p(3 * dim * n), y0(dim * n, 0.0);
  for (int64_t q = 0; q < 3 * dim; ++q)
    for (int64_t i = 0; i < n; ++i) p[q * n + i] = q % 3 == 1 ? 0.0 : 0.5 + 0.01 * (double)i;
  const int schemes[] = {DE_SCHEME_EULER_MARUYAMA, DE_SCHEME_MILSTEIN};
  for (int sc : schemes) {
    de_sde_batch b;
    memset(&b, 0, sizeof b);
    b.family = DE_SDE_ORNSTEIN_UHLENBECK;
    b.scheme = sc;
    b.n = n;
    b.dim = dim;
    b.noise = DE_NOISE_PHILOX;
    b.params = p.data();
    b.y0 = y0.data();
    b.tf = 1.0;
    b.h0 = 1e-2;
    b.h_max = INFINITY;
    b.max_steps = 10000;
    b.seed = 7;
    b.path_stride = 7;
    const int64_t n_rec = 100 / 7;
    std::vector<double> y(dim * n), t(n), path(n_rec * dim * n), path_t(n_rec);
    std::vector<int64_t> steps(n), evals(n);
    std::vector<int32_t> status(n);
    de_sde_result r;
    memset(&r, 0, sizeof r);
    r.y_final = y.data();
    r.t_final = t.data();
    r.steps = steps.data();
    r.evals = evals.data();
    r.status = status.data();
    r.path = path.data();
    r.path_t = path_t.data();
    check(de_oracle_sde_solve(&b, &r, 2) == 0, "sde");
    for (int64_t i = 0; i < n; ++i) check(steps[i] == 100, "sde steps");
  }
}

}  // namespace

int main() {
  lorenz_cases();
  spring_cases();
  sde_cases();
  printf("ok\n");
  return 0;
}
