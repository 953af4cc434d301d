// Device-compiled right-hand-side families.
//
// The reference takes the right-hand side as user code through the ODE trait
// (ODE::diff, src/ode/ode.rs:20-44) and the SDE trait (SDE::{drift,diffusion},
// src/sde/sde.rs:16-67).  A GPU cannot call back into host code, so the
// integrator offers a fixed set of families compiled into the kernels.  Where
// the reference's own tests define a system, the family evaluates it with the
// same operation order (tests/ode/systems.rs, tests/sde/*.rs), so its bits match
// the reference's RHS.
//
// Parameter layout: a family with P parameters reads params[p * n_traj + i]
// for trajectory i (structure of arrays, coalesced across lanes).
#pragma once

#include "de_common.h"

namespace de {

enum Family : int32_t {
  kLorenz = 0,       // dim 3, params (sigma, rho, beta)
  kVanDerPol = 1,    // dim 2, params (mu)
  kSpringChain = 2,  // dim 2N (ragged), params per system (c, k_0..k_N)
  kExponential = 3,  // dim 1, params (k)
  kLinear = 4,       // dim 1, params (a, b)
  kHarmonic = 5,     // dim 2, params (k)
  kLogistic = 6,     // dim 1, params (k, m)
  kRobertson = 7,    // dim 3, no params
  kNumOdeFamilies = 8,
};

enum SdeFamily : int32_t {
  kOrnsteinUhlenbeck = 0,  // dim D, params per component (theta, mu, sigma)
  kLinearDiagonal = 1,     // dim D, params per component (a, b): drift a*y,
                           // diffusion b*y (geometric Brownian motion for D=1)
  kNumSdeFamilies = 2,
};

// Lorenz, tests/ode/systems.rs:85-101.
struct LorenzRhs {
  static constexpr int kDim = 3;
  static constexpr int kParams = 3;
  double sigma, rho, beta;
  DE_HD void load(const double* p, int64_t stride, int64_t i) {
    sigma = p[i];
    rho = p[stride + i];
    beta = p[2 * stride + i];
  }
  DE_HD void diff(double, const double* y, double* dydt) const {
    double x = y[0], yv = y[1], z = y[2];
    dydt[0] = sigma * (yv - x);
    dydt[1] = x * (rho - z) - yv;
    dydt[2] = x * yv - beta * z;
  }
  static constexpr int kFlops = 8;
};

// Van der Pol, tests/ode/systems.rs:66-78.
struct VanDerPolRhs {
  static constexpr int kDim = 2;
  static constexpr int kParams = 1;
  double mu;
  DE_HD void load(const double* p, int64_t, int64_t i) { mu = p[i]; }
  DE_HD void diff(double, const double* y, double* dydt) const {
    double y1 = y[0], y2 = y[1];
    dydt[0] = y2;
    dydt[1] = mu * (1.0 - y1 * y1) * y2 - y1;
  }
  static constexpr int kFlops = 5;
};

// Exponential growth, tests/ode/systems.rs:9-17.
struct ExponentialRhs {
  static constexpr int kDim = 1;
  static constexpr int kParams = 1;
  double k;
  DE_HD void load(const double* p, int64_t, int64_t i) { k = p[i]; }
  DE_HD void diff(double, const double* y, double* dydt) const {
    dydt[0] = k * y[0];
  }
  static constexpr int kFlops = 1;
};

// Linear equation, tests/ode/systems.rs:22-32.
struct LinearRhs {
  static constexpr int kDim = 1;
  static constexpr int kParams = 2;
  double a, b;
  DE_HD void load(const double* p, int64_t stride, int64_t i) {
    a = p[i];
    b = p[stride + i];
  }
  DE_HD void diff(double, const double* y, double* dydt) const {
    dydt[0] = a + b * y[0];
  }
  static constexpr int kFlops = 2;
};

// Harmonic oscillator, tests/ode/systems.rs:37-48.
struct HarmonicRhs {
  static constexpr int kDim = 2;
  static constexpr int kParams = 1;
  double k;
  DE_HD void load(const double* p, int64_t, int64_t i) { k = p[i]; }
  DE_HD void diff(double, const double* y, double* dydt) const {
    dydt[0] = y[1];
    dydt[1] = -k * y[0];
  }
  static constexpr int kFlops = 1;
};

// Logistic equation, tests/ode/systems.rs:53-61.
struct LogisticRhs {
  static constexpr int kDim = 1;
  static constexpr int kParams = 2;
  double k, m;
  DE_HD void load(const double* p, int64_t stride, int64_t i) {
    k = p[i];
    m = p[stride + i];
  }
  DE_HD void diff(double, const double* y, double* dydt) const {
    dydt[0] = k * y[0] * (1.0 - y[0] / m);
  }
  static constexpr int kFlops = 4;
};

// Robertson kinetics, tests/ode/systems.rs:155-173.
struct RobertsonRhs {
  static constexpr int kDim = 3;
  static constexpr int kParams = 0;
  DE_HD void load(const double*, int64_t, int64_t) {}
  DE_HD void diff(double, const double* y, double* dydt) const {
    double y1 = y[0], y2 = y[1], y3 = y[2];
    dydt[0] = -0.04 * y1 + 1.0e4 * y2 * y3;
    dydt[1] = 0.04 * y1 - 1.0e4 * y2 * y3 - 3.0e7 * y2 * y2;
    dydt[2] = 3.0e7 * y2 * y2;
  }
  static constexpr int kFlops = 13;
};

// Damped spring-mass chain with fixed walls (no counterpart in the reference's
// systems; defined here once).  State [x_0..x_{N-1}, v_0..v_{N-1}], unit
// masses, spring k_j between mass j-1 and j (k_0 and k_N attach to the walls):
//   dx_j = v_j
//   dv_j = (k_{j+1} (x_{j+1} - x_j) - k_j (x_j - x_{j-1})) - c v_j,
//   x_{-1} = x_N = 0.
// params = (c, k_0, ..., k_N).  7N flops.
DE_HD double spring_accel(double xm, double x, double xp, double v, double kl,
                          double kr, double c) {
  return (kr * (xp - x) - kl * (x - xm)) - c * v;
}

DE_HD void spring_chain_diff(int n, const double* params, const double* y,
                             double* dydt) {
  double c = params[0];
  const double* k = params + 1;
  const double* x = y;
  const double* v = y + n;
  for (int j = 0; j < n; ++j) dydt[j] = v[j];
  for (int j = 0; j < n; ++j) {
    double xm = j > 0 ? x[j - 1] : 0.0;
    double xp = j + 1 < n ? x[j + 1] : 0.0;
    dydt[n + j] = spring_accel(xm, x[j], xp, v[j], k[j], k[j + 1], c);
  }
}

// Ornstein-Uhlenbeck, per component: drift theta*(mu - y), diffusion sigma.
// params[(3*c + 0/1/2) * n + i] = theta_c, mu_c, sigma_c of path i.
template <int D>
struct OrnsteinUhlenbeckSde {
  static constexpr int kDim = D;
  static constexpr int kParams = 3 * D;
  double theta[D], mu[D], sigma[D];
  DE_HD void load(const double* p, int64_t stride, int64_t i) {
    for (int c = 0; c < D; ++c) {
      theta[c] = p[(3 * c + 0) * stride + i];
      mu[c] = p[(3 * c + 1) * stride + i];
      sigma[c] = p[(3 * c + 2) * stride + i];
    }
  }
  DE_HD void drift(double, const double* y, double* f) const {
    for (int c = 0; c < D; ++c) f[c] = theta[c] * (mu[c] - y[c]);
  }
  DE_HD void diffusion(double, const double*, double* b) const {
    for (int c = 0; c < D; ++c) b[c] = sigma[c];
  }
  static constexpr bool kAdditive = true;
};

// Linear diagonal SDE: drift a*y, diffusion b*y per component.  D=1 is
// geometric Brownian motion; D=2 with a=(-1,-2), b=(0.5,0.1) is the system of
// tests/sde/milstein.rs:19-29, and D=1 with a=0.5, b=0.1 the one of
// tests/sde/from_fn.rs:10-11.
template <int D>
struct LinearDiagonalSde {
  static constexpr int kDim = D;
  static constexpr int kParams = 2 * D;
  double a[D], b[D];
  DE_HD void load(const double* p, int64_t stride, int64_t i) {
    for (int c = 0; c < D; ++c) {
      a[c] = p[(2 * c + 0) * stride + i];
      b[c] = p[(2 * c + 1) * stride + i];
    }
  }
  DE_HD void drift(double, const double* y, double* f) const {
    for (int c = 0; c < D; ++c) f[c] = a[c] * y[c];
  }
  DE_HD void diffusion(double, const double* y, double* g) const {
    for (int c = 0; c < D; ++c) g[c] = b[c] * y[c];
  }
  static constexpr bool kAdditive = false;
};

}  // namespace de
