// Lane-per-path SDE kernels: Euler-Maruyama and Milstein.
//
// Each lane runs solve_sde (src/sde/solve_ivp.rs:135-287) for one path:
//   Euler-Maruyama = ExplicitRungeKutta::euler on an SDE
//                    (src/methods/erk/fixed/stochastic.rs:18-146)
//   Milstein       = derivative-free Milstein (src/methods/milstein.rs:76-180)
// with the reference's operation order (the leading `0 +` of zeros_like()
// combinations and the `1 *` factors of plus_linear_combination, which are
// exact, are omitted).  SDE::noise is the counter-based Philox source of
// de_noise.h keyed by (seed, path_base + path, step, component), a table, or a
// constant.  State lives in registers; the only HBM traffic is inputs, final
// states and the optional recorded path (every path_stride-th state).
#include <hip/hip_runtime.h>

#include "de_launch.h"
#include "de_noise.h"

namespace de {
namespace parity {

template <class S>
struct SdeLane {
  static constexpr int D = S::kDim;
  S sys;
  double y[D], dydt[D];
  double t, h, tf, dir;
  int64_t steps, ev;
  uint64_t path;
  double sqrt_h;  // sqrt(h), formed where h is set (the start and the clipped last step)

  DE_DEV void noise_fill(const SdeArgs& a, double* dw) const {
    uint32_t call = (uint32_t)(steps - 1);
    if (a.noise == kNoisePhilox) {
      const double sq = sqrt_h;
#pragma unroll
      for (int c = 0; c < D; c += 2) {
        u32x4 ctr = {{(uint32_t)(c >> 1), call, (uint32_t)path, (uint32_t)(path >> 32)}};
        u32x4 w = philox4x32_10(ctr, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        double z0, z1;
        box_muller(w, &z0, &z1);
        dw[c] = sq * z0;
        if (c + 1 < D) dw[c + 1] = sq * z1;
      }
    } else if (a.noise == kNoiseTable) {
      for (int c = 0; c < D; ++c)
        dw[c] = (int64_t)call < a.noise_steps ? a.noise_table[(int64_t)call * D + c] : 0.0;
    } else {
      for (int c = 0; c < D; ++c) dw[c] = a.noise_constant;
    }
  }

  DE_DEV void finish(const SdeArgs& a, int64_t i, int64_t acc, int32_t st) const {
    for (int c = 0; c < D; ++c) a.y_final[(int64_t)c * a.n + i] = y[c];
    a.t_final[i] = t;
    a.steps[i] = acc;
    a.evals[i] = ev;
    a.status[i] = st;
  }

  template <int SCHEME>
  DE_DEV void run(const SdeArgs& a, int64_t i) {
    sys.load(a.params, a.n, i);
    path = a.path_base + (uint64_t)i;
    for (int c = 0; c < D; ++c) y[c] = a.y0[(int64_t)c * a.n + i];
    t = a.t0;
    tf = a.tf;
    ev = 0;
    steps = 0;
    int64_t acc = 0;
    int32_t rec_countdown = a.path_stride;
    double s = signum(tf - a.t0);
    dir = s;
    double h0 = a.h0;
    if (h0 == 0.0) h0 = fabs(tf - a.t0) / 100.0;
    if (!(s == 1.0 || s == -1.0) || !validate_step_size(h0, a.h_min, a.h_max, a.t0, tf)) {
      finish(a, i, 0, kBadInput);
      return;
    }
    h = h0;
    sqrt_h = sqrt(h);
    sys.drift(t, y, dydt);
    ev = SCHEME == kSchemeMilstein ? 1 : 2;  // EM init also evaluates the diffusion
    for (;;) {
      if ((t + h - tf) * dir > 0.0) {
        double hn = tf - t;
        if (fabs(hn) < kTenEps) break;
        h = hn;
        sqrt_h = sqrt(h);
      }
      if (steps >= a.max_steps) {
        finish(a, i, acc, kMaxSteps);
        return;
      }
      steps += 1;
      double b[D], dw[D], ynext[D];
      if (SCHEME == kSchemeEulerMaruyama) {
        const double b0h = 1.0 * h;
        double inc[D];
        for (int c = 0; c < D; ++c) inc[c] = b0h * dydt[c];
        sys.diffusion(t, y, b);
        noise_fill(a, dw);
        for (int c = 0; c < D; ++c) ynext[c] = (y[c] + inc[c]) + b[c] * dw[c];
        ev += 2;
      } else {
        sys.diffusion(t, y, b);
        noise_fill(a, dw);
        double yaux[D], baux[D];
        for (int c = 0; c < D; ++c) yaux[c] = y[c] + sqrt_h * b[c];
        sys.diffusion(t, yaux, baux);
        double factor = 1.0 / (2.0 * sqrt_h);
        for (int c = 0; c < D; ++c) {
          double term = (baux[c] - b[c]) * (dw[c] * dw[c] - h) * factor;
          ynext[c] = ((y[c] + dydt[c] * h) + b[c] * dw[c]) + term;
        }
        ev += 3;
      }
      t += h;
      for (int c = 0; c < D; ++c) y[c] = ynext[c];
      sys.drift(t, y, dydt);
      acc += 1;
      if (a.path_stride > 0 && --rec_countdown == 0) {
        rec_countdown = a.path_stride;
        int64_t w = acc / a.path_stride - 1;
        for (int c = 0; c < D; ++c) a.path[(w * D + c) * a.n + i] = y[c];
        if (i == 0 && a.path_t) a.path_t[w] = t;
      }
      if (fabs(tf - t) <= kTenEps) break;
    }
    finish(a, i, acc, kComplete);
  }
};

constexpr int kSdeBlock = 256;

template <class S, int SCHEME>
__global__ __launch_bounds__(kSdeBlock) void sde_kernel(SdeArgs a) {
  int64_t i = (int64_t)blockIdx.x * kSdeBlock + threadIdx.x;
  if (i >= a.n) return;
  SdeLane<S> L;
  L.template run<SCHEME>(a, i);
}

template <class S>
static int launch_scheme(int scheme, const SdeArgs& a, hipStream_t s) {
  int64_t grid = (a.n + kSdeBlock - 1) / kSdeBlock;
  if (a.n <= 0) return 0;
  if (scheme == kSchemeMilstein)
    hipLaunchKernelGGL((sde_kernel<S, kSchemeMilstein>), dim3((unsigned)grid), dim3(kSdeBlock), 0,
                       s, a);
  else
    hipLaunchKernelGGL((sde_kernel<S, kSchemeEulerMaruyama>), dim3((unsigned)grid),
                       dim3(kSdeBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Dimensions built per family: 1, 2, 4, 8.
int launch_sde(int family, int scheme, const SdeArgs& a, hipStream_t s) {
  if (family == kOrnsteinUhlenbeck) {
    switch (a.dim) {
      case 1: return launch_scheme<OrnsteinUhlenbeckSde<1>>(scheme, a, s);
      case 2: return launch_scheme<OrnsteinUhlenbeckSde<2>>(scheme, a, s);
      case 4: return launch_scheme<OrnsteinUhlenbeckSde<4>>(scheme, a, s);
      case 8: return launch_scheme<OrnsteinUhlenbeckSde<8>>(scheme, a, s);
      default: return 1;
    }
  }
  if (family == kLinearDiagonal) {
    switch (a.dim) {
      case 1: return launch_scheme<LinearDiagonalSde<1>>(scheme, a, s);
      case 2: return launch_scheme<LinearDiagonalSde<2>>(scheme, a, s);
      case 4: return launch_scheme<LinearDiagonalSde<4>>(scheme, a, s);
      case 8: return launch_scheme<LinearDiagonalSde<8>>(scheme, a, s);
      default: return 1;
    }
  }
  return 1;
}

}  // namespace parity
}  // namespace de
