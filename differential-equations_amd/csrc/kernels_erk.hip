// Lane-per-trajectory kernels of the embedded explicit Runge-Kutta family
// (rkf45, cash_karp, rkv655e .. rkv989e; src/methods/erk/adaptive) for every
// fixed-dimension family, with scalar or per-component tolerances and with
// the t_eval output compiled in only when it is requested.  Parity build
// (-ffp-contract=off) only.  Compiled five times, DE_ERK_GROUP 0..4, two
// tableaux per object, so that the groups build in parallel.
#include <hip/hip_runtime.h>

#define DE_KNS parity
#include "de_erk_lane.h"
#include "de_launch.h"
#include "de_ode_sched.h"

#ifndef DE_ERK_GROUP
#error "DE_ERK_GROUP must be 0..4"
#endif

namespace de {
namespace parity {

template <class F, class Tab>
static int launch_erk_tab(const OdeArgs& a, int schedule, hipStream_t s) {
  const bool vtol = a.rtol_comp || a.atol_comp;
  if (a.n_eval > 0 || a.pts_mode != kPointsNone)
    return vtol ? launch<ErkLane<F, Tab, true, true>>(a, schedule, s)
                : launch<ErkLane<F, Tab, true, false>>(a, schedule, s);
  return vtol ? launch<ErkLane<F, Tab, false, true>>(a, schedule, s)
              : launch<ErkLane<F, Tab, false, false>>(a, schedule, s);
}

template <class F>
static int launch_erk_method(int method, const OdeArgs& a, int schedule, hipStream_t s) {
  switch (method) {
#if DE_ERK_GROUP == 0
    case kMethodRkf45: return launch_erk_tab<F, erk_tab::Rkf45>(a, schedule, s);
    case kMethodCashKarp: return launch_erk_tab<F, erk_tab::CashKarp>(a, schedule, s);
#elif DE_ERK_GROUP == 1
    case kMethodRkv655e: return launch_erk_tab<F, erk_tab::Rkv655e>(a, schedule, s);
    case kMethodRkv656e: return launch_erk_tab<F, erk_tab::Rkv656e>(a, schedule, s);
#elif DE_ERK_GROUP == 2
    case kMethodRkv766e: return launch_erk_tab<F, erk_tab::Rkv766e>(a, schedule, s);
    case kMethodRkv767e: return launch_erk_tab<F, erk_tab::Rkv767e>(a, schedule, s);
#elif DE_ERK_GROUP == 3
    case kMethodRkv877e: return launch_erk_tab<F, erk_tab::Rkv877e>(a, schedule, s);
    case kMethodRkv878e: return launch_erk_tab<F, erk_tab::Rkv878e>(a, schedule, s);
#else
    case kMethodRkv988e: return launch_erk_tab<F, erk_tab::Rkv988e>(a, schedule, s);
    case kMethodRkv989e: return launch_erk_tab<F, erk_tab::Rkv989e>(a, schedule, s);
#endif
    default: return 1;
  }
}

#define DE_ERK_ENTRY_(g) launch_ode_erk##g
#define DE_ERK_ENTRY(g) DE_ERK_ENTRY_(g)

// Returns 0 on success, 1 if the combination is not in this group, -1 on HIP error.
int DE_ERK_ENTRY(DE_ERK_GROUP)(int family, int method, const OdeArgs& a, int schedule,
                               hipStream_t s) {
  switch (family) {
    case kLorenz: return launch_erk_method<LorenzRhs>(method, a, schedule, s);
    case kVanDerPol: return launch_erk_method<VanDerPolRhs>(method, a, schedule, s);
    case kExponential: return launch_erk_method<ExponentialRhs>(method, a, schedule, s);
    case kLinear: return launch_erk_method<LinearRhs>(method, a, schedule, s);
    case kHarmonic: return launch_erk_method<HarmonicRhs>(method, a, schedule, s);
    case kLogistic: return launch_erk_method<LogisticRhs>(method, a, schedule, s);
    case kRobertson: return launch_erk_method<RobertsonRhs>(method, a, schedule, s);
    default: return 1;
  }
}

}  // namespace parity
}  // namespace de
