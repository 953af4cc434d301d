// Lane-per-trajectory kernels of the embedded explicit Runge-Kutta family
// (rkf45, cash_karp, rkv655e .. rkv989e) with an event solout: IVP::event
// (EventWrappedSolout around the final-state or t_eval output) and
// IVP::crossing / ::hyperplane_crossing (src/ivp.rs:662-705), for every
// fixed-dimension family.  Parity build.  Compiled ten times: DE_ERK_GROUP
// 0..4 (two tableaux each) x DE_EVENT_KIND 1 (event) / 2 (crossings).  The
// lanes take per-component tolerances (scalar ones are the same numbers in
// every component), and event lanes run on the persistent schedule (a PLAIN
// request runs as PERSISTENT_FIFO: same results, trajectories in index order).
#include <hip/hip_runtime.h>

#define DE_KNS parity
#include "de_erk_lane.h"
#include "de_launch.h"
#include "de_ode_sched.h"

#ifndef DE_ERK_GROUP
#error "DE_ERK_GROUP must be 0..4"
#endif
#ifndef DE_EVENT_KIND
#error "DE_EVENT_KIND must be 1 or 2"
#endif

namespace de {
namespace parity {

template <class F, class Tab>
static int launch_erk_event_tab(const OdeArgs& a, int schedule, hipStream_t s) {
#if DE_EVENT_KIND == 1
  return launch<ErkLane<F, Tab, true, true, 1>>(a, schedule, s);
#else
  return launch<ErkLane<F, Tab, false, true, 2>>(a, schedule, s);
#endif
}

template <class F>
static int launch_erk_event_method(int method, const OdeArgs& a, int schedule, hipStream_t s) {
  switch (method) {
#if DE_ERK_GROUP == 0
    case kMethodRkf45: return launch_erk_event_tab<F, erk_tab::Rkf45>(a, schedule, s);
    case kMethodCashKarp: return launch_erk_event_tab<F, erk_tab::CashKarp>(a, schedule, s);
#elif DE_ERK_GROUP == 1
    case kMethodRkv655e: return launch_erk_event_tab<F, erk_tab::Rkv655e>(a, schedule, s);
    case kMethodRkv656e: return launch_erk_event_tab<F, erk_tab::Rkv656e>(a, schedule, s);
#elif DE_ERK_GROUP == 2
    case kMethodRkv766e: return launch_erk_event_tab<F, erk_tab::Rkv766e>(a, schedule, s);
    case kMethodRkv767e: return launch_erk_event_tab<F, erk_tab::Rkv767e>(a, schedule, s);
#elif DE_ERK_GROUP == 3
    case kMethodRkv877e: return launch_erk_event_tab<F, erk_tab::Rkv877e>(a, schedule, s);
    case kMethodRkv878e: return launch_erk_event_tab<F, erk_tab::Rkv878e>(a, schedule, s);
#else
    case kMethodRkv988e: return launch_erk_event_tab<F, erk_tab::Rkv988e>(a, schedule, s);
    case kMethodRkv989e: return launch_erk_event_tab<F, erk_tab::Rkv989e>(a, schedule, s);
#endif
    default: return 1;
  }
}

#define DE_ERK_EV_ENTRY_(g, k) launch_ode_erk##g##_ev##k
#define DE_ERK_EV_ENTRY(g, k) DE_ERK_EV_ENTRY_(g, k)

// Returns 0 on success, 1 if the combination is not in this object, -1 on HIP error.
int DE_ERK_EV_ENTRY(DE_ERK_GROUP, DE_EVENT_KIND)(int family, int method, const OdeArgs& a,
                                                 int schedule, hipStream_t s) {
  if (schedule == kSchedPlain) schedule = kSchedPersistentFifo;
  switch (family) {
    case kLorenz: return launch_erk_event_method<LorenzRhs>(method, a, schedule, s);
    case kVanDerPol: return launch_erk_event_method<VanDerPolRhs>(method, a, schedule, s);
    case kExponential: return launch_erk_event_method<ExponentialRhs>(method, a, schedule, s);
    case kLinear: return launch_erk_event_method<LinearRhs>(method, a, schedule, s);
    case kHarmonic: return launch_erk_event_method<HarmonicRhs>(method, a, schedule, s);
    case kLogistic: return launch_erk_event_method<LogisticRhs>(method, a, schedule, s);
    case kRobertson: return launch_erk_event_method<RobertsonRhs>(method, a, schedule, s);
    default: return 1;
  }
}

}  // namespace parity
}  // namespace de
