// Lane-per-trajectory ODE kernels with an event solout (parity build only).
//
// IVP::event (EventWrappedSolout around the final-state or t_eval output) and
// IVP::crossing / ::hyperplane_crossing (src/ivp.rs:662-705) for every
// fixed-dimension family and method.  The lanes are de_lane.h's with EV = 1
// (Brent-Dekker event) or EV = 2 (Newton crossing), with scalar or
// per-component tolerances (VTOL) for the adaptive methods, and with the
// t_eval output compiled in only when it is requested.  Compiled twice, with
// DE_EVENT_KIND 1 (event) and 2 (crossings), so that the two halves build in
// parallel.  Events always run on the persistent schedule: a PLAIN request
// runs as PERSISTENT_FIFO (same results, trajectories in index order).
#include <hip/hip_runtime.h>

#define DE_KNS parity
#include "de_ode_sched.h"

#ifndef DE_EVENT_KIND
#error "DE_EVENT_KIND must be 1 or 2"
#endif

namespace de {
namespace parity {

template <class F, int EV, bool DENSE>
static int launch_event_lane(int method, const OdeArgs& a, int schedule, hipStream_t s) {
  const bool vtol = a.rtol_comp || a.atol_comp;
  switch (method) {
    case kMethodDopri5:
      return vtol ? launch<Dopri5Lane<F, DENSE, true, EV>>(a, schedule, s)
                  : launch<Dopri5Lane<F, DENSE, false, EV>>(a, schedule, s);
    case kMethodDop853:
      return vtol ? launch<Dop853Lane<F, DENSE, true, EV>>(a, schedule, s)
                  : launch<Dop853Lane<F, DENSE, false, EV>>(a, schedule, s);
    case kMethodRk4: return launch<FixedLane<F, 4, EV>>(a, schedule, s);
    case kMethodEuler: return launch<FixedLane<F, 1, EV>>(a, schedule, s);
    default: return 1;
  }
}

template <class F>
static int launch_event_family(int method, const OdeArgs& a, int schedule, hipStream_t s) {
#if DE_EVENT_KIND == 1
  if (a.n_eval > 0) return launch_event_lane<F, 1, true>(method, a, schedule, s);
  return launch_event_lane<F, 1, false>(method, a, schedule, s);
#else
  return launch_event_lane<F, 2, false>(method, a, schedule, s);
#endif
}

// Returns 0 on success, 1 if the combination is not built, -1 on HIP error.
#if DE_EVENT_KIND == 1
int launch_ode_event(int family, int method, const OdeArgs& a, int schedule, hipStream_t s) {
#else
int launch_ode_crossing(int family, int method, const OdeArgs& a, int schedule, hipStream_t s) {
#endif
  if (schedule == kSchedPlain) schedule = kSchedPersistentFifo;
  switch (family) {
    case kLorenz: return launch_event_family<LorenzRhs>(method, a, schedule, s);
    case kVanDerPol: return launch_event_family<VanDerPolRhs>(method, a, schedule, s);
    case kExponential: return launch_event_family<ExponentialRhs>(method, a, schedule, s);
    case kLinear: return launch_event_family<LinearRhs>(method, a, schedule, s);
    case kHarmonic: return launch_event_family<HarmonicRhs>(method, a, schedule, s);
    case kLogistic: return launch_event_family<LogisticRhs>(method, a, schedule, s);
    case kRobertson: return launch_event_family<RobertsonRhs>(method, a, schedule, s);
    default: return 1;
  }
}

}  // namespace parity
}  // namespace de
