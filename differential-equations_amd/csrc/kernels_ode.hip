// Lane-per-trajectory ODE kernels for fixed-dimension families.
//
// Compiled twice: with -ffp-contract=off into namespace de::parity (results
// bit-identical to the CPU restatement of the reference) and, with
// -DDE_FAST_BUILD -ffp-contract=fast, into namespace de::fast (FMA-contracted;
// same algorithm, different rounding).
//
// The plain and persistent schedules are in de_ode_sched.h.
#include <hip/hip_runtime.h>

#if defined(DE_FAST_BUILD)
#define DE_KNS fast
#else
#define DE_KNS parity
#endif

#include "de_ode_sched.h"

namespace de {
namespace DE_KNS {

template <class F>
static int launch_method(int method, const OdeArgs& a, int schedule, hipStream_t s) {
  switch (method) {
    case kMethodDopri5: {
      const bool vtol = a.rtol_comp || a.atol_comp;
      if (a.n_eval > 0 || a.pts_mode != kPointsNone)
        return vtol ? launch<Dopri5Lane<F, true, true>>(a, schedule, s)
                    : launch<Dopri5Lane<F, true, false>>(a, schedule, s);
      return vtol ? launch<Dopri5Lane<F, false, true>>(a, schedule, s)
                  : launch<Dopri5Lane<F, false, false>>(a, schedule, s);
    }
#if !defined(DE_FAST_BUILD)
    case kMethodRk4: return launch<FixedLane<F, 4>>(a, schedule, s);
    case kMethodEuler: return launch<FixedLane<F, 1>>(a, schedule, s);
    case kMethodDop853: {
      const bool vtol = a.rtol_comp || a.atol_comp;
      if (a.n_eval > 0 || a.pts_mode != kPointsNone)
        return vtol ? launch<Dop853Lane<F, true, true>>(a, schedule, s)
                    : launch<Dop853Lane<F, true, false>>(a, schedule, s);
      return vtol ? launch<Dop853Lane<F, false, true>>(a, schedule, s)
                  : launch<Dop853Lane<F, false, false>>(a, schedule, s);
    }
#endif
    default: return 1;
  }
}

// Returns 0 on success, 1 if the combination is not built, -1 on HIP error.
int launch_ode_lane(int family, int method, const OdeArgs& a, int schedule, hipStream_t s) {
  switch (family) {
    case kLorenz: return launch_method<LorenzRhs>(method, a, schedule, s);
    case kVanDerPol: return launch_method<VanDerPolRhs>(method, a, schedule, s);
    case kExponential: return launch_method<ExponentialRhs>(method, a, schedule, s);
    case kLinear: return launch_method<LinearRhs>(method, a, schedule, s);
    case kHarmonic: return launch_method<HarmonicRhs>(method, a, schedule, s);
    case kLogistic: return launch_method<LogisticRhs>(method, a, schedule, s);
    case kRobertson: return launch_method<RobertsonRhs>(method, a, schedule, s);
    default: return 1;
  }
}

}  // namespace DE_KNS
}  // namespace de
