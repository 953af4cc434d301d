// Serving pools made on first use (their chunk kernels, by translation unit):
//   DE_SERVE_GROUP 0..4   two embedded-family tableaux each, on Lorenz and Van
//                         der Pol lanes (ErkLane, adaptive/ordinary.rs:16-292),
//                         without and with an event solout
//   DE_SERVE_GROUP 5      RK4 and Euler on Lorenz and Van der Pol lanes
//                         (FixedLane) and DOP853 lanes with an event solout,
//                         and every method on spring chains with one lane per
//                         system (the Vec* solvers, their vectors in the
//                         request's device workspace), and DOPRI5 spring
//                         classes with an event solout on the group mapping
// A lane with an event solout finds its roots inline (no deferred brackets):
// EventWrappedSolout (EV 1) keeps the t_eval output, the crossing solouts
// (EV 2) replace it.
#include "../../include/de_api.h"
#include "de_server.h"

#ifndef DE_SERVE_GROUP
#error "DE_SERVE_GROUP selects the pools of this translation unit"
#endif

namespace de {
namespace serve {

template <class L>
static bool lane_kind(Family& f) {
  set_pool_kind<LaneSlot<L>>(f);
  f.launch = launch_lane<L>;
  return true;
}

#if DE_SERVE_GROUP < 5
template <class F, class Tab>
static bool erk_kind_of(int ev, Family& f) {
  if (ev == 1) return lane_kind<ErkLane<F, Tab, true, false, 1>>(f);
  if (ev == 2) return lane_kind<ErkLane<F, Tab, false, false, 2>>(f);
  return lane_kind<ErkLane<F, Tab, true, false>>(f);
}
template <class Tab>
static bool erk_kind(int family, int ev, Family& f) {
  if (family == DE_FAMILY_LORENZ) return erk_kind_of<LorenzRhs, Tab>(ev, f);
  if (family == DE_FAMILY_VAN_DER_POL) return erk_kind_of<VanDerPolRhs, Tab>(ev, f);
  return false;
}

#define DE_SERVE_PAIR(G, M0, T0, M1, T1)                                          \
  bool lazy_pool_kind_##G(int family, int method, int ev, Family& f) {           \
    if (method == M0) return erk_kind<erk_tab::T0>(family, ev, f);               \
    if (method == M1) return erk_kind<erk_tab::T1>(family, ev, f);               \
    return false;                                                                 \
  }

#if DE_SERVE_GROUP == 0
DE_SERVE_PAIR(0, DE_METHOD_RKF45, Rkf45, DE_METHOD_CASH_KARP, CashKarp)
#elif DE_SERVE_GROUP == 1
DE_SERVE_PAIR(1, DE_METHOD_RKV655E, Rkv655e, DE_METHOD_RKV656E, Rkv656e)
#elif DE_SERVE_GROUP == 2
DE_SERVE_PAIR(2, DE_METHOD_RKV766E, Rkv766e, DE_METHOD_RKV767E, Rkv767e)
#elif DE_SERVE_GROUP == 3
DE_SERVE_PAIR(3, DE_METHOD_RKV877E, Rkv877e, DE_METHOD_RKV878E, Rkv878e)
#elif DE_SERVE_GROUP == 4
DE_SERVE_PAIR(4, DE_METHOD_RKV988E, Rkv988e, DE_METHOD_RKV989E, Rkv989e)
#endif

#else  // DE_SERVE_GROUP == 5

template <class V>
static bool vec_kind(Family& f) {
  f.vectors = V::kVectors;
  return lane_kind<V>(f);
}

template <class F, int STAGES>
static bool fixed_kind_of(int ev, Family& f) {
  if (ev == 1) return lane_kind<FixedLane<F, STAGES, 1>>(f);
  if (ev == 2) return lane_kind<FixedLane<F, STAGES, 2>>(f);
  return lane_kind<FixedLane<F, STAGES>>(f);
}
template <class F>
static bool lane_kind_5(int method, int ev, Family& f) {
  if (method == DE_METHOD_RK4) return fixed_kind_of<F, 4>(ev, f);
  if (method == DE_METHOD_EULER) return fixed_kind_of<F, 1>(ev, f);
  if (method == DE_METHOD_DOP853 && ev == 1) return lane_kind<Dop853Lane<F, true, false, 1>>(f);
  if (method == DE_METHOD_DOP853 && ev == 2) return lane_kind<Dop853Lane<F, false, false, 2>>(f);
  return false;
}

template <int G, int C>
static bool group_event_kind(int cap, Family& f) {
  set_pool_kind<LaneSlot<GroupDopri5<G, C, true>>>(f);
  f.launch = launch_group<G, C, true>;
  f.cap = cap;
  f.g = G;
  return true;
}

bool spring_event_class_kind(int cls, Family& f) {
  switch (cls) {  // the configurations of the DOPRI5 class pools
    case 0: return group_event_kind<4, 1>(4, f);
    case 1: return group_event_kind<8, 1>(8, f);
    case 2: return group_event_kind<16, 1>(16, f);
    case 3: return group_event_kind<32, 1>(32, f);
    case 4: return group_event_kind<64, 1>(64, f);
    case 5: return group_event_kind<64, 2>(128, f);
    case 6: return group_event_kind<64, 4>(256, f);
    default: return false;
  }
}

bool lazy_pool_kind_5(int family, int method, int ev, Family& f) {
  if (family == DE_FAMILY_LORENZ) return lane_kind_5<LorenzRhs>(method, ev, f);
  if (family == DE_FAMILY_VAN_DER_POL) return lane_kind_5<VanDerPolRhs>(method, ev, f);
  if (family != DE_FAMILY_SPRING_CHAIN) return false;
  switch (method) {
    case DE_METHOD_DOPRI5: return vec_kind<VecDopri5>(f);
    case DE_METHOD_DOP853: return vec_kind<VecDop853>(f);
    case DE_METHOD_RK4: return vec_kind<VecFixed<4>>(f);
    case DE_METHOD_EULER: return vec_kind<VecFixed<1>>(f);
    case DE_METHOD_RKF45: return vec_kind<VecErk<erk_tab::Rkf45>>(f);
    case DE_METHOD_CASH_KARP: return vec_kind<VecErk<erk_tab::CashKarp>>(f);
    case DE_METHOD_RKV655E: return vec_kind<VecErk<erk_tab::Rkv655e>>(f);
    case DE_METHOD_RKV656E: return vec_kind<VecErk<erk_tab::Rkv656e>>(f);
    case DE_METHOD_RKV766E: return vec_kind<VecErk<erk_tab::Rkv766e>>(f);
    case DE_METHOD_RKV767E: return vec_kind<VecErk<erk_tab::Rkv767e>>(f);
    case DE_METHOD_RKV877E: return vec_kind<VecErk<erk_tab::Rkv877e>>(f);
    case DE_METHOD_RKV878E: return vec_kind<VecErk<erk_tab::Rkv878e>>(f);
    case DE_METHOD_RKV988E: return vec_kind<VecErk<erk_tab::Rkv988e>>(f);
    case DE_METHOD_RKV989E: return vec_kind<VecErk<erk_tab::Rkv989e>>(f);
    default: return false;
  }
}

#endif

}  // namespace serve
}  // namespace de
