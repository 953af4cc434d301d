// Spring-chain ensembles: uneven system sizes (2N in [8, 512]).
//
// Mappings (selected by de_ode_batch.mapping):
//   DE_MAP_LANE   one thread per system, vectors in an HBM workspace (plain
//                 launch; bit-identical to the CPU restatement).  DOPRI5, and
//                 the only mapping of DOP853, RK4, Euler and the embedded
//                 family.
//   DE_MAP_GROUP  G lanes per system with C masses per lane in registers.
//   DE_MAP_AUTO   systems are bucketed into size classes (N <= 4, 8, ..., 256);
//                 for each class an autotuner times every (G, C, block) with
//                 G*C equal to the class cap on a sample of the class and keeps
//                 the fastest.  Classes run concurrently on their own streams,
//                 each as a persistent kernel whose groups take the next system
//                 of the class from a work counter when they finish one.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "de_launch.h"
#include "de_spring.h"
#include "de_spring_solvers.h"

namespace de {
namespace parity {

constexpr int kSpringBlock = 256;

template <class Solver>
__global__ __launch_bounds__(kSpringBlock) void spring_lane_kernel(SpringArgs A) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.o.n) return;
  Solver s;
  if (!s.init(&A, i)) return;
  while (!s.advance()) {
  }
}

// The one-lane solver of each method: its kernel and its workspace doubles
// per state component.
struct LaneSolver {
  void (*kernel)(SpringArgs);
  int vectors;
};
template <class Solver>
static LaneSolver lane_solver_of() {
  return {spring_lane_kernel<Solver>, Solver::kVectors};
}
static LaneSolver lane_solver(int method) {
  switch (method) {
    case DE_METHOD_DOP853: return lane_solver_of<VecDop853>();
    case DE_METHOD_RK4: return lane_solver_of<VecFixed<4>>();
    case DE_METHOD_EULER: return lane_solver_of<VecFixed<1>>();
    case DE_METHOD_RKF45: return lane_solver_of<VecErk<erk_tab::Rkf45>>();
    case DE_METHOD_CASH_KARP: return lane_solver_of<VecErk<erk_tab::CashKarp>>();
    case DE_METHOD_RKV655E: return lane_solver_of<VecErk<erk_tab::Rkv655e>>();
    case DE_METHOD_RKV656E: return lane_solver_of<VecErk<erk_tab::Rkv656e>>();
    case DE_METHOD_RKV766E: return lane_solver_of<VecErk<erk_tab::Rkv766e>>();
    case DE_METHOD_RKV767E: return lane_solver_of<VecErk<erk_tab::Rkv767e>>();
    case DE_METHOD_RKV877E: return lane_solver_of<VecErk<erk_tab::Rkv877e>>();
    case DE_METHOD_RKV878E: return lane_solver_of<VecErk<erk_tab::Rkv878e>>();
    case DE_METHOD_RKV988E: return lane_solver_of<VecErk<erk_tab::Rkv988e>>();
    case DE_METHOD_RKV989E: return lane_solver_of<VecErk<erk_tab::Rkv989e>>();
    default: return lane_solver_of<VecDopri5>();
  }
}

// Occupancy hints: the rarely taken double-double pow fallback raises the
// peak register count past a wave-slot boundary; the hints trade that for a
// few spills outside the step block.
template <int G, int C, bool OUT>
__global__ __launch_bounds__(kSpringBlock) __attribute__((amdgpu_waves_per_eu(C == 1 ? 3 : 2))) void
spring_group_kernel(SpringArgs A) {
  const int lane = __lane_id();
  const int l = lane & (G - 1);
  const int leader_lane = lane - l;
  GroupDopri5<G, C, OUT> S;
  bool have = false;
  bool drained = false;
  const unsigned long long len = (unsigned long long)A.list_len;
  for (;;) {
    if (!drained) {
      bool need = !have;
      unsigned long long mask = __ballot(need && l == 0);
      if (mask) {
        int leader = __ffsll((long long)mask) - 1;
        int cnt = __popcll(mask);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(A.o.counter, (unsigned long long)cnt);
        base = __shfl(base, leader);
        if (base + (unsigned long long)cnt >= len) drained = true;
        if (need) {
          unsigned long long below =
              leader_lane == 0 ? 0ull : (mask & ((1ull << leader_lane) - 1ull));
          unsigned long long my = base + (unsigned long long)__popcll(below);
          if (my < len) have = S.init(A, (int64_t)A.list[my], l);
        }
      }
    }
    if (!__any(have)) {
      if (drained) break;
      continue;
    }
    if (have) {
      if (S.advance(A)) have = false;
    }
  }
}

__global__ void spring_zero_counter(unsigned long long* c) { *c = 0ull; }

typedef void (*GroupKernel)(SpringArgs);

struct Config {
  int g, c;
  GroupKernel k;   // final state and t_eval output
  GroupKernel kx;  // with the point outputs and event solouts
};

#define DE_CFG(G, C) {G, C, spring_group_kernel<G, C, false>, spring_group_kernel<G, C, true>}
static const Config kConfigs[] = {
    DE_CFG(4, 1),  DE_CFG(8, 1),  DE_CFG(16, 1), DE_CFG(32, 1), DE_CFG(64, 1),
    DE_CFG(2, 2),  DE_CFG(4, 2),  DE_CFG(8, 2),  DE_CFG(16, 2), DE_CFG(32, 2), DE_CFG(64, 2),
    DE_CFG(1, 4),  DE_CFG(2, 4),  DE_CFG(4, 4),  DE_CFG(8, 4),  DE_CFG(16, 4), DE_CFG(32, 4),
    DE_CFG(64, 4),
};
#undef DE_CFG
static const int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
static const int kCaps[] = {4, 8, 16, 32, 64, 128, 256};
static const int kNumCaps = 7;
static const int kBlocks[] = {64, 256};

struct Choice {
  int cfg = -1;
  int block = kSpringBlock;
};

// Process-wide state: autotune decisions per size class, streams, workspace.
static std::mutex g_mu;
static std::map<int, Choice> g_tuned;
static hipStream_t g_streams[kNumCaps] = {};
static hipEvent_t g_fork = nullptr, g_join[kNumCaps] = {};
static unsigned long long* g_counters = nullptr;
static int g_num_cu = 0;

static int cap_index(int nm) {
  for (int k = 0; k < kNumCaps; ++k)
    if (nm <= kCaps[k]) return k;
  return -1;
}

static int ensure_runtime() {
  if (g_fork) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return -1;
  for (int k = 0; k < kNumCaps; ++k) {
    if (hipStreamCreateWithFlags(&g_streams[k], hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&g_join[k], hipEventDisableTiming) != hipSuccess) return -1;
  }
  if (hipEventCreateWithFlags(&g_fork, hipEventDisableTiming) != hipSuccess) return -1;
  if (hipMalloc(&g_counters, sizeof(unsigned long long) * (kNumCaps + 1)) != hipSuccess) return -1;
  return 0;
}

static int blocks_per_cu(GroupKernel k, int block) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, block, 0) != hipSuccess || b < 1)
    b = 1;
  return b;
}

// The group kernel of a launch: with the point outputs and event solouts
// compiled in only when the launch asks for one of them.
static GroupKernel class_kernel(const SpringArgs& base, const Config& cf) {
  return base.o.pts_mode != kPointsNone || base.o.ev.solout != kSoloutDefault ? cf.kx : cf.k;
}

static int launch_class(const SpringArgs& base, const int32_t* d_list, int64_t len,
                        int64_t work_share_blocks, const Choice& ch, unsigned long long* counter,
                        hipStream_t s) {
  const Config& cf = kConfigs[ch.cfg];
  SpringArgs A = base;
  A.list = d_list;
  A.list_len = len;
  A.o.counter = counter;
  int64_t groups_per_block = ch.block / cf.g;
  int64_t need = (len + groups_per_block - 1) / groups_per_block;
  int64_t grid = work_share_blocks < need ? work_share_blocks : need;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(spring_zero_counter, dim3(1), dim3(1), 0, s, counter);
  hipLaunchKernelGGL(class_kernel(base, cf), dim3((unsigned)grid), dim3(ch.block), 0, s, A);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Time every candidate for size class k on (up to) `sample` systems.
static Choice autotune(const SpringArgs& base, int k, const int32_t* d_list, int64_t len,
                       hipStream_t s) {
  Choice best;
  float best_ms = 1e30f;
  const int64_t sample = len < 1024 ? len : 1024;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return best;
  for (int ci = 0; ci < kNumConfigs; ++ci) {
    if (kConfigs[ci].g * kConfigs[ci].c != kCaps[k]) continue;
    for (int block : kBlocks) {
      if (block < kConfigs[ci].g) continue;
      Choice ch;
      ch.cfg = ci;
      ch.block = block;
      int64_t full = (int64_t)g_num_cu * blocks_per_cu(class_kernel(base, kConfigs[ci]), block);
      (void)hipEventRecord(e0, s);
      if (launch_class(base, d_list, sample, full, ch, g_counters + kNumCaps, s) != 0) continue;
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) continue;
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best_ms) {
        best_ms = ms;
        best = ch;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// One thread per system: DE_MAP_LANE and every method but DOPRI5 (the group
// mapping is built for DOPRI5, with every output and event solout).
static bool lane_mapping(const de_ode_batch* b) {
  return b->mapping == DE_MAP_LANE || b->method != DE_METHOD_DOPRI5;
}

// Start of each system's point states in points_y (n + 1 values): the slot
// counts times the state lengths, summed over the systems before it.
// Returns 0, 2 for offsets that decrease or are negative, -1 on a HIP error.
static int point_state_offsets(const OdeArgs& a, const int32_t* host_sizes, int64_t n,
                               std::vector<int64_t>& ypos, hipStream_t s) {
  std::vector<int64_t> off((size_t)n + 1);
  if (hipMemcpyAsync(off.data(), a.pts_offset, sizeof(int64_t) * (n + 1),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  ypos.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t slots = off[i + 1] - off[i];
    if (off[i] < 0 || slots < 0) return 2;
    ypos[i + 1] = ypos[i] + slots * 2 * (int64_t)host_sizes[i];
  }
  return 0;
}

int64_t spring_workspace_bytes(const de_ode_batch* b) {
  if (!b) return 0;
  if (lane_mapping(b)) {
    // The one-lane solver's state-length vectors per system (offsets are
    // device memory here, so the bound uses 512 doubles per system), one more
    // for an event solout and, with a point output, n + 1 state offsets into
    // points_y.
    const int vectors =
        lane_solver(b->method).vectors + (b->event_solout != DE_SOLOUT_DEFAULT ? 1 : 0);
    return (int64_t)vectors * 512 * 8 * b->n +
           (b->points != DE_POINTS_NONE ? 8 * (b->n + 1) : 0);
  }
  return (int64_t)sizeof(int32_t) * b->n + (int64_t)sizeof(unsigned long long) * (kNumCaps + 1) +
         (b->points != DE_POINTS_NONE ? 8 * (b->n + 2) : 0);
}

int launch_spring(const de_ode_batch* b, const int32_t* host_sizes, const OdeArgs& a,
                  int sched, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (ensure_runtime() != 0) return -1;
  const int64_t n = b->n;
  SpringArgs base;
  base.o = a;
  base.sizes = b->sizes;
  base.state_offset = b->state_offset;
  base.param_offset = b->param_offset;
  base.work = nullptr;
  base.pts_ypos = nullptr;
  base.ev_work = nullptr;
  base.ystate = nullptr;
  base.list = nullptr;
  base.list_len = 0;
  // Host-side size scan: validity, total state length, classes.
  int64_t state_len = 0;
  std::vector<std::vector<int32_t>> lists(kNumCaps);
  std::vector<int64_t> masses(kNumCaps, 0);
  const bool comp_event =
      b->event_solout != DE_SOLOUT_DEFAULT && a.ev.fn == kEventComponent;
  for (int64_t i = 0; i < n; ++i) {
    int k = cap_index(host_sizes[i]);
    if (host_sizes[i] < 1 || k < 0) return 1;
    if (comp_event && b->event_component >= 2 * host_sizes[i]) return 3;
    lists[k].push_back((int32_t)i);
    masses[k] += host_sizes[i];
    state_len += 2 * (int64_t)host_sizes[i];
  }
  const bool pts = b->points != DE_POINTS_NONE;
  std::vector<int64_t> ypos;
  if (pts) {
    const int r = point_state_offsets(a, host_sizes, n, ypos, s);
    if (r != 0) return r;
  }
  if (lane_mapping(b)) {
    const LaneSolver ls = lane_solver(b->method);
    const bool events = b->event_solout != DE_SOLOUT_DEFAULT;
    const int64_t vec_len = ((int64_t)ls.vectors + (events ? 1 : 0)) * state_len;
    const int64_t need = vec_len + (pts ? n + 1 : 0);
    // The vectors come from the library's workspace pool (within its budget)
    // and return to it, stream-ordered, after the launch.
    double* work = nullptr;
    if (workspace_alloc((void**)&work, sizeof(double) * need, s) != hipSuccess) return -4;
    base.work = work;
    if (events) base.ev_work = work + (int64_t)ls.vectors * state_len;
    if (pts) {
      int64_t* d_ypos = (int64_t*)(work + vec_len);
      if (hipMemcpyAsync(d_ypos, ypos.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice,
                         s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFreeAsync(work, s);
        return -1;
      }
      base.pts_ypos = d_ypos;
    }
    int64_t grid = (n + kSpringBlock - 1) / kSpringBlock;
    hipLaunchKernelGGL(ls.kernel, dim3((unsigned)grid), dim3(kSpringBlock), 0, s, base);
    const hipError_t e = hipGetLastError();
    return hipFreeAsync(work, s) == hipSuccess && e == hipSuccess ? 0 : -1;
  }
  // Group mapping: upload class lists (one buffer, classes back to back).
  std::vector<int32_t> all;
  std::vector<int64_t> start(kNumCaps + 1, 0);
  for (int k = 0; k < kNumCaps; ++k) {
    start[k] = (int64_t)all.size();
    all.insert(all.end(), lists[k].begin(), lists[k].end());
  }
  start[kNumCaps] = (int64_t)all.size();
  // The point states' offsets follow the lists, 8-byte aligned.
  const size_t list_bytes = (sizeof(int32_t) * (all.size() + 1) + 7) & ~(size_t)7;
  int32_t* d_list = nullptr;
  if (workspace_alloc((void**)&d_list, list_bytes + (pts ? sizeof(int64_t) * (n + 1) : 0), s) !=
      hipSuccess)
    return -4;
  if (hipMemcpyAsync(d_list, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice,
                     s) != hipSuccess ||
      (pts && hipMemcpyAsync((char*)d_list + list_bytes, ypos.data(), sizeof(int64_t) * (n + 1),
                             hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFreeAsync(d_list, s);
    return -1;
  }
  if (pts) base.pts_ypos = (const int64_t*)((const char*)d_list + list_bytes);
  // Choose a configuration per class.
  std::vector<Choice> choice(kNumCaps);
  for (int k = 0; k < kNumCaps; ++k) {
    if (lists[k].empty()) continue;
    Choice ch;
    if (b->mapping == DE_MAP_GROUP && b->group_size > 0) {
      for (int ci = 0; ci < kNumConfigs; ++ci)
        if (kConfigs[ci].g == b->group_size && kConfigs[ci].g * kConfigs[ci].c == kCaps[k])
          ch.cfg = ci;
      if (ch.cfg < 0) {
        // group size too small for this class: largest C that covers the cap
        for (int ci = 0; ci < kNumConfigs; ++ci)
          if (kConfigs[ci].g * kConfigs[ci].c == kCaps[k] &&
              (ch.cfg < 0 || kConfigs[ci].c > kConfigs[ch.cfg].c))
            ch.cfg = ci;
      }
    } else {
      auto it = g_tuned.find(k);
      if (it != g_tuned.end()) {
        ch = it->second;
      } else {
        ch = autotune(base, k, d_list + start[k], (int64_t)lists[k].size(), s);
        if (ch.cfg < 0) {
          (void)hipFreeAsync(d_list, s);
          return -1;
        }
        g_tuned[k] = ch;
      }
    }
    choice[k] = ch;
  }
  // Launch classes concurrently: the three small classes first, then the
  // others largest first.  Every class kernel gets a full-device grid (capped
  // by its number of systems): the kernels are persistent and exit when their
  // class is drained, so blocks of the classes launched later are dispatched
  // as those of earlier ones retire, and the dispatcher balances the classes.
  // The small classes (N <= 16) hold few waves whose long step sequences set
  // their time (latency, not throughput): started first, they run beside the
  // large classes instead of after them, and the largest class then starts
  // next rather than running alone at the end (11.8 against 12.7 ms for the
  // 64k-system ensemble, largest first).  (A fixed share per class left the
  // GPU idle whenever classes waited behind others for one of the few
  // hardware queues a process gets; ordering each class's systems by size was
  // slower, 14.9 ms: it scatters the group's memory accesses.)
  static const int kLaunchOrder[kNumCaps] = {0, 1, 2, 6, 5, 4, 3};
  bool ok = hipEventRecord(g_fork, s) == hipSuccess;
  for (int q = 0; q < kNumCaps && ok; ++q) {
    const int k = kLaunchOrder[q];
    if (lists[k].empty()) continue;
    const Choice& ch = choice[k];
    int64_t share = (int64_t)g_num_cu * blocks_per_cu(class_kernel(base, kConfigs[ch.cfg]), ch.block);
    ok = hipStreamWaitEvent(g_streams[k], g_fork, 0) == hipSuccess &&
         launch_class(base, d_list + start[k], (int64_t)lists[k].size(), share, ch,
                      g_counters + k, g_streams[k]) == 0 &&
         hipEventRecord(g_join[k], g_streams[k]) == hipSuccess &&
         hipStreamWaitEvent(s, g_join[k], 0) == hipSuccess;
  }
  // The list buffer goes back to the pool after the classes (stream order);
  // the call returns once they are done (the class counters are shared).
  ok = hipFreeAsync(d_list, s) == hipSuccess && ok;
  ok = hipStreamSynchronize(s) == hipSuccess && ok;
  return ok ? 0 : -1;
}

// Autotuner introspection: chosen (G, C, block) for size class cap, or -1.
int spring_tuned(int cap, int* g, int* c, int* block) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int k = 0; k < kNumCaps; ++k) {
    if (kCaps[k] != cap) continue;
    auto it = g_tuned.find(k);
    if (it == g_tuned.end()) return -1;
    *g = kConfigs[it->second.cfg].g;
    *c = kConfigs[it->second.cfg].c;
    *block = it->second.block;
    return 0;
  }
  return -1;
}

}  // namespace parity
}  // namespace de
