// Serving: the device side shared by the request loop (server.hip) and the
// translation units that build its lazily created pools (server_pools.hip):
// work queues, request descriptors, solver slots and the chunk kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "de_erk_lane.h"
#include "de_spring_solvers.h"

namespace de {
namespace serve {

constexpr int kMaxReq = 8192;     // requests resident at once
constexpr int kMaxRanges = 1 << 20;
constexpr int kChunk = 256;       // attempts per slot per launch (no backlog)
constexpr int kMaxChunk = 4096;   // attempts per slot per launch under a backlog
constexpr int kBlock = 256;

struct Range {
  int64_t start;       // first ticket
  int32_t req;         // request slot
  int32_t first;       // first trajectory (index into list when list != null)
  int32_t count;
  int32_t pad;
  const int32_t* list; // spring: system indices of this size class
};

struct Queue {
  unsigned long long next;  // next ticket to hand out
  int64_t published;        // tickets below this are backed by a range
  int64_t nranges;
};

struct ReqDev {
  SpringArgs a;  // lane families use a.o only
  int32_t remaining;
  int32_t pad;
};

template <class L>
struct LaneSlot {
  L s;
  int64_t ticket;
  int32_t req, traj, state;  // state 0 idle, 1 holding ticket, 2 running
};

DE_DEV int find_range(const Range* r, int64_t n, int64_t ticket) {
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (r[mid].start <= ticket) lo = mid;
    else hi = mid - 1;
  }
  return (int)lo;
}

template <class L>
__global__ __launch_bounds__(kBlock) void chunk_lane(Queue* q, const Range* ranges, ReqDev* reqs,
                                                    LaneSlot<L>* slots, int nslots, int k_iters) {
  const int slot = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = slot < nslots;
  const int lane = __lane_id();
  LaneSlot<L> S;
  if (valid) S = slots[slot];
  else S.state = -1;
  const int64_t published = *(volatile int64_t*)&q->published;
  const int64_t nranges = *(volatile int64_t*)&q->nranges;
  for (int it = 0; it < k_iters; ++it) {
    bool need = S.state == 0;
    unsigned long long mask = __ballot(need);
    if (mask) {
      int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(&q->next, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        S.ticket = (int64_t)(base + __popcll(mask & __lanemask_lt()));
        S.state = 1;
      }
    }
    if (S.state == 1 && S.ticket < published) {
      int r = find_range(ranges, nranges, S.ticket);
      S.req = ranges[r].req;
      S.traj = ranges[r].first + (int32_t)(S.ticket - ranges[r].start);
      bool started;
      if constexpr (std::is_base_of<VecCommon, L>::value)
        started = S.s.init(&reqs[S.req].a, S.traj);  // keeps the descriptor's address
      else
        started = S.s.init(reqs[S.req].a.o, S.traj);
      if (started) {
        S.state = 2;
      } else {
        atomicSub(&reqs[S.req].remaining, 1);
        S.state = 0;
      }
    }
    bool act = S.state == 2;
    if (!__any(act)) {
      if (!__any(S.state == 0 || (S.state == 1 && S.ticket < published))) break;
      continue;
    }
    bool finished = false;
    if constexpr (std::is_base_of<VecCommon, L>::value)
      finished = act && S.s.advance();
    else
      finished = act && S.s.advance(reqs[S.req].a.o, S.traj);
    if (finished) {
      atomicSub(&reqs[S.req].remaining, 1);
      S.state = 0;
    }
  }
  if (valid) slots[slot] = S;
}

// Occupancy hints: the rarely taken double-double pow fallback raises the
// peak register count past a wave-slot boundary; the hints trade that for a
// few spills outside the step block.
template <int G, int C, bool OUT = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(C == 1 ? 3 : 2))) void
chunk_group(Queue* q, const Range* ranges, ReqDev* reqs, LaneSlot<GroupDopri5<G, C, OUT>>* slots,
            int nslots, int k_iters) {
  const int slot = blockIdx.x * kBlock + threadIdx.x;  // one per lane; G lanes per system
  const bool valid = slot < nslots;
  const int lane = __lane_id();
  const int l = lane & (G - 1);
  const int leader_lane = lane - l;
  LaneSlot<GroupDopri5<G, C, OUT>> S;
  if (valid) S = slots[slot];
  else S.state = -1;
  const int64_t published = *(volatile int64_t*)&q->published;
  const int64_t nranges = *(volatile int64_t*)&q->nranges;
  for (int it = 0; it < k_iters; ++it) {
    bool need = S.state == 0;
    unsigned long long mask = __ballot(need && l == 0);
    if (mask) {
      int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(&q->next, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        unsigned long long below =
            leader_lane == 0 ? 0ull : (mask & ((1ull << leader_lane) - 1ull));
        S.ticket = (int64_t)(base + __popcll(below));
        S.state = 1;
      }
    }
    if (S.state == 1 && S.ticket < published) {
      int r = find_range(ranges, nranges, S.ticket);
      S.req = ranges[r].req;
      S.traj = ranges[r].list[ranges[r].first + (int32_t)(S.ticket - ranges[r].start)];
      if (S.s.init(reqs[S.req].a, S.traj, l)) {
        S.state = 2;
      } else {
        if (l == 0) atomicSub(&reqs[S.req].remaining, 1);
        S.state = 0;
      }
    }
    bool act = S.state == 2;
    if (!__any(act)) {
      if (!__any(S.state == 0 || (S.state == 1 && S.ticket < published))) break;
      continue;
    }
    if (act && S.s.advance(reqs[S.req].a)) {
      if (l == 0) atomicSub(&reqs[S.req].remaining, 1);
      S.state = 0;
    }
  }
  if (valid) slots[slot] = S;
}

// Reset of a drained pool (no request with work in it): slots that hold an
// unbacked ticket (state 1) become idle and the ticket and range counters
// restart from zero, so the range table is reused.
template <class Slot>
__global__ void reset_pool_kernel(Slot* s, int n, Queue* q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && s[i].state == 1) {
    s[i].state = 0;
    s[i].ticket = 0;
  }
  if (i == 0) {
    q->next = 0;
    q->published = 0;
    q->nranges = 0;
  }
}

template <class Slot>
__global__ void init_slots(Slot* s, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    s[i].state = 0;
    s[i].ticket = 0;
    s[i].req = 0;
    s[i].traj = 0;
  }
}

// ---------------------------------------------------------------- host -----
typedef void (*ChunkFn)(Queue*, const Range*, ReqDev*, void*, int, int);

struct Family {
  const char* name = nullptr;
  int cap = 0;     // spring: class cap; 0 for lane families
  int g = 1;       // lanes per system (1 for lane families)
  size_t slot_bytes = 0;
  int nslots = 0;  // lanes
  void* slots = nullptr;  // null: a lazy pool not made yet
  Queue* q = nullptr;     // device
  Range* ranges = nullptr;  // device
  bool lazy = false;      // made on first use, its memory carved from the arena
  int vectors = 0;        // one-lane spring solvers: workspace vectors per state
  std::vector<Range> host_ranges;
  int64_t published = 0;
  int64_t active_tickets = 0;  // published but possibly unfinished
  int active_reqs = 0;         // active requests with work in this pool
  int64_t max_ranges = 0;      // capacity of the range table
  bool dirty = false;          // published work not yet committed to the device
  hipStream_t stream = nullptr;
  void (*launch)(Family&, ReqDev*, int, hipStream_t) = nullptr;
  void (*reset)(Family&, hipStream_t) = nullptr;
  void (*init)(Family&, hipStream_t) = nullptr;
};

template <class Slot>
void reset_pool(Family& f, hipStream_t s) {
  hipLaunchKernelGGL(reset_pool_kernel<Slot>, dim3((f.nslots + 255) / 256), dim3(256), 0, s,
                     (Slot*)f.slots, f.nslots, f.q);
}

template <class L>
void launch_lane(Family& f, ReqDev* reqs, int k, hipStream_t s) {
  int grid = (f.nslots + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(chunk_lane<L>, dim3(grid), dim3(kBlock), 0, s, f.q, f.ranges, reqs,
                     (LaneSlot<L>*)f.slots, f.nslots, k);
}

template <int G, int C, bool OUT = false>
void launch_group(Family& f, ReqDev* reqs, int k, hipStream_t s) {
  int grid = (f.nslots + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((chunk_group<G, C, OUT>), dim3(grid), dim3(kBlock), 0, s, f.q, f.ranges,
                     reqs, (LaneSlot<GroupDopri5<G, C, OUT>>*)f.slots, f.nslots, k);
}

// A pool's slot type: its size, slot initialisation and reset.
template <class Slot>
void init_pool(Family& f, hipStream_t s) {
  hipLaunchKernelGGL(init_slots<Slot>, dim3((f.nslots + 255) / 256), dim3(256), 0, s,
                     (Slot*)f.slots, f.nslots);
}
template <class Slot>
void set_pool_kind(Family& f) {
  f.slot_bytes = sizeof(Slot);
  f.reset = reset_pool<Slot>;
  f.init = init_pool<Slot>;
}

// Pools made on first use, by method: the embedded family on Lorenz and Van
// der Pol lanes (group g of two tableaux, server_pools.hip with
// DE_SERVE_GROUP = g), and DOP853, RK4 and Euler lanes and the one-lane
// spring-chain solvers (group 5).  ev: the lane's event solout (0 none, 1
// EventWrappedSolout, 2 crossing / hyperplane; spring chains choose at run
// time).  Sets f's slot kind and launcher; false when (family, method, ev) is
// not in the group.
bool lazy_pool_kind_0(int family, int method, int ev, Family& f);
bool lazy_pool_kind_1(int family, int method, int ev, Family& f);
bool lazy_pool_kind_2(int family, int method, int ev, Family& f);
bool lazy_pool_kind_3(int family, int method, int ev, Family& f);
bool lazy_pool_kind_4(int family, int method, int ev, Family& f);
bool lazy_pool_kind_5(int family, int method, int ev, Family& f);
// Spring chains with DOPRI5 and an event solout on the register group mapping
// of size class cls (caps 4, 8, ..., 256): slot kind, launcher, cap and G.
bool spring_event_class_kind(int cls, Family& f);

}  // namespace serve
}  // namespace de
