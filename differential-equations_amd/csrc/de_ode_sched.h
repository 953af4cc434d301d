// Launch schedules of the lane-per-trajectory ODE kernels (plain and
// persistent with the longest-first order), shared by the translation units
// that instantiate lanes.  The includer defines DE_KNS, the namespace the
// kernels are compiled into (parity or fast).
//
//   plain      one thread per trajectory; a wave lives until its slowest lane
//              finishes.
//   persistent a grid sized to the resident wave slots; each lane integrates a
//              trajectory and, when it finishes, takes the next position from a
//              global counter (one atomic per wave, ballot-aggregated), so
//              lanes stay busy while step counts vary ~10x across trajectories.
//              Positions map to trajectories longest-predicted first (see
//              launch_order) or, for PERSISTENT_FIFO, in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "de_lane.h"
#include "de_launch.h"

#include <type_traits>

#ifndef DE_KNS
#error "define DE_KNS before including de_ode_sched.h"
#endif

namespace de {
namespace DE_KNS {

constexpr int kBlock = 256;
constexpr int64_t kOrderMinN = 4096;

template <class Lane>
__global__ __launch_bounds__(kBlock) void ode_plain_kernel(OdeArgs a) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  Lane L;
  if (!L.begin(a, i)) return;
  int32_t st;
  do {
    st = L.step(a, i);
  } while (st < 0);
  L.finish(a, i, st);
}

// Lanes whose step takes a live flag (Dopri5Lane::kLiveStep).
template <class L, class = void>
struct LiveStep : std::false_type {};
template <class L>
struct LiveStep<L, std::void_t<decltype(L::kLiveStep)>> : std::bool_constant<L::kLiveStep> {};

// Waves per SIMD from the lane (Lane::kWavesPerEu): three for most lanes (the
// DOPRI5 lane needs about 165 VGPRs and the hint keeps the allocation under
// the 168-register limit for 3), two for the DOP853 lane with an event solout.
// The arguments live in device memory and are read through a restrict
// pointer: the fields used only when a trajectory starts or ends are then
// loaded inside those branches instead of occupying SGPRs for the whole loop,
// which leaves SGPRs for the step's constants.
template <class Lane>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(Lane::kWavesPerEu))) void
ode_persistent_kernel(const OdeArgs* __restrict__ ap) {
  const OdeArgs& a = *ap;
  Lane L;
  bool have = false;
  bool drained = false;
  int64_t idx = 0;
  const int lane = __lane_id();
  const unsigned long long n = a.n_dev ? (unsigned long long)*a.n_dev : (unsigned long long)a.n;
  for (;;) {
    if (!drained) {
      bool need = !have;
      unsigned long long mask = __ballot(need);
      if (mask) {
        int leader = __ffsll((long long)mask) - 1;
        int cnt = __popcll(mask);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.counter, (unsigned long long)cnt);
        base = __shfl(base, leader);
        if (base + (unsigned long long)cnt >= n) drained = true;
        if (need) {
          unsigned long long my = base + (unsigned long long)__popcll(mask & __lanemask_lt());
          if (my < n) {
            idx = a.order ? (int64_t)a.order[my] : (int64_t)my;
            have = L.begin(a, idx);
          }
        }
      }
    }
    if (!__any(have)) {
      if (drained) break;
      continue;
    }
    if constexpr (LiveStep<Lane>::value) {
      // Every lane of the wave runs the attempt; a lane without work discards
      // it, so the lane state is never carried around a divergent branch.
      const int32_t st = L.step(a, idx, have);
      if (have && st >= 0) {
        L.finish(a, idx, st);
        have = false;
      }
    } else if (have) {
      int32_t st = L.step(a, idx);
      if (st >= 0) {
        L.finish(a, idx, st);
        have = false;
      }
    }
  }
}

static __global__ void zero_counter_kernel(unsigned long long* c) { *c = 0ull; }

static __global__ void store_args_kernel(OdeArgs a, OdeArgs* dst) { *dst = a; }

// HyperplaneCrossingSolout::new (src/solout/hyperplane.rs:126-132): the normal
// times 1/|normal| when |normal| > f64::EPSILON, |normal| = sqrt of the squares
// summed from zero.  Points the stored arguments at the normalised copy.
template <int D>
__global__ void normalise_normal_kernel(OdeArgs* args, double* out) {
  const double* nrm = args->ev.normal;
  double sq = 0.0;
  for (int c = 0; c < D; ++c) sq = sq + nrm[c] * nrm[c];
  const double norm = sqrt(sq);
  const double inv = 1.0 / norm;
  for (int c = 0; c < D; ++c) out[c] = norm > 2.220446049250313080847e-16 ? nrm[c] * inv : nrm[c];
  args->ev.normal = out;
}

// ---- Deferred event roots -------------------------------------------------
// Root pass: each block takes the brackets of kBlock consecutive trajectories,
// compacted by a scan of their counts in LDS, one bracket per thread at a
// time, so waves stay full whatever the per-trajectory counts.
template <class Lane>
__global__ __launch_bounds__(kBlock) void event_root_kernel(const OdeArgs* __restrict__ ap,
                                                            int64_t count) {
  const OdeArgs& a = *ap;
  __shared__ int32_t off[kBlock + 1];
  const int t = threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.x * kBlock;  // position in the slice
  int32_t cnt = 0;
  if (j0 + t < count) cnt = max(a.ev.bcount[j0 + t], 0);
  off[t + 1] = cnt;
  if (t == 0) off[0] = 0;
  __syncthreads();
  for (int d = 1; d < kBlock; d <<= 1) {  // inclusive scan of off[1..kBlock]
    const int32_t add = t >= d ? off[t + 1 - d] : 0;
    __syncthreads();
    off[t + 1] += add;
    __syncthreads();
  }
  const int32_t total = off[kBlock];
  for (int32_t j = t; j < total; j += kBlock) {
    int lo = 0, hi = kBlock;  // the last lo with off[lo] <= j
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (off[mid] <= j) lo = mid;
      else hi = mid;
    }
    const int64_t slot = (j0 + lo) * a.ev.bmax + (j - off[lo]);
    Lane::entry_root(a, a.ev.base + j0 + lo, a.ev.brk + slot * a.ev.width,
                     a.ev.res + slot * kResWidth);
  }
}

// Bookkeeping pass: one thread per trajectory; the trajectories to solve
// again inline are appended to `rerun`.
template <class Lane>
__global__ __launch_bounds__(kBlock) void event_book_kernel(const OdeArgs* __restrict__ ap,
                                                            int64_t count, uint32_t* rerun,
                                                            uint32_t* nrerun) {
  const OdeArgs& a = *ap;
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= count) return;
  const int64_t i = a.ev.base + j;
  if (!Lane::entry_book(a, i)) rerun[atomicAdd(nrerun, 1u)] = (uint32_t)i;
}

static __global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }
static __global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

// order[j] = first + j: a slice's trajectories in index order.
static __global__ __launch_bounds__(kBlock) void iota_kernel(uint32_t* order, int64_t first,
                                                             int64_t count) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < count) order[j] = (uint32_t)(first + j);
}

// Arguments of the inline solve of the listed trajectories.
static __global__ void rerun_args_kernel(const OdeArgs* src, OdeArgs* dst, const uint32_t* order,
                                         const uint32_t* count) {
  OdeArgs r = *src;
  r.ev.brk = nullptr;
  r.ev.res = nullptr;
  r.ev.bcount = nullptr;
  r.order = order;
  r.n_dev = count;
  *dst = r;
}

// Deferred entries per trajectory: max_events + half again (at least 32 more).
// The regions (entries and result records) of one pass stay within 3/4 of
// the workspace budget (48 GiB by default): when those of the whole ensemble
// would not, the
// ensemble is solved in slices of consecutive trajectories, each a pass of
// its own (integration, roots, bookkeeping, inline reruns), of at least
// kMinDeferSlice trajectories (enough to fill the GPU); only if even that
// does not fit is the number of entries halved, and below 16 the roots are
// found inline.  The environment variable DE_EVENT_DEFER_SLICE, when set to
// a positive count, fixes the slice size instead (for tests and tuning).
constexpr int64_t kMinDeferSlice = int64_t(1) << 18;

template <class Lane>
static int64_t defer_slots(const OdeArgs& a, int64_t* slice) {
  *slice = a.n;
  if constexpr (Lane::kEv == 0) {
    return 0;
  } else {
    if (a.ev.terminate != 0 || a.n > 0x7fffffffLL) return 0;
    int64_t want = std::min<int64_t>(a.n, kMinDeferSlice);
    int64_t forced = 0;
    if (const char* f = getenv("DE_EVENT_DEFER_SLICE")) forced = atoll(f);
    if (forced > 0) want = std::min<int64_t>(a.n, forced);
    const size_t budget = (size_t)(workspace_budget() / 4 * 3);  // the regions of one pass
    for (int64_t bmax = (int64_t)a.ev.max_events + std::max<int64_t>(a.ev.max_events / 2, 32);
         bmax >= 16; bmax /= 2) {
      const size_t per = (size_t)bmax * (Lane::kEntryWidth + kResWidth) * sizeof(double);
      int64_t fit = (int64_t)std::min<size_t>(budget / per, (size_t)a.n);
      if (fit < want) continue;
      if (forced > 0) fit = want;
      const int64_t parts = (a.n + fit - 1) / fit;
      *slice = (a.n + parts - 1) / parts;
      return bmax;
    }
    return 0;
  }
}

// ---- Longest-first order for the persistent schedule ----------------------
// A counting sort of trajectories by cost_bucket(predicted_cost): the resident
// lanes then start the longest trajectories first and finish on the short
// ones, so few lanes idle while their wave waits for its last trajectory.
// Order within a bucket is arbitrary (it follows atomic arrival); results do
// not depend on it, each trajectory being solved on its own.
constexpr int kBuckets = 256;
static_assert(kBlock == kBuckets, "cost_kernel and scatter_kernel use one thread per bucket");

static __global__ __launch_bounds__(kBuckets) void zero_hist_kernel(uint32_t* hist) {
  hist[threadIdx.x] = 0u;
}

template <class Lane>
__global__ __launch_bounds__(kBlock) void cost_kernel(OdeArgs a, int64_t first, int64_t count,
                                                      uint8_t* bucket, uint32_t* hist) {
  __shared__ uint32_t h[kBuckets];
  h[threadIdx.x] = 0u;  // kBlock == kBuckets
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < count) {
    const uint32_t b = cost_bucket(Lane::predicted_cost(a, first + j));
    bucket[j] = (uint8_t)b;
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Exclusive prefix sum of the 256 counts, in place (one block).
static __global__ __launch_bounds__(kBuckets) void scan_hist_kernel(uint32_t* hist) {
  __shared__ uint32_t v[kBuckets];
  const int t = threadIdx.x;
  v[t] = hist[t];
  __syncthreads();
  for (int d = 1; d < kBuckets; d <<= 1) {
    const uint32_t add = t >= d ? v[t - d] : 0u;
    __syncthreads();
    v[t] += add;
    __syncthreads();
  }
  hist[t] = v[t] - hist[t];
}

static __global__ __launch_bounds__(kBlock) void scatter_kernel(int64_t first, int64_t count,
                                                                const uint8_t* bucket,
                                                                uint32_t* offset,
                                                                uint32_t* order) {
  __shared__ uint32_t cnt[kBuckets], base[kBuckets];
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t b = 0u, local = 0u;
  if (j < count) {
    b = bucket[j];
    local = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  if (cnt[threadIdx.x]) base[threadIdx.x] = atomicAdd(&offset[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (j < count) order[base[b] + local] = (uint32_t)(first + j);
}

// Fills order[0 .. count) with trajectories first .. first + count - 1,
// longest predicted first, using `bucket` (count bytes) and `hist` (256 words).
template <class Lane>
static void launch_order(const OdeArgs& a, int64_t first, int64_t count, uint8_t* bucket,
                         uint32_t* hist, uint32_t* order, hipStream_t s) {
  const unsigned grid = (unsigned)((count + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(zero_hist_kernel, dim3(1), dim3(kBuckets), 0, s, hist);
  hipLaunchKernelGGL(cost_kernel<Lane>, dim3(grid), dim3(kBlock), 0, s, a, first, count, bucket,
                     hist);
  hipLaunchKernelGGL(scan_hist_kernel, dim3(1), dim3(kBuckets), 0, s, hist);
  hipLaunchKernelGGL(scatter_kernel, dim3(grid), dim3(kBlock), 0, s, first, count,
                     (const uint8_t*)bucket, hist, order);
}

template <class Lane>
static int launch(const OdeArgs& a, int schedule, hipStream_t s) {
  if (a.n <= 0) return 0;
  if constexpr (Lane::kEv == 0) {  // event lanes run persistent only
    if (schedule == kSchedPlain) {
      int64_t grid = (a.n + kBlock - 1) / kBlock;
      hipLaunchKernelGGL(ode_plain_kernel<Lane>, dim3((unsigned)grid), dim3(kBlock), 0, s, a);
      return hipGetLastError() == hipSuccess ? 0 : -1;
    }
  }
  // Resident blocks of the persistent kernel: queried once per lane type
  // (a function-local static is initialised exactly once, also when several
  // threads launch at the same time).
  struct Occupancy {
    int blocks_per_cu = 0, num_cu = 0;
  };
  static const Occupancy occ = [] {
    Occupancy o;
    int dev = 0, cu = 0, b = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ode_persistent_kernel<Lane>, kBlock, 0) !=
            hipSuccess)
      return o;
    o.num_cu = cu;
    o.blocks_per_cu = b > 0 ? b : 1;
    return o;
  }();
  if (occ.num_cu <= 0) return -1;
  int64_t want = (int64_t)occ.num_cu * occ.blocks_per_cu;
  int64_t need = (a.n + kBlock - 1) / kBlock;
  int64_t grid = want < need ? want : need;
  // Stream-ordered workspace: the launch arguments; for event lanes 256 bytes
  // for the normalised hyperplane normal and, with deferred roots, a second
  // argument block, the per-trajectory bracket counts, the rerun list and its
  // count, the slice's trajectory count and the bracket regions; for the
  // longest-first order (schedule PERSISTENT, batches of a few thousand
  // trajectories or more) or a sliced solve, 256 bucket counts, the order and
  // one bucket byte per trajectory of a slice.
  bool ordered = schedule == kSchedPersistent && a.n >= kOrderMinN && a.n <= 0xffffffffLL;
  const size_t args_bytes = (sizeof(OdeArgs) + 255) & ~(size_t)255;
  const size_t ev_bytes = Lane::kEv != 0 ? 256 : 0;
  int64_t slice = a.n;
  int64_t bmax = defer_slots<Lane>(a, &slice);
  void* work = nullptr;
  size_t order_bytes = 0, list_bytes = 0, defer_bytes = 0;
  for (;;) {
    const bool listed = ordered || slice < a.n;
    order_bytes = listed ? ((kBuckets * sizeof(uint32_t) + (size_t)slice * (sizeof(uint32_t) + 1) +
                             255) & ~(size_t)255)
                         : 0;
    list_bytes = ((size_t)slice * sizeof(uint32_t) + 255) & ~(size_t)255;
    defer_bytes = bmax > 0 ? args_bytes + 2 * list_bytes + 256 +
                                 (size_t)slice * (size_t)bmax * (Lane::kEntryWidth + kResWidth) *
                                     sizeof(double)
                           : 0;
    const size_t total = args_bytes + ev_bytes + order_bytes + defer_bytes;
    if (a.scratch) {
      if (total <= (size_t)a.scratch_bytes) {
        work = a.scratch;
        break;
      }
    } else if (workspace_alloc(&work, total, s) == hipSuccess) {
      break;
    } else {
      (void)hipGetLastError();
    }
    if (bmax > 0) {
      bmax = 0;  // no room for the bracket regions: roots inline, one pass
      slice = a.n;
    } else if (ordered) {
      ordered = false;  // no room for the longest-first order: index order
    } else {
      return -1;
    }
  }
  const bool sliced = slice < a.n;
  char* w = (char*)work;
  OdeArgs* d_args = (OdeArgs*)w;
  uint32_t* hist = order_bytes ? (uint32_t*)(w + args_bytes + ev_bytes) : nullptr;
  uint32_t* order = hist ? hist + kBuckets : nullptr;
  uint8_t* bucket = order ? (uint8_t*)(order + slice) : nullptr;
  OdeArgs* d_rerun_args = nullptr;
  uint32_t *bcount = nullptr, *rerun = nullptr, *nrerun = nullptr, *nslice = nullptr;
  double* res = nullptr;
  if (bmax > 0) {
    char* d = w + args_bytes + ev_bytes + order_bytes;
    d_rerun_args = (OdeArgs*)d;
    bcount = (uint32_t*)(d + args_bytes);
    rerun = (uint32_t*)(d + args_bytes + list_bytes);
    nrerun = (uint32_t*)(d + args_bytes + 2 * list_bytes);
    nslice = nrerun + 1;
    res = (double*)(d + args_bytes + 2 * list_bytes + 256);
  }
  for (int64_t first = 0; first < a.n; first += slice) {
    const int64_t count = std::min<int64_t>(slice, a.n - first);
    OdeArgs pa = a;
    pa.order = nullptr;
    pa.n_dev = nullptr;
    if (order) {
      if (ordered && count >= kOrderMinN)
        launch_order<Lane>(a, first, count, bucket, hist, order, s);
      else
        hipLaunchKernelGGL(iota_kernel, dim3((unsigned)((count + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, order, first, count);
      pa.order = order;
    }
    if (sliced) {
      hipLaunchKernelGGL(set_u32_kernel, dim3(1), dim3(1), 0, s, nslice, (uint32_t)count);
      pa.n_dev = nslice;
    }
    if constexpr (Lane::kEv != 0) {
      pa.ev.brk = nullptr;
      pa.ev.res = nullptr;
      pa.ev.bcount = nullptr;
      pa.ev.base = first;
      if (bmax > 0) {
        pa.ev.res = res;
        pa.ev.brk = res + (size_t)slice * (size_t)bmax * kResWidth;
        pa.ev.bcount = (int32_t*)bcount;
        pa.ev.bmax = (int32_t)bmax;
        pa.ev.width = Lane::kEntryWidth;
      }
    }
    hipLaunchKernelGGL(store_args_kernel, dim3(1), dim3(1), 0, s, pa, d_args);
    if constexpr (Lane::kEv != 0) {
      static_assert(Lane::D * sizeof(double) <= 256, "normal buffer");
      if (a.ev.solout == kSoloutHyperplane)
        hipLaunchKernelGGL(normalise_normal_kernel<Lane::D>, dim3(1), dim3(1), 0, s, d_args,
                           (double*)(w + args_bytes));
    }
    const int64_t blocks = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
    hipLaunchKernelGGL(ode_persistent_kernel<Lane>, dim3((unsigned)std::min(grid, blocks)),
                       dim3(kBlock), 0, s, (const OdeArgs*)d_args);
    if constexpr (Lane::kEv != 0) {
      if (bmax > 0) {
        // Roots of every stored bracket, the push rules per trajectory, and the
        // inline solve of the trajectories listed for it (on the slice's grid;
        // the count is known only on the device).
        const unsigned gx = (unsigned)blocks;
        hipLaunchKernelGGL(event_root_kernel<Lane>, dim3(gx), dim3(kBlock), 0, s,
                           (const OdeArgs*)d_args, count);
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, nrerun);
        hipLaunchKernelGGL(event_book_kernel<Lane>, dim3(gx), dim3(kBlock), 0, s,
                           (const OdeArgs*)d_args, count, rerun, nrerun);
        hipLaunchKernelGGL(rerun_args_kernel, dim3(1), dim3(1), 0, s, (const OdeArgs*)d_args,
                           d_rerun_args, (const uint32_t*)rerun, (const uint32_t*)nrerun);
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
        hipLaunchKernelGGL(ode_persistent_kernel<Lane>, dim3((unsigned)std::min(grid, blocks)),
                           dim3(kBlock), 0, s, (const OdeArgs*)d_rerun_args);
      }
    }
  }
  if (!a.scratch && hipFreeAsync(work, s) != hipSuccess) return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace DE_KNS
}  // namespace de
