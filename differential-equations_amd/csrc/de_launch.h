// Host-side launch entry points of the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "de_lane.h"

namespace de {

// Stream-ordered workspace of the ODE launches, from a pool of the library's
// own that keeps freed blocks instead of handing them back to the driver at
// the next synchronisation: fresh device memory costs ~18 ms per GB to obtain
// again (a 45 GB event workspace measured 809 ms), which a repeated solve
// would pay every time.  One pool per device, created on first use.
//
// Device-memory budget of the workspace: one number (default 64 GiB of the
// 288 GiB HBM, de_set_workspace_budget) bounds every workspace allocation,
// the freed blocks the pool keeps, and (at 3/4 of it) the bracket regions of
// one deferred-root event pass.  A failed allocation first trims the pool's
// kept blocks and tries once more.  The serving arenas are sized by their
// own creation argument; callers' buffers are theirs.
constexpr uint64_t kWorkspaceBudgetDefault = uint64_t(64) << 30;
constexpr int kMaxDev = 64;

struct WorkspacePools {
  std::mutex mu;
  hipMemPool_t pools[kMaxDev] = {};
  uint64_t budget = kWorkspaceBudgetDefault;
};

inline WorkspacePools& workspace_pools() {
  static WorkspacePools w;
  return w;
}

inline uint64_t workspace_budget() {
  WorkspacePools& w = workspace_pools();
  std::lock_guard<std::mutex> lk(w.mu);
  return w.budget;
}

// New budget for every device's pool (its kept blocks are trimmed to it).
inline hipError_t set_workspace_budget(uint64_t bytes) {
  WorkspacePools& w = workspace_pools();
  std::lock_guard<std::mutex> lk(w.mu);
  w.budget = bytes;
  for (int d = 0; d < kMaxDev; ++d) {
    if (!w.pools[d]) continue;
    uint64_t keep = bytes;
    hipError_t e = hipMemPoolSetAttribute(w.pools[d], hipMemPoolAttrReleaseThreshold, &keep);
    if (e == hipSuccess) e = hipMemPoolTrimTo(w.pools[d], bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Hand every kept (freed) block of every device's pool back to the driver.
inline hipError_t workspace_trim() {
  WorkspacePools& w = workspace_pools();
  std::lock_guard<std::mutex> lk(w.mu);
  for (int d = 0; d < kMaxDev; ++d) {
    if (!w.pools[d]) continue;
    hipError_t e = hipMemPoolTrimTo(w.pools[d], 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

inline hipError_t workspace_alloc(void** p, size_t bytes, hipStream_t s) {
  WorkspacePools& w = workspace_pools();
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipMemPool_t pool;
  {
    std::lock_guard<std::mutex> lk(w.mu);
    if (bytes > w.budget) return hipErrorOutOfMemory;
    if (dev < 0 || dev >= kMaxDev) return hipMallocAsync(p, bytes, s);
    if (!w.pools[dev]) {
      hipMemPoolProps props = {};
      props.allocType = hipMemAllocationTypePinned;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      hipMemPool_t created = nullptr;
      e = hipMemPoolCreate(&created, &props);
      if (e != hipSuccess) return e;
      uint64_t keep = w.budget;
      e = hipMemPoolSetAttribute(created, hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) {
        (void)hipMemPoolDestroy(created);
        return e;
      }
      w.pools[dev] = created;
    }
    pool = w.pools[dev];
  }
  e = hipMallocFromPoolAsync(p, bytes, pool, s);
  if (e == hipSuccess) return e;
  (void)hipGetLastError();
  // The pool's kept blocks may be what is missing: release them and retry.
  (void)hipMemPoolTrimTo(pool, 0);
  return hipMallocFromPoolAsync(p, bytes, pool, s);
}

enum : int { kSchedAuto = 0, kSchedPlain = 1, kSchedPersistent = 2, kSchedPersistentFifo = 3 };
enum : int { kMethodDopri5 = 0, kMethodRk4 = 1, kMethodEuler = 2, kMethodDop853 = 3 };
// The embedded family (DE_METHOD_RKF45 .. DE_METHOD_RKV989E).
enum : int {
  kMethodRkf45 = 4, kMethodCashKarp = 5, kMethodRkv655e = 6, kMethodRkv656e = 7,
  kMethodRkv766e = 8, kMethodRkv767e = 9, kMethodRkv877e = 10, kMethodRkv878e = 11,
  kMethodRkv988e = 12, kMethodRkv989e = 13
};
enum : int { kSchemeEulerMaruyama = 0, kSchemeMilstein = 1 };
enum : int { kNoisePhilox = 0, kNoiseTable = 1, kNoiseConstant = 2 };

struct SdeArgs {
  int64_t n;
  int32_t dim;
  int32_t noise;
  const double* params;
  const double* y0;
  double t0, tf, h0, h_min, h_max;
  int64_t max_steps;
  uint64_t seed, path_base;
  const double* noise_table;
  int64_t noise_steps;
  double noise_constant;
  int32_t path_stride;
  double* y_final;
  double* t_final;
  int64_t* steps;
  int64_t* evals;
  int32_t* status;
  double* path;
  double* path_t;
};

// Solution::to_csv of every trajectory (kernels_csv.hip): the solve's
// geometry and result arrays, device pointers.
enum : int { kCsvFinal = 0, kCsvTeval = 1, kCsvEvents = 2, kCsvPoints = 3 };
constexpr int kCsvMaxDim = 8;
struct CsvArgs {
  int64_t n;
  int32_t dim;
  int32_t source;    // kCsv*
  int32_t max_rows;  // points per trajectory at most: 1, n_eval or max_events
  const double* t0;  // per trajectory, or null: t0s
  const double* tf;
  double t0s, tfs;
  const double* t_eval;
  int32_t n_eval;
  const double* t_final;
  const double* y_final;
  const double* dense;
  const int32_t* n_dense;
  const double* event_t;
  const double* event_y;
  const int32_t* n_events;
  const int64_t* pts_offset;  // kCsvPoints
  const double* pts_t;
  const double* pts_y;
  const int64_t* n_pts;
};
// Bytes of the longest line: dim + 1 numbers of at most 24 characters,
// their separators and the newline (the header is shorter).
constexpr int csv_max_slot_bytes(int dim) { return (dim + 1) * 24 + dim + 1; }
// offsets[0 .. n] and, when out is not null, the documents into out[0 .. cap)
// (a block of lines that would pass cap is not written).  0, or -1 on a HIP
// error (hipGetLastError has it).
int launch_csv(const CsvArgs& a, int64_t* offsets, char* out, int64_t cap, hipStream_t s);
// Device self-check of the number formatter: the text of x[i] in out[32 i ..],
// its length in len[i] (negative: a loop cap was reached).
int launch_fmt_diag(const double* x, char* out, int32_t* len, int64_t n, hipStream_t s);

namespace parity {
int launch_ode_lane(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_event(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_crossing(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_sde(int family, int scheme, const SdeArgs& a, hipStream_t s);
// The embedded family, two tableaux per translation unit.
int launch_ode_erk0(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_erk1(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_erk2(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_erk3(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
int launch_ode_erk4(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
// The embedded family with an event solout: group g, kind 1 (event) / 2 (crossings).
#define DE_ERK_EV_DECL(g, k)                                                          \
  int launch_ode_erk##g##_ev##k(int family, int method, const OdeArgs& a, int schedule, \
                                hipStream_t s);
DE_ERK_EV_DECL(0, 1) DE_ERK_EV_DECL(1, 1) DE_ERK_EV_DECL(2, 1) DE_ERK_EV_DECL(3, 1)
DE_ERK_EV_DECL(4, 1) DE_ERK_EV_DECL(0, 2) DE_ERK_EV_DECL(1, 2) DE_ERK_EV_DECL(2, 2)
DE_ERK_EV_DECL(3, 2) DE_ERK_EV_DECL(4, 2)
#undef DE_ERK_EV_DECL
}  // namespace parity
namespace fast {
int launch_ode_lane(int family, int method, const OdeArgs& a, int schedule, hipStream_t s);
}  // namespace fast

}  // namespace de
