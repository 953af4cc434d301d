// C-ABI of the ensemble integrator (include/de_api.h, include/de_workload.h):
// descriptor validation, host-memory staging, schedule selection and the
// synthetic workload generators.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/de_api.h"
#include "../../include/de_workload.h"
#include "de_launch.h"
#include "de_noise.h"
#include "de_spring.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(DE_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

const int kFamilyDim[] = {3, 2, 0, 1, 1, 2, 1, 3};
const int kFamilyParams[] = {3, 1, 0, 1, 2, 1, 2, 0};

// Work counters for persistent launches: a ring of device slots, one per call
// in flight (calls on different streams must not share one).
constexpr int kCounterSlots = 1024;
unsigned long long* g_counters = nullptr;
std::atomic<int> g_counter_next(0);
std::mutex g_counter_mu;

int counter_slot(unsigned long long** out) {
  std::lock_guard<std::mutex> lk(g_counter_mu);
  if (!g_counters) {
    hipError_t e = hipMalloc(&g_counters, sizeof(unsigned long long) * kCounterSlots);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(counters)");
  }
  int k = g_counter_next.fetch_add(1) % kCounterSlots;
  *out = g_counters + k;
  return DE_OK;
}

// Device staging of host-memory descriptors.
struct Staging {
  std::vector<void*> bufs;
  hipError_t err = hipSuccess;
  ~Staging() {
    for (void* p : bufs) (void)hipFree(p);  // teardown: a failure here has no one to report to
  }
  template <class T>
  T* in(const T* host, size_t count) {
    if (!host || count == 0 || err != hipSuccess) return nullptr;
    void* d = nullptr;
    err = hipMalloc(&d, count * sizeof(T));
    if (err != hipSuccess) return nullptr;
    bufs.push_back(d);
    err = hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice);
    return (T*)d;
  }
  template <class T>
  T* out(T* host, size_t count) {
    if (!host || count == 0 || err != hipSuccess) return nullptr;
    void* d = nullptr;
    err = hipMalloc(&d, count * sizeof(T));
    if (err != hipSuccess) return nullptr;
    bufs.push_back(d);
    return (T*)d;
  }
  template <class T>
  void back(T* host, const T* dev, size_t count) {
    if (!host || !dev || count == 0 || err != hipSuccess) return;
    err = hipMemcpy(host, dev, count * sizeof(T), hipMemcpyDeviceToHost);
  }
};

__global__ void fill_nan_kernel(double* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = __longlong_as_double(0x7ff8000000000000LL);
}

int fill_nan(double* p, int64_t n, hipStream_t s) {
  if (!p || n <= 0) return DE_OK;
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(fill_nan_kernel, dim3((unsigned)grid), dim3(256), 0, s, p, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DE_OK : hip_fail(e, "fill_nan");
}

int validate_ode(const de_ode_batch* b, const de_ode_result* o) {
  if (!b || !o) return fail(DE_ERR_INVALID, "null descriptor");
  if (b->n < 0) return fail(DE_ERR_INVALID, "n < 0");
  if (b->family < 0 || b->family > DE_FAMILY_ROBERTSON)
    return fail(DE_ERR_INVALID, "unknown family %d", b->family);
  if (b->method < DE_METHOD_DOPRI5 || b->method > DE_METHOD_RKV989E)
    return fail(DE_ERR_INVALID, "unknown method %d", b->method);
  if (b->family != DE_FAMILY_SPRING_CHAIN && b->dim != kFamilyDim[b->family])
    return fail(DE_ERR_INVALID, "family %d has dimension %d, got dim %d", b->family,
                kFamilyDim[b->family], b->dim);
  if (b->family == DE_FAMILY_SPRING_CHAIN &&
      (!b->sizes || !b->state_offset || !b->param_offset) && b->n > 0)
    return fail(DE_ERR_INVALID, "spring chain needs sizes, state_offset and param_offset");
  if (b->n_eval < 0) return fail(DE_ERR_INVALID, "n_eval < 0");
  if (b->n_eval > 0 && !b->t_eval) return fail(DE_ERR_INVALID, "n_eval > 0 without t_eval");
  if (b->n > 0 && (!b->y0 || !o->y_final || !o->t_final || !o->accepted || !o->rejected ||
                   !o->evals || !o->status))
    return fail(DE_ERR_INVALID, "missing input or output array");
  if (b->n > 0 && b->family != DE_FAMILY_SPRING_CHAIN && kFamilyParams[b->family] > 0 &&
      !b->params)
    return fail(DE_ERR_INVALID, "family %d needs params", b->family);
  if (b->schedule < DE_SCHED_AUTO || b->schedule > DE_SCHED_PERSISTENT_FIFO)
    return fail(DE_ERR_INVALID, "unknown schedule %d", b->schedule);
  if (b->max_steps < 0) return fail(DE_ERR_INVALID, "max_steps < 0");
  if (b->max_rejects < 0) return fail(DE_ERR_INVALID, "max_rejects < 0");
  if (b->event_solout < DE_SOLOUT_DEFAULT || b->event_solout > DE_SOLOUT_HYPERPLANE)
    return fail(DE_ERR_INVALID, "unknown event solout %d", b->event_solout);
  if (b->points < DE_POINTS_NONE || b->points > DE_POINTS_EVEN)
    return fail(DE_ERR_INVALID, "unknown point output %d", b->points);
  if (b->points != DE_POINTS_NONE) {
    if (b->event_solout != DE_SOLOUT_DEFAULT)
      return fail(DE_ERR_UNSUPPORTED, "a point output cannot be wrapped by an event solout");
    if (b->n_eval > 0)
      return fail(DE_ERR_INVALID, "a point output replaces the t_eval output (n_eval > 0)");
    if (b->points == DE_POINTS_DENSE && b->points_n < 1)
      return fail(DE_ERR_INVALID, "IVP::dense(n) needs n >= 1");
    if (b->points == DE_POINTS_EVEN && !(b->points_dt > 0.0 && b->points_dt < INFINITY))
      return fail(DE_ERR_INVALID, "IVP::even(dt) needs a finite dt > 0");
    if (b->n > 0 && (!o->points_offset || !o->n_points))
      return fail(DE_ERR_INVALID, "a point output needs points_offset and n_points");
  }
  if (b->event_solout != DE_SOLOUT_DEFAULT) {
    if (b->event_solout != DE_SOLOUT_EVENT && b->n_eval > 0)
      return fail(DE_ERR_INVALID, "a crossing solout replaces the t_eval output (n_eval > 0)");
    if (b->event_direction < DE_CROSS_BOTH || b->event_direction > DE_CROSS_NEGATIVE)
      return fail(DE_ERR_INVALID, "unknown crossing direction %d", b->event_direction);
    if (b->event_terminate < 0) return fail(DE_ERR_INVALID, "event_terminate < 0");
    if (b->max_events < 0) return fail(DE_ERR_INVALID, "max_events < 0");
    if (b->n > 0 && b->max_events > 0 && (!o->event_t || !o->event_y))
      return fail(DE_ERR_INVALID, "max_events > 0 needs event_t and event_y");
    const int fn = b->event_solout == DE_SOLOUT_CROSSING     ? DE_EVENT_COMPONENT
                   : b->event_solout == DE_SOLOUT_HYPERPLANE ? DE_EVENT_PLANE
                                                              : b->event_fn;
    if (fn < DE_EVENT_COMPONENT || fn > DE_EVENT_PLANE)
      return fail(DE_ERR_INVALID, "unknown event function %d", b->event_fn);
    if (b->family == DE_FAMILY_SPRING_CHAIN && fn == DE_EVENT_PLANE)
      return fail(DE_ERR_UNSUPPORTED,
                  "plane events and the hyperplane solout need one state length: not for "
                  "spring chains");
    if (fn == DE_EVENT_COMPONENT && b->event_component < 0)
      return fail(DE_ERR_INVALID, "event component %d < 0", b->event_component);
    if (fn == DE_EVENT_COMPONENT && b->family != DE_FAMILY_SPRING_CHAIN &&
        b->event_component >= b->dim)
      return fail(DE_ERR_INVALID, "event component %d outside [0, %d)", b->event_component,
                  b->dim);
    if (fn == DE_EVENT_PLANE && (!b->event_point || !b->event_normal))
      return fail(DE_ERR_INVALID, "a plane event needs event_point and event_normal");
  }
  return DE_OK;
}

de::OdeArgs make_args(const de_ode_batch* b, const de_ode_result* o) {
  de::OdeArgs a;
  memset(&a, 0, sizeof a);
  a.n = b->n;
  a.params = b->params;
  a.y0 = b->y0;
  a.t0 = b->t0;
  a.tf = b->tf;
  a.t0s = b->t0_scalar;
  a.tfs = b->tf_scalar;
  a.rtol_traj = b->rtol_traj;
  a.atol_traj = b->atol_traj;
  a.rtol_comp = b->rtol_comp;
  a.atol_comp = b->atol_comp;
  a.rtol = b->rtol;
  a.atol = b->atol;
  a.h0 = b->h0;
  a.h_min = b->h_min;
  a.h_max = b->h_max;
  a.safety = b->safety_factor;
  a.min_scale = b->min_scale;
  a.max_scale = b->max_scale;
  a.max_steps = b->max_steps;
  a.max_rejects = b->max_rejects;
  a.t_eval = b->t_eval;
  a.n_eval = b->n_eval;
  a.y_final = o->y_final;
  a.t_final = o->t_final;
  a.accepted = o->accepted;
  a.rejected = o->rejected;
  a.evals = o->evals;
  a.status = o->status;
  a.dense = b->n_eval > 0 ? o->dense : nullptr;
  a.n_dense = b->n_eval > 0 ? o->n_dense : nullptr;
  if (!a.dense) a.n_eval = 0;
  a.pts_mode = b->points;
  a.pts_n = b->points_n;
  a.pts_dt = b->points_dt;
  a.pts_offset = o->points_offset;
  a.pts_t = o->points_t;
  a.pts_y = o->points_y;
  a.n_pts = o->n_points;
  a.ev.solout = b->event_solout;
  a.ev.fn = b->event_solout == DE_SOLOUT_CROSSING     ? de::kEventComponent
            : b->event_solout == DE_SOLOUT_HYPERPLANE ? de::kEventPlane
                                                       : b->event_fn;
  a.ev.comp = b->event_component;
  a.ev.direction = b->event_direction;
  a.ev.terminate = b->event_terminate;
  a.ev.max_events = b->event_solout != DE_SOLOUT_DEFAULT ? b->max_events : 0;
  a.ev.value = b->event_value;
  a.ev.value_traj = b->event_value_traj;
  a.ev.point = b->event_point;
  a.ev.normal = b->event_normal;
  a.ev.t_out = o->event_t;
  a.ev.y_out = o->event_y;
  a.ev.n_out = o->n_events;
  return a;
}

// The embedded family (parity kernels, with or without DE_FLAG_FAST_MATH):
// two tableaux per translation unit.
int launch_ode_erk(int family, int method, int solout, const de::OdeArgs& a, int sched,
                   hipStream_t s) {
  using namespace de::parity;
  const int g = (method - DE_METHOD_RKF45) / 2;
  if (solout == DE_SOLOUT_EVENT) {
    switch (g) {
      case 0: return launch_ode_erk0_ev1(family, method, a, sched, s);
      case 1: return launch_ode_erk1_ev1(family, method, a, sched, s);
      case 2: return launch_ode_erk2_ev1(family, method, a, sched, s);
      case 3: return launch_ode_erk3_ev1(family, method, a, sched, s);
      case 4: return launch_ode_erk4_ev1(family, method, a, sched, s);
      default: return 1;
    }
  }
  if (solout != DE_SOLOUT_DEFAULT) {
    switch (g) {
      case 0: return launch_ode_erk0_ev2(family, method, a, sched, s);
      case 1: return launch_ode_erk1_ev2(family, method, a, sched, s);
      case 2: return launch_ode_erk2_ev2(family, method, a, sched, s);
      case 3: return launch_ode_erk3_ev2(family, method, a, sched, s);
      case 4: return launch_ode_erk4_ev2(family, method, a, sched, s);
      default: return 1;
    }
  }
  switch (g) {
    case 0: return de::parity::launch_ode_erk0(family, method, a, sched, s);
    case 1: return de::parity::launch_ode_erk1(family, method, a, sched, s);
    case 2: return de::parity::launch_ode_erk2(family, method, a, sched, s);
    case 3: return de::parity::launch_ode_erk3(family, method, a, sched, s);
    case 4: return de::parity::launch_ode_erk4(family, method, a, sched, s);
    default: return 1;
  }
}

// Launch on device-resident arguments.
int run_ode_device(const de_ode_batch* b, const int32_t* host_sizes, de::OdeArgs a,
                   hipStream_t s) {
  int sched = b->schedule == DE_SCHED_AUTO ? de::kSchedPersistent : b->schedule;
  if (b->family == DE_FAMILY_SPRING_CHAIN) {
    std::vector<int32_t> sizes_copy;
    if (!host_sizes) {
      sizes_copy.resize((size_t)b->n);
      hipError_t e = hipMemcpy(sizes_copy.data(), b->sizes, sizeof(int32_t) * b->n,
                               hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "reading spring sizes");
      host_sizes = sizes_copy.data();
    }
    int r = de::parity::launch_spring(b, host_sizes, a, sched, s);
    if (r == 1) return fail(DE_ERR_UNSUPPORTED, "spring-chain mapping not supported");
    if (r == 2) return fail(DE_ERR_INVALID, "points_offset is negative or decreases");
    if (r == 3)
      return fail(DE_ERR_INVALID, "event component %d outside a spring chain's state",
                  b->event_component);
    if (r < 0) return r == -4 ? fail(DE_ERR_NOMEM, "spring workspace allocation failed")
                              : hip_fail(hipGetLastError(), "spring launch");
    return DE_OK;
  }
  if ((sched == de::kSchedPersistent || sched == de::kSchedPersistentFifo ||
       b->event_solout != DE_SOLOUT_DEFAULT) && !a.counter) {
    int rc = counter_slot(&a.counter);
    if (rc != DE_OK) return rc;
  }
  // Dense output needs no pre-fill: each trajectory writes every one of its
  // n_eval rows, with NaN for the points it does not produce.
  // Event solouts run the parity kernels, with or without DE_FLAG_FAST_MATH.
  int r;
  if (b->method >= DE_METHOD_RKF45)
    r = launch_ode_erk(b->family, b->method, b->event_solout, a, sched, s);
  else if (b->event_solout == DE_SOLOUT_EVENT)
    r = de::parity::launch_ode_event(b->family, b->method, a, sched, s);
  else if (b->event_solout != DE_SOLOUT_DEFAULT)
    r = de::parity::launch_ode_crossing(b->family, b->method, a, sched, s);
  else if (b->flags & DE_FLAG_FAST_MATH)
    r = de::fast::launch_ode_lane(b->family, b->method, a, sched, s);
  else
    r = de::parity::launch_ode_lane(b->family, b->method, a, sched, s);
  if (r == 1)
    return fail(DE_ERR_UNSUPPORTED, "family %d / method %d not built%s", b->family, b->method,
                (b->flags & DE_FLAG_FAST_MATH) ? " in the fast build" : "");
  if (r < 0) return hip_fail(hipGetLastError(), "kernel launch");
  return DE_OK;
}

// ------------------------------------------------------------ to_csv ----
// The device CSV writer (kernels_csv.hip): bounded loops only (the digit
// search has fixed caps, the scan is three passes of the library's own with
// no inter-block waits), workspace from the library pool.
constexpr bool kCsvDeviceEnabled = true;

int csv_launch_fail(int r) {
  if (r == -2) return fail(DE_ERR_HIP, "csv formatter reached a loop cap (a bug: report it)");
  return hip_fail(hipGetLastError(), "csv launch");
}

int csv_args(const de_ode_batch* b, const de_ode_result* r, int32_t source, de::CsvArgs* a) {
  if (!b || !r) return fail(DE_ERR_INVALID, "null descriptor");
  if (b->n < 0) return fail(DE_ERR_INVALID, "n < 0");
  if (source < DE_CSV_FINAL || source > DE_CSV_POINTS)
    return fail(DE_ERR_INVALID, "unknown csv source %d", source);
  if (b->family < 0 || b->family > DE_FAMILY_ROBERTSON || b->family == DE_FAMILY_SPRING_CHAIN)
    return fail(DE_ERR_UNSUPPORTED, "csv output is built for the fixed-dimension families");
  if (b->dim != kFamilyDim[b->family])
    return fail(DE_ERR_INVALID, "dim %d does not match family %d", b->dim, b->family);
  memset(a, 0, sizeof *a);
  a->n = b->n;
  a->dim = b->dim;
  a->source = source;
  a->t0 = b->t0;
  a->tf = b->tf;
  a->t0s = b->t0_scalar;
  a->tfs = b->tf_scalar;
  a->t_eval = b->t_eval;
  a->n_eval = b->n_eval;
  const bool any = b->n > 0;
  if (source == DE_CSV_FINAL) {
    if (any && (!r->t_final || !r->y_final)) return fail(DE_ERR_INVALID, "missing final state");
    a->max_rows = 1;
    a->t_final = r->t_final;
    a->y_final = r->y_final;
  } else if (source == DE_CSV_TEVAL) {
    if (b->n_eval <= 0) return fail(DE_ERR_INVALID, "t_eval output needs n_eval > 0");
    if (any && (!b->t_eval || !r->dense || !r->n_dense))
      return fail(DE_ERR_INVALID, "missing t_eval, dense or n_dense");
    a->max_rows = b->n_eval;
    a->dense = r->dense;
    a->n_dense = r->n_dense;
  } else if (source == DE_CSV_POINTS) {
    if (any && (!r->points_offset || !r->n_points || !r->points_t || !r->points_y))
      return fail(DE_ERR_INVALID, "missing points_offset, n_points, points_t or points_y");
    // The stride of the slot grid: the most points any trajectory keeps.
    int64_t widest = 0;
    if (any) {
      std::vector<int64_t> off((size_t)b->n + 1);
      if (b->flags & DE_FLAG_HOST_MEMORY) {
        memcpy(off.data(), r->points_offset, sizeof(int64_t) * off.size());
      } else {
        hipError_t e = hipMemcpy(off.data(), r->points_offset, sizeof(int64_t) * off.size(),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "reading points_offset");
      }
      for (int64_t i = 0; i < b->n; ++i) {
        const int64_t w = off[i + 1] - off[i];
        if (w < 0) return fail(DE_ERR_INVALID, "points_offset decreases at %lld", (long long)i);
        widest = w > widest ? w : widest;
      }
    }
    if (widest > 0x7ffffffe) return fail(DE_ERR_UNSUPPORTED, "more than 2^31 points");
    a->max_rows = (int32_t)widest;
    a->pts_offset = r->points_offset;
    a->pts_t = r->points_t;
    a->pts_y = r->points_y;
    a->n_pts = r->n_points;
  } else {
    if (b->max_events < 0) return fail(DE_ERR_INVALID, "max_events < 0");
    if (any && b->max_events > 0 && (!r->event_t || !r->event_y || !r->n_events))
      return fail(DE_ERR_INVALID, "missing event_t, event_y or n_events");
    a->max_rows = b->max_events;
    a->event_t = r->event_t;
    a->event_y = r->event_y;
    a->n_events = r->n_events;
  }
  return DE_OK;
}

// Device copies of a host-memory request's inputs.
void csv_stage(Staging& st, const de_ode_batch* b, de::CsvArgs* a) {
  const size_t n = (size_t)a->n, d = (size_t)a->dim;
  a->t0 = st.in(a->t0, n);
  a->tf = st.in(a->tf, n);
  a->t_eval = st.in(a->t_eval, (size_t)(b->n_eval > 0 ? b->n_eval : 0));
  a->t_final = st.in(a->t_final, n);
  a->y_final = st.in(a->y_final, d * n);
  a->dense = st.in(a->dense, (size_t)a->max_rows * d * n);
  a->n_dense = st.in(a->n_dense, n);
  a->event_t = st.in(a->event_t, (size_t)a->max_rows * n);
  a->event_y = st.in(a->event_y, (size_t)a->max_rows * d * n);
  a->n_events = st.in(a->n_events, n);
  if (a->source == de::kCsvPoints && n > 0) {
    const int64_t* off = a->pts_offset;  // host memory here
    const size_t slots = (size_t)off[n];
    a->pts_t = st.in(a->pts_t, slots);
    a->pts_y = st.in(a->pts_y, slots * d);
    a->n_pts = st.in(a->n_pts, n);
    a->pts_offset = st.in(off, n + 1);
  }
}

}  // namespace

extern "C" {

const char* de_last_error(void) { return g_err.c_str(); }

const char* de_version(void) { return "de_amd 0.1.0 (gfx950, HIP)"; }

int de_set_workspace_budget(size_t bytes) {
  if (bytes < ((size_t)1 << 20)) return fail(DE_ERR_INVALID, "workspace budget below 1 MiB");
  if (de::set_workspace_budget(bytes) != hipSuccess) return fail(DE_ERR_HIP, "hipMemPool attribute");
  return DE_OK;
}

size_t de_workspace_budget(void) { return (size_t)de::workspace_budget(); }

int de_workspace_trim(void) {
  if (de::workspace_trim() != hipSuccess) return fail(DE_ERR_HIP, "hipMemPoolTrimTo");
  return DE_OK;
}

int de_ode_csv_offsets(const de_ode_batch* b, const de_ode_result* r, int32_t source,
                       int64_t* offsets, void* stream) {
  de::CsvArgs a;
  int rc = csv_args(b, r, source, &a);
  if (rc != DE_OK) return rc;
  if (!offsets) return fail(DE_ERR_INVALID, "null offsets");
  if (!kCsvDeviceEnabled)
    return fail(DE_ERR_UNSUPPORTED, "the device CSV writer is switched off in this build");
  hipStream_t s = (hipStream_t)stream;
  const bool host = (b->flags & DE_FLAG_HOST_MEMORY) != 0;
  Staging st;
  int64_t* d_off = offsets;
  if (host) {
    csv_stage(st, b, &a);
    d_off = st.out(offsets, (size_t)a.n + 1);
    if (st.err != hipSuccess) return hip_fail(st.err, "csv staging");
  }
  if (int r = de::launch_csv(a, d_off, nullptr, 0, s)) return csv_launch_fail(r);
  if (host) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "csv sync");
    st.back(offsets, d_off, (size_t)a.n + 1);
    if (st.err != hipSuccess) return hip_fail(st.err, "csv copy-back");
  }
  return DE_OK;
}

int de_ode_csv_write(const de_ode_batch* b, const de_ode_result* r, int32_t source,
                     const int64_t* offsets, char* out, void* stream) {
  de::CsvArgs a;
  int rc = csv_args(b, r, source, &a);
  if (rc != DE_OK) return rc;
  if (!offsets) return fail(DE_ERR_INVALID, "null offsets");
  if (!kCsvDeviceEnabled)
    return fail(DE_ERR_UNSUPPORTED, "the device CSV writer is switched off in this build");
  hipStream_t s = (hipStream_t)stream;
  const bool host = (b->flags & DE_FLAG_HOST_MEMORY) != 0;
  // The caller's buffer size, offsets[n]; the documents are laid out again and
  // must come to the same total.
  int64_t cap = 0;
  if (host) {
    cap = offsets[a.n];
  } else {
    hipError_t e = hipMemcpyAsync(&cap, offsets + a.n, sizeof cap, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "reading offsets[n]");
  }
  if (cap < 0) return fail(DE_ERR_INVALID, "offsets[n] < 0");
  if (cap > 0 && !out) return fail(DE_ERR_INVALID, "null out");
  Staging st;
  char* d_out = out;
  if (host) {
    csv_stage(st, b, &a);
    d_out = st.out(out, (size_t)cap);
  }
  int64_t* d_off = st.out((int64_t*)offsets, (size_t)a.n + 1);  // the layout, recomputed
  if (st.err != hipSuccess) return hip_fail(st.err, "csv staging");
  if (int r = de::launch_csv(a, d_off, cap > 0 ? d_out : nullptr, cap, s))
    return csv_launch_fail(r);
  int64_t total = 0;
  hipError_t e = hipMemcpyAsync(&total, d_off + a.n, sizeof total, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "csv sync");
  if (total != cap)
    return fail(DE_ERR_INVALID, "offsets[n] = %lld but the documents take %lld bytes",
                (long long)cap, (long long)total);
  if (host) {
    st.back(out, (const char*)d_out, (size_t)cap);
    if (st.err != hipSuccess) return hip_fail(st.err, "csv copy-back");
  }
  return DE_OK;
}

void de_ode_batch_init(de_ode_batch* b) {
  memset(b, 0, sizeof *b);
  b->rtol = 1.0e-6;
  b->atol = 1.0e-6;
  b->h0 = 0.0;
  b->h_min = 0.0;
  b->h_max = INFINITY;
  b->max_steps = 10000;
  b->max_rejects = 100;
  b->safety_factor = 0.9;
  b->min_scale = 0.2;
  b->max_scale = 10.0;
}

void de_sde_batch_init(de_sde_batch* b) {
  memset(b, 0, sizeof *b);
  b->h_min = 0.0;
  b->h_max = INFINITY;
  b->max_steps = 10000;
}

int64_t de_ode_workspace_bytes(const de_ode_batch* b) {
  if (!b) return 0;
  if (b->family == DE_FAMILY_SPRING_CHAIN) return de::parity::spring_workspace_bytes(b);
  // Work counter; for the persistent schedules the launch arguments in device
  // memory and, for the longest-first order, 256 bucket counts plus a 4-byte
  // position and a 1-byte bucket per trajectory.
  const bool events = b->event_solout != DE_SOLOUT_DEFAULT;
  const bool ordered = b->schedule == DE_SCHED_AUTO || b->schedule == DE_SCHED_PERSISTENT;
  const bool persistent = ordered || b->schedule == DE_SCHED_PERSISTENT_FIFO || events;
  return (int64_t)sizeof(unsigned long long) + (persistent ? 512 : 0) + (events ? 256 : 0) +
         (ordered ? 1024 + 5 * b->n : 0);
}

int de_ode_batch_solve(const de_ode_batch* b, de_ode_result* o, void* stream) {
  int rc = validate_ode(b, o);
  if (rc != DE_OK) return rc;
  if (b->n == 0) return DE_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!(b->flags & DE_FLAG_HOST_MEMORY)) return run_ode_device(b, nullptr, make_args(b, o), s);

  // Host-memory call: stage inputs, run, copy results back.
  const int64_t n = b->n;
  const bool ragged = b->family == DE_FAMILY_SPRING_CHAIN;
  int64_t state_len = ragged ? b->state_offset[n] : (int64_t)b->dim * n;
  int64_t param_len = ragged ? b->param_offset[n] : (int64_t)kFamilyParams[b->family] * n;
  Staging st;
  de_ode_batch db = *b;
  db.flags &= ~DE_FLAG_HOST_MEMORY;
  db.sizes = st.in(b->sizes, ragged ? n : 0);
  db.state_offset = st.in(b->state_offset, ragged ? n + 1 : 0);
  db.param_offset = st.in(b->param_offset, ragged ? n + 1 : 0);
  db.params = st.in(b->params, (size_t)param_len);
  db.y0 = st.in(b->y0, (size_t)state_len);
  db.t0 = st.in(b->t0, b->t0 ? n : 0);
  db.tf = st.in(b->tf, b->tf ? n : 0);
  db.rtol_traj = st.in(b->rtol_traj, b->rtol_traj ? n : 0);
  db.atol_traj = st.in(b->atol_traj, b->atol_traj ? n : 0);
  db.rtol_comp = st.in(b->rtol_comp, b->rtol_comp ? (size_t)state_len : 0);
  db.atol_comp = st.in(b->atol_comp, b->atol_comp ? (size_t)state_len : 0);
  db.t_eval = st.in(b->t_eval, b->n_eval);
  const bool events = b->event_solout != DE_SOLOUT_DEFAULT;
  db.event_value_traj = st.in(b->event_value_traj, events && b->event_value_traj ? n : 0);
  db.event_point = st.in(b->event_point, events && b->event_point ? b->dim : 0);
  db.event_normal = st.in(b->event_normal, events && b->event_normal ? b->dim : 0);
  de_ode_result dr;
  memset(&dr, 0, sizeof dr);
  dr.y_final = st.out(o->y_final, state_len);
  dr.t_final = st.out(o->t_final, n);
  dr.accepted = st.out(o->accepted, n);
  dr.rejected = st.out(o->rejected, n);
  dr.evals = st.out(o->evals, n);
  dr.status = st.out(o->status, n);
  int64_t dense_len = (int64_t)b->n_eval * state_len;
  dr.dense = b->n_eval > 0 ? st.out(o->dense, dense_len) : nullptr;
  dr.n_dense = b->n_eval > 0 ? st.out(o->n_dense, n) : nullptr;
  const int64_t ev_cap = events ? (int64_t)b->max_events * n : 0;
  dr.event_t = st.out(o->event_t, ev_cap);
  const int64_t ev_ylen = events ? (int64_t)b->max_events * (ragged ? state_len : b->dim * n) : 0;
  dr.event_y = st.out(o->event_y, ev_ylen);
  dr.n_events = st.out(o->n_events, events ? n : 0);
  const bool pts = b->points != DE_POINTS_NONE;
  const int64_t pts_cap = pts ? o->points_offset[n] : 0;
  if (pts && pts_cap < 0) return fail(DE_ERR_INVALID, "points_offset[n] < 0");
  // points_y: the slots times the state length (per system for spring chains).
  int64_t pts_ylen = pts_cap * b->dim;
  if (pts && ragged) {
    pts_ylen = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t slots = o->points_offset[i + 1] - o->points_offset[i];
      if (o->points_offset[i] < 0 || slots < 0)
        return fail(DE_ERR_INVALID, "points_offset is negative or decreases");
      pts_ylen += slots * 2 * (int64_t)b->sizes[i];
    }
  }
  dr.points_offset = st.in(o->points_offset, pts ? n + 1 : 0);
  dr.points_t = st.out(o->points_t, pts_cap);
  dr.points_y = st.out(o->points_y, pts_ylen);
  dr.n_points = st.out(o->n_points, pts ? n : 0);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging inputs");
  rc = run_ode_device(&db, ragged ? b->sizes : nullptr, make_args(&db, &dr), s);
  if (rc != DE_OK) return rc;
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "kernel execution");
  st.back(o->y_final, dr.y_final, state_len);
  st.back(o->t_final, dr.t_final, n);
  st.back(o->accepted, dr.accepted, n);
  st.back(o->rejected, dr.rejected, n);
  st.back(o->evals, dr.evals, n);
  st.back(o->status, dr.status, n);
  if (dr.dense) st.back(o->dense, dr.dense, dense_len);
  if (dr.n_dense) st.back(o->n_dense, dr.n_dense, n);
  if (dr.event_t) st.back(o->event_t, dr.event_t, ev_cap);
  if (dr.event_y) st.back(o->event_y, dr.event_y, ev_ylen);
  if (dr.n_events) st.back(o->n_events, dr.n_events, n);
  if (dr.points_t) st.back(o->points_t, dr.points_t, pts_cap);
  if (dr.points_y) st.back(o->points_y, dr.points_y, pts_ylen);
  if (dr.n_points) st.back(o->n_points, dr.n_points, n);
  if (st.err != hipSuccess) return hip_fail(st.err, "copying results");
  return DE_OK;
}

// ------------------------------------------------- no-allocation solve ----
// de_ode_batch_solve for callers that cannot allocate (code under graph
// capture, serving loops, harnesses that forbid hipMalloc): the launch's
// scratch memory -- the persistent schedule's work counter, the launch
// arguments, the hyperplane normal and the longest-first order -- comes from
// one file-scope device array per GPU instead of the library's pool.  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoAllocCounterBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

// The device address of this GPU's copy of `scratch` (host code never uses
// the array's name as an address).
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kScratchDevices) return fail(DE_ERR_UNSUPPORTED, "device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DE_OK;
}
}  // namespace

int64_t de_ode_noalloc_scratch_bytes(const de_ode_batch* b) {
  if (!b || b->n < 0) return -1;
  const int sched = b->schedule == DE_SCHED_AUTO ? de::kSchedPersistent : b->schedule;
  const bool events = b->event_solout != DE_SOLOUT_DEFAULT;
  const bool persistent = sched != de::kSchedPlain || events;
  if (!persistent) return 0;
  const size_t args = (sizeof(de::OdeArgs) + 255) & ~(size_t)255;
  const bool ordered = sched == de::kSchedPersistent;
  const size_t order = ordered ? (((size_t)256 * 4 + (size_t)b->n * 5 + 255) & ~(size_t)255) : 0;
  return (int64_t)(kNoAllocCounterBytes + args + (events ? 256 : 0) + order);
}

int de_ode_batch_solve_noalloc(const de_ode_batch* b, de_ode_result* o, void* stream) {
  int rc = validate_ode(b, o);
  if (rc != DE_OK) return rc;
  if (b->flags & DE_FLAG_HOST_MEMORY)
    return fail(DE_ERR_UNSUPPORTED, "the no-allocation solve takes device memory only");
  if (b->family == DE_FAMILY_SPRING_CHAIN)
    return fail(DE_ERR_UNSUPPORTED, "the no-allocation solve is built for the lane families");
  if (b->n == 0) return DE_OK;
  const int64_t need = de_ode_noalloc_scratch_bytes(b);
  if (need < 0 || (size_t)need > kNoAllocScratchBytes)
    return fail(DE_ERR_INVALID,
                "n = %lld needs %lld bytes of scratch, more than the %zu-byte maximum of the "
                "no-allocation solve", (long long)b->n, (long long)need, kNoAllocScratchBytes);
  char* base = nullptr;
  rc = scratch_address((void**)&base);
  if (rc != DE_OK) return rc;
  de::OdeArgs a = make_args(b, o);
  a.counter = (unsigned long long*)base;
  a.scratch = base + kNoAllocCounterBytes;
  a.scratch_bytes = (int64_t)(kNoAllocScratchBytes - kNoAllocCounterBytes);
  return run_ode_device(b, nullptr, a, (hipStream_t)stream);
}

int de_sde_batch_solve(const de_sde_batch* b, de_sde_result* o, void* stream) {
  if (!b || !o) return fail(DE_ERR_INVALID, "null descriptor");
  if (b->n < 0) return fail(DE_ERR_INVALID, "n < 0");
  if (b->family != DE_SDE_ORNSTEIN_UHLENBECK && b->family != DE_SDE_LINEAR_DIAGONAL)
    return fail(DE_ERR_INVALID, "unknown SDE family %d", b->family);
  if (b->scheme != DE_SCHEME_EULER_MARUYAMA && b->scheme != DE_SCHEME_MILSTEIN)
    return fail(DE_ERR_INVALID, "unknown scheme %d", b->scheme);
  if (b->dim != 1 && b->dim != 2 && b->dim != 4 && b->dim != 8)
    return fail(DE_ERR_UNSUPPORTED, "SDE dimension %d not built (1, 2, 4, 8)", b->dim);
  if (b->noise < DE_NOISE_PHILOX || b->noise > DE_NOISE_CONSTANT)
    return fail(DE_ERR_INVALID, "unknown noise source %d", b->noise);
  if (b->noise == DE_NOISE_TABLE && b->noise_steps > 0 && !b->noise_table)
    return fail(DE_ERR_INVALID, "noise table missing");
  if (b->path_stride < 0) return fail(DE_ERR_INVALID, "path_stride < 0");
  if (b->n == 0) return DE_OK;
  if (!b->params || !b->y0 || !o->y_final || !o->t_final || !o->steps || !o->evals ||
      !o->status)
    return fail(DE_ERR_INVALID, "missing input or output array");
  if (b->path_stride > 0 && !o->path)
    return fail(DE_ERR_INVALID, "path_stride > 0 needs a path buffer");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = b->n;
  const int P = b->family == DE_SDE_ORNSTEIN_UHLENBECK ? 3 * b->dim : 2 * b->dim;
  de::SdeArgs a;
  memset(&a, 0, sizeof a);
  a.n = n;
  a.dim = b->dim;
  a.noise = b->noise;
  a.t0 = b->t0;
  a.tf = b->tf;
  a.h0 = b->h0;
  a.h_min = b->h_min;
  a.h_max = b->h_max;
  a.max_steps = b->max_steps;
  a.seed = b->seed;
  a.path_base = b->path_base;
  a.noise_steps = b->noise_steps;
  a.noise_constant = b->noise_constant;
  a.path_stride = b->path_stride;
  int64_t n_rec = b->path_stride > 0 ? b->max_steps / b->path_stride + 1 : 0;
  if (!(b->flags & DE_FLAG_HOST_MEMORY)) {
    a.params = b->params;
    a.y0 = b->y0;
    a.noise_table = b->noise_table;
    a.y_final = o->y_final;
    a.t_final = o->t_final;
    a.steps = o->steps;
    a.evals = o->evals;
    a.status = o->status;
    a.path = o->path;
    a.path_t = o->path_t;
    int r = de::parity::launch_sde(b->family, b->scheme, a, s);
    if (r == 1) return fail(DE_ERR_UNSUPPORTED, "SDE kernel not built");
    if (r < 0) return hip_fail(hipGetLastError(), "SDE launch");
    return DE_OK;
  }
  Staging st;
  a.params = st.in(b->params, (size_t)P * n);
  a.y0 = st.in(b->y0, (size_t)b->dim * n);
  a.noise_table = st.in(b->noise_table, b->noise == DE_NOISE_TABLE ? (size_t)b->noise_steps * b->dim : 0);
  a.y_final = st.out(o->y_final, (size_t)b->dim * n);
  a.t_final = st.out(o->t_final, n);
  a.steps = st.out(o->steps, n);
  a.evals = st.out(o->evals, n);
  a.status = st.out(o->status, n);
  a.path = st.out(o->path, (size_t)(n_rec * b->dim * n));
  a.path_t = st.out(o->path_t, (size_t)n_rec);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging inputs");
  if (a.path) {
    int rc = fill_nan(a.path, n_rec * b->dim * n, s);
    if (rc != DE_OK) return rc;
  }
  if (a.path_t) {
    int rc = fill_nan(a.path_t, n_rec, s);
    if (rc != DE_OK) return rc;
  }
  int r = de::parity::launch_sde(b->family, b->scheme, a, s);
  if (r == 1) return fail(DE_ERR_UNSUPPORTED, "SDE kernel not built");
  if (r < 0) return hip_fail(hipGetLastError(), "SDE launch");
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "SDE kernel execution");
  st.back(o->y_final, a.y_final, (size_t)b->dim * n);
  st.back(o->t_final, a.t_final, n);
  st.back(o->steps, a.steps, n);
  st.back(o->evals, a.evals, n);
  st.back(o->status, a.status, n);
  st.back(o->path, a.path, (size_t)(n_rec * b->dim * n));
  st.back(o->path_t, a.path_t, (size_t)n_rec);
  if (st.err != hipSuccess) return hip_fail(st.err, "copying results");
  return DE_OK;
}

// ------------------------------------------------------------- workloads ---
void de_gen_lorenz(uint64_t seed, int64_t n, double* params, double* y0) {
  for (int64_t i = 0; i < n; ++i) {
    params[i] = 10.0;
    params[n + i] = 50.0 * de::input_uniform(seed, (uint64_t)i, 0, 0);
    params[2 * n + i] = 8.0 / 3.0;
    double z0, z1, z2, z3;
    de::input_normal_pair(seed, (uint64_t)i, 1, 0, &z0, &z1);
    de::input_normal_pair(seed, (uint64_t)i, 1, 1, &z2, &z3);
    y0[i] = 1.0 + 0.1 * z0;
    y0[n + i] = 1.0 + 0.1 * z1;
    y0[2 * n + i] = 1.0 + 0.1 * z2;
  }
}

void de_gen_vdp(uint64_t seed, int64_t n, double* params, double* y0) {
  for (int64_t i = 0; i < n; ++i) {
    double u = de::input_uniform(seed, (uint64_t)i, 2, 0);
    params[i] = exp(log(0.1) + u * log(100.0));
    double z0, z1;
    de::input_normal_pair(seed, (uint64_t)i, 3, 0, &z0, &z1);
    y0[i] = 2.0 + 0.1 * z0;
    y0[n + i] = 0.1 * z1;
  }
}

void de_gen_spring_sizes(uint64_t seed, int64_t n, int32_t* sizes) {
  for (int64_t i = 0; i < n; ++i) {
    double u = de::input_uniform(seed, (uint64_t)i, 4, 0);
    int32_t m = (int32_t)floor(exp(log(4.0) + u * (log(257.0) - log(4.0))));
    sizes[i] = m < 4 ? 4 : (m > 256 ? 256 : m);
  }
}

void de_spring_offsets(int64_t n, const int32_t* sizes, int64_t* so, int64_t* po) {
  so[0] = 0;
  po[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    so[i + 1] = so[i] + 2 * (int64_t)sizes[i];
    po[i + 1] = po[i] + (int64_t)sizes[i] + 2;
  }
}

void de_gen_spring(uint64_t seed, int64_t n, const int32_t* sizes, const int64_t* so,
                   const int64_t* po, double* params, double* y0) {
  for (int64_t i = 0; i < n; ++i) {
    const int m = sizes[i];
    double* p = params + po[i];
    p[0] = 0.1 * de::input_uniform(seed, (uint64_t)i, 5, 0);
    for (int j = 0; j <= m; j += 2) {
      double z0, z1;
      de::input_normal_pair(seed, (uint64_t)i, 6, (uint32_t)(j >> 1), &z0, &z1);
      p[1 + j] = exp(0.5 * z0);
      if (j + 1 <= m) p[2 + j] = exp(0.5 * z1);
    }
    double* y = y0 + so[i];
    for (int j = 0; j < m; j += 2) {
      double z0, z1;
      de::input_normal_pair(seed, (uint64_t)i, 7, (uint32_t)(j >> 1), &z0, &z1);
      y[j] = z0;
      if (j + 1 < m) y[j + 1] = z1;
    }
    for (int j = 0; j < m; ++j) y[m + j] = 0.0;
  }
}

void de_gen_ou(uint64_t seed, int64_t n, int32_t dim, double* params, double* y0) {
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < dim; ++c) {
      de::u32x4 w = de::input_block(seed, (uint64_t)i, 8, (uint32_t)c);
      double u_theta = de::uniform_co(w.v[0], w.v[1]);
      double u_sigma = de::uniform_co(w.v[2], w.v[3]);
      double z0, z1;
      de::input_normal_pair(seed, (uint64_t)i, 9, (uint32_t)c, &z0, &z1);
      params[(int64_t)(3 * c + 0) * n + i] = 0.5 + 1.5 * u_theta;
      params[(int64_t)(3 * c + 1) * n + i] = z0;
      params[(int64_t)(3 * c + 2) * n + i] = 0.1 + 0.9 * u_sigma;
      y0[(int64_t)c * n + i] = 0.0;
    }
  }
}

void de_gen_gbm(int64_t n, double mu, double sigma, double s0, double* params, double* y0) {
  for (int64_t i = 0; i < n; ++i) {
    params[i] = mu;
    params[n + i] = sigma;
    y0[i] = s0;
  }
}

void de_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  de::u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  de::u32x4 w = de::philox4x32_10(c, k0, k1);
  for (int i = 0; i < 4; ++i) out[i] = w.v[i];
}

double de_noise_normal(uint64_t seed, uint64_t path, uint32_t step, uint32_t comp) {
  return de::noise_normal(seed, path, step, comp);
}

int de_spring_tuned(int32_t cap, int32_t* g, int32_t* c, int32_t* block) {
  int gg = 0, cc = 0, bb = 0;
  int r = de::parity::spring_tuned(cap, &gg, &cc, &bb);
  if (r == 0) {
    *g = gg;
    *c = cc;
    *block = bb;
  }
  return r;
}

}  // extern "C"

// ---------------------------------------------------------- self-checks ----
namespace {

__global__ void diag_pow_kernel(const double* x, const double* y, double* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = de::libm_pow(x[i], y[i]);
}

__global__ void diag_arith_kernel(const double* a, const double* b, const double* c,
                                  double* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[3 * i + 0] = a[i] / b[i];
    out[3 * i + 1] = sqrt(a[i]);
    out[3 * i + 2] = fma(a[i], b[i], c[i]);
  }
}

__global__ void diag_noise_kernel(uint64_t seed, const uint64_t* path, const uint32_t* step,
                                  const uint32_t* comp, double* out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = de::noise_normal(seed, path[i], step[i], comp[i]);
}

unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int de_diag_pow(const double* x, const double* y, double* out, int64_t n) {
  if (n <= 0) return DE_OK;
  Staging st;
  const double* dx = st.in(x, n);
  const double* dy = st.in(y, n);
  double* dout = st.out(out, n);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging");
  hipLaunchKernelGGL(diag_pow_kernel, dim3(grid_for(n)), dim3(256), 0, 0, dx, dy, dout, n);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "diag_pow");
  st.back(out, dout, n);
  return st.err == hipSuccess ? DE_OK : hip_fail(st.err, "copy back");
}

int de_diag_arith(const double* a, const double* b, const double* c, double* out, int64_t n) {
  if (n <= 0) return DE_OK;
  Staging st;
  const double* da = st.in(a, n);
  const double* db = st.in(b, n);
  const double* dc = st.in(c, n);
  double* dout = st.out(out, 3 * n);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging");
  hipLaunchKernelGGL(diag_arith_kernel, dim3(grid_for(n)), dim3(256), 0, 0, da, db, dc, dout, n);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "diag_arith");
  st.back(out, dout, 3 * n);
  return st.err == hipSuccess ? DE_OK : hip_fail(st.err, "copy back");
}

int de_diag_fmt(const double* x, char* out, int32_t* len, int64_t n) {
  if (n <= 0) return DE_OK;
  Staging st;
  const double* dx = st.in(x, n);
  char* dout = st.out(out, (size_t)n * 32);
  int32_t* dlen = st.out(len, n);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging");
  if (de::launch_fmt_diag(dx, dout, dlen, n, 0) != 0) return hip_fail(hipGetLastError(), "diag_fmt");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "diag_fmt");
  st.back(out, dout, (size_t)n * 32);
  st.back(len, dlen, n);
  return st.err == hipSuccess ? DE_OK : hip_fail(st.err, "copy back");
}

int de_diag_noise(uint64_t seed, const uint64_t* path, const uint32_t* step,
                  const uint32_t* comp, double* out, int64_t n) {
  if (n <= 0) return DE_OK;
  Staging st;
  const uint64_t* dp = st.in(path, n);
  const uint32_t* ds = st.in(step, n);
  const uint32_t* dc = st.in(comp, n);
  double* dout = st.out(out, n);
  if (st.err != hipSuccess) return hip_fail(st.err, "staging");
  hipLaunchKernelGGL(diag_noise_kernel, dim3(grid_for(n)), dim3(256), 0, 0, seed, dp, ds, dc,
                     dout, n);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "diag_noise");
  st.back(out, dout, n);
  return st.err == hipSuccess ? DE_OK : hip_fail(st.err, "copy back");
}

}  // extern "C"
