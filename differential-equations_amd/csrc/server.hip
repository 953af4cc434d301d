// Serving: continuous batching of ensemble solves on one GPU.
//
// The reference has no serving layer; this is the request loop that admits
// ensemble solves (IVP::ode(...).method(dopri5()).solve() batches) of
// different families, sizes and tolerances into one running device.
//
// Device side.  Each family (Lorenz, Van der Pol, and one per spring-chain size
// class) owns a pool of solver slots whose complete state (a Dopri5Lane or a
// GroupDopri5 per lane) stays in HBM between launches, and a work queue: an
// append-only array of ranges (request, trajectories) whose trajectories are
// numbered by a running ticket.  A chunk kernel runs every slot for up to K
// attempts; a slot that finishes (or is idle) takes the next ticket with one
// wave-aggregated atomic and starts that trajectory in the same launch, so
// requests admitted while others run share the lanes immediately.  A ticket
// beyond the published range is held until the host publishes more work.
//
// Host side.  A scheduler thread admits requests first-come-first-served while
// their device footprint (inputs, outputs, dense output) fits the HBM budget
// (first-fit arena inside one allocation), uploads inputs, appends ranges and
// publishes them, launches one chunk per family with work (concurrent streams,
// optionally replayed from a captured hipGraph), reads back the per-request
// remaining counts, and retires finished requests by copying their results to
// host memory.  The loop is pipelined: iteration k+1 is queued on the device
// before the host reads iteration k's counts, and admissions (uploads on their
// own stream) and retirements (copies on theirs) run while the device
// iterates.  A request's work becomes visible to the chunk kernels only
// through a commit kernel ordered between two iterations.  Every trajectory's
// arithmetic depends only on its own inputs, so a request's results do not
// depend on what else is being served.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/de_api.h"
#include "de_launch.h"
#include "de_server.h"

namespace de {
namespace serve {

// Publishes a pool's new work between two iterations: the chunk kernels read
// the counters once at their start.
__global__ void commit_pool(Queue* q, int64_t published, int64_t nranges) {
  if (threadIdx.x == 0) {
    q->published = published;
    q->nranges = nranges;
  }
}

__global__ void gather_remaining(const ReqDev* reqs, int32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = reqs[i].remaining;
}

// First-fit arena over one allocation: device memory (inputs) or pinned,
// device-mapped host memory that the kernels write results into.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  bool host = false;
  std::map<size_t, size_t> free_;  // offset -> length
  bool init(size_t bytes, bool in_host) {
    host = in_host;
    const hipError_t e = in_host ? hipHostMalloc(&base, bytes, hipHostMallocCoherent)
                                 : hipMalloc(&base, bytes);
    if (e != hipSuccess) {
      base = nullptr;
      return false;
    }
    size = bytes;
    free_[0] = bytes;
    return true;
  }
  void destroy() {
    if (base) (void)(host ? hipHostFree(base) : hipFree(base));
    base = nullptr;
  }
  size_t largest_free() const {
    size_t m = 0;
    for (const auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
  void* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        return base + off;
      }
    }
    return nullptr;
  }
  void release(void* p, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    size_t off = (char*)p - base;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  uint64_t ticket;
  de_ode_batch b;            // host copy; pointers below
  std::vector<double> params, y0, t0, tf, rtol, atol, t_eval;
  std::vector<double> ev_value, ev_point, ev_normal;  // event solout inputs (normal normalised)
  std::vector<int32_t> sizes;
  std::vector<int64_t> so, po;
  int64_t state_len = 0;
  int vectors = 0;  // spring chains on one lane per system: workspace vectors
  // results: written by the kernels into the host arena at these pointers
  de::OdeArgs out;
  // device inputs and host-arena results
  void* dmem = nullptr;
  size_t dbytes = 0;
  void* hmem = nullptr;
  size_t hbytes = 0;
  std::vector<int> pools;  // the pools its work is published to
  int slot = -1;
  int64_t first_iter = 0;  // first iteration whose gathered counts include this request
  bool trimmed = false;    // idle lazy pools were released to make room for it
  bool done = false;
  int error = 0;
  double t_submit = 0, t_done = 0;

  // The host copies of the inputs, once they are on the device.
  void drop_inputs() {
    for (auto* v : {&params, &y0, &t0, &tf, &rtol, &atol, &t_eval, &ev_value, &ev_point, &ev_normal})
      std::vector<double>().swap(*v);
    std::vector<int32_t>().swap(sizes);
    std::vector<int64_t>().swap(so);
    std::vector<int64_t>().swap(po);
  }
};

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace serve
}  // namespace de

using namespace de::serve;

struct de_server {
  size_t budget = 0;
  int32_t flags = 0;
  // Attempts per slot of the next iteration: kChunk while the active work
  // fits the pools' slots, up to kMaxChunk under a backlog (fewer iteration
  // boundaries, where every pool's kernel must end before the next starts).
  // A captured graph keeps kChunk.
  int chunk = de::serve::kChunk;
  int ncu = 256;
  int64_t max_ranges = 0;
  Arena arena;   // device: request inputs
  Arena harena;  // pinned host: request results
  std::vector<Family> fams;
  de::serve::ReqDev* d_reqs = nullptr;
  std::vector<de::serve::ReqDev> h_reqs;
  int32_t* h_counts[2] = {};       // pinned: gathered counts of iterations k (k & 1)
  int32_t* d_remaining = nullptr;  // gathered from d_reqs for one copy per iteration
  hipStream_t upl = nullptr;  // admission uploads
  hipEvent_t ev_counts[2] = {}, ev_upl = nullptr;
  // Pinned staging of an admission phase's uploads (reused once the upload
  // stream is synchronised).
  char* stage = nullptr;
  size_t stage_size = 0, stage_used = 0;
  int64_t launched = 0;            // iterations launched
  std::vector<int> free_slots;
  std::map<int, Request*> active;  // request slot -> request
  std::deque<Request*> pending;
  std::map<uint64_t, Request*> all;      // submitted, not yet delivered
  std::map<uint64_t, double> latency;    // delivered requests' latencies
  uint64_t next_ticket = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::thread thread;
  hipStream_t ctl;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> fam_done;
  hipGraphExec_t graph = nullptr;
  int iterations = 0;
  int64_t range_resets = 0;
  int64_t pools_released = 0;  // idle lazy pools released to the arena
  // Scheduler time split (seconds): admitting (uploads, publishing), the
  // iterations from launch to their counts on the host, and retiring.
  double t_admit = 0, t_iter = 0, t_retire = 0;
  int error = 0;
};

namespace {

enum : int32_t { kServerGraph = 1 };

// Pools made at creation: 0 / 1 Lorenz / Van der Pol with DOPRI5, 2 ..
// kSpringEnd - 1 the spring classes (DOPRI5, group mappings), then Lorenz /
// Van der Pol with DOP853, then DOPRI5 with an event solout
// (EventWrappedSolout, or the crossing / hyperplane solouts), whose roots are
// found inline.  Pools made on first use (server_pools.hip), their memory
// carved from the arena: Lorenz / Van der Pol with RK4, Euler and the
// embedded family at kLaneLazy + 2 method + (Van der Pol), spring chains one
// lane per system at kSpringVec + method (every method but DOPRI5, and DOPRI5
// asking for DE_MAP_LANE), and Lorenz / Van der Pol with an event solout and a
// method other than DOPRI5 at kLaneEvent + 4 method + 2 (crossing) + (Van der
// Pol), and the spring classes with DOPRI5 and an event solout at
// kSpringEvent + class.
constexpr int kSpringEnd = 9;
constexpr int kLorenzDop853 = 9, kVdpDop853 = 10;
constexpr int kLorenzEvent = 11, kLorenzCross = 12, kVdpEvent = 13, kVdpCross = 14;
constexpr int kEagerPools = 15;
constexpr int kMethods = DE_METHOD_RKV989E + 1;
constexpr int kLaneLazy = kEagerPools;
constexpr int kSpringVec = kLaneLazy + 2 * kMethods;
constexpr int kLaneEvent = kSpringVec + kMethods;
constexpr int kSpringEvent = kLaneEvent + 4 * kMethods;  // + size class
constexpr int kPools = kSpringEvent + (kSpringEnd - 2);
constexpr int kLazyRanges = 1 << 16;  // range table of a lazy pool

bool lane_lazy_method(int method) {
  return method == DE_METHOD_RK4 || method == DE_METHOD_EULER || method >= DE_METHOD_RKF45;
}

// The pool of a request (spring chains: of its systems with nm masses), or -1.
// Spring chains asking for one lane per system (DE_MAP_LANE) are served so.
int fam_index_for(de_server* s, int family, int nm, int method, int solout, int mapping) {
  if (method < 0 || method >= kMethods) return -1;
  const bool d853 = method == DE_METHOD_DOP853;
  if (family == DE_FAMILY_SPRING_CHAIN) {
    if (method != DE_METHOD_DOPRI5 || mapping == DE_MAP_LANE) return kSpringVec + method;
    for (int k = 2; k < kSpringEnd; ++k)
      if (nm <= s->fams[k].cap)
        return solout == DE_SOLOUT_DEFAULT ? k : kSpringEvent + (k - 2);
    return -1;
  }
  if (family != DE_FAMILY_LORENZ && family != DE_FAMILY_VAN_DER_POL) return -1;
  const int v = family == DE_FAMILY_VAN_DER_POL;
  if (solout != DE_SOLOUT_DEFAULT) {
    const bool cross = solout != DE_SOLOUT_EVENT;
    if (method != DE_METHOD_DOPRI5) return kLaneEvent + 4 * method + 2 * cross + v;
    return v ? (cross ? kVdpCross : kVdpEvent) : (cross ? kLorenzCross : kLorenzEvent);
  }
  if (method == DE_METHOD_DOPRI5) return v;
  if (d853) return v ? kVdpDop853 : kLorenzDop853;
  return kLaneLazy + 2 * method + v;
}

// The slot kind and chunk kernel of lazy pool k (family and method from its index).
bool lazy_kind(int k, Family& f) {
  int family, method, ev = 0;
  if (k >= kSpringEvent) return spring_event_class_kind(k - kSpringEvent, f);
  if (k >= kLaneEvent) {
    const int rel = k - kLaneEvent;
    family = rel & 1 ? DE_FAMILY_VAN_DER_POL : DE_FAMILY_LORENZ;
    ev = (rel >> 1) & 1 ? 2 : 1;
    method = rel >> 2;
    if (method == DE_METHOD_DOPRI5) return false;
  } else if (k >= kSpringVec) {
    family = DE_FAMILY_SPRING_CHAIN;
    method = k - kSpringVec;
  } else {
    family = (k - kLaneLazy) & 1 ? DE_FAMILY_VAN_DER_POL : DE_FAMILY_LORENZ;
    method = (k - kLaneLazy) >> 1;
    if (!lane_lazy_method(method)) return false;
  }
  return lazy_pool_kind_0(family, method, ev, f) || lazy_pool_kind_1(family, method, ev, f) ||
         lazy_pool_kind_2(family, method, ev, f) || lazy_pool_kind_3(family, method, ev, f) ||
         lazy_pool_kind_4(family, method, ev, f) || lazy_pool_kind_5(family, method, ev, f);
}

// Device bytes of a request's inputs (and spring class lists) and host bytes
// of its results.
void footprint(const Request* r, size_t* dev, size_t* host) {
  const de_ode_batch& b = r->b;
  size_t n = (size_t)b.n;
  size_t d = 0, h = 0;
  d += r->params.size() * 8 + r->y0.size() * 8 + r->t0.size() * 8 + r->tf.size() * 8;
  d += r->rtol.size() * 8 + r->atol.size() * 8 + r->t_eval.size() * 8;
  d += r->sizes.size() * 4 + (r->so.size() + r->po.size()) * 8 + n * 4;  // + class lists
  d += (r->ev_value.size() + r->ev_point.size() + r->ev_normal.size()) * 8;
  // one-lane spring solvers: their vectors, the working state and, with an
  // event solout, the event search's state
  d += (size_t)r->vectors * r->state_len * 8;
  h += (size_t)r->state_len * 8 + n * (8 * 4 + 4 + 4);
  h += (size_t)b.n_eval * r->state_len * 8;
  if (b.event_solout != DE_SOLOUT_DEFAULT)
    h += (size_t)b.max_events * n * 8 + (size_t)b.max_events * r->state_len * 8 + n * 4;
  *dev = d + 64 * 256;
  *host = h + 16 * 256;
}

// Host-to-device copy on the upload stream through the pinned stage (one
// asynchronous DMA, no staging inside the runtime); a full stage first waits
// for the uploads in flight.
void stage_copy(de_server* s, void* dst, const void* src, size_t nb) {
  const size_t need = (nb + 255) & ~size_t(255);
  if (need > s->stage_size) {  // larger than the stage: the runtime stages it
    (void)hipMemcpyAsync(dst, src, nb, hipMemcpyHostToDevice, s->upl);
    (void)hipStreamSynchronize(s->upl);
    return;
  }
  if (s->stage_used + need > s->stage_size) {
    (void)hipStreamSynchronize(s->upl);
    s->stage_used = 0;
  }
  char* h = s->stage + s->stage_used;
  memcpy(h, src, nb);
  s->stage_used += need;
  (void)hipMemcpyAsync(dst, h, nb, hipMemcpyHostToDevice, s->upl);
}

// Upload a request into the arena and build its device descriptor.
bool upload(de_server* s, Request* r, std::vector<std::vector<int32_t>>& class_lists,
            int32_t** d_lists) {
  size_t bytes = 0, hbytes = 0;
  footprint(r, &bytes, &hbytes);
  char* p = (char*)s->arena.alloc(bytes);
  if (!p) return false;
  char* hp = (char*)s->harena.alloc(hbytes);
  if (!hp) {
    s->arena.release(p, bytes);
    return false;
  }
  r->dmem = p;
  r->dbytes = bytes;
  r->hmem = hp;
  r->hbytes = hbytes;
  auto put = [&](const void* src, size_t nb) -> void* {
    if (!nb) return nullptr;
    void* d = p;
    stage_copy(s, d, src, nb);
    p += (nb + 255) & ~size_t(255);
    return d;
  };
  auto reserve = [&](size_t nb) -> double* {  // device workspace, not initialised
    if (!nb) return nullptr;
    void* d = p;
    p += (nb + 255) & ~size_t(255);
    return (double*)d;
  };
  auto take = [&](size_t nb) -> void* {  // results: the host arena
    if (!nb) return nullptr;
    void* d = hp;
    hp += (nb + 255) & ~size_t(255);
    return d;
  };
  const de_ode_batch& b = r->b;
  const int64_t n = b.n;
  ReqDev rd;
  memset(&rd, 0, sizeof rd);
  de::OdeArgs& a = rd.a.o;
  a.n = n;
  a.params = (const double*)put(r->params.data(), r->params.size() * 8);
  a.y0 = (const double*)put(r->y0.data(), r->y0.size() * 8);
  a.t0 = (const double*)put(r->t0.data(), r->t0.size() * 8);
  a.tf = (const double*)put(r->tf.data(), r->tf.size() * 8);
  a.t0s = b.t0_scalar;
  a.tfs = b.tf_scalar;
  a.rtol_traj = (const double*)put(r->rtol.data(), r->rtol.size() * 8);
  a.atol_traj = (const double*)put(r->atol.data(), r->atol.size() * 8);
  a.rtol = b.rtol;
  a.atol = b.atol;
  a.h0 = b.h0;
  a.h_min = b.h_min;
  a.h_max = b.h_max;
  a.safety = b.safety_factor;
  a.min_scale = b.min_scale;
  a.max_scale = b.max_scale;
  a.max_steps = b.max_steps;
  a.max_rejects = b.max_rejects;
  a.t_eval = (const double*)put(r->t_eval.data(), r->t_eval.size() * 8);
  a.n_eval = b.n_eval;
  a.y_final = (double*)take((size_t)r->state_len * 8);
  a.t_final = (double*)take(n * 8);
  a.accepted = (int64_t*)take(n * 8);
  a.rejected = (int64_t*)take(n * 8);
  a.evals = (int64_t*)take(n * 8);
  a.status = (int32_t*)take(n * 4);
  a.n_dense = b.n_eval > 0 ? (int32_t*)take(n * 4) : nullptr;
  a.dense = b.n_eval > 0 ? (double*)take((size_t)b.n_eval * r->state_len * 8) : nullptr;
  if (b.event_solout != DE_SOLOUT_DEFAULT) {
    // The batch path's event arguments (roots inline: no bracket regions).
    a.ev.solout = b.event_solout;
    a.ev.fn = b.event_solout == DE_SOLOUT_CROSSING     ? de::kEventComponent
              : b.event_solout == DE_SOLOUT_HYPERPLANE ? de::kEventPlane
                                                        : b.event_fn;
    a.ev.comp = b.event_component;
    a.ev.direction = b.event_direction;
    a.ev.terminate = b.event_terminate;
    a.ev.max_events = b.max_events;
    a.ev.value = b.event_value;
    a.ev.value_traj = (const double*)put(r->ev_value.data(), r->ev_value.size() * 8);
    a.ev.point = (const double*)put(r->ev_point.data(), r->ev_point.size() * 8);
    a.ev.normal = (const double*)put(r->ev_normal.data(), r->ev_normal.size() * 8);
    a.ev.t_out = (double*)take((size_t)b.max_events * n * 8);
    a.ev.y_out = (double*)take((size_t)b.max_events * r->state_len * 8);
    a.ev.n_out = (int32_t*)take(n * 4);
  }
  if (r->vectors > 0) {  // one lane per system: working state, event state, vectors
    const size_t sl = (size_t)r->state_len * 8;
    const int ev = b.event_solout != DE_SOLOUT_DEFAULT;
    rd.a.ystate = reserve(sl);
    rd.a.ev_work = ev ? reserve(sl) : nullptr;
    rd.a.work = reserve((size_t)(r->vectors - 1 - ev) * sl);
  }
  if (b.family == DE_FAMILY_SPRING_CHAIN) {
    rd.a.sizes = (const int32_t*)put(r->sizes.data(), r->sizes.size() * 4);
    rd.a.state_offset = (const int64_t*)put(r->so.data(), r->so.size() * 8);
    rd.a.param_offset = (const int64_t*)put(r->po.data(), r->po.size() * 8);
    for (size_t k = 0; k < class_lists.size(); ++k)
      d_lists[k] = (int32_t*)put(class_lists[k].data(), class_lists[k].size() * 4);
  }
  rd.remaining = (int32_t)n;
  s->h_reqs[r->slot] = rd;
  r->out = rd.a.o;
  stage_copy(s, s->d_reqs + r->slot, &s->h_reqs[r->slot], sizeof rd);
  return true;
}

void publish(de_server* s, Family& f, int req_slot, int32_t first, int32_t count,
             const int32_t* list) {
  Range rg;
  rg.start = f.published;
  rg.req = req_slot;
  rg.first = first;
  rg.count = count;
  rg.pad = 0;
  rg.list = list;
  int64_t idx = (int64_t)f.host_ranges.size();
  f.host_ranges.push_back(rg);
  stage_copy(s, f.ranges + idx, &rg, sizeof rg);
  f.published += count;
  f.dirty = true;
}

// One chunk of every pool (graph capture), or of the pools that hold work of
// an active request: an idle pool's launch would only read and write back its
// slots.
// With lazy_only, the pools made on first use (outside a captured graph),
// forked from the event recorded before the graph's launch.
void launch_all(de_server* s, bool busy_only = false, bool lazy_only = false) {
  if (!lazy_only) (void)hipEventRecord(s->fork, s->ctl);
  for (size_t k = 0; k < s->fams.size(); ++k) {
    Family& f = s->fams[k];
    if (!f.slots || (lazy_only && !f.lazy)) continue;
    if (busy_only && f.active_reqs == 0) continue;
    (void)hipStreamWaitEvent(f.stream, s->fork, 0);
    f.launch(f, s->d_reqs, s->chunk, f.stream);
    (void)hipEventRecord(s->fam_done[k], f.stream);
    (void)hipStreamWaitEvent(s->ctl, s->fam_done[k], 0);
  }
}

// Capture one iteration (chunk of every family) into a graph.
bool capture(de_server* s) {
  hipGraph_t g;
  if (hipStreamBeginCapture(s->ctl, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  launch_all(s);
  if (hipStreamEndCapture(s->ctl, &g) != hipSuccess) return false;
  bool ok = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0) == hipSuccess;
  (void)hipGraphDestroy(g);
  return ok;
}

// Makes lazy pool k: slots, queue and range table from the arena, slots
// initialised and the queue cleared on the control stream (before any of the
// pool's launches, which fork from it).
bool make_lazy(de_server* s, int k) {
  Family& f = s->fams[k];
  if (!lazy_kind(k, f)) return false;
  const size_t sb = f.slot_bytes * (size_t)f.nslots, rb = sizeof(Range) * (size_t)f.max_ranges;
  void* slots = s->arena.alloc(sb);
  void* q = slots ? s->arena.alloc(sizeof(Queue)) : nullptr;
  void* ranges = q ? s->arena.alloc(rb) : nullptr;
  bool ok = ranges != nullptr;
  if (ok && !f.stream) ok = hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    if (ranges) s->arena.release(ranges, rb);
    if (q) s->arena.release(q, sizeof(Queue));
    if (slots) s->arena.release(slots, sb);
    return false;
  }
  f.slots = slots;
  f.q = (Queue*)q;
  f.ranges = (Range*)ranges;
  f.init(f, s->ctl);
  f.reset(f, s->ctl);  // the queue's counters from zero
  return true;
}

// Releases the lazy pools that hold no active request's work back to the
// arena, after the iterations in flight (which may still run their chunks)
// have ended.  Returns whether any was released.
bool trim_lazy(de_server* s) {
  bool any = false;
  for (auto& f : s->fams)
    any = any || (f.lazy && f.slots && f.active_reqs == 0 && !f.dirty);
  if (!any) return false;
  (void)hipStreamSynchronize(s->ctl);
  for (auto& f : s->fams) {
    if (!(f.lazy && f.slots && f.active_reqs == 0 && !f.dirty)) continue;
    s->arena.release(f.ranges, sizeof(Range) * (size_t)f.max_ranges);
    s->arena.release(f.q, sizeof(Queue));
    s->arena.release(f.slots, f.slot_bytes * (size_t)f.nslots);
    f.slots = nullptr;
    f.q = nullptr;
    f.ranges = nullptr;
    f.host_ranges.clear();
    f.published = 0;
    s->pools_released += 1;
  }
  return true;
}

void scheduler(de_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  int64_t inflight = -1;  // last queued iteration whose counts are not read yet
  std::vector<Request*> admitted;  // inputs in flight on the upload stream
  while (true) {
    // Admission: first come, first served, while the footprint fits.
    const double t_a0 = now_s();
    while (!s->pending.empty() && !s->free_slots.empty()) {
      Request* r = s->pending.front();
      std::vector<std::vector<int32_t>> lists(s->fams.size());
      const de_ode_batch& b = r->b;
      std::vector<int> needs;
      const int k0 = fam_index_for(s, b.family, 0, b.method, b.event_solout, b.mapping);
      // Spring chains with DOPRI5 go to the size-class pools (group mappings).
      const bool vec = k0 >= kSpringVec && k0 < kSpringVec + kMethods;
      const bool classes = b.family == DE_FAMILY_SPRING_CHAIN && !vec;
      if (classes) {
        for (int64_t i = 0; i < b.n; ++i)
          lists[fam_index_for(s, DE_FAMILY_SPRING_CHAIN, r->sizes[i], b.method, b.event_solout,
                              b.mapping)]
              .push_back((int32_t)i);
        for (size_t k = 0; k < lists.size(); ++k)
          if (!lists[k].empty()) needs.push_back((int)k);
      } else {
        needs.push_back(k0);
      }
      // A pool made on first use takes its memory from the arena.
      bool unmade = false;
      for (int k : needs)
        if (!s->fams[k].slots && !make_lazy(s, k)) unmade = true;
      if (unmade) {
        if (!s->active.empty()) break;  // after the running requests release memory
        if (!r->trimmed) {
          r->trimmed = true;
          if (trim_lazy(s)) continue;  // try again with the idle pools released
        }
        s->pending.pop_front();
        r->error = DE_ERR_NOMEM;
        r->done = true;
        r->t_done = now_s();
        s->cv_done.notify_all();
        continue;
      }
      if (vec)
        r->vectors = s->fams[k0].vectors + 1 + (b.event_solout != DE_SOLOUT_DEFAULT ? 1 : 0);
      // A pool whose range table is full takes no new work until it drains;
      // then (no kernel is in flight between iterations) its table restarts.
      bool held = false;
      for (int k : needs) {
        Family& f = s->fams[k];
        if ((int64_t)f.host_ranges.size() + 1 < f.max_ranges) continue;
        if (f.active_reqs > 0) {
          held = true;
          break;
        }
        f.reset(f, s->ctl);
        f.host_ranges.clear();
        f.published = 0;
        s->range_resets += 1;
      }
      if (held) break;
      r->slot = s->free_slots.back();
      std::vector<int32_t*> d_lists(s->fams.size(), nullptr);
      if (!upload(s, r, lists, d_lists.data())) {
        r->slot = -1;
        size_t dev = 0, host = 0;
        footprint(r, &dev, &host);
        const bool never = dev + 255 > s->arena.size || host + 255 > s->harena.size;
        if (!never && s->active.empty() && dev + 255 > s->arena.largest_free() && !r->trimmed) {
          r->trimmed = true;
          if (trim_lazy(s)) continue;  // try again with the idle pools released
        }
        if (never || (s->active.empty() && dev + 255 > s->arena.largest_free())) {
          s->pending.pop_front();
          r->error = DE_ERR_NOMEM;
          r->done = true;
          r->t_done = now_s();
          s->cv_done.notify_all();
          continue;
        }
        break;
      }
      s->free_slots.pop_back();
      s->pending.pop_front();
      s->active[r->slot] = r;
      r->first_iter = s->launched;
      admitted.push_back(r);
      if (classes) {
        for (int k : needs) {
          publish(s, s->fams[k], r->slot, 0, (int32_t)lists[k].size(), d_lists[k]);
          r->pools.push_back(k);
        }
      } else {
        const int k = needs[0];
        publish(s, s->fams[k], r->slot, 0, (int32_t)b.n, nullptr);
        r->pools.push_back(k);
      }
      for (int k : r->pools) s->fams[k].active_reqs += 1;
    }
    // The uploads and range entries are on the device before the commit
    // kernels, which run after the iteration in flight (stream order on ctl).
    bool dirty = false;
    for (auto& f : s->fams) dirty = dirty || f.dirty;
    if (dirty) {
      (void)hipEventRecord(s->ev_upl, s->upl);
      (void)hipStreamSynchronize(s->upl);
      s->stage_used = 0;
      for (Request* r : admitted) r->drop_inputs();
      admitted.clear();
      (void)hipStreamWaitEvent(s->ctl, s->ev_upl, 0);
      for (auto& f : s->fams) {
        if (!f.dirty) continue;
        hipLaunchKernelGGL(commit_pool, dim3(1), dim3(64), 0, s->ctl, f.q, (int64_t)f.published,
                           (int64_t)f.host_ranges.size());
        f.dirty = false;
      }
    }
    const double t_a1 = now_s();
    s->t_admit += t_a1 - t_a0;
    if (s->active.empty()) {
      // Nothing left to run: let the iteration in flight (if any) finish.
      if (inflight >= 0) {
        lk.unlock();
        (void)hipEventSynchronize(s->ev_counts[inflight & 1]);
        lk.lock();
        inflight = -1;
        continue;
      }
      if (s->stop) break;
      s->cv_work.wait(lk);
      continue;
    }
    // Queue iteration k; its counts land in h_counts[k & 1].
    const int64_t k = s->launched++;
    if (s->graph) {
      (void)hipEventRecord(s->fork, s->ctl);
      (void)hipGraphLaunch(s->graph, s->ctl);
      launch_all(s, true, true);
    } else {
      launch_all(s, true);
    }
    // Remaining counts of the request slots up to the highest active one:
    // gathered on the device, then one copy (not one per request).
    const int span = s->active.rbegin()->first + 1;
    hipLaunchKernelGGL(gather_remaining, dim3((span + 255) / 256), dim3(256), 0, s->ctl,
                       (const de::serve::ReqDev*)s->d_reqs, s->d_remaining, span);
    (void)hipMemcpyAsync(s->h_counts[k & 1], s->d_remaining, sizeof(int32_t) * span,
                         hipMemcpyDeviceToHost, s->ctl);
    (void)hipEventRecord(s->ev_counts[k & 1], s->ctl);
    const int64_t prev = inflight;
    inflight = k;
    lk.unlock();
    // The previous iteration's counts (the device has iteration k queued).
    hipError_t e = prev >= 0 ? hipEventSynchronize(s->ev_counts[prev & 1]) : hipSuccess;
    lk.lock();
    const double t_i1 = now_s();
    s->t_iter += t_i1 - t_a1;
    s->iterations += 1;
    if (e != hipSuccess) {
      s->error = DE_ERR_HIP;
      for (auto& kv : s->active) {
        kv.second->error = DE_ERR_HIP;
        kv.second->done = true;
      }
      s->active.clear();
      s->cv_done.notify_all();
      break;
    }
    if (prev < 0) continue;
    // Requests whose count reached zero by iteration prev (counts of requests
    // admitted later are not theirs yet).
    const int32_t* counts = s->h_counts[prev & 1];
    std::vector<int> finished;
    for (auto& kv : s->active)
      if (kv.second->first_iter <= prev && counts[kv.first] <= 0) finished.push_back(kv.first);
    if (!s->graph) {
      // Backlog: remaining trajectories over the slots of the pools in use.
      double remaining = 0.0, slots = 0.0;
      for (auto& kv : s->active)
        remaining += kv.second->first_iter <= prev ? std::max(counts[kv.first], 0)
                                                   : (double)kv.second->b.n;
      for (auto& f : s->fams)
        if (f.active_reqs > 0) slots += (double)f.nslots / f.g;
      if (remaining > slots) s->chunk = kMaxChunk;
      else if (remaining < 0.25 * slots) s->chunk = kChunk;
    }
    if (!finished.empty()) {
      const double t_done = now_s();
      for (int slot : finished) {
        Request* r = s->active[slot];
        for (int q : r->pools) s->fams[q].active_reqs -= 1;
        s->arena.release(r->dmem, r->dbytes);  // the results stay until delivered
        r->dmem = nullptr;
        r->done = true;
        r->t_done = t_done;
        s->active.erase(slot);
        s->free_slots.push_back(slot);
      }
      s->cv_done.notify_all();
    }
    s->t_retire += now_s() - t_i1;
  }
  if (inflight >= 0) (void)hipEventSynchronize(s->ev_counts[inflight & 1]);
}

template <class Slot>
bool make_pool(Family& f, int nslots) {
  f.nslots = nslots;
  set_pool_kind<Slot>(f);
  if (hipMalloc(&f.slots, sizeof(Slot) * (size_t)nslots) != hipSuccess) return false;
  f.init(f, 0);
  if (hipMalloc(&f.q, sizeof(de::serve::Queue)) != hipSuccess) return false;
  if (hipMemset(f.q, 0, sizeof(de::serve::Queue)) != hipSuccess) return false;
  if (hipMalloc(&f.ranges, sizeof(Range) * (size_t)f.max_ranges) != hipSuccess) return false;
  if (hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking) != hipSuccess) return false;
  return true;
}

}  // namespace

extern "C" {

int de_server_create(size_t hbm_budget_bytes, int32_t flags, de_server** out) {
  if (!out) return DE_ERR_INVALID;
  *out = nullptr;
  de_server* s = new de_server();
  s->budget = hbm_budget_bytes;
  s->flags = flags;
  // Range-table capacity per pool: kMaxRanges, or DE_SERVER_MAX_RANGES (at
  // least 4) to exercise the table's reuse with few requests.
  int64_t max_ranges = kMaxRanges;
  if (const char* env = getenv("DE_SERVER_MAX_RANGES")) {
    long long v = atoll(env);
    if (v >= 4 && v <= kMaxRanges) max_ranges = v;
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  s->ncu = ncu;
  s->max_ranges = max_ranges;
  // Slot pools: Lorenz and Van der Pol lanes, one pool per spring class, and
  // the DOP853 lanes of Lorenz and Van der Pol.
  struct Spec {
    const char* name;
    int cap, g;
  };
  const Spec specs[] = {{"lorenz", 0, 1},     {"vdp", 0, 1},       {"spring4", 4, 4},
                        {"spring8", 8, 8},    {"spring16", 16, 16}, {"spring32", 32, 32},
                        {"spring64", 64, 64}, {"spring128", 128, 64}, {"spring256", 256, 64},
                        {"lorenz_dop853", 0, 1}, {"vdp_dop853", 0, 1},
                        {"lorenz_event", 0, 1},  {"lorenz_cross", 0, 1},
                        {"vdp_event", 0, 1},     {"vdp_cross", 0, 1}};
  static_assert(sizeof specs / sizeof specs[0] == kEagerPools, "pool indices");
  s->fams.resize(kPools);
  const int lane_slots = ncu * 4 * 256;  // 4 blocks of 256 lanes per CU
  const int event_slots = ncu * 256;     // event requests: one block per CU
  const int group_lanes = ncu * 2 * 256;
  bool ok = true;
  for (int k = kEagerPools; k < kPools; ++k) {  // made on first use
    Family& f = s->fams[k];
    f.name = k >= kSpringVec ? "spring_lane" : "lane";
    f.lazy = true;
    f.nslots = event_slots;
    f.max_ranges = std::min<int64_t>(max_ranges, kLazyRanges);
  }
  for (size_t k = 0; k < (size_t)kEagerPools && ok; ++k) {
    Family& f = s->fams[k];
    f.name = specs[k].name;
    f.cap = specs[k].cap;
    f.g = specs[k].g;
    f.max_ranges = max_ranges;
    switch (k) {
      case 0: ok = make_pool<LaneSlot<de::Dopri5Lane<de::LorenzRhs, true>>>(f, lane_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::LorenzRhs, true>>; break;
      case 1: ok = make_pool<LaneSlot<de::Dopri5Lane<de::VanDerPolRhs, true>>>(f, lane_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::VanDerPolRhs, true>>; break;
      case kLorenzDop853: ok = make_pool<LaneSlot<de::Dop853Lane<de::LorenzRhs, true>>>(f, lane_slots);
        f.launch = launch_lane<de::Dop853Lane<de::LorenzRhs, true>>; break;
      case kVdpDop853: ok = make_pool<LaneSlot<de::Dop853Lane<de::VanDerPolRhs, true>>>(f, lane_slots);
        f.launch = launch_lane<de::Dop853Lane<de::VanDerPolRhs, true>>; break;
      case kLorenzEvent:
        ok = make_pool<LaneSlot<de::Dopri5Lane<de::LorenzRhs, true, false, 1>>>(f, event_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::LorenzRhs, true, false, 1>>; break;
      case kLorenzCross:
        ok = make_pool<LaneSlot<de::Dopri5Lane<de::LorenzRhs, false, false, 2>>>(f, event_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::LorenzRhs, false, false, 2>>; break;
      case kVdpEvent:
        ok = make_pool<LaneSlot<de::Dopri5Lane<de::VanDerPolRhs, true, false, 1>>>(f, event_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::VanDerPolRhs, true, false, 1>>; break;
      case kVdpCross:
        ok = make_pool<LaneSlot<de::Dopri5Lane<de::VanDerPolRhs, false, false, 2>>>(f, event_slots);
        f.launch = launch_lane<de::Dopri5Lane<de::VanDerPolRhs, false, false, 2>>; break;
      case 2: ok = make_pool<LaneSlot<de::GroupDopri5<4, 1>>>(f, group_lanes);
        f.launch = launch_group<4, 1>; break;
      case 3: ok = make_pool<LaneSlot<de::GroupDopri5<8, 1>>>(f, group_lanes);
        f.launch = launch_group<8, 1>; break;
      case 4: ok = make_pool<LaneSlot<de::GroupDopri5<16, 1>>>(f, group_lanes);
        f.launch = launch_group<16, 1>; break;
      case 5: ok = make_pool<LaneSlot<de::GroupDopri5<32, 1>>>(f, group_lanes);
        f.launch = launch_group<32, 1>; break;
      case 6: ok = make_pool<LaneSlot<de::GroupDopri5<64, 1>>>(f, group_lanes);
        f.launch = launch_group<64, 1>; break;
      case 7: ok = make_pool<LaneSlot<de::GroupDopri5<64, 2>>>(f, group_lanes);
        f.launch = launch_group<64, 2>; break;
      case 8: ok = make_pool<LaneSlot<de::GroupDopri5<64, 4>>>(f, group_lanes);
        f.launch = launch_group<64, 4>; break;
      default: ok = false; break;
    }
  }
  size_t pools = 0;
  for (auto& f : s->fams)  // lazy pools take theirs from the arena when made
    if (!f.lazy) pools += f.slot_bytes * f.nslots + sizeof(Range) * (size_t)f.max_ranges;
  ok = ok && hipMalloc(&s->d_reqs, sizeof(de::serve::ReqDev) * kMaxReq) == hipSuccess;
  for (auto& h : s->h_counts)
    ok = ok && hipHostMalloc(&h, sizeof(int32_t) * kMaxReq, 0) == hipSuccess;
  ok = ok && hipMalloc(&s->d_remaining, sizeof(int32_t) * kMaxReq) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->ctl, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->upl, hipStreamNonBlocking) == hipSuccess;
  for (auto& e : s->ev_counts)
    ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&s->ev_upl, hipEventDisableTiming) == hipSuccess;
  s->stage_size = (size_t)64 << 20;
  ok = ok && hipHostMalloc(&s->stage, s->stage_size, 0) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&s->fork, hipEventDisableTiming) == hipSuccess;
  s->fam_done.resize(s->fams.size());
  for (auto& e : s->fam_done) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  if (ok && hbm_budget_bytes <= pools) ok = false;
  ok = ok && s->arena.init(hbm_budget_bytes - pools, false);
  // Results go to pinned host memory that the kernels write directly, so a
  // finished request is retired without a device copy while the device keeps
  // iterating: min(budget / 8, 4 GiB), or DE_SERVER_HOST_BYTES.
  size_t host_bytes = std::min(hbm_budget_bytes / 8, (size_t)4 << 30);
  if (const char* env = getenv("DE_SERVER_HOST_BYTES")) {
    long long v = atoll(env);
    if (v >= (1 << 20)) host_bytes = (size_t)v;
  }
  ok = ok && s->harena.init(host_bytes, true);
  ok = ok && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    de_server_destroy(s);
    return DE_ERR_NOMEM;
  }
  s->h_reqs.resize(kMaxReq);
  for (int i = kMaxReq - 1; i >= 0; --i) s->free_slots.push_back(i);
  if (flags & kServerGraph) {
    if (!capture(s)) {
      de_server_destroy(s);
      return DE_ERR_HIP;
    }
  }
  s->thread = std::thread(scheduler, s);
  *out = s;
  return DE_OK;
}

int de_server_submit(de_server* s, const de_ode_batch* b, uint64_t* ticket) {
  if (!s || !b || !ticket) return DE_ERR_INVALID;
  if (b->family != DE_FAMILY_LORENZ && b->family != DE_FAMILY_VAN_DER_POL &&
      b->family != DE_FAMILY_SPRING_CHAIN)
    return DE_ERR_UNSUPPORTED;
  // Every method for every served family, with or without an event solout.
  if (b->method < DE_METHOD_DOPRI5 || b->method > DE_METHOD_RKV989E) return DE_ERR_INVALID;
  if (b->n <= 0 || !b->y0 || (b->n_eval > 0 && !b->t_eval)) return DE_ERR_INVALID;
  // Served requests carry scalar or per-trajectory tolerances and the final
  // state / t_eval output, or an event solout (no planes on spring chains),
  // validated as de_ode_batch_solve validates it; per-component vectors and
  // point outputs go through de_ode_batch_solve.
  if (b->rtol_comp || b->atol_comp) return DE_ERR_UNSUPPORTED;
  if (b->points != DE_POINTS_NONE) return DE_ERR_UNSUPPORTED;  // final state, t_eval or events
  const int ev_fn = b->event_solout == DE_SOLOUT_CROSSING     ? DE_EVENT_COMPONENT
                    : b->event_solout == DE_SOLOUT_HYPERPLANE ? DE_EVENT_PLANE
                                                               : b->event_fn;
  if (b->event_solout != DE_SOLOUT_DEFAULT) {
    if (b->event_solout < DE_SOLOUT_EVENT || b->event_solout > DE_SOLOUT_HYPERPLANE)
      return DE_ERR_INVALID;
    if (b->family == DE_FAMILY_SPRING_CHAIN && ev_fn == DE_EVENT_PLANE)
      return DE_ERR_UNSUPPORTED;
    if (b->event_solout != DE_SOLOUT_EVENT && b->n_eval > 0) return DE_ERR_INVALID;
    if (b->event_direction < DE_CROSS_BOTH || b->event_direction > DE_CROSS_NEGATIVE)
      return DE_ERR_INVALID;
    if (b->event_terminate < 0 || b->max_events < 0) return DE_ERR_INVALID;
    if (ev_fn < DE_EVENT_COMPONENT || ev_fn > DE_EVENT_PLANE) return DE_ERR_INVALID;
    if (ev_fn == DE_EVENT_COMPONENT &&
        (b->event_component < 0 ||
         (b->family != DE_FAMILY_SPRING_CHAIN && b->event_component >= b->dim)))
      return DE_ERR_INVALID;
    if (ev_fn == DE_EVENT_PLANE && (!b->event_point || !b->event_normal)) return DE_ERR_INVALID;
  }
  const bool ragged = b->family == DE_FAMILY_SPRING_CHAIN;
  if (ragged && (!b->sizes || !b->state_offset || !b->param_offset)) return DE_ERR_INVALID;
  if (!ragged && b->dim != (b->family == DE_FAMILY_LORENZ ? 3 : 2)) return DE_ERR_INVALID;
  Request* r = new Request();
  r->b = *b;
  const int64_t n = b->n;
  if (ragged) {
    r->sizes.assign(b->sizes, b->sizes + n);
    for (int64_t i = 0; i < n; ++i)
      if (r->sizes[i] < 1 || r->sizes[i] > 256) {
        delete r;
        return DE_ERR_UNSUPPORTED;
      }
    if (b->event_solout != DE_SOLOUT_DEFAULT && ev_fn == DE_EVENT_COMPONENT)
      for (int64_t i = 0; i < n; ++i)
        if (b->event_component >= 2 * r->sizes[i]) {  // past this system's state
          delete r;
          return DE_ERR_INVALID;
        }
    r->so.assign(b->state_offset, b->state_offset + n + 1);
    r->po.assign(b->param_offset, b->param_offset + n + 1);
    r->state_len = r->so[n];
    r->params.assign(b->params, b->params + r->po[n]);
  } else {
    r->state_len = (int64_t)b->dim * n;
    int P = b->family == DE_FAMILY_LORENZ ? 3 : 1;
    r->params.assign(b->params, b->params + P * n);
  }
  r->y0.assign(b->y0, b->y0 + r->state_len);
  if (b->t0) r->t0.assign(b->t0, b->t0 + n);
  if (b->tf) r->tf.assign(b->tf, b->tf + n);
  if (b->rtol_traj) r->rtol.assign(b->rtol_traj, b->rtol_traj + n);
  if (b->atol_traj) r->atol.assign(b->atol_traj, b->atol_traj + n);
  if (b->n_eval > 0) r->t_eval.assign(b->t_eval, b->t_eval + b->n_eval);
  if (b->event_solout != DE_SOLOUT_DEFAULT) {
    if (b->event_value_traj) r->ev_value.assign(b->event_value_traj, b->event_value_traj + n);
    if (ev_fn == DE_EVENT_PLANE) {
      r->ev_point.assign(b->event_point, b->event_point + b->dim);
      r->ev_normal.assign(b->event_normal, b->event_normal + b->dim);
      if (b->event_solout == DE_SOLOUT_HYPERPLANE) {
        // HyperplaneCrossingSolout::new (src/solout/hyperplane.rs:126-132): the
        // normal times 1/|normal| when |normal| > f64::EPSILON, |normal| = sqrt
        // of the squares summed from zero (the batch path's operations).
        double sq = 0.0;
        for (double v : r->ev_normal) sq = sq + v * v;
        const double norm = sqrt(sq), inv = 1.0 / norm;
        if (norm > 2.220446049250313080847e-16)
          for (double& v : r->ev_normal) v = v * inv;
      }
    }
  }
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) {
    delete r;
    return s->error;
  }
  r->ticket = s->next_ticket++;
  s->all[r->ticket] = r;
  s->pending.push_back(r);
  *ticket = r->ticket;
  s->cv_work.notify_one();
  return DE_OK;
}

int de_server_poll(de_server* s, uint64_t ticket, de_ode_result* out, int32_t wait) {
  if (!s) return DE_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return DE_ERR_INVALID;
  Request* r = it->second;
  while (!r->done) {
    if (!wait) return 0;
    s->cv_done.wait(lk);
  }
  // Delivered once: the results leave the host arena and the request is
  // forgotten (its latency stays queryable).
  const int err = r->error;
  if (!err && out) {
    const de::OdeArgs& a = r->out;
    const int64_t n = r->b.n;
    const size_t sl = (size_t)r->state_len;
    auto put = [](void* dst, const void* src, size_t bytes) {
      if (dst && src && bytes) memcpy(dst, src, bytes);
    };
    put(out->y_final, a.y_final, sl * 8);
    put(out->t_final, a.t_final, n * 8);
    put(out->accepted, a.accepted, n * 8);
    put(out->rejected, a.rejected, n * 8);
    put(out->evals, a.evals, n * 8);
    put(out->status, a.status, n * 4);
    if (a.dense) {
      put(out->dense, a.dense, (size_t)r->b.n_eval * sl * 8);
      put(out->n_dense, a.n_dense, n * 4);
    }
    if (a.ev.n_out) {
      const size_t m = (size_t)a.ev.max_events;
      put(out->event_t, a.ev.t_out, m * n * 8);
      put(out->event_y, a.ev.y_out, m * sl * 8);
      put(out->n_events, a.ev.n_out, n * 4);
    }
  }
  s->latency[ticket] = r->t_done - r->t_submit;
  if (r->hmem) s->harena.release(r->hmem, r->hbytes);
  s->all.erase(it);
  delete r;
  s->cv_work.notify_all();  // the freed results may let a waiting request in
  return err ? err : 1;
}

double de_server_latency(de_server* s, uint64_t ticket) {
  if (!s) return -1.0;
  std::lock_guard<std::mutex> lk(s->mu);
  auto d = s->latency.find(ticket);
  if (d != s->latency.end()) return d->second;
  auto it = s->all.find(ticket);
  if (it == s->all.end() || !it->second->done) return -1.0;
  return it->second->t_done - it->second->t_submit;
}

int de_server_stats(de_server* s, int64_t* iterations, int64_t* range_resets) {
  if (!s) return DE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (iterations) *iterations = s->iterations;
  if (range_resets) *range_resets = s->range_resets;
  return DE_OK;
}

int de_server_pools(de_server* s, int32_t* made, int64_t* released) {
  if (!s) return DE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  int32_t m = 0;
  for (auto& f : s->fams) m += f.lazy && f.slots ? 1 : 0;
  if (made) *made = m;
  if (released) *released = s->pools_released;
  return DE_OK;
}

int de_server_timing(de_server* s, double* admit_s, double* iterate_s, double* retire_s) {
  if (!s) return DE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (admit_s) *admit_s = s->t_admit;
  if (iterate_s) *iterate_s = s->t_iter;
  if (retire_s) *retire_s = s->t_retire;
  return DE_OK;
}

void de_server_destroy(de_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->thread.joinable()) s->thread.join();
  (void)hipDeviceSynchronize();
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  for (auto& f : s->fams) {
    if (!f.lazy) {  // lazy pools live in the arena
      if (f.slots) (void)hipFree(f.slots);
      if (f.q) (void)hipFree(f.q);
      if (f.ranges) (void)hipFree(f.ranges);
    }
    if (f.stream) (void)hipStreamDestroy(f.stream);
  }
  for (auto& kv : s->all) delete kv.second;
  s->arena.destroy();
  s->harena.destroy();
  if (s->d_reqs) (void)hipFree(s->d_reqs);
  for (auto h : s->h_counts)
    if (h) (void)hipHostFree(h);
  if (s->d_remaining) (void)hipFree(s->d_remaining);
  for (auto& e : s->fam_done)
    if (e) (void)hipEventDestroy(e);
  if (s->fork) (void)hipEventDestroy(s->fork);
  for (auto e : s->ev_counts)
    if (e) (void)hipEventDestroy(e);
  if (s->ev_upl) (void)hipEventDestroy(s->ev_upl);
  if (s->stage) (void)hipHostFree(s->stage);
  if (s->ctl) (void)hipStreamDestroy(s->ctl);
  if (s->upl) (void)hipStreamDestroy(s->upl);
  delete s;
}

}  // extern "C"
