/* C-ABI of the MI355X ensemble Runge-Kutta integrator.
 *
 * The reference crate solves ONE trajectory per call:
 *   IVP::ode(&sys, t0, tf, y0).method(ExplicitRungeKutta::dopri5()).solve()
 *     -> solve_ode (src/ode/solve_ivp.rs:116-277)
 *   IVP::sde(&mut sys, t0, tf, y0).method(ExplicitRungeKutta::euler(dt) |
 *     Milstein::new(dt)).solve() -> solve_sde (src/sde/solve_ivp.rs:135-287)
 * Each entry point below runs that same call for a whole ensemble of
 * trajectories on one GPU.  Per-trajectory outcomes never fail the call: they
 * come back in the status array (the Ok/Err of each reference solve).  Call
 * failures (bad descriptor, HIP error) return a negative code and set a
 * thread-local message readable with de_last_error().
 *
 * Memory: by default every pointer is device memory (HBM) and the call is
 * asynchronous on `stream` (a hipStream_t, NULL = default stream).  With
 * DE_FLAG_HOST_MEMORY every pointer is host memory; the library stages copies
 * and returns after the results are back on the host.
 *
 * Layouts (structure of arrays, trajectory index fastest, coalesced on HBM):
 *   fixed-dimension families: y0/y_final[c * n + i], params[p * n + i],
 *     dense[(e * dim + c) * n + i]
 *   spring chain (ragged):    system i owns y0/y_final[state_offset[i] ..
 *     state_offset[i+1]) = [x_0..x_{N-1}, v_0..v_{N-1}], params[param_offset[i]
 *     .. param_offset[i+1]) = [c, k_0..k_N], dense[n_eval * state_offset[i] +
 *     e * 2N + c].
 */
#ifndef DE_API_H
#define DE_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Call return codes. */
#define DE_OK 0
#define DE_ERR_INVALID (-1)     /* malformed descriptor */
#define DE_ERR_HIP (-2)         /* HIP runtime failure */
#define DE_ERR_UNSUPPORTED (-3) /* family/method/mapping combination not built */
#define DE_ERR_NOMEM (-4)       /* device allocation or budget exceeded */

/* Per-trajectory status (Solution status / Error variant of one solve,
 * src/status.rs:26, src/error.rs:13-41). */
#define DE_STATUS_COMPLETE 0
#define DE_STATUS_MAX_STEPS 1  /* Error::MaxSteps  */
#define DE_STATUS_STEP_SIZE 2  /* Error::StepSize  */
#define DE_STATUS_STIFFNESS 3  /* Error::Stiffness */
#define DE_STATUS_BAD_INPUT 4  /* Error::BadInput  */
#define DE_STATUS_INTERRUPTED 5 /* Status::Interrupted: an event solout returned Terminate
                                   (src/ode/solve_ivp.rs:254-258) */
#define DE_STATUS_OUT_OF_BOUNDS 6 /* the reference solve panics here: a solout unwraps an
                                     Interpolation::interpolate that returned
                                     Error::OutOfBounds (t outside [t_prev, t] of the step, as
                                     in every interpolation of a backward solve) */

/* ODE families (device-compiled right-hand sides). */
#define DE_FAMILY_LORENZ 0       /* dim 3, params sigma, rho, beta */
#define DE_FAMILY_VAN_DER_POL 1  /* dim 2, params mu */
#define DE_FAMILY_SPRING_CHAIN 2 /* dim 2N ragged, params c, k_0..k_N */
#define DE_FAMILY_EXPONENTIAL 3  /* dim 1, params k */
#define DE_FAMILY_LINEAR 4       /* dim 1, params a, b */
#define DE_FAMILY_HARMONIC 5     /* dim 2, params k */
#define DE_FAMILY_LOGISTIC 6     /* dim 1, params k, m */
#define DE_FAMILY_ROBERTSON 7    /* dim 3, no params */

/* SDE families. */
#define DE_SDE_ORNSTEIN_UHLENBECK 0 /* dim D, params theta_c, mu_c, sigma_c */
#define DE_SDE_LINEAR_DIAGONAL 1    /* dim D, params a_c, b_c (GBM for D=1) */

/* Methods: ExplicitRungeKutta::dopri5() (adaptive), ::rk4(h) and ::euler(h)
 * (fixed step), src/methods/erk. */
#define DE_METHOD_DOPRI5 0
#define DE_METHOD_RK4 1
#define DE_METHOD_EULER 2
#define DE_METHOD_DOP853 3 /* ExplicitRungeKutta::dop853 (dormandprince/mod.rs:46-52) */
/* The embedded family ExplicitRungeKutta<_, Adaptive, ..> (erk/adaptive/mod.rs:43-122):
 * inf-norm error, stiffness = max_rejects consecutive rejections. */
#define DE_METHOD_RKF45 4     /* ::rkf45(), Fehlberg 4(5) */
#define DE_METHOD_CASH_KARP 5 /* ::cash_karp() */
#define DE_METHOD_RKV655E 6   /* ::rkv655e() .. ::rkv989e(): Verner pairs with interpolants */
#define DE_METHOD_RKV656E 7
#define DE_METHOD_RKV766E 8
#define DE_METHOD_RKV767E 9
#define DE_METHOD_RKV877E 10
#define DE_METHOD_RKV878E 11
#define DE_METHOD_RKV988E 12
#define DE_METHOD_RKV989E 13

/* SDE schemes: ExplicitRungeKutta::euler(dt) on an SDE (Euler-Maruyama,
 * src/methods/erk/fixed/stochastic.rs) and Milstein::new(dt)
 * (src/methods/milstein.rs). */
#define DE_SCHEME_EULER_MARUYAMA 0
#define DE_SCHEME_MILSTEIN 1

/* SDE noise sources (what SDE::noise fills dW with). */
#define DE_NOISE_PHILOX 0   /* sqrt(h) * N(0,1) from Philox4x32-10(seed,path,step,c) */
#define DE_NOISE_TABLE 1    /* noise_table[step * dim + c], 0 after the table ends */
#define DE_NOISE_CONSTANT 2 /* noise_constant for every component and step */

/* Launch schedules for the ensemble. */
#define DE_SCHED_AUTO 0       /* library choice (autotuned) */
#define DE_SCHED_PLAIN 1      /* one thread per trajectory, one launch */
#define DE_SCHED_PERSISTENT 2 /* resident lanes refilled from a work counter, longest
                                 predicted trajectories first */
#define DE_SCHED_PERSISTENT_FIFO 3 /* as PERSISTENT, trajectories taken in index order */

/* Mappings of a system onto lanes (spring chain). */
#define DE_MAP_AUTO 0  /* autotuner picks per size class */
#define DE_MAP_LANE 1  /* one lane per system, state in HBM */
#define DE_MAP_GROUP 2 /* `group_size` lanes per system, state in registers */

/* Solouts beside DefaultSolout / TEvalSolout (src/solout, builders src/ivp.rs:662-705). */
#define DE_SOLOUT_DEFAULT 0    /* DefaultSolout, or TEvalSolout when n_eval > 0 */
#define DE_SOLOUT_EVENT 1      /* IVP::event: EventWrappedSolout around the above, root by
                                  Brent-Dekker on the step interpolant (src/solout/event.rs) */
#define DE_SOLOUT_CROSSING 2   /* IVP::crossing: CrossingSolout, Newton on the interpolant
                                  (src/solout/crossing.rs); replaces the t_eval output */
#define DE_SOLOUT_HYPERPLANE 3 /* IVP::hyperplane_crossing with the identity extractor
                                  (src/solout/hyperplane.rs) */

/* Solution points of the output strategies that push a variable number of points
 * (src/ivp.rs:643-655 and the default):
 * de_ode_batch.points selects one; the points come back in de_ode_result.points_*. */
#define DE_POINTS_NONE 0  /* final state only (and t_eval / events when asked) */
#define DE_POINTS_STEPS 1 /* DefaultSolout: (t0, y0) and every accepted step
                             (src/solout/default.rs:71-91) */
#define DE_POINTS_DENSE 2 /* IVP::dense(n): DenseSolout, n - 1 interpolated points inside
                             every accepted step, then the step (src/solout/dense.rs:81-111) */
#define DE_POINTS_EVEN 3  /* IVP::even(dt): EvenSolout, t0, t0 + dt, ... and tf
                             (src/solout/even.rs:83-190) */

/* Event functions g(t, y) (Event::event, src/solout/event.rs) of DE_SOLOUT_EVENT. */
#define DE_EVENT_COMPONENT 0 /* y[event_component] - value */
#define DE_EVENT_TIME 1      /* t - value */
#define DE_EVENT_PLANE 2     /* y.minus(&point).dot(&normal), normal as given */

/* CrossingDirection (src/solout/mod.rs:38-45). */
#define DE_CROSS_BOTH 0
#define DE_CROSS_POSITIVE 1
#define DE_CROSS_NEGATIVE 2

/* Flags. */
#define DE_FLAG_HOST_MEMORY 1u /* all pointers are host memory */
#define DE_FLAG_FAST_MATH 2u   /* FMA-contracted kernels: faster, not bit-equal */

typedef struct de_ode_batch {
  int32_t family; /* DE_FAMILY_* */
  int32_t method; /* DE_METHOD_* */
  int64_t n;      /* number of trajectories */
  int32_t dim;    /* state dimension of fixed-size families (ignored for ragged) */
  int32_t n_eval; /* number of t_eval points (0: final state only) */
  const int32_t* sizes;        /* spring chain: masses N per system */
  const int64_t* state_offset; /* spring chain: n+1 prefix offsets */
  const int64_t* param_offset; /* spring chain: n+1 prefix offsets */
  const double* params;
  const double* y0;
  const double* t0; /* per trajectory, or NULL to use t0_scalar */
  const double* tf; /* per trajectory, or NULL to use tf_scalar */
  double t0_scalar;
  double tf_scalar;
  /* Method settings; de_ode_batch_init fills the reference defaults of
   * ExplicitRungeKutta (src/methods/erk/mod.rs:115-158). */
  double rtol;        /* Tolerance::Scalar(rtol) */
  double atol;
  const double* rtol_traj; /* per-trajectory scalar rtol, or NULL */
  const double* atol_traj;
  double h0;          /* 0: automatic (h_init for DOPRI5, (tf-t0)/100 for RK4) */
  double h_min;
  double h_max;
  int64_t max_steps;  /* counts attempted steps, as the reference does */
  double safety_factor;
  double min_scale;
  double max_scale;
  const double* t_eval; /* n_eval points (TEvalSolout semantics), sorted */
  int32_t schedule;   /* DE_SCHED_* */
  int32_t mapping;    /* DE_MAP_* */
  int32_t group_size; /* lanes per system for DE_MAP_GROUP (0: autotune) */
  uint32_t flags;     /* DE_FLAG_* */
  /* Per-component tolerances, Tolerance::Vector (src/tolerance.rs:11-68): one value
   * per state component of every trajectory, in the layout of y0 ([c*n + i], or at
   * state_offset[i] for spring chains), or NULL.  They take precedence over the
   * per-trajectory and scalar values; the norm uses atol[c] + rtol[c]*max(|y|,|y_new|)
   * (src/traits.rs:394-411) and h_init atol[c] + rtol[c]*|y0| (src/methods/h_init.rs:80). */
  const double* rtol_comp;
  const double* atol_comp;
  /* Event solouts (DE_SOLOUT_*; spring chains: component and time events and the
   * crossing solout).  The event points of each
   * trajectory -- the points the reference's event / crossing solout pushes into
   * Solution -- come back in de_ode_result.event_t / event_y. */
  int32_t event_solout;    /* DE_SOLOUT_* */
  int32_t event_fn;        /* DE_EVENT_*: g(t, y) of DE_SOLOUT_EVENT */
  int32_t event_component; /* component of DE_EVENT_COMPONENT and DE_SOLOUT_CROSSING */
  int32_t event_direction; /* DE_CROSS_*: EventConfig::direction / CrossingDirection */
  int32_t event_terminate; /* EventConfig::terminate: 0 = None, k > 0 = Some(k) */
  int32_t max_events;      /* event points kept per trajectory */
  double event_value;      /* value of DE_EVENT_COMPONENT / DE_EVENT_TIME, crossing threshold */
  const double* event_value_traj; /* per-trajectory value, or NULL for event_value */
  const double* event_point;      /* [dim] point of DE_EVENT_PLANE / DE_SOLOUT_HYPERPLANE */
  const double* event_normal;     /* [dim] normal (normalised by the hyperplane solout,
                                     src/solout/hyperplane.rs:126-132) */
  /* ExplicitRungeKutta::max_rejects of the embedded family (erk/mod.rs:81,140,201):
   * consecutive rejections that end the solve with Error::Stiffness. */
  int64_t max_rejects;
  /* Output strategy with a variable number of points (DE_POINTS_*; without t_eval
   * or an event solout; spring chains: DOPRI5 on the group mappings, the other
   * methods on one lane per system). */
  int32_t points;
  int32_t points_n;  /* DE_POINTS_DENSE: n (points per step, including the step) */
  double points_dt;  /* DE_POINTS_EVEN: dt */
} de_ode_batch;

typedef struct de_ode_result {
  double* y_final;   /* final state (state at failure for error statuses) */
  double* t_final;   /* [n] */
  int64_t* accepted; /* [n] Steps::accepted */
  int64_t* rejected; /* [n] Steps::rejected */
  int64_t* evals;    /* [n] Evals::function */
  int32_t* status;   /* [n] DE_STATUS_* */
  double* dense;     /* t_eval output (NaN where a point was not produced), or NULL */
  int32_t* n_dense;  /* [n] number of t_eval points produced, or NULL */
  double* event_t;   /* [max_events * n] event times, event_t[k * n + i]; NaN past the last */
  double* event_y;   /* [max_events * dim * n] event states, event_y[(k * dim + c) * n + i];
                        spring chains: event_y[max_events * state_offset[i] + k * 2 sizes[i]
                        + c], max_events * state_offset[n] values */
  int32_t* n_events; /* [n] event points pushed (those past max_events are not kept), or NULL */
  /* DE_POINTS_*: trajectory i's Solution points go to slots points_offset[i] ..
   * points_offset[i+1] (n + 1 values, given by the caller): time points_t[s], state
   * points_y[s * dim + c].  Spring chains (state length 2 sizes[i]) keep the states of
   * trajectory i's slots back to back from Y_i = sum_{j<i} (points_offset[j+1] -
   * points_offset[j]) * 2 sizes[j]: slot points_offset[i] + k at points_y[Y_i + k * 2
   * sizes[i] + c], points_y holding Y_n values.  n_points[i] is the number of points the
   * solout pushed; those
   * beyond the trajectory's slots are counted, not kept, and slots past n_points[i] are
   * left unspecified.  Slot counts that always
   * suffice: accepted + 1 (STEPS), n * accepted + 1 (DENSE), from the accepted counts
   * of a solve of the same ensemble (results are deterministic), and
   * floor(|tf - t0| / dt) + 3 (EVEN). */
  const int64_t* points_offset;
  double* points_t;
  double* points_y;
  int64_t* n_points;
} de_ode_result;

typedef struct de_sde_batch {
  int32_t family; /* DE_SDE_* */
  int32_t scheme; /* DE_SCHEME_* */
  int64_t n;      /* number of paths */
  int32_t dim;
  int32_t noise;  /* DE_NOISE_* */
  const double* params; /* [P * n] structure of arrays */
  const double* y0;     /* [dim * n] */
  double t0;
  double tf;
  double h0;            /* dt; 0: (tf - t0)/100 as the reference */
  double h_min;
  double h_max;
  int64_t max_steps;    /* reference default 10 000 */
  uint64_t seed;        /* DE_NOISE_PHILOX */
  uint64_t path_base;   /* global index of path 0 (noise key) */
  const double* noise_table; /* DE_NOISE_TABLE: [noise_steps * dim] */
  int64_t noise_steps;
  double noise_constant;
  int32_t path_stride;  /* >0: record y after every path_stride-th step */
  uint32_t flags;
} de_sde_batch;

typedef struct de_sde_result {
  double* y_final;  /* [dim * n] */
  double* t_final;  /* [n] */
  int64_t* steps;   /* [n] Steps::accepted (every SDE step is accepted) */
  int64_t* evals;   /* [n] Evals::function */
  int32_t* status;  /* [n] */
  double* path;     /* [n_rec * dim * n] with n_rec = steps / path_stride, or NULL */
  double* path_t;   /* [n_rec] times of the recorded states, or NULL */
} de_sde_result;

/* Fill a descriptor with the reference defaults: rtol = atol = 1e-6, h0 = 0,
 * h_min = 0, h_max = inf, max_steps = 10000, max_rejects = 100, safety 0.9, scales [0.2, 10]
 * (ExplicitRungeKutta::default, src/methods/erk/mod.rs:115-158). */
void de_ode_batch_init(de_ode_batch* b);
void de_sde_batch_init(de_sde_batch* b);

/* Batched solve_ode (src/ode/solve_ivp.rs:116-277) with ExplicitRungeKutta
 * dopri5 (src/methods/erk/dormandprince/ordinary.rs:16-336) or rk4
 * (src/methods/erk/fixed/ordinary.rs:16-220). */
int de_ode_batch_solve(const de_ode_batch* batch, de_ode_result* out, void* stream);

/* No counterpart in the reference: de_ode_batch_solve for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as de_ode_batch_solve; device
 * memory only (no DE_FLAG_HOST_MEMORY), fixed-dimension families.  The launch's
 * scratch memory (work counter, launch arguments, the longest-first order: about
 * 2 KiB + 5 bytes per trajectory, de_ode_noalloc_scratch_bytes) comes from one
 * 64 MiB device array per GPU that the library holds until the process ends;
 * larger requests fail with DE_ERR_INVALID.  Event solouts find their roots
 * inline.  Calls that overlap on one GPU share that array and race. */
int de_ode_batch_solve_noalloc(const de_ode_batch* batch, de_ode_result* out, void* stream);
int64_t de_ode_noalloc_scratch_bytes(const de_ode_batch* batch);

/* Batched solve_sde (src/sde/solve_ivp.rs:135-287) with Euler-Maruyama
 * (src/methods/erk/fixed/stochastic.rs:18-193) or Milstein
 * (src/methods/milstein.rs:76-224). */
int de_sde_batch_solve(const de_sde_batch* batch, de_sde_result* out, void* stream);

/* ------------------------------------------------------------ CSV writer --
 * Solution::to_csv (src/solution.rs:190-230) for every trajectory of an ODE
 * solve: one CSV document per trajectory, the header "t,y0,...,y<dim-1>\n"
 * then one "<t>,<y0>,...,<y dim-1>\n" line per point of the trajectory's
 * Solution, every number as Rust's `{:?}` of an f64 (the shortest digits
 * that read back as the same double, the closest of those; "NaN", "inf",
 * "-inf", "-0.0"; d[.ddd]e<exp> outside 1e-4 <= |x| < 1e16, else positional
 * with at least one fractional digit).  The documents are concatenated:
 * trajectory i's is out[offsets[i] .. offsets[i+1]), offsets[0] = 0.  A
 * trajectory without points gets an empty document (the reference's to_csv
 * creates the file, then panics indexing y[0]).
 *
 * `batch` and `result` are the solve's (n, dim, t0 / tf and t_eval, and the
 * result arrays of the chosen source); their pointers and offsets / out are
 * device memory unless batch->flags has DE_FLAG_HOST_MEMORY.  Fixed-dimension
 * families only (dim <= 8). */
#define DE_CSV_FINAL 0  /* the final state: the point (t_final, y_final) */
#define DE_CSV_TEVAL 1  /* the t_eval points produced: n_dense[i] of them, with the
                           t_eval values as times (TEvalSolout, src/solout/t_eval.rs) */
#define DE_CSV_EVENTS 2 /* the event points: n_events[i] of them (event_t / event_y) */
#define DE_CSV_POINTS 3 /* a point output (DE_POINTS_*): the kept points of trajectory i,
                           slots points_offset[i] .. + n_points[i] -- the whole Solution of
                           the reference's default solve with DE_POINTS_STEPS */

/* Document sizes: offsets[0 .. n] (n + 1 int64 values). */
int de_ode_csv_offsets(const de_ode_batch* batch, const de_ode_result* result, int32_t source,
                       int64_t* offsets, void* stream);
/* The documents, out[0 .. offsets[n]), from the offsets of de_ode_csv_offsets. */
int de_ode_csv_write(const de_ode_batch* batch, const de_ode_result* result, int32_t source,
                     const int64_t* offsets, char* out, void* stream);

/* Bytes of device workspace a de_ode_batch_solve call with this descriptor
 * allocates beyond its inputs and outputs (HBM admission budget). */
int64_t de_ode_workspace_bytes(const de_ode_batch* batch);

/* Message of the last failed call on this thread. */
const char* de_last_error(void);

/* Library version string. */
const char* de_version(void);

/* ---- Device-memory budget of the launch workspace ------------------------
 * No counterpart in the reference (it allocates per trajectory on the host).
 * Batch solves take their stream-ordered workspace (launch arguments, the
 * longest-first order, the bracket regions of deferred event roots, the
 * spring chains' class lists and one-lane vectors) from a
 * library pool that keeps freed blocks for the next solve.  One budget, 64 GiB
 * by default, bounds each workspace allocation and the blocks the pool keeps;
 * the bracket regions of one event pass stay within 3/4 of it (an ensemble
 * whose regions would not is solved in slices).  The serving arena is sized by
 * de_server_create's own budget.  Total device memory of the library: the
 * workspace budget plus the live servers' budgets. */
int de_set_workspace_budget(size_t bytes);   /* >= 1 MiB; trims kept blocks to it */
size_t de_workspace_budget(void);
int de_workspace_trim(void);                 /* return every kept block to the driver */

/* ---- Serving: continuous batching of ensemble solves on one GPU ----------
 * Requests (each a de_ode_batch over host memory) are admitted while their
 * inputs fit the HBM budget (slot pools included), their results fit a pinned
 * host arena of min(budget / 8, 4 GiB) (the environment variable
 * DE_SERVER_HOST_BYTES, read at creation, sets another size) and request slots
 * are free; their state stays resident on the device until the request
 * completes, and the kernels write each trajectory's results straight into the
 * host arena.  Requests admitted while others run share the persistent lanes
 * of the running kernels. */
typedef struct de_server de_server;
int de_server_create(size_t hbm_budget_bytes, int32_t flags, de_server** out);
/* Copies the request's inputs; returns a ticket.  Lorenz, Van der Pol and spring
 * chains with every method, with or without an event solout (no planes on spring
 * chains); final state, t_eval or events (no point outputs, no per-component
 * tolerances). */
int de_server_submit(de_server* s, const de_ode_batch* req, uint64_t* ticket);
/* Returns 1 and fills host-memory `out` when the request is done, 0 if it is
 * still running (wait = 0) or blocks until done (wait = 1). */
int de_server_poll(de_server* s, uint64_t ticket, de_ode_result* out, int32_t wait);
/* Submit-to-completion latency of a finished request, in seconds. */
double de_server_latency(de_server* s, uint64_t ticket);
/* Scheduler counters: iterations (chunk launches of the pools with work) and
 * restarts of a drained pool's range table (each pool holds up to 2^20
 * published ranges; the environment variable DE_SERVER_MAX_RANGES, read at
 * creation, lowers that capacity). */
int de_server_stats(de_server* s, int64_t* iterations, int64_t* range_resets);
/* Pools made on first use (RK4, Euler and the embedded family on lane families,
 * spring chains on one lane per system, event lanes of the methods other than
 * DOPRI5, spring chains with DOPRI5 and an event solout on the group mapping of
 * each size class), their memory taken from the budget's arena: how many are made now,
 * and how many times an idle one was released to make room for a request. */
int de_server_pools(de_server* s, int32_t* made, int64_t* released);
/* Scheduler wall time so far, in seconds: admitting requests (uploads and
 * publishing their work), iterating (launch to the remaining counts on the
 * host) and retiring finished requests (result copies). */
int de_server_timing(de_server* s, double* admit_s, double* iterate_s, double* retire_s);
void de_server_destroy(de_server* s);

#ifdef __cplusplus
}
#endif

#endif /* DE_API_H */
