"""ctypes mirror of include/de_api.h and include/de_workload.h."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "_build", "libde_amd.so")

DE_OK = 0
DE_ERR_INVALID = -1
DE_ERR_HIP = -2
DE_ERR_UNSUPPORTED = -3
DE_ERR_NOMEM = -4

DE_STATUS_COMPLETE = 0
DE_STATUS_MAX_STEPS = 1
DE_STATUS_STEP_SIZE = 2
DE_STATUS_STIFFNESS = 3
DE_STATUS_BAD_INPUT = 4
DE_STATUS_INTERRUPTED = 5
DE_STATUS_OUT_OF_BOUNDS = 6

DE_FAMILY_LORENZ = 0
DE_FAMILY_VAN_DER_POL = 1
DE_FAMILY_SPRING_CHAIN = 2
DE_FAMILY_EXPONENTIAL = 3
DE_FAMILY_LINEAR = 4
DE_FAMILY_HARMONIC = 5
DE_FAMILY_LOGISTIC = 6
DE_FAMILY_ROBERTSON = 7

FAMILY_DIM = {0: 3, 1: 2, 3: 1, 4: 1, 5: 2, 6: 1, 7: 3}
FAMILY_PARAMS = {0: 3, 1: 1, 3: 1, 4: 2, 5: 1, 6: 2, 7: 0}

DE_SDE_ORNSTEIN_UHLENBECK = 0
DE_SDE_LINEAR_DIAGONAL = 1

DE_METHOD_DOPRI5 = 0
DE_METHOD_RK4 = 1
DE_METHOD_EULER = 2
DE_METHOD_DOP853 = 3
DE_METHOD_RKF45 = 4
DE_METHOD_CASH_KARP = 5
DE_METHOD_RKV655E = 6
DE_METHOD_RKV656E = 7
DE_METHOD_RKV766E = 8
DE_METHOD_RKV767E = 9
DE_METHOD_RKV877E = 10
DE_METHOD_RKV878E = 11
DE_METHOD_RKV988E = 12
DE_METHOD_RKV989E = 13

DE_SCHEME_EULER_MARUYAMA = 0
DE_SCHEME_MILSTEIN = 1

DE_NOISE_PHILOX = 0
DE_NOISE_TABLE = 1
DE_NOISE_CONSTANT = 2

DE_SCHED_AUTO = 0
DE_SCHED_PLAIN = 1
DE_SCHED_PERSISTENT = 2
DE_SCHED_PERSISTENT_FIFO = 3

DE_MAP_AUTO = 0
DE_MAP_LANE = 1
DE_MAP_GROUP = 2

DE_CSV_FINAL = 0
DE_CSV_TEVAL = 1
DE_CSV_EVENTS = 2
DE_CSV_POINTS = 3
DE_SOLOUT_DEFAULT = 0
DE_SOLOUT_EVENT = 1
DE_SOLOUT_CROSSING = 2
DE_SOLOUT_HYPERPLANE = 3
DE_POINTS_NONE = 0
DE_POINTS_STEPS = 1
DE_POINTS_DENSE = 2
DE_POINTS_EVEN = 3

DE_EVENT_COMPONENT = 0
DE_EVENT_TIME = 1
DE_EVENT_PLANE = 2

DE_CROSS_BOTH = 0
DE_CROSS_POSITIVE = 1
DE_CROSS_NEGATIVE = 2

DE_FLAG_HOST_MEMORY = 1
DE_FLAG_FAST_MATH = 2

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class OdeBatch(C.Structure):
    _fields_ = [
        ("family", C.c_int32), ("method", C.c_int32),
        ("n", C.c_int64),
        ("dim", C.c_int32), ("n_eval", C.c_int32),
        ("sizes", _i32p), ("state_offset", _i64p), ("param_offset", _i64p),
        ("params", C.c_void_p), ("y0", C.c_void_p),
        ("t0", C.c_void_p), ("tf", C.c_void_p),
        ("t0_scalar", C.c_double), ("tf_scalar", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double),
        ("rtol_traj", C.c_void_p), ("atol_traj", C.c_void_p),
        ("h0", C.c_double), ("h_min", C.c_double), ("h_max", C.c_double),
        ("max_steps", C.c_int64),
        ("safety_factor", C.c_double), ("min_scale", C.c_double), ("max_scale", C.c_double),
        ("t_eval", C.c_void_p),
        ("schedule", C.c_int32), ("mapping", C.c_int32), ("group_size", C.c_int32),
        ("flags", C.c_uint32),
        ("rtol_comp", C.c_void_p), ("atol_comp", C.c_void_p),
        ("event_solout", C.c_int32), ("event_fn", C.c_int32),
        ("event_component", C.c_int32), ("event_direction", C.c_int32),
        ("event_terminate", C.c_int32), ("max_events", C.c_int32),
        ("event_value", C.c_double), ("event_value_traj", C.c_void_p),
        ("event_point", C.c_void_p), ("event_normal", C.c_void_p),
        ("max_rejects", C.c_int64),
        ("points", C.c_int32), ("points_n", C.c_int32), ("points_dt", C.c_double),
    ]


class OdeResult(C.Structure):
    _fields_ = [
        ("y_final", C.c_void_p), ("t_final", C.c_void_p),
        ("accepted", C.c_void_p), ("rejected", C.c_void_p), ("evals", C.c_void_p),
        ("status", C.c_void_p), ("dense", C.c_void_p), ("n_dense", C.c_void_p),
        ("event_t", C.c_void_p), ("event_y", C.c_void_p), ("n_events", C.c_void_p),
        ("points_offset", C.c_void_p), ("points_t", C.c_void_p), ("points_y", C.c_void_p),
        ("n_points", C.c_void_p),
    ]


class SdeBatch(C.Structure):
    _fields_ = [
        ("family", C.c_int32), ("scheme", C.c_int32),
        ("n", C.c_int64),
        ("dim", C.c_int32), ("noise", C.c_int32),
        ("params", C.c_void_p), ("y0", C.c_void_p),
        ("t0", C.c_double), ("tf", C.c_double), ("h0", C.c_double),
        ("h_min", C.c_double), ("h_max", C.c_double),
        ("max_steps", C.c_int64),
        ("seed", C.c_uint64), ("path_base", C.c_uint64),
        ("noise_table", C.c_void_p), ("noise_steps", C.c_int64),
        ("noise_constant", C.c_double),
        ("path_stride", C.c_int32), ("flags", C.c_uint32),
    ]


class SdeResult(C.Structure):
    _fields_ = [
        ("y_final", C.c_void_p), ("t_final", C.c_void_p),
        ("steps", C.c_void_p), ("evals", C.c_void_p), ("status", C.c_void_p),
        ("path", C.c_void_p), ("path_t", C.c_void_p),
    ]


def ode_defaults():
    """Reference defaults (src/methods/erk/mod.rs:115-158)."""
    b = OdeBatch()
    b.rtol = 1e-6
    b.atol = 1e-6
    b.h0 = 0.0
    b.h_min = 0.0
    b.h_max = float("inf")
    b.max_steps = 10000
    b.max_rejects = 100
    b.safety_factor = 0.9
    b.min_scale = 0.2
    b.max_scale = 10.0
    return b


def sde_defaults():
    b = SdeBatch()
    b.h_min = 0.0
    b.h_max = float("inf")
    b.max_steps = 10000
    return b


# Every symbol the headers declare, with its ctypes signature.
SIGNATURES = {
    "de_ode_batch_init": (None, [C.POINTER(OdeBatch)]),
    "de_sde_batch_init": (None, [C.POINTER(SdeBatch)]),
    "de_ode_batch_solve": (C.c_int, [C.POINTER(OdeBatch), C.POINTER(OdeResult), C.c_void_p]),
    "de_sde_batch_solve": (C.c_int, [C.POINTER(SdeBatch), C.POINTER(SdeResult), C.c_void_p]),
    "de_ode_workspace_bytes": (C.c_int64, [C.POINTER(OdeBatch)]),
    "de_ode_csv_offsets": (C.c_int, [C.POINTER(OdeBatch), C.POINTER(OdeResult), C.c_int32,
                                     C.c_void_p, C.c_void_p]),
    "de_ode_csv_write": (C.c_int, [C.POINTER(OdeBatch), C.POINTER(OdeResult), C.c_int32,
                                   C.c_void_p, C.c_void_p, C.c_void_p]),
    "de_last_error": (C.c_char_p, []),
    "de_version": (C.c_char_p, []),
    "de_set_workspace_budget": (C.c_int, [C.c_size_t]),
    "de_workspace_budget": (C.c_size_t, []),
    "de_workspace_trim": (C.c_int, []),
    "de_ode_batch_solve_noalloc": (C.c_int, [C.POINTER(OdeBatch), C.POINTER(OdeResult),
                                             C.c_void_p]),
    "de_ode_noalloc_scratch_bytes": (C.c_int64, [C.POINTER(OdeBatch)]),
    "de_server_create": (C.c_int, [C.c_size_t, C.c_int32, C.POINTER(C.c_void_p)]),
    "de_server_submit": (C.c_int, [C.c_void_p, C.POINTER(OdeBatch), C.POINTER(C.c_uint64)]),
    "de_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(OdeResult), C.c_int32]),
    "de_server_latency": (C.c_double, [C.c_void_p, C.c_uint64]),
    "de_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "de_server_pools": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "de_server_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "de_server_destroy": (None, [C.c_void_p]),
    "de_gen_lorenz": (None, [C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p]),
    "de_gen_vdp": (None, [C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p]),
    "de_gen_spring_sizes": (None, [C.c_uint64, C.c_int64, C.c_void_p]),
    "de_spring_offsets": (None, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "de_gen_spring": (None, [C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p]),
    "de_gen_ou": (None, [C.c_uint64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "de_gen_gbm": (None, [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_void_p,
                          C.c_void_p]),
    "de_philox4x32_10": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "de_noise_normal": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
    "de_diag_pow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "de_diag_arith": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "de_spring_tuned": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
    "de_diag_fmt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "de_diag_noise": (C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64]),
}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libde_amd.so is not built (%s); run make -C differential-equations_amd" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != DE_OK:
        msg = load().de_last_error()
        raise RuntimeError("de call failed (%d): %s" % (rc, msg.decode() if msg else ""))
    return rc
