"""The no-allocation entry point (de_ode_batch_solve_noalloc) against
de_ode_batch_solve on device memory: bit-identical results on consecutive
calls (the second smaller than the first, so the shared scratch array still
holds the first call's order and counter), an event solve, the plain
schedule, and the size limit."""
import ctypes as C

import numpy as np
import pytest

import deamd
from deamd import _abi

pytestmark = pytest.mark.gpu


def _device_solve(fn, n, seed, tf, schedule=_abi.DE_SCHED_PERSISTENT, event=False):
    """One solve through `fn` on device buffers allocated with the HIP runtime
    the library links (outputs poisoned: NaN doubles, -1 integers)."""
    from test_gpu_parity import HipBuffers
    mem = HipBuffers()
    try:
        p, y0 = deamd.gen_lorenz(n, seed=seed)
        max_ev = 8 if event else 0
        fields = [("y", 3 * n, np.float64), ("t", n, np.float64), ("acc", n, np.int64),
                  ("rej", n, np.int64), ("ev", n, np.int64), ("st", n, np.int32)]
        if event:
            fields += [("et", max_ev * n, np.float64), ("ey", max_ev * 3 * n, np.float64),
                       ("ne", n, np.int32)]
        out = {k: mem.poisoned(m * np.dtype(t).itemsize) for k, m, t in fields}
        b = _abi.ode_defaults()
        b.family, b.method, b.n, b.dim = _abi.DE_FAMILY_LORENZ, _abi.DE_METHOD_DOPRI5, n, 3
        b.params, b.y0 = mem.put(p.reshape(-1)), mem.put(y0.reshape(-1))
        b.t0_scalar, b.tf_scalar = 0.0, tf
        b.rtol, b.atol, b.max_steps = 1e-8, 1e-10, 1000000
        b.schedule = schedule
        r = _abi.OdeResult()
        r.y_final, r.t_final = out["y"], out["t"]
        r.accepted, r.rejected = out["acc"], out["rej"]
        r.evals, r.status = out["ev"], out["st"]
        if event:
            b.event_solout, b.event_fn, b.event_component = (_abi.DE_SOLOUT_EVENT,
                                                             _abi.DE_EVENT_COMPONENT, 2)
            b.event_value_traj = mem.put(p[1] - 1.0)
            b.max_events = max_ev
            r.event_t, r.event_y, r.n_events = out["et"], out["ey"], out["ne"]
        _abi.check(fn(C.byref(b), C.byref(r), None))
        assert mem.hip.hipDeviceSynchronize() == 0
        return {k: mem.get(out[k], (m,), t) for k, m, t in fields}
    finally:
        mem.free()


def _same(a, b):
    for k in a:
        x, y = a[k], b[k]
        if x.dtype.kind == "f":
            assert np.array_equal(np.isnan(x), np.isnan(y)), k
            assert np.array_equal(np.nan_to_num(x), np.nan_to_num(y)), k
        else:
            assert np.array_equal(x, y), k


def test_noalloc_matches_batch_solve_on_consecutive_calls():
    lib = deamd.lib()
    for n, seed, tf in ((200000, 3, 4.0), (5000, 4, 6.0)):   # the second call is smaller
        want = _device_solve(lib.de_ode_batch_solve, n, seed, tf)
        got = _device_solve(lib.de_ode_batch_solve_noalloc, n, seed, tf)
        assert (got["st"] == 0).all()
        _same(got, want)


def test_noalloc_event_and_plain_schedule():
    lib = deamd.lib()
    want = _device_solve(lib.de_ode_batch_solve, 3000, 5, 5.0, event=True)
    got = _device_solve(lib.de_ode_batch_solve_noalloc, 3000, 5, 5.0, event=True)
    assert want["ne"].sum() > 0
    _same(got, want)
    want = _device_solve(lib.de_ode_batch_solve, 4096, 6, 3.0, schedule=_abi.DE_SCHED_PLAIN)
    got = _device_solve(lib.de_ode_batch_solve_noalloc, 4096, 6, 3.0, schedule=_abi.DE_SCHED_PLAIN)
    _same(got, want)


def test_noalloc_rejects_requests_beyond_its_scratch():
    lib = deamd.lib()
    b = _abi.ode_defaults()
    b.family, b.method, b.dim, b.n = _abi.DE_FAMILY_LORENZ, _abi.DE_METHOD_DOPRI5, 3, 20 << 20
    assert lib.de_ode_noalloc_scratch_bytes(C.byref(b)) > 64 << 20
    b.schedule = _abi.DE_SCHED_PERSISTENT_FIFO
    assert lib.de_ode_noalloc_scratch_bytes(C.byref(b)) < 4096
    b.schedule = _abi.DE_SCHED_PERSISTENT
    dummy = np.zeros(4)
    b.params = b.y0 = C.c_void_p(dummy.ctypes.data)   # never read: the size check comes first
    r = _abi.OdeResult()
    for f in ("y_final", "t_final", "accepted", "rejected", "evals", "status"):
        setattr(r, f, C.c_void_p(dummy.ctypes.data))
    assert lib.de_ode_batch_solve_noalloc(C.byref(b), C.byref(r), None) == _abi.DE_ERR_INVALID
    assert b"scratch" in lib.de_last_error()
