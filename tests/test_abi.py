"""C-ABI library: loads, exports every declared symbol, validates descriptors
and generates the synthetic workloads deterministically (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import deamd
from deamd import _abi

HEADERS = [os.path.join(_abi.REPO, "include", h) for h in ("de_api.h", "de_workload.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(de_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = deamd.lib()
    names = declared_functions()
    assert "de_ode_batch_solve" in names and "de_sde_batch_solve" in names
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.SIGNATURES, name


def test_version_and_defaults():
    lib = deamd.lib()
    assert b"gfx950" in lib.de_version()
    b = _abi.OdeBatch()
    lib.de_ode_batch_init(C.byref(b))
    assert (b.rtol, b.atol, b.h0, b.h_min, b.max_steps) == (1e-6, 1e-6, 0.0, 0.0, 10000)
    assert (b.safety_factor, b.min_scale, b.max_scale) == (0.9, 0.2, 10.0)
    assert b.h_max == float("inf")


def test_workspace_budget_setters():
    lib = deamd.lib()
    assert lib.de_workspace_budget() == 64 << 30
    assert lib.de_set_workspace_budget(1000) == _abi.DE_ERR_INVALID
    assert lib.de_workspace_budget() == 64 << 30
    assert lib.de_set_workspace_budget(8 << 30) == _abi.DE_OK
    assert lib.de_workspace_budget() == 8 << 30
    assert lib.de_set_workspace_budget(64 << 30) == _abi.DE_OK
    assert lib.de_workspace_trim() == _abi.DE_OK


def _solve(b):
    res = deamd.Solution(max(int(b.n), 1), max(int(b.n), 1) * max(b.dim, 1), 0)
    return deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), None)


@pytest.mark.parametrize("mutate,code", [
    (lambda b: setattr(b, "family", 99), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "dim", 4), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "method", 14), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "n", -1), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "n_eval", 3), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "schedule", 9), _abi.DE_ERR_INVALID),
    (lambda b: setattr(b, "y0", None), _abi.DE_ERR_INVALID),
])
def test_invalid_descriptors_are_rejected_before_any_device_work(mutate, code):
    ens = deamd.IVP.ode(deamd.Lorenz(10.0, 28.0, 8.0 / 3.0), 0.0, 1.0, [1.0, 1.0, 1.0], n=4)
    b, res, keep = ens.batch()
    mutate(b)
    rc = deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), None)
    assert rc == code
    assert deamd.lib().de_last_error()


def _event_case(kind):
    e = deamd.IVP.ode(deamd.Lorenz(10.0, 28.0, 8.0 / 3.0), 0.0, 1.0, [1.0, 1.0, 1.0], n=4)
    return e.event(deamd.Event.component(2, 27.0)) if kind == "event" else e.crossing(0, 0.0)


@pytest.mark.parametrize("kind,mutate,code", [
    ("event", lambda b: setattr(b, "event_solout", 7), _abi.DE_ERR_INVALID),
    ("event", lambda b: setattr(b, "event_fn", 5), _abi.DE_ERR_INVALID),
    ("event", lambda b: setattr(b, "event_component", 3), _abi.DE_ERR_INVALID),
    ("event", lambda b: setattr(b, "event_direction", 3), _abi.DE_ERR_INVALID),
    ("event", lambda b: setattr(b, "event_terminate", -1), _abi.DE_ERR_INVALID),
    ("event", lambda b: setattr(b, "max_events", -2), _abi.DE_ERR_INVALID),
    ("event", lambda b: (setattr(b, "event_fn", _abi.DE_EVENT_PLANE),
                         setattr(b, "event_normal", None)), _abi.DE_ERR_INVALID),
    ("crossing", lambda b: setattr(b, "event_component", -1), _abi.DE_ERR_INVALID),
    # a crossing solout replaces the t_eval output (src/ivp.rs:682-692)
    ("crossing", lambda b: (setattr(b, "n_eval", 1), setattr(b, "t_eval", b.y0)),
     _abi.DE_ERR_INVALID),
])
def test_invalid_event_descriptors_are_rejected(kind, mutate, code):
    b, res, keep = _event_case(kind).batch()
    mutate(b)
    assert deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), None) == code
    assert deamd.lib().de_last_error()


def test_plane_events_are_refused_for_spring_chains():
    # Spring chains take component and time events and the crossing solout;
    # a plane needs one state length for the whole ensemble.
    sysm, y0 = deamd.gen_spring(8, seed=1)
    mk = lambda: deamd.IVP.ode(sysm, 0.0, 1.0, y0)
    for ens in (mk().event(deamd.Event.plane(np.zeros(8), np.ones(8))),
                mk().hyperplane_crossing(np.zeros(8), np.ones(8))):
        b, res, keep = ens.batch()
        assert deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), None) == \
            _abi.DE_ERR_UNSUPPORTED


def test_empty_batch_is_a_no_op():
    ens = deamd.IVP.ode(deamd.Lorenz(10.0, 28.0, 8.0 / 3.0), 0.0, 1.0, np.zeros((3, 0)))
    b, res, keep = ens.batch()
    assert deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), None) == 0


def test_lorenz_workload_generator():
    p, y0 = deamd.gen_lorenz(4096, seed=0x5EED2026)
    assert (p[0] == 10.0).all() and (p[2] == 8.0 / 3.0).all()
    assert p[1].min() >= 0.0 and p[1].max() < 50.0
    assert abs(p[1].mean() - 25.0) < 1.0
    assert abs(y0.mean() - 1.0) < 0.01 and abs(y0.std() - 0.1) < 0.005
    p2, y02 = deamd.gen_lorenz(4096, seed=0x5EED2026)
    assert np.array_equal(p, p2) and np.array_equal(y0, y02)
    p3, _ = deamd.gen_lorenz(4096, seed=1)
    assert not np.array_equal(p, p3)
    # prefix property: trajectory i does not depend on the ensemble size
    p4, y04 = deamd.gen_lorenz(100, seed=0x5EED2026)
    assert np.array_equal(p4, p[:, :100]) and np.array_equal(y04, y0[:, :100])


def test_spring_workload_generator():
    sysm, y0 = deamd.gen_spring(4096)
    s = sysm.sizes
    assert s.min() >= 4 and s.max() <= 256
    # log-uniform: about a third below 16, about two thirds below 64
    assert 0.28 < (s < 16).mean() < 0.39 and 0.6 < (s < 64).mean() < 0.72
    assert len(y0) == 2 * s.sum()
    assert len(sysm.params_packed) == (s + 2).sum()
    c = sysm.params_packed[sysm.param_offset[:-1]]
    assert c.min() >= 0.0 and c.max() < 0.1
    k = np.concatenate([sysm.params_packed[sysm.param_offset[i] + 1:sysm.param_offset[i + 1]]
                        for i in range(200)])
    assert abs(np.log(k).std() - 0.5) < 0.03
    v = np.concatenate([y0[sysm.state_offset[i] + s[i]:sysm.state_offset[i + 1]]
                        for i in range(200)])
    assert (v == 0.0).all()


def test_ou_and_gbm_generators():
    sysm, y0 = deamd.gen_ou(2048, 8)
    p = sysm.params.reshape(3 * 8, 2048)
    theta, mu, sigma = p[0::3], p[1::3], p[2::3]
    assert theta.min() >= 0.5 and theta.max() < 2.0
    assert sigma.min() >= 0.1 and sigma.max() < 1.0
    assert abs(mu.mean()) < 0.05 and abs(mu.std() - 1.0) < 0.05
    assert (y0 == 0).all()
    g, gy = deamd.gen_gbm(16)
    assert (g.params.reshape(2, 16)[0] == 0.05).all() and (gy == 100.0).all()


def test_noise_normal_matches_oracle(orc):
    lib = deamd.lib()
    olib = orc.load()
    for p in range(20):
        for s in range(20):
            for c in range(8):
                assert lib.de_noise_normal(9, p, s, c) == olib.de_oracle_noise_normal(9, p, s, c)


def test_csv_entry_points_validate_before_any_device_work():
    # The device CSV writer validates its descriptor first: an unknown source
    # or a spring-chain batch is refused without device work.
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    ens = deamd.IVP.ode(deamd.Exponential(1.0), 0.0, 1.0, [1.0], n=1)
    b, res, keep = ens.batch()
    r = res.struct()
    off = np.zeros(2, dtype=np.int64)
    lib = deamd.lib()
    assert lib.de_ode_csv_offsets(C.byref(b), C.byref(r), 7, off.ctypes.data, None) == \
        _abi.DE_ERR_INVALID
    b.family = _abi.DE_FAMILY_SPRING_CHAIN
    assert lib.de_ode_csv_offsets(C.byref(b), C.byref(r), _abi.DE_CSV_FINAL, off.ctypes.data,
                                  None) == _abi.DE_ERR_UNSUPPORTED
