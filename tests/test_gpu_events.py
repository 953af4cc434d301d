"""GPU parity of the event solouts: the HIP lanes against the CPU oracle,
bit for bit (event times and states, counts, final state, step counts and
statuses), for every method, solout kind and crossing direction.
"""
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_SCHED_PERSISTENT, DE_SCHED_PERSISTENT_FIFO, DE_SCHED_PLAIN,
                   DE_STATUS_COMPLETE, DE_STATUS_INTERRUPTED, DE_STATUS_OUT_OF_BOUNDS,
                   CrossingDirection, Event, EventConfig, ExplicitRungeKutta, Exponential,
                   Harmonic, IVP, Linear, Logistic, Lorenz, Robertson, VanDerPol)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
DIRS = [CrossingDirection.Both, CrossingDirection.Positive, CrossingDirection.Negative]


def same(a, b):
    if a is None or b is None:
        return a is None and b is None
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def assert_events_equal(got, want, dense=False):
    assert np.array_equal(got.status, want.status)
    assert np.array_equal(got.accepted, want.accepted)
    assert np.array_equal(got.rejected, want.rejected)
    assert np.array_equal(got.evals, want.evals)
    assert same(got.t, want.t)
    assert same(got.y, want.y)
    assert same(got.n_events, want.n_events)
    assert same(got.event_t, want.event_t)
    assert same(got.event_y, want.event_y)
    if dense:
        assert np.array_equal(got.n_dense, want.n_dense)
        assert same(got.dense, want.dense)


def method(name, h=0.01):
    return {"dopri5": ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10),
            "dop853": ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11),
            "rk4": ExplicitRungeKutta.rk4(h),
            "euler": ExplicitRungeKutta.euler(h)}[name].max_steps(1000000)


def lorenz(n, tf, seed):
    p, y0 = deamd.gen_lorenz(n, seed=seed)
    return p, IVP.ode(Lorenz(*p), 0.0, tf, y0)


@pytest.mark.parametrize("name", ["dopri5", "dop853", "rk4", "euler"])
@pytest.mark.parametrize("d", DIRS)
def test_lorenz_poincare_section_event(orc, name, d):
    # z = rho - 1 per trajectory (the plane through the two non-trivial fixed
    # points), EventWrappedSolout around the final-state output.
    p, ens = lorenz(2048, 6.0, seed=41)
    ens = ens.method(method(name)).event(
        Event.component(2, p[1] - 1.0, EventConfig(d))).options(max_events=24)
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert (want.status == DE_STATUS_COMPLETE).mean() > 0.9
    assert want.n_events.max() > 0
    assert_events_equal(got, want)


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT,
                                      DE_SCHED_PERSISTENT_FIFO])
def test_terminal_plane_event_with_t_eval_all_schedules(orc, schedule):
    # Plane event with terminate_after(3) around a t_eval output: dense rows,
    # event points and the Interrupted statuses; 8192 trajectories engage the
    # longest-first order.
    p, ens = lorenz(8192, 8.0, seed=43)
    ens = ens.t_eval(np.linspace(0.0, 8.0, 17)).method(method("dopri5")).event(
        Event.plane([0.0, 0.0, 20.0], [0.2, -0.1, 1.0], EventConfig().terminate_after(3))) \
        .options(schedule=schedule, max_events=4)
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert (want.status == DE_STATUS_INTERRUPTED).sum() > 100
    assert_events_equal(got, want, dense=True)


@pytest.mark.parametrize("name", ["dopri5", "dop853", "rk4", "euler"])
@pytest.mark.parametrize("kind", ["crossing", "hyperplane"])
def test_crossing_and_hyperplane_solouts(orc, name, kind):
    p, ens = lorenz(2048, 5.0, seed=47)
    ens = ens.method(method(name)).options(max_events=16)
    if kind == "crossing":
        ens = ens.crossing(0, 0.0, CrossingDirection.Both)
    else:
        ens = ens.hyperplane_crossing([0.0, 0.0, 25.0], [1.0, 1.0, 3.0],
                                      CrossingDirection.Positive)
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert want.n_events.sum() > 0
    assert_events_equal(got, want)


def test_small_families_time_and_component_events(orc):
    rng = np.random.default_rng(53)
    n = 1024
    cases = [
        IVP.ode(Exponential(rng.uniform(0.1, 1.0, n)), 0.0, 10.0, rng.uniform(0.5, 2.0, n).reshape(1, -1))
        .event(Event.time(rng.uniform(1.0, 9.0, n), EventConfig().terminal())),
        IVP.ode(Logistic(rng.uniform(0.5, 2.0, n), np.full(n, 10.0)), 0.0, 20.0,
                np.full((1, n), 0.1)).event(Event.component(0, 9.0, EventConfig().terminal())),
        IVP.ode(Harmonic(rng.uniform(0.5, 4.0, n)), 0.0, 15.0,
                np.vstack([np.ones(n), rng.uniform(-1, 1, n)])).event(Event.component(1, 0.0)),
        IVP.ode(VanDerPol(rng.uniform(0.1, 5.0, n)), 0.0, 15.0,
                np.vstack([np.full(n, 2.0), np.zeros(n)])).crossing(0, 0.5, CrossingDirection.Negative),
        IVP.ode(Linear(np.full(n, -1.0), np.zeros(n)), 0.0, 20.0, np.full((1, n), 10.0))
        .event(Event.component(0, 0.0)),
        IVP.ode(Robertson(), 0.0, 5.0, np.vstack([np.ones(n), np.zeros(n), np.zeros(n)]))
        .event(Event.component(2, rng.uniform(0.01, 0.5, n))),
    ]
    for ens in cases:
        for name in ("dopri5", "dop853", "rk4", "euler"):
            e = ens.method(method(name, h=0.05)).options(max_events=12)
            want = orc.solve_ode(e, threads=THREADS)
            got = e.solve()
            assert_events_equal(got, want)


def test_backward_solves_out_of_bounds(orc):
    # Backward t_eval and crossing solves end with DE_STATUS_OUT_OF_BOUNDS
    # (the reference's unwrap of a failed interpolation); backward events find
    # no root.  Also the plain (no event) dense lanes.
    rng = np.random.default_rng(59)
    n = 512
    k = rng.uniform(0.5, 3.0, n)
    y0 = np.vstack([np.ones(n), rng.uniform(-1, 1, n)])

    def base():  # the builders configure the ensemble in place: one per case
        return IVP.ode(Harmonic(k), 12.0, 0.0, y0)

    cases = [lambda: base().t_eval([12.0, 9.0, 3.0]).method(method("dopri5")),
             lambda: base().t_eval([12.0, 9.0, 3.0]).method(method("dop853")),
             lambda: base().t_eval([12.0, 9.0, 3.0]).method(method("rk4", -0.05)),
             lambda: base().method(method("dopri5")).crossing(0, 0.0),
             lambda: base().method(method("euler", -0.05)).event(Event.component(0, 0.0)),
             lambda: base().t_eval([12.0, 9.0]).method(method("dopri5"))
             .event(Event.component(0, 0.0))]
    for mk in cases:
        ens = mk()
        want = orc.solve_ode(ens, threads=THREADS)
        got = ens.solve()
        assert_events_equal(got, want, dense=ens._t_eval is not None)
    assert (orc.solve_ode(cases[0]()).status == DE_STATUS_OUT_OF_BOUNDS).all()


def test_events_with_per_component_tolerances_and_fast_flag(orc):
    # Event lanes carry per-component tolerances; DE_FLAG_FAST_MATH does not
    # change an event solve (it runs the parity kernels).
    p, ens = lorenz(1024, 4.0, seed=61)
    ens = ens.method(ExplicitRungeKutta.dopri5().rtol([1e-8, 1e-7, 1e-9])
                     .atol([1e-10, 1e-9, 1e-11]).max_steps(100000)) \
        .event(Event.component(2, p[1] - 1.0)).options(max_events=16, fast_math=True)
    assert_events_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))


def test_server_refuses_out_of_range_event_components():
    # Served event solouts take every method; a component outside the state
    # is refused at submission.
    p, ens = lorenz(16, 1.0, seed=67)
    with deamd.Server(8 << 30) as srv:
        with pytest.raises(RuntimeError):
            srv.submit(ens.method(method("dop853")).event(Event.component(3, 20.0)))
        p, bad = lorenz(16, 1.0, seed=67)
        with pytest.raises(RuntimeError):
            srv.submit(bad.method(method("dopri5")).crossing(3, 0.0))


@pytest.mark.parametrize("name", ["dopri5", "dop853", "rk4"])
def test_deferred_regions_overflow_to_inline_solves(orc, name):
    # Without a terminate count the roots are searched after the integration,
    # from per-trajectory regions of max_events + max(max_events / 2, 32)
    # brackets.  Oscillators of widely spread frequencies give 2 to ~250 zero
    # crossings, so some trajectories outgrow their region and are solved
    # again with inline roots; the results must not show which.
    rng = np.random.default_rng(71)
    n = 4096
    k = np.exp(rng.uniform(np.log(0.5), np.log(400.0), n))
    y0 = np.vstack([np.ones(n), rng.uniform(-1, 1, n)])
    for cap in (0, 8):
        ens = IVP.ode(Harmonic(k), 0.0, 40.0, y0).method(method(name, h=0.002)) \
            .event(Event.component(0, 0.0)).options(max_events=cap)
        want = orc.solve_ode(ens, threads=THREADS)
        assert (want.n_events > 40 + cap).any() and (want.n_events < 32).any()
        assert_events_equal(ens.solve(), want)
        ens = IVP.ode(Harmonic(k), 0.0, 40.0, y0).method(method(name, h=0.002)) \
            .crossing(1, 0.0).options(max_events=cap)
        assert_events_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))


@pytest.mark.parametrize("name,n,slice_", [("dopri5", 2048, 700), ("dop853", 9000, 4500),
                                           ("rkv878e", 2048, 1000), ("rkf45", 9000, 4500)])
def test_deferred_roots_in_slices(orc, monkeypatch, name, n, slice_):
    # Ensembles whose bracket regions exceed the deferral budget are solved in
    # slices of consecutive trajectories, each with its own region, root and
    # bookkeeping passes and inline reruns; DE_EVENT_DEFER_SLICE forces small
    # slices (in index order below 4096 trajectories, longest-first above).
    monkeypatch.setenv("DE_EVENT_DEFER_SLICE", str(slice_))
    m = method(name) if name in ("dopri5", "dop853") else \
        getattr(ExplicitRungeKutta, name)().rtol(1e-8).atol(1e-10).max_steps(1000000)
    # regions of 6 + 32 brackets: about a third of the trajectories outgrow them
    p, ens = lorenz(n, 12.0, seed=61)
    ens = ens.method(m).event(Event.component(2, p[1] - 1.0)).options(max_events=6)
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.n_events > 38).any() and (want.n_events < 38).any()
    assert_events_equal(ens.solve(), want)
    ens = ens.t_eval(np.linspace(0.0, 12.0, 9)).options(schedule=DE_SCHED_PERSISTENT_FIFO)
    assert_events_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS), dense=True)


@pytest.mark.parametrize("budget_mib,name", [(2, "dopri5"), (64, "dop853")])
def test_workspace_budget_bounds_the_event_regions(orc, budget_mib, name):
    # A small workspace budget leaves no room for the bracket regions (2 MiB:
    # roots inline; 64 MiB: fewer entries per trajectory, the overflowing ones
    # rerun inline); the bits do not change.
    lib = deamd.lib()
    old = lib.de_workspace_budget()
    p, ens = lorenz(6000, 8.0, seed=71)
    ens = ens.method(method(name)).event(Event.component(2, p[1] - 1.0)).options(max_events=40)
    want = orc.solve_ode(ens, threads=THREADS)
    try:
        deamd._abi.check(lib.de_set_workspace_budget(budget_mib << 20))
        got = ens.solve()
    finally:
        deamd._abi.check(lib.de_set_workspace_budget(old))
        deamd._abi.check(lib.de_workspace_trim())
    assert_events_equal(got, want)
