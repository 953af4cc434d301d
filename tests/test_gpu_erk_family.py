"""GPU parity of the embedded explicit Runge-Kutta family (rkf45, cash_karp,
rkv655e .. rkv989e): the HIP lanes against the CPU oracle, equal as IEEE
values (final states and times, dense output, step and evaluation counts,
statuses, event points) for every tableau, schedule, tolerance form, family
and event solout.
"""
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_SCHED_PERSISTENT, DE_SCHED_PERSISTENT_FIFO, DE_SCHED_PLAIN,
                   DE_STATUS_BAD_INPUT, DE_STATUS_COMPLETE, DE_STATUS_OUT_OF_BOUNDS,
                   DE_STATUS_INTERRUPTED, DE_STATUS_STIFFNESS, CrossingDirection, Event,
                   EventConfig, ExplicitRungeKutta, Exponential, Harmonic, IVP,
                   Linear, Logistic, Lorenz, Robertson, VanDerPol)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
FAMILY = ["rkf45", "cash_karp", "rkv655e", "rkv656e", "rkv766e", "rkv767e", "rkv877e",
          "rkv878e", "rkv988e", "rkv989e"]


def same(a, b):
    if a is None or b is None:
        return a is None and b is None
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def assert_equal(got, want, dense=False):
    assert np.array_equal(got.status, want.status)
    assert np.array_equal(got.accepted, want.accepted)
    assert np.array_equal(got.rejected, want.rejected)
    assert np.array_equal(got.evals, want.evals)
    assert same(got.t, want.t)
    assert same(got.y, want.y)
    if dense:
        assert np.array_equal(got.n_dense, want.n_dense)
        assert same(got.dense, want.dense)


def method(name, rtol=1e-8, atol=1e-10):
    return getattr(ExplicitRungeKutta, name)().rtol(rtol).atol(atol).max_steps(200000)


@pytest.mark.parametrize("name", FAMILY)
def test_lorenz_final_state(orc, name):
    p, y0 = deamd.gen_lorenz(4096, seed=81)
    ens = IVP.ode(Lorenz(*p), 0.0, 4.0, y0).method(method(name))
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).all() and want.rejected.sum() > 0
    assert_equal(ens.solve(), want)


@pytest.mark.parametrize("name", FAMILY)
def test_lorenz_dense_output(orc, name):
    # Horner on the interpolation rows for the Verner pairs, cubic Hermite for
    # rkf45 and cash_karp; points at the ends, between steps and repeated.
    p, y0 = deamd.gen_lorenz(1024, seed=83)
    pts = [0.0, 0.013, 0.5, 0.5, 1.0, 1.7, 2.2, 3.0]
    ens = IVP.ode(Lorenz(*p), 0.0, 3.0, y0).t_eval(pts).method(method(name))
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.n_dense == len(pts)).all()
    assert_equal(ens.solve(), want, dense=True)


@pytest.mark.parametrize("name", ["rkf45", "rkv656e", "rkv878e", "rkv989e"])
def test_per_component_and_per_trajectory_tolerances(orc, name):
    p, y0 = deamd.gen_lorenz(1024, seed=87)
    rng = np.random.default_rng(87)
    ens = IVP.ode(Lorenz(*p), 0.0, 2.0, y0).method(
        getattr(ExplicitRungeKutta, name)().rtol([1e-8, 1e-6, 1e-9]).atol([1e-10, 1e-8, 1e-11])
        .max_steps(200000))
    assert_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))
    ens = IVP.ode(Lorenz(*p), 0.0, 2.0, y0).method(method(name))
    ens.rtol_traj = 10.0 ** rng.uniform(-10, -5, 1024)
    ens.atol_traj = 10.0 ** rng.uniform(-12, -7, 1024)
    assert_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT,
                                      DE_SCHED_PERSISTENT_FIFO])
def test_schedules(orc, schedule):
    p, y0 = deamd.gen_lorenz(8192, seed=89)
    for name in ("cash_karp", "rkv767e"):
        ens = IVP.ode(Lorenz(*p), 0.0, 2.0, y0).t_eval(np.linspace(0.0, 2.0, 9)) \
            .method(method(name, 1e-7, 1e-9)).options(schedule=schedule)
        assert_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS), dense=True)


def test_small_families(orc):
    rng = np.random.default_rng(91)
    n = 512
    cases = [
        IVP.ode(Exponential(rng.uniform(-2.0, 1.0, n)), 0.0, 10.0, rng.uniform(0.5, 2.0, n).reshape(1, -1)),
        IVP.ode(Linear(rng.uniform(-1.0, 1.0, n), rng.uniform(-1.0, 1.0, n)), 0.0, 10.0,
                rng.uniform(-1.0, 1.0, n).reshape(1, -1)),
        IVP.ode(Harmonic(rng.uniform(0.5, 4.0, n)), 0.0, 15.0,
                np.vstack([np.ones(n), rng.uniform(-1, 1, n)])),
        IVP.ode(Logistic(rng.uniform(0.5, 2.0, n), np.full(n, 10.0)), 0.0, 10.0, np.full((1, n), 0.1)),
        IVP.ode(VanDerPol(rng.uniform(0.1, 5.0, n)), 0.0, 15.0,
                np.vstack([np.full(n, 2.0), np.zeros(n)])),
        IVP.ode(Robertson(), 0.0, 0.4, np.vstack([np.ones(n), np.zeros(n), np.zeros(n)])),
    ]
    for ens in cases:
        for name in FAMILY:
            e = ens.method(method(name, 1e-7, 1e-9)).t_eval([0.0, 0.25, 0.4])
            want = orc.solve_ode(e, threads=THREADS)
            assert_equal(e.solve(), want, dense=True)


def test_reference_accuracy_configuration(orc):
    # tests/ode/accuracy.rs defaults (rtol = atol = 1e-6, h0 computed) on the
    # scalar problems, forwards and backwards.
    for name in FAMILY:
        for t0, tf, y0 in [(0.0, 10.0, 1.0), (0.0, -10.0, 22026.46579479)]:
            ens = IVP.ode(Exponential(1.0), t0, tf, [y0], n=1).method(
                getattr(ExplicitRungeKutta, name)())
            assert_equal(ens.solve(), orc.solve_ode(ens))


@pytest.mark.parametrize("name,h0", [("rkf45", 0.5), ("rkv655e", 0.5), ("rkv988e", 0.2),
                                     ("rkv988e", 0.5)])
def test_max_rejects_and_nan_steps(orc, name, h0):
    # max_rejects consecutive rejections end in Stiffness; from h0 = 0.5
    # rkv988e's stages overflow and the NaN error norm accepts the step
    # (max() ignores NaN), as in the reference.
    n = 256
    mus = np.linspace(5.0, 40.0, n)
    for max_rejects in (0, 1, 2, 3, 100):
        ens = IVP.ode(VanDerPol(mus), 0.0, 8.0, np.vstack([np.full(n, 2.0), np.zeros(n)])) \
            .method(getattr(ExplicitRungeKutta, name)().h0(h0).rtol(1e-10).atol(1e-10)
                    .max_rejects(max_rejects).max_steps(50000))
        want = orc.solve_ode(ens, threads=THREADS)
        if max_rejects == 0 and not (name == "rkv988e" and h0 == 0.5):
            assert (want.status == DE_STATUS_STIFFNESS).any()
        assert_equal(ens.solve(), want)
    if name == "rkv988e" and h0 == 0.5:
        assert np.isnan(want.y).any()


def test_backward_and_bad_input(orc):
    rng = np.random.default_rng(97)
    n = 256
    k = rng.uniform(0.5, 3.0, n)
    y0 = np.vstack([np.ones(n), rng.uniform(-1, 1, n)])
    for name in FAMILY:
        # backward solve; a backward t_eval ends in OutOfBounds (the
        # reference unwraps the failed interpolation of interior points)
        ens = IVP.ode(Harmonic(k), 12.0, 0.0, y0).method(method(name))
        assert_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))
        ens = IVP.ode(Harmonic(k), 12.0, 0.0, y0).t_eval([12.0, 9.0, 3.0]).method(method(name))
        want = orc.solve_ode(ens, threads=THREADS)
        assert (want.status == DE_STATUS_OUT_OF_BOUNDS).all()
        assert_equal(ens.solve(), want, dense=True)
        # |h0| > |tf - t0| and tf == t0 (tests/ode/errors.rs)
        ens = IVP.ode(Harmonic(k), 0.0, 1.0, y0).method(method(name).h0(10.0))
        got = ens.solve()
        assert (got.status == DE_STATUS_BAD_INPUT).all() and (got.evals == 0).all()
        assert_equal(got, orc.solve_ode(ens, threads=THREADS))


def assert_events_equal(got, want, dense=False):
    assert_equal(got, want, dense)
    assert same(got.n_events, want.n_events)
    assert same(got.event_t, want.event_t)
    assert same(got.event_y, want.event_y)


@pytest.mark.parametrize("name", FAMILY)
def test_lorenz_section_events(orc, name):
    # z = rho - 1 per trajectory, both directions, EventWrappedSolout around
    # the final-state output; roots from the dense output of each tableau
    # (Horner over the bi rows, or the cubic Hermite of rkf45 / cash_karp).
    p, y0 = deamd.gen_lorenz(2048, seed=41)
    ens = IVP.ode(Lorenz(*p), 0.0, 6.0, y0).method(method(name)) \
        .event(Event.component(2, p[1] - 1.0)).options(max_events=24)
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).mean() > 0.9 and want.n_events.max() > 0
    assert_events_equal(ens.solve(), want)


@pytest.mark.parametrize("name", ["rkf45", "rkv656e", "rkv878e", "rkv989e"])
@pytest.mark.parametrize("kind", ["crossing", "hyperplane"])
def test_crossing_and_hyperplane_solouts(orc, name, kind):
    p, y0 = deamd.gen_lorenz(2048, seed=47)
    ens = IVP.ode(Lorenz(*p), 0.0, 5.0, y0).method(method(name)).options(max_events=16)
    if kind == "crossing":
        ens = ens.crossing(0, 0.0, CrossingDirection.Both)
    else:
        ens = ens.hyperplane_crossing([0.0, 0.0, 25.0], [1.0, 1.0, 3.0],
                                      CrossingDirection.Positive)
    want = orc.solve_ode(ens, threads=THREADS)
    assert want.n_events.sum() > 0
    assert_events_equal(ens.solve(), want)


@pytest.mark.parametrize("name", ["cash_karp", "rkv655e", "rkv767e", "rkv988e"])
def test_terminal_plane_event_with_t_eval(orc, name):
    # terminate_after(3) around a t_eval output: dense rows, event points and
    # the Interrupted statuses (the last dense row is the event's time).
    p, y0 = deamd.gen_lorenz(4096, seed=43)
    ens = IVP.ode(Lorenz(*p), 0.0, 8.0, y0).t_eval(np.linspace(0.0, 8.0, 17)) \
        .method(method(name)).event(Event.plane([0.0, 0.0, 20.0], [0.2, -0.1, 1.0],
                                                EventConfig().terminate_after(3))) \
        .options(schedule=DE_SCHED_PERSISTENT, max_events=4)
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_INTERRUPTED).sum() > 100
    assert_events_equal(ens.solve(), want, dense=True)


@pytest.mark.parametrize("name", ["rkf45", "rkv655e", "rkv877e"])
def test_deferred_regions_overflow_to_inline_solves(orc, name):
    # Oscillators of widely spread frequencies: 2 to ~250 zero crossings, so
    # some trajectories outgrow their deferred-root region and are solved
    # again with inline roots; the results must not show which.
    rng = np.random.default_rng(71)
    n = 4096
    k = np.exp(rng.uniform(np.log(0.5), np.log(400.0), n))
    y0 = np.vstack([np.ones(n), rng.uniform(-1, 1, n)])
    for cap in (0, 8):
        ens = IVP.ode(Harmonic(k), 0.0, 40.0, y0).method(method(name, 1e-7, 1e-9)) \
            .event(Event.component(0, 0.0)).options(max_events=cap)
        want = orc.solve_ode(ens, threads=THREADS)
        assert (want.n_events > 40 + cap).any() and (want.n_events < 32).any()
        assert_events_equal(ens.solve(), want)


def test_small_family_time_events_and_backward_solves(orc):
    rng = np.random.default_rng(53)
    n = 1024
    cases = [
        IVP.ode(Exponential(rng.uniform(0.1, 1.0, n)), 0.0, 10.0,
                rng.uniform(0.5, 2.0, n).reshape(1, -1))
        .event(Event.time(rng.uniform(1.0, 9.0, n), EventConfig().terminal())),
        IVP.ode(Logistic(rng.uniform(0.5, 2.0, n), np.full(n, 10.0)), 0.0, 20.0,
                np.full((1, n), 0.1)).event(Event.component(0, 9.0, EventConfig().terminal())),
        IVP.ode(Harmonic(rng.uniform(0.5, 3.0, n)), 0.0, -12.0,
                np.vstack([np.ones(n), np.zeros(n)])).crossing(0, 0.0, CrossingDirection.Negative)
        .options(max_events=6),
    ]
    for name in ("cash_karp", "rkv766e", "rkv989e"):
        for ens in cases:
            ens = ens.method(method(name))
            assert_events_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))
