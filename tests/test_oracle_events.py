"""The oracle's event solouts (IVP::event, ::crossing, ::hyperplane_crossing)
pinned against closed forms, the reference's own event test setup
(tests/ode/errors.rs: g = t - 10, EventConfig::default().terminal()), cases
whose outcome follows from reading the reference's code, and an independent
Python transcription (pyref_events) bit for bit.
"""
import math

import numpy as np
import pytest

import pyref_dopri5
import pyref_events as PE
from deamd import (DE_STATUS_BAD_INPUT, DE_STATUS_COMPLETE, DE_STATUS_INTERRUPTED,
                   DE_STATUS_OUT_OF_BOUNDS, CrossingDirection, Event, EventConfig,
                   ExplicitRungeKutta, Exponential, Harmonic, IVP, Linear, Lorenz)

DP5 = dict(S=7, I=7, O=5, C=pyref_dopri5.C, A=pyref_dopri5.A, B=pyref_dopri5.B,
           ER=pyref_dopri5.ER, BH=None, BI=[[0.0] * 7 for _ in range(5)])
DIRS = {CrossingDirection.Both: "both", CrossingDirection.Positive: "positive",
        CrossingDirection.Negative: "negative"}


def cosine(tf=20.0, rtol=1e-10, atol=1e-12):
    """y'' = -y, y(0) = (1, 0): y = cos t, zero at pi/2 + j pi."""
    return IVP.ode(Harmonic(1.0), 0.0, tf, np.array([[1.0], [0.0]])).method(
        ExplicitRungeKutta.dopri5().rtol(rtol).atol(atol).max_steps(100000))


ZEROS = math.pi / 2 + math.pi * np.arange(7)


@pytest.mark.parametrize("d", list(DIRS))
def test_event_crossing_hyperplane_find_the_zeros_of_cosine(orc, d):
    # Both: all six zeros in [0, 20]; Negative (g from > 0 to <= 0): the even
    # ones (cos falls through pi/2, 5pi/2, ...); Positive: the odd ones.
    want = ZEROS[:6] if d == CrossingDirection.Both else (
        ZEROS[0:6:2] if d == CrossingDirection.Negative else ZEROS[1:6:2])
    cases = [cosine().event(Event.component(0, 0.0, EventConfig(d))),
             cosine().crossing(0, 0.0, d),
             cosine().hyperplane_crossing([0.0, 5.0], [2.0, 0.0], d)]
    for ens in cases:
        sol = orc.solve_ode(ens)
        assert sol.status[0] == DE_STATUS_COMPLETE
        assert sol.n_events[0] == len(want)
        t, y = sol.events(0)
        assert np.abs(t - want).max() < 1e-8
        assert np.abs(y[:, 0]).max() < 1e-8                       # cos = 0 there
        assert np.abs(np.abs(y[:, 1]) - 1.0).max() < 1e-8         # |sin| = 1


@pytest.mark.parametrize("method", ["dopri5", "dop853", "rk4", "euler"])
def test_time_event_terminal_like_the_reference_test(orc, method):
    # tests/ode/errors.rs: SimpleODE y' = y with g = t - 10 and
    # EventConfig::default().terminal().  g is linear in t, so the first secant
    # step of brent_dekker lands on 10 up to rounding; the solve then stops
    # with Status::Interrupted at the end of the step that crossed t = 10.
    m = {"dopri5": ExplicitRungeKutta.dopri5().rtol(1e-10).atol(1e-10),
         "dop853": ExplicitRungeKutta.dop853().rtol(1e-10).atol(1e-10),
         "rk4": ExplicitRungeKutta.rk4(0.3), "euler": ExplicitRungeKutta.euler(0.3)}[method]
    ens = IVP.ode(Exponential(1.0), 0.0, 20.0, [1.0], n=1).method(m.max_steps(100000)) \
        .event(Event.time(10.0, EventConfig().terminal()))
    sol = orc.solve_ode(ens)
    assert sol.status[0] == DE_STATUS_INTERRUPTED
    assert sol.n_events[0] == 1
    t, y = sol.events(0)
    assert abs(t[0] - 10.0) <= 1e-11
    assert sol.t[0] >= 10.0
    if method == "dopri5":
        assert abs(y[0, 0] / math.exp(10.0) - 1.0) < 1e-7
    if method == "dop853":
        # the step end is within 1e-10 of e^t; the reference's DOP853 fold
        # (s and s1 in the opposite phase from Hairer's contd8) leaves the
        # interpolated point ~6e-7 off here
        assert abs(sol.y[0] / math.exp(sol.t[0]) - 1.0) < 1e-9
        assert abs(y[0, 0] / math.exp(10.0) - 1.0) < 1e-5
    # the same solve without the event runs on past t = 10
    free = orc.solve_ode(IVP.ode(Exponential(1.0), 0.0, 20.0, [1.0], n=1).method(m))
    assert free.status[0] == DE_STATUS_COMPLETE and free.accepted[0] > sol.accepted[0]


def test_bad_input_with_event_like_the_reference_test(orc):
    # tests/ode/errors.rs invalid_time_span / initial_step_size_too_big: the
    # Err(BadInput) comes before any solout call; no event points.
    for t0, tf, m in [(0.0, 0.0, ExplicitRungeKutta.dopri5()),
                      (0.0, 1.0, ExplicitRungeKutta.dopri5().h0(10.0)),
                      (0.0, 1.0, ExplicitRungeKutta.rk4(10.0))]:
        ens = IVP.ode(Exponential(1.0), t0, tf, [1.0], n=1).method(m) \
            .event(Event.time(10.0, EventConfig().terminal()))
        sol = orc.solve_ode(ens)
        assert sol.status[0] == DE_STATUS_BAD_INPUT and sol.n_events[0] == 0
        assert np.isnan(sol.event_t).all()


def test_terminate_after_three(orc):
    ens = cosine().event(Event.component(0, 0.0, EventConfig().terminate_after(3)))
    sol = orc.solve_ode(ens)
    assert sol.status[0] == DE_STATUS_INTERRUPTED and sol.n_events[0] == 3
    assert ZEROS[2] <= sol.t[0] < ZEROS[3]
    assert np.isnan(sol.event_t[3:]).all()


def test_event_on_a_step_end_and_the_solution_tail(orc):
    # Euler with h = 1 on y' = 0: the steps end exactly on t = 1, 2, ...  With
    # g = t - 10, the step ending at 10 has g = 0: brent_dekker returns b = 10
    # at once (fb == 0).  DefaultSolout pushed (10, y) just before, so the
    # event point is within abs_tol of the Solution's last time and is not
    # pushed (event.rs:379-385) -- yet it counts, and terminal() stops the
    # solve.  With a t_eval output the last point is 0.5 and it is pushed.
    def ens():
        return IVP.ode(Exponential(0.0), 0.0, 20.0, [2.0], n=1).method(
            ExplicitRungeKutta.euler(1.0)).event(Event.time(10.0, EventConfig().terminal()))
    sol = orc.solve_ode(ens())
    assert sol.status[0] == DE_STATUS_INTERRUPTED and sol.t[0] == 10.0
    assert sol.n_events[0] == 0 and sol.accepted[0] == 10
    e = IVP.ode(Exponential(0.0), 0.0, 20.0, [2.0], n=1).t_eval([0.5]).method(
        ExplicitRungeKutta.euler(1.0)).event(Event.time(10.0, EventConfig().terminal()))
    sol = orc.solve_ode(e)
    assert sol.status[0] == DE_STATUS_INTERRUPTED and sol.n_events[0] == 1
    assert sol.events(0)[0][0] == 10.0 and sol.n_dense[0] == 1


def test_zero_at_a_step_end_and_the_sign_of_zero(orc):
    # y' = -1, y(0) = 10, Euler h = 1: y is exactly 9, 8, ..., +0 at t = 10,
    # then -1.  f64::signum(+0) = +1, so g = y goes 1 -> +0 without a sign
    # change; the change is seen on the next step (+0 -> -1), where Negative
    # needs g_prev > 0 and misses it, while Both finds the root b = 10 at once
    # (fb == 0) and pushes it (the Solution's last time is 11).
    def ens(d):
        return IVP.ode(Linear(-1.0, 0.0), 0.0, 20.0, [10.0], n=1).method(
            ExplicitRungeKutta.euler(1.0)).event(Event.component(0, 0.0, EventConfig(d)))
    both = orc.solve_ode(ens(CrossingDirection.Both))
    assert both.n_events[0] == 1 and both.events(0)[0][0] == 10.0
    assert orc.solve_ode(ens(CrossingDirection.Negative)).n_events[0] == 0
    assert orc.solve_ode(ens(CrossingDirection.Positive)).n_events[0] == 0


def test_backward_solves_and_failed_interpolations(orc):
    # Interpolation::interpolate rejects t < t_prev or t > t (ordinary.rs:
    # 303-309), and in a backward solve t < t_prev, so every interior point
    # fails.  TEvalSolout and the crossing solouts unwrap it (the reference
    # panics: DE_STATUS_OUT_OF_BOUNDS); brent_dekker's `.ok()?` gives no root,
    # so a backward event solve runs to the end without event points.
    back = IVP.ode(Harmonic(1.0), 20.0, 0.0, np.array([[1.0], [0.0]])).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(100000))
    sol = orc.solve_ode(back.t_eval([20.0, 15.0, 10.0]))
    assert sol.status[0] == DE_STATUS_OUT_OF_BOUNDS
    assert sol.n_dense[0] == 1 and np.isnan(sol.dense[2:]).all()
    assert 10.0 < sol.t[0] < 15.0  # the step that passed over 15
    sol = orc.solve_ode(back.crossing(0, 0.0))
    assert sol.status[0] == DE_STATUS_OUT_OF_BOUNDS and sol.n_events[0] == 0
    sol = orc.solve_ode(back.event(Event.component(0, 0.0)))
    assert sol.status[0] == DE_STATUS_COMPLETE and sol.n_events[0] == 0


def test_max_events_keeps_the_first_points(orc):
    ens = cosine().event(Event.component(0, 0.0)).options(max_events=2)
    sol = orc.solve_ode(ens)
    assert sol.n_events[0] == 6
    t, _ = sol.events(0)
    assert len(t) == 2 and np.abs(t - ZEROS[:2]).max() < 1e-8


# ------------------------------------------- bit equality with pyref_events --
def _lorenz_f(rho):
    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z
    return f


def _harm_f(k):
    def f(t, y, dy):
        dy[0] = y[1]
        dy[1] = -k * y[0]
    return f


def _tab(method):
    if method == "dopri5":
        return DP5
    from test_oracle_math import _scipy_dop853
    return _scipy_dop853()


CASES = []
for _m in ("dopri5", "dop853"):
    for _kind in ("event", "event_teval", "event_plane_term", "crossing", "hyperplane"):
        for _d in DIRS:
            CASES.append((_m, _kind, _d))


@pytest.mark.parametrize("method,kind,d", CASES)
def test_oracle_events_bit_equal_python_restatement(orc, method, kind, d):
    rng = np.random.default_rng(hash((method, kind, d)) % 2**32)
    lorenz = kind in ("event", "event_plane_term", "hyperplane")
    if lorenz:
        rho = float(rng.uniform(20, 40))
        y0 = list(1.0 + 0.1 * rng.standard_normal(3))
        f, sysm, tf = _lorenz_f(rho), Lorenz(10.0, rho, 8.0 / 3.0), 3.0
    else:
        k = float(rng.uniform(0.5, 2.0))
        y0 = [1.0, float(rng.uniform(-0.5, 0.5))]
        f, sysm, tf = _harm_f(k), Harmonic(k), 12.0
    rtol, atol = 1e-9, 1e-11
    m = (ExplicitRungeKutta.dopri5() if method == "dopri5" else ExplicitRungeKutta.dop853()) \
        .rtol(rtol).atol(atol).max_steps(100000)
    ens = IVP.ode(sysm, 0.0, tf, np.array(y0).reshape(-1, 1)).method(m).options(max_events=64)
    sol = PE.Solution()
    dname = DIRS[d]
    if kind == "event":
        ens = ens.event(Event.component(2, rho - 1.0, EventConfig(d)))
        so = PE.EventWrapped(PE.DefaultSolout(sol), sol, lambda t, y: y[2] - (rho - 1.0),
                             0.0, tf, dname, None)
    elif kind == "event_teval":
        pts = list(np.linspace(0.0, tf, 29))
        ens = ens.t_eval(pts).event(Event.component(0, 0.25, EventConfig(d)))
        so = PE.EventWrapped(PE.TEvalSolout(sol, pts, 0.0, tf), sol, lambda t, y: y[0] - 0.25,
                             0.0, tf, dname, None)
    elif kind == "event_plane_term":
        p, nrm = [0.0, 0.0, rho - 1.0], [0.3, -0.2, 1.0]
        ens = ens.event(Event.plane(p, nrm, EventConfig(d).terminate_after(2)))
        g = PE.signed_distance(p, nrm)
        so = PE.EventWrapped(PE.DefaultSolout(sol), sol, lambda t, y: g(y), 0.0, tf, dname, 2)
    elif kind == "crossing":
        ens = ens.crossing(1, 0.1, d)
        so = PE.NewtonCrossing(sol, lambda y: y[1] - 0.1, dname)
    else:
        p, nrm = [1.0, 1.0, rho - 1.0], [0.5, 0.5, 2.0]
        ens = ens.hyperplane_crossing(p, nrm, d)
        so = PE.NewtonCrossing(sol, PE.signed_distance(p, PE.normalise(nrm)), dname,
                               hyperplane=True)
    want = pyref_dopri5.solve_dp(f, 0.0, tf, y0, rtol, atol, _tab(method), solout=so)
    got = orc.solve_ode(ens, pow_mode=1)
    pts_want = so.root_points if kind.startswith("event") else sol.points
    t, y = got.events(0)
    assert got.n_events[0] == len(pts_want)
    assert t.tolist() == [p_[0] for p_ in pts_want]
    assert y.tolist() == [p_[1] for p_ in pts_want]
    assert list(got.y) == want[0] and got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == want[2:6]
