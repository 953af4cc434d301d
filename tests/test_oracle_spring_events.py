"""Event solouts on spring-chain ensembles in the oracle (CPU).

Each trajectory's event states are kept at event_y[max_events * state_offset[i]
+ k * dim + c].  Pins that are not a second transcription of the solouts:
  * a one-mass undamped chain with k_0 + k_1 = 1 is x'' = -x: from (1, 0) the
    component event and the crossing solout find cos t's zeros, pi/2 + j pi,
    in the configured directions, with |v| = 1 there;
  * the ragged layout: every system's event points equal those of the same
    system solved alone (where the layout is trivial);
  * a time event g = t - value (linear in t) lands on value; terminal() stops
    the solve Interrupted after the step that crosses it, and the event state
    equals the t_eval output at the event time;
  * the push rule: an event within 1e-14 of the Solution's last point is
    counted but not pushed, which under a t_eval output is the t_eval point
    just pushed and under the default output the step's end.
"""
import math

import numpy as np
import pytest

import deamd
from deamd import (DE_STATUS_COMPLETE, DE_STATUS_INTERRUPTED, CrossingDirection, Event,
                   EventConfig, ExplicitRungeKutta, IVP, SpringChain)

METHODS = {
    "dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-10).atol(1e-12),
    "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-10).atol(1e-12),
    "rk4": lambda: ExplicitRungeKutta.rk4(0.005),
    "rkv767e": lambda: ExplicitRungeKutta.rkv767e().rtol(1e-10).atol(1e-12),
}
ZEROS = math.pi / 2 + math.pi * np.arange(7)


def cosine_chain(extra_sizes=(), seed=5):
    """System 0: one mass, c = 0, k = (0.25, 0.75): x'' = -x from (1, 0); then
    random chains of the given sizes."""
    params = [np.array([0.0, 0.25, 0.75])]
    y0 = [np.array([1.0, 0.0])]
    sizes = [1]
    if extra_sizes:
        rs, ry = deamd.gen_spring(len(extra_sizes), seed=seed,
                                  sizes=np.array(extra_sizes, dtype=np.int32))
        for j in range(len(extra_sizes)):
            params.append(rs.params_packed[rs.param_offset[j]:rs.param_offset[j + 1]])
            y0.append(ry[rs.state_offset[j]:rs.state_offset[j + 1]])
            sizes.append(extra_sizes[j])
    return SpringChain(np.array(sizes, dtype=np.int32), np.concatenate(params)), np.concatenate(y0)


def one(sysm, y0, i):
    a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
    pa, pb = sysm.param_offset[i], sysm.param_offset[i + 1]
    return SpringChain(sysm.sizes[i:i + 1], sysm.params_packed[pa:pb]), y0[a:b]


@pytest.mark.parametrize("name", list(METHODS))
@pytest.mark.parametrize("d", [CrossingDirection.Both, CrossingDirection.Positive,
                               CrossingDirection.Negative])
def test_cosine_zeros_on_a_one_mass_chain(orc, name, d):
    sysm, y0 = cosine_chain()
    want = ZEROS[:6] if d == CrossingDirection.Both else (
        ZEROS[0:6:2] if d == CrossingDirection.Negative else ZEROS[1:6:2])
    mk = lambda: IVP.ode(sysm, 0.0, 20.0, y0).method(METHODS[name]().max_steps(100000))
    for ens in (mk().event(Event.component(0, 0.0, EventConfig(d))), mk().crossing(0, 0.0, d)):
        sol = orc.solve_ode(ens)
        assert sol.status[0] == DE_STATUS_COMPLETE
        assert sol.n_events[0] == len(want)
        t, y = sol.events(0)
        # DOP853: the reference's dense-output fold (s and s1 in the opposite
        # phase from Hairer's contd8) leaves interpolated points ~5e-7 off
        tol = {"rk4": 1e-7, "dop853": 1e-6}.get(name, 1e-8)
        assert np.abs(t - want).max() < tol
        assert np.abs(y[:, 0]).max() < tol
        assert np.abs(np.abs(y[:, 1]) - 1.0).max() < tol


@pytest.mark.parametrize("name", list(METHODS))
@pytest.mark.parametrize("kind", ["event", "crossing", "event_teval", "time_terminal"])
def test_ragged_layout_equals_single_system_solves(orc, name, kind):
    sysm, y0 = cosine_chain(extra_sizes=(3, 8, 2, 5))
    n = len(sysm.sizes)

    def build(s, y):
        e = IVP.ode(s, 0.0, 6.0, y).method(METHODS[name]().max_steps(100000))
        if kind == "event":
            return e.event(Event.component(1, 0.0)).options(max_events=12)
        if kind == "crossing":
            return e.crossing(0, 0.0, CrossingDirection.Positive).options(max_events=12)
        if kind == "event_teval":
            return e.t_eval(np.linspace(0.0, 6.0, 13)).event(
                Event.component(0, 0.0)).options(max_events=12)
        return e.event(Event.time(2.5, EventConfig().terminal())).options(max_events=2)

    sol = orc.solve_ode(build(sysm, y0))
    assert len(sol.event_y) == sol.max_events * int(sysm.state_offset[-1])
    for i in range(n):
        s1, y1 = one(sysm, y0, i)
        alone = orc.solve_ode(build(s1, y1))
        for f in ("status", "accepted", "rejected", "evals", "n_events"):
            assert getattr(sol, f)[i] == getattr(alone, f)[0], (f, i)
        t, y = sol.events(i)
        ta, ya = alone.events(0)
        assert y.shape == (len(t), 2 * sysm.sizes[i])
        assert np.array_equal(t, ta) and np.array_equal(y, ya)
        k = min(int(sol.n_events[i]), sol.max_events)
        lo = sol.max_events * int(sysm.state_offset[i])
        hi = sol.max_events * int(sysm.state_offset[i + 1])
        assert np.isnan(sol.event_y[lo + k * 2 * sysm.sizes[i]:hi]).all()   # past the last
        if kind == "time_terminal":
            assert sol.status[i] == DE_STATUS_INTERRUPTED and sol.n_events[i] == 1
            assert abs(t[0] - 2.5) <= 1e-12 and sol.t[i] >= 2.5
            te = orc.solve_ode(IVP.ode(s1, 0.0, 6.0, y1).method(
                METHODS[name]().max_steps(100000)).t_eval([t[0]]))
            assert np.array_equal(te.dense, y[0])


def test_event_on_a_point_already_pushed_is_counted_not_pushed(orc):
    # A time event at t = 1.5 under a t_eval output that has 1.5: TEvalSolout
    # pushes (1.5, y) first, so the root (within 1e-14 of it) is not pushed but
    # counted (terminate_after(1) still stops the solve).  Under the default
    # output the Solution's last point is the step's end, not 1.5: pushed.
    sysm, y0 = cosine_chain(extra_sizes=(4,))
    mk = lambda: IVP.ode(sysm, 0.0, 4.0, y0).method(METHODS["dopri5"]())
    ev = Event.time(1.5, EventConfig().terminal())
    with_te = orc.solve_ode(mk().t_eval([0.5, 1.5, 3.0]).event(ev))
    bare = orc.solve_ode(mk().event(ev))
    for i in range(2):
        assert with_te.status[i] == DE_STATUS_INTERRUPTED and with_te.n_events[i] == 0
        assert bare.status[i] == DE_STATUS_INTERRUPTED and bare.n_events[i] == 1
        assert abs(bare.events(i)[0][0] - 1.5) <= 1e-12
        assert with_te.accepted[i] == bare.accepted[i]
