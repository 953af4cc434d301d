"""Point outputs of the oracle (CPU): DefaultSolout (every accepted step, the
reference's default), IVP::dense(n) and IVP::even(dt).

Pins that are not a second transcription of the solouts:
  * the points land where the solouts' definitions put them: every step end
    (A + 1 points with t0), n - 1 equal subdivisions of every step, the grid
    t0 + k dt accumulated as EvenSolout accumulates it plus tf, with a grid
    point within 1e-12 dt of tf replaced by tf itself;
  * their states equal the TEvalSolout output at the same times (the t_eval
    path is pinned elsewhere): y_curr at a step end, the step's interpolant
    inside a step;
  * the final point is the final state; a backward solve panics (OutOfBounds)
    as soon as a dense or even point needs an interpolation, while the step
    output does not interpolate; slots beyond a trajectory's capacity are
    counted, not written.
"""
import numpy as np
import pytest

import deamd
from deamd import (DE_STATUS_COMPLETE, DE_STATUS_OUT_OF_BOUNDS, ExplicitRungeKutta, IVP,
                   Lorenz, Harmonic)

METHODS = {
    "dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10),
    "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11),
    "rk4": lambda: ExplicitRungeKutta.rk4(0.01),
    "rkf45": lambda: ExplicitRungeKutta.rkf45().rtol(1e-8).atol(1e-10),
    "rkv878e": lambda: ExplicitRungeKutta.rkv878e().rtol(1e-9).atol(1e-11),
}


def lorenz(n=6, tf=2.0, seed=91):
    p, y0 = deamd.gen_lorenz(n, seed=seed)
    return p, y0, (lambda: IVP.ode(Lorenz(*p), 0.0, tf, y0))


def teval_states(orc, mk, name, i, times):
    """TEvalSolout output of trajectory i at `times` (a one-trajectory solve)."""
    ens = mk().method(METHODS[name]().max_steps(100000)).t_eval(times)
    sol = orc.solve_ode(ens)
    dim = len(sol.y) // sol.n
    return sol.dense.reshape(len(times), dim, sol.n)[:, :, i]


def same(a, b):
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("name", list(METHODS))
def test_all_steps_are_the_step_ends(orc, name):
    p, y0, mk = lorenz()
    sol = orc.solve_ode(mk().method(METHODS[name]().max_steps(100000)).all_steps())
    assert (sol.status == DE_STATUS_COMPLETE).all()
    assert np.array_equal(sol.n_points, sol.accepted + 1)
    for i in range(sol.n):
        t, y = sol.points(i)
        assert t[0] == 0.0 and np.array_equal(y[0], y0[:, i])
        assert np.all(np.diff(t) > 0) and t[-1] == sol.t[i]
        assert np.array_equal(y[-1], sol.y.reshape(3, -1)[:, i])
        # the step ends as t_eval points: y_curr exactly
        assert same(teval_states(orc, mk, name, i, t), y)


@pytest.mark.parametrize("name", list(METHODS))
def test_dense_subdivides_every_step(orc, name):
    p, y0, mk = lorenz()
    steps = orc.solve_ode(mk().method(METHODS[name]().max_steps(100000)).all_steps())
    sol = orc.solve_ode(mk().method(METHODS[name]().max_steps(100000)).dense(4))
    assert np.array_equal(sol.n_points, 4 * sol.accepted + 1)
    for i in range(sol.n):
        t, y = sol.points(i)
        ts, ys = steps.points(i)
        assert np.array_equal(t[::4], ts) and np.array_equal(y[::4], ys)
        for k in range(1, 4):   # t_prev + k h / 4 (the reference's operation order)
            want = ts[:-1] + (k * (ts[1:] - ts[:-1])) / 4.0
            assert np.array_equal(t[k::4], want)
        assert same(teval_states(orc, mk, name, i, t), y)


@pytest.mark.parametrize("name", list(METHODS))
@pytest.mark.parametrize("dt,tf", [(0.05, 2.0), (0.1, 1.0), (0.3, 2.0)])
def test_even_grid(orc, name, dt, tf):
    p, y0, mk = lorenz(tf=tf)
    sol = orc.solve_ode(mk().method(METHODS[name]().max_steps(100000)).even(dt))
    assert (sol.status == DE_STATUS_COMPLETE).all()
    tol = dt * 1e-12 + np.finfo(float).eps * 10
    grid = [0.0]
    while grid[-1] + dt <= tf:
        grid.append(grid[-1] + dt)
    if abs(grid[-1] - tf) <= tol:
        grid[-1] = tf        # Solution::pop, then the exact final point
    elif grid[-1] != tf:
        grid.append(tf)
    grid = np.array(grid)
    for i in range(sol.n):
        t, y = sol.points(i)
        assert np.array_equal(t, grid)
        assert np.array_equal(y[-1], sol.y.reshape(3, -1)[:, i])
        assert same(teval_states(orc, mk, name, i, t[:-1]), y[:-1])
    if (dt, tf) == (0.1, 1.0):
        assert len(grid) == 11   # 0.1 added ten times is 0.9999999999999999: replaced by 1.0


def test_backward_solve_panics_when_a_point_needs_an_interpolation(orc):
    # Interpolants reject t outside [t_prev, t] of a backward step
    # (ordinary.rs:303-309), and the dense and even solouts unwrap.
    mk = lambda: IVP.ode(Harmonic(1.0), 0.0, -3.0, [1.0, 0.0], n=2).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10))
    steps = orc.solve_ode(mk().all_steps())
    assert (steps.status == DE_STATUS_COMPLETE).all()
    assert np.array_equal(steps.n_points, steps.accepted + 1)
    dense = orc.solve_ode(mk().dense(3))          # the first step interpolates
    assert (dense.status == DE_STATUS_OUT_OF_BOUNDS).all()
    assert (dense.accepted == 1).all() and (dense.n_points == 1).all()   # (t0, y0) only
    even = orc.solve_ode(mk().even(0.5))          # the first step that passes t = -0.5
    assert (even.status == DE_STATUS_OUT_OF_BOUNDS).all() and (even.n_points == 1).all()
    assert (even.accepted > 1).all() and (even.t <= -0.5).all()


def test_points_beyond_the_slots_are_counted_not_written(orc):
    p, y0, mk = lorenz(n=3)
    ens = mk().method(METHODS["dopri5"]().max_steps(100000)).all_steps()
    full = orc.solve_ode(ens)
    off = np.array([0, 5, 5, 12], dtype=np.int64)       # 5, 0 and 7 slots
    b, res, keep = ens.batch(off)
    lib = orc.load()
    import ctypes as C
    assert lib.de_oracle_ode_solve(C.byref(b), C.byref(res.struct()), 0, 1) == 0
    assert np.array_equal(res.n_points, full.n_points)
    for i in range(3):
        t, y = res.points(i)
        k = len(t)
        assert k == off[i + 1] - off[i]
        ft, fy = full.points(i)
        assert np.array_equal(t, ft[:k]) and np.array_equal(y, fy[:k])
