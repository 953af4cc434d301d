"""Point outputs of spring-chain ensembles in the oracle (CPU): DefaultSolout,
IVP::dense(n) and IVP::even(dt) over systems of different state lengths.

Each trajectory's point states are kept back to back from
Y_i = sum_{j<i} slots_j * 2 sizes[j].  Pins that are not a second
transcription of the solouts:
  * the ragged layout: every system's points equal those of the same system
    solved alone (one trajectory, where the layout is trivial);
  * the points land where the solouts put them (step ends, n - 1 equal
    subdivisions of every step, the EvenSolout grid) and their states equal
    the t_eval output at the same times, the final point the final state;
  * slots beyond a trajectory's capacity are counted, not written, and the
    neighbouring trajectories' states are not touched.
"""
import ctypes as C

import numpy as np
import pytest

import deamd
from deamd import DE_STATUS_COMPLETE, ExplicitRungeKutta, IVP, SpringChain

SIZES = np.array([3, 1, 8, 5, 2, 13], dtype=np.int32)
METHODS = {
    "dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9),
    "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-8).atol(1e-10),
    "rk4": lambda: ExplicitRungeKutta.rk4(0.02),
    "rkv767e": lambda: ExplicitRungeKutta.rkv767e().rtol(1e-8).atol(1e-10),
}


def chains():
    sysm, y0 = deamd.gen_spring(len(SIZES), seed=77, sizes=SIZES.copy())
    return sysm, y0


def one(sysm, y0, i):
    """System i of the ensemble as a one-trajectory ensemble."""
    a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
    pa, pb = sysm.param_offset[i], sysm.param_offset[i + 1]
    return SpringChain(sysm.sizes[i:i + 1], sysm.params_packed[pa:pb]), y0[a:b]


def teval_states(orc, sysm, y0, name, i, times):
    """TEvalSolout rows of system i at `times`."""
    s1, y1 = one(sysm, y0, i)
    sol = orc.solve_ode(IVP.ode(s1, 0.0, 2.0, y1).method(METHODS[name]()).t_eval(times))
    return sol.dense.reshape(len(times), -1)


def same(a, b):
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("name", list(METHODS))
@pytest.mark.parametrize("out", ["steps", "dense", "even"])
def test_ragged_layout_equals_single_system_solves(orc, name, out):
    sysm, y0 = chains()
    pick = {"steps": lambda e: e.all_steps(), "dense": lambda e: e.dense(3),
            "even": lambda e: e.even(0.15)}[out]
    sol = orc.solve_ode(pick(IVP.ode(sysm, 0.0, 2.0, y0).method(METHODS[name]())))
    assert (sol.status == DE_STATUS_COMPLETE).all()
    assert len(sol.points_y) == int(sol.points_ypos[-1])
    for i in range(len(SIZES)):
        s1, y1 = one(sysm, y0, i)
        alone = orc.solve_ode(pick(IVP.ode(s1, 0.0, 2.0, y1).method(METHODS[name]())))
        t, y = sol.points(i)
        ta, ya = alone.points(0)
        assert y.shape == (len(t), 2 * SIZES[i])
        assert np.array_equal(t, ta) and np.array_equal(y, ya)
        assert sol.n_points[i] == alone.n_points[0]


@pytest.mark.parametrize("name", list(METHODS))
def test_steps_dense_and_even_points(orc, name):
    sysm, y0 = chains()
    mk = lambda: IVP.ode(sysm, 0.0, 2.0, y0).method(METHODS[name]())
    steps = orc.solve_ode(mk().all_steps())
    dense = orc.solve_ode(mk().dense(4))
    even = orc.solve_ode(mk().even(0.25))
    assert np.array_equal(steps.n_points, steps.accepted + 1)
    assert np.array_equal(dense.n_points, 4 * dense.accepted + 1)
    grid = np.arange(9) * 0.25          # 0.25 accumulates exactly to tf = 2
    for i in range(len(SIZES)):
        a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
        t, y = steps.points(i)
        assert t[0] == 0.0 and np.array_equal(y[0], y0[a:b])
        assert np.all(np.diff(t) > 0) and t[-1] == steps.t[i]
        assert np.array_equal(y[-1], steps.y[a:b])
        assert same(teval_states(orc, sysm, y0, name, i, t), y)
        td, yd = dense.points(i)
        assert np.array_equal(td[::4], t) and np.array_equal(yd[::4], y)
        for k in range(1, 4):
            assert np.array_equal(td[k::4], t[:-1] + (k * (t[1:] - t[:-1])) / 4.0)
        assert same(teval_states(orc, sysm, y0, name, i, td), yd)
        te, ye = even.points(i)
        assert np.array_equal(te, grid)
        assert np.array_equal(ye[-1], even.y[a:b])
        assert same(teval_states(orc, sysm, y0, name, i, te[:-1]), ye[:-1])


def test_points_beyond_the_slots_are_counted_not_written(orc):
    sysm, y0 = chains()
    ens = IVP.ode(sysm, 0.0, 2.0, y0).method(METHODS["dopri5"]()).all_steps()
    full = orc.solve_ode(ens)
    off = np.array([0, 4, 4, 9, 10, 13, 15], dtype=np.int64)
    b, res, keep = ens.batch(off)
    assert orc.load().de_oracle_ode_solve(C.byref(b), C.byref(res.struct()), 0, 1) == 0
    assert np.array_equal(res.n_points, full.n_points)
    assert not np.isnan(res.points_y).any()      # every kept slot written, none past
    for i in range(len(SIZES)):
        t, y = res.points(i)
        k = off[i + 1] - off[i]
        assert len(t) == k and y.shape == (k, 2 * SIZES[i])
        ft, fy = full.points(i)
        assert np.array_equal(t, ft[:k]) and np.array_equal(y, fy[:k])


def test_dataframes_of_ragged_systems(orc):
    # Solution::to_polars of each spring chain (a pandas DataFrame here):
    # columns t, y0 .. y{2N-1} of that system's own length, rows its points,
    # its t_eval points produced, or its final state.
    sysm, y0 = chains()
    base = lambda: IVP.ode(sysm, 0.0, 2.0, y0).method(METHODS["dopri5"]())
    pts = orc.solve_ode(base().all_steps())
    te = np.linspace(0.0, 2.0, 5)
    tev = orc.solve_ode(base().t_eval(te))
    fin = orc.solve_ode(base())
    for i in range(len(SIZES)):
        n = 2 * int(SIZES[i])
        a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
        df = pts.to_dataframe(i)
        t, y = pts.points(i)
        assert list(df.columns) == ["t"] + ["y%d" % c for c in range(n)]
        assert np.array_equal(df.to_numpy(), np.column_stack([t, y]))
        df = tev.to_dataframe(i)
        rows = tev.dense[len(te) * a:len(te) * b].reshape(len(te), n)
        assert np.array_equal(df["t"].to_numpy(), te)
        assert np.array_equal(df.to_numpy()[:, 1:], rows)
        df = fin.to_dataframe(i)
        assert df.shape == (1, n + 1) and df["t"][0] == fin.t[i]
        assert np.array_equal(df.to_numpy()[0, 1:], fin.y[a:b])
