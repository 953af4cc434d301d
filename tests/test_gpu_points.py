"""Point outputs on the GPU against the oracle, bit for bit: DefaultSolout
(every accepted step), IVP::dense(n) and IVP::even(dt) for the DOPRI5, DOP853,
RK4, Euler and embedded-family lanes, a backward solve, per-trajectory time
spans, both persistent schedules and slots that run out."""
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_SCHED_PERSISTENT_FIFO, DE_SCHED_PLAIN, DE_STATUS_COMPLETE,
                   DE_STATUS_OUT_OF_BOUNDS, ExplicitRungeKutta, Harmonic, IVP, Lorenz)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)

METHODS = {
    "dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10),
    "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11),
    "rk4": lambda: ExplicitRungeKutta.rk4(0.01),
    "euler": lambda: ExplicitRungeKutta.euler(0.002),
    "rkf45": lambda: ExplicitRungeKutta.rkf45().rtol(1e-8).atol(1e-10),
    "rkv878e": lambda: ExplicitRungeKutta.rkv878e().rtol(1e-9).atol(1e-11),
}
OUTPUTS = {"steps": lambda e: e.all_steps(), "dense3": lambda e: e.dense(3),
           "even": lambda e: e.even(0.05)}


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def assert_points_equal(got, want):
    for f in ("status", "accepted", "rejected", "evals", "n_points", "points_offset"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    assert same(got.t, want.t) and same(got.y, want.y)
    for i in range(got.n):   # the kept points (slots past n_points are not written)
        gt, gy = got.points(i)
        wt, wy = want.points(i)
        assert same(gt, wt) and same(gy, wy), i


@pytest.mark.parametrize("out", list(OUTPUTS))
@pytest.mark.parametrize("name", list(METHODS))
def test_point_outputs_bit_exact(orc, name, out):
    p, y0 = deamd.gen_lorenz(512, seed=93)
    tf = np.linspace(0.5, 2.0, 512)        # uneven spans: uneven point counts
    ens = OUTPUTS[out](IVP.ode(Lorenz(*p), 0.0, tf, y0).method(METHODS[name]().max_steps(100000)))
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).all()
    assert_points_equal(ens.solve(), want)


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT_FIFO])
def test_point_outputs_under_other_schedules(orc, schedule):
    p, y0 = deamd.gen_lorenz(300, seed=94)
    ens = IVP.ode(Lorenz(*p), 0.0, 1.5, y0).method(METHODS["dopri5"]().max_steps(100000)) \
        .dense(5).options(schedule=schedule)
    assert_points_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))


def test_backward_point_outputs(orc):
    for mk in (lambda e: e.all_steps(), lambda e: e.dense(3), lambda e: e.even(0.5)):
        ens = mk(IVP.ode(Harmonic(1.0), 0.0, -3.0, [1.0, 0.0], n=64).method(
            METHODS["dopri5"]()))
        want = orc.solve_ode(ens)
        assert_points_equal(ens.solve(), want)
    assert (want.status == DE_STATUS_OUT_OF_BOUNDS).all()


def test_point_slots_that_run_out(orc):
    import ctypes as C
    p, y0 = deamd.gen_lorenz(200, seed=95)
    ens = IVP.ode(Lorenz(*p), 0.0, 1.0, y0).method(METHODS["dopri5"]().max_steps(100000)) \
        .all_steps()
    full = orc.solve_ode(ens, threads=THREADS)
    cap = np.minimum(full.n_points, np.arange(200) % 40)       # some trajectories short of slots
    off = np.zeros(201, dtype=np.int64)
    np.cumsum(cap, out=off[1:])
    results = []
    for run in ("gpu", "oracle"):
        b, res, keep = ens.batch(off)
        if run == "gpu":
            deamd._abi.check(deamd.lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()),
                                                              None))
        else:
            assert orc.load().de_oracle_ode_solve(C.byref(b), C.byref(res.struct()), 0, 4) == 0
        results.append(res)
    assert_points_equal(*results)
    assert np.array_equal(results[0].n_points, full.n_points)


def test_points_to_csv_and_dataframe(orc):
    # Solution::to_csv of the reference's default solve: every step, formatted
    # on the GPU, byte for byte the oracle's; and to_polars' columns.
    from deamd import DE_CSV_POINTS
    from test_gpu_csv import assert_bytes_equal
    p, y0 = deamd.gen_lorenz(300, seed=96)
    ens = IVP.ode(Lorenz(*p), 0.0, 1.0, y0).method(METHODS["dopri5"]().max_steps(100000)) \
        .all_steps()
    sol = ens.solve()
    data, offsets = sol.csv()
    want, woff = orc.ode_csv(ens, sol, DE_CSV_POINTS)
    assert np.array_equal(offsets, woff)
    assert_bytes_equal(data, want)
    t, y = sol.points(7)
    df = sol.to_dataframe(7)
    assert list(df.columns) == ["t", "y0", "y1", "y2"]
    assert np.array_equal(df["t"].to_numpy(), t) and np.array_equal(df[["y0", "y1", "y2"]].to_numpy(), y)
    lines = data[offsets[7]:offsets[8]].decode().split("\n")
    assert lines[0] == "t,y0,y1,y2" and len(lines) == len(t) + 2
