"""The oracle's Solution::to_csv restatement (src/solution.rs:190-230).

Numbers: Rust's `{:?}` of an f64 is the shortest decimal that reads back as
the same double (closest of those), positional for 1e-4 <= |x| < 1e16 with at
least one fractional digit, else d[.ddd]e<exp>.  CPython's repr() is an
independent implementation of the same digits (David Gay's shortest mode)
with the same switch points and only a different exponent spelling ("e-07",
"e+16"), so the oracle is pinned against it on random bit patterns and every
power of two and ten, plus the fixed strings Rust prints for named values.

Documents: the header, one line per point, and for t_eval output exactly the
points TEvalSolout pushes (checked against which dense rows the oracle's own
solve produced); for a point output (every step, dense(n), even(dt)) the kept
points in order.  to_dataframe (Solution::to_polars) holds the same rows as
the document, parsed back.
"""
import math
import re

import numpy as np
import pytest

from deamd import (DE_CSV_EVENTS, DE_CSV_FINAL, DE_CSV_POINTS, DE_CSV_TEVAL, CrossingDirection, ExplicitRungeKutta,
                   Exponential, Harmonic, IVP, Lorenz)


def python_debug(x):
    """Rust's {:?} spelled from CPython's repr()."""
    if x != x:
        return "NaN"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    r = repr(float(x))
    m = re.match(r"^(-?[0-9.]+)e([+-])(\d+)$", r)
    if m:
        return "%se%s%d" % (m.group(1), "-" if m.group(2) == "-" else "", int(m.group(3)))
    return r


# Strings Rust prints for `format!("{:?}", x)`.
NAMED = [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.0, "-1.0"), (0.1, "0.1"), (0.3, "0.3"),
    (2.0 / 3.0, "0.6666666666666666"), (100.0, "100.0"), (123456789.0, "123456789.0"),
    (0.0001, "0.0001"), (0.00001, "1e-5"), (1e-7, "1e-7"), (1.5e-7, "1.5e-7"),
    (1e15, "1000000000000000.0"), (1e16, "1e16"), (9999999999999998.0, "9999999999999998.0"),
    (1.7976931348623157e308, "1.7976931348623157e308"),
    (2.2250738585072014e-308, "2.2250738585072014e-308"), (5e-324, "5e-324"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (-float("inf"), "-inf"),
    (22026.465794806718, "22026.465794806718"), (-2.5e20, "-2.5e20"),
]


@pytest.mark.parametrize("x,text", NAMED)
def test_named_values(orc, x, text):
    assert orc.fmt_debug(x) == text


def test_negative_nan_prints_without_sign(orc):
    assert orc.fmt_debug(-float("nan")) == "NaN"


def test_random_bits_equal_python_repr(orc):
    rng = np.random.default_rng(7)
    xs = rng.integers(0, 2 ** 64, 40000, dtype=np.uint64).view(np.float64)
    bad = [float(x).hex() for x in xs if orc.fmt_debug(x) != python_debug(x)]
    assert not bad, bad[:5]


def test_powers_and_neighbours_equal_python_repr(orc):
    xs = [math.ldexp(1.0, e) for e in range(-1074, 1024)]
    xs += [10.0 ** e for e in range(-323, 309)]
    xs += list(np.nextafter(np.array([10.0 ** e for e in range(-300, 300)]), 0.0))
    xs += list(np.nextafter(np.array([10.0 ** e for e in range(-300, 300)]), np.inf))
    xs += [float(i) for i in range(1, 3000)] + [float(2 ** 53 - i) for i in range(50)]
    bad = [float(x).hex() for x in xs if orc.fmt_debug(x) != python_debug(x)]
    assert not bad, bad[:5]


def _doc(data, offsets, i):
    return data[offsets[i]:offsets[i + 1]].decode()


def test_final_state_documents(orc):
    p_y0 = np.array([[1.0], [2.0], [3.0]])
    ens = IVP.ode(Lorenz(10.0, 28.0, 8.0 / 3.0), 0.0, 0.5, p_y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10))
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_FINAL)
    lines = _doc(data, offsets, 0).split("\n")
    assert lines[0] == "t,y0,y1,y2" and lines[-1] == "" and len(lines) == 3
    fields = lines[1].split(",")
    assert fields[0] == python_debug(sol.t[0])
    assert fields[1:] == [python_debug(v) for v in sol.y]


def _produced_rows(dense, n, dim, n_eval, i):
    rows = dense.reshape(n_eval, dim, n)[:, :, i]
    return [e for e in range(n_eval) if not np.isnan(rows[e]).any()]


@pytest.mark.parametrize("pts", [
    [0.0, 0.25, 0.5, 1.0],             # t0 is the first point
    [-1.0, 0.0, 0.0, 0.5, 0.9],        # points before t0 and a repeated t0 are passed over
    [0.0, 0.0, 0.3, 0.3, 2.0, 3.0],    # repeated t0 after index 0; points past tf never come
    [0.1, 0.2],
])
def test_t_eval_documents_hold_the_produced_points(orc, pts):
    n = 5
    k = np.linspace(0.5, 2.0, n)
    ens = IVP.ode(Harmonic(k), 0.0, 1.0, np.vstack([np.ones(n), np.zeros(n)])).t_eval(pts) \
        .method(ExplicitRungeKutta.dopri5())
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
    t_sorted = ens._t_eval
    for i in range(n):
        rows = _produced_rows(sol.dense, n, 2, len(t_sorted), i)
        assert len(rows) == sol.n_dense[i]
        lines = _doc(data, offsets, i).split("\n")[1:-1]
        assert [float(ln.split(",")[0]) for ln in lines] == [t_sorted[e] for e in rows]
        for ln, e in zip(lines, rows):
            y = sol.dense.reshape(len(t_sorted), 2, n)[e, :, i]
            assert ln.split(",")[1:] == [python_debug(v) for v in y]


def test_event_documents_and_empty_t_eval(orc):
    n = 3
    k = np.array([1.0, 2.0, 3.0])
    y0 = np.vstack([np.ones(n), np.zeros(n)])
    ens = IVP.ode(Harmonic(k), 0.0, 6.0, y0).method(ExplicitRungeKutta.dopri5()) \
        .crossing(0, 0.0, CrossingDirection.Both).options(max_events=8)
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_EVENTS)
    for i in range(n):
        t, y = sol.events(i)
        lines = _doc(data, offsets, i).split("\n")
        assert lines[0] == "t,y0,y1" and len(lines) == len(t) + 2
        assert [ln.split(",")[0] for ln in lines[1:-1]] == [python_debug(v) for v in t]
    # no points: an empty document
    ens = IVP.ode(Exponential(1.0), 0.0, 1.0, [1.0], n=1).t_eval([5.0]) \
        .method(ExplicitRungeKutta.dopri5())
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
    assert sol.n_dense[0] == 0 and offsets.tolist() == [0, 0]


@pytest.mark.parametrize("mk", [lambda e: e.all_steps(), lambda e: e.dense(3),
                                lambda e: e.even(0.1)], ids=["steps", "dense3", "even"])
def test_point_documents_and_dataframes(orc, mk):
    p = np.array([[10.0, 9.0], [28.0, 27.0], [8.0 / 3.0, 2.5]])
    y0 = np.array([[1.0, 0.5], [2.0, 1.0], [3.0, 2.0]])
    ens = mk(IVP.ode(Lorenz(*p), 0.0, 0.7, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10)))
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_POINTS)
    for i in range(2):
        t, y = sol.points(i)
        lines = _doc(data, offsets, i).split("\n")
        assert lines[0] == "t,y0,y1,y2" and lines[-1] == "" and len(lines) == len(t) + 2
        for ln, tk, yk in zip(lines[1:-1], t, y):
            assert ln.split(",") == [python_debug(v) for v in [tk, *yk]]
        df = sol.to_dataframe(i)
        assert list(df.columns) == ["t", "y0", "y1", "y2"]
        parsed = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:-1]])
        assert np.array_equal(df.to_numpy(), parsed)


def test_dataframes_of_t_eval_and_final_outputs(orc):
    n = 3
    k = np.array([1.0, 2.0, 3.0])
    y0 = np.vstack([np.ones(n), np.zeros(n)])
    for pts in ([0.0, 0.25, 0.5, 1.0], [-1.0, 0.0, 0.0, 0.5, 0.9]):
        ens = IVP.ode(Harmonic(k), 0.0, 1.0, y0).t_eval(pts).method(ExplicitRungeKutta.dopri5())
        sol = orc.solve_ode(ens)
        data, offsets = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
        for i in range(n):
            lines = _doc(data, offsets, i).split("\n")[1:-1]
            parsed = np.array([[float(v) for v in ln.split(",")] for ln in lines])
            df = sol.to_dataframe(i)
            assert list(df.columns) == ["t", "y0", "y1"] and len(df) == sol.n_dense[i]
            assert np.array_equal(df.to_numpy(), parsed)
    ens = IVP.ode(Harmonic(k), 0.0, 1.0, y0).method(ExplicitRungeKutta.dopri5())
    sol = orc.solve_ode(ens)
    data, offsets = orc.ode_csv(ens, sol, DE_CSV_FINAL)
    for i in range(n):
        line = _doc(data, offsets, i).split("\n")[1]
        assert np.array_equal(sol.to_dataframe(i).to_numpy()[0], [float(v) for v in line.split(",")])
