"""The CPU oracle against the reference crate's own known-answer tests.

Each case is a row of the reference's tests for the explicit Runge-Kutta path
(expected values are the SciPy DOP853 references quoted in
tests/ode/accuracy.rs:1), run through the oracle with the same settings and the
same tolerance.  With both pow modes: the project pow and the C library pow
(what a Rust build of the reference calls).
"""
import math

import numpy as np
import pytest

from deamd import (ExplicitRungeKutta, Exponential, Harmonic, IVP, Linear, Logistic, Robertson,
                   EulerMaruyama, Milstein, LinearDiagonal, DE_NOISE_CONSTANT, DE_NOISE_TABLE,
                   DE_STATUS_BAD_INPUT, DE_STATUS_COMPLETE)

POW_MODES = [0, 1]

# (name, system, t0, tf, y0, expected, [(method, tolerance), ...])
# tests/ode/accuracy.rs:70-103, 177-213, 286-321, 395-430, 503-539, 622-671 (DOP853
# rows at :78, 187, 296, 405, 514, 666)
ACCURACY = [
    ("exponential_growth_positive_direction", Exponential(1.0), 0.0, 10.0, [1.0],
     [22026.46579479],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-12).atol(1e-12), 1e-3), (ExplicitRungeKutta.dopri5, 1e1),
      (lambda: ExplicitRungeKutta.rk4(0.01), 1e3), (lambda: ExplicitRungeKutta.euler(0.01), 2e3)]),
    ("exponential_growth_negative_direction", Exponential(1.0), 0.0, -10.0, [22026.46579479],
     [1.0],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-12).atol(1e-12), 1e-3), (ExplicitRungeKutta.dopri5, 1e-2),
      (lambda: ExplicitRungeKutta.rk4(-0.01), 1e-2), (lambda: ExplicitRungeKutta.euler(-0.01), 1e-1)]),
    ("linear_equation", Linear(1.0, 1.0), 0.0, 10.0, [1.0], [44051.93158958],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-12).atol(1e-12), 1e-3), (ExplicitRungeKutta.dopri5, 1e1),
      (lambda: ExplicitRungeKutta.rk4(0.01), 1e3), (lambda: ExplicitRungeKutta.euler(0.01), 1e4)]),
    ("harmonic_oscillator", Harmonic(1.0), 0.0, 10.0, [1.0, 0.0], [-0.83907153, 0.54402111],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-12).atol(1e-12), 1e-3), (ExplicitRungeKutta.dopri5, 1e-2),
      (lambda: ExplicitRungeKutta.rk4(0.01), 1e-1), (lambda: ExplicitRungeKutta.euler(0.01), 1e-1)]),
    ("logistic_equation", Logistic(1.0, 10.0), 0.0, 10.0, [0.1], [9.95525518],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-12).atol(1e-12), 1e-3), (ExplicitRungeKutta.dopri5, 1e-3),
      (lambda: ExplicitRungeKutta.rk4(0.01), 1e3), (lambda: ExplicitRungeKutta.euler(0.01), 1e-2)]),
    ("robertson_stiff_problem", Robertson(), 0.0, 0.4, [1.0, 0.0, 0.0],
     [0.9851721139, 3.3863953790e-05, 0.0147940222],
     [(lambda: ExplicitRungeKutta.dop853().rtol(1e-10).atol(1e-12), 1e-1),
      (lambda: ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10), 5e-1)]),
]


@pytest.mark.parametrize("pow_mode", POW_MODES)
@pytest.mark.parametrize("case", ACCURACY, ids=[c[0] for c in ACCURACY])
def test_accuracy_rows(orc, case, pow_mode):
    name, system, t0, tf, y0, expected, methods = case
    for make_method, tol in methods:
        ens = IVP.ode(system, t0, tf, y0, n=1).method(make_method())
        sol = orc.solve_ode(ens, pow_mode=pow_mode)
        assert sol.status[0] == DE_STATUS_COMPLETE, (name, sol.status)
        for c, e in enumerate(expected):
            assert abs(sol.y[c] - e) < tol, (name, c, sol.y[c], e)


@pytest.mark.parametrize("make_method", [ExplicitRungeKutta.dop853, ExplicitRungeKutta.dopri5,
                                         lambda: ExplicitRungeKutta.rk4(0.01)])
def test_interpolation(orc, make_method):
    # tests/ode/interpolation.rs:22-72: y' = y on [0, 2], t_eval [0.5, 1.0, 1.69]
    ens = IVP.ode(Exponential(1.0), 0.0, 2.0, [1.0], n=1).t_eval([0.5, 1.0, 1.69]) \
        .method(make_method())
    sol = orc.solve_ode(ens)
    assert sol.n_dense[0] == 3
    assert abs(sol.dense[0] - math.exp(0.5)) < 1e-3
    # the remaining points are not checked by the reference; they are by us
    assert abs(sol.dense[1] - math.exp(1.0)) < 1e-3
    assert abs(sol.dense[2] - math.exp(1.69)) < 1e-3


@pytest.mark.parametrize("make_method", [ExplicitRungeKutta.dop853, ExplicitRungeKutta.dopri5,
                                         lambda: ExplicitRungeKutta.rk4(0.1),
                                         lambda: ExplicitRungeKutta.euler(0.1)])
def test_invalid_time_span(orc, make_method):
    # tests/ode/errors.rs:87-107: tf == t0 -> BadInput
    ens = IVP.ode(Exponential(1.0), 0.0, 0.0, [1.0], n=1).method(make_method())
    assert orc.solve_ode(ens).status[0] == DE_STATUS_BAD_INPUT


@pytest.mark.parametrize("make_method", [lambda: ExplicitRungeKutta.dop853().h0(10.0),
                                         lambda: ExplicitRungeKutta.dopri5().h0(10.0),
                                         lambda: ExplicitRungeKutta.rk4(10.0),
                                         lambda: ExplicitRungeKutta.euler(10.0)])
def test_initial_step_size_too_big(orc, make_method):
    # tests/ode/errors.rs:110-130: |h0| > |tf - t0| -> BadInput
    ens = IVP.ode(Exponential(1.0), 0.0, 1.0, [1.0], n=1).method(make_method())
    assert orc.solve_ode(ens).status[0] == DE_STATUS_BAD_INPUT


def test_ode_from_fn_euler(orc):
    # tests/ode/from_fn.rs:3-18: Euler h=0.1 on y'=y over [0,1] -> 1.1^10
    ens = IVP.ode(Exponential(1.0), 0.0, 1.0, [1.0], n=1).method(ExplicitRungeKutta.euler(0.1))
    sol = orc.solve_ode(ens)
    assert abs(sol.y[0] - 2.5937) < 1e-3
    assert abs(sol.y[0] - 1.1 ** 10) < 1e-12


def test_sde_from_fn(orc):
    # tests/sde/from_fn.rs:4-23: drift 0.5y, diffusion 0.1y, dW = 0.1, EM h=0.01
    sys_ = LinearDiagonal(1, [0.5, 0.1])
    ens = IVP.sde(sys_, 0.0, 0.1, [1.0], noise=DE_NOISE_CONSTANT, noise_constant=0.1) \
        .method(EulerMaruyama(0.01))
    sol = orc.solve_sde(ens)
    assert sol.status[0] == DE_STATUS_COMPLETE
    assert abs(sol.t[0] - 0.1) < 1e-10


def test_multidimensional_milstein(orc):
    # tests/sde/milstein.rs:43-65
    dw = np.array([[0.1, -0.1], [-0.05, 0.2], [0.15, -0.05], [0.0, 0.1]])
    sys_ = LinearDiagonal(2, [-1.0, 0.5, -2.0, 0.1])
    h = (1.0 - 0.0) / 4
    ens = IVP.sde(sys_, 0.0, 1.0, [1.0, 1.0], noise=DE_NOISE_TABLE, noise_table=dw) \
        .method(Milstein(h))
    sol = orc.solve_sde(ens)
    assert sol.steps[0] == 4          # sol.t.len() == steps + 1
    assert sol.y[0] != 1.0 and sol.y[1] != 1.0
    # closed form of the derivative-free Milstein step for a linear diagonal
    # SDE: y <- y (1 + a h + b dW + b^2 (dW^2 - h) / 2), exact arithmetic.
    a = [-1.0, -2.0]
    b = [0.5, 0.1]
    for c in range(2):
        y = 1.0
        for n in range(4):
            d = dw[n, c]
            y = y * (1 + a[c] * h + b[c] * d + b[c] ** 2 * (d * d - h) / 2)
        assert abs(sol.y[c] - y) < 1e-14 * abs(y)
