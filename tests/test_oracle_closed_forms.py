"""Closed-form pins of the oracle's step-size machinery (CPU).

Each check compares the oracle with a value fixed by mathematics or by the
reference's own unit tests, not with a second transcription of the oracle:

  * constrain_step_size / validate_step_size_parameters: the exact values and
    failure reasons of the reference's tests (src/utils.rs:165-289);
  * the initial step (src/methods/h_init.rs:45-135) on y' = lam*y, where every
    quantity of the heuristic has a closed form (der2 = lam^2 |y0| / sk);
  * the error norm and controller (src/methods/erk/dormandprince/
    ordinary.rs:125-157, 249-267; src/traits.rs:394-411): on a linear system
    one DOPRI5 attempt maps y to R(hA) y and its error estimate is
    (R - R_hat)(hA) y, with R and R_hat the stability polynomials of the
    5th-order weights and of the published embedded 4th-order weights
    (Dormand & Prince 1980), built here in exact rationals from the
    tableau; the step-size sequence of the textbook controller
    h' = h * clamp(0.9 err^(-1/5), 0.2, 10), min(., 1) after a rejection, with
    err = |h| sqrt(sum((e_i/sk_i)^2)/n), sk_i = atol + rtol max(|y_i|, |y_new_i|),
    is computed at 40 digits and compared with the oracle's times;
  * the stiffness test (ordinary.rs:163-194) on y' = lam*y with a fixed step,
    where the reference's quotient h*sqrt(|yseg - k_7|^2 / |f(y_new) - y_7|^2)
    is h |lam| |(R-1)/z - R| / (|R| |lam - 1|), z = h lam.

test_planted_mistakes_are_caught builds the oracle with one plausible mistake
at a time and checks that these pins fail on it.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction as Fr

import numpy as np
import pytest

from deamd import (DE_STATUS_COMPLETE, DE_STATUS_MAX_STEPS, DE_STATUS_STIFFNESS,
                   ExplicitRungeKutta, Exponential, Harmonic, IVP)
from test_oracle_math import DP_A

mpmath = pytest.importorskip("mpmath")

# Dormand-Prince 5(4) weights (src/tableau/dorman_prince.rs:44-101) and the
# published embedded 4th-order weights b_hat (Hairer, Norsett, Wanner I,
# Table II.5.2); the reference stores er = b - b_hat.
DP_B = [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), Fr(0)]
DP_BHAT = [Fr(5179, 57600), Fr(0), Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200),
           Fr(187, 2100), Fr(1, 40)]


def _stage_polys():
    """P_i(z) with y_stage_i = P_i(h lam) y on y' = lam y (coefficient lists)."""
    P = []
    for i in range(7):
        p = [Fr(0)] * 8
        p[0] = Fr(1)
        for j in range(i):
            for d, c in enumerate(P[j]):
                if c and d + 1 < 8:
                    p[d + 1] += DP_A[i][j] * c
        P.append(p)
    return P


def _weights_poly(w):
    r = [Fr(0)] * 8
    r[0] = Fr(1)
    for j, P in enumerate(_stage_polys()):
        for d, c in enumerate(P):
            if c and d + 1 < 8:
                r[d + 1] += w[j] * c
    return r


R_POLY = _weights_poly(DP_B)
RHAT_POLY = _weights_poly(DP_BHAT)


def _ev(coef, z):
    return sum(mpmath.mpf(c.numerator) / c.denominator * z ** d for d, c in enumerate(coef))


def test_stability_polynomials():
    # R is the textbook DOPRI5 polynomial; R - R_hat = O(z^5).
    assert R_POLY == [1, 1, Fr(1, 2), Fr(1, 6), Fr(1, 24), Fr(1, 120), Fr(1, 600), 0]
    assert all(R_POLY[d] == RHAT_POLY[d] for d in range(5)) and R_POLY[5] != RHAT_POLY[5]


# ----------------------------------------------------- utils.rs:165-289 ----
def test_constrain_step_size_reference_values(orc):
    # src/utils.rs:169-179
    for h, want in [(0.5, 0.5), (-0.5, -0.5), (0.05, 0.1), (-0.05, -0.1), (1.5, 1.0),
                    (-1.5, -1.0)]:
        assert orc.constrain_step_size(h, 0.1, 1.0) == want


def test_validate_step_size_reference_cases(orc):
    # src/utils.rs:182-288; the number is the check that fails (1 = tf == t0,
    # 2 = wrong sign, 3 = h_min < 0, 4 = h_max < 0, 5 = h_min > h_max,
    # 6 = |h0| < h_min, 7 = |h0| > h_max, 8 = |h0| > |tf - t0|, 9 = h0 == 0)
    cases = [((0.1, 0.01, 1.0, 0.0, 10.0), 0), ((-0.1, 0.01, 1.0, 10.0, 0.0), 0),
             ((0.1, 0.01, 1.0, 5.0, 5.0), 1), ((-0.1, 0.01, 1.0, 0.0, 10.0), 2),
             ((0.1, 0.01, 1.0, 10.0, 0.0), 2), ((0.1, -0.01, 1.0, 0.0, 10.0), 3),
             ((0.1, 0.01, -1.0, 0.0, 10.0), 4), ((0.5, 1.0, 0.1, 0.0, 10.0), 5),
             ((0.001, 0.01, 1.0, 0.0, 10.0), 6), ((2.0, 0.01, 1.0, 0.0, 10.0), 7),
             ((0.5, 0.01, 1.0, 0.0, 0.2), 8), ((0.0, 0.0, 1.0, 0.0, 10.0), 9)]
    for args, want in cases:
        assert orc.validate_step_size(*args) == want, args


# -------------------------------------------------------- h_init.rs:45-135 ----
def _h_init_exponential(lam, y0, t0, tf, rtol, atol, h_min=0.0, h_max=np.inf):
    """The heuristic's closed form on y' = lam y (order 5)."""
    mp = mpmath.mpf
    lam, y0, rtol, atol = mp(lam), mp(y0), mp(rtol), mp(atol)
    posneg = 1 if tf > t0 else -1
    sk = atol + rtol * abs(y0)
    dnf = (lam * y0 / sk) ** 2
    dny = (y0 / sk) ** 2
    h = mp("1e-6") if (dnf <= mp("1e-10") or dny <= mp("1e-10")) else mpmath.sqrt(dny / dnf) * mp("0.01")
    h = min(h, mp(h_max))
    der2 = lam ** 2 * abs(y0) / sk          # |f(t0 + h, y0 + h f0) - f0| / (sk |h|)
    der12 = max(mpmath.sqrt(dnf), der2)
    h1 = h * mp("1e-3") if der12 <= mp("1e-15") else (mp("0.01") / der12) ** (mp(1) / 5)
    h = min(max(min(min(h * 100, h1), mp(h_max)), mp(h_min)), mp(abs(tf - t0)))
    return posneg * h


H_INIT_CASES = [  # (lam, y0, tf, rtol, atol, h_min, h_max): the branch each one takes
    (-1.3, 1.0, 10.0, 1e-9, 1e-12, 0.0, np.inf),     # h1 = (0.01/der12)^(1/5)
    (100.0, 1.0, 10.0, 1e-2, 1e-2, 0.0, np.inf),     # 100 |h|, h = 0.01 sqrt(dny/dnf)
    (0.0, 1.0, 10.0, 1e-6, 1e-6, 0.0, np.inf),       # dnf = 0: h = 1e-6, h1 = 1e-3 |h|
    (-1.3, 1.0, 10.0, 1e-9, 1e-12, 0.0, 1e-3),       # h_max
    (-1.3, 1.0, 10.0, 1e-9, 1e-12, 0.02, np.inf),    # h_min
    (-1.3, 1.0, 2e-3, 1e-9, 1e-12, 0.0, np.inf),     # |tf - t0|
    (2.0, 3.0, -10.0, 1e-8, 1e-10, 0.0, np.inf),     # backward
]


@pytest.mark.parametrize("case", H_INIT_CASES)
def test_initial_step_closed_form(orc, case):
    lam, y0, tf, rtol, atol, h_min, h_max = case
    want = _h_init_exponential(lam, y0, 0.0, tf, rtol, atol, h_min, h_max)
    m = ExplicitRungeKutta.dopri5().rtol(rtol).atol(atol).h_min(h_min).h_max(h_max).max_steps(1)
    sol = orc.solve_ode(IVP.ode(Exponential(lam), 0.0, tf, [y0], n=1).method(m))
    # One attempt, accepted at t = h0, then MaxSteps (or Complete at tf).
    assert sol.accepted[0] == 1 and sol.rejected[0] == 0
    assert sol.status[0] in (DE_STATUS_MAX_STEPS, DE_STATUS_COMPLETE)
    assert abs(mpmath.mpf(float(sol.t[0])) - want) <= 1e-13 * abs(want), (float(want), sol.t[0])


# -------------------------------------------- norm and controller ----------
def _controller_times(system, w0, h0, rtol, atol, attempts):
    """Time after `attempts` attempts of the textbook DOPRI5 controller on a
    linear system; `system(z_coef)` maps a stability polynomial evaluated at
    h*A (A the system matrix) onto a state."""
    mp = mpmath.mpf
    t, w, h, rejected = mp(0), w0, mp(h0), False
    rtol, atol = mp(rtol), mp(atol)
    for _ in range(attempts):
        y_new = system(R_POLY, w, h)
        e = [a - b for a, b in zip(y_new, system(RHAT_POLY, w, h))]   # h * err_state
        sk = [atol + rtol * max(abs(a), abs(b)) for a, b in zip(w, y_new)]
        err = mpmath.sqrt(sum((ei / si) ** 2 for ei, si in zip(e, sk)) / len(w))
        scale = min(max(mp("0.9") * err ** (-mp(1) / 5), mp("0.2")), mp(10))
        if err <= 1:
            t, w = t + h, y_new
            if rejected:
                scale = min(scale, mp(1))
            rejected = False
        else:
            rejected = True
        h = h * scale
    return t


def _harmonic(poly, w, h):
    # y' = (y1, -y0): A acts on y0 + i y1 as multiplication by -i.
    c = _ev(poly, mpmath.mpc(0, -1) * h) * mpmath.mpc(w[0], w[1])
    return [c.real, c.imag]


def _exponential(lam):
    return lambda poly, w, h: [_ev(poly, lam * h) * w[0]]


CONTROLLER_CASES = [  # (system, y0, h0, rtol, atol, attempts, what it exercises)
    ("harmonic", [1.0, 0.0], 0.3, 1e-3, 1e-5, 3, "three accepted, unclamped scales"),
    ("harmonic", [1.0, 0.0], 1.0, 1e-4, 1e-6, 3, "rejected, then two accepted"),
    ("harmonic", [1.0, 0.0], 1e-3, 1e-6, 1e-8, 2, "max_scale clamp"),
    ("harmonic", [1.0, 0.0], 1.5, 1e-6, 1e-8, 3, "min_scale clamp, two rejections"),
    ("exp-1", [1.0], 3.0, 1e-2, 1e-2, 3, "rejected, accepted with scale > 1 -> min(scale, 1)"),
]


@pytest.mark.parametrize("case", CONTROLLER_CASES, ids=[c[6] for c in CONTROLLER_CASES])
def test_norm_and_controller_closed_form(orc, case):
    name, y0, h0, rtol, atol, attempts, _ = case
    if name == "harmonic":
        system, sysfn = Harmonic(1.0), _harmonic
    else:
        system, sysfn = Exponential(-1.0), _exponential(mpmath.mpf(-1))
    want = _controller_times(sysfn, [mpmath.mpf(v) for v in y0], h0, rtol, atol, attempts)
    m = ExplicitRungeKutta.dopri5().rtol(rtol).atol(atol).h0(h0).max_steps(attempts)
    sol = orc.solve_ode(IVP.ode(system, 0.0, 100.0, y0, n=1).method(m))
    assert sol.status[0] == DE_STATUS_MAX_STEPS
    assert sol.accepted[0] + sol.rejected[0] == attempts
    # The oracle's error estimate is a cancelling sum (about 1e-12 relative at
    # these step sizes); a wrong exponent, norm or clamp moves t by percents.
    assert abs(mpmath.mpf(float(sol.t[0])) - want) <= 1e-9 * want, (float(want), sol.t[0])


# ------------------------------------------------------------- stiffness ----
def _h_lamb(lam, h):
    z = lam * h
    r = _ev(R_POLY, z)
    return h * abs(lam) * abs((r - 1) / z - r) / (abs(r) * abs(lam - 1))


def _h_for(lam, target):
    return mpmath.findroot(lambda h: _h_lamb(lam, h) - target, mpmath.mpf("0.1"))


@pytest.mark.parametrize("side", ["above", "below"])
def test_stiffness_quotient_closed_form(orc, side):
    # Fixed step (h_min = h_max = h0, loose tolerances: every attempt is
    # accepted), so the check on every 100th attempt sees the same quotient.
    # Just above 6.1 it counts 15 times and stops at attempt 1500 with the
    # state of attempt 1499; just below the solve completes.  lam near 1 keeps
    # the quotient's squares finite (y grows by e^(lam h) per step).
    lam = mpmath.mpf("0.999")
    h = float(_h_for(lam, mpmath.mpf("6.1") * (mpmath.mpf("1.001") if side == "above"
                                                 else mpmath.mpf("0.999"))))
    tf = 1600.5 * h
    m = ExplicitRungeKutta.dopri5().rtol(1e3).atol(1e3).h0(h).h_min(h).h_max(h).max_steps(10 ** 6)
    sol = orc.solve_ode(IVP.ode(Exponential(float(lam)), 0.0, tf, [1.0], n=1).method(m))
    assert sol.rejected[0] == 0
    if side == "above":
        assert sol.status[0] == DE_STATUS_STIFFNESS
        assert sol.accepted[0] == 1499
        t = 0.0
        for _ in range(1499):
            t += h
        assert sol.t[0] == t
    else:
        assert sol.status[0] == DE_STATUS_COMPLETE
        assert sol.accepted[0] == 1601


# ------------------------------------------------------- planted mistakes ----
MUTANTS = {
    # name: (original text, planted mistake, pins expected to fail)
    "exponent_quarter": ("double error_exponent = 1.0 / 5.0;", "double error_exponent = 1.0 / 4.0;",
                         ["controller"]),
    "norm_without_n": ("err = fabs(h) * err * sqrt(1.0 / (deno * (double)n));",
                       "err = fabs(h) * err * sqrt(1.0 / deno);", ["controller"]),
    "no_min_after_reject": ("        scale = fmin(scale, 1.0);\n", "\n", ["controller"]),
    "h_init_without_sqrt": ("h = sqrt(dny / dnf) * 0.01;", "h = (dny / dnf) * 0.01;", ["h_init"]),
    "stiffness_hairer_quotient": ("double stden = diff_norm_squared(dydt, ysti, n);",
                                  "double stden = diff_norm_squared(dydt, k[6], n);", ["stiffness"]),
    "stiffness_period_50": ("if (steps % 100 == 0) {", "if (steps % 50 == 0) {", ["stiffness"]),
}


def _pins(orc, group):
    if group == "controller":
        for case in CONTROLLER_CASES:
            test_norm_and_controller_closed_form(orc, case)
    elif group == "h_init":
        for case in H_INIT_CASES:
            test_initial_step_closed_form(orc, case)
    else:
        for side in ("above", "below"):
            test_stiffness_quotient_closed_form(orc, side)


def _build_mutant(src_dir, out_dir, name, old, new):
    src = open(os.path.join(src_dir, "de_oracle.cpp")).read()
    assert src.count(old) >= 1, name
    path = os.path.join(out_dir, name + ".cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new, 1))
    lib = os.path.join(out_dir, name + ".so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-mfma", "-fPIC",
                           "-shared", "-pthread", "-I", src_dir, path, "-o", lib])
    return lib


def test_planted_mistakes_are_caught(orc):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    # The mutants include the shared headers by paths relative to oracle/ (-I).
    work = os.path.join(src_dir, "_build", "mutants")
    os.makedirs(work, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        libs = dict(zip(MUTANTS, ex.map(
            lambda kv: _build_mutant(src_dir, work, kv[0], kv[1][0], kv[1][1]), MUTANTS.items())))
    saved_lib, saved_path = orc._lib, orc.LIB_PATH
    try:
        for name, (_, _, groups) in MUTANTS.items():
            orc._lib, orc.LIB_PATH = None, libs[name]
            for group in groups:
                with pytest.raises(AssertionError):
                    _pins(orc, group)
    finally:
        orc._lib, orc.LIB_PATH = saved_lib, saved_path


def test_screened_stiffness_decision_equals_the_reference_expression(orc):
    # The kernels decide h sqrt(num / den) > 6.1 by comparing h^2 num with
    # 6.1^2 den when the two differ by more than 1e-12 (relative): the
    # decision must equal the rounded expression's everywhere, in particular
    # within a few ulps of the threshold and at the edges of the screen's range.
    rng = np.random.default_rng(11)
    n = 400000
    den = np.exp(rng.uniform(-300, 300, n))
    h = np.exp(rng.uniform(-20, 5, n))
    ratio = (6.1 / h) ** 2 * (1.0 + rng.normal(0, 1e-12, n) * rng.choice([1, 1e3, 1e-3], n))
    num = ratio * den
    num[:1000] = np.nextafter(num[:1000], np.inf)
    edges = np.array([float.fromhex(v) for v in ("0x1p-300", "0x1p300", "0x1.0000000000001p-300",
                                                 "0x1.fffffffffffffp299")])
    h[1000:1004], num[1004:1008], den[1008:1012] = edges, edges, edges
    h[1012:1020] = [-1.0, 0.0, np.inf, np.nan, 1e-310, 1.0, 1.0, 1.0]
    num[1017:1020] = [0.0, np.inf, np.nan]
    h, num, den = (np.ascontiguousarray(v, dtype=np.float64) for v in (h, num, den))
    agree = np.zeros(len(h), dtype=np.uint8)
    orc.load().de_oracle_stiff_check(h.ctypes.data, num.ctypes.data, den.ctypes.data,
                                     agree.ctypes.data, len(h))
    assert agree.all(), np.flatnonzero(agree == 0)[:10]
    # the sample really straddles the threshold
    lhs = h * np.sqrt(num / den)
    assert ((lhs > 6.1) & (lhs < 6.1 * (1 + 1e-11))).sum() > 1000
    assert ((lhs <= 6.1) & (lhs > 6.1 * (1 - 1e-11))).sum() > 1000
