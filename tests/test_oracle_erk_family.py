"""The oracle's embedded explicit Runge-Kutta family (rkf45, cash_karp,
rkv655e .. rkv989e; src/methods/erk/adaptive/ordinary.rs) pinned four ways:

  * the reference's own known-answer rows (tests/ode/accuracy.rs,
    tests/ode/errors.rs, tests/ode/interpolation.rs);
  * the Butcher order conditions over every rooted tree, evaluated on the
    doubles the oracle uses: b to the method's order, bh to one less and not
    more, the continuous extension to the interpolant order each constructor's
    documentation states (src/methods/erk/adaptive/mod.rs:43-122) and not more;
  * one step on y' = lam y against R(z) computed exactly from those doubles,
    and the evaluation count identity of ordinary.rs;
  * bit equality with an independent pure-Python restatement
    (tests/pyref_erk.py) on Lorenz, with and without dense output.
"""
import itertools
import math
from fractions import Fraction as Fr
from functools import lru_cache

import numpy as np
import pytest

import pyref_erk
from deamd import (DE_STATUS_BAD_INPUT, DE_STATUS_COMPLETE, DE_STATUS_INTERRUPTED,
                   DE_STATUS_STIFFNESS, CrossingDirection, Event, EventConfig, ExplicitRungeKutta,
                   Exponential, Harmonic, IVP, Linear, Logistic, Lorenz, VanDerPol)
from deamd import _abi

FAMILY = ["rkf45", "cash_karp", "rkv655e", "rkv656e", "rkv766e", "rkv767e", "rkv877e",
          "rkv878e", "rkv988e", "rkv989e"]
METHOD_ID = {name: getattr(_abi, "DE_METHOD_" + name.upper()) for name in FAMILY}
# (method order, interpolant order or None for cubic Hermite): mod.rs:43-122
ORDERS = {"rkf45": (5, None), "cash_karp": (5, None), "rkv655e": (6, 5), "rkv656e": (6, 6),
          "rkv766e": (7, 6), "rkv767e": (7, 7), "rkv877e": (8, 7), "rkv878e": (8, 8),
          "rkv988e": (9, 8), "rkv989e": (9, 9)}


def build(name):
    return getattr(ExplicitRungeKutta, name)()


# ---------------------------------------------------------- reference KATs --
# tests/ode/accuracy.rs: (system, t0, tf, y0, expected, {method: tolerance}).
# The "RKV65" rows run rkv655e; rkv656e has no row and gets rkv655e's.
_V = ["rkv655e", "rkv656e", "rkv766e", "rkv767e", "rkv877e", "rkv878e", "rkv988e", "rkv989e"]
ACCURACY = [
    ("exponential_growth_positive_direction", Exponential(1.0), 0.0, 10.0, [1.0],
     [22026.46579479], dict({"rkf45": 1e3, "cash_karp": 1e3}, **{v: 1e-1 for v in _V})),  # :70-176
    ("exponential_growth_negative_direction", Exponential(1.0), 0.0, -10.0, [22026.46579479],
     [1.0], dict({"rkf45": 1e-2, "cash_karp": 1e-2}, **{v: 1e-2 for v in _V})),  # :177-285
    ("linear_equation", Linear(1.0, 1.0), 0.0, 10.0, [1.0], [44051.93158958],
     dict({"rkf45": 1e3, "cash_karp": 1e3}, **{v: 1e1 for v in _V})),  # :286-394
    ("harmonic_oscillator", Harmonic(1.0), 0.0, 10.0, [1.0, 0.0], [-0.83907153, 0.54402111],
     dict({"rkf45": 1e-3, "cash_karp": 1e-3}, **{v: 1e-3 for v in _V})),  # :395-503
    ("logistic_equation", Logistic(1.0, 10.0), 0.0, 10.0, [0.1], [9.95525518],
     dict({"rkf45": 1e-2, "cash_karp": 1e-2}, **{v: 1e-3 for v in _V})),  # :504-612
]


@pytest.mark.parametrize("pow_mode", [0, 1])
@pytest.mark.parametrize("case", ACCURACY, ids=[c[0] for c in ACCURACY])
def test_accuracy_rows(orc, case, pow_mode):
    label, system, t0, tf, y0, expected, tols = case
    for name in FAMILY:
        ens = IVP.ode(system, t0, tf, y0, n=1).method(build(name))
        sol = orc.solve_ode(ens, pow_mode=pow_mode)
        assert sol.status[0] == DE_STATUS_COMPLETE, (label, name, sol.status)
        for c, e in enumerate(expected):
            assert abs(sol.y[c] - e) < tols[name], (label, name, c, sol.y[c], e)


@pytest.mark.parametrize("name", FAMILY)
def test_errors_rows(orc, name):
    # tests/ode/errors.rs:87-130: tf == t0, and |h0| > |tf - t0|
    for t0, tf, h0 in [(0.0, 0.0, 0.1), (0.0, 1.0, 10.0)]:
        ens = IVP.ode(Exponential(1.0), t0, tf, [1.0], n=1).method(build(name).h0(h0))
        sol = orc.solve_ode(ens)
        assert sol.status[0] == DE_STATUS_BAD_INPUT
        assert sol.accepted[0] == 0 and sol.evals[0] == 0


@pytest.mark.parametrize("name", FAMILY)
def test_interpolation_rows(orc, name):
    # tests/ode/interpolation.rs:60-73: t_eval [0.5, 1.0, 1.69] on y' = y over
    # [0, 2] within 1e-3 (the reference checks the first point; all are here).
    ens = IVP.ode(Exponential(1.0), 0.0, 2.0, [1.0], n=1).t_eval([0.5, 1.0, 1.69]) \
        .method(build(name))
    sol = orc.solve_ode(ens)
    assert sol.n_dense[0] == 3
    for got, t in zip(sol.dense[:3], [0.5, 1.0, 1.69]):
        assert abs(got - math.exp(t)) < 1e-3, (name, t, got)


# ------------------------------------------------------- order conditions --
@lru_cache(None)
def rooted_trees(n):
    """Rooted trees of n nodes, each a sorted tuple of its subtrees."""
    if n == 1:
        return ((),)

    def partitions(rem, largest):
        if rem == 0:
            yield ()
            return
        for p in range(min(rem, largest), 0, -1):
            for rest in partitions(rem - p, p):
                yield (p,) + rest

    out = set()
    for sizes in partitions(n - 1, n - 1):
        for kids in itertools.product(*[rooted_trees(p) for p in sizes]):
            out.add(tuple(sorted(kids)))
    return tuple(sorted(out))


def _gamma(t):
    return (1 + sum(_size(c) for c in t)) * math.prod(_gamma(c) for c in t)


def _size(t):
    return 1 + sum(_size(c) for c in t)


def _stage_weights(t, A):
    v = np.ones(A.shape[0])
    for c in t:
        v = v * (A @ _stage_weights(c, A))
    return v


def worst_residual(weights, A, n, theta=1.0):
    """max over trees of n nodes of |w . Phi(t) - theta^n / gamma(t)|."""
    return max(abs(weights @ _stage_weights(t, A) - theta ** n / _gamma(t))
               for t in rooted_trees(n))


def test_rooted_tree_counts():
    # OEIS A000081: 1, 1, 2, 4, 9, 20, 48, 115, 286, 719
    assert [len(rooted_trees(n)) for n in range(1, 11)] == [1, 1, 2, 4, 9, 20, 48, 115, 286, 719]


@pytest.mark.parametrize("name", FAMILY)
def test_tableau_order_conditions(orc, name):
    # The coefficient literals meet the conditions to ~1e-9 at worst (rkv877e's
    # are printed to about ten digits); a misread literal or index misses by far
    # more, and the first unmet order misses by > 1e-6.
    tab = orc.erk_tableau(METHOD_ID[name])
    p, q = ORDERS[name]
    S, I, O = tab["S"], tab["I"], tab["O"]
    assert O == p and tab["has_bi"] == (q is not None)
    assert tab["fsal"] == (name in ("rkv655e", "rkv656e"))
    assert np.abs(tab["A"].sum(1) - tab["C"]).max() < 1e-11          # row sums
    assert not np.triu(tab["A"]).any()                                  # explicit
    A = tab["A"][:S, :S]
    for n in range(1, p + 1):
        assert worst_residual(tab["B"], A, n) < 1e-8, ("b", n)
    for n in range(1, p):
        assert worst_residual(tab["BH"], A, n) < 1e-8, ("bh", n)
    assert worst_residual(tab["BH"], A, p) > 1e-6
    if tab["fsal"]:  # the last stage is the new point: a row equal to b, c = 1
        assert tab["C"][S - 1] == 1.0 and list(tab["A"][S - 1][:S]) == list(tab["B"])
    if q is None:
        assert not tab["BI"].any()
        return
    for theta in (0.3, 0.7):
        w = np.array([sum(tab["BI"][i][j] * theta ** (j + 1) for j in range(O))
                      for i in range(I)])
        for n in range(1, q + 1):
            assert worst_residual(w, tab["A"], n, theta) < 1e-8, ("dense", theta, n)
        assert worst_residual(w, tab["A"], q + 1, theta) > 1e-7, ("dense", theta, q + 1)


# ----------------------------------------------------- one step, counts ----
def _R(tab, z, mag=abs):
    """R(z) = 1 + z b^T (I - zA)^{-1} 1, exactly, from the tableau's doubles;
    with mag=abs on every coefficient and z, the sum of the magnitudes of the
    terms, which bounds the rounding of the floating-point step."""
    S = tab["S"]
    f = (lambda v: Fr(float(v))) if mag is None else (lambda v: abs(Fr(float(v))))
    A = [[f(x) for x in row[:S]] for row in tab["A"][:S]]
    x = [Fr(0)] * S
    for i in range(S):
        x[i] = Fr(1) + z * sum(A[i][j] * x[j] for j in range(i))
    return Fr(1) + z * sum(f(tab["B"][i]) * x[i] for i in range(S))


@pytest.mark.parametrize("name", FAMILY)
def test_single_step_is_R_of_z(orc, name):
    tab = orc.erk_tableau(METHOD_ID[name])
    for lam, h in [(1.0, 0.1), (-2.0, 0.05)]:
        ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1) \
            .method(build(name).h0(h).rtol(1.0).atol(1.0))
        sol = orc.solve_ode(ens)
        assert sol.accepted[0] == 1 and sol.rejected[0] == 0
        z = Fr(lam) * Fr(h)
        R = float(_R(tab, z, mag=None))
        # each stage sum rounds at most S times, each term to its own magnitude
        bound = 2.0 * tab["S"] * 2.0 ** -52 * float(_R(tab, abs(z)))
        assert abs(sol.y[0] - R) <= bound, (name, lam, sol.y[0] - R, bound)
        assert abs(sol.y[0] - math.exp(lam * h)) < 1e-9


@pytest.mark.parametrize("name", FAMILY)
def test_evaluation_count_identity(orc, name):
    # ordinary.rs: 5 evaluations before the first step (h0 computed), S - 1 per
    # attempt, and per accepted step one more unless FSAL plus I - S dense stages.
    tab = orc.erk_tableau(METHOD_ID[name])
    p, y0 = __import__("deamd").gen_lorenz(16, seed=5)
    ens = IVP.ode(Lorenz(*p), 0.0, 3.0, y0).method(build(name).rtol(1e-9).atol(1e-11)
                                                   .max_steps(100000))
    sol = orc.solve_ode(ens)
    assert (sol.status == DE_STATUS_COMPLETE).all()
    assert sol.rejected.sum() > 0
    per_acc = (0 if tab["fsal"] else 1) + (tab["I"] - tab["S"] if tab["has_bi"] else 0)
    want = 5 + (tab["S"] - 1) * (sol.accepted + sol.rejected) + per_acc * sol.accepted
    assert np.array_equal(sol.evals, want)


# --------------------------------------------- Python restatement, bits ----
def _lorenz_rhs(rho):
    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z
    return f


@pytest.mark.parametrize("name", FAMILY)
def test_oracle_bit_equals_python_restatement(orc, name):
    tab = orc.erk_tableau(METHOD_ID[name])
    rng = np.random.default_rng(300 + METHOD_ID[name])
    rho = float(rng.uniform(0, 50))
    y0 = list(1.0 + 0.1 * rng.standard_normal(3))
    pts = sorted(float(x) for x in rng.uniform(0.0, 1.2, 9)) + [1.2]
    rtol, atol = [1e-8, 1e-6, 1e-9], 1e-10
    want = pyref_erk.solve_erk(_lorenz_rhs(rho), 0.0, 1.2, y0, rtol, atol, tab, t_eval=pts)
    ens = IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, 1.2, np.array(y0).reshape(3, 1)) \
        .t_eval(pts).method(build(name).rtol(rtol).atol(atol).max_steps(100000))
    got = orc.solve_ode(ens, pow_mode=1)
    assert want[3] > 0  # the restatement exercises rejections
    assert list(got.y) == want[0] and got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == tuple(want[2:6])
    assert got.n_dense[0] == len(want[6]) == len(pts)
    assert got.dense.reshape(len(pts), 3).tolist() == want[6]


@pytest.mark.parametrize("name,h0", [("rkf45", 0.5), ("rkv655e", 0.5), ("rkv988e", 0.2)])
def test_max_rejects_gives_stiffness(orc, name, h0):
    # ordinary.rs:182-197: max_rejects consecutive rejections end the solve
    # with Error::Stiffness; the failing step's evaluations are not counted.
    tab = orc.erk_tableau(METHOD_ID[name])
    f = lambda t, y, dy: dy.__setitem__(slice(None), [y[1], 25.0 * (1 - y[0] * y[0]) * y[1] - y[0]])
    for max_rejects in (0, 1, 2, 3):
        want = pyref_erk.solve_erk(f, 0.0, 8.0, [2.0, 0.0], 1e-10, 1e-10, tab, h0=h0,
                                   max_rejects=max_rejects)
        ens = IVP.ode(VanDerPol(25.0), 0.0, 8.0, [[2.0], [0.0]]) \
            .method(build(name).h0(h0).rtol(1e-10).atol(1e-10).max_rejects(max_rejects))
        got = orc.solve_ode(ens, pow_mode=1)
        if max_rejects <= 2:
            assert want[5] == DE_STATUS_STIFFNESS and want[3] == max(max_rejects - 1, 0)
        assert got.status[0] == want[5], max_rejects
        assert list(got.y) == want[0] and got.t[0] == want[1]
        assert (got.accepted[0], got.rejected[0], got.evals[0]) == tuple(want[2:5])


def test_nan_error_is_accepted(orc):
    # State::error_norm_inf folds with max(), which ignores NaN: a step whose
    # stages overflow to NaN has error norm 0 and is accepted (then scale = 10).
    # rkv988e from h0 = 0.5 on a stiff Van der Pol does exactly that.
    tab = orc.erk_tableau(METHOD_ID["rkv988e"])
    f = lambda t, y, dy: dy.__setitem__(slice(None), [y[1], 25.0 * (1 - y[0] * y[0]) * y[1] - y[0]])
    want = pyref_erk.solve_erk(f, 0.0, 8.0, [2.0, 0.0], 1e-10, 1e-10, tab, h0=0.5)
    ens = IVP.ode(VanDerPol(25.0), 0.0, 8.0, [[2.0], [0.0]]) \
        .method(build("rkv988e").h0(0.5).rtol(1e-10).atol(1e-10))
    got = orc.solve_ode(ens, pow_mode=1)
    assert got.status[0] == want[5] == DE_STATUS_COMPLETE
    assert np.isnan(got.y).all() and all(v != v for v in want[0])
    assert (got.accepted[0], got.rejected[0], got.evals[0]) == tuple(want[2:5]) == (3, 0, 64)


# --- event solouts over the family's dense output ---------------------------

COS_ZEROS = math.pi / 2 + math.pi * np.arange(6)


# Tableaux whose interpolant carries rounding noise above the crossing
# solouts' Newton acceptance (|offset| < 100 eps, crossing.rs:268-318) near
# some zeros of cos t: large bi coefficients cancel, and samples 5e-5 apart
# scatter by ~1e-12 (rkv656e) and more for the order 8/9 pairs on steps of
# ~0.33.  There find_crossing_newton returns None and the solout pushes the
# chord's zero (crossing.rs:225-231).
NEWTON_FALLS_BACK = {"rkv656e", "rkv877e", "rkv878e", "rkv988e"}


def _cosine(name):
    return IVP.ode(Harmonic(1.0), 0.0, 20.0, np.array([[1.0], [0.0]])).method(
        getattr(ExplicitRungeKutta, name)().rtol(1e-10).atol(1e-12).max_steps(100000))


@pytest.mark.parametrize("name", FAMILY)
def test_events_find_the_zeros_of_cosine(orc, name):
    # y'' = -y, y(0) = (1, 0): y0 = cos t, six zeros in [0, 20].  The event
    # solout (Brent on the interpolant) finds each one through the tableau's
    # interpolant (Horner over the bi rows, or the cubic Hermite of rkf45 /
    # cash_karp); Negative keeps the falling zeros (pi/2, 5pi/2, ...),
    # Positive the rising ones.  The crossing and hyperplane solouts push
    # Newton's root where it converges, else the chord's zero.
    fell_back = False
    for d in (CrossingDirection.Both, CrossingDirection.Negative, CrossingDirection.Positive):
        want = COS_ZEROS if d == CrossingDirection.Both else (
            COS_ZEROS[0::2] if d == CrossingDirection.Negative else COS_ZEROS[1::2])
        sol = orc.solve_ode(_cosine(name).event(Event.component(0, 0.0, EventConfig(d))))
        assert sol.status[0] == DE_STATUS_COMPLETE
        t, y = sol.events(0)
        assert len(t) == len(want)
        assert np.abs(t - want).max() < 1e-9
        assert np.abs(y[:, 0]).max() < 1e-9
        assert np.abs(np.abs(y[:, 1]) - 1.0).max() < 1e-9
        for ens in (_cosine(name).crossing(0, 0.0, d),
                    _cosine(name).hyperplane_crossing([0.0, 5.0], [2.0, 0.0], d)):
            sol = orc.solve_ode(ens)
            assert sol.status[0] == DE_STATUS_COMPLETE
            t2, y2 = sol.events(0)
            assert len(t2) == len(want)
            newton = np.abs(y2[:, 0]) < 100 * 2.0 ** -52 * 10
            assert np.abs(t2 - want)[newton].max(initial=0.0) < 1e-9
            assert np.abs(t2 - want)[~newton].max(initial=0.0) < 1e-2
            fell_back |= bool((~newton).any())
    assert fell_back == (name in NEWTON_FALLS_BACK)


@pytest.mark.parametrize("name", FAMILY)
def test_terminal_time_event_like_the_reference_test(orc, name):
    # tests/ode/errors.rs: y' = y with g = t - 10 and a terminal event: g is
    # linear in t, so the root is 10 up to rounding, and the solve stops with
    # Interrupted at the end of the step that crossed it.
    m = getattr(ExplicitRungeKutta, name)().rtol(1e-10).atol(1e-10).max_steps(100000)
    ens = IVP.ode(Exponential(1.0), 0.0, 20.0, [1.0], n=1).method(m) \
        .event(Event.time(10.0, EventConfig().terminal()))
    sol = orc.solve_ode(ens)
    assert sol.status[0] == DE_STATUS_INTERRUPTED and sol.n_events[0] == 1
    t, y = sol.events(0)
    assert abs(t[0] - 10.0) <= 1e-11 and sol.t[0] >= 10.0
    assert abs(y[0, 0] / math.exp(10.0) - 1.0) < 1e-6
    free = orc.solve_ode(IVP.ode(Exponential(1.0), 0.0, 20.0, [1.0], n=1).method(m))
    assert free.status[0] == DE_STATUS_COMPLETE and free.accepted[0] > sol.accepted[0]
    # the steps before the event are the free solve's steps
    assert sol.evals[0] < free.evals[0]


# Bit equality of the oracle's family event paths with the Python restatements
# (pyref_erk's step and interpolant, pyref_events' solouts): every tableau,
# every solout kind, the three directions taken in turn.
_EV_KINDS = ("event", "event_teval", "event_plane_term", "crossing", "hyperplane")
_EV_DIRS = [CrossingDirection.Both, CrossingDirection.Positive, CrossingDirection.Negative]
_EV_DNAME = {CrossingDirection.Both: "both", CrossingDirection.Positive: "positive",
             CrossingDirection.Negative: "negative"}
EV_CASES = [(name, kind, _EV_DIRS[(a + b) % 3]) for a, name in enumerate(FAMILY)
            for b, kind in enumerate(_EV_KINDS)]


@pytest.mark.parametrize("name,kind,d", EV_CASES)
def test_oracle_events_bit_equal_python_restatement(orc, name, kind, d):
    import pyref_events as PE
    tab = orc.erk_tableau(METHOD_ID[name])
    rng = np.random.default_rng(700 + 10 * METHOD_ID[name] + _EV_KINDS.index(kind))
    lorenz = kind in ("event", "event_plane_term", "hyperplane")
    if lorenz:
        rho = float(rng.uniform(20, 40))
        y0 = list(1.0 + 0.1 * rng.standard_normal(3))
        f, sysm, tf = _lorenz_rhs(rho), Lorenz(10.0, rho, 8.0 / 3.0), 3.0
    else:
        k = float(rng.uniform(0.5, 2.0))
        y0 = [1.0, float(rng.uniform(-0.5, 0.5))]
        f = lambda t, y, dy, k=k: dy.__setitem__(slice(None), [y[1], -k * y[0]])
        sysm, tf = Harmonic(k), 12.0
    rtol, atol = 1e-9, 1e-11
    ens = IVP.ode(sysm, 0.0, tf, np.array(y0).reshape(-1, 1)) \
        .method(build(name).rtol(rtol).atol(atol).max_steps(100000)).options(max_events=64)
    sol = PE.Solution()
    dname = _EV_DNAME[d]
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
    want = pyref_erk.solve_erk(f, 0.0, tf, y0, rtol, atol, tab, max_steps=100000, solout=so)
    got = orc.solve_ode(ens, pow_mode=1)
    pts_want = so.root_points if kind.startswith("event") else sol.points
    t, y = got.events(0)
    assert len(pts_want) > 0
    assert got.n_events[0] == len(pts_want)
    assert t.tolist() == [p_[0] for p_ in pts_want]
    assert y.tolist() == [p_[1] for p_ in pts_want]
    assert list(got.y) == want[0] and got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == tuple(want[2:6])
