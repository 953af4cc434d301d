"""Pins of the CPU oracle by mathematics and independent restatement.

Each check is chosen so that a plausible mistake in the oracle (a wrong
coefficient, a dropped term, a transposed index, a wrong operation order)
fails it:
  * DOPRI5 tableau: classical order conditions (order 5 for b, order 4 for the
    embedded weights b - er) in exact rationals;
  * one step on y' = lam*y equals R(z) y0 with R the DOPRI5 stability
    polynomial 1 + z + ... + z^5/120 + z^6/600 (textbook);
  * RK4 / Euler steps equal the Taylor polynomials of degree 4 / 1;
  * dense output on a single DOPRI5 step is the cubic Hermite interpolant of
    the step's end values and derivatives (the reference's bi[4] row is zero);
  * evaluation-count invariants of the driver;
  * bit equality with an independent pure-Python restatement (tests/pyref_dopri5.py)
    on Lorenz trajectories, using the C library pow as the reference does;
  * SDE: closed-form recurrences for linear SDEs, Milstein == Euler-Maruyama
    for additive noise, the clipped last step of [0, 1] at dt = 1e-3;
  * Philox4x32-10 known-answer vectors (Random123), normal-transform moments;
  * pow: equal to the host C library's pow (glibc) bit for bit, under 0.52 ulp
    against mpmath, and the IEEE special values.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

import pyref_dopri5
from deamd import (DE_NOISE_PHILOX, DE_NOISE_TABLE, DE_STATUS_COMPLETE, EulerMaruyama,
                   ExplicitRungeKutta, Exponential, IVP, Lorenz, LinearDiagonal, Milstein,
                   OrnsteinUhlenbeck)

# DOPRI5 rationals (Dormand & Prince 1980; src/tableau/dorman_prince.rs:44-101)
DP_C = [Fr(0), Fr(1, 5), Fr(3, 10), Fr(4, 5), Fr(8, 9), Fr(1), Fr(1)]
DP_A = [[Fr(0)] * 7 for _ in range(7)]
DP_A[1][0] = Fr(1, 5)
DP_A[2][:2] = [Fr(3, 40), Fr(9, 40)]
DP_A[3][:3] = [Fr(44, 45), Fr(-56, 15), Fr(32, 9)]
DP_A[4][:4] = [Fr(19372, 6561), Fr(-25360, 2187), Fr(64448, 6561), Fr(-212, 729)]
DP_A[5][:5] = [Fr(9017, 3168), Fr(-355, 33), Fr(46732, 5247), Fr(49, 176), Fr(-5103, 18656)]
DP_A[6][:6] = [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84)]
DP_B = [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), Fr(0)]
DP_BHAT = [Fr(5179, 57600), Fr(0), Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200),
           Fr(187, 2100), Fr(1, 40)]


def order_conditions(b, A, c, order):
    """Butcher order conditions up to `order` (all rooted trees up to order 5)."""
    s = len(b)
    Ac = [sum(A[i][j] * c[j] for j in range(s)) for i in range(s)]
    Ac2 = [sum(A[i][j] * c[j] ** 2 for j in range(s)) for i in range(s)]
    AAc = [sum(A[i][j] * Ac[j] for j in range(s)) for i in range(s)]
    Ac3 = [sum(A[i][j] * c[j] ** 3 for j in range(s)) for i in range(s)]
    cAc = [c[i] * Ac[i] for i in range(s)]
    AcAc = [sum(A[i][j] * cAc[j] for j in range(s)) for i in range(s)]
    AAc2 = [sum(A[i][j] * Ac2[j] for j in range(s)) for i in range(s)]
    AAAc = [sum(A[i][j] * AAc[j] for j in range(s)) for i in range(s)]
    dot = lambda u, v: sum(x * y for x, y in zip(u, v))
    conds = [
        (1, dot(b, [1] * s), Fr(1)),
        (2, dot(b, c), Fr(1, 2)),
        (3, dot(b, [x ** 2 for x in c]), Fr(1, 3)),
        (3, dot(b, Ac), Fr(1, 6)),
        (4, dot(b, [x ** 3 for x in c]), Fr(1, 4)),
        (4, dot(b, [c[i] * Ac[i] for i in range(s)]), Fr(1, 8)),
        (4, dot(b, Ac2), Fr(1, 12)),
        (4, dot(b, AAc), Fr(1, 24)),
        (5, dot(b, [x ** 4 for x in c]), Fr(1, 5)),
        (5, dot(b, [c[i] ** 2 * Ac[i] for i in range(s)]), Fr(1, 10)),
        (5, dot(b, [c[i] * Ac2[i] for i in range(s)]), Fr(1, 15)),
        (5, dot(b, [c[i] * AAc[i] for i in range(s)]), Fr(1, 30)),
        (5, dot(b, [Ac[i] ** 2 for i in range(s)]), Fr(1, 20)),
        (5, dot(b, Ac3), Fr(1, 20)),
        (5, dot(b, AcAc), Fr(1, 40)),
        (5, dot(b, AAc2), Fr(1, 60)),
        (5, dot(b, AAAc), Fr(1, 120)),
    ]
    return [(o, got, want) for (o, got, want) in conds if o <= order]


def test_dopri5_rationals_satisfy_order_conditions():
    for o, got, want in order_conditions(DP_B, DP_A, DP_C, 5):
        assert got == want, (o, got, want)
    for o, got, want in order_conditions(DP_BHAT, DP_A, DP_C, 4):
        assert got == want, (o, got, want)
    # row sums = c
    for i in range(7):
        assert sum(DP_A[i]) == DP_C[i]


def test_oracle_tableau_is_dopri5(orc):
    c, a, b, er, bi0 = orc.dopri5_tableau()
    for i in range(7):
        assert c[i] == float(DP_C[i])
        assert b[i] == float(DP_B[i])
        assert er[i] == float(DP_B[i] - DP_BHAT[i])
        for j in range(7):
            assert a[i, j] == float(DP_A[i][j])


def stability_poly(b, A):
    """R(z) = 1 + z b^T (I - zA)^{-1} 1 as exact coefficients (explicit A)."""
    s = len(b)
    coeffs = [Fr(1)]
    # b^T A^k 1 for k = 0..s-1
    v = [Fr(1)] * s
    for k in range(s):
        coeffs.append(sum(b[i] * v[i] for i in range(s)))
        v = [sum(A[i][j] * v[j] for j in range(s)) for i in range(s)]
    return coeffs


def test_dopri5_stability_polynomial_is_textbook():
    R = stability_poly(DP_B, DP_A)
    assert R == [Fr(1), Fr(1), Fr(1, 2), Fr(1, 6), Fr(1, 24), Fr(1, 120), Fr(1, 600), Fr(0)]


@pytest.mark.parametrize("lam,h", [(1.0, 0.1), (-2.0, 0.05), (0.5, 0.3)])
def test_dopri5_single_step_is_R_of_z(orc, lam, h):
    # tf = t0 + h with h0 = h: exactly one attempt; loose tolerances accept it.
    ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1) \
        .method(ExplicitRungeKutta.dopri5().h0(h).rtol(1.0).atol(1.0))
    sol = orc.solve_ode(ens)
    assert sol.accepted[0] == 1 and sol.rejected[0] == 0
    z = Fr(lam) * Fr(h)
    R = sum(c * z ** k for k, c in enumerate(stability_poly(DP_B, DP_A)))
    assert abs(sol.y[0] - float(R)) <= 4e-16 * abs(float(R))


@pytest.mark.parametrize("method,taylor", [
    (lambda h: ExplicitRungeKutta.rk4(h), 4), (lambda h: ExplicitRungeKutta.euler(h), 1)])
def test_fixed_single_step_is_taylor(orc, method, taylor):
    lam, h = -1.5, 0.1
    ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1).method(method(h))
    sol = orc.solve_ode(ens)
    assert sol.accepted[0] == 1
    z = Fr(lam) * Fr(h)
    R = sum(z ** k / math.factorial(k) for k in range(taylor + 1))
    assert abs(sol.y[0] - float(R)) <= 4e-16


def test_dopri5_dense_output_is_cubic_hermite(orc):
    lam, h = 0.7, 0.2
    pts = [0.0, 0.05, 0.1, 0.13, 0.2]
    ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1).t_eval(pts) \
        .method(ExplicitRungeKutta.dopri5().h0(h).rtol(1.0).atol(1.0))
    sol = orc.solve_ode(ens)
    assert sol.accepted[0] == 1 and sol.n_dense[0] == len(pts)
    y0 = Fr(1)
    y1 = Fr(sol.y[0])
    f0, f1 = Fr(lam) * y0, Fr(lam) * y1
    H = Fr(h)
    for e, tp in enumerate(pts):
        s = Fr(tp) / H
        herm = ((2 * s ** 3 - 3 * s ** 2 + 1) * y0 + (s ** 3 - 2 * s ** 2 + s) * H * f0 +
                (-2 * s ** 3 + 3 * s ** 2) * y1 + (s ** 3 - s ** 2) * H * f1)
        assert abs(sol.dense[e] - float(herm)) <= 1e-15, (tp, sol.dense[e], float(herm))


def lorenz_ensemble(n, tf, rtol=1e-8, atol=1e-10, seed=7):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0, 50, n)
    y0 = 1.0 + 0.1 * rng.standard_normal((3, n))
    return IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, tf, y0) \
        .method(ExplicitRungeKutta.dopri5().rtol(rtol).atol(atol).max_steps(100000))


def test_evaluation_count_invariants(orc):
    sol = orc.solve_ode(lorenz_ensemble(64, 2.0))
    assert (sol.status == DE_STATUS_COMPLETE).all()
    assert (sol.evals == 3 + 6 * (sol.accepted + sol.rejected) + sol.accepted).all()
    assert (np.abs(sol.t - 2.0) <= 10 * 2.220446049250313e-16).all()
    ens = lorenz_ensemble(16, 1.0).method(ExplicitRungeKutta.rk4(0.01))
    sol = orc.solve_ode(ens)
    assert (sol.evals == 1 + 4 * sol.accepted).all()
    ens = lorenz_ensemble(16, 1.0).method(ExplicitRungeKutta.euler(0.01))
    sol = orc.solve_ode(ens)
    assert (sol.evals == 1 + sol.accepted).all()


@pytest.mark.parametrize("traj", range(6))
def test_oracle_bit_equals_python_restatement(orc, traj):
    rng = np.random.default_rng(100 + traj)
    rho = float(rng.uniform(0, 50))
    y0 = list(1.0 + 0.1 * rng.standard_normal(3))
    tf = 1.5

    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z

    want = pyref_dopri5.solve(f, 0.0, tf, y0, 1e-8, 1e-10)
    ens = IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, tf, np.array(y0).reshape(3, 1)) \
        .method(ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(100000))
    got = orc.solve_ode(ens, pow_mode=1)
    assert list(got.y) == want[0]
    assert got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == want[2:]


@pytest.mark.parametrize("traj", range(3))
def test_oracle_per_component_tolerance_bit_equals_python_restatement(orc, traj):
    # Tolerance::Vector: atol[i] + rtol[i]*max(|y_i|, |y_new_i|) in the norm and
    # atol[i] + rtol[i]*|y0_i| in h_init (src/traits.rs:400, src/methods/h_init.rs:80).
    rng = np.random.default_rng(200 + traj)
    rho = float(rng.uniform(0, 50))
    y0 = list(1.0 + 0.1 * rng.standard_normal(3))
    rtol = [1e-8, 1e-5, 1e-9]
    atol = [1e-10, 1e-7, 1e-12]

    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z

    want = pyref_dopri5.solve(f, 0.0, 1.5, y0, rtol, atol)
    ens = IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, 1.5, np.array(y0).reshape(3, 1)) \
        .method(ExplicitRungeKutta.dopri5().rtol(rtol).atol(atol).max_steps(100000))
    got = orc.solve_ode(ens, pow_mode=1)
    assert list(got.y) == want[0]
    assert got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == want[2:]
    # A tighter component must cost steps: the scalar loosest setting takes fewer.
    loose = pyref_dopri5.solve(f, 0.0, 1.5, y0, 1e-5, 1e-7)
    assert loose[2] < want[2]


def test_oracle_vector_tolerance_of_equal_entries_is_the_scalar(orc):
    import deamd
    p, y0 = deamd.gen_lorenz(64, seed=3)
    mk = lambda r, a: IVP.ode(Lorenz(*p), 0.0, 5.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(r).atol(a).max_steps(100000))
    a = orc.solve_ode(mk(1e-7, 1e-9))
    b = orc.solve_ode(mk([1e-7] * 3, [1e-9] * 3))
    assert np.array_equal(a.y, b.y) and np.array_equal(a.accepted, b.accepted)
    assert np.array_equal(a.rejected, b.rejected) and np.array_equal(a.t, b.t)


def test_sde_last_step_is_clipped(orc):
    # [0, 1] at dt = 1e-3: 1000 steps, the last one h = 1 - t_999.
    t = 0.0
    for _ in range(999):
        t += 1e-3
    h_last = 1.0 - t
    assert h_last == 0.0009999999999992237
    sys_ = LinearDiagonal(1, [0.0, 0.0])
    ens = IVP.sde(sys_, 0.0, 1.0, [1.0], noise=DE_NOISE_PHILOX, seed=1) \
        .method(EulerMaruyama(1e-3)).record_every(1)
    sol = orc.solve_sde(ens)
    assert sol.steps[0] == 1000 and sol.t[0] == 1.0
    assert sol.path_t[998] == t and sol.path_t[999] == 1.0
    assert sol.evals[0] == 2 + 2 * 1000
    sol = orc.solve_sde(ens.method(Milstein(1e-3)))
    assert sol.evals[0] == 1 + 3 * 1000


def test_milstein_equals_euler_maruyama_for_additive_noise(orc):
    n, d = 64, 8
    rng = np.random.default_rng(3)
    p = np.stack([rng.uniform(0.5, 2, (d, n)), rng.standard_normal((d, n)),
                  rng.uniform(0.1, 1, (d, n))], axis=1).reshape(3 * d, n)
    sys_ = OrnsteinUhlenbeck(d, p)
    y0 = np.zeros(d * n)
    em = orc.solve_sde(IVP.sde(sys_, 0.0, 0.5, y0, seed=11).method(EulerMaruyama(1e-2)))
    mil = orc.solve_sde(IVP.sde(sys_, 0.0, 0.5, y0, seed=11).method(Milstein(1e-2)))
    assert np.array_equal(em.y, mil.y)


def test_gbm_euler_maruyama_closed_form(orc):
    # y_{n+1} = y_n (1 + mu h + sigma dW_n) in exact arithmetic
    mu, sigma, h = 0.05, 0.2, 0.01
    dw = np.random.default_rng(5).standard_normal((50, 1)) * math.sqrt(h)
    sys_ = LinearDiagonal(1, [mu, sigma])
    ens = IVP.sde(sys_, 0.0, 0.5, [100.0], noise=DE_NOISE_TABLE, noise_table=dw) \
        .method(EulerMaruyama(h))
    sol = orc.solve_sde(ens)
    assert sol.steps[0] == 50
    y = Fr(100)
    for k in range(50):
        y = y * (1 + Fr(mu) * Fr(h) + Fr(sigma) * Fr(dw[k, 0]))
    assert abs(sol.y[0] - float(y)) <= 1e-13 * float(y)


KAT_PHILOX = [
    ([0, 0, 0, 0], 0, 0, [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, 0xffffffff, 0xffffffff, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], 0xa4093822, 0x299f31d0,
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,k0,k1,want", KAT_PHILOX)
def test_philox_known_answers(orc, ctr, k0, k1, want):
    assert orc.philox(ctr, k0, k1).tolist() == want


def _host_pow_is_glibc_fma():
    """glibc >= 2.28 on x86-64 with FMA runs the fused build of the table-driven
    pow that de::libm_pow restates; other C libraries round differently."""
    import platform
    lib, ver = platform.libc_ver()
    if platform.machine() != "x86_64" or lib != "glibc":
        return False
    if tuple(int(v) for v in ver.split(".")[:2]) < (2, 28):
        return False
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        return False
    return " fma" in flags and " avx2" in flags


def _pow_arguments(rng, n):
    """Arguments over every regime of pow: the controllers' exponents on any
    positive x, subnormal x, x near 1, results near overflow and underflow,
    negative x with integer y, and raw bit patterns."""
    bits = lambda: rng.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) \
        + rng.integers(0, 2, n, dtype=np.int64).astype(np.uint64)
    xs, ys = [], []
    ctrl = np.array([-0.2, 0.2, -1.0 / 6.0, -1.0 / 7.0, -0.125, -1.0 / 9.0, 0.125])
    xs.append(np.exp(rng.uniform(np.log(1e-300), np.log(1e300), n)))
    ys.append(rng.choice(ctrl, n))
    xs.append(1.0 + rng.integers(-2 ** 40, 2 ** 40, n) * 2.0 ** -60)
    ys.append(rng.choice(ctrl, n))
    xs.append(rng.integers(1, 2 ** 52, n).astype(np.float64) * 2.0 ** -1074)
    ys.append(rng.uniform(-3, 3, n))
    xs.append(np.exp(rng.uniform(-700, 700, n)))
    ys.append(rng.uniform(-4, 4, n) * np.where(rng.integers(0, 2, n) == 1, 1.0, 200.0))
    xs.append(-np.exp(rng.uniform(-20, 20, n)))
    ys.append(rng.integers(-60, 60, n).astype(np.float64))
    xs.append(bits().view(np.float64))
    ys.append(bits().view(np.float64))
    return np.concatenate(xs), np.concatenate(ys)


@pytest.mark.skipif(not _host_pow_is_glibc_fma(), reason="host pow is not glibc's fused x86-64 build")
def test_pow_equals_host_c_library(orc):
    # de::libm_pow (run by the GPU kernels and oracle mode 0) against the
    # C library pow that a Rust build of the reference calls on this host.
    rng = np.random.default_rng(2026)
    x, y = _pow_arguments(rng, 400000)
    special = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 3.0, math.inf, -math.inf,
                        math.nan, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                        1e-300, 1e300, 0.2, -0.2, 2.0 ** -66, 2.0 ** 63, -(2.0 ** 63),
                        1.0000000000000002, 0.9999999999999999])
    sx, sy = np.meshgrid(special, special)
    x = np.concatenate([x, sx.ravel()])
    y = np.concatenate([y, sy.ravel()])
    got = orc.pow_many(x, y)
    want = orc.pow_many(x, y, host=True)
    same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, [(x[i].hex(), y[i].hex(), got[i].hex(), want[i].hex()) for i in bad[:5]]


@pytest.mark.parametrize("y", [-0.2, 0.2, -1.0 / 6.0, -0.125, -1.0 / 9.0, 2.5])
def test_pow_error_below_0_52_ulp(orc, y):
    # The method's documented accuracy: under 0.52 ulp everywhere (not
    # correctly rounded).  A wrong table entry, ln2 split or polynomial
    # coefficient costs at least the subinterval's share of accuracy and
    # shows up here independently of any C library.
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 160
    rng = np.random.default_rng(7 + int(100 * abs(y)))
    lim = min(460.0, 700.0 / abs(y))  # keep x^y finite and normal
    xs = np.concatenate([np.exp(rng.uniform(-lim, lim, 3000)),
                         1.0 + rng.integers(-2 ** 45, 2 ** 45, 1000) * 2.0 ** -52,
                         # every log subinterval (top 7 bits of bits(x) - 0x3fe6955500000000)
                         (0x3fe6955500000000 + (np.arange(128, dtype=np.uint64) << np.uint64(45))
                          + np.uint64(12345)).view(np.float64)])
    got = orc.pow_many(xs, y)
    worst = 0.0
    for x, g in zip(xs, got):
        exact = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(y))
        ulp = mpmath.mpf(float(np.spacing(abs(g))))
        worst = max(worst, float(abs(mpmath.mpf(float(g)) - exact) / ulp))
    assert worst < 0.52, worst


def test_pow_special_cases(orc):
    # C99 Annex F / IEEE 754 pow special values.
    cases = [(0.0, -0.2, math.inf), (0.0, 0.2, 0.0), (math.inf, -0.2, 0.0), (math.inf, 0.2, math.inf),
             (1.0, -0.2, 1.0), (1.0, math.nan, 1.0), (math.nan, 0.0, 1.0), (-8.0, 3.0, -512.0),
             (-0.0, -3.0, -math.inf), (-0.0, 3.0, -0.0), (-2.0, -1.0, -0.5), (0.5, math.inf, 0.0),
             (2.0, -math.inf, 0.0), (-1.0, math.inf, 1.0), (2.0, 1024.0, math.inf),
             (2.0, -1075.0, 0.0), (2.0, -1074.0, 5e-324), (4.0, 0.5, 2.0)]
    for x, y, want in cases:
        got = orc.pow_many(np.array([x]), y)[0]
        assert got == want and math.copysign(1, got) == math.copysign(1, want), (x, y, got)
    for x, y in [(math.nan, -0.2), (-1.0, -0.2), (-2.0, 0.5), (2.0, math.nan)]:
        assert math.isnan(orc.pow_many(np.array([x]), y)[0])


def test_noise_normals_are_standard_normal(orc):
    from scipy import stats
    lib = orc.load()
    z = np.array([lib.de_oracle_noise_normal(0x5EED2026, p, s, c)
                  for p in range(50) for s in range(100) for c in range(8)])
    assert abs(z.mean()) < 0.03
    assert abs(z.var() - 1.0) < 0.04
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # deterministic, and keyed by (seed, path, step, component)
    assert lib.de_oracle_noise_normal(1, 2, 3, 4) == lib.de_oracle_noise_normal(1, 2, 3, 4)
    assert lib.de_oracle_noise_normal(1, 2, 3, 4) != lib.de_oracle_noise_normal(1, 2, 3, 5)


# Normals of the noise stream at fixed keys (seed, path, step, component), as
# the stream has been defined since the SDE kernels' first bit-exact runs: a
# change to Philox, the uniform conversion or the Box-Muller arithmetic that
# alters any bit fails here, on the CPU (the device draws the same normals).
NOISE_GOLDEN = [
    ((0x5EED2026, 0, 0, 0), "-0x1.6bde92f837ce2p-2"),
    ((0x5EED2026, 0, 0, 1), "0x1.0b9d12597893fp-2"),
    ((0x5EED2026, 1, 0, 0), "-0x1.030d167799bb5p+0"),
    ((0x5EED2026, 0, 999, 7), "0x1.3e7c07e4aa601p-5"),
    ((0x5EED2026, 4194303, 999, 6), "0x1.fa4cc38dad201p-1"),
    ((1, 2, 3, 4), "-0x1.b2e24a544ce7ap+0"),
    ((0xFFFFFFFFFFFFFFFF, 2 ** 40 + 5, 123456, 3), "-0x1.e38474b2845f5p-3"),
    ((0, 0, 0, 0), "-0x1.f1bcd549837c0p-4"),
]


@pytest.mark.parametrize("key,want", NOISE_GOLDEN)
def test_noise_normals_golden(orc, key, want):
    assert orc.load().de_oracle_noise_normal(*key) == float.fromhex(want)


# ------------------------------------------------------------------ DOP853 --
def _scipy_dop853():
    """SciPy's independent copy of the published DOP853 coefficients, as the
    tableau dict of pyref_dopri5.solve_dp (bh = b - E3, er = E5, bi rows 4..7
    = D; c[12] is unused: k[12] is the derivative at the new point)."""
    d = pytest.importorskip("scipy.integrate._ivp.dop853_coefficients")
    A = [[float(x) for x in row] for row in d.A[:16, :16]]
    A[12] = [0.0] * 16
    BI = [[0.0] * 16 for _ in range(16)]
    for r in range(4):
        BI[4 + r] = [float(x) for x in d.D[r, :16]]
    return dict(S=12, I=16, O=8, C=[float(x) for x in d.C[:16]], A=A,
                B=[float(x) for x in d.B], ER=[float(x) for x in d.E5[:12]],
                BH=[float(x) for x in (d.B - d.E3[:12])], BI=BI)


def test_dop853_tableau_equals_published_coefficients(orc):
    c, a, b, bh, er, bi = orc.dop853_tableau()
    ref = _scipy_dop853()
    for i in range(16):
        for j in range(16):
            assert bi[i, j] == ref["BI"][i][j], (i, j)
        if i == 12:  # unused by the reference (k[12] = f(t_new, y_new)); SciPy's is b
            assert c[i] == 0.0 and not a[i].any()
            continue
        assert c[i] == ref["C"][i], i
        for j in range(16):
            assert a[i, j] == ref["A"][i][j], (i, j)
    assert list(b) == ref["B"] and list(er) == ref["ER"] and list(bh) == ref["BH"]


def test_dop853_tableau_order_conditions(orc):
    # Quadrature and tall-tree conditions of order 8 (b^T c^q = 1/(q+1),
    # b^T A^k 1 = 1/(k+1)!), row sums = c, the 5th-order embedded weights
    # b - er to order 5; in floating point the decimal coefficients meet them
    # to ~1e-15, a transcription error by far more.
    c, a, b, bh, er, bi = orc.dop853_tableau()
    A, C = a[:12, :12], c[:12]
    for q in range(8):
        assert abs(b @ C ** q - 1.0 / (q + 1)) < 1e-14, q
    v = np.ones(12)
    for k in range(8):
        assert abs(b @ v - 1.0 / math.factorial(k + 1)) < 1e-14, k
        v = A @ v
    assert np.abs(a.sum(1)[:12] - C).max() < 1e-14
    assert np.abs(a.sum(1)[13:] - c[13:]).max() < 1e-14
    b5 = b - er
    for q in range(5):
        assert abs(b5 @ C ** q - 1.0 / (q + 1)) < 1e-14, q


def test_dop853_single_step_is_R_of_z(orc):
    # One accepted step on y' = lam y equals R(z) y0 with R computed exactly
    # from the tableau's doubles: R(z) = 1 + z b^T (I - zA)^{-1} 1.
    c, a, b, bh, er, bi = orc.dop853_tableau()
    Af = [[Fr(float(x)) for x in row] for row in a[:12, :12]]
    bf = [Fr(float(x)) for x in b]
    for lam, h in [(1.0, 0.1), (-2.0, 0.05)]:
        ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1) \
            .method(ExplicitRungeKutta.dop853().h0(h).rtol(1.0).atol(1.0))
        sol = orc.solve_ode(ens)
        assert sol.accepted[0] == 1 and sol.rejected[0] == 0
        z = Fr(lam) * Fr(h)
        R = Fr(1) + z * sum(bf[i] * x for i, x in enumerate(_apply_resolvent(Af, z)))
        assert abs(sol.y[0] - float(R)) <= 4e-16 * abs(float(R))
        assert abs(sol.y[0] - math.exp(lam * h)) < 1e-14


def _apply_resolvent(A, z):
    """(I - zA)^{-1} 1 for strictly lower-triangular A, exactly."""
    s = len(A)
    x = [Fr(0)] * s
    for i in range(s):
        x[i] = Fr(1) + z * sum(A[i][j] * x[j] for j in range(i))
    return x


def test_dop853_dense_output_bit_equals_python_restatement(orc):
    # The continuous extension (three extra stages, cont[4..8) from bi rows
    # 4..7) and the reference's interpolation fold, bit for bit against the
    # Python transcription fed SciPy's coefficients.  The fold alternates s and
    # s1 in the high-order part in the opposite phase from Hairer's contd8, so
    # the interpolant is less accurate than DOP853's published 7th-order one;
    # it is the reference's.
    rho = 28.0
    y0 = [1.0, 1.0, 1.0]
    pts = list(np.linspace(0.0, 1.0, 23))

    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z

    want = pyref_dopri5.solve_dp(f, 0.0, 1.0, y0, 1e-9, 1e-11, _scipy_dop853(), t_eval=pts)
    ens = IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, 1.0, np.array(y0).reshape(3, 1)).t_eval(pts) \
        .method(ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(100000))
    got = orc.solve_ode(ens, pow_mode=1)
    assert got.n_dense[0] == len(want[6]) == len(pts)
    assert got.dense.reshape(len(pts), 3).tolist() == want[6]
    lam, h = 0.7, 0.2   # and the step's ends on y' = lam y
    ens = IVP.ode(Exponential(lam), 0.0, h, [1.0], n=1).t_eval([0.0, 0.1, 0.2]) \
        .method(ExplicitRungeKutta.dop853().h0(h).rtol(1.0).atol(1.0))
    sol = orc.solve_ode(ens)
    assert sol.dense[0] == 1.0 and sol.dense[2] == sol.y[0]
    assert abs(sol.dense[1] - math.exp(lam * 0.1)) < 1e-8


def test_dop853_evaluation_counts(orc):
    # 11 evaluations per attempt, plus f(t_new, y_new) and the three
    # dense-output stages on every accepted step; 3 for init with h_init.
    sol = orc.solve_ode(lorenz_ensemble(64, 2.0).method(
        ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(100000)))
    assert (sol.status == DE_STATUS_COMPLETE).all()
    assert (sol.evals == 3 + 11 * (sol.accepted + sol.rejected) + 4 * sol.accepted).all()


@pytest.mark.parametrize("traj", range(4))
def test_oracle_dop853_bit_equals_python_restatement(orc, traj):
    rng = np.random.default_rng(300 + traj)
    rho = float(rng.uniform(0, 50))
    y0 = list(1.0 + 0.1 * rng.standard_normal(3))

    def f(t, y, dy):
        x, yy, z = y
        dy[0] = 10.0 * (yy - x)
        dy[1] = x * (rho - z) - yy
        dy[2] = x * yy - (8.0 / 3.0) * z

    want = pyref_dopri5.solve_dp(f, 0.0, 2.0, y0, 1e-9, 1e-11, _scipy_dop853())
    ens = IVP.ode(Lorenz(10.0, rho, 8.0 / 3.0), 0.0, 2.0, np.array(y0).reshape(3, 1)) \
        .method(ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(100000))
    got = orc.solve_ode(ens, pow_mode=1)
    assert list(got.y) == want[0]
    assert got.t[0] == want[1]
    assert (got.accepted[0], got.rejected[0], got.evals[0], got.status[0]) == want[2:]


def test_generic_dormand_prince_step_reproduces_dopri5(orc):
    # The oracle's generic Dormand-Prince class with the DOPRI5 tableau gives
    # the dedicated DOPRI5 class's bits (dense output included).
    ens = lorenz_ensemble(48, 3.0).t_eval(np.linspace(0, 3, 17))
    a = orc.solve_ode(ens)
    orc.set_generic_dopri5(True)
    try:
        b = orc.solve_ode(ens)
    finally:
        orc.set_generic_dopri5(False)
    for f in ("y", "t", "accepted", "rejected", "evals", "status", "n_dense"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(a.dense, b.dense)
