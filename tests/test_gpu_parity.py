"""GPU parity: the HIP kernels against the CPU oracle on the same inputs.

The bar is bit equality (final state, final time, accepted/rejected/evals
counts, status, dense output) for every trajectory: the kernels use the
reference's operation order with contraction off and the same pow, so any
difference is a bug.  The FMA-contracted fast build is checked separately
against looser, documented bounds.
"""
import math
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_NOISE_CONSTANT, DE_NOISE_PHILOX, DE_NOISE_TABLE, DE_SCHED_PERSISTENT,
                   DE_SCHED_PLAIN, DE_STATUS_BAD_INPUT, DE_STATUS_COMPLETE, DE_STATUS_MAX_STEPS,
                   EulerMaruyama, ExplicitRungeKutta, IVP, Lorenz, Milstein, VanDerPol)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def same(a, b):
    """Bitwise equality of float arrays (NaN == NaN; +0 == -0 as values)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def assert_ode_equal(got, want, dense=False):
    assert np.array_equal(got.status, want.status)
    assert np.array_equal(got.accepted, want.accepted)
    assert np.array_equal(got.rejected, want.rejected)
    assert np.array_equal(got.evals, want.evals)
    assert same(got.t, want.t)
    assert same(got.y, want.y)
    if dense:
        assert np.array_equal(got.n_dense, want.n_dense)
        assert same(got.dense, want.dense)


# ------------------------------------------------------- device arithmetic --
def test_device_pow_matches_host(orc):
    # The device pow against its host build and against this host's C library
    # pow (the reference binary's powf) over every regime of arguments.
    from test_oracle_math import _host_pow_is_glibc_fma, _pow_arguments
    lib = deamd.lib()
    rng = np.random.default_rng(1)
    x, y = _pow_arguments(rng, 1 << 18)
    out = np.empty_like(x)
    deamd._abi.check(lib.de_diag_pow(deamd._ptr(x), deamd._ptr(y), deamd._ptr(out), len(x)))
    assert same(out, orc.pow_many(x, y))
    if _host_pow_is_glibc_fma():
        assert same(out, orc.pow_many(x, y, host=True))


def test_device_division_sqrt_fma_are_ieee():
    lib = deamd.lib()
    rng = np.random.default_rng(2)
    n = 1 << 20
    a = np.exp(rng.uniform(-700, 700, n)) * rng.choice([-1, 1], n)
    b = np.exp(rng.uniform(-700, 700, n)) * rng.choice([-1, 1], n)
    c = rng.standard_normal(n)
    a[:4] = [0.0, 1.0, 4.0, 2.0]
    b[:4] = [1.0, 3.0, 7.0, 3.0]
    out = np.empty(3 * n)
    deamd._abi.check(lib.de_diag_arith(deamd._ptr(a), deamd._ptr(b), deamd._ptr(c),
                                       deamd._ptr(out), n))
    out = out.reshape(n, 3)
    assert np.array_equal(out[:, 0], a / b)
    pos = a > 0
    assert np.array_equal(out[pos, 1], np.sqrt(a[pos]))
    import mpmath  # fma reference: exact product-sum rounded once
    for i in range(0, n, n // 2000):
        want = float(mpmath.fadd(mpmath.fmul(a[i], b[i], exact=True), c[i], prec=53,
                                 rounding="n"))
        assert out[i, 2] == want


def test_device_noise_matches_host():
    lib = deamd.lib()
    n = 200000
    rng = np.random.default_rng(3)
    path = rng.integers(0, 1 << 40, n, dtype=np.uint64)
    step = rng.integers(0, 1 << 20, n, dtype=np.uint32)
    comp = rng.integers(0, 9, n, dtype=np.uint32)
    out = np.empty(n)
    deamd._abi.check(lib.de_diag_noise(77, deamd._ptr(path), deamd._ptr(step), deamd._ptr(comp),
                                       deamd._ptr(out), n))
    want = np.array([lib.de_noise_normal(77, int(p), int(s), int(c))
                     for p, s, c in zip(path[:5000], step[:5000], comp[:5000])])
    assert np.array_equal(out[:5000], want)


# ------------------------------------------------------------ Lorenz (C1) --
def lorenz_c1(n=1024, tf=20.0):
    p, y0 = deamd.gen_lorenz(n)
    return IVP.ode(Lorenz(*p), 0.0, tf, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000))


@pytest.fixture(scope="module")
def c1_oracle(orc):
    return orc.solve_ode(lorenz_c1(), threads=THREADS)


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT])
def test_lorenz_c1_bit_exact(c1_oracle, schedule):
    got = lorenz_c1().options(schedule=schedule).solve()
    assert (got.status == DE_STATUS_COMPLETE).all()
    assert_ode_equal(got, c1_oracle)


def test_lorenz_c1_equals_reference_arithmetic(orc):
    # The oracle with this host's C library pow is the reference binary's
    # arithmetic; the GPU matches it bit for bit on every trajectory.
    from test_oracle_math import _host_pow_is_glibc_fma
    if not _host_pow_is_glibc_fma():
        pytest.skip("host pow is not glibc's fused x86-64 build")
    want = orc.solve_ode(lorenz_c1(), pow_mode=orc.POW_HOST, threads=THREADS)
    assert_ode_equal(lorenz_c1().solve(), want)


def test_lorenz_c2_full_ensemble(orc):
    # The 1 048 576-trajectory Lorenz ensemble over t in [0, 50] (rho ~ U[0, 50],
    # rtol 1e-8, atol 1e-10), solved whole on the GPU: every 64th trajectory
    # bit for bit against the oracle with this host's C library pow (the
    # reference binary's arithmetic), and on every trajectory Complete,
    # |t - tf| <= 10 eps and evals = 3 + 6 (A + R) + A.
    from test_oracle_math import _host_pow_is_glibc_fma
    n = 1 << 20
    p, y0 = deamd.gen_lorenz(n)
    mk = lambda p, y0: IVP.ode(Lorenz(*p), 0.0, 50.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000))
    got = mk(p, y0).solve()
    assert (got.status == DE_STATUS_COMPLETE).all()
    assert np.all(np.abs(got.t - 50.0) <= 10 * np.finfo(float).eps)
    assert np.array_equal(got.evals, 3 + 6 * (got.accepted + got.rejected) + got.accepted)
    idx = np.arange(0, n, 64)
    mode = orc.POW_HOST if _host_pow_is_glibc_fma() else orc.POW_RESTATED
    want = orc.solve_ode(mk(p[:, idx], y0[:, idx]), pow_mode=mode, threads=THREADS)
    assert np.array_equal(got.status[idx], want.status)
    assert np.array_equal(got.accepted[idx], want.accepted)
    assert np.array_equal(got.rejected[idx], want.rejected)
    assert np.array_equal(got.evals[idx], want.evals)
    assert same(got.t[idx], want.t)
    assert same(got.y.reshape(3, n)[:, idx].ravel(), want.y)


def test_workspace_budget_below_the_order_falls_back_to_index_order():
    # 1 MiB of workspace cannot hold the longest-first order of 2^20
    # trajectories (5 MiB): the persistent kernel runs them in index order,
    # with the same bits.
    lib = deamd.lib()
    p, y0 = deamd.gen_lorenz(1 << 20, seed=12)
    mk = lambda: IVP.ode(Lorenz(*p), 0.0, 0.5, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(100000))
    want = mk().solve()
    old = lib.de_workspace_budget()
    try:
        deamd._abi.check(lib.de_set_workspace_budget(1 << 20))
        got = mk().solve()
    finally:
        deamd._abi.check(lib.de_set_workspace_budget(old))
    assert_ode_equal(got, want)


def test_lorenz_persistent_large_batch_matches_plain():
    # 65 536 trajectories (several refills per lane) over a short span.
    p, y0 = deamd.gen_lorenz(65536, seed=11)
    mk = lambda s: IVP.ode(Lorenz(*p), 0.0, 2.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(100000)).options(schedule=s)
    a = mk(DE_SCHED_PLAIN).solve()
    b = mk(DE_SCHED_PERSISTENT).solve()
    assert_ode_equal(b, a)


class HipBuffers:
    """Device buffers through the HIP runtime the library itself links (a
    second runtime, such as the one bundled with torch, is not initialised
    after the library's in the same process)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def poisoned(self, nbytes):
        p = self.alloc(nbytes)
        assert self.hip.hipMemset(p, 0xFF, nbytes) == 0   # NaN doubles, -1 integers
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_persistent_orders_cover_every_trajectory():
    # PERSISTENT takes trajectories longest-predicted first through a device
    # counting sort; PERSISTENT_FIFO in index order.  Outputs are device
    # buffers poisoned beforehand, so a trajectory that a wrong permutation
    # skipped (or solved twice in place of another) cannot match PLAIN.
    import ctypes as C
    from deamd import _abi
    lib = deamd.lib()
    mem = HipBuffers()
    n = 50000
    p, y0 = deamd.gen_lorenz(n, seed=17)
    d_p, d_y0 = mem.put(p.reshape(-1)), mem.put(y0.reshape(-1))
    d_tf = mem.put(np.linspace(0.5, 8.0, n))   # uneven lengths
    fields = [("y", 3 * n, np.float64), ("t", n, np.float64), ("acc", n, np.int64),
              ("rej", n, np.int64), ("ev", n, np.int64), ("st", n, np.int32)]
    results = {}
    try:
        for sched in (DE_SCHED_PLAIN, DE_SCHED_PERSISTENT, _abi.DE_SCHED_PERSISTENT_FIFO):
            out = {k: mem.poisoned(m * np.dtype(t).itemsize) for k, m, t in fields}
            b = _abi.ode_defaults()
            b.family, b.method, b.n, b.dim = _abi.DE_FAMILY_LORENZ, _abi.DE_METHOD_DOPRI5, n, 3
            b.params, b.y0, b.tf = d_p, d_y0, d_tf
            b.rtol, b.atol, b.max_steps, b.schedule = 1e-8, 1e-10, 1000000, sched
            r = _abi.OdeResult()
            r.y_final, r.t_final, r.accepted = out["y"], out["t"], out["acc"]
            r.rejected, r.evals, r.status = out["rej"], out["ev"], out["st"]
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), None))
            results[sched] = {k: mem.get(out[k], m, t) for k, m, t in fields}
    finally:
        mem.free()
    want = results[DE_SCHED_PLAIN]
    assert (want["st"] == DE_STATUS_COMPLETE).all()
    for sched in (DE_SCHED_PERSISTENT, _abi.DE_SCHED_PERSISTENT_FIFO):
        for k in want:
            assert np.array_equal(results[sched][k], want[k]), (sched, k)


# ------------------------------------------------ reference test systems ----
from test_oracle_reference_kats import ACCURACY  # noqa: E402


@pytest.mark.parametrize("case", ACCURACY, ids=[c[0] for c in ACCURACY])
def test_reference_systems_bit_exact(orc, case):
    name, system, t0, tf, y0, expected, methods = case
    for make_method, tol in methods:
        ens = IVP.ode(system, t0, tf, y0, n=3).method(make_method())
        want = orc.solve_ode(ens)
        got = ens.solve()
        assert_ode_equal(got, want)
        for c, e in enumerate(expected):
            assert abs(got.y[c * 3] - e) < tol


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT])
def test_dop853_lorenz_bit_exact(orc, schedule):
    p, y0 = deamd.gen_lorenz(8192, seed=41)
    ens = IVP.ode(Lorenz(*p), 0.0, 10.0, y0).method(
        ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(1000000)) \
        .options(schedule=schedule)
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).all()
    assert_ode_equal(ens.solve(), want)
    # early stops and per-component tolerances
    ens = IVP.ode(Lorenz(*p[:, :1024]), 0.0, 10.0, y0[:, :1024]).method(
        ExplicitRungeKutta.dop853().rtol([1e-9, 1e-7, 1e-10]).atol([1e-11, 1e-9, 1e-12])
        .max_steps(120)).options(schedule=schedule)
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_MAX_STEPS).any()
    assert_ode_equal(ens.solve(), want)


@pytest.mark.parametrize("make_method", [
    lambda: ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(1000000),
    lambda: ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000),
    lambda: ExplicitRungeKutta.rk4(0.01).max_steps(1000000),
    lambda: ExplicitRungeKutta.euler(0.001).max_steps(1000000)])
def test_dense_output_bit_exact(orc, make_method):
    p, y0 = deamd.gen_lorenz(256, seed=5)
    # Points before t0 and after tf are passed over (NaN rows), as is t0 when
    # it is not the first point (the idx == 0 rule of t_eval.rs:106).
    pts = np.concatenate([[-0.5, -0.1], np.linspace(0.0, 5.0, 64), [5.5]])
    ens = IVP.ode(Lorenz(*p), 0.0, 5.0, y0).t_eval(pts).method(make_method())
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert_ode_equal(got, want, dense=True)


def test_vdp_backward_and_per_trajectory_tolerances(orc):
    mu, y0 = deamd.gen_vdp(512, seed=9)
    ens = IVP.ode(VanDerPol(*mu), 5.0, -3.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9).max_steps(200000))
    ens.rtol_traj = np.where(np.arange(512) % 2 == 0, 1e-6, 1e-8)
    ens.atol_traj = ens.rtol_traj * 1e-2
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert_ode_equal(got, want)


@pytest.mark.parametrize("schedule", [DE_SCHED_PLAIN, DE_SCHED_PERSISTENT])
def test_per_component_tolerances_bit_exact(orc, schedule):
    # Tolerance::Vector (src/tolerance.rs): per component of every trajectory,
    # and the builder's one vector for all trajectories, with dense output.
    n = 8192
    p, y0 = deamd.gen_lorenz(n, seed=29)
    rng = np.random.default_rng(29)
    ens = IVP.ode(Lorenz(*p), 0.0, 3.0, y0).method(
        ExplicitRungeKutta.dopri5().max_steps(1000000)).options(schedule=schedule)
    ens.rtol_comp = 10.0 ** rng.uniform(-9, -5, 3 * n)
    ens.atol_comp = ens.rtol_comp * 10.0 ** rng.uniform(-3, 0, 3 * n)
    assert_ode_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS))
    ens = IVP.ode(Lorenz(*p[:, :512]), 0.0, 3.0, y0[:, :512]).t_eval(np.linspace(0, 3, 31)) \
        .method(ExplicitRungeKutta.dopri5().rtol([1e-8, 1e-6, 1e-9]).atol([1e-10, 1e-8, 1e-12])
                .max_steps(1000000)).options(schedule=schedule)
    assert_ode_equal(ens.solve(), orc.solve_ode(ens, threads=THREADS), dense=True)


def test_error_statuses_bit_exact(orc):
    p, y0 = deamd.gen_lorenz(256, seed=13)
    p[1, :8] = np.nan                      # NaN rho: rejected forever -> StepSize
    y0[:, 8:16] = np.inf
    ens = IVP.ode(Lorenz(*p), 0.0, 20.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(300))
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert_ode_equal(got, want)
    assert (got.status == DE_STATUS_MAX_STEPS).sum() > 100
    # BadInput: h0 larger than the interval, and tf == t0 per trajectory
    tf = np.full(256, 1.0)
    tf[::3] = 0.0
    ens = IVP.ode(Lorenz(*p), 0.0, tf, y0).method(ExplicitRungeKutta.dopri5().h0(0.5))
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.solve()
    assert_ode_equal(got, want)
    assert (got.status[::3] == DE_STATUS_BAD_INPUT).all()


def test_stiffness_detection_bit_exact(orc):
    # The reference's stiffness test (ordinary.rs:165-194) compares
    # |yseg - k6|^2 with |f(y_new) - y_stage6|^2; for f(y) = y the second is
    # rounding-level, so h_lamb exceeds 6.1 on most checks and the solve stops
    # with Error::Stiffness after 15 of them.
    ens = IVP.ode(deamd.Exponential(1.0), 0.0, 40.0, [1.0], n=8).method(
        ExplicitRungeKutta.dopri5().rtol(1e-13).atol(1e-13).max_steps(200000))
    want = orc.solve_ode(ens)
    got = ens.solve()
    assert (want.status == deamd.DE_STATUS_STIFFNESS).all()
    assert_ode_equal(got, want)


def test_stiffness_check_after_rejected_attempt(orc):
    # The stiffness test runs on accepted attempts whose number is a multiple
    # of 100 (ordinary.rs:167).  With safety factor 1 and an absolute-only
    # tolerance the controller rejects about two attempts in three, so some
    # multiples of 100 fall on rejected attempts; the checks at later
    # multiples must still happen (a countdown that only restarted on accepted
    # attempts missed them and let these solves run to the end).
    n = 256
    rng = np.random.default_rng(5)
    ens = IVP.ode(deamd.Exponential(1.0), 0.0, 20.0, np.ones((1, n))).method(
        ExplicitRungeKutta.dopri5().rtol(0.0).atol(1e-9).safety_factor(1.0).max_steps(20000))
    ens.atol_traj = 1e-9 * np.exp(rng.uniform(-3, 3, n))
    ens.rtol_traj = np.zeros(n)
    want = orc.solve_ode(ens)
    got = ens.solve()
    assert (want.status == deamd.DE_STATUS_STIFFNESS).sum() > n // 4
    assert (want.status == deamd.DE_STATUS_COMPLETE).sum() > n // 4
    assert_ode_equal(got, want)


# -------------------------------------------------------------------- SDE ---
@pytest.mark.parametrize("scheme", [EulerMaruyama, Milstein])
@pytest.mark.parametrize("family", ["ou8", "gbm"])
def test_sde_bit_exact(orc, scheme, family):
    n = 4096
    if family == "ou8":
        sysm, y0 = deamd.gen_ou(n, 8, seed=21)
    else:
        sysm, y0 = deamd.gen_gbm(n)
    ens = IVP.sde(sysm, 0.0, 1.0, y0, noise=DE_NOISE_PHILOX, seed=0x5EED2026, path_base=1000) \
        .method(scheme(1e-3).max_steps(2000)).record_every(100)
    want = orc.solve_sde(ens, threads=THREADS)
    got = ens.solve()
    assert np.array_equal(got.status, want.status)
    assert (got.status == DE_STATUS_COMPLETE).all()
    assert np.array_equal(got.steps, want.steps) and (got.steps == 1000).all()
    assert np.array_equal(got.evals, want.evals)
    assert same(got.t, want.t) and same(got.y, want.y)
    assert same(got.path, want.path) and same(got.path_t, want.path_t)


def test_sde_table_and_constant_noise(orc):
    dw = np.array([[0.1, -0.1], [-0.05, 0.2], [0.15, -0.05], [0.0, 0.1]])
    sysm = deamd.LinearDiagonal(2, np.tile(np.array([-1.0, 0.5, -2.0, 0.1])[:, None], 8))
    ens = IVP.sde(sysm, 0.0, 1.0, np.ones(16), noise=DE_NOISE_TABLE, noise_table=dw) \
        .method(Milstein(0.25))
    want = orc.solve_sde(ens)
    got = ens.solve()
    assert same(got.y, want.y) and (got.steps == 4).all()
    sysm = deamd.LinearDiagonal(1, np.tile(np.array([0.5, 0.1])[:, None], 4))
    ens = IVP.sde(sysm, 0.0, 0.1, np.ones(4), noise=DE_NOISE_CONSTANT, noise_constant=0.1) \
        .method(EulerMaruyama(0.01))
    want = orc.solve_sde(ens)
    got = ens.solve()
    assert same(got.y, want.y) and (np.abs(got.t - 0.1) < 1e-10).all()


# -------------------------------------------------------------- fast build --
def test_fast_build_close_on_nonchaotic_family(orc):
    mu, y0 = deamd.gen_vdp(2048, seed=4)
    ens = IVP.ode(VanDerPol(*mu), 0.0, 10.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000))
    want = orc.solve_ode(ens, threads=THREADS)
    got = ens.options(fast_math=True).solve()
    assert (got.status == DE_STATUS_COMPLETE).all()
    rel = np.abs(got.y - want.y) / np.maximum(np.abs(want.y), 1e-3)
    assert np.percentile(rel, 99) < 1e-6
