"""GPU parity of the CSV writer (Solution::to_csv, src/solution.rs:190-230):
the documents the HIP kernels write are byte-identical to the oracle's for
every source (final state, t_eval points, event points), for special and
random doubles, in host- and device-memory modes, across chunk boundaries.
"""
import ctypes as C
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_CSV_EVENTS, DE_CSV_FINAL, DE_CSV_TEVAL, CrossingDirection,
                   ExplicitRungeKutta, Exponential, Harmonic, IVP, Lorenz, _abi)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def assert_bytes_equal(got, want):
    """Byte equality with the first difference in the message (a plain
    `assert got == want` on megabytes makes pytest compute a diff for minutes)."""
    if got == want:
        return
    a = np.frombuffer(got, dtype=np.uint8)
    b = np.frombuffer(want, dtype=np.uint8)
    m = min(len(a), len(b))
    d = np.flatnonzero(a[:m] != b[:m])
    i = int(d[0]) if len(d) else m
    raise AssertionError("documents differ at byte %d of %d / %d: got %r, want %r"
                         % (i, len(got), len(want), got[max(0, i - 60):i + 60],
                            want[max(0, i - 60):i + 60]))


def test_device_formatter_equals_oracle(orc):
    # The number formatter alone, on the values of the document tests below.
    n = 40000
    rng = np.random.default_rng(101)
    vals = rng.integers(0, 2 ** 64, 5 * n, dtype=np.uint64).view(np.float64).copy()
    vals[:16] = [0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 5e-324, 1e-4, 1e16,
                 9999999999999998.0, 1.7976931348623157e308, 2.2250738585072014e-308,
                 0.1, 1.0, 123456789.0, 1e-5]
    out = np.zeros(32 * len(vals), dtype=np.uint8)
    ln = np.zeros(len(vals), dtype=np.int32)
    _abi.check(deamd.lib().de_diag_fmt(vals.ctypes.data, out.ctypes.data, ln.ctypes.data,
                                       len(vals)))
    assert (ln > 0).all()
    bad = []
    for i, x in enumerate(vals):
        got = out[32 * i:32 * i + ln[i]].tobytes().decode()
        want = orc.fmt_debug(x)
        if got != want:
            bad.append((x.hex() if np.isfinite(x) else repr(x), got, want))
            if len(bad) >= 8:
                break
    assert not bad, bad


def lorenz(n, tf, seed):
    p, y0 = deamd.gen_lorenz(n, seed=seed)
    return IVP.ode(Lorenz(*p), 0.0, tf, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(100000))


def test_final_state_special_and_random_doubles(orc):
    n = 40000
    ens = lorenz(n, 0.2, 101)
    sol = ens.solve()
    rng = np.random.default_rng(101)
    bits = rng.integers(0, 2 ** 64, 4 * n, dtype=np.uint64)
    vals = bits.view(np.float64).copy()
    special = np.array([0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 5e-324, 1e-4, 1e16,
                        9999999999999998.0, 1.7976931348623157e308, 2.2250738585072014e-308,
                        0.1, 1.0, 123456789.0, 1e-5])
    vals[:len(special)] = special
    vals[n:n + len(special)] = special[::-1]
    sol.t[:] = vals[:n]
    sol.y[:] = vals[n:]
    got, goff = sol.csv("final")
    want, woff = orc.ode_csv(ens, sol, DE_CSV_FINAL)
    assert np.array_equal(goff, woff)
    assert_bytes_equal(got, want)


@pytest.mark.parametrize("pts", [[0.0, 0.5, 1.0, 2.0, 4.0], [-1.0, 0.0, 0.0, 0.7, 3.9, 9.0]])
def test_t_eval_documents(orc, pts):
    n = 4096
    rng = np.random.default_rng(103)
    k = rng.uniform(0.5, 3.0, n)
    ens = IVP.ode(Harmonic(k), 0.0, 4.0, np.vstack([np.ones(n), rng.uniform(-1, 1, n)])) \
        .t_eval(pts).method(ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11))
    sol = ens.solve()
    got, goff = sol.csv("t_eval")
    want, woff = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
    assert np.array_equal(goff, woff)
    assert_bytes_equal(got, want)


def test_event_documents_and_empty_ones(orc):
    ens = lorenz(2048, 3.0, 107).crossing(0, 0.0, CrossingDirection.Both).options(max_events=16)
    sol = ens.solve()
    assert (sol.n_events == 0).any() and (sol.n_events > 0).any()
    got, goff = sol.csv("events")
    want, woff = orc.ode_csv(ens, sol, DE_CSV_EVENTS)
    assert np.array_equal(goff, woff)
    assert_bytes_equal(got, want)
    # t_eval points past tf: no points, empty documents
    ens = IVP.ode(Exponential(np.ones(64)), 0.0, 1.0, np.ones((1, 64))).t_eval([5.0]) \
        .method(ExplicitRungeKutta.dopri5())
    sol = ens.solve()
    data, offsets = sol.csv()
    assert data == b"" and not offsets.any()


def test_to_csv_writes_the_reference_file(tmp_path):
    ens = IVP.ode(Exponential(np.array([1.0, 2.0])), 0.0, 1.0, np.ones((1, 2))) \
        .t_eval([0.0, 0.5, 1.0]).method(ExplicitRungeKutta.dopri5())
    sol = ens.solve()
    path = tmp_path / "results" / "run.csv"
    sol.to_csv(str(path), i=1)
    text = path.read_text()
    lines = text.split("\n")
    assert lines[0] == "t,y0" and lines[1] == "0.0,1.0" and lines[-1] == ""
    assert float(lines[3].split(",")[1]) == pytest.approx(np.exp(2.0), rel=1e-5)
    sol.to_csv_all(str(tmp_path / "all" / "traj_{i}.csv"))
    assert (tmp_path / "all" / "traj_0.csv").read_bytes() == sol.csv()[0][:sol.csv()[1][1]]


def test_device_memory_mode_and_offsets_check(orc):
    # Device buffers through the HIP runtime the library links.
    from test_gpu_parity import HipBuffers
    n = 3000
    ens = lorenz(n, 1.0, 109).t_eval([0.0, 0.25, 0.5, 1.0])
    sol = ens.solve()
    want, woff = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
    b, _, keep = ens.batch()
    mem = HipBuffers()
    try:
        b.flags = b.flags & ~_abi.DE_FLAG_HOST_MEMORY
        b.t_eval = mem.put(ens._t_eval)
        b.t0 = b.tf = None
        r = _abi.OdeResult()
        r.dense = mem.put(sol.dense)
        r.n_dense = mem.put(sol.n_dense)
        d_off = mem.put(np.zeros(n + 1, dtype=np.int64))
        lib = deamd.lib()
        _abi.check(lib.de_ode_csv_offsets(C.byref(b), C.byref(r), DE_CSV_TEVAL, d_off, None))
        assert mem.hip.hipDeviceSynchronize() == 0
        assert np.array_equal(mem.get(d_off, (n + 1,), np.int64), woff)
        total = int(woff[-1])
        d_out = mem.alloc(total + 64)
        assert mem.hip.hipMemset(d_out, 0x5A, total + 64) == 0
        _abi.check(lib.de_ode_csv_write(C.byref(b), C.byref(r), DE_CSV_TEVAL, d_off, d_out, None))
        assert mem.hip.hipDeviceSynchronize() == 0
        out = mem.get(d_out, (total + 64,), np.uint8)
        assert_bytes_equal(out[:total].tobytes(), want)
        assert (out[total:] == 0x5A).all()
        # offsets that do not match the result: refused, nothing written past them
        half = np.array([total // 2], dtype=np.int64)
        assert mem.hip.hipMemcpy(C.c_void_p(d_off.value + 8 * n), half.ctypes.data, 8, 1) == 0
        assert mem.hip.hipMemset(d_out, 0x5A, total + 64) == 0
        rc = lib.de_ode_csv_write(C.byref(b), C.byref(r), DE_CSV_TEVAL, d_off, d_out, None)
        assert mem.hip.hipDeviceSynchronize() == 0
        assert rc == _abi.DE_ERR_INVALID
        assert (mem.get(d_out, (total + 64,), np.uint8)[total // 2:] == 0x5A).all()
    finally:
        mem.free()


def test_chunked_layout_matches_single_trajectory_documents(orc):
    # 70 000 trajectories x 1 001 slots > 2^26 slots: two chunks; documents of
    # trajectories on both sides of the cut equal the oracle's for that
    # trajectory solved alone.
    n, m = 70000, 1000
    rng = np.random.default_rng(113)
    k = rng.uniform(0.5, 3.0, n)
    y1 = rng.uniform(-1, 1, n)
    pts = np.linspace(0.0, 5.0, m)

    def ensemble(kk, yy):
        return IVP.ode(Harmonic(kk), 0.0, 5.0, np.vstack([np.ones(len(kk)), yy])).t_eval(pts) \
            .method(ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9))

    sol = ensemble(k, y1).solve()
    data, offsets = sol.csv("t_eval")
    cut = (1 << 26) // (m + 1)
    for i in [0, 1, cut - 1, cut, cut + 1, n - 1]:
        e1 = ensemble(k[i:i + 1], y1[i:i + 1])
        want, _ = orc.ode_csv(e1, orc.solve_ode(e1), DE_CSV_TEVAL)
        assert_bytes_equal(data[offsets[i]:offsets[i + 1]], want)
