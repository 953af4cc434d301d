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

pytestmark = [pytest.mark.gpu, pytest.mark.skip(
    reason="the device CSV writer is switched off: its first GPU run did not return and no "
           "cause has been shown (a CPU emulation of the kernels with the same inputs is "
           "correct and in bounds); the entry points refuse with DE_ERR_UNSUPPORTED")]
THREADS = min(16, os.cpu_count() or 1)


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
    assert got == want


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
    assert np.array_equal(goff, woff) and got == want


def test_event_documents_and_empty_ones(orc):
    ens = lorenz(2048, 3.0, 107).crossing(0, 0.0, CrossingDirection.Both).options(max_events=16)
    sol = ens.solve()
    assert (sol.n_events == 0).any() and (sol.n_events > 0).any()
    got, goff = sol.csv("events")
    want, woff = orc.ode_csv(ens, sol, DE_CSV_EVENTS)
    assert np.array_equal(goff, woff) and got == want
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
    torch = pytest.importorskip("torch")
    n = 3000
    ens = lorenz(n, 1.0, 109).t_eval([0.0, 0.25, 0.5, 1.0])
    sol = ens.solve()
    want, woff = orc.ode_csv(ens, sol, DE_CSV_TEVAL)
    b, _, keep = ens.batch()
    dev = torch.device("cuda:0")
    d_t_eval = torch.from_numpy(ens._t_eval.copy()).to(dev)
    d_dense = torch.from_numpy(sol.dense.copy()).to(dev)
    d_nd = torch.from_numpy(sol.n_dense.copy()).to(dev)
    b.flags = b.flags & ~_abi.DE_FLAG_HOST_MEMORY
    b.t_eval = C.c_void_p(d_t_eval.data_ptr())
    b.t0 = b.tf = None
    r = _abi.OdeResult()
    r.dense = C.c_void_p(d_dense.data_ptr())
    r.n_dense = C.c_void_p(d_nd.data_ptr())
    d_off = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    lib = deamd.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _abi.check(lib.de_ode_csv_offsets(C.byref(b), C.byref(r), DE_CSV_TEVAL,
                                      C.c_void_p(d_off.data_ptr()), stream))
    torch.cuda.synchronize()
    assert np.array_equal(d_off.cpu().numpy(), woff)
    total = int(woff[-1])
    d_out = torch.full((total + 64,), 0x5A, dtype=torch.uint8, device=dev)
    _abi.check(lib.de_ode_csv_write(C.byref(b), C.byref(r), DE_CSV_TEVAL,
                                    C.c_void_p(d_off.data_ptr()), C.c_void_p(d_out.data_ptr()),
                                    stream))
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    assert out[:total].tobytes() == want and (out[total:] == 0x5A).all()
    # offsets that do not match the result: refused, nothing written past them
    d_off[n] = total // 2
    d_out.fill_(0x5A)
    rc = lib.de_ode_csv_write(C.byref(b), C.byref(r), DE_CSV_TEVAL,
                              C.c_void_p(d_off.data_ptr()), C.c_void_p(d_out.data_ptr()), stream)
    torch.cuda.synchronize()
    assert rc == _abi.DE_ERR_INVALID
    assert (d_out.cpu().numpy()[total // 2:] == 0x5A).all()


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
        assert data[offsets[i]:offsets[i + 1]] == want, i
