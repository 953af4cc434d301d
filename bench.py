#!/usr/bin/env python3
"""Benchmark: accepted steps x trajectories per second, FP64 DOPRI5, one MI355X.

Default workload (the configuration the metric is quoted on): 1 048 576 Lorenz
trajectories (sigma = 10, beta = 8/3, rho ~ U[0, 50], y0 ~ N((1,1,1), 0.1^2 I)),
DOPRI5 rtol = 1e-8, atol = 1e-10, t in [0, 50], final state and step
statistics per trajectory.  One "step" of the benchmark is one full ensemble
solve with inputs already resident in HBM.  The batched (persistent
lane-refilling) schedule is the headline `value`; the plain one-thread-per-
trajectory launch is reported beside it, with p50/p99 per-solve latency.

Other workloads (--workload): events (the Lorenz ensemble with a Poincare
section event z = rho - 1 per trajectory, and the crossing / hyperplane
solouts), erk (the Lorenz ensemble with each tableau of the embedded family,
rkf45 .. rkv989e), spring (65 536 uneven spring chains, dense output), sde
(OU-8 + GBM paths, Euler-Maruyama and Milstein), serve (Poisson request stream
through the serving loop).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "differential-equations_amd"))

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector/matrix peak (AMD spec), FMA = 2 flops
HBM_PEAK_GBS = 8000.0
LORENZ_FLOP_ACC = 263     # 7F + 69n, F = 8, n = 3 (per accepted step)
LORENZ_FLOP_REJ = 255     # 6F + 69n (per rejected step)
SEED = 0x5EED2026


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lorenz",
                    choices=["lorenz", "events", "erk", "spring", "sde", "serve"])
    ap.add_argument("--n", type=int, default=0, help="override ensemble size")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="lorenz: skip the spring-chain, SDE and serving legs")
    ap.add_argument("--only", default=None,
                    help="run only one variant (profiling): persistent, fifo, plain, fast "
                         "(lorenz); event, crossing (events); a tableau name (erk)")
    ap.add_argument("--cpu-sample-stride", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="requests per second")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window")
    ap.add_argument("--serve-graph", action="store_true", help="replay chunks from a hipGraph")
    ap.add_argument("--serve-offline", action="store_true",
                    help="also solve the stream's work as one batch per family")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC traffic calibration)")
    return ap.parse_args()


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


class Device:
    """Device buffers (torch owns HBM allocations and the stream)."""

    def __init__(self):
        import torch
        if not torch.cuda.is_available():
            raise SystemExit("bench.py needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.stream = torch.cuda.current_stream()

    def put(self, a, dtype=None):
        import numpy as np
        t = self.torch.from_numpy(np.ascontiguousarray(a))
        return t.to(self.dev)

    def empty(self, n, dtype):
        return self.torch.empty(n, dtype=dtype, device=self.dev)

    def event(self):
        return self.torch.cuda.Event(enable_timing=True)

    def sync(self):
        self.torch.cuda.synchronize()


def calibrate_copy(dv):
    t = dv.torch
    src = t.ones(1 << 27, dtype=t.float64, device=dv.dev)   # 1 GiB
    dst = t.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    dv.sync()


# --------------------------------------------------------------- Lorenz ---
def bench_lorenz(args, dv):
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    torch = dv.torch
    n = args.n or (1 << 20)
    p, y0 = deamd.gen_lorenz(n, seed=SEED)
    d_p = dv.put(p.reshape(-1))
    d_y0 = dv.put(y0.reshape(-1))
    out = dict(y=dv.empty(3 * n, torch.float64), t=dv.empty(n, torch.float64),
               acc=dv.empty(n, torch.int64), rej=dv.empty(n, torch.int64),
               ev=dv.empty(n, torch.int64), st=dv.empty(n, torch.int32))
    lib = deamd.lib()
    stream = C.c_void_p(dv.stream.cuda_stream)

    def descriptor(schedule, fast, method=_abi.DE_METHOD_DOPRI5):
        b = _abi.ode_defaults()
        b.family = _abi.DE_FAMILY_LORENZ
        b.method = method
        b.n = n
        b.dim = 3
        b.params = C.c_void_p(d_p.data_ptr())
        b.y0 = C.c_void_p(d_y0.data_ptr())
        b.t0_scalar = 0.0
        b.tf_scalar = 50.0
        b.rtol = 1e-8
        b.atol = 1e-10
        b.max_steps = 1000000
        b.schedule = schedule
        b.flags = _abi.DE_FLAG_FAST_MATH if fast else 0
        r = _abi.OdeResult()
        r.y_final = C.c_void_p(out["y"].data_ptr())
        r.t_final = C.c_void_p(out["t"].data_ptr())
        r.accepted = C.c_void_p(out["acc"].data_ptr())
        r.rejected = C.c_void_p(out["rej"].data_ptr())
        r.evals = C.c_void_p(out["ev"].data_ptr())
        r.status = C.c_void_p(out["st"].data_ptr())
        return b, r

    def run(schedule, fast, steps, warmup, method=_abi.DE_METHOD_DOPRI5):
        b, r = descriptor(schedule, fast, method)
        for _ in range(warmup):
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
        dv.sync()
        evs = [(dv.event(), dv.event()) for _ in range(steps)]
        dv.sync()
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(dv.stream)
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
            e1.record(dv.stream)
        dv.sync()
        wall = time.perf_counter() - t0
        lat = [e0.elapsed_time(e1) for e0, e1 in evs]   # ms, on the launch stream
        A = int(out["acc"].sum().item())
        R = int(out["rej"].sum().item())
        status_ok = bool((out["st"] == 0).all().item())
        return dict(wall=wall, lat=lat, A=A, R=R, ok=status_ok)

    if args.only:
        sched = {"plain": _abi.DE_SCHED_PLAIN,
                 "fifo": _abi.DE_SCHED_PERSISTENT_FIFO}.get(args.only, _abi.DE_SCHED_PERSISTENT)
        r = run(sched, args.only == "fast", args.steps, args.warmup)
        return {"variant": args.only, "accepted": r["A"], "rejected": r["R"],
                "kernel_ms": sum(r["lat"]) / len(r["lat"])}, (p, y0, n, None)
    res = run(_abi.DE_SCHED_PERSISTENT, False, args.steps, args.warmup)
    gpu_out = {k: out[k].cpu().numpy() for k in ("y", "t", "acc", "rej", "st")}
    plain = None if args.no_plain else run(_abi.DE_SCHED_PLAIN, False, max(3, args.steps // 2), 1)
    fifo = run(_abi.DE_SCHED_PERSISTENT_FIFO, False, max(3, args.steps // 2), 1)
    d853 = run(_abi.DE_SCHED_PERSISTENT, False, max(3, args.steps // 2), 1, _abi.DE_METHOD_DOP853)
    fast = run(_abi.DE_SCHED_PERSISTENT, True, max(3, args.steps // 2), 1)
    if args.calibrate_copy:
        calibrate_copy(dv)

    def summarize(r):
        ms = 1000.0 * r["wall"] / len(r["lat"])
        return dict(value=r["A"] / (ms / 1000.0), ms_per_step=ms,
                    p50_ms=percentile(r["lat"], 50), p99_ms=percentile(r["lat"], 99))

    main = summarize(res)
    kernel_ms = sum(res["lat"]) / len(res["lat"])
    flops = LORENZ_FLOP_ACC * res["A"] + LORENZ_FLOP_REJ * res["R"]
    achieved = flops / (kernel_ms / 1000.0) / 1e12
    line = {
        "metric": "accepted steps*traj/s FP64 DOPRI5, batched vs plain, 1 GPU; p50/p99 latency",
        "value": main["value"],
        "unit": "accepted_steps*traj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox4x32-10 inputs, seed 0x5EED2026)",
        "config": {"workload": "lorenz_dopri5_1M", "n_traj": n, "t_span": [0.0, 50.0],
                   "rtol": 1e-8, "atol": 1e-10, "schedule": "persistent",
                   "build": "parity (bit-exact, no FMA contraction)",
                   "accepted_steps_per_solve": res["A"], "rejected_steps_per_solve": res["R"],
                   "all_complete": res["ok"], "parallelism": "none (1 GPU)"},
        "latency_ms": {"p50": main["p50_ms"], "p99": main["p99_ms"]},
        "roofline": {"bound": "fp64_vector", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": load_traffic("lorenz_dopri5_1M"),
                     "flop_model": "263 per accepted + 255 per rejected step (7F+69n, 6F+69n)",
                     "kernel_ms": kernel_ms,
                     "kernel_ms_note": "HIP events around each solve on its stream: the "
                                       "longest-first ordering pre-pass plus the persistent "
                                       "kernel"},
    }
    fi = summarize(fifo)
    line["persistent_fifo"] = {"value": fi["value"], "ms_per_step": fi["ms_per_step"],
                               "note": "same kernel, trajectories in index order"}
    d8 = summarize(d853)
    line["dop853"] = {"value": d8["value"], "ms_per_step": d8["ms_per_step"],
                      "accepted_steps_per_solve": d853["A"], "all_complete": d853["ok"],
                      "note": "ExplicitRungeKutta::dop853 on the same ensemble and tolerances "
                              "(parity build)"}
    if plain is not None:
        ps = summarize(plain)
        line["plain"] = {"value": ps["value"], "ms_per_step": ps["ms_per_step"],
                         "p50_ms": ps["p50_ms"], "p99_ms": ps["p99_ms"]}
        line["batched_over_plain"] = main["value"] / ps["value"]
    fs = summarize(fast)
    fast_ms = sum(fast["lat"]) / len(fast["lat"])
    fast_flops = LORENZ_FLOP_ACC * fast["A"] + LORENZ_FLOP_REJ * fast["R"]
    line["fast_math"] = {"value": fs["value"], "ms_per_step": fs["ms_per_step"],
                         "tflops": fast_flops / (fast_ms / 1000.0) / 1e12,
                         "frac": fast_flops / (fast_ms / 1000.0) / 1e12 / FP64_PEAK_TFLOPS,
                         "note": "FMA-contracted build: same algorithm, not bit-equal"}
    return line, (p, y0, n, gpu_out)


def cpu_baseline_lorenz(args, p, y0, n, gpu_out):
    """The oracle with this host's C library pow (the reference binary's
    arithmetic) on every k-th trajectory, timed on all host cores, and the
    GPU's results on the same trajectories compared with it."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import deamd
    import oracle
    k = args.cpu_sample_stride
    idx = np.arange(0, n, k)
    ens = deamd.IVP.ode(deamd.Lorenz(*p[:, idx]), 0.0, 50.0, y0[:, idx]).method(
        deamd.ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000))
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    walls = []
    for _ in range(3):   # median of 3 repeats
        t0 = time.perf_counter()
        sol = oracle.solve_ode(ens, pow_mode=oracle.POW_HOST, threads=threads)
        walls.append(time.perf_counter() - t0)
    dt = sorted(walls)[1]
    base = {"value": float(sol.accepted.sum()) / dt, "unit": "accepted_steps*traj/s",
            "cores": threads, "kind": "port",
            "sample": "every %d-th trajectory of the 1M Lorenz ensemble (%d trajectories, "
                      "%d accepted steps, median %.2f s wall of 3 runs on %d threads); C++ "
                      "restatement of the reference solver with the C library pow"
                      % (k, len(idx), int(sol.accepted.sum()), dt, threads)}
    # Parity with the reference's arithmetic on the sampled trajectories.
    gy = gpu_out["y"].reshape(3, n)[:, idx]
    wy = sol.y.reshape(3, len(idx))
    counts = (gpu_out["acc"][idx] == sol.accepted) & (gpu_out["rej"][idx] == sol.rejected) \
        & (gpu_out["st"][idx] == sol.status)
    rel = np.max(np.abs(gy - wy) / np.maximum(np.abs(wy), 1e-300), axis=0)
    bits = np.all(gy.view(np.uint64) == wy.view(np.uint64), axis=0) \
        & (gpu_out["t"][idx].view(np.uint64) == sol.t.view(np.uint64))
    chaotic = p[1, idx] >= 24.74

    def share(mask, sel):
        return float(mask[sel].mean()) if sel.any() else None

    everything = np.ones(len(idx), dtype=bool)
    parity = {"trajectories": int(len(idx)), "host_pow": "C library pow of the host"}
    for name, sel in (("all", everything), ("rho_lt_24_74", ~chaotic), ("rho_ge_24_74", chaotic)):
        parity[name] = {"step_counts_equal": share(counts, sel),
                        "final_within_1e-10": share(rel <= 1e-10, sel),
                        "bit_identical": share(bits & counts, sel)}
    return base, parity


# ---------------------------------------------------- embedded ERK family ---
ERK_FAMILY = ["rkf45", "cash_karp", "rkv655e", "rkv656e", "rkv766e", "rkv767e", "rkv877e",
              "rkv878e", "rkv988e", "rkv989e"]


def erk_tableau_counts():
    """Stage counts and nonzero coefficients of each tableau, read from the
    generated header the kernels are built from: {name: (S, I, fsal, nnz of
    a[i][j] for i < S, nnz b, nnz bh)}."""
    import re
    text = open(os.path.join(REPO, "differential-equations_amd", "csrc",
                             "de_erk_tables.h")).read()
    out = {}
    for name in ERK_FAMILY:
        cls = "".join(w.capitalize() for w in name.split("_"))
        body = text[text.index("struct %s {" % cls):]
        body = body[:body.index("\n};\n")]
        S, I = (int(x) for x in re.search(r"S = (\d+), I = (\d+)", body).groups())
        fsal = "kFsal = true" in body

        def nums(tag):
            seg = body[body.index(tag):]
            seg = seg[seg.index("{") + 1:]
            depth, end = 1, 0
            for end, ch in enumerate(seg):
                depth += ch == "{"
                depth -= ch == "}"
                if depth == 0:
                    break
            return [float.fromhex(v) if "0x" in v else float(v)
                    for v in re.findall(r"-?0x[0-9a-fA-F.p+-]+|-?\d+\.\d+", seg[:end])]
        a = nums("A[%d][%d]" % (I, I))
        nnz_a = sum(1 for i in range(S) for j in range(i) if a[i * I + j] != 0.0)
        nnz_b = sum(1 for v in nums("B[%d]" % S) if v != 0.0)
        nnz_bh = sum(1 for v in nums("BH[%d]" % S) if v != 0.0)
        out[name] = (S, I, fsal, nnz_a, nnz_b, nnz_bh)
    return out


def erk_flops(counts, n=3, F=8):
    """Flops the lane executes per attempt and per accepted step (Lorenz: F =
    8 per right-hand side, n = 3): S - 1 right-hand sides and stage times; per
    nonzero a[i][j], b[j], bh[j] one scalar product and n multiply-adds; the
    error and its norm (9n); per accepted step f(t_new, y_new) unless FSAL."""
    S, I, fsal, nnz_a, nnz_b, nnz_bh = counts
    attempt = (S - 1) * (F + 2) + (nnz_a + nnz_b + nnz_bh) * (1 + 2 * n) + 9 * n
    return attempt, (0 if fsal else F + 1)


def bench_erk(args, dv):
    """The 1M Lorenz ensemble of the headline workload (t in [0, 50], rtol
    1e-8, atol 1e-10, persistent schedule, final state) solved with each
    tableau of the embedded family.  `value` is the family's aggregate: the
    accepted steps of the ten solves over their summed time."""
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    torch = dv.torch
    n = args.n or (1 << 20)
    p, y0 = deamd.gen_lorenz(n, seed=SEED)
    d_p = dv.put(p.reshape(-1))
    d_y0 = dv.put(y0.reshape(-1))
    out = dict(y=dv.empty(3 * n, torch.float64), t=dv.empty(n, torch.float64),
               acc=dv.empty(n, torch.int64), rej=dv.empty(n, torch.int64),
               ev=dv.empty(n, torch.int64), st=dv.empty(n, torch.int32))
    lib = deamd.lib()
    stream = C.c_void_p(dv.stream.cuda_stream)
    counts = erk_tableau_counts()
    names = [args.only] if args.only else ERK_FAMILY

    def run(name):
        b = _abi.ode_defaults()
        b.family = _abi.DE_FAMILY_LORENZ
        b.method = getattr(_abi, "DE_METHOD_" + name.upper())
        b.n, b.dim = n, 3
        b.params = C.c_void_p(d_p.data_ptr())
        b.y0 = C.c_void_p(d_y0.data_ptr())
        b.t0_scalar, b.tf_scalar = 0.0, 50.0
        b.rtol, b.atol, b.max_steps = 1e-8, 1e-10, 1000000
        b.schedule = _abi.DE_SCHED_PERSISTENT
        r = _abi.OdeResult()
        r.y_final = C.c_void_p(out["y"].data_ptr())
        r.t_final = C.c_void_p(out["t"].data_ptr())
        r.accepted = C.c_void_p(out["acc"].data_ptr())
        r.rejected = C.c_void_p(out["rej"].data_ptr())
        r.evals = C.c_void_p(out["ev"].data_ptr())
        r.status = C.c_void_p(out["st"].data_ptr())
        for _ in range(args.warmup):
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
        dv.sync()
        evs = [(dv.event(), dv.event()) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(dv.stream)
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
            e1.record(dv.stream)
        dv.sync()
        wall = time.perf_counter() - t0
        kernel_ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
        A = int(out["acc"].sum().item())
        R = int(out["rej"].sum().item())
        ok = bool((out["st"] == 0).all().item())
        ms = 1000.0 * wall / args.steps
        per_att, per_acc = erk_flops(counts[name])
        tflops = ((A + R) * per_att + A * per_acc) / (kernel_ms / 1000.0) / 1e12
        print("erk %-9s %8.2f ms  %.3e acc*traj/s  %.2f TFLOP/s" % (name, ms, A / (ms / 1e3),
                                                                  tflops),
              file=sys.stderr, flush=True)
        return dict(ms_per_step=ms, kernel_ms=kernel_ms, accepted=A, rejected=R,
                    value=A / (ms / 1000.0), tflops=tflops, all_complete=ok,
                    flops_per_attempt=per_att, flops_per_accepted_extra=per_acc)

    res = {name: run(name) for name in names}
    ms = sum(r["ms_per_step"] for r in res.values())
    A = sum(r["accepted"] for r in res.values())
    kern = sum(r["kernel_ms"] for r in res.values())
    fl = sum(r["tflops"] * r["kernel_ms"] for r in res.values()) / kern
    line = {
        "metric": "accepted steps*traj/s FP64, embedded ERK family (rkf45 .. rkv989e), 1 GPU",
        "value": A / (ms / 1000.0),
        "unit": "accepted_steps*traj/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox4x32-10 inputs, seed 0x5EED2026)",
        "config": {"workload": "lorenz_erk_family_1M", "n_traj": n, "t_span": [0.0, 50.0],
                   "rtol": 1e-8, "atol": 1e-10, "schedule": "persistent",
                   "tableaux": names, "parallelism": "none (1 GPU)",
                   "step": "one solve of the ensemble with each tableau in turn"},
        "methods": res,
        "roofline": {"bound": "fp64_vector", "achieved": fl, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": fl / FP64_PEAK_TFLOPS,
                     "traffic": load_traffic("lorenz_erk_family_1M"),
                     "flop_model": "per attempt (S-1)(F+2) + nnz(a, b, bh)(1+2n) + 9n, per "
                                   "accepted step F+1 unless FSAL; F = 8, n = 3; time-weighted "
                                   "over the tableaux"},
    }
    if not args.no_cpu_baseline and not args.only:
        line["cpu_baseline"] = cpu_baseline_erk(args, p, y0, n)
    return line


def cpu_baseline_erk(args, p, y0, n):
    """The oracle (C library pow) on every k-th trajectory, each tableau."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import deamd
    import oracle
    k = max(1, args.cpu_sample_stride // 4)   # about 20 s of CPU work over the ten tableaux
    idx = np.arange(0, n, k)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    A, dt = 0, 0.0
    for name in ERK_FAMILY:
        ens = deamd.IVP.ode(deamd.Lorenz(*p[:, idx]), 0.0, 50.0, y0[:, idx]).method(
            getattr(deamd.ExplicitRungeKutta, name)().rtol(1e-8).atol(1e-10).max_steps(1000000))
        t0 = time.perf_counter()
        sol = oracle.solve_ode(ens, pow_mode=oracle.POW_HOST, threads=threads)
        dt += time.perf_counter() - t0
        A += int(sol.accepted.sum())
    return {"value": A / dt, "unit": "accepted_steps*traj/s", "cores": threads, "kind": "port",
            "sample": "every %d-th trajectory of the 1M Lorenz ensemble (%d trajectories) with "
                      "each of the ten tableaux, %d accepted steps in %.2f s on %d threads; C++ "
                      "restatement of the reference solver with the C library pow"
                      % (k, len(idx), A, dt, threads)}


# --------------------------------------------------------------- events ---
def bench_events(args, dv):
    """The Lorenz ensemble of the headline workload with event solouts:
    IVP::event with g = z - (rho - 1) (the Poincare section through the two
    non-trivial fixed points, one value per trajectory, both directions),
    IVP::crossing(x = 0) and IVP::hyperplane_crossing; the same solve without
    a solout is timed beside them for the overhead."""
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    torch = dv.torch
    n = args.n or (1 << 20)
    cap = 128
    p, y0 = deamd.gen_lorenz(n, seed=SEED)
    d_p = dv.put(p.reshape(-1))
    d_y0 = dv.put(y0.reshape(-1))
    d_val = dv.put(np.ascontiguousarray(p[1] - 1.0))
    d_pt = dv.put(np.array([0.0, 0.0, 25.0]))
    d_nrm = dv.put(np.array([1.0, 1.0, 3.0]))
    out = dict(y=dv.empty(3 * n, torch.float64), t=dv.empty(n, torch.float64),
               acc=dv.empty(n, torch.int64), rej=dv.empty(n, torch.int64),
               ev=dv.empty(n, torch.int64), st=dv.empty(n, torch.int32),
               et=dv.empty(cap * n, torch.float64), ey=dv.empty(cap * 3 * n, torch.float64),
               ne=dv.empty(n, torch.int32))
    lib = deamd.lib()
    stream = C.c_void_p(dv.stream.cuda_stream)

    def descriptor(kind, method=_abi.DE_METHOD_DOPRI5):
        b = _abi.ode_defaults()
        b.family = _abi.DE_FAMILY_LORENZ
        b.method = method
        b.n, b.dim = n, 3
        b.params = C.c_void_p(d_p.data_ptr())
        b.y0 = C.c_void_p(d_y0.data_ptr())
        b.t0_scalar, b.tf_scalar = 0.0, 50.0
        b.rtol, b.atol, b.max_steps = 1e-8, 1e-10, 1000000
        b.schedule = _abi.DE_SCHED_PERSISTENT
        if kind == "event":
            b.event_solout = _abi.DE_SOLOUT_EVENT
            b.event_fn = _abi.DE_EVENT_COMPONENT
            b.event_component = 2
            b.event_value_traj = C.c_void_p(d_val.data_ptr())
        elif kind == "crossing":
            b.event_solout = _abi.DE_SOLOUT_CROSSING
            b.event_component = 0
        elif kind == "hyperplane":
            b.event_solout = _abi.DE_SOLOUT_HYPERPLANE
            b.event_point = C.c_void_p(d_pt.data_ptr())
            b.event_normal = C.c_void_p(d_nrm.data_ptr())
        b.max_events = cap if kind != "none" else 0
        r = _abi.OdeResult()
        r.y_final = C.c_void_p(out["y"].data_ptr())
        r.t_final = C.c_void_p(out["t"].data_ptr())
        r.accepted = C.c_void_p(out["acc"].data_ptr())
        r.rejected = C.c_void_p(out["rej"].data_ptr())
        r.evals = C.c_void_p(out["ev"].data_ptr())
        r.status = C.c_void_p(out["st"].data_ptr())
        if kind != "none":
            r.event_t = C.c_void_p(out["et"].data_ptr())
            r.event_y = C.c_void_p(out["ey"].data_ptr())
            r.n_events = C.c_void_p(out["ne"].data_ptr())
        return b, r

    def run(kind, steps, warmup, method=_abi.DE_METHOD_DOPRI5):
        b, r = descriptor(kind, method)
        for _ in range(warmup):
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
        dv.sync()
        evs = [(dv.event(), dv.event()) for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(dv.stream)
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
            e1.record(dv.stream)
        dv.sync()
        wall = time.perf_counter() - t0
        lat = [e0.elapsed_time(e1) for e0, e1 in evs]
        A = int(out["acc"].sum().item())
        R = int(out["rej"].sum().item())
        E = int(out["ne"].sum().item()) if kind != "none" else 0
        ok = bool((out["st"] == 0).all().item())
        ms = 1000.0 * wall / steps
        return dict(value=A / (ms / 1000.0), ms_per_step=ms, kernel_ms=sum(lat) / len(lat),
                    A=A, R=R, events=E, events_per_s=E / (ms / 1000.0), ok=ok)

    if args.only in ("event", "crossing"):
        r = run(args.only, args.steps, args.warmup)
        return {"variant": args.only, "accepted": r["A"], "events": r["events"],
                "kernel_ms": r["kernel_ms"]}
    if args.only in ("rkf45", "rkv656e", "rkv878e"):  # the event solve with a family tableau
        r = run("event", args.steps, args.warmup, getattr(_abi, "DE_METHOD_" + args.only.upper()))
        return {"variant": "event/" + args.only, "accepted": r["A"], "events": r["events"],
                "kernel_ms": r["kernel_ms"]}
    ev = run("event", args.steps, args.warmup)
    base = run("none", max(3, args.steps // 2), 1)
    cr = run("crossing", max(3, args.steps // 2), 1)
    hp = run("hyperplane", max(3, args.steps // 2), 1)
    flops = LORENZ_FLOP_ACC * ev["A"] + LORENZ_FLOP_REJ * ev["R"]
    achieved = flops / (ev["kernel_ms"] / 1000.0) / 1e12
    # The embedded family with the same event: the cubic Hermite interpolant
    # (rkf45) and the bi interpolants with extra stages (rkv656e, rkv878e).
    family = {}
    for name in ("rkf45", "rkv656e", "rkv878e"):
        m = getattr(_abi, "DE_METHOD_" + name.upper())
        fe = run("event", max(3, args.steps // 2), 1, m)
        fb = run("none", max(3, args.steps // 2), 1, m)
        family[name] = {"value": fe["value"], "ms_per_step": fe["ms_per_step"],
                        "kernel_ms": fe["kernel_ms"], "events_per_solve": fe["events"],
                        "all_complete": fe["ok"],
                        "event_overhead": fe["ms_per_step"] / fb["ms_per_step"] - 1.0}
    return {
        "metric": "accepted steps*traj/s FP64 DOPRI5 with an event solout, 1 GPU",
        "value": ev["value"],
        "unit": "accepted_steps*traj/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ev["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox4x32-10 inputs, seed 0x5EED2026)",
        "config": {"workload": "lorenz_event_dopri5_1M", "n_traj": n, "t_span": [0.0, 50.0],
                   "rtol": 1e-8, "atol": 1e-10, "schedule": "persistent",
                   "solout": "IVP::event, g = z - (rho - 1), both directions",
                   "max_events": cap, "events_per_solve": ev["events"],
                   "accepted_steps_per_solve": ev["A"], "all_complete": ev["ok"],
                   "parallelism": "none (1 GPU)"},
        "events_per_s": ev["events_per_s"],
        "without_solout": {"value": base["value"], "ms_per_step": base["ms_per_step"],
                           "event_overhead": ev["ms_per_step"] / base["ms_per_step"] - 1.0},
        "crossing_x0": {"value": cr["value"], "ms_per_step": cr["ms_per_step"],
                        "events_per_solve": cr["events"], "all_complete": cr["ok"]},
        "hyperplane": {"value": hp["value"], "ms_per_step": hp["ms_per_step"],
                       "events_per_solve": hp["events"], "all_complete": hp["ok"]},
        "embedded_family": family,
        "roofline": {"bound": "fp64_vector", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": load_traffic("lorenz_event_dopri5_1M"),
                     "flop_model": "the step's 263 / 255 flops per accepted / rejected step; "
                                   "event evaluation and root search not counted",
                     "kernel_ms": ev["kernel_ms"]},
    }


# --------------------------------------------------------------- spring ---
def bench_spring(args, dv):
    """65 536 damped spring chains, N log-uniform in [4, 256], DOPRI5 rtol 1e-6
    atol 1e-9, t in [0, 10], dense output at 100 uniform times."""
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    torch = dv.torch
    n = args.n or 65536
    sysm, y0 = deamd.gen_spring(n, seed=SEED)
    pts = np.linspace(0.0, 10.0, 100)
    state_len = int(sysm.state_offset[-1])
    d = dict(sizes=dv.put(sysm.sizes), so=dv.put(sysm.state_offset), po=dv.put(sysm.param_offset),
             p=dv.put(sysm.params_packed), y0=dv.put(y0), te=dv.put(pts))
    out = dict(y=dv.empty(state_len, torch.float64), t=dv.empty(n, torch.float64),
               acc=dv.empty(n, torch.int64), rej=dv.empty(n, torch.int64),
               ev=dv.empty(n, torch.int64), st=dv.empty(n, torch.int32),
               dense=dv.empty(100 * state_len, torch.float64), nd=dv.empty(n, torch.int32))
    lib = deamd.lib()
    stream = C.c_void_p(dv.stream.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())

    def descriptor(mapping):
        b = _abi.ode_defaults()
        b.family = _abi.DE_FAMILY_SPRING_CHAIN
        b.method = _abi.DE_METHOD_DOPRI5
        b.n = n
        b.sizes = C.cast(ptr(d["sizes"]), C.POINTER(C.c_int32))
        b.state_offset = C.cast(ptr(d["so"]), C.POINTER(C.c_int64))
        b.param_offset = C.cast(ptr(d["po"]), C.POINTER(C.c_int64))
        b.params, b.y0, b.t_eval = ptr(d["p"]), ptr(d["y0"]), ptr(d["te"])
        b.n_eval = 100
        b.t0_scalar, b.tf_scalar = 0.0, 10.0
        b.rtol, b.atol = 1e-6, 1e-9
        b.max_steps = 1000000
        b.mapping = mapping
        r = _abi.OdeResult()
        r.y_final, r.t_final = ptr(out["y"]), ptr(out["t"])
        r.accepted, r.rejected = ptr(out["acc"]), ptr(out["rej"])
        r.evals, r.status = ptr(out["ev"]), ptr(out["st"])
        r.dense, r.n_dense = ptr(out["dense"]), ptr(out["nd"])
        return b, r

    def run(mapping, steps, warmup):
        b, r = descriptor(mapping)
        for _ in range(warmup):
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
        dv.sync()
        evs = [(dv.event(), dv.event()) for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(dv.stream)
            _abi.check(lib.de_ode_batch_solve(C.byref(b), C.byref(r), stream))
            e1.record(dv.stream)
        dv.sync()
        wall = time.perf_counter() - t0
        lat = [e0.elapsed_time(e1) for e0, e1 in evs]
        acc = out["acc"].cpu().numpy()
        rej = out["rej"].cpu().numpy()
        return dict(wall=wall, lat=lat, acc=acc, rej=rej,
                    ok=bool((out["st"] == 0).all().item()))

    res = run(_abi.DE_MAP_AUTO, args.steps, max(1, args.warmup))
    tuned = {}
    for cap in (4, 8, 16, 32, 64, 128, 256):
        g, c, blk = C.c_int32(), C.c_int32(), C.c_int32()
        if lib.de_spring_tuned(cap, C.byref(g), C.byref(c), C.byref(blk)) == 0:
            tuned[str(cap)] = {"lanes_per_system": g.value, "masses_per_lane": c.value,
                               "block": blk.value}
    plain = None if args.no_plain else run(_abi.DE_MAP_LANE, 2, 1)
    N = sysm.sizes.astype(np.float64)
    ms = 1000.0 * res["wall"] / len(res["lat"])
    A = float(res["acc"].sum())
    flops = float((187 * N * res["acc"] + 180 * N * res["rej"] + 12 * N * res["acc"]).sum()
                  + 12 * N.sum() * 100)
    dense_bytes = 100 * state_len * 8.0
    kern_s = sum(res["lat"]) / len(res["lat"]) / 1000.0
    tf = flops / kern_s / 1e12
    gbs = dense_bytes / kern_s / 1e9
    line = {
        "metric": "accepted steps*traj/s FP64 DOPRI5, batched vs plain, 1 GPU; p50/p99 latency",
        "value": A / (ms / 1000.0), "unit": "accepted_steps*traj/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox4x32-10 inputs, seed 0x5EED2026)",
        "config": {"workload": "spring_chain_dopri5_64k", "n_systems": n,
                   "masses": [int(sysm.sizes.min()), int(sysm.sizes.max())],
                   "mean_masses": float(N.mean()), "t_span": [0.0, 10.0], "rtol": 1e-6,
                   "atol": 1e-9, "n_eval": 100, "mapping": "auto (autotuned per size class)",
                   "autotuned": tuned, "all_complete": res["ok"],
                   "accepted_steps_per_solve": int(A)},
        "latency_ms": {"p50": percentile(res["lat"], 50), "p99": percentile(res["lat"], 99)},
        "roofline": {"bound": "fp64_vector" if tf / FP64_PEAK_TFLOPS > gbs / HBM_PEAK_GBS else "hbm",
                     "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tf / FP64_PEAK_TFLOPS, "traffic": load_traffic("spring_chain_dopri5_64k"),
                     "dense_write_GBps": gbs, "dense_write_frac": gbs / HBM_PEAK_GBS,
                     "flop_model": "187N per accepted, 180N per rejected step, +12N dense prep "
                                   "per accepted step, +12N per output point"},
    }
    if plain is not None:
        pms = 1000.0 * plain["wall"] / len(plain["lat"])
        line["plain"] = {"value": float(plain["acc"].sum()) / (pms / 1000.0), "ms_per_step": pms,
                         "mapping": "one thread per system, state in HBM"}
        line["batched_over_plain"] = line["value"] / line["plain"]["value"]
    return line


# ------------------------------------------------------------------ sde ---
# FP64 operations (an fma counts 2) of one OU-8 Euler-Maruyama path-step, read
# off the source: the step y + h f(y) + sigma dW and the drift theta (mu - y) at
# the new state, 6 per component; and one standard normal of the noise, half
# of a Box-Muller pair (two uniforms 3 each, log 27, -2 log and sqrt 2, angle
# reduction 5, sin/cos polynomials 34, the two products 2: 76) plus its
# scaling by sqrt(h).
SDE_STEP_FLOPS = 48
SDE_NORMAL_FLOPS = 39


def bench_sde(args, dv):
    """4 194 304 paths each of OU-8 and GBM, Euler-Maruyama and Milstein, dt 1e-3,
    1000 steps, Philox noise; every 100th state recorded."""
    import ctypes as C
    import numpy as np
    import deamd
    from deamd import _abi
    torch = dv.torch
    n = args.n or (1 << 22)
    ou, ou_y0 = deamd.gen_ou(n, 8, seed=SEED)
    gbm, gbm_y0 = deamd.gen_gbm(n)
    lib = deamd.lib()
    stream = C.c_void_p(dv.stream.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    stride = 100
    n_rec = 2000 // stride + 1
    runs = []
    for fam, sysm, y0, dim in ((_abi.DE_SDE_ORNSTEIN_UHLENBECK, ou, ou_y0, 8),
                               (_abi.DE_SDE_LINEAR_DIAGONAL, gbm, gbm_y0, 1)):
        dp_, dy0 = dv.put(sysm.params), dv.put(y0)
        o = dict(y=dv.empty(dim * n, torch.float64), t=dv.empty(n, torch.float64),
                 steps=dv.empty(n, torch.int64), ev=dv.empty(n, torch.int64),
                 st=dv.empty(n, torch.int32), path=dv.empty(n_rec * dim * n, torch.float64),
                 pt=dv.empty(n_rec, torch.float64))
        for scheme in (_abi.DE_SCHEME_EULER_MARUYAMA, _abi.DE_SCHEME_MILSTEIN):
            b = _abi.sde_defaults()
            b.family, b.scheme, b.n, b.dim = fam, scheme, n, dim
            b.noise = _abi.DE_NOISE_PHILOX
            b.params, b.y0 = ptr(dp_), ptr(dy0)
            b.t0, b.tf, b.h0 = 0.0, 1.0, 1e-3
            b.max_steps = 2000
            b.seed = SEED
            b.path_stride = stride
            r = _abi.SdeResult()
            r.y_final, r.t_final, r.steps = ptr(o["y"]), ptr(o["t"]), ptr(o["steps"])
            r.evals, r.status, r.path, r.path_t = ptr(o["ev"]), ptr(o["st"]), ptr(o["path"]), ptr(o["pt"])
            runs.append((fam, scheme, dim, b, r, o, (dp_, dy0)))

    def one_pass():
        for (_, _, _, b, r, _, _) in runs:
            _abi.check(lib.de_sde_batch_solve(C.byref(b), C.byref(r), stream))

    for _ in range(max(1, args.warmup)):
        one_pass()
    dv.sync()
    per = {}
    evs = [[(dv.event(), dv.event()) for _ in runs] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        for j, (_, _, _, b, r, _, _) in enumerate(runs):
            evs[k][j][0].record(dv.stream)
            _abi.check(lib.de_sde_batch_solve(C.byref(b), C.byref(r), stream))
            evs[k][j][1].record(dv.stream)
    dv.sync()
    wall = time.perf_counter() - t0
    total_steps = 0
    names = {(0, 0): "ou8_em", (0, 1): "ou8_milstein", (1, 0): "gbm_em", (1, 1): "gbm_milstein"}
    for j, (fam, scheme, dim, b, r, o, _) in enumerate(runs):
        lat = [evs[k][j][0].elapsed_time(evs[k][j][1]) for k in range(args.steps)]
        steps = int(o["steps"].sum().item())
        total_steps += steps
        kms = sum(lat) / len(lat)
        path_bytes = (steps // n // stride) * dim * n * 8.0
        per[names[(fam, scheme)]] = {"kernel_ms": kms, "steps_paths_per_s": steps / (kms / 1000),
                                     "path_write_GBps": path_bytes / (kms / 1000) / 1e9,
                                     "all_complete": bool((o["st"] == 0).all().item())}
    ms = 1000.0 * wall / args.steps
    ou_em = per["ou8_em"]
    ou_tf = ou_em["steps_paths_per_s"] * (SDE_STEP_FLOPS + 8 * SDE_NORMAL_FLOPS) / 1e12
    return {
        "metric": "SDE path-steps*paths/s FP64 (Euler-Maruyama, Milstein), 1 GPU",
        "value": total_steps / (ms / 1000.0), "unit": "steps*paths/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox4x32-10 noise keyed by seed, path, step, component)",
        "config": {"workload": "sde_ou8_gbm_4M", "n_paths": n, "dt": 1e-3, "steps": 1000,
                   "record_every": stride, "per_run": per},
        "roofline": {"bound": "fp64_vector", "achieved": ou_tf, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": ou_tf / FP64_PEAK_TFLOPS,
                     "traffic": load_traffic("sde_ou8_em_4M"),
                     "flop_model": "OU-8 Euler-Maruyama, per path-step: %d for the step (drift, "
                                   "increment, diffusion term) + 8 x %d for the normals of the "
                                   "noise (Box-Muller: uniforms, log, sqrt, sin/cos polynomials, "
                                   "scaling by sqrt(h)); fma = 2; the Philox rounds' integer "
                                   "products and xors are not counted" % (SDE_STEP_FLOPS,
                                                                           SDE_NORMAL_FLOPS),
                     "update_only_frac": ou_em["steps_paths_per_s"] * SDE_STEP_FLOPS / 1e12
                     / FP64_PEAK_TFLOPS,
                     "path_write_GBps": ou_em["path_write_GBps"],
                     "path_write_frac": ou_em["path_write_GBps"] / HBM_PEAK_GBS},
    }


# ---------------------------------------------------------------- serve ---
def bench_serve(args, dv):
    """Poisson request stream through the serving loop.  Each request: one
    family (Lorenz, Van der Pol mu ~ LogUniform[0.1, 10], spring chain
    N in [4, 256]), ensemble size log-uniform in [1, 65 536], rtol in
    {1e-6, 1e-8} with atol = rtol/100, t_final ~ U[1, 50]."""
    import threading
    import numpy as np
    import deamd
    stream = deamd.poisson_requests(args.serve_rate, args.serve_seconds, seed=SEED)
    n_req = len(stream)
    arrivals = [r[0] for r in stream]
    reqs = [r[1:] for r in stream]
    srv = deamd.Server(128 << 30, graph=args.serve_graph)
    tickets = [None] * n_req
    t_start = time.perf_counter()

    def submitter():
        for k, (fam, n, e) in enumerate(reqs):
            delay = t_start + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[k] = srv.submit(e)

    th = threading.Thread(target=submitter)
    th.start()
    steps = 0
    lat = []
    fam_lat = {0: [], 1: [], 2: []}
    for k in range(n_req):
        while tickets[k] is None:
            time.sleep(0.0005)
        sol = srv.poll(tickets[k])
        steps += int(sol.accepted.sum())
        l = srv.latency(tickets[k])
        lat.append(l)
        fam_lat[reqs[k][0]].append(l)
    t_end = time.perf_counter()
    th.join()
    iterations, _ = srv.stats()
    t_admit, t_iter, t_retire = srv.timing()
    srv.close()
    offline = serve_offline(stream) if args.serve_offline else None
    elapsed = t_end - t_start
    ms = lambda v: 1000.0 * v if v is not None else None
    traj = sum(r[1] for r in reqs)
    return {
        "metric": "served accepted steps*traj/s FP64 DOPRI5 with p50/p99 request latency, 1 GPU",
        "value": steps / elapsed, "unit": "accepted_steps*traj/s", "n_gpus": 1,
        "steps": n_req, "warmup": 0, "ms_per_step": 1000.0 * elapsed / n_req,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Poisson arrivals and Philox4x32-10 inputs)",
        "config": {"workload": "serve_poisson", "rate_req_per_s": args.serve_rate,
                   "arrival_window_s": args.serve_seconds, "requests": n_req,
                   "trajectories": traj, "graph": bool(args.serve_graph),
                   "wall_s": elapsed, "accepted_steps": steps, "iterations": iterations,
                   "offline": offline,
                   "scheduler_s": {"admit": t_admit, "iterate": t_iter, "retire": t_retire},
                   "latency_ms_by_family": {
                       name: {"p50": ms(percentile(fam_lat[i], 50)),
                              "p99": ms(percentile(fam_lat[i], 99)), "n": len(fam_lat[i])}
                       for i, name in ((0, "lorenz"), (1, "vdp"), (2, "spring"))}},
        "latency_ms": {"p50": ms(percentile(lat, 50)), "p99": ms(percentile(lat, 99))},
    }


def serve_offline(stream):
    """The stream's work as three batch solves (one per family, every
    trajectory with its request's t_final and tolerances), host memory in and
    out: the throughput the same work gets without arrivals."""
    import numpy as np
    import deamd
    by = {0: [], 1: [], 2: []}
    for (_, fam, _, e) in stream:
        by[fam].append(e)
    total_s, steps, per = 0.0, 0, {}
    for fam, es in by.items():
        if not es:
            continue
        tf = np.concatenate([np.full(e.n, float(e.tf)) for e in es])
        rt = np.concatenate([np.full(e.n, float(e._method.settings["rtol"])) for e in es])
        at = np.concatenate([np.full(e.n, float(e._method.settings["atol"])) for e in es])
        if fam < 2:
            dim = es[0].system.dim
            params = [np.concatenate([np.broadcast_to(e.system.params[j], (e.n,)) for e in es])
                      for j in range(len(es[0].system.params))]
            y0 = np.concatenate([e.y0.reshape(dim, e.n) for e in es], axis=1)
            sysm = type(es[0].system)(*params)
        else:
            sysm = deamd.SpringChain(np.concatenate([e.system.sizes for e in es]),
                                     np.concatenate([e.system.params_packed for e in es]))
            y0 = np.concatenate([e.y0 for e in es])
        big = deamd.IVP.ode(sysm, 0.0, tf, y0).method(
            deamd.ExplicitRungeKutta.dopri5().max_steps(1000000))
        big.rtol_traj, big.atol_traj = rt, at
        t0 = time.perf_counter()
        sol = big.solve()
        dt = time.perf_counter() - t0
        total_s += dt
        steps += int(sol.accepted.sum())
        per[("lorenz", "vdp", "spring")[fam]] = {"s": dt, "trajectories": big.n}
    return {"s": total_s, "accepted_steps": steps, "value": steps / total_s, "per_family": per}


def other_configs(args, dv):
    """The other configurations, each in a short run of its own workload: the
    uneven spring-chain ensemble (batched vs plain), the SDE paths and a
    one-second window of the Poisson serving stream."""
    def sub(**kw):
        a = argparse.Namespace(**vars(args))
        a.n = 0
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    keep = ("value", "ms_per_step", "latency_ms", "roofline", "plain", "batched_over_plain")
    out = {}
    spring = bench_spring(sub(steps=5, warmup=1), dv)
    out["spring_chain_dopri5_64k"] = dict({k: spring[k] for k in keep if k in spring},
                                          all_complete=spring["config"]["all_complete"])
    sde = bench_sde(sub(steps=2, warmup=1), dv)
    out["sde_ou8_gbm_4M"] = dict({k: sde[k] for k in keep if k in sde},
                                 per_run=sde["config"]["per_run"])
    serve = bench_serve(sub(serve_seconds=1.0), dv)
    out["serve_poisson"] = dict({k: serve[k] for k in keep if k in serve},
                                rate_req_per_s=serve["config"]["rate_req_per_s"],
                                arrival_window_s=serve["config"]["arrival_window_s"],
                                requests=serve["config"]["requests"],
                                trajectories=serve["config"]["trajectories"],
                                latency_ms_by_family=serve["config"]["latency_ms_by_family"])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported "
              "(an ensemble would shard by trajectory index ranges, one range per GPU)"
              % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    dv = Device()
    if args.workload == "lorenz":
        line, (p, y0, n, gpu_out) = bench_lorenz(args, dv)
        if not args.no_cpu_baseline and not args.only:
            line["cpu_baseline"], line["parity_vs_reference_arith"] = \
                cpu_baseline_lorenz(args, p, y0, n, gpu_out)
        if not args.only and not args.no_other_configs:
            line["other_configs"] = other_configs(args, dv)
    elif args.workload == "events":
        line = bench_events(args, dv)
    elif args.workload == "erk":
        line = bench_erk(args, dv)
    elif args.workload == "spring":
        line = bench_spring(args, dv)
    elif args.workload == "sde":
        line = bench_sde(args, dv)
    elif args.workload == "serve":
        line = bench_serve(args, dv)
    else:
        raise SystemExit("workload %s is not available yet" % args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
