"""Pure-Python restatement of the reference's embedded explicit Runge-Kutta
solve (ExplicitRungeKutta<Ordinary, Adaptive, ..>: rkf45, cash_karp and the
Verner pairs) for tiny cases.

Written directly from the reference's Rust source, in its operation order
(Python floats are IEEE doubles and Python never fuses a*b+c), to bit-compare
the C++ oracle against:
  solve loop             src/ode/solve_ivp.rs:116-277
  init / step            src/methods/erk/adaptive/ordinary.rs:16-210
  interpolate            src/methods/erk/adaptive/ordinary.rs:238-292
  cubic Hermite          src/interpolate.rs:40-60
  State::error_norm_inf  src/traits.rs:597-610 (max starts at zero; max()
                         ignores a NaN, so a NaN error is not a rejection)
`powf` is the C library's pow (what a Rust build calls on linux-gnu): math.pow,
save that pow(+-0, y < 0) is +inf there, where math.pow raises.
The tableau is a dict of S, I, O, fsal, has_bi, C, A, B, BH, BI.
"""
import math

from pyref_dopri5 import EPS10, add_scaled, fmax, fmin, h_init, signum, tol


def c_pow(x, y):
    if x == 0.0 and y < 0.0:
        return math.inf
    return math.pow(x, y)


def hermite(t_prev, t_curr, y_prev, y_curr, d_prev, d_curr, tt):
    h = t_curr - t_prev
    s = (tt - t_prev) / h
    s2 = s * s
    s3 = s2 * s
    h00 = 2.0 * s3 - 3.0 * s2 + 1.0
    h10 = s3 - 2.0 * s2 + s
    h01 = -2.0 * s3 + 3.0 * s2
    h11 = s3 - s2
    out = [0.0] * len(y_prev)
    add_scaled(out, h00, y_prev)
    add_scaled(out, h10 * h, d_prev)
    add_scaled(out, h01, y_curr)
    add_scaled(out, h11 * h, d_curr)
    return out


def interpolate(tab, st, tt):
    """ordinary.rs:241-292; None for Err(OutOfBounds)."""
    if tt < st["t_prev"] or tt > st["t"]:
        return None
    if tab["has_bi"]:
        I, O = tab["I"], tab["O"]
        s = (tt - st["t_prev"]) / st["h_prev"]
        out = list(st["y_prev"])
        for i in range(I):
            ci = float(tab["BI"][i][O - 1])
            for j in range(O - 2, -1, -1):
                ci = ci * s + float(tab["BI"][i][j])
            ci *= s
            add_scaled(out, ci * st["h_prev"], st["k"][i])
        return out
    return hermite(st["t_prev"], st["t"], st["y_prev"], st["y"], st["dydt_prev"], st["dydt"], tt)


def solve_erk(f, t0, tf, y0, rtol, atol, tab, h0=0.0, max_steps=10000, max_rejects=100,
              powf=c_pow, t_eval=None, solout=None):
    """Returns (y, t, accepted, rejected, evals, status, dense); status 0
    complete, 1 max steps, 2 step size, 3 stiffness (max_rejects consecutive
    rejections), 4 bad input.  `t_eval` takes increasing points."""
    S, I, O = tab["S"], tab["I"], tab["O"]
    C = [float(x) for x in tab["C"]]
    A = [[float(x) for x in row] for row in tab["A"]]
    B = [float(x) for x in tab["B"]]
    BH = [float(x) for x in tab["BH"]]
    n = len(y0)
    direction = signum(tf - t0)
    if direction not in (1.0, -1.0):
        return list(y0), t0, 0, 0, 0, 4, []
    evals = 0
    h = h0
    if h == 0.0:
        h = h_init(f, t0, tf, y0, O, rtol, atol, 0.0, math.inf)
        evals += 2 + 2  # InitialStepSize::compute counts two, init adds two more
    if abs(h) > abs(tf - t0) or h * direction <= 0.0:
        return list(y0), t0, 0, 0, evals, 4, []
    st = dict(t=t0, y=list(y0), dydt=[0.0] * n, k=[[0.0] * n for _ in range(I)],
              t_prev=t0, h_prev=0.0)
    f(t0, st["y"], st["dydt"])
    evals += 1
    st["y_prev"] = list(st["y"])
    st["dydt_prev"] = list(st["dydt"])
    steps = stiff = 0
    rejected_status = False
    acc = rej = 0
    dense = []
    idx = 0
    if t_eval is not None and len(t_eval) and t_eval[0] == t0:
        dense.append(list(y0))
        idx = 1
    # `solout(t_curr, t_prev, y_curr, interp)` after init and after every
    # accepted step (solve_ivp.rs:152-183,242-259); "terminate" ends the solve
    # Interrupted (5), anything else but "continue" with status 6.
    if solout is not None:
        flag = solout(t0, t0, list(st["y"]), lambda tt: None)
        if flag != "continue":
            return st["y"], t0, acc, rej, evals, 5 if flag == "terminate" else 6, dense
    while True:
        t, y, k = st["t"], st["y"], st["k"]
        if (t + h - tf) * direction > 0.0:
            h_new = tf - t
            if abs(h_new) < EPS10:
                break
            h = h_new
        if abs(h) < abs(st["h_prev"]) * 1e-14:
            return y, t, acc, rej, evals, 2, dense
        if steps >= max_steps:
            return y, t, acc, rej, evals, 1, dense
        steps += 1
        sev = 0  # the step's evaluations; an Err return drops them (solve_ivp.rs:211-229)
        k[0] = list(st["dydt"])
        for i in range(1, S):
            y_stage = list(y)
            for j in range(i):
                add_scaled(y_stage, A[i][j] * h, k[j])
            f(t + C[i] * h, y_stage, k[i])
        sev += S - 1
        y_high = list(y)
        for i in range(S):
            add_scaled(y_high, B[i] * h, k[i])
        y_low = list(y)
        for i in range(S):
            add_scaled(y_low, BH[i] * h, k[i])
        err = list(y_high)
        add_scaled(err, -1.0, y_low)
        err_norm = 0.0
        for i in range(n):
            sk = tol(atol, i) + tol(rtol, i) * fmax(abs(y[i]), abs(y_high[i]))
            err_norm = fmax(err_norm, abs(err[i] / sk))
        scale = 0.9 * powf(err_norm, -(1.0 / float(O)))
        scale = fmin(fmax(scale, 0.2), 10.0)
        if err_norm <= 1.0:
            st["t_prev"] = t
            st["y_prev"] = list(y)
            st["dydt_prev"] = list(k[0])
            st["h_prev"] = h
            if rejected_status:
                stiff = 0
                rejected_status = False
                scale = fmin(scale, 1.0)
            if tab["has_bi"]:
                for i in range(S, I):
                    y_stage = list(y)
                    for j in range(i):
                        add_scaled(y_stage, A[i][j] * h, k[j])
                    f(t + C[i] * h, y_stage, k[i])
                sev += I - S
            st["t"] = t + h
            st["y"] = y_high
            if tab["fsal"]:
                st["dydt"] = list(k[S - 1])
            else:
                st["dydt"] = [0.0] * n
                f(st["t"], st["y"], st["dydt"])
                sev += 1
            evals += sev
            acc += 1
            if t_eval is not None:
                while idx < len(t_eval) and st["t_prev"] < t_eval[idx] <= st["t"]:
                    te = t_eval[idx]
                    dense.append(list(st["y"]) if te == st["t"] else interpolate(tab, st, te))
                    idx += 1
            if solout is not None:
                snap = dict(st, k=[list(r) for r in st["k"]], y=list(st["y"]),
                            dydt=list(st["dydt"]))
                flag = solout(st["t"], st["t_prev"], list(st["y"]),
                              lambda tt, sn=snap: interpolate(tab, sn, tt))
                if flag != "continue":
                    return (st["y"], st["t"], acc, rej, evals, 5 if flag == "terminate" else 6,
                            dense)
        else:
            rejected_status = True
            stiff += 1
            if stiff >= max_rejects:
                return y, t, acc, rej, evals, 3, dense
            evals += sev
            rej += 1
        h *= scale
        a_h = abs(h)
        if a_h < 0.0:
            h = signum(h) * 0.0
        elif a_h > math.inf:
            h = signum(h) * math.inf
        if rejected_status:
            continue
        if abs(tf - st["t"]) <= EPS10:
            break
    return st["y"], st["t"], acc, rej, evals, 0, dense
