"""Pure-Python restatement of the reference DOPRI5 solve for tiny cases.

Written directly from the reference's Rust source, in its operation order
(Python floats are IEEE doubles and Python never fuses a*b+c), as a second,
independent restatement to bit-compare the C++ oracle against:
  solve_ode              src/ode/solve_ivp.rs:116-277
  DOPRI5 init/step       src/methods/erk/dormandprince/ordinary.rs:16-270
  h_init                 src/methods/h_init.rs:45-135
  State::add_scaled      src/traits.rs:316-320 ; error_norm :394-411
  solve_dp: the same step for any Dormand-Prince tableau (DOP853 uses the
  second error estimate bh and three dense-output stages)
`powf` is math.pow, which CPython forwards to the C library's pow -- the
function a Rust build of the reference calls on linux-gnu.
"""
import math

EPS10 = 2.220446049250313e-16 * 10.0

C = [0.0, 0.2, 0.3, 0.8, 8.0 / 9.0, 1.0, 1.0]
A = [[0.0] * 7 for _ in range(7)]
A[1][0] = 0.2
A[2][0] = 3.0 / 40.0
A[2][1] = 9.0 / 40.0
A[3][0] = 44.0 / 45.0
A[3][1] = -56.0 / 15.0
A[3][2] = 32.0 / 9.0
A[4][0] = 19372.0 / 6561.0
A[4][1] = -25360.0 / 2187.0
A[4][2] = 64448.0 / 6561.0
A[4][3] = -212.0 / 729.0
A[5][0] = 9017.0 / 3168.0
A[5][1] = -355.0 / 33.0
A[5][2] = 46732.0 / 5247.0
A[5][3] = 49.0 / 176.0
A[5][4] = -5103.0 / 18656.0
A[6][0] = 35.0 / 384.0
A[6][1] = 0.0
A[6][2] = 500.0 / 1113.0
A[6][3] = 125.0 / 192.0
A[6][4] = -2187.0 / 6784.0
A[6][5] = 11.0 / 84.0
B = [35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0]
ER = [71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0, -17253.0 / 339200.0,
      22.0 / 525.0, -1.0 / 40.0]


def signum(x):
    if x != x:
        return x
    return math.copysign(1.0, x)


def fmax(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def fmin(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a < b else b


def add_scaled(v, alpha, w):
    for i in range(len(v)):
        v[i] += alpha * w[i]


def tol(v, i):
    """Tolerance indexing (src/tolerance.rs:33-51): a scalar for every
    component, or the component's entry of a vector."""
    return v[i] if isinstance(v, (list, tuple)) else v


def h_init(f, t0, tf, y0, order, rtol, atol, h_min, h_max):
    posneg = signum(tf - t0)
    n = len(y0)
    f0 = [0.0] * n
    f1 = [0.0] * n
    f(t0, y0, f0)
    dnf = 0.0
    dny = 0.0
    sk = []
    for i in range(n):
        s = tol(atol, i) + tol(rtol, i) * abs(y0[i])
        sk.append(s)
        dnf += (f0[i] / s) * (f0[i] / s)
        dny += (y0[i] / s) * (y0[i] / s)
    if dnf <= 1.0e-10 or dny <= 1.0e-10:
        h = 1.0e-6
    else:
        h = math.sqrt(dny / dnf) * 0.01
    h = fmin(h, h_max)
    h *= posneg
    y1 = list(y0)
    add_scaled(y1, h, f0)
    f(t0 + h, y1, f1)
    der2 = 0.0
    for i in range(n):
        d = (f1[i] - f0[i]) / sk[i]
        der2 += d * d
    der2 = math.sqrt(der2) / abs(h)
    der12 = fmax(math.sqrt(dnf), der2)
    if der12 <= 1.0e-15:
        h1 = abs(h) * fmax(1.0e-3, 1.0e-6)
    else:
        h1 = math.pow(0.01 / der12, 1.0 / float(order))
    interval = abs(tf - t0)
    h = fmin(fmax(fmin(fmin(abs(h) * 100.0, h1), h_max), h_min), interval)
    return h * posneg


def solve(f, t0, tf, y0, rtol, atol, max_steps=100000, powf=math.pow):
    """Returns (y, t, accepted, rejected, evals, status) with status 0 on success."""
    n = len(y0)
    direction = signum(tf - t0)
    if direction not in (1.0, -1.0):
        return list(y0), t0, 0, 0, 0, 4
    evals = 0
    h = h_init(f, t0, tf, y0, 5, rtol, atol, 0.0, math.inf)
    evals += 2
    t = t0
    y = list(y0)
    k = [[0.0] * n for _ in range(7)]
    f(t, y, k[0])
    evals += 1
    h_prev = 0.0
    steps = 0
    stiff = 0
    nonstiff = 0
    rejected_status = False
    acc = rej = 0
    while True:
        if (t + h - tf) * direction > 0.0:
            h_new = tf - t
            if abs(h_new) < EPS10:
                break
            h = h_new
        if abs(h) < abs(h_prev) * 1e-14:
            return y, t, acc, rej, evals, 2
        if steps >= max_steps:
            return y, t, acc, rej, evals, 1
        steps += 1
        y_stage = [0.0] * n
        for i in range(1, 7):
            y_stage = list(y)
            for j in range(i):
                add_scaled(y_stage, A[i][j] * h, k[j])
            f(t + C[i] * h, y_stage, k[i])
        ysti = list(y_stage)
        yseg = [0.0] * n
        for i in range(7):
            add_scaled(yseg, B[i], k[i])
        y_new = list(y)
        add_scaled(y_new, h, yseg)
        t_new = t + h
        step_evals = 6
        err_state = [0.0] * n
        for j in range(7):
            add_scaled(err_state, ER[j], k[j])
        err = 0.0
        for i in range(n):
            sk = tol(atol, i) + tol(rtol, i) * fmax(abs(y[i]), abs(y_new[i]))
            e = err_state[i] / sk
            err += e * e
        deno = err + 0.01 * 0.0
        if deno <= 0.0:
            deno = 1.0
        err = abs(h) * err * math.sqrt(1.0 / (deno * float(n)))
        scale = 0.9 * powf(err, -(1.0 / 5.0))
        scale = fmin(fmax(scale, 0.2), 10.0)
        if err <= 1.0:
            dydt = [0.0] * n
            f(t_new, y_new, dydt)
            step_evals += 1
            if steps % 100 == 0:
                stdnum = 0.0
                stden = 0.0
                for i in range(n):
                    d = yseg[i] - k[6][i]
                    stdnum += d * d
                    d = dydt[i] - ysti[i]
                    stden += d * d
                if stden > 0.0:
                    if h * math.sqrt(stdnum / stden) > 6.1:
                        nonstiff = 0
                        stiff += 1
                        if stiff == 15:
                            return y, t, acc, rej, evals, 3
                else:
                    nonstiff += 1
                    if nonstiff == 6:
                        stiff = 0
            h_prev = h
            t = t_new
            y = y_new
            k[0] = dydt
            if rejected_status:
                rejected_status = False
                scale = fmin(scale, 1.0)
            evals += step_evals
            acc += 1
        else:
            rejected_status = True
            evals += step_evals
            rej += 1
        h *= scale
        a_h = abs(h)
        if a_h < 0.0:
            h = signum(h) * 0.0
        elif a_h > math.inf:
            h = signum(h) * math.inf
        if rejected_status:
            continue
        if abs(tf - t) <= EPS10:
            break
    return y, t, acc, rej, evals, 0


def dp_interpolate(cont, t_prev, h_prev, t_interp):
    """Interpolation of dormandprince/ordinary.rs:301-336: a fold from the last
    coefficient down to cont[1]; indices >= 4 take s1 when their distance to
    the last index is odd, indices 1..3 take s1 when odd; then cont[0] + s*acc."""
    s = (t_interp - t_prev) / h_prev
    s1 = 1.0 - s
    ilast = len(cont) - 1
    acc = list(cont[ilast])
    for i in range(ilast - 1, 0, -1):
        if i >= 4:
            factor = s1 if (ilast - i) % 2 == 1 else s
        else:
            factor = s1 if i % 2 == 1 else s
        acc = [a * factor for a in acc]
        add_scaled(acc, 1.0, cont[i])
    out = list(cont[0])
    add_scaled(out, s, acc)
    return out


def solve_dp(f, t0, tf, y0, rtol, atol, tab, max_steps=100000, powf=math.pow,
             t_eval=None, solout=None):
    """The reference's generic Dormand-Prince step (dormandprince/ordinary.rs:
    13-270, ExplicitRungeKutta<Ordinary, DormandPrince, T, Y, O, S, I>) with a
    tableau dict: S stages, I stages with the dense-output ones, order O,
    lists C, A, B, ER, BI and BH (None when the method has no second error
    estimate).  Same return value as solve(), plus the t_eval values (increasing
    points, TEvalSolout of src/solout/t_eval.rs:83-134) when t_eval is given.
    `solout(t_curr, t_prev, y_curr, interp)` is called after init and after
    every accepted step, where solve_ivp.rs:152-183,242-259 call Solout::solout;
    interp(t) is Interpolation::interpolate (None for Err(OutOfBounds)).  It
    returns "continue", "terminate" (status 5, Interrupted) or "panic"
    (status 6: the reference unwraps a failed interpolation)."""
    S, I, O = tab["S"], tab["I"], tab["O"]
    Cc, Aa, Bb, ERr, BH, BI = tab["C"], tab["A"], tab["B"], tab["ER"], tab["BH"], tab["BI"]
    n = len(y0)
    direction = signum(tf - t0)
    if direction not in (1.0, -1.0):
        return list(y0), t0, 0, 0, 0, 4
    evals = 0
    h = h_init(f, t0, tf, y0, O, rtol, atol, 0.0, math.inf)
    evals += 2
    t = t0
    y = list(y0)
    k = [[0.0] * n for _ in range(I)]
    f(t, y, k[0])
    evals += 1
    h_prev = 0.0
    steps = 0
    stiff = 0
    nonstiff = 0
    rejected_status = False
    acc = rej = 0
    dense = []
    idx = 0
    if t_eval is not None and len(t_eval) and t_eval[0] == t:
        dense.append(list(y))
        idx = 1
    if solout is not None:
        flag = solout(t, t, list(y), lambda tt: None)
        if flag != "continue":
            return y, t, acc, rej, evals, 5 if flag == "terminate" else 6
    while True:
        if (t + h - tf) * direction > 0.0:
            h_new = tf - t
            if abs(h_new) < EPS10:
                break
            h = h_new
        if abs(h) < abs(h_prev) * 1e-14:
            return y, t, acc, rej, evals, 2
        if steps >= max_steps:
            return y, t, acc, rej, evals, 1
        steps += 1
        y_stage = [0.0] * n
        for i in range(1, S):
            y_stage = list(y)
            for j in range(i):
                add_scaled(y_stage, Aa[i][j] * h, k[j])
            f(t + Cc[i] * h, y_stage, k[i])
        ysti = list(y_stage)
        yseg = [0.0] * n
        for i in range(S):
            add_scaled(yseg, Bb[i], k[i])
        y_new = list(y)
        add_scaled(y_new, h, yseg)
        t_new = t + h
        step_evals = S - 1

        def norm(e_state):
            s_ = 0.0
            for i in range(n):
                sk = tol(atol, i) + tol(rtol, i) * fmax(abs(y[i]), abs(y_new[i]))
                e = e_state[i] / sk
                s_ += e * e
            return s_

        err_state = [0.0] * n
        for j in range(S):
            add_scaled(err_state, ERr[j], k[j])
        err = norm(err_state)
        err2 = 0.0
        if BH is not None:
            err2_state = list(yseg)
            for j in range(S):
                add_scaled(err2_state, -BH[j], k[j])
            err2 = norm(err2_state)
        deno = err + 0.01 * err2
        if deno <= 0.0:
            deno = 1.0
        err = abs(h) * err * math.sqrt(1.0 / (deno * float(n)))
        scale = 0.9 * powf(err, -(1.0 / float(O)))
        scale = fmin(fmax(scale, 0.2), 10.0)
        if err <= 1.0:
            dydt = [0.0] * n
            f(t_new, y_new, dydt)
            step_evals += 1
            if steps % 100 == 0:
                stdnum = 0.0
                stden = 0.0
                for i in range(n):
                    d = yseg[i] - k[S - 1][i]
                    stdnum += d * d
                    d = dydt[i] - ysti[i]
                    stden += d * d
                if stden > 0.0:
                    if h * math.sqrt(stdnum / stden) > 6.1:
                        nonstiff = 0
                        stiff += 1
                        if stiff == 15:
                            return y, t, acc, rej, evals, 3
                else:
                    nonstiff += 1
                    if nonstiff == 6:
                        stiff = 0
            cont = [list(y)]
            ydiff = list(y_new)
            add_scaled(ydiff, -1.0, y)
            cont.append(list(ydiff))
            bspl = [0.0] * n
            add_scaled(bspl, h, k[0])
            add_scaled(bspl, -1.0, ydiff)
            cont.append(list(bspl))
            cont3 = list(ydiff)
            add_scaled(cont3, -h, dydt)
            add_scaled(cont3, -1.0, bspl)
            cont.append(cont3)
            if I > S:  # dense-output stages, evaluated on every accepted step
                k[S] = list(dydt)
                for i in range(S + 1, I):
                    y_stage = list(y)
                    for j in range(i):
                        add_scaled(y_stage, Aa[i][j] * h, k[j])
                    f(t + Cc[i] * h, y_stage, k[i])
                    step_evals += 1
            for i in range(4, O):
                ci = [0.0] * n
                for j in range(I):
                    add_scaled(ci, BI[i][j], k[j])
                cont.append([v * h for v in ci])
            if t_eval is not None:
                while idx < len(t_eval) and t < t_eval[idx] <= t_new:
                    te = t_eval[idx]
                    dense.append(list(y_new) if te == t_new else dp_interpolate(cont, t, h, te))
                    idx += 1
            h_prev = h
            h_prev_t = t
            t = t_new
            y = y_new
            k[0] = dydt
            if rejected_status:
                rejected_status = False
                scale = fmin(scale, 1.0)
            evals += step_evals
            acc += 1
            if solout is not None:
                def interp(tt, t_lo=h_prev_t, h_lo=h_prev, cont_lo=cont, t_hi=t):
                    if tt < t_lo or tt > t_hi:
                        return None
                    return dp_interpolate(cont_lo, t_lo, h_lo, tt)

                flag = solout(t, h_prev_t, list(y), interp)
                if flag != "continue":
                    return y, t, acc, rej, evals, 5 if flag == "terminate" else 6
        else:
            rejected_status = True
            evals += step_evals
            rej += 1
        h *= scale
        a_h = abs(h)
        if a_h < 0.0:
            h = signum(h) * 0.0
        elif a_h > math.inf:
            h = signum(h) * math.inf
        if rejected_status:
            continue
        if abs(tf - t) <= EPS10:
            break
    if t_eval is not None:
        return y, t, acc, rej, evals, 0, dense
    return y, t, acc, rej, evals, 0
