"""Pure-Python restatement of the reference's event solouts for tiny cases.

A second transcription, written from the Rust source independently of the C++
oracle, used to bit-compare it (both run the same solve through
pyref_dopri5.solve_dp's solout hook):
  EventWrappedSolout   src/solout/event.rs:320-523 (detect_event :335-393,
                       brent_dekker :395-490, solout :499-522)
  DefaultSolout        src/solout/default.rs (pushes every step)
  TEvalSolout          src/solout/t_eval.rs:83-134
  CrossingSolout       src/solout/crossing.rs:181-320
  HyperplaneCrossing   src/solout/hyperplane.rs:126-340
Each solout object is a callable (t_curr, t_prev, y_curr, interp) -> flag and
keeps the Solution's points in .points (list of (t, y)).
"""
import math

EPS = 2.220446049250313e-16


def signum(x):
    """f64::signum: NaN for NaN, else +-1.0 by the sign bit."""
    if x != x:
        return x
    return math.copysign(1.0, x)


def differ(a, b):
    """signum(a) != signum(b) with IEEE !=, so a NaN signum differs from all."""
    sa, sb = signum(a), signum(b)
    return not (sa == sb)


class Solution:
    def __init__(self):
        self.points = []

    def push(self, t, y):
        self.points.append((t, list(y)))

    def last_t(self):
        return self.points[-1][0] if self.points else None


class DefaultSolout:
    def __init__(self, sol):
        self.sol = sol

    def __call__(self, t_curr, t_prev, y_curr, interp):
        self.sol.push(t_curr, y_curr)
        return "continue"


class TEvalSolout:
    def __init__(self, sol, t_evals, t0, tf):
        self.sol = sol
        self.dir = signum(tf - t0)
        self.t_evals = sorted(t_evals, reverse=self.dir < 0)
        self.next = 0

    def __call__(self, t_curr, t_prev, y_curr, interp):
        idx = self.next
        while idx < len(self.t_evals):
            te = self.t_evals[idx]
            if self.dir > 0:
                in_range = (te == t_prev and idx == 0) or (t_prev < te <= t_curr)
            else:
                in_range = (te == t_prev and idx == 0) or (t_prev > te >= t_curr)
            if in_range:
                if te == t_curr:
                    self.sol.push(te, y_curr)
                else:
                    y = interp(te)
                    if y is None:
                        return "panic"  # interpolate(t_eval).unwrap()
                    self.sol.push(te, y)
                idx += 1
            else:
                if (self.dir > 0 and te > t_curr) or (self.dir < 0 and te < t_curr):
                    break
                idx += 1
        self.next = idx
        return "continue"


class EventWrapped:
    """EventWrappedSolout with an event function g(t, y) and EventConfig
    (direction 'both' | 'positive' | 'negative', terminate None | k)."""

    def __init__(self, base, sol, g, t0, tf, direction="both", terminate=None):
        self.base = base
        self.sol = sol
        self.g = g
        self.cfg_dir = direction
        self.terminate = terminate
        self.last_g = None
        self.count = 0
        self.direction = signum(tf - t0)
        self.rel_tol = 1e-12
        self.abs_tol = 1e-14
        self.root_points = []  # the event points this solout pushed

    def brent(self, a, b, fa, fb, interp):
        if abs(fa) < abs(fb):
            a, b = b, a
            fa, fb = fb, fa
        c, fc = a, fa
        d = b - a
        e = d
        for _ in range(50):
            if fb == 0.0:
                return b
            if signum(fa) == signum(fb):
                a, fa = c, fc
                c, fc = b, fb
                d = b - a
                e = d
            if abs(fa) < abs(fb):
                c = b
                b = a
                a = c
                fc = fb
                fb = fa
                fa = fc
            tol = max_rust(self.abs_tol, self.rel_tol * abs(b))
            m = 0.5 * (a - b)
            if abs(m) <= tol or fb == 0.0:
                return b
            bis = True
            if abs(e) > tol and abs(fa) > abs(fb):
                s = fb / fa
                if a == c:
                    p = 2.0 * m * s
                    q = 1.0 - s
                else:
                    q1 = fa / fc
                    r = fb / fc
                    p = s * (2.0 * m * q1 * (q1 - r) - (b - a) * (r - 1.0))
                    q = (q1 - 1.0) * (r - 1.0) * (s - 1.0)
                if q > 0.0:
                    p = -p
                else:
                    q = -q
                if abs(2.0 * p) < (3.0 * m * q - abs(tol * q)) and p < abs(e * 0.5 * q):
                    e = d
                    d = p / q
                    bis = False
            if bis:
                d = m
                e = m
            a, fa = b, fb
            if abs(d) > tol:
                b = b + d
            else:
                b = b + (tol if m > 0.0 else -tol)
            yb = interp(b)
            if yb is None:
                return None
            fb = self.g(b, yb)
            c, fc = a, fa
        return None

    def __call__(self, t_curr, t_prev, y_curr, interp):
        flag = self.base(t_curr, t_prev, y_curr, interp)
        if flag == "terminate" or flag == "panic":
            return flag
        g_curr = self.g(t_curr, y_curr)
        if self.last_g is None:
            self.last_g = g_curr
            return flag
        g_prev = self.last_g
        change = differ(g_prev, g_curr)
        if self.cfg_dir == "positive":
            ok = change and g_prev < 0.0 and g_curr >= 0.0
        elif self.cfg_dir == "negative":
            ok = change and g_prev > 0.0 and g_curr <= 0.0
        else:
            ok = change
        if ok:
            a, b, fa, fb = t_prev, t_curr, g_prev, g_curr
            if (self.direction > 0 and a > b) or (self.direction < 0 and a < b):
                a, b = b, a
                fa, fb = fb, fa
            if fa * fb <= 0.0:
                te = self.brent(a, b, fa, fb, interp)
                if te is not None:
                    ye = interp(te)
                    if ye is None:
                        return "panic"
                    last = self.sol.last_t()
                    if last is None or abs(te - last) > self.abs_tol:
                        self.sol.push(te, ye)
                        self.root_points.append((te, list(ye)))
                    self.count += 1
                    if self.terminate is not None and self.count >= self.terminate:
                        self.last_g = g_curr
                        return "terminate"
        self.last_g = g_curr
        return flag


class NewtonCrossing:
    """CrossingSolout (offset = y[c] - threshold) or, with hyperplane=True,
    HyperplaneCrossingSolout (offset = signed distance with the normalised
    normal); `offset(y)` computes it."""

    def __init__(self, sol, offset, direction="both", hyperplane=False):
        self.sol = sol
        self.offset = offset
        self.cfg_dir = direction
        self.hyper = hyperplane
        self.last = None

    def newton(self, interp, tl, tu, ol, ou):
        t = tl - ol * (tu - tl) / (ou - ol)
        tolerance = EPS * 100.0
        for _ in range(10):
            yt = interp(t)
            if yt is None:
                raise OutOfBounds
            off = self.offset(yt)
            if abs(off) < tolerance:
                return t
            dt = (tu - tl) * 1e-6
            tp = t + dt
            yp = interp(tp)
            if yp is None:
                raise OutOfBounds
            der = (self.offset(yp) - off) / dt
            if abs(der) < (EPS if self.hyper else EPS * 10.0):
                break
            tn = t - off / der
            if tn < tl or tn > tu:
                t = (tl + tu) / 2.0
            else:
                if not self.hyper and abs(tn - t) < tolerance * 0.1:
                    t = tn
                    break
                t = tn
        yt = interp(t)
        if yt is None:
            raise OutOfBounds
        return t if abs(self.offset(yt)) < tolerance * 10.0 else None

    def __call__(self, t_curr, t_prev, y_curr, interp):
        off = self.offset(y_curr)
        if self.last is not None:
            lo = self.last
            cross = differ(lo, off)
            if self.hyper:
                cross = cross or (lo == 0.0 and off != 0.0) or (lo != 0.0 and off == 0.0)
            if cross:
                if self.cfg_dir == "positive":
                    rec = lo < 0.0 and off >= 0.0
                elif self.cfg_dir == "negative":
                    rec = lo > 0.0 and off <= 0.0
                else:
                    rec = True
                if rec:
                    try:
                        tc = self.newton(interp, t_prev, t_curr, lo, off)
                    except OutOfBounds:
                        return "panic"
                    if tc is None:
                        frac = -lo / (off - lo)
                        tc = t_prev + frac * (t_curr - t_prev)
                    yc = interp(tc)
                    if yc is None:
                        return "panic"
                    self.sol.push(tc, yc)
        self.last = off
        return "continue"


class OutOfBounds(Exception):
    pass


def max_rust(a, b):
    """f64::max: the other operand when one is NaN."""
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def normalise(normal):
    """HyperplaneCrossingSolout::new (hyperplane.rs:126-132)."""
    sq = 0.0
    for x in normal:
        sq = sq + x * x
    norm = math.sqrt(sq)
    if norm > EPS:
        inv = 1.0 / norm
        return [x * inv for x in normal]
    return list(normal)


def signed_distance(point, normal):
    def f(y):
        s = 0.0
        for c in range(len(y)):
            s = s + (y[c] + (-1.0) * point[c]) * normal[c]
        return s
    return f
