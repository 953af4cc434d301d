"""Python host side of the MI355X ensemble integrator.

Mirrors the reference crate's builder so ensemble calls read like the crate's
own tests (src/ivp.rs:277-317, 632-706, 781-895):

    sol = IVP.ode(Lorenz(sigma, rho, beta), 0.0, 20.0, y0) \\
             .method(ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10)) \\
             .solve()

where every argument that is per trajectory in an ensemble is an array over
trajectories.  `solve()` runs the HIP library (include/de_api.h); there is no
CPU fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import *  # noqa: F401,F403


def lib():
    return _abi.load()


# ----------------------------------------------------------------- methods --
class ExplicitRungeKutta:
    """ExplicitRungeKutta::{dopri5, dop853, rk4, euler} and the embedded family
    (rkf45, cash_karp, rkv655e .. rkv989e) with the builder setters of
    src/methods/erk/mod.rs:165-228 and the defaults of :115-158."""

    def __init__(self, method, h0=0.0):
        self.method_id = method
        self.settings = dict(rtol=1e-6, atol=1e-6, h0=h0, h_min=0.0, h_max=float("inf"),
                             max_steps=10000, max_rejects=100, safety_factor=0.9,
                             min_scale=0.2, max_scale=10.0)

    @classmethod
    def dopri5(cls):
        return cls(DE_METHOD_DOPRI5)

    @classmethod
    def dop853(cls):
        """ExplicitRungeKutta::dop853 (src/methods/erk/dormandprince/mod.rs:46-52)."""
        return cls(DE_METHOD_DOP853)

    # The embedded family, src/methods/erk/adaptive/mod.rs:43-122.
    @classmethod
    def rkf45(cls):
        return cls(DE_METHOD_RKF45)

    @classmethod
    def cash_karp(cls):
        return cls(DE_METHOD_CASH_KARP)

    @classmethod
    def rkv655e(cls):
        return cls(DE_METHOD_RKV655E)

    @classmethod
    def rkv656e(cls):
        return cls(DE_METHOD_RKV656E)

    @classmethod
    def rkv766e(cls):
        return cls(DE_METHOD_RKV766E)

    @classmethod
    def rkv767e(cls):
        return cls(DE_METHOD_RKV767E)

    @classmethod
    def rkv877e(cls):
        return cls(DE_METHOD_RKV877E)

    @classmethod
    def rkv878e(cls):
        return cls(DE_METHOD_RKV878E)

    @classmethod
    def rkv988e(cls):
        return cls(DE_METHOD_RKV988E)

    @classmethod
    def rkv989e(cls):
        return cls(DE_METHOD_RKV989E)

    @classmethod
    def rk4(cls, h0):
        return cls(DE_METHOD_RK4, h0)

    @classmethod
    def euler(cls, h0):
        return cls(DE_METHOD_EULER, h0)

    def _set(self, k, v):
        self.settings[k] = v
        return self

    def rtol(self, v):
        """Tolerance::Scalar (a number) or Tolerance::Vector (one value per state
        component, src/tolerance.rs:11-68)."""
        return self._set("rtol", v)

    def atol(self, v):
        """As rtol."""
        return self._set("atol", v)

    def h0(self, v):
        return self._set("h0", v)

    def h_min(self, v):
        return self._set("h_min", v)

    def h_max(self, v):
        return self._set("h_max", v)

    def max_steps(self, v):
        return self._set("max_steps", int(v))

    def max_rejects(self, v):
        """ExplicitRungeKutta::max_rejects (erk/mod.rs:201), the embedded family's
        stiffness limit."""
        return self._set("max_rejects", int(v))

    def safety_factor(self, v):
        return self._set("safety_factor", v)

    def min_scale(self, v):
        return self._set("min_scale", v)

    def max_scale(self, v):
        return self._set("max_scale", v)


class Milstein:
    """Milstein::new(h) (src/methods/milstein.rs:37-73)."""

    def __init__(self, h0):
        self.scheme = DE_SCHEME_MILSTEIN
        self.settings = dict(h0=h0, h_min=0.0, h_max=float("inf"), max_steps=10000)

    def h_min(self, v):
        self.settings["h_min"] = v
        return self

    def h_max(self, v):
        self.settings["h_max"] = v
        return self

    def max_steps(self, v):
        self.settings["max_steps"] = int(v)
        return self


class EulerMaruyama(Milstein):
    """ExplicitRungeKutta::euler(h) applied to an SDE (stochastic.rs)."""

    def __init__(self, h0):
        super().__init__(h0)
        self.scheme = DE_SCHEME_EULER_MARUYAMA


# ---------------------------------------------------------------- families --
def _soa(*cols, n=None):
    cols = [np.broadcast_to(np.asarray(c, dtype=np.float64), (n,)) for c in cols]
    return np.ascontiguousarray(np.stack(cols)) if cols else np.zeros((0, n))


class _Fixed:
    family = None
    dim = None

    def __init__(self, *params):
        self.params = params

    def pack(self, n):
        return _soa(*self.params, n=n).reshape(-1)


class Lorenz(_Fixed):
    family, dim = DE_FAMILY_LORENZ, 3


class VanDerPol(_Fixed):
    family, dim = DE_FAMILY_VAN_DER_POL, 2


class Exponential(_Fixed):
    family, dim = DE_FAMILY_EXPONENTIAL, 1


class Linear(_Fixed):
    family, dim = DE_FAMILY_LINEAR, 1


class Harmonic(_Fixed):
    family, dim = DE_FAMILY_HARMONIC, 2


class Logistic(_Fixed):
    family, dim = DE_FAMILY_LOGISTIC, 1


class Robertson(_Fixed):
    family, dim = DE_FAMILY_ROBERTSON, 3

    def pack(self, n):
        return np.zeros(1)


class SpringChain:
    """Ragged damped spring-mass chains: sizes[i] masses, params packed per
    system as (c, k_0..k_N), state as (x_0..x_{N-1}, v_0..v_{N-1})."""
    family = DE_FAMILY_SPRING_CHAIN
    dim = None

    def __init__(self, sizes, params_packed):
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        n = len(self.sizes)
        self.state_offset = np.zeros(n + 1, dtype=np.int64)
        self.param_offset = np.zeros(n + 1, dtype=np.int64)
        self.state_offset[1:] = np.cumsum(2 * self.sizes.astype(np.int64))
        self.param_offset[1:] = np.cumsum(self.sizes.astype(np.int64) + 2)
        self.params_packed = np.ascontiguousarray(params_packed, dtype=np.float64)


class OrnsteinUhlenbeck:
    family = DE_SDE_ORNSTEIN_UHLENBECK

    def __init__(self, dim, params):
        self.dim = dim
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)


class LinearDiagonal:
    family = DE_SDE_LINEAR_DIAGONAL

    def __init__(self, dim, params):
        self.dim = dim
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):  # torch tensor (device memory)
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ events --
class CrossingDirection:
    """CrossingDirection (src/solout/mod.rs:38-55)."""
    Both = DE_CROSS_BOTH
    Positive = DE_CROSS_POSITIVE
    Negative = DE_CROSS_NEGATIVE


class EventConfig:
    """EventConfig (src/solout/event.rs:8-50): crossing direction and the
    number of events after which the solve terminates (None: never)."""

    def __init__(self, direction=CrossingDirection.Both, terminate=None):
        self._direction = direction
        self._terminate = terminate

    def direction(self, d):
        self._direction = d
        return self

    def terminate_after(self, count):
        self._terminate = int(count)
        return self

    def terminal(self):
        return self.terminate_after(1)


class Event:
    """An event function g(t, y) built into the device kernels (the Event
    trait, src/solout/event.rs:52-60), with its EventConfig.

    Event.component(c, value): y[c] - value
    Event.time(value):         t - value
    Event.plane(point, normal): y.minus(&point).dot(&normal)
    `value` may be one number or one per trajectory."""

    def __init__(self, fn, component=0, value=0.0, point=None, normal=None, config=None):
        self.fn = fn
        self.comp = int(component)
        self.value = value
        self.point = point
        self.normal = normal
        self.cfg = config if config is not None else EventConfig()

    @classmethod
    def component(cls, c, value, config=None):
        return cls(DE_EVENT_COMPONENT, component=c, value=value, config=config)

    @classmethod
    def time(cls, value, config=None):
        return cls(DE_EVENT_TIME, value=value, config=config)

    @classmethod
    def plane(cls, point, normal, config=None):
        return cls(DE_EVENT_PLANE, point=point, normal=normal, config=config)

    def config(self, cfg):
        self.cfg = cfg
        return self


# ---------------------------------------------------------------- problems --
class OdeEnsemble:
    """IVP::ode for an ensemble.  Arrays live in host memory (numpy)."""

    def __init__(self, system, t0, tf, y0, n=None):
        self.system = system
        self.t0 = t0
        self.tf = tf
        if isinstance(system, SpringChain):
            self.n = len(system.sizes)
            self.y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1)
        else:
            y0 = np.asarray(y0, dtype=np.float64)
            if y0.ndim == 1 and system.dim == len(y0) and n is not None:
                y0 = np.repeat(y0[:, None], n, axis=1)
            elif y0.ndim == 1:
                y0 = y0.reshape(system.dim, -1)
            self.n = y0.shape[1]
            self.y0 = np.ascontiguousarray(y0).reshape(-1)
        self._method = ExplicitRungeKutta.dopri5()
        self._t_eval = None
        self.schedule = DE_SCHED_AUTO
        self.mapping = DE_MAP_AUTO
        self.group_size = 0
        self.flags = 0
        self.rtol_traj = None
        self.atol_traj = None
        # Per-component tolerances of every trajectory, in the layout of y0.
        self.rtol_comp = None
        self.atol_comp = None
        # (DE_SOLOUT_*, Event) of an event / crossing solout, or None.
        self._event = None
        self.max_events = 16
        # Variable-count point output: (DE_POINTS_*, dense n, even dt).
        self._points = (DE_POINTS_NONE, 0, 0.0)

    def method(self, m):
        self._method = m
        return self

    def all_steps(self):
        """DefaultSolout, the reference's default output: (t0, y0) and every
        accepted step (src/solout/default.rs:71-91).  The points come back in
        Solution.points(i)."""
        return self._set_points(DE_POINTS_STEPS, 0, 0.0)

    def dense(self, n):
        """IVP::dense(n) (src/ivp.rs:652-655): DenseSolout, n - 1 interpolated
        points inside every accepted step, then the step itself."""
        return self._set_points(DE_POINTS_DENSE, int(n), 0.0)

    def even(self, dt):
        """IVP::even(dt) (src/ivp.rs:645-649): EvenSolout, the points t0, t0 + dt,
        ... in the integration interval, and tf."""
        return self._set_points(DE_POINTS_EVEN, 0, float(dt))

    def _set_points(self, mode, n, dt):
        self._points = (mode, n, dt)
        self._t_eval = None   # with_solout replaces the whole solout
        self._event = None
        return self

    def t_eval(self, points):
        self._points = (DE_POINTS_NONE, 0, 0.0)
        pts = np.sort(np.asarray(points, dtype=np.float64))
        t0 = float(np.asarray(self.t0).reshape(-1)[0])
        tf = float(np.asarray(self.tf).reshape(-1)[0])
        if tf < t0:
            pts = pts[::-1].copy()
        self._t_eval = np.ascontiguousarray(pts)
        self._event = None  # with_solout replaces the whole solout (src/ivp.rs:656-659)
        return self

    def event(self, ev):
        """IVP::event (src/ivp.rs:662-678): wrap the current output (final state
        or t_eval) with event detection."""
        self._event = (DE_SOLOUT_EVENT, ev)
        return self

    def crossing(self, component_idx, threshold, direction=CrossingDirection.Both):
        """IVP::crossing (src/ivp.rs:682-692): CrossingSolout replaces the output."""
        self._t_eval = None
        self._points = (DE_POINTS_NONE, 0, 0.0)
        self._event = (DE_SOLOUT_CROSSING,
                       Event.component(component_idx, threshold, EventConfig(direction)))
        return self

    def hyperplane_crossing(self, point, normal, direction=CrossingDirection.Both):
        """IVP::hyperplane_crossing (src/ivp.rs:695-705) with the identity
        extractor: the hyperplane lies in the state space."""
        self._t_eval = None
        self._points = (DE_POINTS_NONE, 0, 0.0)
        self._event = (DE_SOLOUT_HYPERPLANE,
                       Event.plane(point, normal, EventConfig(direction)))
        return self

    def options(self, schedule=None, mapping=None, group_size=None, fast_math=None,
                max_events=None):
        if max_events is not None:
            self.max_events = int(max_events)
        if schedule is not None:
            self.schedule = schedule
        if mapping is not None:
            self.mapping = mapping
        if group_size is not None:
            self.group_size = group_size
        if fast_math is not None:
            self.flags = (self.flags | DE_FLAG_FAST_MATH) if fast_math else (
                self.flags & ~DE_FLAG_FAST_MATH)
        return self

    def batch(self, points_offset=None):
        """The de_ode_batch descriptor (host pointers) plus result buffers;
        with a point output, `points_offset` (n + 1 slot offsets) sizes it."""
        s = self.system
        b = ode_defaults()
        b.family = s.family
        b.method = self._method.method_id
        b.n = self.n
        keep = [self.y0]
        if isinstance(s, SpringChain):
            b.dim = 0
            b.sizes = s.sizes.ctypes.data_as(C.POINTER(C.c_int32))
            b.state_offset = s.state_offset.ctypes.data_as(C.POINTER(C.c_int64))
            b.param_offset = s.param_offset.ctypes.data_as(C.POINTER(C.c_int64))
            params = s.params_packed
            state_len = int(s.state_offset[-1])
            keep += [s.sizes, s.state_offset, s.param_offset]
        else:
            b.dim = s.dim
            params = s.pack(self.n)
            state_len = s.dim * self.n
        keep.append(params)
        b.params = _ptr(params)
        b.y0 = _ptr(self.y0)
        for name in ("t0", "tf"):
            v = np.asarray(getattr(self, name), dtype=np.float64)
            if v.ndim == 0:
                setattr(b, name + "_scalar", float(v))
            else:
                v = np.ascontiguousarray(v)
                keep.append(v)
                setattr(b, name, _ptr(v))
                setattr(b, name + "_scalar", float(v[0]))
        comp = {}
        for k, v in self._method.settings.items():
            if k in ("rtol", "atol") and np.ndim(v) > 0:
                vec = np.asarray(v, dtype=np.float64).reshape(-1)
                if isinstance(s, SpringChain) or len(vec) != s.dim:
                    raise ValueError("a %s vector needs one value per component of a fixed-size "
                                     "system; give spring chains %s_comp in the state layout"
                                     % (k, k))
                comp[k + "_comp"] = np.repeat(vec[:, None], self.n, axis=1).reshape(-1)
                continue
            setattr(b, k, v)
        for name in ("rtol_comp", "atol_comp"):
            v = getattr(self, name)
            v = comp.get(name) if v is None else v
            if v is not None:
                v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                if len(v) != state_len:
                    raise ValueError("%s needs %d values (one per state component)"
                                     % (name, state_len))
                keep.append(v)
                setattr(b, name, _ptr(v))
        for name in ("rtol_traj", "atol_traj"):
            v = getattr(self, name)
            if v is not None:
                v = np.ascontiguousarray(v, dtype=np.float64)
                keep.append(v)
                setattr(b, name, _ptr(v))
        n_eval = 0 if self._t_eval is None else len(self._t_eval)
        b.n_eval = n_eval
        if n_eval:
            keep.append(self._t_eval)
            b.t_eval = _ptr(self._t_eval)
        b.schedule = self.schedule
        b.mapping = self.mapping
        b.group_size = self.group_size
        b.flags = self.flags | DE_FLAG_HOST_MEMORY
        cap = None
        if self._event is not None:
            kind, ev = self._event
            cap = self.max_events
            b.event_solout = kind
            b.event_fn = ev.fn
            b.event_component = ev.comp
            b.event_direction = ev.cfg._direction
            b.event_terminate = 0 if ev.cfg._terminate is None else ev.cfg._terminate
            b.max_events = cap
            if np.ndim(ev.value) > 0:
                v = np.ascontiguousarray(ev.value, dtype=np.float64).reshape(-1)
                if len(v) != self.n:
                    raise ValueError("a per-trajectory event value needs %d entries" % self.n)
                keep.append(v)
                b.event_value_traj = _ptr(v)
            else:
                b.event_value = float(ev.value)
            for name in ("point", "normal"):
                v = getattr(ev, name)
                if v is not None:
                    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                    keep.append(v)
                    setattr(b, "event_" + name, _ptr(v))
        res = Solution(self.n, state_len, n_eval, cap)
        if isinstance(s, SpringChain):
            res.state_offset = s.state_offset
        mode, pn, pdt = self._points
        if mode != DE_POINTS_NONE and points_offset is not None:
            b.points, b.points_n, b.points_dt = mode, pn, pdt
            res.set_points(np.ascontiguousarray(points_offset, dtype=np.int64),
                           2 * s.sizes.astype(np.int64) if isinstance(s, SpringChain)
                           else s.dim)
        return b, res, keep

    def points_offsets(self, accepted=None):
        """Slot offsets (n + 1) that hold every point of the point output:
        accepted + 1 per trajectory (DefaultSolout), n * accepted + 1
        (DenseSolout), floor(|tf - t0| / dt) + 3 (EvenSolout)."""
        mode, pn, pdt = self._points
        if mode == DE_POINTS_EVEN:
            span = np.abs(np.broadcast_to(np.asarray(self.tf, dtype=np.float64), (self.n,)) -
                          np.broadcast_to(np.asarray(self.t0, dtype=np.float64), (self.n,)))
            cnt = np.floor(span / pdt).astype(np.int64) + 3
        else:
            acc = np.asarray(accepted, dtype=np.int64)
            cnt = acc + 1 if mode == DE_POINTS_STEPS else pn * acc + 1
        off = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(cnt, out=off[1:])
        return off

    def solve(self, stream=None, solver=None):
        """The ensemble's solve.  With a point output from all_steps() or
        dense(n), a first solve counts each trajectory's accepted steps (the
        results are deterministic) and a second one writes the points."""
        solver = solver or (lambda e, off: e._solve_once(stream, off))
        mode = self._points[0]
        offsets = None
        if mode in (DE_POINTS_STEPS, DE_POINTS_DENSE):
            saved = self._points
            self._points = (DE_POINTS_NONE, 0, 0.0)
            try:
                first = solver(self, None)
            finally:
                self._points = saved
            offsets = self.points_offsets(first.accepted)
        elif mode == DE_POINTS_EVEN:
            offsets = self.points_offsets()
        return solver(self, offsets)

    def _solve_once(self, stream, points_offset):
        b, res, keep = self.batch(points_offset)
        _abi.check(lib().de_ode_batch_solve(C.byref(b), C.byref(res.struct()), stream))
        res._solved_by = (b, keep)  # the descriptor, for the CSV writer
        return res


def _write_file(filename, data):
    parent = os.path.dirname(filename)
    if parent:
        os.makedirs(parent, exist_ok=True)
    with open(filename, "wb") as f:
        f.write(data)


class Solution:
    """Per-trajectory results (Solution / Error of each reference solve)."""

    def __init__(self, n, state_len, n_eval, max_events=None):
        self.n = n
        self.max_events = max_events or 0
        self.y = np.full(state_len, np.nan)
        self.t = np.full(n, np.nan)
        self.accepted = np.zeros(n, dtype=np.int64)
        self.rejected = np.zeros(n, dtype=np.int64)
        self.evals = np.zeros(n, dtype=np.int64)
        self.status = np.full(n, -1, dtype=np.int32)
        self.dense = np.full(state_len * n_eval, np.nan) if n_eval else None
        self.n_dense = np.zeros(n, dtype=np.int32) if n_eval else None
        # Event points: event_t[k * n + i], event_y[(k * dim + c) * n + i].
        self.event_t = np.full(max_events * n, np.nan) if max_events else None
        self.event_y = np.full(max_events * state_len, np.nan) if max_events else None
        self.n_events = np.zeros(n, dtype=np.int32) if max_events is not None else None
        self.points_offset = self.points_t = self.points_y = self.n_points = None
        self.points_ypos = None
        # Spring chains: trajectory i's state is y[state_offset[i]:state_offset[i + 1]].
        self.state_offset = None

    def set_points(self, offsets, dim):
        """Buffers of a point output: slots offsets[i] .. offsets[i + 1] of
        trajectory i, states of `dim` components.  With per-trajectory state
        lengths (an array: spring chains) each trajectory's slot states are
        kept back to back from points_ypos[i]."""
        self.points_offset = offsets
        cap = int(offsets[-1])
        self.points_t = np.full(max(cap, 1), np.nan)
        if np.ndim(dim):
            self.points_ypos = np.zeros(self.n + 1, dtype=np.int64)
            np.cumsum(np.diff(offsets) * np.asarray(dim, dtype=np.int64),
                      out=self.points_ypos[1:])
            ylen = int(self.points_ypos[-1])
        else:
            ylen = cap * dim
        self.points_y = np.full(max(ylen, 1), np.nan)
        self.n_points = np.zeros(self.n, dtype=np.int64)

    def _dim(self, i):
        if self.state_offset is not None:
            return int(self.state_offset[i + 1] - self.state_offset[i])
        return len(self.y) // self.n

    def final_state(self, i):
        """The final state of trajectory i (the state at failure for an error status)."""
        if self.state_offset is not None:
            return self.y[self.state_offset[i]:self.state_offset[i + 1]].copy()
        return self.y.reshape(self._dim(i), self.n)[:, i].copy()

    def points(self, i):
        """(t[k], y[k, dim]) of trajectory i's Solution points (the kept ones)."""
        dim = self._dim(i)
        lo, hi = int(self.points_offset[i]), int(self.points_offset[i + 1])
        k = min(int(self.n_points[i]), hi - lo)
        y0 = int(self.points_ypos[i]) if self.points_ypos is not None else lo * dim
        return (self.points_t[lo:lo + k].copy(),
                self.points_y[y0:y0 + k * dim].reshape(k, dim).copy())

    def events(self, i):
        """(t[k], y[k, dim]) of the event points kept for trajectory i."""
        k = min(int(self.n_events[i]), self.max_events)
        dim = self._dim(i)
        if k == 0:
            return np.empty(0), np.empty((0, dim))
        t = self.event_t.reshape(self.max_events, self.n)[:k, i]
        if self.state_offset is not None:
            # spring chains: point k at event_y[max_events * state_offset[i] + k * dim + c]
            lo = self.max_events * int(self.state_offset[i])
            y = self.event_y[lo:lo + k * dim].reshape(k, dim)
        else:
            y = self.event_y.reshape(self.max_events, dim, self.n)[:k, :, i]
        return t.copy(), y.copy()

    _CSV_SOURCES = {"final": DE_CSV_FINAL, "t_eval": DE_CSV_TEVAL, "events": DE_CSV_EVENTS,
                    "points": DE_CSV_POINTS}

    def csv(self, source=None):
        """Solution::to_csv of every trajectory (src/solution.rs:190-230),
        formatted on the GPU: (bytes, offsets), trajectory i's document being
        bytes[offsets[i]:offsets[i + 1]].  `source`: "points" (the point output
        of all_steps(), dense(n) or even(dt): the reference's whole Solution;
        the default for such a solve), "t_eval" (the t_eval points; the
        default when the solve had a t_eval output), "final" (the final state;
        the default otherwise) or "events" (the event points)."""
        if not hasattr(self, "_solved_by"):
            raise RuntimeError("csv() needs a Solution returned by OdeEnsemble.solve()")
        if source is None:
            source = ("points" if self.n_points is not None else
                      "t_eval" if self.n_dense is not None else "final")
        src = self._CSV_SOURCES[source]
        b, keep = self._solved_by
        r = self.struct()
        offsets = np.zeros(self.n + 1, dtype=np.int64)
        _abi.check(lib().de_ode_csv_offsets(C.byref(b), C.byref(r), src, offsets.ctypes.data,
                                            None))
        out = np.zeros(max(1, int(offsets[-1])), dtype=np.uint8)
        _abi.check(lib().de_ode_csv_write(C.byref(b), C.byref(r), src, offsets.ctypes.data,
                                          out.ctypes.data, None))
        return out[:int(offsets[-1])].tobytes(), offsets

    def to_csv(self, filename, i=0, source=None):
        """Solution::to_csv(filename) of trajectory i: creates the parent
        directories, then writes the document (an empty file for a trajectory
        without points, where the reference panics after creating it)."""
        data, offsets = self.csv(source)
        _write_file(filename, data[offsets[i]:offsets[i + 1]])

    def to_dataframe(self, i=0):
        """Solution::to_polars of trajectory i (src/solution.rs:239-280): columns
        t, y0 .. y<dim-1> over its Solution points (the point output's, else
        the t_eval points produced, else the final state), as a pandas
        DataFrame (polars is not available here)."""
        import pandas as pd
        dim = self._dim(i)
        if self.n_points is not None:
            t, y = self.points(i)
        elif self.n_dense is not None:
            # TEvalSolout's points: index 0 when it equals t0, then the run of
            # points beyond t0 (src/solout/t_eval.rs:100-135), n_dense[i] in all.
            te = self._t_eval_values()
            b, _ = self._solved_by
            t0 = (np.ctypeslib.as_array(C.cast(b.t0, C.POINTER(C.c_double)), (self.n,))[i]
                  if b.t0 else b.t0_scalar)
            tf = (np.ctypeslib.as_array(C.cast(b.tf, C.POINTER(C.c_double)), (self.n,))[i]
                  if b.tf else b.tf_scalar)
            head = int(len(te) > 0 and te[0] == t0)
            beyond = (te > t0) if tf > t0 else (te < t0)
            run = int(np.argmax(beyond)) if beyond.any() else len(te)
            idx = ([0] if head else []) + list(range(run, len(te)))
            idx = idx[:int(self.n_dense[i])]
            if self.state_offset is not None:
                ne = len(te)
                so = self.state_offset
                rows = self.dense[ne * so[i]:ne * so[i + 1]].reshape(ne, dim)
            else:
                rows = self.dense.reshape(-1, dim, self.n)[:, :, i]
            t, y = te[idx], rows[idx]
        else:
            t = self.t[i:i + 1]
            y = self.final_state(i)[None, :]
        cols = {"t": np.asarray(t, dtype=np.float64)}
        for c in range(dim):
            cols["y%d" % c] = np.asarray(y[:, c], dtype=np.float64)
        return pd.DataFrame(cols)

    def _t_eval_values(self):
        b, keep = self._solved_by
        return np.ctypeslib.as_array(C.cast(b.t_eval, C.POINTER(C.c_double)), (b.n_eval,))

    def to_csv_all(self, pattern, source=None):
        """One to_csv file per trajectory: pattern.format(i=i)."""
        data, offsets = self.csv(source)
        for i in range(self.n):
            _write_file(pattern.format(i=i), data[offsets[i]:offsets[i + 1]])

    def struct(self):
        r = OdeResult()
        r.y_final = _ptr(self.y)
        r.t_final = _ptr(self.t)
        r.accepted = _ptr(self.accepted)
        r.rejected = _ptr(self.rejected)
        r.evals = _ptr(self.evals)
        r.status = _ptr(self.status)
        r.dense = _ptr(self.dense)
        r.n_dense = _ptr(self.n_dense)
        r.event_t = _ptr(self.event_t)
        r.event_y = _ptr(self.event_y)
        r.n_events = _ptr(self.n_events)
        r.points_offset = _ptr(self.points_offset)
        r.points_t = _ptr(self.points_t)
        r.points_y = _ptr(self.points_y)
        r.n_points = _ptr(self.n_points)
        self._struct = r
        return r


class SdeEnsemble:
    """IVP::sde for an ensemble of paths."""

    def __init__(self, system, t0, tf, y0, noise=DE_NOISE_PHILOX, seed=0, noise_table=None,
                 noise_constant=0.0, path_base=0):
        self.system = system
        self.t0 = t0
        self.tf = tf
        y0 = np.asarray(y0, dtype=np.float64)
        self.n = y0.size // system.dim
        self.y0 = np.ascontiguousarray(y0).reshape(-1)
        self.noise = noise
        self.seed = seed
        self.noise_table = None if noise_table is None else np.ascontiguousarray(
            noise_table, dtype=np.float64).reshape(-1)
        self.noise_constant = noise_constant
        self.path_base = path_base
        self._method = EulerMaruyama(0.0)
        self.path_stride = 0
        self.flags = 0

    def method(self, m):
        self._method = m
        return self

    def record_every(self, stride):
        self.path_stride = stride
        return self

    def batch(self):
        s = self.system
        b = sde_defaults()
        b.family = s.family
        b.scheme = self._method.scheme
        b.n = self.n
        b.dim = s.dim
        b.noise = self.noise
        b.params = _ptr(s.params)
        b.y0 = _ptr(self.y0)
        b.t0 = self.t0
        b.tf = self.tf
        for k, v in self._method.settings.items():
            setattr(b, k, v)
        b.seed = self.seed
        b.path_base = self.path_base
        keep = [s.params, self.y0]
        if self.noise_table is not None:
            keep.append(self.noise_table)
            b.noise_table = _ptr(self.noise_table)
            b.noise_steps = len(self.noise_table) // s.dim
        b.noise_constant = self.noise_constant
        b.path_stride = self.path_stride
        b.flags = self.flags | DE_FLAG_HOST_MEMORY
        n_rec = 0
        if self.path_stride > 0:
            n_rec = int(b.max_steps) // self.path_stride + 1
        res = SdeSolution(self.n, s.dim, n_rec)
        return b, res, keep

    def solve(self, stream=None):
        b, res, keep = self.batch()
        _abi.check(lib().de_sde_batch_solve(C.byref(b), C.byref(res.struct()), stream))
        return res


class SdeSolution:
    def __init__(self, n, dim, n_rec):
        self.y = np.full(dim * n, np.nan)
        self.t = np.full(n, np.nan)
        self.steps = np.zeros(n, dtype=np.int64)
        self.evals = np.zeros(n, dtype=np.int64)
        self.status = np.full(n, -1, dtype=np.int32)
        self.path = np.full(n_rec * dim * n, np.nan) if n_rec else None
        self.path_t = np.full(n_rec, np.nan) if n_rec else None

    def struct(self):
        r = SdeResult()
        r.y_final = _ptr(self.y)
        r.t_final = _ptr(self.t)
        r.steps = _ptr(self.steps)
        r.evals = _ptr(self.evals)
        r.status = _ptr(self.status)
        r.path = _ptr(self.path)
        r.path_t = _ptr(self.path_t)
        self._struct = r
        return r


class IVP:
    """IVP::ode / IVP::sde (src/ivp.rs:277, 502) for ensembles."""

    @staticmethod
    def ode(system, t0, tf, y0, n=None):
        return OdeEnsemble(system, t0, tf, y0, n=n)

    @staticmethod
    def sde(system, t0, tf, y0, **noise):
        return SdeEnsemble(system, t0, tf, y0, **noise)


# ---------------------------------------------------------------- workloads --
def gen_lorenz(n, seed=0x5EED2026):
    p = np.empty(3 * n)
    y0 = np.empty(3 * n)
    lib().de_gen_lorenz(seed, n, _ptr(p), _ptr(y0))
    return p.reshape(3, n), y0.reshape(3, n)


def gen_vdp(n, seed=0x5EED2026):
    p = np.empty(n)
    y0 = np.empty(2 * n)
    lib().de_gen_vdp(seed, n, _ptr(p), _ptr(y0))
    return p.reshape(1, n), y0.reshape(2, n)


def gen_spring(n, seed=0x5EED2026, sizes=None):
    if sizes is None:
        sizes = np.empty(n, dtype=np.int32)
        lib().de_gen_spring_sizes(seed, n, _ptr(sizes))
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    so = np.empty(n + 1, dtype=np.int64)
    po = np.empty(n + 1, dtype=np.int64)
    lib().de_spring_offsets(n, _ptr(sizes), _ptr(so), _ptr(po))
    params = np.empty(int(po[-1]))
    y0 = np.empty(int(so[-1]))
    lib().de_gen_spring(seed, n, _ptr(sizes), _ptr(so), _ptr(po), _ptr(params), _ptr(y0))
    return SpringChain(sizes, params), y0


def gen_ou(n, dim=8, seed=0x5EED2026):
    p = np.empty(3 * dim * n)
    y0 = np.empty(dim * n)
    lib().de_gen_ou(seed, n, dim, _ptr(p), _ptr(y0))
    return OrnsteinUhlenbeck(dim, p), y0


def gen_gbm(n, mu=0.05, sigma=0.2, s0=100.0):
    p = np.empty(2 * n)
    y0 = np.empty(n)
    lib().de_gen_gbm(n, mu, sigma, s0, _ptr(p), _ptr(y0))
    return LinearDiagonal(1, p), y0


def poisson_requests(rate, seconds, seed=0x5EED2026, big=65536):
    """A seeded Poisson request stream: (arrival time s, family, size,
    OdeEnsemble) per request, arrivals at `rate` per second over `seconds`.
    Each request: one family (0 Lorenz, 1 Van der Pol mu ~ LogUniform[0.1, 10],
    2 spring chains N in [4, 256]), ensemble size log-uniform in [1, big], a
    slice of one pool of `big` generated systems, rtol in {1e-6, 1e-8} with
    atol = rtol / 100, t_final ~ U[1, 50], DOPRI5."""
    rng = np.random.default_rng(seed)
    lp, ly0 = gen_lorenz(big, seed=seed)
    vp, vy0 = gen_vdp(big, seed=seed)
    sp, sy0 = gen_spring(big, seed=seed)
    n_req = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for k in range(n_req):
        fam = int(rng.integers(0, 3))
        n = int(np.exp(rng.uniform(0, np.log(big))))
        a = int(rng.integers(0, big - n + 1))
        rtol = float(rng.choice([1e-6, 1e-8]))
        tf = float(rng.uniform(1.0, 50.0))
        m = ExplicitRungeKutta.dopri5().rtol(rtol).atol(rtol * 1e-2).max_steps(1000000)
        if fam == 0:
            e = IVP.ode(Lorenz(*lp[:, a:a + n]), 0.0, tf, ly0[:, a:a + n]).method(m)
        elif fam == 1:
            e = IVP.ode(VanDerPol(*vp[:, a:a + n]), 0.0, tf, vy0[:, a:a + n]).method(m)
        else:
            sizes = sp.sizes[a:a + n]
            params = sp.params_packed[sp.param_offset[a]:sp.param_offset[a + n]]
            y0 = sy0[sp.state_offset[a]:sp.state_offset[a + n]]
            e = IVP.ode(SpringChain(sizes, params), 0.0, tf, y0).method(m)
        out.append((float(arrivals[k]), fam, n, e))
    return out


# ----------------------------------------------------------------- serving --
DE_SERVER_GRAPH = 1


class Server:
    """de_server_*: continuous batching of ensemble solves (host memory in/out)."""

    def __init__(self, hbm_budget_bytes, graph=False):
        h = C.c_void_p()
        _abi.check(lib().de_server_create(int(hbm_budget_bytes), DE_SERVER_GRAPH if graph else 0,
                                          C.byref(h)))
        self.h = h
        self._pending = {}

    def submit(self, ensemble):
        b, res, keep = ensemble.batch()
        t = C.c_uint64()
        _abi.check(lib().de_server_submit(self.h, C.byref(b), C.byref(t)))
        self._pending[t.value] = res
        return t.value

    def poll(self, ticket, wait=True):
        res = self._pending[ticket]
        rc = lib().de_server_poll(self.h, ticket, C.byref(res.struct()), 1 if wait else 0)
        if rc < 0:
            _abi.check(rc)
        if rc == 0:
            return None
        del self._pending[ticket]
        return res

    def latency(self, ticket):
        return lib().de_server_latency(self.h, ticket)

    def stats(self):
        """(iterations, range-table restarts) of the scheduler."""
        it, rr = C.c_int64(), C.c_int64()
        _abi.check(lib().de_server_stats(self.h, C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def pools(self):
        """(pools made on first use that are made now, releases of idle ones)."""
        m, r = C.c_int32(), C.c_int64()
        _abi.check(lib().de_server_pools(self.h, C.byref(m), C.byref(r)))
        return m.value, r.value

    def timing(self):
        """Scheduler seconds spent admitting, iterating and retiring."""
        a, i, r = C.c_double(), C.c_double(), C.c_double()
        _abi.check(lib().de_server_timing(self.h, C.byref(a), C.byref(i), C.byref(r)))
        return a.value, i.value, r.value

    def close(self):
        if self.h:
            lib().de_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
