"""Spring-chain ensembles (uneven sizes, dense output) on the GPU.

DE_MAP_LANE (one thread per system) must be bit-identical to the oracle.  The
register-resident group mappings sum the error norm as a butterfly across the
lanes of a group instead of the reference's sequential loop, so their results
are held to the stated bound: final state within 1e-10 relative error, step
counts equal on at least 99.9% of systems.
"""
import os

import numpy as np
import pytest

import deamd
from deamd import (DE_MAP_AUTO, DE_MAP_GROUP, DE_MAP_LANE, DE_STATUS_COMPLETE, CrossingDirection,
                   Event, EventConfig, ExplicitRungeKutta, IVP)

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
PTS = np.linspace(0.0, 10.0, 100)


def chains(n, seed):
    sysm, y0 = deamd.gen_spring(n, seed=seed)
    ens = IVP.ode(sysm, 0.0, 10.0, y0).t_eval(PTS).method(
        ExplicitRungeKutta.dopri5().rtol(1e-6).atol(1e-9).max_steps(1000000))
    return sysm, ens


@pytest.fixture(scope="module")
def case(orc):
    sysm, ens = chains(2048, seed=17)
    return sysm, ens, orc.solve_ode(ens, threads=THREADS)


def rel_state_error(sysm, got, want):
    errs = []
    for i in range(len(sysm.sizes)):
        a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
        ref = np.abs(want.y[a:b]).max()
        errs.append(np.abs(got.y[a:b] - want.y[a:b]).max() / max(ref, 1e-300))
    return np.array(errs)


def test_lane_mapping_bit_exact(case):
    sysm, ens, want = case
    got = ens.options(mapping=DE_MAP_LANE).solve()
    assert (got.status == DE_STATUS_COMPLETE).all()
    for f in ("status", "accepted", "rejected", "evals", "t", "y", "n_dense"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
    assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))


@pytest.mark.parametrize("mapping,group", [(DE_MAP_GROUP, 1), (DE_MAP_GROUP, 4),
                                           (DE_MAP_GROUP, 16), (DE_MAP_GROUP, 64),
                                           (DE_MAP_AUTO, 0)])
def test_group_mappings_within_bound(case, mapping, group):
    sysm, ens, want = case
    got = ens.options(mapping=mapping, group_size=group).solve()
    assert (got.status == DE_STATUS_COMPLETE).all()
    err = rel_state_error(sysm, got, want)
    assert err.max() <= 1e-10, err.max()
    same_counts = (got.accepted == want.accepted) & (got.rejected == want.rejected)
    assert same_counts.mean() >= 0.999, same_counts.mean()
    assert np.array_equal(got.n_dense, want.n_dense)
    dense_err = np.nanmax(np.abs(got.dense - want.dense))
    assert dense_err <= 1e-9, dense_err


def test_autotuner_reports_its_choices(case):
    sysm, ens, want = case
    ens.options(mapping=DE_MAP_AUTO).solve()
    import ctypes as C
    lib = deamd.lib()
    caps = sorted({c for c in (4, 8, 16, 32, 64, 128, 256)
                   if ((sysm.sizes <= c) & (sysm.sizes > c // 2 if c > 4 else True)).any()})
    for cap in caps:
        g, c, b = C.c_int32(), C.c_int32(), C.c_int32()
        assert lib.de_spring_tuned(cap, C.byref(g), C.byref(c), C.byref(b)) == 0
        assert g.value * c.value == cap and b.value in (64, 256)


def test_mixed_sizes_results_do_not_depend_on_batch_composition(case):
    # A system's result under the group mapping is a function of the system
    # alone: solving a subset gives the same bits as inside the full batch.
    sysm, ens, want = case
    full = ens.options(mapping=DE_MAP_GROUP, group_size=16).solve()
    idx = np.arange(0, len(sysm.sizes), 7)
    sub_sizes = sysm.sizes[idx]
    sub_params = np.concatenate([sysm.params_packed[sysm.param_offset[i]:sysm.param_offset[i + 1]]
                                 for i in idx])
    sub_y0 = np.concatenate([ens.y0[sysm.state_offset[i]:sysm.state_offset[i + 1]] for i in idx])
    sub = deamd.SpringChain(sub_sizes, sub_params)
    sens = IVP.ode(sub, 0.0, 10.0, sub_y0).t_eval(PTS).method(
        ExplicitRungeKutta.dopri5().rtol(1e-6).atol(1e-9).max_steps(1000000)) \
        .options(mapping=DE_MAP_GROUP, group_size=16)
    part = sens.solve()
    for k, i in enumerate(idx):
        a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
        sa, sb = sub.state_offset[k], sub.state_offset[k + 1]
        assert np.array_equal(part.y[sa:sb], full.y[a:b])
        assert part.accepted[k] == full.accepted[i]


def test_per_component_tolerances(orc):
    # Tolerance::Vector on ragged systems: one (rtol, atol) per state component
    # in the state layout; velocities held 100x tighter than displacements.
    sysm, ens = chains(512, seed=23)
    rt = np.empty(len(ens.y0))
    at = np.empty(len(ens.y0))
    for i, nm in enumerate(sysm.sizes):
        a = sysm.state_offset[i]
        rt[a:a + nm], rt[a + nm:a + 2 * nm] = 1e-6, 1e-8
        at[a:a + nm], at[a + nm:a + 2 * nm] = 1e-9, 1e-11
    ens.rtol_comp, ens.atol_comp = rt, at
    want = orc.solve_ode(ens, threads=THREADS)
    scalar = orc.solve_ode(chains(512, seed=23)[1], threads=THREADS)
    assert (want.accepted > scalar.accepted).mean() > 0.5   # the tighter components bind
    got = ens.options(mapping=DE_MAP_LANE).solve()
    for f in ("status", "accepted", "rejected", "evals", "t", "y", "n_dense"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    got = ens.options(mapping=DE_MAP_GROUP, group_size=16).solve()
    assert rel_state_error(sysm, got, want).max() <= 1e-10
    same_counts = (got.accepted == want.accepted) & (got.rejected == want.rejected)
    assert same_counts.mean() >= 0.999


def test_dense_rows_of_unfinished_systems_are_nan(orc):
    # max_steps stops the solves part-way: points never reached stay NaN (the
    # reference pushes nothing for them), in both mappings.
    sysm, y0 = deamd.gen_spring(256, seed=31)
    ens = IVP.ode(sysm, 0.0, 10.0, y0).t_eval(PTS).method(
        ExplicitRungeKutta.dopri5().rtol(1e-6).atol(1e-9).max_steps(60))
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == deamd.DE_STATUS_MAX_STEPS).mean() > 0.5
    assert (want.n_dense < len(PTS)).mean() > 0.5
    got = ens.options(mapping=DE_MAP_LANE).solve()
    for f in ("status", "accepted", "rejected", "evals", "t", "y", "n_dense"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
    assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))
    got = ens.options(mapping=DE_MAP_GROUP, group_size=8).solve()
    assert np.array_equal(got.n_dense, want.n_dense)
    assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))


@pytest.mark.parametrize("method", ["dop853", "rk4", "euler"])
@pytest.mark.parametrize("dense", [False, True])
def test_other_methods_bit_exact(orc, method, dense):
    # DOP853, RK4 and Euler on spring chains run one lane per system (whatever
    # the mapping asked for) and equal the oracle bit for bit, with and
    # without t_eval output, under automatic and given initial steps.
    sysm, y0 = deamd.gen_spring(600, seed=31)
    m = {"dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-7).atol(1e-9).max_steps(200000),
         "rk4": lambda: ExplicitRungeKutta.rk4(0.01),
         "euler": lambda: ExplicitRungeKutta.euler(0.002).max_steps(100000)}[method]()
    ens = IVP.ode(sysm, 0.0, 3.0, y0).method(m)
    if dense:
        ens = ens.t_eval(np.linspace(-0.5, 3.5, 41))   # points before t0 and past tf
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).all()
    for mapping in (DE_MAP_AUTO, DE_MAP_LANE):
        got = ens.options(mapping=mapping).solve()
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), (f, mapping)
        if dense:
            assert np.array_equal(got.n_dense, want.n_dense)
            assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
            assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))


def test_dop853_spring_errors_and_stiffness(orc):
    # Status paths of the one-lane DOP853 solver: max_steps, a step size that
    # underflows h_min's bound (BadInput), per-trajectory tolerances.
    sysm, y0 = deamd.gen_spring(64, seed=32)
    n = len(sysm.sizes)
    for mk in (lambda e: e.method(ExplicitRungeKutta.dop853().max_steps(3)),
               lambda e: e.method(ExplicitRungeKutta.dop853().h0(-0.1)),
               lambda e: e.method(ExplicitRungeKutta.dop853().h_max(0.05))):
        ens = mk(IVP.ode(sysm, 0.0, 2.0, y0))
        ens.rtol_traj = np.geomspace(1e-9, 1e-5, n)
        ens.atol_traj = np.geomspace(1e-11, 1e-7, n)
        want = orc.solve_ode(ens, threads=THREADS)
        got = ens.solve()
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), f


def test_backward_t_eval_ends_out_of_bounds(orc):
    # A backward solve with t_eval points ends OutOfBounds at the first point
    # that needs the interpolant (the reference's interpolants reject t outside
    # [t_prev, t] of a backward step and TEvalSolout unwraps): every mapping
    # and method stops there with the oracle's status, counts and rows.
    sysm, y0 = deamd.gen_spring(300, seed=33)
    pts = np.linspace(0.0, -2.0, 9)
    for m, mappings in ((ExplicitRungeKutta.dopri5, (DE_MAP_LANE, DE_MAP_GROUP, DE_MAP_AUTO)),
                        (ExplicitRungeKutta.dop853, (DE_MAP_LANE,)),
                        (lambda: ExplicitRungeKutta.rk4(-0.01), (DE_MAP_LANE,))):
        ens = IVP.ode(sysm, 0.0, -2.0, y0).method(m()).t_eval(pts)
        want = orc.solve_ode(ens, threads=THREADS)
        assert (want.status == deamd.DE_STATUS_OUT_OF_BOUNDS).all()
        for mapping in mappings:
            got = ens.options(mapping=mapping, group_size=16 if mapping == DE_MAP_GROUP else 0).solve()
            for f in ("status", "accepted", "rejected", "n_dense"):
                assert np.array_equal(getattr(got, f), getattr(want, f)), (f, mapping)
            if mapping == DE_MAP_LANE:
                assert np.array_equal(got.y, want.y) and np.array_equal(got.t, want.t)
                assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
                assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))
            else:
                assert (rel_state_error(sysm, got, want) <= 1e-10).all()


FAMILY = ["rkf45", "cash_karp", "rkv655e", "rkv656e", "rkv766e", "rkv767e", "rkv877e", "rkv878e",
          "rkv988e", "rkv989e"]


@pytest.mark.parametrize("name", FAMILY)
def test_embedded_family_equals_oracle(orc, name):
    # The embedded family on spring chains (one lane per system) equals the
    # oracle as IEEE values (a zero's sign may differ), with and without t_eval
    # output, automatic initial step.
    sysm, y0 = deamd.gen_spring(200, seed=34)
    for dense in (False, True):
        ens = IVP.ode(sysm, 0.0, 2.0, y0).method(
            getattr(ExplicitRungeKutta, name)().rtol(1e-7).atol(1e-9).max_steps(200000))
        if dense:
            ens = ens.t_eval(np.linspace(-0.25, 2.25, 21))
        want = orc.solve_ode(ens, threads=THREADS)
        assert (want.status == DE_STATUS_COMPLETE).all()
        got = ens.solve()
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), (f, dense)
        if dense:
            assert np.array_equal(got.n_dense, want.n_dense)
            assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
            assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))


def same_points(got, want):
    """The kept points of every trajectory, as bits (NaN where the oracle has NaN)."""
    assert np.array_equal(got.n_points, want.n_points)
    assert np.array_equal(got.points_offset, want.points_offset)
    for i in range(got.n):
        gt, gy = got.points(i)
        wt, wy = want.points(i)
        assert np.array_equal(gt, wt), i
        assert np.array_equal(np.isnan(gy), np.isnan(wy)), i
        assert np.array_equal(np.nan_to_num(gy), np.nan_to_num(wy)), i


def points_within_bound(sysm, got, want, even=False):
    # The group mappings' butterfly norm rounds differently from the
    # reference's sequential sum, so step sizes differ in their last bits and
    # the step ends drift apart by up to ~1e-10 over a solve.  Checked: equal
    # statuses; step counts equal on at least 99.9% of systems; where they are
    # and the solve completed, the final state within 1e-10 relative error
    # (the mappings' bound), the same number of points, the last one the
    # solve's own final state, the even grid's times exactly the oracle's
    # with states within 1e-9, and the step ends (and their subdivisions)
    # within 1e-8 of the oracle's times with states within 1e-7 (a point at a
    # slightly different time moves by |y'| times that difference).  A solve
    # cut short (a backward one at its first interpolation) is held to the
    # counts only: one ended 1.6e-9 apart after six steps with equal counts.
    assert (got.status == want.status).all()
    same = (got.accepted == want.accepted) & (got.rejected == want.rejected)
    assert same.mean() >= 0.999
    assert np.array_equal(got.n_points[same], want.n_points[same])
    done = same & (got.status == DE_STATUS_COMPLETE)
    assert (rel_state_error(sysm, got, want)[done] <= 1e-10).all()
    t_tol, y_tol = (0.0, 1e-9) if even else (1e-8, 1e-7)
    for i in np.nonzero(done)[0]:
        gt, gy = got.points(i)
        wt, wy = want.points(i)
        assert gt[-1] == got.t[i] and np.array_equal(gy[-1], got.final_state(i))
        assert np.abs(gt - wt).max(initial=0.0) <= t_tol * max(1.0, np.abs(wt).max(initial=0.0))
        scale = max(np.abs(wy).max(initial=0.0), 1e-300)
        assert np.abs(gy - wy).max(initial=0.0) <= y_tol * scale


@pytest.mark.parametrize("out", ["steps", "dense", "even"])
@pytest.mark.parametrize("method", ["dopri5", "dop853", "rk4", "rkv878e"])
def test_point_outputs_bit_exact(orc, method, out):
    # DefaultSolout, IVP::dense(3) and IVP::even(dt) on spring chains (DOPRI5
    # on the register group mappings or one lane per system, the other
    # methods one lane per system; each system's point states back to back)
    # equal the oracle bit for bit on one lane per system, and within the
    # group mappings' bound there.
    sysm, y0 = deamd.gen_spring(400, seed=35)
    m = {"dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9),
         "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-8).atol(1e-10),
         "rk4": lambda: ExplicitRungeKutta.rk4(0.01),
         "rkv878e": lambda: ExplicitRungeKutta.rkv878e().rtol(1e-8).atol(1e-10)}[method]
    pick = {"steps": lambda e: e.all_steps(), "dense": lambda e: e.dense(3),
            "even": lambda e: e.even(0.07)}[out]
    ens = pick(IVP.ode(sysm, 0.0, 2.0, y0).method(m().max_steps(200000)))
    want = orc.solve_ode(ens, threads=THREADS)
    assert (want.status == DE_STATUS_COMPLETE).all()
    mappings = [(DE_MAP_AUTO, 0), (DE_MAP_LANE, 0)]
    if method == "dopri5":
        mappings.append((DE_MAP_GROUP, 16))
    for mapping, group in mappings:
        got = ens.options(mapping=mapping, group_size=group).solve()
        if method == "dopri5" and mapping != DE_MAP_LANE:
            points_within_bound(sysm, got, want, even=out == "even")
            continue
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), (f, mapping)
        same_points(got, want)


def test_point_outputs_backward_and_short_slots(orc):
    # A backward solve: the step output completes, the dense and even outputs
    # end OutOfBounds at their first interpolation; slots short of the points
    # keep the first ones and count the rest.
    sysm, y0 = deamd.gen_spring(300, seed=36)
    mk = lambda: IVP.ode(sysm, 0.0, -1.5, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9))
    for even, ens in ((False, mk().all_steps()), (False, mk().dense(2)), (True, mk().even(0.2))):
        want = orc.solve_ode(ens, threads=THREADS)
        for mapping in (DE_MAP_AUTO, DE_MAP_LANE):
            got = ens.options(mapping=mapping).solve()
            if mapping == DE_MAP_AUTO:
                points_within_bound(sysm, got, want, even=even)
                continue
            for f in ("status", "accepted", "rejected", "evals", "t", "y"):
                assert np.array_equal(getattr(got, f), getattr(want, f)), (f, mapping)
            same_points(got, want)
    for m in (ExplicitRungeKutta.dop853().rtol(1e-8).atol(1e-10),     # one lane per system
              ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9)):     # group mappings
        ens = IVP.ode(sysm, 0.0, 2.0, y0).method(m).all_steps()
        full = orc.solve_ode(ens, threads=THREADS)
        cap = np.minimum(full.accepted + 1, np.arange(len(sysm.sizes)) % 7)
        off = np.zeros(len(cap) + 1, dtype=np.int64)
        np.cumsum(cap, out=off[1:])
        got = ens._solve_once(None, off)
        same = (got.accepted == full.accepted) & (got.rejected == full.rejected)
        assert same.mean() >= 0.999
        assert np.array_equal(got.n_points[same], full.n_points[same])
        for i in np.nonzero(same)[0]:
            t, y = got.points(i)
            ft, fy = full.points(i)
            assert len(t) == cap[i]
            assert np.abs(t - ft[:cap[i]]).max(initial=0.0) <= 1e-8 * max(1.0, abs(ft).max())
            assert np.abs(y - fy[:cap[i]]).max(initial=0.0) <= 1e-7 * max(
                np.abs(fy).max(initial=0.0), 1e-300)


def same_events(got, want):
    assert np.array_equal(got.n_events, want.n_events)
    assert np.array_equal(np.isnan(got.event_t), np.isnan(want.event_t))
    assert np.array_equal(np.nan_to_num(got.event_t), np.nan_to_num(want.event_t))
    assert np.array_equal(np.isnan(got.event_y), np.isnan(want.event_y))
    assert np.array_equal(np.nan_to_num(got.event_y), np.nan_to_num(want.event_y))


def events_within_bound(sysm, got, want):
    # DOPRI5 on the group mappings (butterfly norm: step sizes differ in their
    # last bits): equal statuses, step counts and event counts equal on at
    # least 99.9% of systems, and where both are and the solve completed, the
    # final state within 1e-10 and every kept event point's time within 1e-8
    # (relative to |tf|), its state within 1e-7.
    assert (got.status == want.status).all()
    same = (got.accepted == want.accepted) & (got.rejected == want.rejected)
    assert same.mean() >= 0.999
    same_ev = same & (got.n_events == want.n_events)
    assert same_ev.mean() >= 0.999
    done = same_ev & (got.status == DE_STATUS_COMPLETE)
    assert (rel_state_error(sysm, got, want)[done] <= 1e-10).all()
    for i in np.nonzero(same_ev)[0]:
        gt, gy = got.events(i)
        wt, wy = want.events(i)
        assert gt.shape == wt.shape
        assert np.abs(gt - wt).max(initial=0.0) <= 1e-8 * max(1.0, np.abs(wt).max(initial=0.0))
        assert np.abs(gy - wy).max(initial=0.0) <= 1e-7 * max(np.abs(wy).max(initial=0.0),
                                                                1e-300)


EVENT_KINDS = {
    "component": lambda e: e.event(Event.component(1, 0.0)).options(max_events=16),
    "terminate": lambda e: e.event(Event.component(
        0, 0.0, EventConfig(CrossingDirection.Both).terminate_after(2))).options(max_events=4),
    "positive": lambda e: e.event(Event.component(
        1, 0.0, EventConfig(CrossingDirection.Positive))).options(max_events=4),
    "time_terminal_teval": lambda e: e.t_eval(np.linspace(0.0, 2.0, 9)).event(
        Event.time(1.25, EventConfig().terminal())).options(max_events=2),
    "teval_component": lambda e: e.t_eval(np.linspace(-0.5, 2.5, 31)).event(
        Event.component(0, 0.0, EventConfig(CrossingDirection.Negative))).options(max_events=8),
    "crossing": lambda e: e.crossing(0, 0.0, CrossingDirection.Both).options(max_events=1),
}


@pytest.mark.parametrize("kind", list(EVENT_KINDS))
@pytest.mark.parametrize("method", ["dopri5", "dop853", "rk4", "rkv767e"])
def test_event_solouts_bit_exact(orc, method, kind):
    # Event and crossing solouts on spring chains (roots searched inline,
    # event states back to back per system) equal the oracle bit for bit on
    # one lane per system, and within the group mappings' bound there
    # (DOPRI5); a few slots per system, so some points are counted and not
    # kept.
    sysm, y0 = deamd.gen_spring(300, seed=37)
    m = {"dopri5": lambda: ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9),
         "dop853": lambda: ExplicitRungeKutta.dop853().rtol(1e-8).atol(1e-10),
         "rk4": lambda: ExplicitRungeKutta.rk4(0.01),
         "rkv767e": lambda: ExplicitRungeKutta.rkv767e().rtol(1e-8).atol(1e-10)}[method]
    ens = EVENT_KINDS[kind](IVP.ode(sysm, 0.0, 2.0, y0).method(m().max_steps(200000)))
    want = orc.solve_ode(ens, threads=THREADS)
    if kind == "time_terminal_teval":
        # the root lands on the t_eval point 1.25 just pushed: counted, mostly
        # not pushed, and the solve stops there
        assert (want.status == deamd.DE_STATUS_INTERRUPTED).all()
    else:
        assert (want.n_events > 0).mean() > 0.25
    for mapping in (DE_MAP_AUTO, DE_MAP_LANE):
        got = ens.options(mapping=mapping).solve()
        if method == "dopri5" and mapping == DE_MAP_AUTO:
            events_within_bound(sysm, got, want)
            continue
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), (f, mapping)
        same_events(got, want)
        if kind in ("time_terminal_teval", "teval_component"):
            assert np.array_equal(got.n_dense, want.n_dense)
            assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
            assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))


def test_event_solouts_backward_and_refusals(orc):
    # Backward: brent's interpolations fail (no root) and the crossing solout's
    # unwrap ends the solve OutOfBounds, as the oracle; plane events and an
    # event component past a system's state are refused.
    sysm, y0 = deamd.gen_spring(200, seed=38)
    mk = lambda: IVP.ode(sysm, 0.0, -1.5, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9))
    for ens in (mk().event(Event.component(1, 0.0)).options(max_events=8),
                mk().crossing(1, 0.0).options(max_events=8)):
        want = orc.solve_ode(ens, threads=THREADS)
        events_within_bound(sysm, ens.options(mapping=DE_MAP_AUTO).solve(), want)
        got = ens.options(mapping=DE_MAP_LANE).solve()
        for f in ("status", "accepted", "rejected", "evals", "t", "y"):
            assert np.array_equal(getattr(got, f), getattr(want, f)), f
        same_events(got, want)
    with pytest.raises(RuntimeError):
        mk().event(Event.plane(np.zeros(8), np.ones(8))).solve()
    with pytest.raises(RuntimeError):
        mk().event(Event.component(2 * int(sysm.sizes.min()), 0.0)).solve()
