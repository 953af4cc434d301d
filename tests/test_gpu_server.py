"""Serving loop: continuous batching of ensemble requests on one GPU.

Each request's results must not depend on what else is being served: the
same request gives the same bits alone and inside a mixed concurrent stream,
and (for the lane-mapped families) the same bits as a direct batch solve.
"""
import numpy as np
import pytest

import deamd
from deamd import ExplicitRungeKutta, IVP, Lorenz, Server, VanDerPol

pytestmark = pytest.mark.gpu
GIB = 1 << 30


def lorenz_req(n, seed, tf=5.0, rtol=1e-8):
    p, y0 = deamd.gen_lorenz(n, seed=seed)
    return IVP.ode(Lorenz(*p), 0.0, tf, y0).method(
        ExplicitRungeKutta.dopri5().rtol(rtol).atol(rtol * 1e-2).max_steps(1000000))


def vdp_req(n, seed, tf=8.0, rtol=1e-6):
    mu, y0 = deamd.gen_vdp(n, seed=seed)
    return IVP.ode(VanDerPol(*mu), 0.0, tf, y0).method(
        ExplicitRungeKutta.dopri5().rtol(rtol).atol(rtol * 1e-2).max_steps(1000000))


def spring_req(n, seed, tf=3.0):
    sysm, y0 = deamd.gen_spring(n, seed=seed)
    return IVP.ode(sysm, 0.0, tf, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-6).atol(1e-8).max_steps(1000000))


FIELDS = ("status", "accepted", "rejected", "evals", "t", "y")


def assert_same(a, b):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_server_matches_direct_solve():
    req = lorenz_req(3000, seed=1)
    want = req.solve()
    with Server(8 * GIB) as srv:
        t = srv.submit(req)
        got = srv.poll(t)
        assert srv.latency(t) > 0
    assert_same(got, want)


def test_served_dense_output_matches_direct_solve():
    # t_eval on the lane-mapped families: the served request writes the same
    # dense rows (NaN where a point is not produced) as a direct solve.
    pts = np.linspace(-1.0, 6.0, 29)        # the first points precede t0
    req = lorenz_req(1500, seed=3).t_eval(pts)
    want = req.solve()
    with Server(8 * GIB) as srv:
        got = srv.poll(srv.submit(req))
    assert_same(got, want)
    assert np.array_equal(got.n_dense, want.n_dense)
    assert np.array_equal(np.isnan(got.dense), np.isnan(want.dense))
    assert np.array_equal(np.nan_to_num(got.dense), np.nan_to_num(want.dense))
    rows = want.dense.reshape(29, 3, -1)
    assert np.isnan(rows[:4]).all() and np.isnan(rows[-4:]).all()   # before t0, after tf
    assert (want.n_dense > 0).all()


@pytest.mark.parametrize("graph", [False, True])
def test_result_alone_equals_result_in_mix(graph):
    probes = [lorenz_req(700, seed=11), vdp_req(500, seed=12), spring_req(300, seed=13)]
    alone = []
    with Server(8 * GIB, graph=graph) as srv:
        for r in probes:
            alone.append(srv.poll(srv.submit(r)))
    rng = np.random.default_rng(5)
    with Server(8 * GIB, graph=graph) as srv:
        tickets = []
        for k in range(24):
            kind = k % 3
            n = int(np.exp(rng.uniform(0, np.log(4096))))
            mk = (lorenz_req, vdp_req, spring_req)[kind]
            tickets.append(srv.submit(mk(max(n, 1), seed=100 + k)))
            if k in (5, 11, 17):
                tickets.append(("probe", k // 6, srv.submit(probes[k // 6])))
        results = {}
        for t in tickets:
            if isinstance(t, tuple):
                results[t[1]] = srv.poll(t[2])
            else:
                sol = srv.poll(t)
                assert (sol.status == 0).all()
    for i in range(3):
        assert_same(results[i], alone[i])


def test_spring_requests_within_bound(orc):
    req = spring_req(400, seed=21)
    want = orc.solve_ode(req, threads=16)
    with Server(8 * GIB) as srv:
        got = srv.poll(srv.submit(req))
    assert (got.status == 0).all()
    err = np.abs(got.y - want.y).max() / np.abs(want.y).max()
    assert err <= 1e-10
    assert ((got.accepted == want.accepted) & (got.rejected == want.rejected)).mean() >= 0.999


def test_admission_respects_budget():
    # Budget a little above the slot pools (about 0.6 GB), so a results arena
    # of 128 MiB: a 4M-trajectory request (250 MB of results) can never fit
    # and is rejected; small ones are admitted progressively as earlier ones
    # retire, and all complete.
    with Server(1 * GIB) as srv:
        big = lorenz_req(1 << 22, seed=3, tf=0.1)
        t_big = srv.submit(big)
        with pytest.raises(RuntimeError):
            srv.poll(t_big)
        ts = [srv.submit(lorenz_req(20000, seed=40 + k, tf=1.0)) for k in range(12)]
        for t in ts:
            assert (srv.poll(t).status == 0).all()


def test_served_dop853_matches_direct_solve():
    # DOP853 requests of the lane-mapped families run in their own slot pools
    # beside the DOPRI5 ones: each gives the bits of a direct solve, alone and
    # in a mixed stream, with and without t_eval output.
    p, y0 = deamd.gen_lorenz(2000, seed=21)
    d853 = IVP.ode(Lorenz(*p), 0.0, 5.0, y0).method(
        ExplicitRungeKutta.dop853().rtol(1e-9).atol(1e-11).max_steps(1000000))
    mu, v0 = deamd.gen_vdp(1500, seed=22)
    vdp853 = IVP.ode(VanDerPol(*mu), 0.0, 8.0, v0).method(
        ExplicitRungeKutta.dop853().rtol(1e-7).atol(1e-9).max_steps(1000000)) \
        .t_eval(np.linspace(0.0, 8.0, 11))
    want_l, want_v = d853.solve(), vdp853.solve()
    with Server(8 * GIB) as srv:
        assert_same(srv.poll(srv.submit(d853)), want_l)
        ts = [srv.submit(lorenz_req(1000, seed=23)), srv.submit(d853), srv.submit(vdp853),
              srv.submit(vdp_req(800, seed=24))]
        got = [srv.poll(t) for t in ts]
    assert_same(got[1], want_l)
    assert_same(got[2], want_v)
    assert np.array_equal(got[2].n_dense, want_v.n_dense)
    assert np.array_equal(np.nan_to_num(got[2].dense), np.nan_to_num(want_v.dense))
    assert_same(got[0], lorenz_req(1000, seed=23).solve())


def _same_outputs(got, want):
    assert_same(got, want)
    for f in ("dense", "event_t", "event_y"):
        x, y = getattr(got, f), getattr(want, f)
        assert (x is None) == (y is None), f
        if x is not None:
            assert np.array_equal(np.isnan(x), np.isnan(y)) and \
                np.array_equal(np.nan_to_num(x), np.nan_to_num(y)), f
    for f in ("n_dense", "n_events"):
        x, y = getattr(got, f), getattr(want, f)
        assert (x is None) == (y is None) and (x is None or np.array_equal(x, y)), f


@pytest.mark.parametrize("graph", [False, True])
def test_served_methods_match_direct_solves(graph):
    # Every method is served: RK4, Euler and the embedded family on Lorenz and
    # Van der Pol lanes, and every method (and DOPRI5 with an event solout) on
    # spring chains with one lane per system, in pools made on first use.
    # Each request gives the bits of its direct solve, in a mixed stream.
    from deamd import CrossingDirection, Event
    p, y0 = deamd.gen_lorenz(1200, seed=31)
    mu, v0 = deamd.gen_vdp(900, seed=32)
    sysm, sy0 = deamd.gen_spring(150, seed=33)
    lor = lambda m: IVP.ode(Lorenz(*p), 0.0, 3.0, y0).method(m.max_steps(1000000))
    vdp = lambda m: IVP.ode(VanDerPol(*mu), 0.0, 6.0, v0).method(m.max_steps(1000000))
    spr = lambda m: IVP.ode(sysm, 0.0, 2.0, sy0).method(m.max_steps(1000000))
    E = ExplicitRungeKutta
    reqs = [
        lor(E.rkv878e().rtol(1e-9).atol(1e-11)).t_eval(np.linspace(0.0, 3.0, 7)),
        vdp(E.rkf45().rtol(1e-7).atol(1e-9)),
        lor(E.rk4(0.002)),
        vdp(E.euler(0.001)).t_eval(np.linspace(0.0, 6.0, 5)),
        spr(E.dop853().rtol(1e-8).atol(1e-10)),
        spr(E.rk4(0.01)).t_eval(np.linspace(0.0, 2.0, 9)),
        spr(E.rkv767e().rtol(1e-8).atol(1e-10)),
        # served on one lane per system: its direct solve on the same mapping
        spr(E.dopri5().rtol(1e-7).atol(1e-9)).event(Event.component(1, 0.0))
        .options(max_events=4, mapping=deamd.DE_MAP_LANE),
        spr(E.dop853().rtol(1e-8).atol(1e-10)).crossing(0, 0.0, CrossingDirection.Positive)
        .options(max_events=2),
    ]
    want = [r.solve() for r in reqs]
    with Server(8 * GIB, graph=graph) as srv:
        tickets = []
        for k, r in enumerate(reqs):
            tickets.append(srv.submit(r))
            if k % 3 == 0:
                tickets.append(srv.submit(lorenz_req(500, seed=40 + k)))
        got = [srv.poll(t) for t in tickets]
        again = srv.poll(srv.submit(reqs[4]))   # a second request to a made pool
    idx = 0
    for k in range(len(reqs)):
        _same_outputs(got[idx], want[k])
        idx += 2 if k % 3 == 0 else 1
    _same_outputs(again, want[4])


def test_server_refusals():
    # Event solouts are served for every method; spring chains' states have
    # no single length for a plane, and an event component must lie in every
    # trajectory's state.
    from deamd import Event
    sysm, y0 = deamd.gen_spring(8, seed=1)
    p, ly0 = deamd.gen_lorenz(8, seed=1)
    with Server(8 * GIB) as srv:
        for req in (IVP.ode(sysm, 0.0, 1.0, y0).event(Event.plane(np.zeros(8), np.ones(8))),
                    IVP.ode(Lorenz(*p), 0.0, 1.0, ly0).method(ExplicitRungeKutta.rkf45())
                    .event(Event.component(3, 0.0)),
                    IVP.ode(sysm, 0.0, 1.0, y0).event(
                        Event.component(2 * int(sysm.sizes.min()), 0.0))):
            with pytest.raises(RuntimeError):
                srv.submit(req)


def _event_fields(a, b):
    assert np.array_equal(a.n_events, b.n_events)
    for f in ("event_t", "event_y"):
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(np.isnan(x), np.isnan(y)) and \
            np.array_equal(np.nan_to_num(x), np.nan_to_num(y)), f


def test_served_event_solouts_match_direct_solves():
    # DOPRI5 with an event solout on the lane-mapped families: served with
    # inline roots, the same bits as a direct solve (deferred roots), alone
    # and inside a mixed stream.
    from deamd import CrossingDirection, Event, EventConfig
    p, y0 = deamd.gen_lorenz(1500, seed=31)
    base = lambda: IVP.ode(Lorenz(*p), 0.0, 6.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-8).atol(1e-10).max_steps(1000000))
    mu, v0 = deamd.gen_vdp(1200, seed=32)
    reqs = [
        base().event(Event.component(2, p[1] - 1.0)).options(max_events=16),
        base().t_eval(np.linspace(0.0, 6.0, 13)).event(
            Event.plane([0.0, 0.0, 20.0], [0.2, -0.1, 1.0], EventConfig().terminate_after(3)))
        .options(max_events=4),
        base().crossing(0, 0.0, CrossingDirection.Both).options(max_events=12),
        base().hyperplane_crossing([0.0, 0.0, 25.0], [1.0, 1.0, 3.0],
                                   CrossingDirection.Positive).options(max_events=12),
        IVP.ode(VanDerPol(*mu), 0.0, 10.0, v0).method(
            ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9).max_steps(1000000))
        .crossing(0, 0.0, CrossingDirection.Negative).options(max_events=8),
    ]
    wants = [r.solve() for r in reqs]
    assert all(w.n_events.sum() > 0 for w in wants)
    with Server(8 * GIB) as srv:
        alone = srv.poll(srv.submit(reqs[0]))
        ts = [srv.submit(lorenz_req(900, seed=33))] + [srv.submit(r) for r in reqs]
        got = [srv.poll(t) for t in ts]
    for g, w in zip([alone] + got[1:], [wants[0]] + wants):
        assert_same(g, w)
        _event_fields(g, w)
    assert np.array_equal(got[2].n_dense, wants[1].n_dense)
    assert np.array_equal(np.nan_to_num(got[2].dense), np.nan_to_num(wants[1].dense))


@pytest.mark.parametrize("method", ["dop853", "rk4", "euler", "rkf45", "rkv878e"])
def test_served_event_solouts_of_other_methods(method):
    # Event solouts with DOP853, RK4, Euler and the embedded family on Lorenz
    # and Van der Pol lanes (pools made on first use, roots inline): the bits
    # of a direct solve, t_eval and terminate counts included.
    from deamd import CrossingDirection, Event, EventConfig
    E = ExplicitRungeKutta
    m = {"dop853": lambda: E.dop853().rtol(1e-9).atol(1e-11),
         "rk4": lambda: E.rk4(0.002), "euler": lambda: E.euler(0.0005),
         "rkf45": lambda: E.rkf45().rtol(1e-8).atol(1e-10),
         "rkv878e": lambda: E.rkv878e().rtol(1e-9).atol(1e-11)}[method]
    p, y0 = deamd.gen_lorenz(800, seed=41)
    mu, v0 = deamd.gen_vdp(600, seed=42)
    lor = lambda: IVP.ode(Lorenz(*p), 0.0, 4.0, y0).method(m().max_steps(1000000))
    reqs = [
        lor().event(Event.component(2, p[1] - 1.0)).options(max_events=8),
        lor().t_eval(np.linspace(0.0, 4.0, 9)).event(
            Event.component(0, 0.0, EventConfig().terminate_after(2))).options(max_events=4),
        lor().hyperplane_crossing([0.0, 0.0, 25.0], [1.0, 1.0, 3.0],
                                  CrossingDirection.Positive).options(max_events=8),
        IVP.ode(VanDerPol(*mu), 0.0, 8.0, v0).method(m().max_steps(1000000))
        .crossing(0, 0.0, CrossingDirection.Negative).options(max_events=6),
    ]
    wants = [r.solve() for r in reqs]
    assert all(w.n_events.sum() > 0 for w in wants)
    with Server(8 * GIB) as srv:
        ts = [srv.submit(r) for r in reqs] + [srv.submit(lorenz_req(600, seed=43))]
        got = [srv.poll(t) for t in ts]
    for g, w in zip(got, wants):
        _same_outputs(g, w)


@pytest.mark.parametrize("graph", [False, True])
def test_range_table_restarts_when_pool_drains(monkeypatch, graph):
    # With room for 8 ranges per pool, 40 requests to one pool fill the table
    # several times: admission holds until the pool drains, the table
    # restarts, and every request still completes with the direct solve's bits.
    monkeypatch.setenv("DE_SERVER_MAX_RANGES", "8")
    reqs = [lorenz_req(64 + 7 * k, seed=100 + k, tf=1.0) for k in range(40)]
    wants = [r.solve() for r in reqs]
    with Server(4 * GIB, graph=graph) as srv:
        tickets = [srv.submit(r) for r in reqs]
        got = [srv.poll(t) for t in tickets]
        _, resets = srv.stats()
    assert resets >= 40 // 8
    for g, w in zip(got, wants):
        assert_same(g, w)


def test_poisson_stream_results_equal_direct_solves():
    # The serving workload's request stream (families, sizes up to 65 536,
    # rtol, t_final as the serving benchmark draws them) at 2 000 requests/s,
    # submitted in real time over a 40 ms window: every Lorenz and Van der Pol
    # request returns its direct solve's bits; spring-chain requests (their
    # slot pools group chains by size class, so the reductions' order can differ
    # from a direct solve's) agree to 1e-10 of the largest state, with equal
    # step counts on at least 99.9% of the chains.
    import threading
    import time
    stream = deamd.poisson_requests(2000.0, 0.04, seed=2026)
    assert len(stream) == 80 and {f for _, f, _, _ in stream} == {0, 1, 2}
    tickets = [None] * len(stream)
    with Server(32 * GIB) as srv:
        t0 = time.perf_counter()

        def submitter():
            for k, (at, _, _, e) in enumerate(stream):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[k] = srv.submit(e)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        got = [srv.poll(t) for t in tickets]
    same_counts = total = 0
    for (_, fam, n, e), g in zip(stream, got):
        assert (g.status == 0).all()
        want = e.solve()
        if fam < 2:
            assert_same(g, want)
        else:
            err = np.abs(g.y - want.y).max() / np.abs(want.y).max()
            assert err <= 1e-10
            same_counts += int(((g.accepted == want.accepted) &
                                (g.rejected == want.rejected)).sum())
            total += n
    assert total > 0 and same_counts >= 0.999 * total


def test_idle_lazy_pools_are_released_under_a_small_budget():
    # A budget a little above the pools made at creation leaves an arena of a
    # few hundred MB: the pools made on first use for 27 (family, method)
    # kinds do not all fit, so idle ones are released for the next, and a
    # released kind made again still gives the bits of a direct solve.
    E = ExplicitRungeKutta
    names = ["rk4", "euler", "rkf45", "cash_karp", "rkv655e", "rkv656e", "rkv766e",
             "rkv767e", "rkv877e", "rkv878e", "rkv988e", "rkv989e"]
    mk = {"rk4": lambda: E.rk4(0.01), "euler": lambda: E.euler(0.002)}
    method = lambda nm: mk[nm]() if nm in mk else getattr(E, nm)().rtol(1e-7).atol(1e-9)
    p, y0 = deamd.gen_lorenz(64, seed=51)
    sysm, sy0 = deamd.gen_spring(16, seed=52)
    reqs = [IVP.ode(Lorenz(*p), 0.0, 1.0, y0).method(method(nm).max_steps(100000))
            for nm in names]
    reqs += [IVP.ode(sysm, 0.0, 0.5, sy0).method(method(nm).max_steps(100000))
             for nm in names + ["dop853"]]
    reqs.append(reqs[0])                     # a released kind, made again
    with Server(1 * GIB) as srv:
        for r in reqs:
            _same_outputs(srv.poll(srv.submit(r)), r.solve())
        made, released = srv.pools()
        assert released > 0 and 0 < made < 25


@pytest.mark.parametrize("graph", [False, True])
def test_served_spring_events_on_the_group_mapping(orc, graph):
    # Spring chains with DOPRI5 and an event or crossing solout are served on
    # the register group mapping of each size class (pools made on first
    # use): against the oracle, the group mappings' bound -- step and event
    # counts equal on at least 99.9% of systems, and where they are, final
    # states within 1e-10 and event times within 1e-8, states within 1e-7.
    from deamd import CrossingDirection, Event
    sysm, y0 = deamd.gen_spring(400, seed=61)
    base = lambda: IVP.ode(sysm, 0.0, 3.0, y0).method(
        ExplicitRungeKutta.dopri5().rtol(1e-7).atol(1e-9).max_steps(1000000))
    reqs = [base().event(Event.component(1, 0.0)).options(max_events=6),
            base().crossing(0, 0.0, CrossingDirection.Positive).options(max_events=3)]
    with Server(8 * GIB, graph=graph) as srv:
        ts = [srv.submit(r) for r in reqs]
        got = [srv.poll(t) for t in ts]
    for g, r in zip(got, reqs):
        w = orc.solve_ode(r, threads=16)
        assert (w.n_events > 0).mean() > 0.5
        assert (g.status == w.status).all()
        same = (g.accepted == w.accepted) & (g.rejected == w.rejected) & (g.n_events == w.n_events)
        assert same.mean() >= 0.999
        for i in np.nonzero(same)[0]:
            a, b = sysm.state_offset[i], sysm.state_offset[i + 1]
            assert np.abs(g.y[a:b] - w.y[a:b]).max() <= 1e-10 * np.abs(w.y[a:b]).max()
            gt, gy = g.events(i)
            wt, wy = w.events(i)
            assert np.abs(gt - wt).max(initial=0.0) <= 1e-8 * 3.0
            assert np.abs(gy - wy).max(initial=0.0) <= 1e-7 * max(np.abs(wy).max(initial=0.0),
                                                                    1e-300)
