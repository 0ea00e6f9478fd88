"""Serving loop: requests admitted together give the routes they get alone;
overlapping requests run in arrival order, as a one-at-a-time replay does."""
import numpy as np
import pytest

from instantgr_amd import serve


def _design(nets=6000):
    from instantgr_amd.host import Design
    return Design.preset(0, nets=nets).build(8)


def _routed(d):
    from instantgr_amd import gpu
    r = gpu.Router.from_design(d)
    r.run(d, form=gpu.FORM_BATCHED, rrr=1, threads=8)
    return r


def _window_nets(d, x0, y0, w, limit):
    cx, cy = serve.subnet_centres(d)
    sel = np.nonzero((cx >= x0) & (cx < x0 + w) & (cy >= y0) & (cy < y0 + w))[0]
    return sel[:limit].astype(np.int32)


def test_request_generator_shapes():
    d = _design(3000)
    reqs = serve.make_requests(d, 20, rate=50.0, window=50, kmax=200, seed=3)
    assert len(reqs) == 20
    ts = [t for t, _ in reqs]
    assert all(b > a for a, b in zip(ts, ts[1:]))
    cx, cy = serve.subnet_centres(d)
    for _, nets in reqs:
        assert 1 <= len(nets) <= 200
        assert len(np.unique(nets)) == len(nets)
        # one 50x50 window holds every picked subnet's centre
        assert cx[nets].max() - cx[nets].min() < 50 and cy[nets].max() - cy[nets].min() < 50


@pytest.mark.gpu
def test_alone_and_in_the_mix_route_the_same():
    from instantgr_amd import gpu
    d = _design()
    corners = [(10, 10), (150, 10), (10, 150), (150, 150)]
    reqs = [_window_nets(d, x, y, 30, 40) for x, y in corners]
    assert all(len(q) > 5 for q in reqs)
    # all four together, gathered into one tick
    r = _routed(d)
    with gpu.Server(r, d, hold_ms=300) as srv:
        ids = [srv.submit(q) for q in reqs]
        res = [srv.wait(i) for i in ids]
    assert len({x["tick"] for x in res}) == 1, "disjoint requests were not admitted together"
    assert res[0]["tick_nets"] == sum(len(q) for q in reqs)
    mix = [[r.route(int(n)) for n in q] for q in reqs]
    r.close()
    # each alone, from the same routed state
    for k, q in enumerate(reqs):
        ra = _routed(d)
        with gpu.Server(ra, d) as srv:
            srv.wait(srv.submit(q))
        alone = [ra.route(int(n)) for n in q]
        ra.close()
        same = sum(np.array_equal(a, b) for a, b in zip(alone, mix[k]))
        # Tolerance: every subnet's route identical.  (Only FP64 rounding of
        # prefix sums over tracks shared with another request could differ.)
        assert same == len(q), "request %d: %d of %d routes differ" % (k, len(q) - same, len(q))


@pytest.mark.gpu
def test_overlapping_requests_follow_arrival_order():
    from instantgr_amd import gpu
    d = _design()
    a = _window_nets(d, 60, 60, 40, 60)
    b = _window_nets(d, 70, 70, 40, 60)       # overlaps a
    c = _window_nets(d, 150, 10, 30, 30)      # apart from both
    # arrival order (default): b waits for a, and c for b
    r = _routed(d)
    with gpu.Server(r, d, hold_ms=300) as srv:
        ids = [srv.submit(q) for q in (a, b, c)]
        res = [srv.wait(i) for i in ids]
    assert res[0]["tick"] < res[1]["tick"] <= res[2]["tick"]
    in_order = r.routes_flat()
    r.close()
    r2 = _routed(d)
    with gpu.Server(r2, d) as srv:
        for q in (a, b, c):
            srv.wait(srv.submit(q))
    assert gpu.routes_equal(in_order, r2.routes_flat())
    r2.close()
    # overtaking: c, apart from both, joins a's tick; b still follows a
    r = _routed(d)
    with gpu.Server(r, d, hold_ms=300, overtake=True) as srv:
        ids = [srv.submit(q) for q in (a, b, c)]
        res = [srv.wait(i) for i in ids]
    assert res[0]["tick"] < res[1]["tick"], "an overlapping later request overtook an earlier one"
    assert res[2]["tick"] == res[0]["tick"], "a disjoint request waited behind an overlapping pair"
    served = r.routes_flat()
    r.close()
    # offline: one request at a time in arrival order, c between a and b as
    # it was routed alongside a
    r2 = _routed(d)
    with gpu.Server(r2, d) as srv:
        for q in (a, c, b):
            srv.wait(srv.submit(q))
    assert gpu.routes_equal(served, r2.routes_flat())
    r2.close()


@pytest.mark.gpu
def test_request_validation():
    from instantgr_amd import gpu
    d = _design(500)
    r = _routed(d)
    with gpu.Server(r, d) as srv:
        with pytest.raises(RuntimeError):
            srv.submit(np.array([], np.int32))
        with pytest.raises(RuntimeError):
            srv.submit(np.array([0, 0], np.int32))
        with pytest.raises(RuntimeError):
            srv.submit(np.array([r.num_nets], np.int32))
        with pytest.raises(RuntimeError):
            srv.wait(12345)
        res = srv.wait(srv.submit(np.array([1, 2, 3], np.int32)))
        assert res["nets"] == 3 and res["latency_s"] >= 0
    r.close()


@pytest.mark.gpu
def test_requests_on_the_800x800_design_match_the_oracle():
    """Serving on the routed 800x800 design (500k nets): each reroute request,
    served alone in its tick, leaves the demand, the timestamps and every
    route the oracle's rip-up-and-reroute pass gives when replayed from the
    same state (view, DAGs of the request's subnets by decreasing HPWL, one
    DAG pass: src/Lshape_route_detour.hpp:726-1112)."""
    import orc
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    d = Design.preset(1).build(16)
    r = gpu.Router.from_design(d)
    r.run(d, form=gpu.FORM_BATCHED, rrr=3, threads=16)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.set_state(r.demand(), r.timestamp(), r.stamp, r.routes_flat())
    hp = d.array("hpwl")
    reqs = serve.make_requests(d, 4, rate=10.0, window=50, kmax=1500, seed=21)
    assert any(len(q) > 100 for _, q in reqs)
    for _, q in reqs:
        with gpu.Server(r, d, threads=16) as srv:
            res = srv.wait(srv.submit(q))
        assert res["tick_nets"] == len(q)
        xs, ys = o.congestion_view()
        nets = q[np.argsort(-hp[q], kind="stable")].astype(np.int32)
        orc.dag_pass(o, nets, d.build_dags(nets, xs, ys, 0, 16), 0)
        assert np.array_equal(r.demand().view(np.uint32), o.demand().view(np.uint32))
        assert np.array_equal(r.timestamp(), o.timestamp())
        for n in q:
            assert np.array_equal(r.route(int(n)), o.route(int(n))), int(n)
    assert gpu.routes_equal(r.routes_flat(), o.routes_flat())
    r.close()
