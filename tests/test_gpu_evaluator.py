"""GPU evaluator (igr_evaluate) against the host evaluator, which the CPU
tests pin to the reference evaluator's golden scores: the same routes give
the same wirelength and via cost exactly, the same overflow cost to 1e-9
relative (FP64 sums in another order), and the same open-net count, on a
routed design and on routes with planted opens and a malformed line."""
import os
import tempfile

import numpy as np
import pytest

from instantgr_amd import gpu

pytestmark = pytest.mark.gpu


def _host_score(design, off, data):
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "r.out")
        design.write_routes(p, off, data)
        return design.evaluate(p)


def _same(g, h):
    assert g["wirelength"] == h["wirelength"]
    assert g["via"] == h["via"]
    # Tolerance: 1e-9 relative; the overflow cost is a sum of exp terms taken
    # in another order than the host evaluator's.
    assert abs(g["overflow"] - h["overflow"]) <= 1e-9 * abs(h["overflow"])
    assert abs(g["total"] - h["total"]) <= 1e-9 * abs(h["total"])
    assert g["opens"] == h["opens"]


@pytest.fixture(scope="module")
def routed(design_c1):
    d = design_c1
    r = gpu.Router.from_design(d)
    r.run(d, form=gpu.FORM_BATCHED, rrr=1, threads=8)
    yield d, r
    r.close()


def test_context_routes_score_as_the_route_file(routed):
    d, r = routed
    off, data = r.routes_flat()
    h = _host_score(d, off, data)
    assert h["opens"] == 0
    _same(r.evaluate(d), h)
    # the same routes passed in from the host
    _same(r.evaluate(d, (off, data)), h)


def test_planted_opens_and_malformed_line(routed):
    d, r = routed
    off, data = r.routes_flat()
    data = data.copy()
    net_sub_off, net_sub = d.array("net_sub_off"), d.array("net_sub")
    pin_off = d.array("net_pin_off")
    # nets with at least three pins and two subnets with route pairs
    cand = [n for n in range(d.info.nets)
            if pin_off[n + 1] - pin_off[n] >= 3 and net_sub_off[n + 1] - net_sub_off[n] >= 2
            and all(data[off[s]] >= 3 for s in net_sub[net_sub_off[n]:net_sub_off[n + 1]])]
    assert len(cand) >= 20
    rng = np.random.default_rng(5)
    pick = rng.choice(cand, size=20, replace=False)
    # drop the whole route of one subnet of 10 nets: each loses a pin
    for n in pick[:10]:
        s = net_sub[net_sub_off[n]]
        data[off[s]] = 1
    # a diagonal line in one subnet of another 10 nets (a malformed line)
    L, X, Y = d.info.L, d.info.X, d.info.Y
    for n in pick[10:]:
        s = net_sub[net_sub_off[n] + 1]
        a = int(data[off[s] + 1])
        z, x, y = a // (X * Y), a // Y % X, a % Y
        x2, y2 = (x + 1) % X, (y + 1) % Y
        data[off[s] + 2] = z * X * Y + x2 * Y + y2
    h = _host_score(d, off, data)
    assert h["opens"] >= 5
    _same(r.evaluate(d, (off, data)), h)


def test_rejects_bad_route_slices(routed):
    d, r = routed
    off, data = r.routes_flat()
    data = data.copy()
    data[off[0]] = 2        # an odd number of cells
    with pytest.raises(RuntimeError):
        r.evaluate(d, (off, data))


def test_heavy_nets_score_as_the_route_file():
    # The configuration-2 generator at 20k nets: its 100-1000-pin nets have
    # more route lines than one wave takes, so they go through the
    # workgroup-per-net kernel.
    from instantgr_amd.host import Design
    d = Design.preset(2, nets=20000).build(8)
    r = gpu.Router.from_design(d)
    r.run(d, form=gpu.FORM_BATCHED, rrr=1, threads=8)
    off, data = r.routes_flat()
    net_sub_off, net_sub = d.array("net_sub_off"), d.array("net_sub")
    lines = np.array([sum((data[off[s]] - 1) // 2 for s in net_sub[net_sub_off[n]:net_sub_off[n + 1]])
                      for n in range(d.info.nets)])
    heavy = np.nonzero(lines > 600)[0]
    assert len(heavy) >= 10
    h = _host_score(d, off, data)
    _same(r.evaluate(d), h)
    # drop one subnet route of 5 heavy nets
    data = data.copy()
    for n in heavy[:5]:
        s = net_sub[net_sub_off[n] + 1]
        data[off[s]] = 1
    h = _host_score(d, off, data)
    assert h["opens"] >= 1
    _same(r.evaluate(d, (off, data)), h)
    r.close()


def test_planted_oracle_routes():
    # Oracle-routed small designs with the planted changes of the CPU check of
    # these rules (dropped lines, stray wires and vias, wires touching end to
    # end, malformed lines), passed in from the host.
    import orc
    from instantgr_amd.host import Design
    from test_eval_model import _plant
    for seed in (5, 6):
        d = Design.custom(layers=6, xsize=30, ysize=28, nets=160, pin_cap=8, sigma=4.0, seed=seed).build(2)
        o = orc.Oracle.from_design(d)
        o.set_nets_from_design(d)
        o.pattern_route(d.array("order"))
        off, data = o.routes_flat()
        off2, data2 = _plant(d, off, data)
        r = gpu.Router.from_design(d)
        r.set_nets_from_design(d)
        for o_, d_ in ((off, data), (off2, data2)):
            _same(r.evaluate(d, (o_, d_)), _host_score(d, o_, d_))
        r.close()
