"""GPU parity of the rip-up-and-reroute passes: three passes after the
pattern pass on benchmark configuration 0, every route and every demand
value bit-identical to the CPU oracle, same evaluator score."""
import os
import tempfile

import numpy as np
import pytest

import orc
from instantgr_amd import gpu
from instantgr_amd.host import DAG_POINTS, DAG_STRUCTURE

pytestmark = pytest.mark.gpu


def _routes(h):
    off, data = h.routes_flat()
    return [data[off[i]:off[i] + data[off[i]]] for i in range(len(off) - 1)]


@pytest.mark.parametrize("form,kind,overlap", [(gpu.FORM_PLAIN, 0, False), (gpu.FORM_BATCHED, 0, False),
                                               (gpu.FORM_BATCHED, 2, False), (gpu.FORM_BATCHED, 0, True)],
                         ids=["plain", "batched", "batched-release-driven", "batched-overlapped"])
def test_rrr_three_passes_c1(design_c1, form, kind, overlap):
    """'overlapped': the DAG structure is built by the pass's completion
    callback while the scheduler runs (igr_dag_pass_overlapped)."""
    d = design_c1
    r = gpu.Router.from_design(d)
    if kind:
        r.set_tuning(sched_kind=kind)
    o = orc.Oracle.from_design(d)
    r.set_nets_from_design(d)
    o.set_nets_from_design(d)
    order = d.array("order")
    r.pattern_route(order, form=form)
    o.pattern_route(order)
    hp = d.array("hpwl")
    for it in range(3):
        xg, yg = r.congestion_view()
        xo, yo = o.congestion_view()
        assert np.array_equal(xg, xo) and np.array_equal(yg, yo)
        ig = r.ripup_mark(0)
        assert np.array_equal(ig, o.ripup_mark(0))
        nets = np.nonzero(ig)[0].astype(np.int32)
        nets = nets[np.argsort(-hp[nets], kind="stable")]
        if overlap:
            pts = d.build_dags(nets, xg, yg, 0, 8, DAG_POINTS)
            st = r.dag_pass_overlapped(nets, pts, lambda: d.build_dags(nets, xg, yg, 0, 6, DAG_STRUCTURE))
            dg = d.build_dags(nets, xg, yg, 0, 8)
        else:
            dg = d.build_dags(nets, xg, yg, 0, 8)
            st = r.dag_pass(nets, dg, form=form)
        nb = orc.dag_pass(o, nets, dg)
        assert st["batches"] == nb, it
        assert np.array_equal(r.demand().view(np.uint32), o.demand().view(np.uint32)), it
        assert np.array_equal(r.timestamp(), o.timestamp()), it
        for a, b in zip(r.costs(), o.costs()):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), it
        rg = _routes(r)
        ro = [o.route(i) for i in range(o.num_nets)]
        for i in range(len(rg)):
            assert np.array_equal(rg[i], ro[i]), (it, i)
    with tempfile.TemporaryDirectory() as tmp:
        scores = []
        for h in (r, o):
            off, data = (h.routes_flat() if h is r else o.routes_flat())
            p = os.path.join(tmp, "out")
            d.write_routes(p, off, data)
            scores.append(d.evaluate(p))
        assert scores[0]["opens"] == 0
        assert scores[0]["total"] == scores[1]["total"]
    r.close()


def test_overlapped_pass_reports_a_failed_completion(design_c1):
    """A completion that fails (here it raises on the host) fails the pass with
    that error once the scheduler's launches have finished; the context then
    routes a full pass normally."""
    d = design_c1
    r = gpu.Router.from_design(d)
    r.set_nets_from_design(d)
    r.pattern_route(d.array("order"), form=gpu.FORM_BATCHED)
    xg, yg = r.congestion_view()
    nets = np.nonzero(r.ripup_mark(0))[0].astype(np.int32)
    pts = d.build_dags(nets, xg, yg, 0, 8, DAG_POINTS)

    def broken():
        raise ValueError("no DAGs today")

    with pytest.raises(ValueError, match="no DAGs today"):
        r.dag_pass_overlapped(nets, pts, broken)
    with pytest.raises(ValueError):
        r.dag_pass_overlapped(nets[:-1], pts, broken)
    st = r.dag_pass(nets, d.build_dags(nets, xg, yg, 0, 8), form=gpu.FORM_BATCHED)
    assert st["batches"] > 0
    r.close()
