"""GPU parity of the entry points users call, and of the layer count of the
largest configuration:

- igr_run (pattern pass plus three rip-up-and-reroute passes, the sequence of
  src/main.cpp:22-73 with the RRR step repeated) in both forms against the
  oracle running the same sequence: every route, demand, timestamps, costs;
- the igr_route command line (README.md:20-22, src/main.cpp:91-102) on .cap/.net
  files of configuration 0: the route file it writes is byte-identical to the
  one written from the oracle's routes, and scores the same;
- a 12-metal-layer design (11 routing layers, the layer count of the
  2500x2500 configuration; the reference fixes 10 metal layers,
  src/database.hpp:292): igr_run in both forms bit-exact against the oracle.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import orc
from instantgr_amd import gpu
from instantgr_amd.host import Design

pytestmark = pytest.mark.gpu


def oracle_run(d, rrr):
    """The oracle through the sequence igr_run performs: pattern pass, then per
    pass congestion view, rip-up marking, DAGs of the marked subnets in
    decreasing HPWL order (stable), one DAG pass.  DAG mode by subnet count
    (src/main.cpp:39-40)."""
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    n = o.num_nets
    mode = 2 if n > 20000000 else (1 if n > 5000000 else 0)
    hp = d.array("hpwl")
    for _ in range(rrr):
        xs, ys = o.congestion_view()
        nets = np.nonzero(o.ripup_mark(0))[0].astype(np.int32)
        if len(nets) == 0:
            break
        nets = nets[np.argsort(-hp[nets], kind="stable")]
        orc.dag_pass(o, nets, d.build_dags(nets, xs, ys, mode, 8), mode)
    return o


def _same_state(r, o):
    assert np.array_equal(r.demand().view(np.uint32), o.demand().view(np.uint32))
    assert np.array_equal(r.timestamp(), o.timestamp())
    for a, b in zip(r.costs(), o.costs()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    goff, gdata = r.routes_flat()
    ooff, odata = o.routes_flat()
    for i in range(r.num_nets):
        assert np.array_equal(gdata[goff[i]:goff[i] + gdata[goff[i]]], odata[ooff[i]:ooff[i + 1]]), i


@pytest.fixture(scope="module")
def c1_oracle(design_c1):
    return oracle_run(design_c1, 3)


@pytest.mark.parametrize("form", [gpu.FORM_PLAIN, gpu.FORM_BATCHED], ids=["plain", "batched"])
def test_igr_run_matches_oracle_c1(design_c1, c1_oracle, form):
    d = design_c1
    r = gpu.Router.from_design(d)
    st = r.run(d, form=form, rrr=3, threads=8)
    assert st["iterations"] == 3
    _same_state(r, c1_oracle)
    # the device score of those routes equals the host evaluator's on the
    # route file written from the oracle's routes
    g = r.evaluate(d)
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "o.out")
        d.write_routes(p, *c1_oracle.routes_flat())
        h = d.evaluate(p)
    assert g["opens"] == h["opens"] == 0
    assert g["wirelength"] == h["wirelength"] and g["via"] == h["via"]
    assert abs(g["overflow"] - h["overflow"]) <= 1e-9 * abs(h["overflow"])
    r.close()


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_command_line_c1(design_c1, form):
    from instantgr_amd import build as igr_build
    exe = os.path.join(os.path.dirname(igr_build.__file__), "lib", "igr_route")
    with tempfile.TemporaryDirectory() as tmp:
        cap, net, out, ref = (os.path.join(tmp, f) for f in ("d.cap", "d.net", "d.out", "ref.out"))
        design_c1.save(cap, net)
        proc = subprocess.run([exe, "-cap", cap, "-net", net, "-out", out, "-form", form, "-rrr", "3",
                               "-threads", "8"], capture_output=True, text=True, timeout=300)
        assert proc.returncode == 0, proc.stderr
        assert "rrr3" in proc.stdout
        # the same design read back from the files, routed by the oracle
        d = Design.load(cap, net).build(8)
        o = oracle_run(d, 3)
        d.write_routes(ref, *o.routes_flat())
        with open(out, "rb") as f1, open(ref, "rb") as f2:
            assert f1.read() == f2.read()
        s = d.evaluate(out)
        assert s["opens"] == 0 and s["total"] == d.evaluate(ref)["total"]


@pytest.fixture(scope="module")
def design_l12():
    """12 metal layers (11 routing layers), 120x100 gcells, 4000 nets."""
    return Design.custom(layers=12, xsize=120, ysize=100, nets=4000, pin_cap=30, sigma=6.0, seed=12).build(8)


@pytest.mark.parametrize("form", [gpu.FORM_PLAIN, gpu.FORM_BATCHED], ids=["plain", "batched"])
def test_twelve_metal_layers_bitexact(design_l12, form):
    d = design_l12
    assert d.info.L == 11
    o = oracle_run(d, 3)
    r = gpu.Router.from_design(d)
    st = r.run(d, form=form, rrr=3, threads=8)
    assert st["iterations"] == 3 and all(k > 0 for k in st["ripped"])
    _same_state(r, o)
    g = r.evaluate(d)
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "o.out")
        d.write_routes(p, *o.routes_flat())
        h = d.evaluate(p)
    assert g["opens"] == h["opens"] == 0
    assert g["wirelength"] == h["wirelength"] and g["via"] == h["via"]
    assert abs(g["overflow"] - h["overflow"]) <= 1e-9 * abs(h["overflow"])
    r.close()
