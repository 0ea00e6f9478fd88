"""Writes the evaluator fixtures under tests/golden/.

For each fixture: a small synthetic design (.cap/.net in the reference
format), its routes (.out) after the CPU oracle's pattern pass and one
rip-up-and-reroute pass, and the scores printed by the reference evaluator
(run/evaluator.cpp, compiled from its source by `make -C oracle ref`) in a
.json file.  Run from the repository root in a tree where oracle/_ref/evaluator
has been built:

    python tests/golden/make_eval_golden.py
"""
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

FIXTURES = {
    # name: generator parameters
    "eval_small": dict(layers=6, xsize=40, ysize=36, nets=260, pin_cap=12, sigma=4.0, seed=11),
    "eval_dense": dict(layers=5, xsize=24, ysize=20, nets=220, pin_cap=8, sigma=3.0, seed=12),
}


def parse_ref(text):
    def num(label):
        m = re.search(r"^%s\s+([-0-9.eE+]+)\s*$" % re.escape(label), text, re.M)
        if not m:
            raise RuntimeError("reference evaluator printed no '%s'" % label)
        return float(m.group(1))
    opens = int(re.search(r"Number of open nets : (\d+)", text).group(1))
    return {"wirelength": num("wirelength cost"), "via": num("via cost"), "overflow": num("overflow cost"),
            "total": num("total cost"), "opens": opens}


def route(d):
    import orc
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    xs, ys = o.congestion_view()
    isof = o.ripup_mark(0)
    sel = np.nonzero(isof)[0].astype(np.int32)
    hp = d.array("hpwl")
    sel = sel[np.argsort(-hp[sel], kind="stable")]
    if len(sel):
        dg = d.build_dags(sel, xs, ys, 0, 1)
        orc.dag_pass(o, sel, dg, 0)
    return o.routes_flat(), len(sel)


def main():
    from instantgr_amd.host import Design
    ev = os.path.join(ROOT, "oracle", "_ref", "evaluator")
    if not os.path.exists(ev):
        raise SystemExit("build the reference evaluator first: make -C oracle ref")
    for name, params in FIXTURES.items():
        base = os.path.join(HERE, name)
        Design.custom(**params).save(base + ".cap", base + ".net")
        d = Design.load(base + ".cap", base + ".net").build(1)
        (off, data), ripped = route(d)
        d.write_routes(base + ".out", off, data)
        # The reference evaluator's main() returns 1 after a successful run
        # and 0 on an input error (run/evaluator.cpp:267-277).
        res = subprocess.run([ev, base + ".cap", base + ".net", base + ".out"], capture_output=True, text=True)
        if res.returncode != 1:
            raise RuntimeError("reference evaluator failed on %s: %s" % (name, res.stdout[-400:]))
        out = res.stdout
        score = parse_ref(out)
        score["generator"] = params
        score["ripped_in_rrr"] = ripped
        with open(base + ".json", "w") as f:
            json.dump(score, f, indent=1, sort_keys=True)
        print(name, score)


if __name__ == "__main__":
    main()
