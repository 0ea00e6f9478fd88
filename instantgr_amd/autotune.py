"""Launch-shape autotuner of the batched form.

The batched form has a few launch-shape choices that change speed and never
results (include/igr.h, igr_tuning):

  sched_group  lanes per net in the scheduler's claim and viability kernels
  sched_chunk  scheduler rounds enqueued between host checks
  graph        one hipGraph per pass for the batch sequence, or stream launches
  dag_order    workgroup order of a DAG batch (deepest DAG first, or routing order)
  dag_waves    register budget of the DAG kernel (2 or 4 waves per SIMD)
  track_k      cells per thread of the fused track kernel (4, 8, 16)
  dag_edges    layer-change costs: a wave per edge with LDS staging, or a thread per layer pair
  dag_seq      node count from which a DAG level runs a node per lane instead of 16 lanes per node
  sched_kind   scheduler rounds: active-list (work follows the active nets) or release-driven
               (work follows the released cells)
  dag_grid     persistent workgroups of the DAG kernel pulling nets from the batch's work queue
  lshape_groups  nets (16-lane groups) per workgroup of the pattern DP, which sets the LDS slab a
               workgroup holds (their layer-range tables)

In the terms of a tile shape: nets per workgroup (lshape_groups, dag_grid),
LDS slab size (lshape_groups times the node tables of the node-count class),
and layers per wave (dag_seq: 16 lanes per node, one layer per lane, four
nodes per wave; or a node per lane holding all its layers, 64 per wave).

`tune` measures them on a design (normally a reduced sample of the production
design from the same generator) by coordinate descent: each field in turn takes
the candidate value with the smallest measured time of the pass stage it
affects, the others held at their best so far.  Every candidate run routes the
same nets; the tuner checks that the routes do not change with the shape.

The result is a plain dict that `save` / `load` keep as JSON, and that
`apply` hands to a Router.
"""
import json
import os
import time

from . import gpu

CANDIDATES = {
    "sched_group": [4, 8, 16],
    "sched_chunk": [8, 16, 32, 64],
    "graph": [1, 2],
    "dag_order": [1, 2],
    "dag_waves": [2, 4],
    "track_k": [4, 8, 16],
    "dag_edges": [1, 2],
    "dag_seq": [4, 8, 16, 32, 1 << 30],
    "sched_kind": [1, 2],
    "dag_grid": [1024, 2048, 4096],
    "lshape_groups": [1, 2, 4],
}

# Fields of the plain form only (the tuner measures the batched form).
PLAIN_ONLY = {"presum_k"}

# Which measured stage each field affects.
STAGE = {"sched_group": "schedule", "sched_chunk": "schedule", "graph": "route", "dag_order": "route",
         "dag_waves": "route", "track_k": "route", "dag_edges": "route", "dag_seq": "route",
         "sched_kind": "schedule", "dag_grid": "route", "lshape_groups": "route"}


def _measure(design, tuning, rrr, threads):
    """One pattern pass + rrr rip-up-and-reroute passes; stage times and routes."""
    r = gpu.Router.from_design(design)
    try:
        r.set_tuning(**tuning)
        t0 = time.perf_counter()
        st = r.run(design, form=gpu.FORM_BATCHED, rrr=rrr, threads=threads)
        wall = time.perf_counter() - t0
        sched = st["pattern_schedule_s"] + sum(st["rrr_schedule_s"])
        route = st["pattern_s"] - st["pattern_schedule_s"] + sum(st["rrr_route_s"])
        return {"schedule": sched, "route": route, "wall": wall}, r.routes_flat()
    finally:
        r.close()


def tune(design, rrr=1, threads=16, reps=1, candidates=None, log=None):
    """Coordinate descent over the tuning fields on a built design.

    Returns {"tuning": {...}, "trials": [...]}; every trial records the shape,
    its stage times (best of reps) and whether its routes matched the first
    trial's."""
    cand = dict(CANDIDATES)
    if candidates:
        cand.update(candidates)
    best = {k: 0 for k in cand}       # 0: built-in choice
    trials = []
    ref_routes = None
    for field, values in cand.items():
        scores = {}
        for v in values:
            shape = dict(best)
            shape[field] = v
            t_best = None
            for _ in range(reps):
                times, routes = _measure(design, shape, rrr, threads)
                if ref_routes is None:
                    ref_routes = routes
                if not gpu.routes_equal(routes, ref_routes):
                    raise RuntimeError("routes changed with launch shape %r" % (shape,))
                if t_best is None or times[STAGE[field]] < t_best[STAGE[field]]:
                    t_best = times
            scores[v] = t_best[STAGE[field]]
            trials.append({"shape": shape, "times": t_best, "routes_match": True})
            if log:
                log("autotune %s=%s: %s %.4f s" % (field, v, STAGE[field], scores[v]))
        best[field] = min(values, key=lambda v: (scores[v], values.index(v)))
    return {"tuning": best, "trials": trials}


def key(design):
    """Cache key of a design: grid shape and net count."""
    i = design.info
    return "L%d_X%d_Y%d_nets%d" % (i.L, i.X, i.Y, i.nets)


def save(path, design_key, result):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[design_key] = result
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def load(path, design_key):
    """The tuning stored for design_key, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    res = data.get(design_key)
    return res["tuning"] if res else None


def apply(router, tuning):
    if tuning:
        router.set_tuning(**tuning)


def main(argv=None):
    """Tune on a sample of a benchmark design and store the result:
    python -m instantgr_amd.autotune --config 2 --nets 200000 --key config2"""
    import argparse
    import sys
    from .host import Design
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--config", type=int, default=2, help="benchmark design the sample is drawn from")
    ap.add_argument("--nets", type=int, default=200000, help="nets of the sample")
    ap.add_argument("--rrr", type=int, default=1)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json"))
    ap.add_argument("--key", default="", help="cache key (default: configN, as bench.py looks it up)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    d = Design.preset(a.config, nets=a.nets).build(a.threads)
    log("sample of %d nets built in %.1f s" % (a.nets, time.perf_counter() - t0))
    res = tune(d, rrr=a.rrr, threads=a.threads, reps=a.reps, log=log)
    res["sample"] = "%d-net sample of configuration %d, %d rip-up pass" % (a.nets, a.config, a.rrr)
    save(a.out, a.key or "config%d" % a.config, res)
    print(json.dumps(res["tuning"], sort_keys=True))


if __name__ == "__main__":
    main()
