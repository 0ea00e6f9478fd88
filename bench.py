"""Benchmark: nets routed per second end to end on one MI355X.

One step routes the whole synthetic design from scratch:
  tree build (host) -> pattern pass -> 3 rip-up-and-reroute passes (GPU),
on a fresh device context.  Both execution forms run in the same process:
'plain' (the reference's launch structure: one launch per kernel per batch,
per-depth DAG launches, host synchronisation per step) and 'batched'; their
routes must be identical and their evaluator scores equal.

Prints one JSON line (rank 0).  Only one GPU is supported.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes per grid cell of one launch of each streaming kernel.
BYTES_PER_CELL = {
    "cost_update": 16,   # read capacity + demand, write wcost + vcost
    "presum": 12,        # read wcost (f32), write presum (f64)
    "wire_commit": 16,   # read difference array, read/write demand, clear
    "track_fused": 36,   # commit + cost + prefix of a track in one pass
}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, help="benchmark configuration index (1 = 800x800, 500k nets)")
    ap.add_argument("--nets", type=int, default=0, help="override the net count (testing)")
    ap.add_argument("--rrr", type=int, default=3)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--forms", default="plain,batched")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def one_step(design, form, rrr, threads, profile=False):
    from instantgr_amd import gpu
    t0 = time.perf_counter()
    design.build(threads)
    t1 = time.perf_counter()
    r = gpu.Router.from_design(design)
    if profile:
        r.set_profiling(True)
    st = r.run(design, form=form, rrr=rrr, threads=threads)
    t2 = time.perf_counter()
    st["tree_s"] = t1 - t0
    st["route_total_s"] = t2 - t1
    st["step_s"] = t2 - t0
    return r, st


def roofline(kt, cells):
    """Dominant kernel class and its achieved bandwidth on algorithmic bytes."""
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    name, (ms, launches) = dom
    out = {"kernel": name, "total_ms": ms, "launches": launches}
    if name in BYTES_PER_CELL and launches:
        per_launch = BYTES_PER_CELL[name] * cells
        achieved = per_launch * launches / (ms * 1e-3) / 1e9
        out.update({"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "bytes_per_launch": per_launch, "avg_launch_us": ms * 1e3 / launches})
    else:
        out.update({"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                    "traffic": None})
    # Streaming grid kernels, reported beside the dominant one.
    grid = {}
    for k, per in BYTES_PER_CELL.items():
        ms_k, n_k = kt.get(k, (0.0, 0))
        if n_k and ms_k > 0:
            grid[k] = {"GB/s": per * cells * n_k / (ms_k * 1e-3) / 1e9, "avg_launch_us": ms_k * 1e3 / n_k,
                       "launches": n_k}
    out["grid_kernels"] = grid
    return out


def cpu_baseline():
    """The CPU oracle (sequential restatement of the reference logic, one
    thread) routing a bounded sample of the same workload: the generator's
    configuration-1 distributions at the same net density on a 128x128 grid
    (12,800 nets), pattern pass plus 3 rip-up-and-reroute passes."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from instantgr_amd.host import Design
    nets = 12800
    d = Design.custom(layers=10, xsize=128, ysize=128, nets=nets, pin_cap=64, seed=2)
    t0 = time.perf_counter()
    d.build(1)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    hp = d.array("hpwl")
    for _ in range(3):
        xs, ys = o.congestion_view()
        isof = o.ripup_mark(0)
        sel = np.nonzero(isof)[0].astype(np.int32)
        if len(sel) == 0:
            break
        sel = sel[np.argsort(-hp[sel], kind="stable")]
        dg = d.build_dags(sel, xs, ys, 0, 1)
        orc.dag_pass(o, sel, dg, 0)
    dt = time.perf_counter() - t0
    return {"value": nets / dt, "unit": "nets/s", "cores": 1, "kind": "port",
            "sample": "10 layers, 128x128 gcells, 12800 nets (configuration-1 generator and density), "
                      "tree build + pattern pass + 3 RRR passes; %.1f s" % dt}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the routing path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    design = Design.preset(a.config, nets=a.nets)
    forms = [f.strip() for f in a.forms.split(",") if f.strip()]
    form_id = {"plain": gpu.FORM_PLAIN, "batched": gpu.FORM_BATCHED}
    results = {}
    routes = {}
    for f in forms:
        for _ in range(a.warmup):
            r, _st = one_step(design, form_id[f], a.rrr, a.threads)
            r.close()
        times = []
        last = None
        for k in range(a.steps):
            r, st = one_step(design, form_id[f], a.rrr, a.threads, profile=(k == a.steps - 1))
            times.append(st["step_s"])
            if last is not None:
                last.close()
            last = (r, st)
        r, st = last
        kt = r.kernel_times()
        routes[f] = r.routes_flat()
        results[f] = {"step_s": times, "stats": st, "kernels": kt}
        print("[bench] %s: %s" % (f, json.dumps({"step_s": times, "pattern_batches": st["pattern_batches"],
                                                   "ripped": st["ripped"], "tree_s": st["tree_s"],
                                                   "pattern_s": st["pattern_s"],
                                                   "rrr_dag_build_s": st["rrr_dag_build_s"],
                                                   "rrr_route_s": st["rrr_route_s"],
                                                   "rrr_schedule_s": st["rrr_schedule_s"],
                                                   "rrr_prep_s": st["rrr_prep_s"],
                                                   "pattern_schedule_s": st["pattern_schedule_s"],
                                                   "kernel_ms": {k: round(v[0], 1) for k, v in kt.items() if v[1]}})),
              file=sys.stderr, flush=True)
        r.close()

    info = design.info
    nets_in = info.nets
    cells = info.L * info.X * info.Y
    # Identical routes and score across forms.
    import numpy as np
    same = None
    scores = {}
    with tempfile.TemporaryDirectory() as tmp:
        for f in forms:
            off, data = routes[f]
            p = os.path.join(tmp, f + ".out")
            design.write_routes(p, off, data)
            scores[f] = design.evaluate(p)
    if len(forms) == 2:
        (o1, d1), (o2, d2) = routes[forms[0]], routes[forms[1]]
        same = all(np.array_equal(d1[o1[i]:o1[i] + d1[o1[i]]], d2[o2[i]:o2[i] + d2[o2[i]]])
                   for i in range(len(o1) - 1))
    head = "batched" if "batched" in results else forms[-1]
    step_ms = 1e3 * sum(results[head]["step_s"]) / len(results[head]["step_s"])
    value = nets_in / (step_ms * 1e-3)
    out = {
        "metric": "nets routed/sec end-to-end (batched vs plain, same GPU) at identical eval score",
        "value": value, "unit": "nets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "configuration %d: %d layers, %dx%d gcells, %d nets; tree build + pattern pass + "
                               "%d rip-up-and-reroute passes" % (a.config, info.metal_layers, info.X, info.Y,
                                                                 nets_in, a.rrr),
                   "subnets": int(info.subnets), "form": head},
        "forms": {f: {"nets_per_s": nets_in / (sum(v["step_s"]) / len(v["step_s"])),
                      "ms_per_step": 1e3 * sum(v["step_s"]) / len(v["step_s"]),
                      "score": scores[f]} for f, v in results.items()},
        "routes_identical": same,
        "roofline": roofline(results[head]["kernels"], cells),
    }
    if "plain" in results and "batched" in results:
        out["speedup_batched_vs_plain"] = out["forms"]["plain"]["ms_per_step"] / out["forms"]["batched"]["ms_per_step"]
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
