"""Benchmark: nets routed per second end to end on one MI355X.

One step routes the whole synthetic design from scratch:
  tree build (host) -> pattern pass -> 3 rip-up-and-reroute passes (GPU),
on a fresh device context.  Both execution forms run in the same process:
'plain' (the reference's launch structure: one launch per kernel per batch,
per-depth DAG launches, host synchronisation per step) and 'batched'; their
routes must be identical and their evaluator scores equal.

Prints one JSON line.  Only one GPU is supported.

Helpers for the committed profile summaries (no GPU needed):
  --summarize-stats DIR             kernel table from rocprofv3 --kernel-trace --stats
  --summarize-pmc FETCH_DIR WRITE_DIR  calibrated HBM bytes per dispatch from two
                                    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE)
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Kernel class -> kernel symbol (the name rocprofv3 reports).
KERNEL_SYMBOL = {"dag_dp": "k_dag_net", "dag_edges": "k_dag_edges_wave", "track_fused": "k_track_fused",
                 "lshape_dp": "k_lshape_plain", "cost_update": "k_update_wcost", "presum": "k_presum",
                 "wire_commit": "k_commit_edges"}
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2,
                    help="benchmark configuration index (0 = 200x200 20k nets, 1 = 800x800 500k nets, "
                         "2 = 1500x1500 3M nets with heavy nets and hotspots, 3 = 12 layers 2500x2500 8M nets)")
    ap.add_argument("--budget", type=float, default=380.0,
                    help="wall seconds (from process start) within which warmup and timed steps of the headline "
                         "configuration start; at least one timed step always runs")
    ap.add_argument("--compare-config", type=int, default=1,
                    help="also run both forms one step each on this configuration in the same process "
                         "(-1: off)")
    ap.add_argument("--nets", type=int, default=0, help="override the net count (testing)")
    ap.add_argument("--rrr", type=int, default=3)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--forms", default="batched",
                    help="forms timed on the headline configuration (plain,batched: both, with route identity)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip the untimed kernel-timing step")
    ap.add_argument("--serve", metavar="RATES",
                    help="serving benchmark instead: comma-separated Poisson request rates (req/s) against the "
                         "routed configuration-1 design")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests per rate (serving benchmark)")
    ap.add_argument("--tuning", default="",
                    help='launch shapes of the batched form as JSON, e.g. {"graph": 2} (stream launches, as '
                         'profiler runs use)')
    ap.add_argument("--cli", action="store_true",
                    help="also time the command-line stage (files in, route file out) on the same design")
    ap.add_argument("--tuning-file", default=os.path.join(ROOT, "instantgr_amd", "tuning.json"),
                    help="launch shapes chosen by the autotuner for this design (autotune.save format), used "
                         "when present and neither --tuning nor --autotune-nets is given")
    ap.add_argument("--autotune-nets", type=int, default=0,
                    help="before timing, tune the batched launch shapes on a sample design of this many nets "
                         "from the same generator (untimed)")
    ap.add_argument("--host-score", action="store_true",
                    help="also score the written route file with the host evaluator on designs above "
                         "configuration 1 (always done up to configuration 1)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (bandwidth reference, counter calibration kernel)")
    ap.add_argument("--summarize-stats", metavar="DIR",
                    help="print a summary of the rocprofv3 kernel statistics found under DIR and exit")
    ap.add_argument("--summarize-lds", metavar="DIR",
                    help="print the LDS bank-conflict rate per kernel from a rocprofv3 --pmc run of "
                         "SQ_LDS_BANK_CONFLICT and SQ_LDS_IDX_ACTIVE and exit")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the calibrated HBM traffic per dispatch from two rocprofv3 --pmc runs and exit")
    return ap.parse_args()


def _progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def one_step(design, form, rrr, threads, profile=False, tuning=None):
    from instantgr_amd import gpu
    t0 = time.perf_counter()
    design.build(threads)
    t1 = time.perf_counter()
    _progress("trees and pattern inputs built in %.3f s" % (t1 - t0))
    r = gpu.Router.from_design(design)
    if tuning:
        r.set_tuning(**tuning)
    if profile:
        r.set_profiling(True)
    st = r.run(design, form=form, rrr=rrr, threads=threads)
    t2 = time.perf_counter()
    st["tree_s"] = t1 - t0
    st["route_total_s"] = t2 - t1
    st["step_s"] = t2 - t0
    return r, st


def _traffic(kernel, config):
    """HBM bytes per dispatch of 'kernel' on benchmark configuration 'config'
    from the committed counter summary (lines 'traffic <kernel> <bytes>'),
    and the summary's name; (None, None) when there is none."""
    name = "traffic_config%d_batched.txt" % config
    path = os.path.join(PROFILES, name)
    if not kernel or not os.path.exists(path):
        return None, None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
            return float(f[2]), "profiles/" + name
    return None, None


def roofline(kt, kb, config):
    """Dominant kernel (by device time in the timed step, among the kernel
    classes with an algorithmic byte model) and its achieved bandwidth on
    those bytes; every class is listed beside it, with its counter traffic
    where a summary holds one.  The scheduler class is the sequence of round
    kernels of one scheduler call, timed as one range; its release-driven
    rounds count the bytes they move on the device when kernel timing is on.
    A class without a byte model that holds the most time is named in
    'largest_class'."""
    if not kt or all(v[0] <= 0 for v in kt.values()):
        return {"kernel": None, "bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "achieved": None, "frac": None,
                "traffic": None, "note": "kernel timing off (--no-profile)"}
    priced = {k: v for k, v in kt.items() if kb.get(k, 0) > 0 and v[0] > 0}
    name, (ms, launches) = max((priced or kt).items(), key=lambda kv: kv[1][0])
    out = {"kernel": name, "symbol": KERNEL_SYMBOL.get(name), "total_ms": ms, "launches": launches,
           "bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    top = max(kt.items(), key=lambda kv: kv[1][0])[0]
    if top != name:
        out["largest_class"] = top
    b = kb.get(name, 0.0)
    if b > 0 and ms > 0:
        achieved = b / (ms * 1e-3) / 1e9
        out.update({"achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": b / launches,
                    "avg_launch_us": ms * 1e3 / launches})
    else:
        out.update({"achieved": None, "frac": None})
    out["traffic"], src = _traffic(KERNEL_SYMBOL.get(name), config)
    if src:
        out["traffic_source"] = src
    classes = {}
    for k, (ms_k, n_k) in kt.items():
        if n_k and ms_k > 0:
            e = {"ms": round(ms_k, 3), "launches": n_k, "avg_launch_us": round(ms_k * 1e3 / n_k, 2)}
            if kb.get(k, 0) > 0:
                e["GB/s"] = round(kb[k] / (ms_k * 1e-3) / 1e9, 1)
                e["bytes_per_launch"] = kb[k] / n_k
            tr, _ = _traffic(KERNEL_SYMBOL.get(k), config)
            if tr:
                e["traffic"] = tr
            classes[k] = e
    out["classes"] = classes
    return out


def cpu_baseline():
    """The CPU oracle (sequential restatement of the reference logic, one
    thread) routing configuration 0 in full (10 metal layers, 200x200 gcells,
    20k nets): tree build + pattern pass + 3 rip-up-and-reroute passes."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from instantgr_amd.host import Design
    d = Design.preset(0)
    t0 = time.perf_counter()
    d.build(1)
    nets = d.info.nets
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
            "sample": "configuration 0 in full (10 metal layers, 200x200 gcells, %d nets), tree build + pattern pass "
                      "+ 3 RRR passes, one thread: %.1f s on %s (%d host cores)"
                      % (nets, dt, _cpu_model(), os.cpu_count() or 0)}


def _find(root, suffix):
    hits = []
    for dp, _, files in os.walk(root):
        hits += [os.path.join(dp, f) for f in files if f.endswith(suffix)]
    return sorted(hits)


def _short(name, with_template=False):
    """Kernel symbol out of a demangled name (namespaces and parameters
    dropped; template arguments kept on request)."""
    import re
    m = re.search(r"\b(k_[A-Za-z0-9_]+|__amd_[A-Za-z0-9_]+)(<[^()]*?>)?\(", name)
    if not m:
        return name[:40]
    return m.group(1) + ((m.group(2) or "") if with_template else "")


def summarize_stats(root):
    import csv
    rows = []
    for path in _find(root, "kernel_stats.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("%-44s %8s %12s %10s %7s" % ("kernel", "calls", "total_ms", "avg_us", "pct"))
    for r in rows[:30]:
        print("%-44s %8s %12.3f %10.2f %7.2f" % (_short(r["Name"], True)[:44], r["Calls"],
                                                  float(r["TotalDurationNs"]) / 1e6,
                                                  float(r["AverageNs"]) / 1e3,
                                                  100 * float(r["TotalDurationNs"]) / tot))
    print("%-44s %8s %12.3f" % ("all kernels", "", tot / 1e6))


def summarize_pmc(fetch_root, write_root, copy_bytes=1 << 30):
    """Average FETCH_SIZE / WRITE_SIZE (KiB) per dispatch of each kernel, each
    scaled by what the same counter reads for k_copy_probe, which reads and
    writes exactly copy_bytes per dispatch."""
    import csv
    from collections import defaultdict

    def load(root, counter):
        acc = defaultdict(list)
        for path in _find(root, "counter_collection.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") == counter:
                        acc[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}
    fetch, nf = load(fetch_root, "FETCH_SIZE")
    write, _ = load(write_root, "WRITE_SIZE")
    if "k_copy_probe" not in fetch or "k_copy_probe" not in write:
        raise SystemExit("k_copy_probe missing from the counter runs (run bench.py with --copy-probe)")
    ff = copy_bytes / (fetch["k_copy_probe"] * 1024.0)
    fw = copy_bytes / (write["k_copy_probe"] * 1024.0)
    print("# calibration: k_copy_probe reads and writes %d B per dispatch; FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB"
          " -> factors %.4f (fetch), %.4f (write)" % (copy_bytes, fetch["k_copy_probe"], write["k_copy_probe"],
                                                      ff, fw))
    print("%-24s %10s %16s %16s %18s" % ("kernel", "dispatches", "FETCH_KiB_avg", "WRITE_KiB_avg",
                                         "traffic_B_per_disp"))
    lines = []
    for k in sorted(fetch, key=lambda k: -fetch[k] * nf[k]):
        t = fetch[k] * 1024.0 * ff + write.get(k, 0.0) * 1024.0 * fw
        print("%-24s %10d %16.1f %16.1f %18.0f" % (k, nf[k], fetch[k], write.get(k, 0.0), t))
        lines.append("traffic %s %.0f" % (k, t))
    print("\n".join(lines))


def summarize_lds(root):
    """LDS bank-conflict rate per kernel from a rocprofv3 --pmc run of
    SQ_LDS_BANK_CONFLICT and SQ_LDS_IDX_ACTIVE: conflict cycles over active
    LDS cycles, summed over the kernel's dispatches."""
    import csv
    from collections import defaultdict
    acc = defaultdict(lambda: [0.0, 0.0, 0])
    for path in _find(root, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                k = _short(r["Kernel_Name"])
                if r.get("Counter_Name") == "SQ_LDS_BANK_CONFLICT":
                    acc[k][0] += float(r["Counter_Value"])
                    acc[k][2] += 1
                elif r.get("Counter_Name") == "SQ_LDS_IDX_ACTIVE":
                    acc[k][1] += float(r["Counter_Value"])
    print("%-24s %10s %18s %18s %12s" % ("kernel", "dispatches", "bank_conflict_cyc", "lds_active_cyc", "conflict"))
    for k, (c, t, n) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print("%-24s %10d %18.0f %18.0f %12s" % (k, n, c, t, "%.4f" % (c / t) if t > 0 else "no LDS"))


def cost_prefix_sweeps(design, reps=20, track_k=0):
    """Full-grid cost update + prefix sums in both forms, on the benchmark grid
    and on the largest grid the generator defines (12 metal layers,
    2500x2500).  GB/s over the algorithmic bytes per cell: 28 for the plain
    form's three kernels (cost update: read capacity and demand, write wire
    and via cost; prefix sums: read wire cost, write the FP64 sums), 24 for
    the fused kernel, which does not read the wire cost back."""
    import numpy as np
    from instantgr_amd import gpu
    out = []
    grids = [("benchmark grid", None), ("12 metal layers, 2500x2500", (11, 2500, 2500))]
    for name, shape in grids:
        if shape is None:
            r = gpu.Router.from_design(design)
        else:
            L, X, Y = shape
            rng = np.random.default_rng(7)
            cap = rng.integers(2, 7, size=L * X * Y).astype(np.float32)
            r = gpu.Router(L, X, Y, 1, 1.0, 4.0, np.ones(L), np.ones(X - 1, np.int32), np.ones(Y - 1, np.int32), cap)
        cells = r.n
        if track_k:
            r.set_tuning(track_k=track_k)
        row = {"grid": name, "cells": cells}
        for form, key, per in ((gpu.FORM_PLAIN, "plain", 28.0), (gpu.FORM_BATCHED, "fused", 24.0)):
            ms = r.cost_sweep_ms(form, reps)
            gbs = per * cells / (ms * 1e-3) / 1e9
            row.update({key + "_us": round(ms * 1e3, 2), key + "_GBs": round(gbs, 1),
                        key + "_frac": round(gbs / 8000.0, 4)})
        r.close()
        out.append(row)
        _progress("cost + prefix sweep, %s: %s" % (name, json.dumps(row)))
    return out


def serve_bench(a):
    """Serving: the design stays routed on the device; Poisson reroute requests
    of 1-5000 subnets from random 50x50 windows at each rate; p50/p99 latency,
    nets/s, and the final score against a one-request-at-a-time replay."""
    from instantgr_amd import gpu, serve
    from instantgr_amd.host import Design
    design = Design.preset(a.config, nets=a.nets).build(a.threads)
    rates = [float(x) for x in a.serve.split(",") if x.strip()]
    lines = []
    with tempfile.TemporaryDirectory() as tmp:
        for k, rate in enumerate(rates):
            reqs = serve.make_requests(design, a.serve_requests, rate, window=50, kmax=5000, seed=100 + k)
            r = gpu.Router.from_design(design)
            r.run(design, form=gpu.FORM_BATCHED, rrr=a.rrr, threads=a.threads)
            _progress("serving %.0f req/s: design routed" % rate)
            res = serve.serve(r, design, reqs, threads=a.threads, max_tick_nets=20000, hold_ms=0)
            p = os.path.join(tmp, "served.out")
            design.write_routes(p, *r.routes_flat())
            s_served = design.evaluate(p)["total"]
            r.close()
            r2 = gpu.Router.from_design(design)
            r2.run(design, form=gpu.FORM_BATCHED, rrr=a.rrr, threads=a.threads)
            serve.replay(r2, design, reqs, threads=a.threads)
            design.write_routes(p, *r2.routes_flat())
            s_replay = design.evaluate(p)["total"]
            r2.close()
            res.pop("results")
            res.update({"rate": rate, "score": s_served, "replay_score": s_replay,
                        "score_rel_diff": abs(s_served - s_replay) / max(1e-300, abs(s_replay))})
            _progress("serving %.0f req/s: %s" % (rate, json.dumps(res)))
            lines.append(res)
    print(json.dumps({"metric": "reroute request latency and throughput, resident routed design",
                      "config": {"workload": "configuration %d design, Poisson reroute requests of "
                                             "floor(exp(U(0, ln 5000))) subnets from 50x50 windows" % a.config,
                                 "requests_per_rate": a.serve_requests},
                      "rates": lines}), flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def run_form(design, form, name, a, tuning, budget_s, t_origin):
    """Warmup steps (the first of them with kernel timing on: the kernel
    times and algorithmic bytes of the roofline come from it, untimed), then
    timed steps, each a full pipeline on a fresh context.  A step starts only
    while the wall time since t_origin plus one step's time stays inside
    budget_s (at least one timed step runs).  Returns the measurements and the
    last timed step's router (routes, score)."""
    from instantgr_amd import gpu
    tun = tuning if name == "batched" else None
    est = None
    warm = 0
    kt = kb = None
    for w in range(a.warmup):
        if est is not None and time.perf_counter() - t_origin + 2 * est > budget_s:
            break
        prof = w == 0 and not a.no_profile
        r, st = one_step(design, form, a.rrr, a.threads, profile=prof, tuning=tun)
        if prof:
            kt, kb = r.kernel_times(), r.kernel_bytes()
            prof_routes = r.routes_flat()
        r.close()
        est = st["step_s"]
        warm += 1
        _progress("%s warmup %d: %.3f s%s" % (name, w, st["step_s"], " (kernel timing on)" if prof else ""))
    times, last = [], None
    for k in range(a.steps):
        if k > 0 and est is not None and time.perf_counter() - t_origin + est > budget_s:
            break
        r, st = one_step(design, form, a.rrr, a.threads, tuning=tun)
        times.append(st["step_s"])
        est = st["step_s"]
        free, total = r.device_memory()
        _progress("%s step %d: %.3f s (device memory in use %.1f of %.1f GB)"
                  % (name, k, st["step_s"], (total - free) / 1e9, total / 1e9))
        if last is not None:
            last[0].close()
        last = (r, st)
    r, st = last
    routes = r.routes_flat()
    t_ev = time.perf_counter()
    gscore = r.evaluate(design)
    gscore["eval_s"] = time.perf_counter() - t_ev
    _progress("%s GPU evaluator: %s" % (name, json.dumps(gscore)))
    r.close()   # before another context takes the device memory of a whole step
    if kt is None and not a.no_profile:
        # no warmup ran: one extra, untimed step with kernel timing
        r2, _ = one_step(design, form, a.rrr, a.threads, profile=True, tuning=tun)
        kt, kb = r2.kernel_times(), r2.kernel_bytes()
        prof_routes = r2.routes_flat()
        r2.close()
    if kt is not None and not gpu.routes_equal(prof_routes, routes):
        raise RuntimeError("the kernel-timing step routed differently")
    print("[bench] %s: %s" % (name, json.dumps({"step_s": times, "warmup": warm, "pattern_batches": st["pattern_batches"],
                                               "ripped": st["ripped"], "rrr_batches": st["rrr_batches"],
                                               "tree_s": st["tree_s"], "pattern_s": st["pattern_s"],
                                               "pattern_schedule_s": st["pattern_schedule_s"],
                                               "rrr_dag_build_s": st["rrr_dag_build_s"],
                                               "rrr_route_s": st["rrr_route_s"],
                                               "rrr_schedule_s": st["rrr_schedule_s"],
                                               "kernel_ms": {k: round(v[0], 1) for k, v in (kt or {}).items() if v[1]}})),
          file=sys.stderr, flush=True)
    return {"step_s": times, "warmup": warm, "stats": st, "kernels": kt or {}, "bytes": kb or {}, "gpu_score": gscore,
            "routes": routes}


def host_scores(design, results):
    """Each form's routes written as the route file and scored by the host
    evaluator, checked against the GPU evaluator's score."""
    scores = {}
    with tempfile.TemporaryDirectory() as tmp:
        for f, res in results.items():
            p = os.path.join(tmp, f + ".out")
            t_h = time.perf_counter()
            design.write_routes(p, *res["routes"])
            h = design.evaluate(p)
            h["write_eval_s"] = time.perf_counter() - t_h
            g = res["gpu_score"]
            if h["opens"] != g["opens"] or abs(h["total"] - g["total"]) > 1e-9 * abs(h["total"]):
                raise RuntimeError("GPU and host evaluators disagree: %s vs %s" % (g, h))
            scores[f] = dict(g, host=h)
    return scores


def same_routes(results):
    from instantgr_amd import gpu
    if len(results) != 2:
        return None
    (fa, ra), (fb, rb) = results.items()
    same = gpu.routes_equal(ra["routes"], rb["routes"])
    ga, gb = ra["gpu_score"], rb["gpu_score"]
    if same and (ga["opens"] != gb["opens"] or abs(ga["total"] - gb["total"]) > 1e-9 * abs(gb["total"])):
        raise RuntimeError("identical routes scored differently")
    return same


def workload_name(cfg, info, rrr):
    return ("configuration %d: %d metal layers, %dx%d gcells, %d nets; tree build + pattern pass + %d "
            "rip-up-and-reroute passes" % (cfg, info.metal_layers, info.X, info.Y, info.nets, rrr))


def comparison_run(cfg, a, t_origin):
    """Both forms on another configuration in the same process, one step each:
    plain vs batched time, identical routes, equal GPU and host scores."""
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    import argparse as _ap
    design = Design.preset(cfg)
    b = _ap.Namespace(**vars(a))
    b.warmup, b.steps, b.no_profile = 0, 1, True
    res = {}
    for name, form in (("plain", gpu.FORM_PLAIN), ("batched", gpu.FORM_BATCHED)):
        res[name] = run_form(design, form, name, b, None, float("inf"), t_origin)
    same = same_routes(res)
    sc = host_scores(design, res)
    info = design.info
    out = {"workload": workload_name(cfg, info, a.rrr), "subnets": int(info.subnets), "routes_identical": same,
           "plain_ms_per_step": 1e3 * res["plain"]["step_s"][0], "batched_ms_per_step": 1e3 * res["batched"]["step_s"][0],
           "speedup_batched_vs_plain": res["plain"]["step_s"][0] / res["batched"]["step_s"][0],
           "score": {f: sc[f] for f in sc}}
    _progress("comparison on configuration %d: %s" % (cfg, json.dumps({k: v for k, v in out.items() if k != "score"})))
    return out


def committed_plain_run():
    """The plain form's full step on the headline configuration from the
    committed summary of a run with both forms in one process (lines
    'plain_ms_per_step V', 'batched_ms_per_step V', 'routes_identical V')."""
    path = os.path.join(PROFILES, "bench_config2_both_forms.txt")
    if not os.path.exists(path):
        return None
    out = {"source": "profiles/bench_config2_both_forms.txt"}
    for line in open(path):
        f = line.split()
        if len(f) == 2 and f[0] in ("plain_ms_per_step", "batched_ms_per_step"):
            out[f[0]] = float(f[1])
        elif len(f) == 2 and f[0] == "routes_identical":
            out[f[0]] = f[1] == "true"
    if "plain_ms_per_step" in out and "batched_ms_per_step" in out:
        out["speedup_batched_vs_plain"] = out["plain_ms_per_step"] / out["batched_ms_per_step"]
    return out


def tuning_key(config, nets=0):
    """Key of a benchmark design in the autotuner's cache file."""
    return "config%d" % config + ("_nets%d" % nets if nets else "")


def main():
    a = parse()
    if a.summarize_stats:
        summarize_stats(a.summarize_stats)
        return
    if a.summarize_pmc:
        summarize_pmc(*a.summarize_pmc)
        return
    if a.summarize_lds:
        summarize_lds(a.summarize_lds)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: the routing path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("IGR_PROGRESS", "1")
    t_origin = time.perf_counter()
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    if a.serve:
        serve_bench(a)
        return
    copy_gbs = gpu.copy_probe(1 << 30, 20) if a.copy_probe else None
    t_gen = time.perf_counter()
    design = Design.preset(a.config, nets=a.nets)
    _progress("design generated in %.1f s" % (time.perf_counter() - t_gen))
    tuning = json.loads(a.tuning) if a.tuning else None
    tuning_source = "command line" if tuning else None
    if tuning is None and not a.autotune_nets and a.tuning_file and os.path.exists(a.tuning_file):
        with open(a.tuning_file) as f:
            cached = json.load(f).get(tuning_key(a.config, a.nets))
        if cached:
            tuning = cached["tuning"]
            tuning_source = "autotuner (%s)" % cached.get("sample", "cached")
            _progress("launch shapes from the autotuner's cache: %s" % json.dumps(tuning))
    if a.autotune_nets:
        from instantgr_amd import autotune
        t_at = time.perf_counter()
        sample = Design.preset(a.config, nets=a.autotune_nets).build(a.threads)
        tuning = autotune.tune(sample, rrr=1, threads=a.threads, log=_progress)["tuning"]
        _progress("autotuned on %d nets in %.1f s: %s" % (a.autotune_nets, time.perf_counter() - t_at,
                                                        json.dumps(tuning)))
        tuning_source = "autotuner (%d-net sample, this run)" % a.autotune_nets
    forms = [f.strip() for f in a.forms.split(",") if f.strip()]
    form_id = {"plain": gpu.FORM_PLAIN, "batched": gpu.FORM_BATCHED}
    results = {}
    for f in forms:
        results[f] = run_form(design, form_id[f], f, a, tuning, a.budget, t_origin)
    info = design.info
    nets_in = info.nets
    same = same_routes(results)
    # Host evaluator on the written route file beside the GPU evaluator (the
    # check that the file the command line writes scores the same); on the
    # large designs only when asked (two 11M-subnet files take minutes).
    scores = {f: results[f]["gpu_score"] for f in forms}
    if a.host_score or a.config <= 1:
        scores = host_scores(design, results)
    head = "batched" if "batched" in results else forms[-1]
    # Same-run plain vs batched on the 800x800 design, and the plain form's
    # full step on this design from the committed both-forms run.
    comparison = None
    if a.compare_config >= 0 and a.compare_config != a.config and not ("plain" in results and "batched" in results):
        comparison = comparison_run(a.compare_config, a, t_origin)
    sweeps = cost_prefix_sweeps(design, track_k=(tuning or {}).get("track_k", 0))
    cli = None
    if a.cli:
        cli = cli_run(design, a)
    step_ms = 1e3 * sum(results[head]["step_s"]) / len(results[head]["step_s"])
    value = nets_in / (step_ms * 1e-3)
    steps_run = len(results[head]["step_s"])
    out = {
        "metric": "nets routed/sec end-to-end (batched vs plain, same GPU) at identical eval score",
        "value": value, "unit": "nets/s", "n_gpus": 1, "steps": steps_run, "warmup": results[head]["warmup"],
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload_name(a.config, info, a.rrr), "subnets": int(info.subnets), "form": head,
                   "tuning": tuning, "tuning_source": tuning_source},
        "forms": {f: {"nets_per_s": nets_in / (sum(v["step_s"]) / len(v["step_s"])),
                      "ms_per_step": 1e3 * sum(v["step_s"]) / len(v["step_s"]),
                      "score": scores[f]} for f, v in results.items()},
        "routes_identical": same,
    }
    if steps_run != a.steps or results[head]["warmup"] != a.warmup:
        out["steps_requested"] = a.steps
        out["warmup_requested"] = a.warmup
        out["note"] = ("one step is the whole pipeline on this design (%.0f s); %d warmup and %d timed steps fit the "
                       "%.0f s budget (--budget), of the %d and %d asked for"
                       % (step_ms / 1e3, results[head]["warmup"], steps_run, a.budget, a.warmup, a.steps))
    if "plain" in results and "batched" in results:
        out["speedup_batched_vs_plain"] = out["forms"]["plain"]["ms_per_step"] / out["forms"]["batched"]["ms_per_step"]
    else:
        committed = committed_plain_run() if a.config == 2 else None
        if committed:
            out["plain_on_this_design"] = committed
    if comparison:
        out["same_run_comparison"] = comparison
    out["cli"] = cli
    out["cost_prefix_sweep"] = sweeps
    out["roofline"] = roofline(results[head]["kernels"], results[head]["bytes"], a.config)
    if copy_gbs is not None:
        out["roofline"]["copy_probe_GBs"] = copy_gbs
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline()
    print(json.dumps(out), flush=True)


def cli_run(design, a):
    """The drop-in command line: parse .cap/.net, trees, route, write the
    route file; wall time from process start to exit, I/O included."""
    import subprocess
    from instantgr_amd import build as igr_build
    exe = os.path.join(os.path.dirname(igr_build.__file__), "lib", "igr_route")
    with tempfile.TemporaryDirectory() as tmp:
        cap, net, outp = (os.path.join(tmp, n) for n in ("d.cap", "d.net", "d.out"))
        design.save(cap, net)
        t_cli = time.perf_counter()
        proc = subprocess.run([exe, "-cap", cap, "-net", net, "-out", outp, "-form", "batched",
                               "-rrr", str(a.rrr), "-threads", str(a.threads)], capture_output=True, text=True)
        wall = time.perf_counter() - t_cli
        if proc.returncode != 0:
            raise RuntimeError("igr_route failed: " + proc.stderr)
        cli = {"wall_s": wall, "files_mb": (os.path.getsize(cap) + os.path.getsize(net) +
                                            os.path.getsize(outp)) / 1e6,
               "score": design.evaluate(outp)["total"], "stages": proc.stdout.strip()}
        _progress("command line: %.3f s (%s)" % (wall, cli["stages"]))
    return cli


if __name__ == "__main__":
    main()
