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
    ap.add_argument("--config", type=int, default=1, help="benchmark configuration index (1 = 800x800, 500k nets)")
    ap.add_argument("--nets", type=int, default=0, help="override the net count (testing)")
    ap.add_argument("--rrr", type=int, default=3)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--forms", default="plain,batched")
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


def _traffic(kernel):
    """HBM bytes per dispatch of 'kernel' from the committed counter summary
    (lines 'traffic <kernel> <bytes>'), or None."""
    path = os.path.join(PROFILES, "traffic_config1_batched.txt")
    if not kernel or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
            return float(f[2])
    return None


def roofline(kt, kb):
    """Dominant kernel (by device time in the timed step, among the kernel
    classes with an algorithmic byte model) and its achieved bandwidth on
    those bytes; every class is listed beside it.  The scheduler class is a
    sequence of small round kernels timed as one group, with no byte model:
    when it holds the most time it is named in 'largest_class'."""
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
    out["traffic"] = _traffic(KERNEL_SYMBOL.get(name))
    classes = {}
    for k, (ms_k, n_k) in kt.items():
        if n_k and ms_k > 0:
            e = {"ms": round(ms_k, 3), "launches": n_k, "avg_launch_us": round(ms_k * 1e3 / n_k, 2)}
            if kb.get(k, 0) > 0:
                e["GB/s"] = round(kb[k] / (ms_k * 1e-3) / 1e9, 1)
            classes[k] = e
    out["classes"] = classes
    return out


def cpu_baseline():
    """The CPU oracle (sequential restatement of the reference logic, one
    thread) routing a bounded sample of the same workload: the generator's
    configuration-1 distributions at the same net density on a 176x176 grid
    (24,200 nets), tree build + pattern pass + 3 rip-up-and-reroute passes."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from instantgr_amd.host import Design
    nets = 24200
    d = Design.custom(layers=10, xsize=176, ysize=176, nets=nets, pin_cap=64, seed=2)
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
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": nets / dt, "unit": "nets/s", "cores": 1, "kind": "port",
            "sample": "10 layers, 176x176 gcells, 24200 nets (configuration-1 generator and net density), "
                      "tree build + pattern pass + 3 RRR passes, one thread; %.1f s on %s (%d host cores)"
                      % (dt, cpu or "unknown CPU", os.cpu_count() or 0)}


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


def cost_prefix_sweeps(design, reps=20):
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
    if a.autotune_nets:
        from instantgr_amd import autotune
        t_at = time.perf_counter()
        sample = Design.preset(a.config, nets=a.autotune_nets).build(a.threads)
        tuning = autotune.tune(sample, rrr=1, threads=a.threads, log=_progress)["tuning"]
        _progress("autotuned on %d nets in %.1f s: %s" % (a.autotune_nets, time.perf_counter() - t_at,
                                                        json.dumps(tuning)))
    forms = [f.strip() for f in a.forms.split(",") if f.strip()]
    form_id = {"plain": gpu.FORM_PLAIN, "batched": gpu.FORM_BATCHED}
    results = {}
    routes = {}
    for f in forms:
        for w in range(a.warmup):
            r, _st = one_step(design, form_id[f], a.rrr, a.threads, tuning=tuning if f == "batched" else None)
            r.close()
            _progress("%s warmup %d: %.3f s" % (f, w, _st["step_s"]))
        times = []
        last = None
        for k in range(a.steps):
            r, st = one_step(design, form_id[f], a.rrr, a.threads, tuning=tuning if f == "batched" else None)
            times.append(st["step_s"])
            _progress("%s step %d: %.3f s" % (f, k, st["step_s"]))
            if last is not None:
                last[0].close()
            last = (r, st)
        r, st = last
        routes[f] = r.routes_flat()
        # Score of the routes on the device (GPU evaluator), before the
        # context goes.
        t_ev = time.perf_counter()
        gscore = r.evaluate(design)
        gscore["eval_s"] = time.perf_counter() - t_ev
        _progress("%s GPU evaluator: %s" % (f, json.dumps(gscore)))
        if not a.no_profile:
            # Kernel timing (events on the stream, no graph) in one extra,
            # untimed step.
            r.close()
            r, pst = one_step(design, form_id[f], a.rrr, a.threads, profile=True,
                              tuning=tuning if f == "batched" else None)
            _progress("%s profiled step: %.3f s" % (f, pst["step_s"]))
            if not gpu.routes_equal(r.routes_flat(), routes[f]):
                raise RuntimeError("profiled step routed differently")
        kt = r.kernel_times()
        kb = r.kernel_bytes()
        results[f] = {"step_s": times, "stats": st, "kernels": kt, "bytes": kb, "gpu_score": gscore}
        print("[bench] %s: %s" % (f, json.dumps({"step_s": times, "pattern_batches": st["pattern_batches"],
                                                   "ripped": st["ripped"], "rrr_batches": st["rrr_batches"],
                                                   "tree_s": st["tree_s"],
                                                   "pattern_s": st["pattern_s"],
                                                   "rrr_dag_build_s": st["rrr_dag_build_s"],
                                                   "rrr_route_s": st["rrr_route_s"],
                                                   "rrr_schedule_s": st["rrr_schedule_s"],
                                                   "rrr_prep_s": st["rrr_prep_s"],
                                                   "rrr_prep_parts": st["rrr_prep_parts"],
                                                   "pattern_schedule_s": st["pattern_schedule_s"],
                                                   "kernel_ms": {k: round(v[0], 1) for k, v in kt.items() if v[1]}})),
              file=sys.stderr, flush=True)
        r.close()

    info = design.info
    nets_in = info.nets
    # Identical routes and score across forms.  The score is the GPU
    # evaluator's; the host evaluator scores the written route file beside it
    # (the check that the file the command line writes scores the same),
    # skipped on designs above configuration 1 unless asked for.
    same = None
    scores = {f: results[f]["gpu_score"] for f in forms}
    if a.host_score or a.config <= 1:
        with tempfile.TemporaryDirectory() as tmp:
            for f in forms:
                off, data = routes[f]
                p = os.path.join(tmp, f + ".out")
                t_h = time.perf_counter()
                design.write_routes(p, off, data)
                h = design.evaluate(p)
                h["write_eval_s"] = time.perf_counter() - t_h
                g = scores[f]
                if h["opens"] != g["opens"] or abs(h["total"] - g["total"]) > 1e-9 * abs(h["total"]):
                    raise RuntimeError("GPU and host evaluators disagree: %s vs %s" % (g, h))
                scores[f] = dict(g, host=h)
    if len(forms) == 2:
        same = gpu.routes_equal(routes[forms[0]], routes[forms[1]])
        g0, g1 = (results[f]["gpu_score"] for f in forms)
        if same and (g0["opens"] != g1["opens"] or abs(g0["total"] - g1["total"]) > 1e-9 * abs(g1["total"])):
            raise RuntimeError("identical routes scored differently")
    sweeps = cost_prefix_sweeps(design)
    cli = None
    if a.cli:
        # The drop-in command line: parse .cap/.net, trees, route, write the
        # route file; wall time from process start to exit, I/O included.
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
    head = "batched" if "batched" in results else forms[-1]
    step_ms = 1e3 * sum(results[head]["step_s"]) / len(results[head]["step_s"])
    value = nets_in / (step_ms * 1e-3)
    out = {
        "metric": "nets routed/sec end-to-end (batched vs plain, same GPU) at identical eval score",
        "value": value, "unit": "nets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "configuration %d: %d metal layers, %dx%d gcells, %d nets; tree build + pattern "
                               "pass + %d rip-up-and-reroute passes" % (a.config, info.metal_layers, info.X, info.Y,
                                                                        nets_in, a.rrr),
                   "subnets": int(info.subnets), "form": head, "tuning": tuning},
        "forms": {f: {"nets_per_s": nets_in / (sum(v["step_s"]) / len(v["step_s"])),
                      "ms_per_step": 1e3 * sum(v["step_s"]) / len(v["step_s"]),
                      "score": scores[f]} for f, v in results.items()},
        "routes_identical": same,
        "cli": cli,
        "cost_prefix_sweep": sweeps,
        "roofline": roofline(results[head]["kernels"], results[head]["bytes"]),
    }
    if copy_gbs is not None:
        out["roofline"]["copy_probe_GBs"] = copy_gbs
    if "plain" in results and "batched" in results:
        out["speedup_batched_vs_plain"] = out["forms"]["plain"]["ms_per_step"] / out["forms"]["batched"]["ms_per_step"]
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
