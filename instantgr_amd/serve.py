"""Serving workload: Poisson reroute requests against a resident routed design.

A request rips up and reroutes k subnets, k = floor(exp(U(0, ln kmax))), picked
from the subnets whose pin-box centre lies in a random window x window gcell
region.  Requests arrive as a Poisson stream at a given rate; `serve` submits
them at their arrival times to a Server (instantgr_amd.gpu.Server) and reports
per-request latency (completion minus arrival), p50 / p99 and nets/s.
`replay` routes the same request sequence one request at a time, the offline
reference the served result is scored against.
"""
import math
import threading
import time

import numpy as np

from . import gpu


def subnet_centres(design):
    """Per subnet the (x, y) centre of its pins' box."""
    i = design.info
    XY, Y = i.X * i.Y, i.Y
    off = design.array("sub_pin_off")
    pins = design.array("sub_pins")
    pos = pins % XY
    x, y = pos // Y, pos % Y
    n = len(off) - 1
    seg = np.repeat(np.arange(n), np.diff(off))
    x0 = np.full(n, np.iinfo(np.int32).max)
    x1 = np.full(n, -1)
    y0 = np.full(n, np.iinfo(np.int32).max)
    y1 = np.full(n, -1)
    np.minimum.at(x0, seg, x)
    np.maximum.at(x1, seg, x)
    np.minimum.at(y0, seg, y)
    np.maximum.at(y1, seg, y)
    return (x0 + x1) // 2, (y0 + y1) // 2


def make_requests(design, count, rate, window=50, kmax=5000, seed=0):
    """count requests: (arrival offset in s, subnet array), Poisson at rate/s."""
    rng = np.random.default_rng(seed)
    cx, cy = subnet_centres(design)
    X, Y = design.info.X, design.info.Y
    order = np.lexsort((cy, cx))
    sx, sy = cx[order], cy[order]
    out = []
    t = 0.0
    while len(out) < count:
        t += rng.exponential(1.0 / rate)
        k = int(math.floor(math.exp(rng.uniform(0.0, math.log(kmax)))))
        wx, wy = int(rng.integers(0, max(1, X - window))), int(rng.integers(0, max(1, Y - window)))
        lo, hi = np.searchsorted(sx, wx), np.searchsorted(sx, wx + window)
        cand = order[lo:hi][(sy[lo:hi] >= wy) & (sy[lo:hi] < wy + window)]
        if len(cand) == 0:
            continue
        pick = rng.choice(cand, size=min(k, len(cand)), replace=False)
        out.append((t, np.sort(pick).astype(np.int32)))
    return out


def serve(router, design, requests, threads=8, max_tick_nets=20000, scratch_bytes=0, hold_ms=0, overtake=False):
    """Submit requests at their arrival offsets; latency and throughput."""
    srv = gpu.Server(router, design, threads=threads, max_tick_nets=max_tick_nets, scratch_bytes=scratch_bytes,
                     hold_ms=hold_ms, overtake=overtake)
    ids = []
    t0 = time.perf_counter()
    err = []

    def feeder():
        try:
            for off, nets in requests:
                dt = t0 + off - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append((srv.submit(nets), time.perf_counter()))
        except Exception as e:  # surfaced below
            err.append(e)

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    results = []
    try:
        for rid, _ in ids:
            results.append(srv.wait(rid))
        stats = srv.stats()
    finally:
        srv.close()
    if err:
        raise err[0]
    t_end = max(r["done_s"] for r in results)
    t_first = min(r["submit_s"] for r in results)
    lat = np.array([r["latency_s"] for r in results])
    nets = sum(r["nets"] for r in results)
    return {"requests": len(results), "nets": int(nets),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "nets_per_s": nets / max(1e-9, t_end - t_first),
            "ticks": stats["ticks"], "max_tick_nets": stats["max_tick_nets"],
            "requests_per_tick": len(results) / max(1, stats["ticks"]), "results": results}


def replay(router, design, requests, threads=8):
    """The same requests one at a time, in arrival order."""
    srv = gpu.Server(router, design, threads=threads)
    try:
        for _, nets in requests:
            srv.wait(srv.submit(nets))
    finally:
        srv.close()
