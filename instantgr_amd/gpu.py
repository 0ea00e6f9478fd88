"""ctypes binding of libigr.so, the HIP routing library.

Raises at load time when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_lib = None

FORM_PLAIN = 0
FORM_BATCHED = 1


class GridDesc(C.Structure):
    _fields_ = [("L", C.c_int), ("X", C.c_int), ("Y", C.c_int), ("DIR", C.c_int),
                ("unit_wire", C.c_double), ("unit_via", C.c_double),
                ("of_weight", C.c_void_p), ("xlen", C.c_void_p), ("ylen", C.c_void_p), ("capacity", C.c_void_p)]


class NetDesc(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                                      ("node_off", "nodes", "parents", "pt_off", "pts", "hs_off", "hs", "vs_off", "vs")]


class DagDesc(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                                      ("node_off", "nodes", "par_off", "par", "slot", "child_num", "depth",
                                       "pt_off", "pts")]


class PassStats(C.Structure):
    _fields_ = [("batches", C.c_int), ("schedule_s", C.c_double), ("route_s", C.c_double), ("nets", C.c_int64)]


MAX_RRR = 8
KCLASSES = ["cost_update", "presum", "lshape_dp", "wire_commit", "via_commit", "dag_edges", "dag_dp",
            "dag_extract", "scheduler", "track_fused", "ripup"]


class RunOpts(C.Structure):
    _fields_ = [("form", C.c_int), ("rrr_iterations", C.c_int), ("threads", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("graph", C.c_int), ("sched_group", C.c_int), ("sched_chunk", C.c_int),
                ("dag_order", C.c_int), ("dag_waves", C.c_int), ("track_k", C.c_int), ("dag_edges", C.c_int),
                ("dag_seq", C.c_int), ("sched_kind", C.c_int), ("presum_k", C.c_int), ("dag_grid", C.c_int),
                ("lshape_groups", C.c_int)]


TUNING_FIELDS = [f for f, _ in Tuning._fields_]


class EvalNets(C.Structure):
    _fields_ = [("nets", C.c_int64), ("sub_off", C.c_void_p), ("subs", C.c_void_p), ("pin_off", C.c_void_p),
                ("pins", C.c_void_p), ("extra_off", C.c_void_p), ("extra", C.c_void_p), ("dpin_off", C.c_void_p),
                ("ap_off", C.c_void_p), ("ap", C.c_void_p), ("dir", C.c_void_p), ("of_weight", C.c_void_p),
                ("capacity", C.c_void_p)]


class Score(C.Structure):
    _fields_ = [("wirelength", C.c_double), ("via", C.c_double), ("overflow", C.c_double), ("total", C.c_double),
                ("opens", C.c_int64)]


class ServeOpts(C.Structure):
    _fields_ = [("threads", C.c_int), ("max_tick_nets", C.c_int64), ("scratch_bytes", C.c_int64),
                ("hold_ms", C.c_int), ("overtake", C.c_int)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("nets", C.c_int64), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("done_s", C.c_double), ("latency_s", C.c_double), ("tick", C.c_int), ("tick_nets", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("ticks", C.c_int64), ("requests", C.c_int64), ("nets", C.c_int64), ("max_tick_nets", C.c_int64),
                ("prep_s", C.c_double), ("route_s", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("upload_s", C.c_double), ("pattern_s", C.c_double), ("pattern_schedule_s", C.c_double),
                ("total_s", C.c_double), ("pattern_batches", C.c_int), ("iterations", C.c_int),
                ("subnets", C.c_int64), ("ripped", C.c_int64 * MAX_RRR), ("rrr_batches", C.c_int * MAX_RRR),
                ("rrr_dag_nodes", C.c_int64 * MAX_RRR), ("rrr_view_s", C.c_double * MAX_RRR),
                ("rrr_dag_build_s", C.c_double * MAX_RRR), ("rrr_route_s", C.c_double * MAX_RRR),
                ("rrr_schedule_s", C.c_double * MAX_RRR), ("rrr_prep_s", C.c_double * MAX_RRR),
                ("rrr_prep_parts", (C.c_double * 4) * MAX_RRR), ("rrr_dag_rest_s", C.c_double * MAX_RRR)]


# Every entry point of include/igr.h with its ctypes signature.
SIGNATURES = {
    "igr_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    "igr_destroy": (None, [C.c_void_p]),
    "igr_last_error": (C.c_char_p, []),
    "igr_set_demand": (C.c_int, [C.c_void_p, C.c_void_p]),
    "igr_get_demand": (C.c_int, [C.c_void_p, C.c_void_p]),
    "igr_get_timestamp": (C.c_int, [C.c_void_p, C.c_void_p]),
    "igr_get_costs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "igr_update_cost": (C.c_int, [C.c_void_p]),
    "igr_compute_presum": (C.c_int, [C.c_void_p]),
    "igr_set_nets": (C.c_int, [C.c_void_p, C.POINTER(NetDesc)]),
    "igr_schedule": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p,
                               C.POINTER(C.c_int)]),
    "igr_route_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    "igr_pattern_route": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(PassStats)]),
    "igr_dag_pass": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.POINTER(DagDesc),
                               C.POINTER(PassStats)]),
    "igr_dag_pass_overlapped": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(DagDesc),
                                          C.c_void_p, C.c_void_p, C.POINTER(PassStats)]),
    "igr_get_route": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "igr_get_routes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "igr_evaluate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "igr_routes_capacity": (C.c_int64, [C.c_void_p]),
    "igr_ripup_mark": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "igr_congestion_view": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "igr_congestion_view_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "igr_stamp": (C.c_int, [C.c_void_p]),
    "igr_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "igr_device_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "igr_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "igr_kernel_bytes": (C.c_int, [C.c_void_p, C.c_void_p]),
    "igr_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RunOpts), C.POINTER(RunStats)]),
    "igr_expf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "igr_copy_probe": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "igr_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "igr_cost_sweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "igr_server_open": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(ServeOpts), C.POINTER(C.c_void_p)]),
    "igr_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "igr_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RequestResult)]),
    "igr_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "igr_server_close": (C.c_int, [C.c_void_p]),
    "igr_server_error": (C.c_char_p, []),
    "igr_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
}


def lib_path():
    return os.path.join(_LIB_DIR, "libigr.so")


def lib():
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libigr.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("libigr: %s (status %d)" % (lib().igr_last_error().decode(), rc))


def _ptr(a):
    return None if a is None else a.ctypes.data


def expf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    _check(lib().igr_expf(x.ctypes.data, out.ctypes.data, x.size))
    return out


def copy_probe(nbytes=1 << 30, reps=20):
    """GB/s (read + write) of a plain device copy of nbytes."""
    g = C.c_double()
    _check(lib().igr_copy_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def routes_words(flat):
    """All route words of an (off, data) route set in net order, and the
    per-net lengths (data[off[i]] words for net i, the length word included)."""
    off, data = flat
    lens = data[off[:-1]].astype(np.int64)
    starts = off[:-1].astype(np.int64)
    idx = np.repeat(starts - np.concatenate(([0], np.cumsum(lens)[:-1])), lens) + np.arange(lens.sum())
    return data[idx], lens


def routes_equal(a, b):
    """Route-by-route equality of two (off, data) route sets."""
    (wa, la), (wb, lb) = routes_words(a), routes_words(b)
    return bool(np.array_equal(la, lb) and np.array_equal(wa, wb))


class Router:
    """One routing context on the current GPU (grid state + nets)."""

    def __init__(self, L, X, Y, DIR, unit_wire, unit_via, w, xlen, ylen, cap):
        self._keep = [np.ascontiguousarray(w, np.float64), np.ascontiguousarray(xlen, np.int32),
                      np.ascontiguousarray(ylen, np.int32), np.ascontiguousarray(cap, np.float32)]
        d = GridDesc(L, X, Y, DIR, unit_wire, unit_via, *[k.ctypes.data for k in self._keep])
        h = C.c_void_p()
        _check(lib().igr_create(C.byref(d), C.byref(h)))
        self.h = h
        self.L, self.X, self.Y, self.DIR = L, X, Y, DIR
        self.n = L * X * Y
        self.num_nets = 0

    @classmethod
    def from_design(cls, d):
        i = d.info
        cap, w, xlen, ylen = d.grid_arrays()
        return cls(i.L, i.X, i.Y, i.DIR, i.unit_wire, i.unit_via, w, xlen, ylen, cap)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.igr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_demand(self, d):
        d = np.ascontiguousarray(d, np.float32)
        if d.size != self.n:
            raise ValueError("demand needs L*X*Y = %d values, got %d" % (self.n, d.size))
        _check(lib().igr_set_demand(self.h, d.ctypes.data))

    def demand(self):
        out = np.empty(self.n, np.float32)
        _check(lib().igr_get_demand(self.h, out.ctypes.data))
        return out

    def timestamp(self):
        out = np.empty(self.n, np.int32)
        _check(lib().igr_get_timestamp(self.h, out.ctypes.data))
        return out

    def costs(self):
        wc = np.empty(self.n, np.float32)
        vc = np.empty(self.n, np.float32)
        ps = np.empty(self.n, np.float64)
        _check(lib().igr_get_costs(self.h, wc.ctypes.data, vc.ctypes.data, ps.ctypes.data))
        return wc, vc, ps

    def update_cost(self):
        _check(lib().igr_update_cost(self.h))

    def compute_presum(self):
        _check(lib().igr_compute_presum(self.h))

    def set_nets(self, node_off, node, parent, pt_off=None, pt=None, hs_off=None, hs=None, vs_off=None, vs=None):
        arrs = [np.ascontiguousarray(node_off, np.int64), np.ascontiguousarray(node, np.int32),
                np.ascontiguousarray(parent, np.int32)]
        for off, v in ((pt_off, pt), (hs_off, hs), (vs_off, vs)):
            arrs += [None if off is None else np.ascontiguousarray(off, np.int64),
                     None if v is None else np.ascontiguousarray(v, np.int32)]
        self._nets = arrs
        d = NetDesc(len(arrs[0]) - 1, *[_ptr(a) for a in arrs])
        _check(lib().igr_set_nets(self.h, C.byref(d)))
        self.num_nets = len(arrs[0]) - 1

    def set_nets_from_design(self, d):
        a = d.array
        self.set_nets(a("node_off"), a("node"), a("parent"), a("pt_off"), a("pt"), a("hs_off"), a("hs"),
                      a("vs_off"), a("vs"))

    def schedule(self, order, mode=0, form=FORM_PLAIN):
        order = np.ascontiguousarray(order, np.int32)
        out = np.empty(len(order), np.int32)
        nb = C.c_int()
        _check(lib().igr_schedule(self.h, order.ctypes.data, len(order), mode, form, out.ctypes.data,
                                  C.byref(nb)))
        return nb.value, out

    def route_batch(self, nets, form=FORM_PLAIN):
        nets = np.ascontiguousarray(nets, np.int32)
        _check(lib().igr_route_batch(self.h, nets.ctypes.data, len(nets), form))

    def pattern_route(self, order, form=FORM_PLAIN):
        order = np.ascontiguousarray(order, np.int32)
        st = PassStats()
        _check(lib().igr_pattern_route(self.h, order.ctypes.data, len(order), form, C.byref(st)))
        return {"batches": st.batches, "schedule_s": st.schedule_s, "route_s": st.route_s, "nets": st.nets}

    def dag_pass(self, nets, dags, mode=0, form=FORM_PLAIN):
        """One rip-up-and-reroute pass over the DAGs returned by
        instantgr_amd.host.Design.build_dags for the same net list."""
        nets = np.ascontiguousarray(nets, np.int32)
        keys = (("dag_node_off", np.int64), ("dag_node", np.int32), ("dag_par_off", np.int64),
                ("dag_par", np.int32), ("dag_slot", np.int32), ("dag_child_num", np.int32),
                ("dag_depth", np.int32), ("dag_pt_off", np.int64), ("dag_pt", np.int32))
        arrs = [np.ascontiguousarray(dags[k], dt) for k, dt in keys]
        node_off, node, par_off, par, slot, cnum, depth, pt_off, pt = arrs
        n = len(nets)
        if len(node_off) != n + 1 or len(pt_off) != n + 1:
            raise ValueError("dag_node_off and dag_pt_off need one entry per net plus one")
        nodes = int(node_off[-1])
        if len(node) < nodes or len(cnum) < nodes or len(depth) < nodes or len(par_off) != nodes + 1:
            raise ValueError("DAG node arrays shorter than dag_node_off says")
        if len(par) < par_off[-1] or len(slot) < par_off[-1] or len(pt) < pt_off[-1]:
            raise ValueError("DAG edge or point arrays shorter than their offsets say")
        d = DagDesc(n, *[a.ctypes.data for a in arrs])
        st = PassStats()
        _check(lib().igr_dag_pass(self.h, nets.ctypes.data, len(nets), mode, form, C.byref(d), C.byref(st)))
        return {"batches": st.batches, "schedule_s": st.schedule_s, "route_s": st.route_s, "nets": st.nets}

    def dag_pass_overlapped(self, nets, points, complete, mode=0):
        """igr_dag_pass_overlapped (batched form): 'points' holds dag_pt_off and
        dag_pt; complete() is called once, on the pass's helper thread while
        the scheduler runs, and returns the other DAG arrays (the dict of
        Design.build_dags with part DAG_STRUCTURE).  An exception in it fails
        the pass."""
        nets = np.ascontiguousarray(nets, np.int32)
        pt_off = np.ascontiguousarray(points["dag_pt_off"], np.int64)
        pt = np.ascontiguousarray(points["dag_pt"], np.int32)
        n = len(nets)
        if len(pt_off) != n + 1 or len(pt) < pt_off[-1]:
            raise ValueError("dag_pt_off needs one entry per net plus one, dag_pt as many points as it says")
        keys = (("dag_node_off", np.int64), ("dag_node", np.int32), ("dag_par_off", np.int64),
                ("dag_par", np.int32), ("dag_slot", np.int32), ("dag_child_num", np.int32),
                ("dag_depth", np.int32))
        held = []
        errors = []

        def fill(_user, desc):
            try:
                got = complete()
                arrs = [np.ascontiguousarray(got[k], dt) for k, dt in keys]
                node_off, node, par_off, par, slot, cnum, depth = arrs
                nodes = int(node_off[-1])
                if len(node_off) != n + 1 or len(par_off) != nodes + 1 or len(node) < nodes or \
                        len(cnum) < nodes or len(depth) < nodes or len(par) < par_off[-1] or len(slot) < par_off[-1]:
                    raise ValueError("DAG arrays shorter than their offsets say")
                held.extend(arrs)
                d = desc.contents
                (d.node_off, d.nodes, d.par_off, d.par, d.slot, d.child_num, d.depth) = [a.ctypes.data for a in arrs]
                return 0
            except Exception as e:  # reported after the call; the pass fails
                errors.append(e)
                return -1

        cb = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(DagDesc))(fill)
        d = DagDesc(n, None, None, None, None, None, None, None, pt_off.ctypes.data, pt.ctypes.data)
        st = PassStats()
        rc = lib().igr_dag_pass_overlapped(self.h, nets.ctypes.data, n, mode, C.byref(d),
                                           C.cast(cb, C.c_void_p), None, C.byref(st))
        if errors:
            raise errors[0]
        _check(rc)
        return {"batches": st.batches, "schedule_s": st.schedule_s, "route_s": st.route_s, "nets": st.nets}

    def route(self, net):
        buf = np.empty(65536, np.int32)
        ln = C.c_int()
        _check(lib().igr_get_route(self.h, net, buf.ctypes.data, buf.size, C.byref(ln)))
        return buf[:ln.value].copy()

    def routes_flat(self):
        """(off, data): net i's route is data[off[i] : off[i] + data[off[i]]]."""
        cap = lib().igr_routes_capacity(self.h)
        off = np.empty(self.num_nets + 1, np.int64)
        data = np.empty(max(cap, 1), np.int32)
        _check(lib().igr_get_routes(self.h, off.ctypes.data, data.ctypes.data, data.size))
        return off, data

    def evaluate(self, design, routes=None):
        """Score of the routes on the device (igr_evaluate): the routes the
        context holds, or routes=(off, data) in routes_flat's layout.  The
        same definitions as design.evaluate(route_file)."""
        keep = []

        def ptr(name, dt):
            a = np.ascontiguousarray(design.array(name), dt)
            keep.append(a)
            return a.ctypes.data

        e = EvalNets(design.info.nets, ptr("net_sub_off", np.int64), ptr("net_sub", np.int32),
                     ptr("pin_off", np.int64), ptr("pins", np.int32), ptr("extra_off", np.int64),
                     ptr("extra", np.int32), ptr("net_pin_off", np.int64), ptr("ap_off", np.int64),
                     ptr("ap", np.int32), ptr("metal_dir", np.int32), ptr("metal_of_weight", np.float64),
                     ptr("metal_cap", np.float64))
        s = Score()
        if routes is None:
            _check(lib().igr_evaluate(self.h, C.byref(e), None, None, C.byref(s)))
        else:
            off = np.ascontiguousarray(routes[0], np.int64)
            data = np.ascontiguousarray(routes[1], np.int32)
            _check(lib().igr_evaluate(self.h, C.byref(e), off.ctypes.data, data.ctypes.data, C.byref(s)))
        return {"wirelength": s.wirelength, "via": s.via, "overflow": s.overflow, "total": s.total,
                "opens": int(s.opens)}

    def ripup_mark(self, th=0):
        out = np.empty(self.num_nets, np.uint8)
        _check(lib().igr_ripup_mark(self.h, th, out.ctypes.data))
        return out

    def congestion_view(self):
        xs = np.empty(self.X * self.Y, np.float32)
        ys = np.empty(self.X * self.Y, np.float32)
        _check(lib().igr_congestion_view(self.h, xs.ctypes.data, ys.ctypes.data))
        return xs, ys

    def congestion_view_noalloc(self):
        """congestion_view through the entry point that allocates no device
        memory (igr_congestion_view_noalloc)."""
        xs = np.empty(self.X * self.Y, np.float32)
        ys = np.empty(self.X * self.Y, np.float32)
        _check(lib().igr_congestion_view_noalloc(self.h, xs.ctypes.data, ys.ctypes.data))
        return xs, ys

    def set_profiling(self, on=True):
        _check(lib().igr_set_profiling(self.h, 1 if on else 0))

    def device_memory(self):
        """(free, total) bytes of the context's GPU."""
        free, total = C.c_size_t(), C.c_size_t()
        _check(lib().igr_device_memory(self.h, C.byref(free), C.byref(total)))
        return free.value, total.value

    def kernel_times(self):
        ms = (C.c_double * len(KCLASSES))()
        ln = (C.c_int64 * len(KCLASSES))()
        _check(lib().igr_kernel_times(self.h, ms, ln))
        return {k: (ms[i], ln[i]) for i, k in enumerate(KCLASSES)}

    def kernel_bytes(self):
        """Algorithmic HBM bytes per kernel class since profiling was enabled."""
        b = (C.c_double * len(KCLASSES))()
        _check(lib().igr_kernel_bytes(self.h, b))
        return {k: b[i] for i, k in enumerate(KCLASSES)}

    def set_tuning(self, **kw):
        """Launch-shape overrides of the batched form (igr_tuning fields; 0 = built-in)."""
        cur = {f: 0 for f in TUNING_FIELDS}
        cur.update(getattr(self, "_tuning", {}))
        for k, v in kw.items():
            if k not in cur:
                raise KeyError(k)
            cur[k] = int(v)
        t = Tuning(*[cur[f] for f in TUNING_FIELDS])
        _check(lib().igr_set_tuning(self.h, C.byref(t)))
        self._tuning = cur

    def tuning(self):
        t = Tuning()
        _check(lib().igr_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f in TUNING_FIELDS}

    def cost_sweep_ms(self, form=FORM_BATCHED, reps=20):
        """Milliseconds per full-grid cost update + prefix sums (28 algorithmic bytes per cell)."""
        ms = C.c_double()
        _check(lib().igr_cost_sweep(self.h, form, reps, C.byref(ms)))
        return ms.value

    def run(self, design, form=FORM_PLAIN, rrr=3, threads=8):
        """Pattern pass plus 'rrr' rip-up-and-reroute passes on a built design."""
        o = RunOpts(form, rrr, threads)
        st = RunStats()
        _check(lib().igr_run(self.h, design.h, C.byref(o), C.byref(st)))
        self.num_nets = st.subnets
        it = st.iterations
        return {"upload_s": st.upload_s, "pattern_s": st.pattern_s, "pattern_schedule_s": st.pattern_schedule_s,
                "total_s": st.total_s, "pattern_batches": st.pattern_batches, "iterations": it,
                "subnets": st.subnets, "ripped": list(st.ripped[:it]), "rrr_batches": list(st.rrr_batches[:it]),
                "rrr_dag_nodes": list(st.rrr_dag_nodes[:it]), "rrr_view_s": list(st.rrr_view_s[:it]),
                "rrr_dag_build_s": list(st.rrr_dag_build_s[:it]), "rrr_route_s": list(st.rrr_route_s[:it]),
                "rrr_schedule_s": list(st.rrr_schedule_s[:it]), "rrr_prep_s": list(st.rrr_prep_s[:it]),
                "rrr_prep_parts": [list(st.rrr_prep_parts[k]) for k in range(it)],
                "rrr_dag_rest_s": list(st.rrr_dag_rest_s[:it])}

    @property
    def stamp(self):
        return lib().igr_stamp(self.h)


def _scheck(rc):
    if rc != 0:
        raise RuntimeError("igr server: %s (status %d)" % (lib().igr_server_error().decode(), rc))


class Server:
    """Serving loop over a routed Router (include/igr.h, igr_server_*).

    While open, the worker thread owns the router's context: call nothing else
    on the router until close()."""

    def __init__(self, router, design, threads=8, max_tick_nets=0, scratch_bytes=0, hold_ms=0, overtake=False):
        self.router = router
        self.design = design
        o = ServeOpts(threads, max_tick_nets, scratch_bytes, hold_ms, 1 if overtake else 0)
        h = C.c_void_p()
        _scheck(lib().igr_server_open(router.h, design.h, C.byref(o), C.byref(h)))
        self.h = h

    def submit(self, nets):
        a = np.ascontiguousarray(nets, dtype=np.int32)
        rid = C.c_int64()
        _scheck(lib().igr_server_submit(self.h, a.ctypes.data, a.size, C.byref(rid)))
        return rid.value

    def wait(self, rid):
        r = RequestResult()
        _scheck(lib().igr_server_wait(self.h, rid, C.byref(r)))
        return {f: getattr(r, f) for f, _ in RequestResult._fields_}

    def stats(self):
        st = ServerStats()
        _scheck(lib().igr_server_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _scheck(lib().igr_server_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
