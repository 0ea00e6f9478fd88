"""ctypes binding of libigr_host.so: synthetic designs, .cap/.net I/O, subnets,
trees, routing inputs, the route writer and the evaluator."""
import ctypes as C
import os

import numpy as np

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_lib = None

# igrh_build_dags_part parts (include/igr_host.h) and the arrays each writes
DAG_ALL, DAG_POINTS, DAG_STRUCTURE = 0, 1, 2
DAG_POINT_KEYS = ("dag_pt_off", "dag_pt")
DAG_STRUCTURE_KEYS = ("dag_node_off", "dag_node", "dag_par_off", "dag_par", "dag_slot", "dag_child_num", "dag_depth")


class GenParams(C.Structure):
    _fields_ = [("layers", C.c_int), ("xsize", C.c_int), ("ysize", C.c_int), ("nets", C.c_int64),
                ("pin_cap", C.c_int), ("geom_p", C.c_double), ("sigma", C.c_double),
                ("blockage_frac", C.c_double), ("heavy_frac", C.c_double), ("heavy_min", C.c_int),
                ("heavy_max", C.c_int), ("heavy_window", C.c_int), ("hotspots", C.c_int),
                ("hotspot_scale", C.c_double), ("seed", C.c_uint64)]


class GridInfo(C.Structure):
    _fields_ = [("L", C.c_int), ("X", C.c_int), ("Y", C.c_int), ("DIR", C.c_int),
                ("unit_wire", C.c_double), ("unit_via", C.c_double),
                ("w", C.POINTER(C.c_double)), ("xlen", C.POINTER(C.c_int)), ("ylen", C.POINTER(C.c_int)),
                ("cap", C.POINTER(C.c_float)), ("metal_layers", C.c_int), ("nets", C.c_int64),
                ("subnets", C.c_int64), ("max_nodes", C.c_int)]


class Score(C.Structure):
    _fields_ = [("wirelength", C.c_double), ("via", C.c_double), ("overflow", C.c_double),
                ("total", C.c_double), ("opens", C.c_int64)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_LIB_DIR, "libigr_host.so")
        if not os.path.exists(path):
            raise RuntimeError("libigr_host.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.igrh_error.restype = C.c_char_p
        L.igrh_generate.restype = C.c_void_p
        L.igrh_generate.argtypes = [C.c_int, C.c_int64, C.c_uint64]
        L.igrh_generate_custom.restype = C.c_void_p
        L.igrh_generate_custom.argtypes = [C.POINTER(GenParams)]
        L.igrh_load.restype = C.c_void_p
        L.igrh_load.argtypes = [C.c_char_p, C.c_char_p]
        L.igrh_save.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        L.igrh_free.argtypes = [C.c_void_p]
        L.igrh_build.argtypes = [C.c_void_p, C.c_int]
        L.igrh_build_seconds.argtypes = [C.c_void_p]
        L.igrh_build_seconds.restype = C.c_double
        L.igrh_grid.argtypes = [C.c_void_p, C.POINTER(GridInfo)]
        L.igrh_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.igrh_array.restype = C.c_int64
        L.igrh_write_routes.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.igrh_build_dags.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.igrh_build_dags_part.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                           C.c_int]
        L.igrh_evaluate.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(Score)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().igrh_error().decode())


_DTYPES = {"node_off": np.int64, "pt_off": np.int64, "hs_off": np.int64, "vs_off": np.int64,
           "sub_pin_off": np.int64, "tree_node_off": np.int64, "tree_adj_off": np.int64, "net_sub_off": np.int64,
           "pin_off": np.int64, "extra_off": np.int64, "ap_off": np.int64, "net_pin_off": np.int64,
           "dag_node_off": np.int64, "dag_par_off": np.int64, "dag_pt_off": np.int64,
           "metal_cap": np.float64, "metal_of_weight": np.float64}


def gen_params(layers=10, xsize=64, ysize=64, nets=500, pin_cap=30, geom_p=0.35, sigma=8.0,
               blockage_frac=0.05, heavy_frac=0.0, heavy_min=100, heavy_max=1000, heavy_window=300,
               hotspots=0, hotspot_scale=0.7, seed=1):
    return GenParams(layers, xsize, ysize, nets, pin_cap, geom_p, sigma, blockage_frac, heavy_frac,
                     heavy_min, heavy_max, heavy_window, hotspots, hotspot_scale, seed)


class Design:
    """A design (metal-layer .cap/.net contents) and, once built, its routing
    grid, subnets, trees and pattern-routing inputs."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError(lib().igrh_error().decode())
        self.h = C.c_void_p(handle)
        self._arrays = {}
        self.info = None

    @classmethod
    def preset(cls, idx, nets=0, seed=0):
        return cls(lib().igrh_generate(idx, nets, seed))

    @classmethod
    def custom(cls, **kw):
        p = gen_params(**kw)
        return cls(lib().igrh_generate_custom(C.byref(p)))

    @classmethod
    def load(cls, cap, net):
        return cls(lib().igrh_load(cap.encode(), net.encode()))

    def save(self, cap, net):
        _check(lib().igrh_save(self.h, cap.encode(), net.encode()))

    def build(self, threads=8):
        _check(lib().igrh_build(self.h, threads))
        info = GridInfo()
        _check(lib().igrh_grid(self.h, C.byref(info)))
        self.info = info
        self._arrays = {}
        return self

    @property
    def build_seconds(self):
        return lib().igrh_build_seconds(self.h)

    def array(self, name):
        if name not in self._arrays:
            p = C.c_void_p()
            n = lib().igrh_array(self.h, name.encode(), C.byref(p))
            if n < 0:
                raise KeyError(name)
            dt = _DTYPES.get(name, np.int32)
            if n == 0:
                self._arrays[name] = np.zeros(0, dtype=dt)
            else:
                ct = {np.int64: C.c_int64, np.float64: C.c_double}.get(dt, C.c_int32)
                self._arrays[name] = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,))
        return self._arrays[name]

    def grid_arrays(self):
        i = self.info
        n = i.L * i.X * i.Y
        cap = np.ctypeslib.as_array(i.cap, shape=(n,))
        w = np.ctypeslib.as_array(i.w, shape=(i.L,))
        xlen = np.ctypeslib.as_array(i.xlen, shape=(i.X - 1,))
        ylen = np.ctypeslib.as_array(i.ylen, shape=(i.Y - 1,))
        return cap, w, xlen, ylen

    def build_dags(self, nets, xsum, ysum, mode=0, threads=8, part=DAG_ALL):
        """igrh_build_dags_part: the arrays the part writes (DAG_POINTS: dag_pt_off
        and dag_pt; DAG_STRUCTURE, after DAG_POINTS on the same nets: the
        others; DAG_ALL: all nine)."""
        nets = np.ascontiguousarray(nets, np.int32)
        xsum = np.ascontiguousarray(xsum, np.float32)
        ysum = np.ascontiguousarray(ysum, np.float32)
        _check(lib().igrh_build_dags_part(self.h, nets.ctypes.data, len(nets), xsum.ctypes.data, ysum.ctypes.data,
                                          mode, threads, part))
        for k in [k for k in self._arrays if k.startswith("dag_")]:
            del self._arrays[k]
        keys = {DAG_ALL: DAG_POINT_KEYS + DAG_STRUCTURE_KEYS, DAG_POINTS: DAG_POINT_KEYS,
                DAG_STRUCTURE: DAG_STRUCTURE_KEYS}[part]
        return {k: self.array(k).copy() for k in keys}

    def write_routes(self, path, off, data):
        off = np.ascontiguousarray(off, dtype=np.int64)
        data = np.ascontiguousarray(data, dtype=np.int32)
        _check(lib().igrh_write_routes(self.h, path.encode(), off.ctypes.data, data.ctypes.data, data.size))

    def evaluate(self, path):
        s = Score()
        _check(lib().igrh_evaluate(self.h, path.encode(), C.byref(s)))
        return {"wirelength": s.wirelength, "via": s.via, "overflow": s.overflow, "total": s.total,
                "opens": s.opens}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.igrh_free(self.h)
            self.h = None
