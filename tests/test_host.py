"""Host-side logic (no GPU): file formats, the synthetic generator, Steiner
trees and subnets, pattern-routing inputs, detour DAGs, the C ABI's exported
symbols and the command line's argument handling."""
import filecmp
import hashlib
import itertools
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import orc
from instantgr_amd.host import Design

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _md5(path):
    return hashlib.md5(open(path, "rb").read()).hexdigest()


# ---------------------------------------------------------------------------
# .cap / .net formats and the generator
# ---------------------------------------------------------------------------
def test_cap_net_roundtrip(tmp_path):
    d = Design.custom(layers=6, xsize=33, ysize=27, nets=150, pin_cap=14, seed=21)
    c1, n1, c2, n2 = (str(tmp_path / f) for f in ("a.cap", "a.net", "b.cap", "b.net"))
    d.save(c1, n1)
    d2 = Design.load(c1, n1)
    d2.save(c2, n2)
    assert filecmp.cmp(c1, c2, shallow=False) and filecmp.cmp(n1, n2, shallow=False)
    d.build(2)
    d2.build(2)
    for k in ("node_off", "node", "parent", "pt", "hs", "vs", "order", "hpwl", "pins", "ap"):
        assert np.array_equal(d.array(k), d2.array(k)), k


def test_load_rejects_malformed(tmp_path):
    bad = tmp_path / "bad.cap"
    bad.write_text("3 10\n")
    net = tmp_path / "x.net"
    net.write_text("")
    with pytest.raises(RuntimeError):
        Design.load(str(bad), str(net))
    with pytest.raises(RuntimeError):
        Design.load(str(tmp_path / "missing.cap"), str(net))


def test_generator_deterministic(tmp_path):
    paths = []
    for k in range(2):
        c, n = str(tmp_path / f"{k}.cap"), str(tmp_path / f"{k}.net")
        Design.custom(layers=5, xsize=40, ysize=30, nets=400, seed=5).save(c, n)
        paths.append((c, n))
    assert _md5(paths[0][0]) == _md5(paths[1][0]) and _md5(paths[0][1]) == _md5(paths[1][1])
    c, n = str(tmp_path / "o.cap"), str(tmp_path / "o.net")
    Design.custom(layers=5, xsize=40, ysize=30, nets=400, seed=6).save(c, n)
    assert _md5(c) != _md5(paths[0][0])


def test_generator_distributions():
    """Configuration 0: 10 metal layers, 200x200, 20k nets; pins per net
    2 + geometric(0.35) (mean 3.857), capacities uniform{2..6} on the routing
    layers with about 5% blocked, metal1 without capacity."""
    d = Design.preset(0).build(8)
    i = d.info
    assert (i.metal_layers, i.X, i.Y, i.nets) == (10, 200, 200, 20000)
    pins = np.diff(d.array("net_pin_off"))
    assert abs(pins.mean() - (2 + 0.65 / 0.35)) < 0.07
    assert pins.min() >= 2 and pins.max() <= 30
    cap, w, xlen, ylen = d.grid_arrays()
    cap = cap.reshape(i.L, i.X, i.Y)
    vals = set(np.unique(cap).tolist())
    assert vals <= {0.0, 2.0, 3.0, 4.0, 5.0, 6.0}
    zero = (cap == 0).mean()
    assert 0.02 < zero < 0.08
    assert np.all(w == 1.0) and np.all(xlen == 1) and np.all(ylen == 1)


# ---------------------------------------------------------------------------
# Steiner trees, subnets, pattern inputs
# ---------------------------------------------------------------------------
def _subnet_trees(d):
    i = d.info
    XY = i.X * i.Y
    noff, node = d.array("tree_node_off"), d.array("tree_node")
    aoff, adj = d.array("tree_adj_off"), d.array("tree_adj")
    poff, pins = d.array("sub_pin_off"), d.array("sub_pins")
    for s in range(len(noff) - 1):
        nodes = node[noff[s]:noff[s + 1]]
        xy = [(v % XY // i.Y, v % i.Y) for v in nodes]
        edges = set()
        for k in range(len(nodes)):
            g = noff[s] + k
            for j in adj[aoff[g]:aoff[g + 1]]:
                edges.add((min(k, int(j)), max(k, int(j))))
        p = pins[poff[s]:poff[s + 1]]
        pxy = sorted({(v % XY // i.Y, v % i.Y) for v in p})
        yield xy, edges, pxy


def _mst_len(pts):
    if len(pts) <= 1:
        return 0
    pts = list(pts)
    inf = 1 << 60
    best = [inf] * len(pts)
    used = [False] * len(pts)
    best[0] = 0
    tot = 0
    for _ in range(len(pts)):
        k = min((b, j) for j, b in enumerate(best) if not used[j])[1]
        used[k] = True
        tot += best[k]
        for j in range(len(pts)):
            if not used[j]:
                best[j] = min(best[j], abs(pts[j][0] - pts[k][0]) + abs(pts[j][1] - pts[k][1]))
    return tot


def _rsmt_brute(pts):
    """Exact rectilinear Steiner minimal tree length: MST over the pins plus
    the best subset of at most n-2 Hanan-grid points."""
    xs = sorted({p[0] for p in pts})
    ys = sorted({p[1] for p in pts})
    hanan = [(x, y) for x in xs for y in ys if (x, y) not in set(pts)]
    best = _mst_len(pts)
    for k in range(1, len(pts) - 1):
        for extra in itertools.combinations(hanan, k):
            best = min(best, _mst_len(list(pts) + list(extra)))
    return best


def test_steiner_trees_connect_and_bounds(small_design):
    d = small_design
    checked = {2: 0, 3: 0, 4: 0, 5: 0}
    for xy, edges, pxy in _subnet_trees(d):
        n = len(xy)
        # a tree: n-1 edges, connected, covering every pin position
        assert len(edges) == max(0, n - 1)
        seen, stack = {0}, [0]
        nb = {k: [] for k in range(n)}
        for a, b in edges:
            nb[a].append(b)
            nb[b].append(a)
        while stack:
            v = stack.pop()
            for u in nb[v]:
                if u not in seen:
                    seen.add(u)
                    stack.append(u)
        assert len(seen) == n
        assert set(pxy) <= set(xy)
        length = sum(abs(xy[a][0] - xy[b][0]) + abs(xy[a][1] - xy[b][1]) for a, b in edges)
        hp = (max(p[0] for p in pxy) - min(p[0] for p in pxy)) + (max(p[1] for p in pxy) - min(p[1] for p in pxy))
        assert hp <= length <= _mst_len(pxy)
        k = len(pxy)
        if k in (2, 3):
            assert length == hp  # closed form: RSMT of 2 or 3 pins is the half perimeter
            checked[k] += 1
        elif k == 4 and checked[k] < 200:
            assert length == _rsmt_brute(pxy)  # exact for four pins, as FLUTE's length is
            checked[k] += 1
        elif k == 5 and checked[k] < 60:
            assert length == _rsmt_brute(pxy)  # exact up to nine pins, as FLUTE's length is
            checked[k] += 1
    assert checked[2] > 0 and checked[3] > 0 and checked[4] > 0 and checked[5] > 0


def test_subnet_split_covers_pins():
    d = Design.custom(layers=6, xsize=60, ysize=60, nets=40, pin_cap=40, geom_p=0.05, seed=9).build(2)
    i = d.info
    XY = i.X * i.Y
    orig = d.array("orig")
    poff, pins = d.array("sub_pin_off"), d.array("sub_pins")
    soff, sub = d.array("net_sub_off"), d.array("net_sub")
    npo, npins = d.array("pin_off"), d.array("pins")
    split = 0
    for net in range(len(soff) - 1):
        subs = sub[soff[net]:soff[net + 1]]
        assert np.all(orig[subs] == net)
        covered = set()
        for s in subs:
            ps = pins[poff[s]:poff[s + 1]]
            assert len(ps) >= 1
            covered |= {int(v) % XY for v in ps}
        want = {int(v) % XY for v in npins[npo[net]:npo[net + 1]]}
        assert want <= covered
        # only nets above MAX_PIN_SIZE = 12 pins are split
        # (src/database_cuda.hpp:2111-2120)
        if npo[net + 1] - npo[net] <= 12:
            assert len(subs) == 1
        split += len(subs) > 1
    assert split > 0


def test_pattern_inputs_order(small_design):
    d = small_design
    noff, par = d.array("node_off"), d.array("parent")
    for s in range(len(noff) - 1):
        p = par[noff[s]:noff[s + 1]]
        assert p[0] == -1 and np.all(p[1:] >= 0) and np.all(p[1:] < np.arange(1, len(p)))
    order, hp = d.array("order"), d.array("hpwl")
    h = hp[order]
    assert np.all(np.diff(h) <= 0)
    ties = np.nonzero(np.diff(h) == 0)[0]
    assert np.all(order[ties] < order[ties + 1])  # stable
    assert sorted(order.tolist()) == list(range(len(hp)))


# ---------------------------------------------------------------------------
# Detour DAGs (src/database_cuda.hpp:826-889, :1429-2086)
# ---------------------------------------------------------------------------
def test_dag_builder_invariants(small_design):
    d = small_design
    i = d.info
    XY = i.X * i.Y
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    xs, ys = o.congestion_view()
    nets = np.arange(d.info.subnets, dtype=np.int32)
    for mode in (0, 1):
        g = d.build_dags(nets, xs, ys, mode, 4)
        noff, node = g["dag_node_off"], g["dag_node"]
        poff, par, slot = g["dag_par_off"], g["dag_par"], g["dag_slot"]
        cn, depth = g["dag_child_num"], g["dag_depth"]
        ptoff, pt = g["dag_pt_off"], g["dag_pt"]
        toff, tnode = d.array("node_off"), d.array("node")
        for k, net in enumerate(nets):
            b, e = noff[k], noff[k + 1]
            assert e > b
            for j in range(b, e):
                ps = par[poff[j]:poff[j + 1]]
                assert (j == b) == (len(ps) == 0)
                assert np.all(ps < j - b) and np.all(ps >= 0)
                assert cn[j] <= 10
                for q in range(poff[j], poff[j + 1]):
                    assert 0 <= slot[q] < cn[b + par[q]]
                    assert depth[j] > depth[b + par[q]]
            # the tree's pins survive with their layers; every DAG position is
            # a scheduler point of the net
            dn = set(node[b:e].tolist())
            for v in tnode[toff[net]:toff[net + 1]]:
                if v // XY < i.L:
                    assert int(v) in dn
            assert {int(v) % XY for v in node[b:e]} <= set(pt[ptoff[k]:ptoff[k + 1]].tolist())


def test_dag_build_in_two_steps_equals_one(small_design):
    """The points build and the structure build after it give the arrays of
    one full build, in every mode; a structure build over a different net
    count than the points before it is refused."""
    from instantgr_amd.host import DAG_ALL, DAG_POINTS, DAG_STRUCTURE, DAG_POINT_KEYS, DAG_STRUCTURE_KEYS
    d = small_design
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    xs, ys = o.congestion_view()
    hp = d.array("hpwl")
    nets = np.arange(d.info.subnets, dtype=np.int32)
    nets = nets[np.argsort(-hp[nets], kind="stable")]
    for mode in (0, 1, 2):
        whole = d.build_dags(nets, xs, ys, mode, 4, DAG_ALL)
        pts = d.build_dags(nets, xs, ys, mode, 3, DAG_POINTS)
        assert set(pts) == set(DAG_POINT_KEYS)
        rest = d.build_dags(nets, xs, ys, mode, 2, DAG_STRUCTURE)
        assert set(rest) == set(DAG_STRUCTURE_KEYS)
        for k in DAG_POINT_KEYS:
            assert np.array_equal(pts[k], whole[k]), (mode, k)
        for k in DAG_STRUCTURE_KEYS:
            assert np.array_equal(rest[k], whole[k]), (mode, k)
        # the structure build leaves the points as they are
        assert np.array_equal(d.array("dag_pt"), whole["dag_pt"])
    with pytest.raises(RuntimeError):
        d.build_dags(nets[:-1], xs, ys, 0, 2, DAG_STRUCTURE)
    with pytest.raises(RuntimeError):
        d.build_dags(nets, xs, ys, 0, 2, 7)


# ---------------------------------------------------------------------------
# C ABI and command line
# ---------------------------------------------------------------------------
def _declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    body = txt.split('extern "C" {', 1)[1] if 'extern "C" {' in txt else txt
    body = body.split("#ifdef __cplusplus\n}")[0]
    names = re.findall(r"\b([a-z][a-z0-9_]*)\s*\(", body)
    return {n for n in names if n.startswith(("igr_", "igrh_")) and not n.endswith("_t")}


def test_abi_exports_every_declared_symbol():
    lib_dir = os.path.join(ROOT, "instantgr_amd", "lib")
    for header, so in (("igr.h", "libigr.so"), ("igr_host.h", "libigr_host.so")):
        names = _declared(os.path.join(ROOT, "include", header))
        assert len(names) > 5, header
        out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(lib_dir, so)], capture_output=True,
                             text=True, check=True).stdout
        exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
        missing = sorted(n for n in names if n not in exported)
        assert not missing, (header, missing)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_abi_headers_are_plain_c():
    # The C boundary: both headers compile as C99 with no warnings, so a cgo,
    # JNI or ctypes binding can include them unchanged.
    for header in ("igr.h", "igr_host.h"):
        p = subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-std=c99", "-pedantic", "-x", "c",
                            os.path.join(ROOT, "include", header)], capture_output=True, text=True)
        assert p.returncode == 0, (header, p.stderr)


def test_cli_argument_errors(tmp_path):
    exe = os.path.join(ROOT, "instantgr_amd", "lib", "igr_route")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, "-cap", "a", "-bogus", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = subprocess.run([exe, "-cap", str(tmp_path / "none.cap"), "-net", str(tmp_path / "none.net"), "-out",
                        str(tmp_path / "o.out")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stderr.startswith("input:")
