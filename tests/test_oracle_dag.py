"""Pins of the oracle's DAG rip-up-and-reroute routing (no GPU).

The DAG DP (src/Lshape_route_detour.hpp:123-195, :281-406, :520-694) is
checked where the mathematics fixes its answer:
  * a DAG that is a tree with straight edges too short for a layer change
    reduces to the L-shape DP's model, itself pinned by brute force, so the
    route it produces costs the L-shape optimum (to the FP32 rounding of the
    packed keys);
  * a two-node DAG with a long straight edge is brute-forced over the end
    layers, the switch layers and the sampled switch points.
  * a multi-parent DAG whose alternatives share a slot of their parent (both
    L corners of every diagonal tree edge, the child a child of each corner)
    is routed at the cost of its best selection tree: the minimum over every
    choice of one corner per diagonal edge, each routed as a tree.  In all
    ten cases the choice changes the cost, and in six the best is not the
    first corner.
  * more generally, every detour DAG the builder makes (L corners, Z shapes,
    mirror trunks with duplicated pin nodes) is routed at the cost of its
    best selection tree, one candidate child per slot of each node reached
    from the source, each routed as a tree: checked on 25 nets per design
    and builder mode, with at least 5 carrying mirror trunks."""
import math

import numpy as np
import pytest

import orc

from test_oracle import _grid


def _costs(L, X, Y, DIR, seed, dem):
    o, cap, w, xlen, ylen = _grid(L, X, Y, DIR, seed=seed)
    o.set_demand(dem)
    o.update_cost()
    o.compute_presum()
    _, vc, ps = (a.reshape(L, X, Y) for a in o.costs())
    return o, vc, ps


def _route_cost(route, L, X, Y, ps, vc):
    """Wires by prefix differences, vias by the full cut sum."""
    XY = X * Y
    dec = lambda v: (v // XY, v // Y % X, v % Y)
    cost = 0.0
    for k in range(1, route[0], 2):
        a, b = dec(route[k]), dec(route[k + 1])
        if a[0] == b[0]:
            cost += ps[a[0], max(a[1], b[1]), max(a[2], b[2])] - ps[a[0], min(a[1], b[1]), min(a[2], b[2])]
        else:
            assert a[1:] == b[1:]
            cost += sum(float(vc[z, a[1], a[2]]) for z in range(min(a[0], b[0]), max(a[0], b[0])))
    return cost


def _tree_dag(nodes, parents, X, Y):
    n = len(nodes)
    children = [[] for _ in range(n)]
    for i in range(1, n):
        children[parents[i]].append(i)
    depth = [0] * n
    for i in range(1, n):
        depth[i] = depth[parents[i]] + 1
    par_off = [0]
    par, slot = [], []
    for i in range(n):
        if i > 0:
            par.append(parents[i])
            slot.append(children[parents[i]].index(i))
        par_off.append(len(par))
    XY = X * Y
    pts = sorted({v % XY for v in nodes})
    return {"dag_node_off": np.array([0, n], np.int64), "dag_node": np.array(nodes, np.int32),
            "dag_par_off": np.array(par_off, np.int64), "dag_par": np.array(par, np.int32),
            "dag_slot": np.array(slot, np.int32), "dag_child_num": np.array([len(c) for c in children], np.int32),
            "dag_depth": np.array(depth, np.int32), "dag_pt_off": np.array([0, len(pts)], np.int64),
            "dag_pt": np.array(pts, np.int32)}


def _straight_tree(rng, L, X, Y, n, maxlen):
    """Random tree whose edges are axis-aligned and at most maxlen long."""
    XY = X * Y
    while True:
        pos = [(int(rng.integers(0, X)), int(rng.integers(0, Y)))]
        parents = [-1]
        ok = True
        while len(pos) < n and ok:
            p = int(rng.integers(0, len(pos)))
            px, py = pos[p]
            for _ in range(50):
                step = int(rng.integers(1, maxlen + 1)) * (1 if rng.random() < 0.5 else -1)
                x, y = (px + step, py) if rng.random() < 0.5 else (px, py + step)
                if 0 <= x < X and 0 <= y < Y and (x, y) not in pos:
                    pos.append((x, y))
                    parents.append(p)
                    break
            else:
                ok = False
        if ok:
            break
    layers = [int(rng.choice([0, 1, L])) for _ in range(n)]
    layers[0] = int(rng.choice([0, 1]))
    # DFS pre-order
    order, stack = [], [0]
    kids = [[] for _ in range(n)]
    for i in range(1, n):
        kids[parents[i]].append(i)
    while stack:
        v = stack.pop()
        order.append(v)
        stack.extend(reversed(kids[v]))
    ren = {v: i for i, v in enumerate(order)}
    nodes = [layers[v] * XY + pos[v][0] * Y + pos[v][1] for v in order]
    par = [-1] + [ren[parents[v]] for v in order[1:]]
    return nodes, par


@pytest.mark.parametrize("seed", range(10))
def test_dag_tree_route_costs_lshape_optimum(seed):
    rng = np.random.default_rng(500 + seed)
    L, X, Y, DIR = 4 + seed % 2, 9, 8, seed % 2
    dem = (rng.integers(0, 10, L * X * Y) * 0.5).astype(np.float32)
    ref, vc, ps = _costs(L, X, Y, DIR, seed, dem)
    nodes, parents = _straight_tree(rng, L, X, Y, 2 + seed % 4, 4)
    ref.set_nets(np.array([0, len(nodes)], np.int64), np.array(nodes, np.int32), np.array(parents, np.int32))
    best, lroute = ref.lshape_probe(0)
    o, _, _ = _costs(L, X, Y, DIR, seed, dem)
    o.set_nets(np.array([0, len(nodes)], np.int64), np.array(nodes, np.int32), np.array(parents, np.int32))
    nb = orc.dag_pass(o, np.array([0], np.int32), _tree_dag(nodes, parents, X, Y))
    assert nb == 1
    route = o.route(0)
    got = _route_cost(route, L, X, Y, ps, vc)
    assert math.isclose(got, best, rel_tol=1e-6), (got, best)
    assert math.isclose(_route_cost(lroute, L, X, Y, ps, vc), best, rel_tol=1e-12)


def _two_node_brute(L, X, Y, DIR, ps, vc, a, b):
    """Minimum cost of connecting node b (child) to node a (root) along one
    straight edge: same-layer wire, or (for an edge longer than 4) a layer
    change between layers of the same parity at a sampled switch point,
    plus the via stacks at both ends over the layers meeting there."""
    XY = X * Y
    (la, ax, ay), (lb, bx, by) = [(v // XY, v // Y % X, v % Y) for v in (a, b)]
    dirl = lambda l: (l & 1) ^ DIR
    vert = ax == bx
    dist = abs(ax - bx) + abs(ay - by)
    wseg = lambda l, p, q: ps[l, max(p[0], q[0]), max(p[1], q[1])] - ps[l, min(p[0], q[0]), min(p[1], q[1])]
    stack = lambda ls, x, y: sum(float(vc[z, x, y]) for z in range(min(ls), max(ls)))
    best = math.inf
    for fl in range(L):           # layer at the child end
        if dirl(fl) != (1 if vert else 0):
            continue
        for tl in range(L):       # layer at the parent end
            if (tl & 1) != (fl & 1):
                continue
            if fl == tl:
                edge = wseg(fl, (bx, by), (ax, ay))
            elif dist > 4:
                dd = dist // min(dist, 20)
                s0, s1 = (min(ay, by), max(ay, by)) if vert else (min(ax, bx), max(ax, bx))
                edge = math.inf
                for sp in range(s0 + dd, s1, dd):
                    sw = (bx, sp) if vert else (sp, by)
                    c = stack([fl, tl], *sw) + wseg(fl, (bx, by), sw) + wseg(tl, sw, (ax, ay))
                    edge = min(edge, c)
            else:
                continue
            lsb = [fl] + ([lb] if lb < L else [])
            lsa = [tl] + ([la] if la < L else [])
            best = min(best, edge + stack(lsb, bx, by) + stack(lsa, ax, ay))
    return best


@pytest.mark.parametrize("seed", range(12))
def test_dag_two_node_layer_change_brute_force(seed):
    rng = np.random.default_rng(700 + seed)
    L, X, Y, DIR = 5, 30, 26, seed % 2
    # congest the lower layers so that switching layers pays off
    dem = (rng.integers(0, 6, (L, X, Y)) * 0.5).astype(np.float32)
    dem[:2] += rng.integers(3, 9, (2, X, Y)).astype(np.float32)
    ref, vc, ps = _costs(L, X, Y, DIR, seed, dem.ravel())
    XY = X * Y
    ax, ay = int(rng.integers(0, X)), int(rng.integers(0, Y))
    if seed % 4 < 2:
        bx, by = ax, int((ay + rng.integers(5, Y - 1)) % Y)
    else:
        bx, by = int((ax + rng.integers(5, X - 1)) % X), ay
    la, lb = int(rng.choice([0, 1, L])), int(rng.choice([0, 1, L]))
    nodes, parents = [la * XY + ax * Y + ay, lb * XY + bx * Y + by], [-1, 0]
    brute = _two_node_brute(L, X, Y, DIR, ps, vc, nodes[0], nodes[1])
    o, _, _ = _costs(L, X, Y, DIR, seed, dem.ravel())
    o.set_nets(np.array([0, 2], np.int64), np.array(nodes, np.int32), np.array(parents, np.int32))
    orc.dag_pass(o, np.array([0], np.int32), _tree_dag(nodes, parents, X, Y))
    got = _route_cost(o.route(0), L, X, Y, ps, vc)
    assert math.isclose(got, brute, rel_tol=1e-6), (got, brute)


def _diag_tree(rng, L, X, Y, n):
    """Random tree of n nodes at distinct gcells, parents before children,
    with at least one diagonal edge (both coordinates differ)."""
    XY = X * Y
    while True:
        pos, parents = [], []
        while len(pos) < n:
            p = (int(rng.integers(0, X)), int(rng.integers(0, Y)))
            if p not in pos:
                parents.append(-1 if not pos else int(rng.integers(0, len(pos))))
                pos.append(p)
        if any(pos[i][0] != pos[parents[i]][0] and pos[i][1] != pos[parents[i]][1] for i in range(1, n)):
            break
    layers = [int(rng.choice([0, 1, L])) for _ in range(n)]
    layers[0] = int(rng.choice([0, 1]))
    return [layers[i] * XY + pos[i][0] * Y + pos[i][1] for i in range(n)], [-1] + parents[1:], pos


def _lcorner_dag(nodes, parents, pos, X, Y, L, pick=None):
    """DAG of a tree whose diagonal edges get both L corners as alternatives,
    as the detour builder makes them: Steiner nodes at (parent x, child y) and
    (child x, parent y), each in the parent's slot of that child and each a
    parent of the child through slot 0, so the child has two parents.  With
    pick, diagonal edge e keeps only corner pick[e] (a selection tree)."""
    XY, n = X * Y, len(nodes)
    kids = [[] for _ in range(n)]
    for i in range(1, n):
        kids[parents[i]].append(i)
    out_nodes, out_par, child_num = [], [], []     # out_par[i]: list of (parent, slot)
    diag = 0

    def emit(v, par_list):
        out_nodes.append(v)
        out_par.append(par_list)
        child_num.append(0)
        return len(out_nodes) - 1

    def visit(v, par_list):
        nonlocal diag
        me = emit(nodes[v], par_list)
        child_num[me] = len(kids[v])
        for s, c in enumerate(kids[v]):
            (px, py), (cx, cy) = pos[v], pos[c]
            if px != cx and py != cy:
                corners = [(px, cy), (cx, py)]
                keep = [0, 1] if pick is None else [pick[diag]]
                diag += 1
                tl = []
                for k in keep:
                    t = emit(L * XY + corners[k][0] * Y + corners[k][1], [(me, s)])
                    child_num[t] = 1
                    tl.append(t)
                visit(c, [(t, 0) for t in tl])
            else:
                visit(c, [(me, s)])
    visit(0, [])
    depth = [0] * len(out_nodes)
    for i in range(1, len(out_nodes)):
        depth[i] = max(depth[p] + 1 for p, _ in out_par[i])
    par_off, par, slot = [0], [], []
    for pl in out_par:
        for p, s in pl:
            par.append(p)
            slot.append(s)
        par_off.append(len(par))
    pts = sorted({v % XY for v in out_nodes})
    return {"dag_node_off": np.array([0, len(out_nodes)], np.int64), "dag_node": np.array(out_nodes, np.int32),
            "dag_par_off": np.array(par_off, np.int64), "dag_par": np.array(par, np.int32),
            "dag_slot": np.array(slot, np.int32), "dag_child_num": np.array(child_num, np.int32),
            "dag_depth": np.array(depth, np.int32), "dag_pt_off": np.array([0, len(pts)], np.int64),
            "dag_pt": np.array(pts, np.int32)}, diag


@pytest.mark.parametrize("seed", range(10))
def test_multi_parent_dag_costs_the_best_selection_tree(seed):
    # A node with several parents (one per alternative in the same slot of
    # its grandparent) is routed through the best alternative: the DAG's
    # route costs the minimum over every choice of one corner per diagonal
    # edge, each of those trees routed by the tree DP (pinned above).
    import itertools
    rng = np.random.default_rng(900 + seed)
    L, X, Y, DIR = 4 + seed % 2, 12, 11, seed % 2
    dem = (rng.integers(0, 12, L * X * Y) * 0.5).astype(np.float32)
    _, vc, ps = _costs(L, X, Y, DIR, seed, dem)
    nodes, parents, pos = _diag_tree(rng, L, X, Y, 2 + seed % 3)
    tree = (np.array([0, len(nodes)], np.int64), np.array(nodes, np.int32), np.array(parents, np.int32))

    def routed_cost(pick):
        o, _, _ = _costs(L, X, Y, DIR, seed, dem)
        o.set_nets(*tree)
        dag, _ = _lcorner_dag(nodes, parents, pos, X, Y, L, pick)
        assert orc.dag_pass(o, np.array([0], np.int32), dag) == 1
        return _route_cost(o.route(0), L, X, Y, ps, vc)
    _, k = _lcorner_dag(nodes, parents, pos, X, Y, L)
    assert dag_has_two_parents(_lcorner_dag(nodes, parents, pos, X, Y, L)[0])
    sel = [routed_cost(list(p)) for p in itertools.product([0, 1], repeat=k)]
    got = routed_cost(None)
    assert math.isclose(got, min(sel), rel_tol=1e-6), (got, sel)


def dag_has_two_parents(dag):
    return bool(np.any(np.diff(dag["dag_par_off"]) >= 2))


def _net_dag(a, i):
    no, po = a("dag_node_off"), a("dag_par_off")
    b, e = int(no[i]), int(no[i + 1])
    nodes = [int(v) for v in a("dag_node")[b:e]]
    pars = [[(int(a("dag_par")[k]), int(a("dag_slot")[k])) for k in range(po[b + v], po[b + v + 1])]
            for v in range(e - b)]
    return nodes, pars, [int(c) for c in a("dag_child_num")[b:e]]


def _dag_dict(nodes, pars, cn, XY):
    depth = [0] * len(nodes)
    for v in range(1, len(nodes)):
        depth[v] = max(depth[p] + 1 for p, _ in pars[v])
    par_off, par, slot = [0], [], []
    for pl in pars:
        par += [p for p, _ in pl]
        slot += [s for _, s in pl]
        par_off.append(len(par))
    pts = sorted({v % XY for v in nodes})
    return {"dag_node_off": np.array([0, len(nodes)], np.int64), "dag_node": np.array(nodes, np.int32),
            "dag_par_off": np.array(par_off, np.int64), "dag_par": np.array(par, np.int32),
            "dag_slot": np.array(slot, np.int32), "dag_child_num": np.array(cn, np.int32),
            "dag_depth": np.array(depth, np.int32), "dag_pt_off": np.array([0, len(pts)], np.int64),
            "dag_pt": np.array(pts, np.int32)}


def _selection_trees(nodes, pars, cn, cap):
    """Every tree that keeps one candidate child per slot of each node it
    reaches from the source (edges (child, parent, slot)), or None when
    there are fewer than 2 or more than cap of them."""
    import itertools
    kids = [[[] for _ in range(cn[v])] for v in range(len(nodes))]
    for c in range(1, len(nodes)):
        for p, s in pars[c]:
            kids[p][s].append(c)
    memo = {}

    def count(v):
        if v not in memo:
            r = 1
            for s in range(cn[v]):
                r *= sum(count(c) for c in kids[v][s])
            memo[v] = r
        return memo[v]
    if not 2 <= count(0) <= cap:
        return None

    def sel(v):
        per_slot = [[[(c, v, s)] + sub for c in kids[v][s] for sub in sel(c)] for s in range(cn[v])]
        return [sum(combo, []) for combo in itertools.product(*per_slot)]
    return sel(0)


def _tree_of(nodes, cn, edges):
    ch = {}
    for c, p, s in edges:
        ch.setdefault(p, []).append((s, c))
    order, stack = [], [0]
    while stack:
        v = stack.pop()
        order.append(v)
        stack += [c for _, c in sorted(ch.get(v, []), reverse=True)]
    ren = {v: i for i, v in enumerate(order)}
    parent = {c: (p, s) for c, p, s in edges}
    return ([nodes[v] for v in order], [[]] + [[(ren[parent[v][0]], parent[v][1])] for v in order[1:]],
            [cn[v] for v in order])


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("seed", [5, 6])
def test_builder_dags_cost_the_best_selection_tree(mode, seed):
    # Detour DAGs as the builder makes them (L corners, Z shapes, mirror
    # trunks with duplicated pin nodes) from a congested pattern-routed
    # design: each net's DAG route costs the minimum over its selection trees
    # (one candidate child per slot), each routed as a tree.
    from instantgr_amd.host import Design
    d = Design.custom(layers=6, xsize=30, ysize=28, nets=220, pin_cap=8, sigma=3.0, seed=seed).build(2)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    xs, ys = o.congestion_view()
    nsub = len(d.array("hpwl"))
    d.build_dags(np.arange(nsub, dtype=np.int32), xs, ys, mode, 2)
    a, info = d.array, d.info
    L, X, Y = info.L, info.X, info.Y
    dem = (np.random.default_rng(seed).integers(0, 12, L * X * Y) * 0.5).astype(np.float32)

    def fresh():
        q = orc.Oracle.from_design(d)
        q.set_nets_from_design(d)
        q.set_demand(dem)
        q.update_cost()
        q.compute_presum()
        return q
    _, vc, ps = (x.reshape(L, X, Y) for x in fresh().costs())

    def routed(i, nodes, pars, cn):
        q = fresh()
        orc.dag_pass(q, np.array([i], np.int32), _dag_dict(nodes, pars, cn, X * Y))
        return _route_cost(q.route(i), L, X, Y, ps, vc)
    spo = a("sub_pin_off")
    tested = dup = not_first = 0
    for i in range(nsub):
        nodes, pars, cn = _net_dag(a, i)
        sels = _selection_trees(nodes, pars, cn, 256)
        if sels is None:
            continue
        costs = [routed(i, *_tree_of(nodes, cn, e)) for e in sels]
        got = routed(i, nodes, pars, cn)
        assert math.isclose(got, min(costs), rel_tol=1e-6), (i, got, min(costs))
        tested += 1
        dup += sum(v // (X * Y) < L for v in nodes) > spo[i + 1] - spo[i]
        not_first += costs[0] > min(costs) * (1 + 1e-9)
        if tested == 25:
            break
    assert tested == 25 and dup >= 5 and not_first >= 10, (tested, dup, not_first)
