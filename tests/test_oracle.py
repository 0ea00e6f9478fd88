"""Pins of the CPU oracle (no GPU): each oracle function is checked against
what the mathematics fixes -- closed forms, textbook library routines,
invariants and brute force on tiny inputs -- so that a dropped term, a wrong
sign, index or operand order fails one of these tests.  The reference holds no
tests or golden vectors for this path."""
import itertools
import math

import numpy as np
import pytest

import orc


# ---------------------------------------------------------------------------
# Routing exp (replaces CUDA __expf at src/graph.hpp:24-30)
# ---------------------------------------------------------------------------
def _ulp_err(got, x):
    ref = np.exp(x.astype(np.float64))
    got64 = got.astype(np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got64 - ref) / spacing


def test_expf_ulp_bound_against_libm():
    rng = np.random.default_rng(1)
    x = np.concatenate([np.linspace(-87.0, 88.7, 400001, dtype=np.float32),
                        rng.uniform(-87, 88.7, 200000).astype(np.float32),
                        rng.uniform(-2, 2, 200000).astype(np.float32)])
    err = _ulp_err(orc.expf(x), x)
    assert err.max() <= 2.0, float(err.max())


def test_expf_special_values():
    x = np.array([0.0, -0.0, 1.0, 88.72, 88.73, 100.0, -87.0, -87.01, -1000.0, np.inf, -np.inf, np.nan],
                 np.float32)
    y = orc.expf(x)
    assert y[0] == 1.0 and y[1] == 1.0
    assert abs(float(y[2]) - math.e) <= 2 * float(np.spacing(np.float32(math.e)))
    assert np.isfinite(y[3]) and y[4] == np.inf and y[5] == np.inf
    assert y[6] > 0 and y[7] == 0 and y[8] == 0
    assert y[9] == np.inf and y[10] == 0 and np.isnan(y[11])


def test_expf_monotone():
    x = np.sort(np.random.default_rng(2).uniform(-87, 88.7, 300000).astype(np.float32))
    y = orc.expf(x)
    assert np.all(np.diff(y) >= 0)


# ---------------------------------------------------------------------------
# Cost update (src/graph.hpp:23-71) and prefix sums (:405-430)
# ---------------------------------------------------------------------------
def _grid(L=4, X=7, Y=5, DIR=1, seed=0, unit_wire=1.0, unit_via=4.0):
    rng = np.random.default_rng(seed)
    cap = rng.integers(0, 5, L * X * Y).astype(np.float32)
    w = rng.uniform(0.5, 2.0, L)
    xlen = rng.integers(1, 4, X - 1)
    ylen = rng.integers(1, 4, Y - 1)
    o = orc.Oracle(L, X, Y, DIR, unit_wire, unit_via, w, xlen, ylen, cap)
    return o, cap.reshape(L, X, Y), w, xlen, ylen


def _f(c, d):
    """Logistic overflow cost of one edge in FP64 (src/graph.hpp:23-32)."""
    if c > 0.001:
        return math.exp(0.5 * (d - c))
    if d > 0:
        return math.exp(1.5 * d)
    return 0.0


def test_cost_closed_forms():
    L, X, Y, DIR = 3, 5, 4, 0
    o, cap, w, xlen, ylen = _grid(L, X, Y, DIR, seed=3)
    cap3 = cap.copy()
    # a cell with demand == capacity: increment of one track is (e^0.5 - 1) w
    # plus the unit wire cost times the edge length; a cell without capacity
    # and demand costs e^1.5 w
    dem = np.zeros((L, X, Y), np.float32)
    cap3[1, 2, 1] = 3.0
    dem[1, 2, 1] = 3.0
    cap3[2, 1, 3] = 0.0
    o2 = orc.Oracle(L, X, Y, DIR, 1.0, 4.0, w, xlen, ylen, cap3.ravel())
    o2.set_demand(dem.ravel())
    o2.update_cost()
    wc, vc, _ = (a.reshape(L, X, Y) for a in o2.costs())
    # layer 1 is vertical (dir = (1 & 1) ^ 0): edge length ylen[y]
    expect = (math.exp(0.5) - 1.0) * w[1] + ylen[1]
    assert math.isclose(wc[1, 2, 1], expect, rel_tol=2e-6)
    # layer 2 is horizontal: edge length xlen[x]
    expect = math.exp(1.5) * w[2] + xlen[1]
    assert math.isclose(wc[2, 1, 3], expect, rel_tol=2e-6)
    # the last cell of a track has no length term
    assert math.isclose(wc[1, 2, Y - 1], (_f(cap3[1, 2, Y - 1], 1) - _f(cap3[1, 2, Y - 1], 0)) * w[1], rel_tol=2e-6)


def test_cost_against_definition_random():
    L, X, Y = 4, 7, 5
    for DIR in (0, 1):
        o, cap, w, xlen, ylen = _grid(L, X, Y, DIR, seed=4 + DIR)
        rng = np.random.default_rng(9)
        dem = (rng.integers(0, 16, (L, X, Y)) * 0.5).astype(np.float32)
        o.set_demand(dem.ravel())
        o.update_cost()
        wc, vc, _ = (a.reshape(L, X, Y) for a in o.costs())
        for l in range(L):
            d = (l & 1) ^ DIR
            n = Y if d else X
            for a in range(X if d else Y):
                for p in range(n):
                    x, y = (a, p) if d else (p, a)
                    inc = lambda xx, yy, k: min(1e12, (_f(cap[l, xx, yy], dem[l, xx, yy] + k) -
                                                       _f(cap[l, xx, yy], dem[l, xx, yy])) * w[l])
                    ew = inc(x, y, 1)
                    if p < n - 1:
                        ew += (ylen[p] if d else xlen[p])
                    assert math.isclose(wc[l, x, y], ew, rel_tol=5e-6), (l, x, y)
                    prev = (x, y - 1) if d else (x - 1, y)
                    if p == 0:
                        ev = inc(x, y, 1)
                    elif p == n - 1:
                        ev = inc(*prev, 1)
                    else:
                        ev = inc(x, y, 0.5) + inc(*prev, 0.5)
                    assert math.isclose(vc[l, x, y], ev + 4.0, rel_tol=5e-6), (l, x, y)


def test_cost_invariants_and_clamp():
    L, X, Y = 3, 6, 6
    o, cap, w, xlen, ylen = _grid(L, X, Y, 1, seed=5)
    wcs = []
    for level in (0.0, 2.0, 4.0, 8.0):
        o.set_demand(np.full(L * X * Y, level, np.float32))
        o.update_cost()
        wc, vc, _ = o.costs()
        assert np.all(vc >= 4.0)
        wcs.append(wc.astype(np.float64))
    for a, b in zip(wcs, wcs[1:]):
        assert np.all(b >= a)  # monotone in demand
    # huge demand: the increment is clamped to 1e12 (inf - inf absorbed)
    o.set_demand(np.full(L * X * Y, 1e6, np.float32))
    o.update_cost()
    wc, vc, _ = (a.reshape(L, X, Y) for a in o.costs())
    assert wc[1, 0, 0] == np.float32(1e12 + xlen[0])  # layer 1 horizontal with DIR = 1
    assert vc[1, 0, 0] == np.float32(np.float32(1e12) + 4.0)


def test_presum_closed_form_exact():
    # every wcost equal (clamped): presum of position i is exactly i * W
    L, X, Y = 3, 37, 11
    o, cap, w, xlen, ylen = _grid(L, X, Y, 1, seed=6)
    o = orc.Oracle(L, X, Y, 1, 1.0, 4.0, w, np.ones(X - 1), np.ones(Y - 1), cap.ravel())
    o.set_demand(np.full(L * X * Y, 1e6, np.float32))
    o.update_cost()
    o.compute_presum()
    wc, _, ps = (a.reshape(L, X, Y) for a in o.costs())
    for l in range(L):
        d = (l & 1) ^ 1
        n = Y if d else X
        W = float(wc[l, 0, 0])
        line = ps[l, 0, :] if d else ps[l, :, 0]
        assert np.array_equal(line, np.arange(n, dtype=np.float64) * W)


def _sklansky_numpy(v):
    s = v.astype(np.float64).copy()
    n = len(s)
    d = 0
    while (1 << d) < n:
        idx = np.arange(n)
        sel = (idx >> d) & 1 == 1
        src = ((idx >> d) << d) - 1
        s = np.where(sel, s + s[np.maximum(src, 0)], s)
        d += 1
    return s


@pytest.mark.parametrize("DIR", [0, 1])
def test_presum_matches_sklansky_and_sequential_bound(DIR):
    L, X, Y = 4, 45, 23
    o, cap, w, xlen, ylen = _grid(L, X, Y, DIR, seed=7)
    rng = np.random.default_rng(8)
    o.set_demand((rng.integers(0, 12, L * X * Y) * 0.5).astype(np.float32))
    o.update_cost()
    o.compute_presum()
    wc, _, ps = (a.reshape(L, X, Y) for a in o.costs())
    for l in range(L):
        d = (l & 1) ^ DIR
        for a in range(X if d else Y):
            wline = wc[l, a, :] if d else wc[l, :, a]
            pline = ps[l, a, :] if d else ps[l, :, a]
            v = np.concatenate([[0.0], wline[:-1].astype(np.float64)])
            # level-by-level (vectorised) Sklansky order: bit-exact
            assert np.array_equal(pline, _sklansky_numpy(v))
            # sequential FP64 sum within n * eps * sum
            seq = np.cumsum(v)
            bound = len(v) * np.finfo(np.float64).eps * np.abs(v).sum()
            assert np.all(np.abs(pline - seq) <= bound)


# ---------------------------------------------------------------------------
# L-shape DP (src/Lshape_route.hpp:9-153): brute force on tiny trees
# ---------------------------------------------------------------------------
def _enum_cost(o, L, X, Y, DIR, nodes, parents, ps, vc):
    """Minimum over every L-shape choice and layer assignment of the DP's cost
    model: wires by prefix differences, a via stack over [min, max] of the
    layers meeting at each node (and its pin layer), corner vias over the
    layers strictly between the two wire layers."""
    XY = X * Y
    pos = [(v // XY, v // Y % X, v % Y) for v in nodes]
    dirl = lambda l: (l & 1) ^ DIR

    def wseg(l, x0, x1, y0, y1):
        return ps[l, max(x0, x1), max(y0, y1)] - ps[l, min(x0, x1), min(y0, y1)]

    options = []
    for i in range(1, len(nodes)):
        _, x, y = pos[i]
        _, px, py = pos[parents[i]]
        opts = []
        if x == px or y == py:
            for l in range(L):
                if dirl(l) == 0 and y != py:
                    continue
                if dirl(l) == 1 and x != px:
                    continue
                opts.append((l, l, wseg(l, x, px, y, py)))
        else:
            for cl in range(L):
                for pl in range(L):
                    if (cl & 1) == (pl & 1):
                        continue
                    if dirl(cl) == 1:
                        corner = (x, py)
                        cost = wseg(cl, x, x, y, py) + wseg(pl, x, px, py, py)
                    else:
                        corner = (px, y)
                        cost = wseg(cl, x, px, y, y) + wseg(pl, px, px, y, py)
                    cost += sum(float(vc[k, corner[0], corner[1]]) for k in range(min(cl, pl) + 1, max(cl, pl)))
                    opts.append((cl, pl, cost))
        options.append(opts)
    best = math.inf
    for combo in itertools.product(*options):
        layers = [set() for _ in nodes]
        for i, pl_i in enumerate(nodes):
            if pos[i][0] < L:
                layers[i].add(pos[i][0])
        cost = 0.0
        for k, (cl, pl, c) in enumerate(combo):
            i = k + 1
            layers[i].add(cl)
            layers[parents[i]].add(pl)
            cost += c
        for i in range(len(nodes)):
            if layers[i]:
                lo, hi = min(layers[i]), max(layers[i])
                cost += sum(float(vc[k, pos[i][1], pos[i][2]]) for k in range(lo, hi))
        best = min(best, cost)
    return best


def _route_cost(route, L, X, Y, DIR, nodes, parents, ps, vc):
    """Cost of a route under the same model, parsed in the order the DP emits
    it (per node: its via stack, then the edge to its parent)."""
    XY = X * Y
    dec = lambda v: (v // XY, v // Y % X, v % Y)
    pairs = [(route[k], route[k + 1]) for k in range(1, route[0], 2)]
    pos = [dec(v) for v in nodes]
    cost, k = 0.0, 0
    for i in range(len(nodes)):
        if k < len(pairs):
            a, b = dec(pairs[k][0]), dec(pairs[k][1])
            if a[0] != b[0] and a[1:] == pos[i][1:]:
                cost += sum(float(vc[z, a[1], a[2]]) for z in range(min(a[0], b[0]), max(a[0], b[0])))
                k += 1
        if i == 0:
            continue
        _, x, y = pos[i]
        _, px, py = pos[parents[i]]
        npieces = 1 if (x == px or y == py) else 3
        for j in range(npieces):
            a, b = dec(pairs[k][0]), dec(pairs[k][1])
            k += 1
            if a[0] == b[0]:
                cost += ps[a[0], max(a[1], b[1]), max(a[2], b[2])] - ps[a[0], min(a[1], b[1]), min(a[2], b[2])]
            else:
                assert j == 1
                cost += sum(float(vc[z, a[1], a[2]]) for z in range(min(a[0], b[0]) + 1, max(a[0], b[0])))
    assert k == len(pairs)
    return cost


def _random_tree(rng, L, X, Y, n):
    XY = X * Y
    pts = set()
    nodes, parents = [], []
    while len(nodes) < n:
        x, y = int(rng.integers(0, X)), int(rng.integers(0, Y))
        if (x, y) in pts:
            continue
        pts.add((x, y))
        l = int(rng.choice([0, 1, L]))  # pin on layer 0 or 1, or a Steiner point
        nodes.append(l * XY + x * Y + y)
        parents.append(-1 if not parents else int(rng.integers(0, len(parents))))
    return nodes, parents


@pytest.mark.parametrize("seed", range(12))
def test_lshape_dp_equals_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    L, X, Y, DIR = (3 + seed % 2), 6, 5, seed % 2
    o, cap, w, xlen, ylen = _grid(L, X, Y, DIR, seed=seed)
    o.set_demand((rng.integers(0, 10, L * X * Y) * 0.5).astype(np.float32))
    o.update_cost()
    o.compute_presum()
    _, vc, ps = (a.reshape(L, X, Y) for a in o.costs())
    trees = [_random_tree(rng, L, X, Y, n) for n in (2, 3, 3, 4)]
    # aligned edges (straight wires) in one more tree
    XY = X * Y
    trees.append(([L * XY + 1 * Y + 1, 0 * XY + 1 * Y + 4, 1 * XY + 4 * Y + 1], [-1, 0, 0]))
    off = np.cumsum([0] + [len(t[0]) for t in trees])
    o.set_nets(off, np.concatenate([t[0] for t in trees]), np.concatenate([t[1] for t in trees]))
    for k, (nodes, parents) in enumerate(trees):
        best, route = o.lshape_probe(k)
        brute = _enum_cost(o, L, X, Y, DIR, nodes, parents, ps, vc)
        assert math.isclose(best, brute, rel_tol=1e-12), (k, best, brute)
        assert math.isclose(_route_cost(route, L, X, Y, DIR, nodes, parents, ps, vc), best, rel_tol=1e-12)


# ---------------------------------------------------------------------------
# Scheduler (src/database_cuda.hpp:288-631): invariants
# ---------------------------------------------------------------------------
def test_schedule_invariants(small_design):
    d = small_design
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    order = d.array("order")
    nb, batch = o.schedule(order)
    n = len(order)
    assert nb >= 1 and batch.min() == 0 and batch.max() == nb - 1
    assert len(np.unique(batch)) == nb  # no empty batch
    # the first net of the order (largest hpwl) wins every cell it claims
    assert batch[0] == 0
    pt_off, pt = d.array("pt_off"), d.array("pt")
    for b in range(nb):
        cells = np.concatenate([pt[pt_off[order[i]]:pt_off[order[i] + 1]] for i in np.nonzero(batch == b)[0]])
        # no two nets of a batch share a point cell (a net's own duplicates
        # are collapsed first)
        per_net = [np.unique(pt[pt_off[order[i]]:pt_off[order[i] + 1]]) for i in np.nonzero(batch == b)[0]]
        allc = np.concatenate(per_net)
        assert len(np.unique(allc)) == len(allc), b
        assert len(cells) >= len(per_net)
    # deterministic and independent of the call
    nb2, batch2 = o.schedule(order)
    assert nb2 == nb and np.array_equal(batch, batch2)


def _sched_case(points, hsegs, X, Y, L=3):
    """Nets given by their point cells (x, y) and horizontal segments
    ((x0, y), (x1, y)); listed in priority order (first = largest hpwl)."""
    n = len(points)
    node_off = np.arange(n + 1, dtype=np.int64)
    node = np.array([L * X * Y + p[0][0] * Y + p[0][1] for p in points], np.int32)
    parent = np.full(n, -1, np.int32)
    pt_off = np.cumsum([0] + [len(p) for p in points]).astype(np.int64)
    pt = np.array([x * Y + y for p in points for (x, y) in p], np.int32)
    hs_off = np.cumsum([0] + [len(h) for h in hsegs]).astype(np.int64)
    hs = np.array([v for h in hsegs for (a, b) in h for v in (a[0] * Y + a[1], b[0] * Y + b[1])], np.int32)
    vs_off = np.zeros(n + 1, np.int64)
    o = orc.Oracle(L, X, Y, 1, 1.0, 4.0, np.ones(L), np.ones(X - 1), np.ones(Y - 1), np.ones(L * X * Y, np.float32))
    o.set_nets(node_off, node, parent, pt_off, pt, hs_off, hs, vs_off, np.zeros(0, np.int32))
    return o, np.arange(n, dtype=np.int32)


def test_schedule_known_answer_neighbours():
    """Hand-derived from src/database_cuda.hpp:288-434: a net that owns its
    point cells also claims the neighbours (x, y-1), (x-1, y), (x+1, y) --
    not (x, y+1) -- so lower nets with a point there wait for the next
    batch."""
    A, B, C, D, E, F = [(2, 2)], [(3, 2)], [(2, 3)], [(1, 2)], [(2, 1)], [(2, 4)]
    o, order = _sched_case([A, B, C, D, E, F], [[]] * 6, X=8, Y=8)
    nb, batch = o.schedule(order)
    assert nb == 2
    assert list(batch) == [0, 1, 0, 1, 1, 0]


def test_schedule_known_answer_segment_sampling():
    """Segment cells are claimed every 'sparsity' cells from the start: 15
    in mode 0 and 2 in modes 1 and 2 (src/database_cuda.hpp:323-345,
    :2238-2246)."""
    G = ([(0, 5), (20, 5)], [((0, 5), (20, 5))])
    H = ([(15, 5)], [])   # on a mode-0 sample only (15 is odd)
    I = ([(8, 5)], [])    # on a mode-1 sample only
    J = ([(11, 5)], [])   # on no sample and no neighbour of another point
    o, order = _sched_case([G[0], H[0], I[0], J[0]], [G[1], H[1], I[1], J[1]], X=30, Y=12)
    nb, batch = o.schedule(order, 0)
    assert list(batch) == [0, 1, 0, 0]
    nb, batch = o.schedule(order, 1)
    assert list(batch) == [0, 0, 1, 0]


def test_schedule_sparsity_mode(small_design):
    d = small_design
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    order = d.array("order")
    nb0, _ = o.schedule(order, 0)
    nb1, _ = o.schedule(order, 1)
    # denser segment sampling can only add conflicts
    assert nb1 >= nb0


# ---------------------------------------------------------------------------
# Pattern pass commits, rip-up marking, congestion view (recounts)
# ---------------------------------------------------------------------------
def _pattern(d):
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    return o


def _wires(o, L, X, Y):
    XY = X * Y
    cnt = np.zeros((L, X, Y), np.int64)
    per_net = []
    for i in range(o.num_nets):
        r = o.route(i)
        segs = []
        for k in range(1, r[0], 2):
            a, b = r[k], r[k + 1]
            la, lb = a // XY, b // XY
            if la != lb:
                continue
            x0, y0, x1, y1 = a // Y % X, a % Y, b // Y % X, b % Y
            if x0 != x1:
                cnt[la, min(x0, x1):max(x0, x1), y0] += 1
            elif y0 != y1:
                cnt[la, x0, min(y0, y1):max(y0, y1)] += 1
            segs.append((la, x0, y0, x1, y1))
        per_net.append(segs)
    return cnt, per_net


def test_commit_demand_recount(small_design):
    d = small_design
    i = d.info
    L, X, Y = i.L, i.X, i.Y
    o = _pattern(d)
    dem = o.demand().reshape(L, X, Y).astype(np.float64)
    wires, _ = _wires(o, L, X, Y)
    extra = dem - wires
    # wires add whole tracks; vias add halves (whole at a track end) and are
    # skipped where the commit stamp is already set, so the remainder is a
    # non-negative multiple of 1/2 bounded by the via cuts
    assert np.all(extra >= 0)
    assert np.all((extra * 2) == np.round(extra * 2))
    XY = X * Y
    cuts = 0
    for n in range(o.num_nets):
        r = o.route(n)
        for k in range(1, r[0], 2):
            cuts += abs(r[k] // XY - r[k + 1] // XY)
    assert extra.sum() <= cuts
    assert extra.sum() > 0


def test_ripup_mark_recount(small_design):
    d = small_design
    i = d.info
    L, X, Y = i.L, i.X, i.Y
    o = _pattern(d)
    dem = o.demand().reshape(L, X, Y)
    cap, _, _, _ = d.grid_arrays()
    cap = cap.reshape(L, X, Y)
    full = cap <= dem  # threshold 0 (src/graph.hpp:432-438)
    for l in range(L):
        if ((l & 1) ^ i.DIR) == 0:
            full[l, X - 1, :] = False
        else:
            full[l, :, Y - 1] = False
    _, per_net = _wires(o, L, X, Y)
    expect = np.zeros(o.num_nets, np.uint8)
    for n, segs in enumerate(per_net):
        c = 0
        for (l, x0, y0, x1, y1) in segs:
            if x0 != x1:
                c += int(full[l, min(x0, x1):max(x0, x1), y0].sum())
            elif y0 != y1:
                c += int(full[l, x0, min(y0, y1):max(y0, y1)].sum())
        expect[n] = 1 if c > 10 else 0
    got = o.ripup_mark(0)
    assert np.array_equal(got, expect)
    # a negative threshold marks at least as many nets
    assert o.ripup_mark(-1).sum() >= got.sum()


def test_congestion_view_recount(small_design):
    d = small_design
    i = d.info
    L, X, Y = i.L, i.X, i.Y
    o = _pattern(d)
    dem = o.demand().reshape(L, X, Y)
    cap, _, _, _ = d.grid_arrays()
    v = np.maximum(dem - cap.reshape(L, X, Y), 0).max(axis=0).astype(np.float64)
    xs, ys = o.congestion_view()
    xs, ys = xs.reshape(X, Y), ys.reshape(X, Y)
    # exclusive prefix sums of the view along x (per y) and along y (per x)
    # (src/graph.hpp:558-584); exact for these half-integer values
    assert np.array_equal(xs.astype(np.float64), np.cumsum(v, axis=0) - v)
    assert np.array_equal(ys.astype(np.float64), np.cumsum(v, axis=1) - v)
    assert v.sum() > 0
