"""The GPU evaluator's rules give the host evaluator's score.

The GPU evaluator (csrc/hip/eval_kernels.hip) does not search cells as the
evaluator program does (run/evaluator.cpp:458-633,635-675): it turns each
route-file line into an item (a wire, a via column or a dead line), counts a
via's stacked demand once per net and cell where none of the net's wires is
present, and joins items whose cells are equal or search neighbours, with
union-find.  That this lifting gives the same opens and the same score is
checked here on the CPU: a restatement of those rules, item by item as the
kernels apply them, against the host evaluator (which the evaluator tests pin
to the reference's golden scores) on oracle-routed designs with planted
changes: dropped lines, stray wires and vias, wires that only touch end to
end, malformed lines.  The GPU tests compare the kernels with the host
evaluator on routed designs."""
import os
import tempfile

import numpy as np
import pytest

import orc
from instantgr_amd.host import Design


def _decode(p, q, ML, X, Y, dirs):
    if p > q:
        p, q = q, p
    XY = X * Y
    if p < 0 or q >= ML * XY:
        return None
    za, xa, ya = p // XY, p // Y % X, p % Y
    zb, xb, yb = q // XY, q // Y % X, q % Y
    if za != zb:
        return (xa, xa, ya, ya, za, zb) if (xa == xb and ya == yb) else None
    if za == 0:
        return None
    if dirs[za] == 0:
        return (xa, xb, ya, ya, za, za) if (xb > xa and yb == ya) else None
    return (xa, xa, ya, yb, za, za) if (yb > ya and xb == xa) else None


def _joined(s, t, dirs, lateral=True):
    xo = s[0] <= t[1] and t[0] <= s[1]
    yo = s[2] <= t[3] and t[2] <= s[3]
    zl, zh = max(s[4], t[4]), min(s[5], t[5])
    if zl > zh:
        return (s[5] + 1 == t[4] or t[5] + 1 == s[4]) and xo and yo
    if xo and yo:
        return True
    if not lateral:
        return False
    xn = s[0] <= t[1] + 1 and t[0] <= s[1] + 1
    yn = s[2] <= t[3] + 1 and t[2] <= s[3] + 1
    return any((yo and xn) if dirs[z] == 0 else (xo and yn) for z in range(max(zl, 1), zh + 1))


def _contains(t, x, y, z):
    return t[4] <= z <= t[5] and t[0] <= x <= t[1] and t[2] <= y <= t[3]


def model_score(d, off, data, lateral=True, wire_check=True):
    # lateral / wire_check: switches for the planted mistakes below
    a = d.array
    i = d.info
    X, Y, L = i.X, i.Y, i.L
    ML, XY = L + 1, X * Y
    dirs = a("metal_dir")
    cap = a("metal_cap")
    w = a("metal_of_weight")
    _, _, xlen, ylen = d.grid_arrays()
    sub_off, subs = a("net_sub_off"), a("net_sub")
    pin_off, pins = a("pin_off"), a("pins")
    ex_off, ex = a("extra_off"), a("extra")
    dp_off, ap_off, ap = a("net_pin_off"), a("ap_off"), a("ap")
    wire = np.zeros(ML * XY, np.int64)
    svia = np.zeros(ML * XY, np.int64)
    vias = np.zeros(ML, np.int64)
    opens = 0
    for n in range(i.nets):
        lines = [(int(p), int(p) + XY) for p in pins[pin_off[n]:pin_off[n + 1]] if p < XY]
        lines += [(int(ex[e]), int(ex[e + 1])) for e in range(ex_off[n], ex_off[n + 1] - 1, 2)]
        for s in subs[sub_off[n]:sub_off[n + 1]]:
            r = data[off[s]:off[s] + data[off[s]]]
            lines += [(int(r[j]) + XY, int(r[j + 1]) + XY) for j in range(1, len(r) - 1, 2)]
        items = [_decode(p, q, ML, X, Y, dirs) for p, q in lines]
        bad = any(t is None for t in items)
        items = [t for t in items if t is not None]
        for t in items:   # demand and via counts (k_ev_wires)
            if t[4] != t[5]:
                vias[t[4]:t[5]] += 1
            elif t[1] > t[0]:
                wire[t[4] * XY + np.arange(t[0], t[1]) * Y + t[2]] += 1
            else:
                wire[t[4] * XY + t[0] * Y + np.arange(t[2], t[3])] += 1
        for k, c in enumerate(items):   # stacked vias (k_ev_net)
            if c[4] >= c[5]:
                continue
            for z in range(max(c[4], 1), c[5]):
                skip = any((j < k and t[4] < t[5] and t[0] == c[0] and t[2] == c[2] and t[4] <= z < t[5])
                           or (wire_check and t[4] == t[5] and (t[1] > t[0] or t[3] > t[2]) and t[4] == z
                               and _contains(t, c[0], c[2], z)) for j, t in enumerate(items))
                if skip:
                    continue
                x, y = c[0], c[2]
                base = z * XY
                if dirs[z] == 0:
                    if 0 < x < X - 1:
                        svia[base + (x - 1) * Y + y] += 1
                        svia[base + x * Y + y] += 1
                    elif x > 0:
                        svia[base + (x - 1) * Y + y] += 2
                    elif x < X - 1:
                        svia[base + x * Y + y] += 2
                else:
                    if 0 < y < Y - 1:
                        svia[base + x * Y + y - 1] += 1
                        svia[base + x * Y + y] += 1
                    elif y > 0:
                        svia[base + x * Y + y - 1] += 2
                    elif y < Y - 1:
                        svia[base + x * Y + y] += 2
        p0, p1 = dp_off[n], dp_off[n + 1]
        if p1 - p0 < 2:
            opens += bad
            continue
        par = list(range(len(items)))

        def find(x):
            while par[x] != x:
                x = par[x]
            return x
        for k in range(len(items)):
            for j in range(k + 1, len(items)):
                if _joined(items[k], items[j], dirs, lateral):
                    rk, rj = find(k), find(j)
                    if rk != rj:
                        par[max(rk, rj)] = min(rk, rj)

        def apc(v):
            return v // Y % X, v % Y, v // XY
        traced = set()
        for v in ap[ap_off[p0]:ap_off[p0 + 1]]:
            x, y, z = apc(int(v))
            traced |= {find(j) for j, t in enumerate(items) if _contains(t, x, y, z)}
        ok = not bad
        for p in range(p0 + 1, p1):
            if not ok:
                break
            ok = any(_contains(t, *apc(int(v))) and find(j) in traced
                     for v in ap[ap_off[p]:ap_off[p + 1]] for j, t in enumerate(items))
        opens += not ok
    # score (k_ev_score)
    z = np.repeat(np.arange(ML), XY)
    x = np.tile(np.repeat(np.arange(X), Y), ML)
    y = np.tile(np.arange(Y), ML * X)
    m = z >= 1
    demand = ((2 * wire + svia) & 0xFFFF).astype(np.float64)
    c = np.asarray(cap, np.float64)
    with np.errstate(over="ignore"):
        term = np.where(c > 0.001, np.exp((demand - 2 * c) / 2 * 0.5),
                        np.where((c >= 0) & (demand > 0), np.exp(1.5 * demand * 0.5), 0.0))
    of = float(np.sum((term * np.asarray(w)[z])[m]))
    xl = np.append(np.asarray(xlen, np.int64), 0)
    yl = np.append(np.asarray(ylen, np.int64), 0)
    ln = np.where(np.asarray(dirs)[z] == 0, xl[x], yl[y])
    wl = int(np.sum((wire * ln)[m]))
    return {"wirelength": wl * i.unit_wire, "via": float(np.sum(vias)) * i.unit_via, "overflow": of,
            "opens": opens}


def _host(d, off, data):
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "r.out")
        d.write_routes(p, off, data)
        return d.evaluate(p)


def _check(m, h):
    assert m["wirelength"] == h["wirelength"]
    assert m["via"] == h["via"]
    assert abs(m["overflow"] - h["overflow"]) <= 1e-9 * abs(h["overflow"])
    assert m["opens"] == h["opens"]


@pytest.fixture(scope="module", params=[5, 6])
def routed(request):
    d = Design.custom(layers=6, xsize=30, ysize=28, nets=160, pin_cap=8, sigma=4.0, seed=request.param).build(2)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    o.pattern_route(d.array("order"))
    off, data = o.routes_flat()
    return d, off, data


def test_routed_design(routed):
    d, off, data = routed
    h = _host(d, off, data)
    assert h["opens"] == 0
    _check(model_score(d, off, data), h)


def _rebuild(slices):
    off = np.zeros(len(slices) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in slices])
    return off, np.concatenate(slices).astype(np.int32)


def _plant(d, off, data):
    i = d.info
    X, Y, L = i.X, i.Y, i.L
    dirs = d.array("metal_dir")   # metal layers; routing layer l is metal l + 1
    rng = np.random.default_rng(3)
    slices = [data[off[s]:off[s] + data[off[s]]].copy() for s in range(len(off) - 1)]
    for s in range(len(slices)):
        r = list(slices[s][1:])
        pairs = [(r[k], r[k + 1]) for k in range(0, len(r), 2)]
        u = rng.random()
        if u < 0.15 and pairs:
            pairs.pop(int(rng.integers(len(pairs))))           # a dropped line
        elif u < 0.35:
            l = int(rng.integers(L))
            x, y = int(rng.integers(X)), int(rng.integers(Y))
            c = l * X * Y + x * Y + y
            if rng.random() < 0.5 and l + 1 < L:
                pairs.append((c, c + X * Y))                    # a stray via
            elif dirs[l + 1] == 0 and x + 2 < X:
                pairs.append((c, c + 2 * Y))                    # a stray wire
                if x + 5 < X:
                    pairs.append((c + 3 * Y, c + 5 * Y))        # touching it end to end, one cell apart
            elif dirs[l + 1] == 1 and y + 2 < Y:
                pairs.append((c, c + 2))
                if y + 5 < Y:
                    pairs.append((c + 3, c + 5))
        elif u < 0.40 and pairs:
            a, b = pairs[0]
            x, y = a // Y % X, a % Y
            pairs[0] = (a, a - x * Y - y + ((x + 1) % X) * Y + (y + 1) % Y)   # a malformed (diagonal) line
        flat = [1 + 2 * len(pairs)] + [v for p in pairs for v in p]
        slices[s] = np.array(flat, np.int32)
    return _rebuild(slices)


def test_planted_changes(routed):
    d, off, data = routed
    off2, data2 = _plant(d, off, data)
    h = _host(d, off2, data2)
    assert h["opens"] > 0
    _check(model_score(d, off2, data2), h)


def test_comparison_catches_planted_mistakes(routed):
    # Each of these rules matters on the planted routes: joining items only
    # where they share a cell (no neighbour rule along the layer direction)
    # changes the open count, and counting stacked vias under the net's own
    # wires changes the overflow cost.
    d, off, data = routed
    off2, data2 = _plant(d, off, data)
    h = _host(d, off2, data2)
    assert model_score(d, off2, data2, lateral=False)["opens"] != h["opens"]
    m = model_score(d, off2, data2, wire_check=False)
    assert abs(m["overflow"] - h["overflow"]) > 1e-6 * abs(h["overflow"])
