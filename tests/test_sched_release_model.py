"""The release-driven rounds of the batched scheduler give the reference's
batches.

The batched scheduler's later rounds follow the changes (csrc/hip/
sched_batched.hip, kernels k_r_*): a released cell goes to the first active
net of its priority-sorted list, scanned from a per-cell cursor; only those
new owners are tested for viability; the revert lists the cells the nets
committed in this round own and blocks the active nets in their lists; the
end of a round releases the cells of blocked nets; round 0 of a batch tests
only the candidates the previous batch's end listed (the nets viable in its
round 0 and still unassigned, and the new first listers of cells whose list
head moved).  That this equals the
reference's round-by-round claims (src/database_cuda.hpp:288-434,437-631) is
an argument about the rules, so it is checked here on the CPU: an executable
restatement of those kernels' steps, one step per kernel, against the oracle's
scheduler on random designs, orders and footprint modes.  The GPU tests check
the kernels themselves against the oracle on the benchmark design."""
import numpy as np
import pytest

import orc
from instantgr_amd.host import Design

ROUNDS = 30


def release_schedule(order, pt_off, pt, hs_off, hs, vs_off, vs, X, Y, mode, skip_block=False):
    n = len(order)
    sp = 2 if mode in (1, 2) else 15
    ids = [int(order[n - 1 - i]) for i in range(n)]          # scheduler index -> net id
    pts = [[int(p) for p in pt[pt_off[ids[i]]:pt_off[ids[i] + 1]]] for i in range(n)]
    XY = X * Y
    lists = [[] for _ in range(XY)]                          # per cell, highest index first
    for i in range(n):
        for p in pts[i]:
            lists[p].append(i)
    for q in range(XY):
        lists[q].sort(reverse=True)
    head = [0] * XY
    cur = [0] * XY
    cmap = [lists[q][0] if lists[q] else -1 for q in range(XY)]
    base = list(cmap)                                        # first unassigned lister
    last_compact = 0
    batch = [-1] * n
    status = [1] * n
    unassigned = list(range(n))
    cand0 = list(range(n))                                   # round-0 candidates

    def active(i):
        return status[i] == 1 and batch[i] == -1

    def rest_cells(i):
        net = ids[i]
        out = []
        for p in pts[i]:
            x, y = divmod(p, Y)
            if y - 1 >= 0:
                out.append(x * Y + y - 1)
            if x - 1 >= 0:
                out.append((x - 1) * Y + y)
            if x + 1 < X:
                out.append((x + 1) * Y + y)
        for k in range(hs_off[net], hs_off[net + 1]):
            s, e = int(hs[2 * k]), int(hs[2 * k + 1])
            out += [x * Y + s % Y for x in range(s // Y, e // Y + 1, sp)]
        for k in range(vs_off[net], vs_off[net + 1]):
            s, e = int(vs[2 * k]), int(vs[2 * k + 1])
            out += [(s // Y) * Y + y for y in range(s % Y, e % Y + 1, sp)]
        return out

    current, placed = 0, 0
    while True:
        have = False
        rel = []
        for rnd in range(ROUNDS):
            # k_r_claim: each released cell to its first active lister
            if rnd == 0:
                test = list(cand0)
            else:
                test, seen = [], set()
                for q in rel:
                    lst = lists[q]
                    for k in range(cur[q], len(lst)):
                        j = lst[k]
                        if active(j):
                            cmap[q] = j
                            cur[q] = k
                            if j not in seen:
                                seen.add(j)
                                test.append(j)
                            break
            # k_r_viable
            viable = [i for i in test if active(i) and all(cmap[p] == i for p in pts[i])]
            if rnd == 0:
                # the candidates hold every net viable in round 0
                assert set(viable) == {i for i in unassigned if all(cmap[p] == i for p in pts[i])}
                v0 = list(viable)
            # k_b_claim_rest
            for v in viable:
                for q in rest_cells(v):
                    cmap[q] = max(cmap[q], v)
            # k_r_commit
            committed = [v for v in viable if all(cmap[p] == v for p in pts[v])]
            for v in committed:
                batch[v] = current
            placed += len(committed)
            have = have or bool(committed)
            # k_r_hits, k_r_block
            if committed and not skip_block:
                hits = set()
                for c in committed:
                    for q in pts[c] + rest_cells(c):
                        if cmap[q] == c:
                            hits.add(q)
                for q in hits:
                    for j in lists[q][cur[q]:]:
                        if active(j):
                            status[j] = 2
                    cur[q] = len(lists[q])                   # all inactive from the cursor on
            # k_r_clear_end
            if not committed or rnd + 1 >= ROUNDS:
                break
            rel = [q for q in range(XY) if cmap[q] >= 0 and status[cmap[q]] == 2]
            for q in rel:
                cmap[q] = -1
        cand0, seen = [], set()
        # the lists drop their assigned entries every n/64 placed nets
        compact = (placed - last_compact) * 64 >= n
        if compact:
            last_compact = placed
        for q in range(XY):
            if compact:
                lists[q] = [j for j in lists[q][head[q]:] if batch[j] == -1]
                h = 0
            else:
                lst, h = lists[q], head[q]
                while h < len(lst) and batch[lst[h]] != -1:
                    h += 1
            head[q] = cur[q] = h
            cmap[q] = lists[q][h] if h < len(lists[q]) else -1
            moved = cmap[q] != base[q]
            base[q] = cmap[q]
            if moved and cmap[q] >= 0 and cmap[q] not in seen:
                seen.add(cmap[q])
                cand0.append(cmap[q])
        for i in v0:
            if batch[i] == -1 and i not in seen:
                seen.add(i)
                cand0.append(i)
        left = []
        for i in unassigned:
            status[i] = 1 if batch[i] == -1 else 0
            if batch[i] == -1:
                left.append(i)
        unassigned = left
        if not have:
            break
        current += 1
        if placed == n:
            break
    out = np.empty(n, np.int32)
    for i in range(n):
        out[n - 1 - i] = batch[i]
    return current, out


SHAPES = {"40x36, 260 nets": dict(layers=6, xsize=40, ysize=36, nets=260, pin_cap=10, sigma=5.0),
          "dense 24x24, 400 nets": dict(layers=6, xsize=24, ysize=24, nets=400, pin_cap=10, sigma=4.0)}


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_release_rounds_give_the_reference_batches(shape, seed, mode):
    # Later rounds and released-cell claims are frequent here: the dense
    # shape runs about 100 later rounds and 4 000 claims per schedule.
    d = Design.custom(seed=seed, **SHAPES[shape]).build(2)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    a = d.array
    X, Y = d.info.X, d.info.Y
    rng = np.random.default_rng(seed * 10 + mode)
    full = a("order")
    # the routing order, a random order, and a random subset
    for order in (full, rng.permutation(full), rng.permutation(full)[: len(full) // 2]):
        order = np.ascontiguousarray(order, np.int32)
        nb_o, b_o = o.schedule(order, mode)
        nb_m, b_m = release_schedule(order, a("pt_off"), a("pt"), a("hs_off"), a("hs"), a("vs_off"), a("vs"),
                                     X, Y, mode)
        assert nb_m == nb_o
        assert np.array_equal(b_m, b_o)
        assert (b_o >= 0).all()


def test_comparison_catches_a_dropped_block_step():
    # The comparison is not vacuous: the same restatement without the
    # blocking step gives other batches than the reference.
    d = Design.custom(layers=6, xsize=40, ysize=36, nets=260, pin_cap=10, sigma=5.0, seed=1).build(2)
    o = orc.Oracle.from_design(d)
    o.set_nets_from_design(d)
    a = d.array
    order = np.ascontiguousarray(a("order"), np.int32)
    nb_o, b_o = o.schedule(order, 0)
    nb_m, b_m = release_schedule(order, a("pt_off"), a("pt"), a("hs_off"), a("hs"), a("vs_off"), a("vs"),
                                 d.info.X, d.info.Y, 0, skip_block=True)
    assert nb_m != nb_o or not np.array_equal(b_m, b_o)
