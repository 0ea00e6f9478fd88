"""Known answers derived by hand from the reference sources, for the parts of
the routing logic that no closed form or brute force pins elsewhere:

* demand commits (src/graph.hpp:104-137 wire difference array and
  commit_all_edge, :468-495 commit_wire_demand, :496-543 commit_via_demand):
  the timestamp gate, the half/border split of via demand, the segment end
  cell's timestamp, the rip-up asymmetry;
* the batch scheduler's multi-round rules (src/database_cuda.hpp:288-434,
  :437-631): claims, neighbour claims of viable nets, revert, clear, the
  30-round cap;
* the detour-DAG builder (src/database_cuda.hpp:1429-2086): L corners, Z-shape
  ranking and selection (:1872-1928), subdivision (:1662-1722), mirror places
  (:1733-1771) with their corner nodes and duplicated pins (:1807-1853,
  :1949-1966).

Every expected value below was worked out from the cited lines, step by
step, in the comments beside it; none is produced by the code under test.
"""
import os

import numpy as np
import pytest

import orc
from instantgr_amd.host import Design


# ---------------------------------------------------------------------------
# Demand commits.  Grid: 3 routing layers, X = 6, Y = 5, DIR = 0, so layer 0
# is horizontal (edge (x,y)-(x+1,y) stored at (x,y)), layer 1 vertical, layer
# 2 horizontal.  Cell (l, x, y) = l*30 + x*5 + y.
# ---------------------------------------------------------------------------
L3, X6, Y5 = 3, 6, 5


def cell(l, x, y):
    return l * X6 * Y5 + x * Y5 + y


def route(*pairs):
    r = [1 + 2 * len(pairs)]
    for a, b in pairs:
        r += [a, b]
    return np.array(r, np.int32)


def fresh():
    cap = np.full(L3 * X6 * Y5, 4.0, np.float32)
    o = orc.Oracle(L3, X6, Y5, 0, 1.0, 4.0, np.ones(L3), np.ones(X6 - 1, np.int32), np.ones(Y5 - 1, np.int32), cap)
    return o


def demand_dict(o):
    d = o.demand()
    return {(int(i) // 30, int(i) // 5 % 6, int(i) % 5): float(d[i]) for i in np.nonzero(d)[0]}


def test_via_interior_halves():
    # Via (2,2) from layer 0 to layer 2: cuts l = 0 and l = 1 (graph.hpp:517).
    # l = 0 is horizontal, 0 < x = 2 < X-1: 0.5 on (0,1,2) and (0,2,2)
    # (graph.hpp:533-535).  l = 1 is vertical, 0 < y = 2 < Y-1: 0.5 on
    # (1,2,1) and (1,2,2) (graph.hpp:524-526).
    o = fresh()
    o.commit_routes([route((cell(0, 2, 2), cell(2, 2, 2)))])
    assert demand_dict(o) == {(0, 1, 2): 0.5, (0, 2, 2): 0.5, (1, 2, 1): 0.5, (1, 2, 2): 0.5}
    assert not o.timestamp().any()     # a via commit writes no timestamp (graph.hpp:539)


def test_via_border_whole_units():
    # Border cells take the whole unit on the one edge inside the grid
    # (graph.hpp:520-523, 528-532).
    o = fresh()
    o.commit_routes([
        route((cell(0, 0, 3), cell(1, 0, 3))),   # l=0 horizontal, x=0: +1 on (0,0,3)
        route((cell(0, 5, 1), cell(1, 5, 1))),   # l=0, x=X-1=5: +1 on (0,4,1)
        route((cell(1, 3, 0), cell(2, 3, 0))),   # l=1 vertical, y=0: +1 on (1,3,0)
        route((cell(1, 3, 4), cell(2, 3, 4))),   # l=1, y=Y-1=4: +1 on (1,3,3)
    ])
    assert demand_dict(o) == {(0, 0, 3): 1.0, (0, 4, 1): 1.0, (1, 3, 0): 1.0, (1, 3, 3): 1.0}


def test_same_batch_wire_gates_via_cuts():
    # One route: a wire on layer 0 from x=1 to x=4 at y=2, a via at x=2 (on
    # the wire), a via at x=4 (the wire's end cell) up to layer 2, and a via
    # at x=0 (off the wire).
    # Wire: pre[1] += 1, pre[4] -= 1 and ts[4] = s (graph.hpp:106-116); the
    # exclusive scan gives edges x = 1, 2, 3 coverage 1, so demand +1 there
    # and ts = s on x = 1, 2, 3 (graph.hpp:121-125).  ts = s on x = 1..4.
    # Via x=2, cut l=0: ts = s, not < s -> skipped (graph.hpp:518).
    # Via x=4, cut l=0: ts = s (end cell) -> skipped; cut l=1: ts = 0 ->
    # vertical, y=2 -> 0.5 on (1,4,1) and (1,4,2).
    # Via x=0, cut l=0: ts = 0 -> horizontal border x=0 -> +1 on (0,0,2).
    o = fresh()
    o.commit_routes([route((cell(0, 1, 2), cell(0, 4, 2)), (cell(0, 2, 2), cell(1, 2, 2)),
                           (cell(0, 4, 2), cell(2, 4, 2)), (cell(0, 0, 2), cell(1, 0, 2)))])
    assert demand_dict(o) == {(0, 0, 2): 1.0, (0, 1, 2): 1.0, (0, 2, 2): 1.0, (0, 3, 2): 1.0,
                              (1, 4, 1): 0.5, (1, 4, 2): 0.5}
    ts = o.timestamp()
    s = o.stamp
    assert [int(ts[cell(0, x, 2)]) for x in range(6)] == [0, s, s, s, s, 0]
    assert int(np.count_nonzero(ts)) == 4


def test_earlier_batch_wire_does_not_gate():
    # Batch 1 commits the wire (ts = s1 on x = 1..4); batch 2's via at x=2
    # sees ts = s1 < s2 and adds 0.5 on (0,1,2) and (0,2,2).
    o = fresh()
    o.commit_routes([route((cell(0, 1, 2), cell(0, 4, 2)))])
    o.commit_routes([route((cell(0, 2, 2), cell(1, 2, 2)))])
    assert demand_dict(o) == {(0, 1, 2): 1.5, (0, 2, 2): 1.5, (0, 3, 2): 1.0}


def test_overlapping_vertical_segments():
    # Layer 1 is vertical: wires x=2, y 0..4 and y 1..3 in one batch.
    # Differences pre[0]+1, pre[4]-1, pre[1]+1, pre[3]-1; exclusive scan along
    # y gives edges y=0:1, y=1:2, y=2:2, y=3:1 (graph.hpp:74-99, 126-130).
    o = fresh()
    o.commit_routes([route((cell(1, 2, 0), cell(1, 2, 4))), route((cell(1, 2, 1), cell(1, 2, 3)))])
    assert demand_dict(o) == {(1, 2, 0): 1.0, (1, 2, 1): 2.0, (1, 2, 2): 2.0, (1, 2, 3): 1.0}
    ts = o.timestamp()
    assert all(int(ts[cell(1, 2, y)]) == o.stamp for y in range(5))


def test_ripup_mirrors_commit_when_stamps_agree():
    # Rip-up of the gated route above (graph.hpp:468-495 with K = -1, then
    # :496-543): the wire gives -1 on x = 1..3 and ts = s' on x = 1..3 and on
    # the corners (1,2), (4,2); the via at x=2 and the l=0 cut at x=4 are
    # skipped again (ts = s'), the l=1 cut at x=4 and the x=0 via subtract
    # what they added.  Everything returns to 0.
    o = fresh()
    r = route((cell(0, 1, 2), cell(0, 4, 2)), (cell(0, 2, 2), cell(1, 2, 2)),
              (cell(0, 4, 2), cell(2, 4, 2)), (cell(0, 0, 2), cell(1, 0, 2)))
    o.commit_routes([r])
    o.commit_routes([r], K=-1)
    assert demand_dict(o) == {}


def test_ripup_stamps_the_corner_cells():
    # A wire written from x=4 down to x=1 (routes keep the DP's endpoint
    # order).  Commit: edges 1..3 get demand, ts = s on x = 1..4, the via at
    # x=4 (cut l=0) is gated.  Rip-up (graph.hpp:478-490): the loop over
    # x in [1, 4) stamps x = 1..3; only the corner assignment stamps (4,2), the
    # wire's first cell, which gates the via's l=0 cut again.  Everything
    # returns to 0; without the corner stamp the cut would subtract 0.5 from
    # (0,3,2) and (0,4,2).
    o = fresh()
    r = route((cell(0, 4, 2), cell(0, 1, 2)), (cell(0, 4, 2), cell(1, 4, 2)))
    o.commit_routes([r])
    assert demand_dict(o) == {(0, 1, 2): 1.0, (0, 2, 2): 1.0, (0, 3, 2): 1.0}
    o.commit_routes([r], K=-1)
    assert demand_dict(o) == {}
    assert int(o.timestamp()[cell(0, 4, 2)]) == o.stamp


def test_ripup_of_a_gated_via_is_not_gated():
    # Batch A: P = wire x=1..4 on layer 0, Q = via at x=2, same batch: Q's
    # cut is gated (ts = sA), so it adds nothing.  Ripping Q alone later:
    # Q has no wire, so ts(0,2,2) stays sA < s_rip and the cut subtracts 0.5
    # from (0,1,2) and (0,2,2) (graph.hpp:518, 533-535) that were never
    # added: the reference's commit/rip-up asymmetry.
    o = fresh()
    p = route((cell(0, 1, 2), cell(0, 4, 2)))
    q = route((cell(0, 2, 2), cell(1, 2, 2)))
    o.commit_routes([p, q])
    assert demand_dict(o) == {(0, 1, 2): 1.0, (0, 2, 2): 1.0, (0, 3, 2): 1.0}
    o.commit_routes([q], K=-1)
    assert demand_dict(o) == {(0, 1, 2): 0.5, (0, 2, 2): 0.5, (0, 3, 2): 1.0}


# ---------------------------------------------------------------------------
# Scheduler.  Nets with points only (no segments); order[0] has the highest
# priority, which the scheduler gives the largest internal index
# (database_cuda.hpp:444-447).  Neighbour claims of a viable net are
# (x, y-1), (x-1, y), (x+1, y) of each point (database_cuda.hpp:301-312).
# ---------------------------------------------------------------------------
def _sched(X, Y, nets_points, mode=0):
    n = len(nets_points)
    o = orc.Oracle(2, X, Y, 0, 1.0, 4.0, np.ones(2), np.ones(X - 1, np.int32), np.ones(Y - 1, np.int32),
                   np.full(2 * X * Y, 4.0, np.float32))
    node_off = np.arange(n + 1, dtype=np.int64)
    node = np.array([2 * X * Y + p[0][0] * Y + p[0][1] for p in nets_points], np.int32)
    parent = np.full(n, -1, np.int32)
    pt_off = np.zeros(n + 1, np.int64)
    pt_off[1:] = np.cumsum([len(p) for p in nets_points])
    pt = np.array([x * Y + y for p in nets_points for (x, y) in p], np.int32)
    zero = np.zeros(n + 1, np.int64)
    o.set_nets(node_off, node, parent, pt_off, pt, zero, np.zeros(0, np.int32), zero, np.zeros(0, np.int32))
    nb, batch_of = o.schedule(np.arange(n, dtype=np.int32), mode)
    return nb, [int(b) for b in batch_of]


def test_scheduler_rounds_commit_block_clear():
    # A > B > C > D in priority; internal indices A=3, B=2, C=1, D=0.
    # Batch 0, round 0: point claims give (5,5)=A, (6,5)=B, (2,2)=max(B,C)=B,
    # (5,4)=D.  Viable: A, B, D.  Their neighbour claims: A takes (5,4) from D
    # and (6,5) from B; B's claims do not reach A's or D's points.  Commit:
    # A only.  Revert: B ((6,5) owned by committed A) and D ((5,4) owned by
    # A) get status 2; C's point is owned by B, not committed, so C stays.
    # Clear: every cell owned by B or D goes to -1, among them (2,2).
    # Round 1: C alone claims (2,2), owns it and commits.  Round 2 commits
    # nothing, so batch 0 = {A, C}.  Batch 1: B and D claim disjoint points
    # and neighbour cells that are not each other's points: both commit.
    pts = [[(5, 5)], [(6, 5), (2, 2)], [(2, 2)], [(5, 4)]]
    nb, b = _sched(10, 10, pts)
    assert (nb, b) == (2, [0, 1, 0, 1])


def test_scheduler_round_cap():
    # Chain P0 > P1 > ... > P61, P_i on cells c_i and c_{i+1}, cells three
    # gcells apart (no neighbour claim reaches another cell).  Batch 0 round
    # r: P_{2r} owns both its cells and commits, P_{2r+1} is blocked by it
    # and its cells are cleared, which frees c_{2r+2} for P_{2r+2}.  Rounds
    # 0..29 commit P0, P2, ..., P58; the 30-round cap (database_cuda.hpp:555)
    # ends the batch although P60 could commit next.  Batch 1: every odd
    # P_{2j-1} (j = 1..30) owns both its cells; P60 loses c_60 to P59 and is
    # blocked, which clears c_61 for P61 (round 1).  Batch 2: P60.
    m = 62
    cells = [(3 * (k % 8), 3 * (k // 8)) for k in range(m + 1)]
    pts = [[cells[i], cells[i + 1]] for i in range(m)]
    nb, b = _sched(24, 27, pts)
    want = [0 if i % 2 == 0 else 1 for i in range(m)]
    want[60] = 2
    want[61] = 1
    assert nb == 3 and b == want


# ---------------------------------------------------------------------------
# Detour-DAG builder.  Designs written as .cap/.net files: 10 metal layers
# (9 routing layers), 100 x 100 gcells, capacity 4 everywhere above metal 1.
# Pins on metal 2 land on routing layer 0; DAG nodes that are not pins carry
# layer L = 9.
# ---------------------------------------------------------------------------
X100 = Y100 = 100


def _write_design(tmp, nets):
    M = 10
    cap = os.path.join(tmp, "d.cap")
    net = os.path.join(tmp, "d.net")
    with open(cap, "w") as f:
        f.write("%d %d %d\n1 4 %s\n" % (M, X100, Y100, " ".join(["1"] * M)))
        f.write(" ".join(["1"] * (X100 - 1)) + "\n" + " ".join(["1"] * (Y100 - 1)) + "\n")
        for m in range(M):
            f.write("metal%d %d 1\n" % (m + 1, m & 1))
            row = " ".join(["0" if m == 0 else "4"] * X100) + "\n"
            f.write(row * Y100)
    with open(net, "w") as f:
        for i, pins in enumerate(nets):
            f.write("n%d\n(\n" % i)
            for (x, y) in pins:
                f.write("[(1, %d, %d)]\n" % (x, y))
            f.write(")\n")
    return Design.load(cap, net).build(1)


def _dag(d, sub, xsum, ysum, mode=0):
    g = d.build_dags(np.array([sub], np.int32), xsum, ysum, mode, 1)
    n0, n1 = int(g["dag_node_off"][0]), int(g["dag_node_off"][1])
    node = g["dag_node"][n0:n1]
    po = g["dag_par_off"][n0:n1 + 1]
    par = g["dag_par"][po[0]:po[-1]]
    slot = g["dag_slot"][po[0]:po[-1]]
    XY = X100 * Y100

    def key(v):
        return (int(v) // XY, int(v) // Y100 % X100, int(v) % Y100)
    edges = []
    for i in range(n1 - n0):
        for k in range(po[i] - po[0], po[i + 1] - po[0]):
            edges.append((key(node[par[k]]), key(node[i]), int(slot[k])))
    depth = g["dag_depth"][n0:n1]
    # renumbering invariants: parents first, depth = longest path
    for i in range(n1 - n0):
        ps = [int(par[k]) for k in range(po[i] - po[0], po[i + 1] - po[0])]
        assert all(p < i for p in ps)
        assert depth[i] == (max(depth[p] for p in ps) + 1 if ps else 0)
    pts = g["dag_pt"][g["dag_pt_off"][0]:g["dag_pt_off"][1]]
    return sorted(key(v) for v in node), sorted(edges), sorted(int(p) for p in pts)


def _layer_of_pin(d, sub, x, y):
    off = d.array("tree_node_off")
    for v in d.array("tree_node")[off[sub]:off[sub + 1]]:
        if int(v) % (X100 * Y100) == x * Y100 + y:
            return int(v) // (X100 * Y100)
    raise AssertionError("pin not in the tree")


def _root_first(d, sub, a, b):
    """The tree node list of a two-pin subnet starts with the root that
    select_root_net picks (database_cuda.hpp:826-853: a breadth-first walk
    from the leaves in index order, the last popped entry wins, which for two
    nodes is node 0)."""
    off = d.array("tree_node_off")
    first = int(d.array("tree_node")[off[sub]]) % (X100 * Y100)
    return (a, b) if first == a[0] * Y100 + a[1] else (b, a)


@pytest.fixture(scope="module")
def dag_design(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("dag"))
    return _write_design(tmp, [[(20, 30), (26, 34)],    # L-shaped: corners and Z shapes
                               [(20, 50), (40, 50)],    # straight, mirrored when row 50 is congested
                               [(10, 75), (70, 75)],    # straight, 60 gcells: one midpoint
                               [(0, 80), (60, 80)],     # 60 gcells from the grid edge
                               [(1, 95), (99, 95)],     # 98 gcells: one midpoint
                               [(20, 60), (60, 60)]])   # straight, 40 gcells, mirrored


def _sub_of(d, xy):
    """Subnet whose tree holds gcell xy."""
    off = d.array("tree_node_off")
    tn = d.array("tree_node")
    for s in range(len(off) - 1):
        if any(int(v) % (X100 * Y100) == xy[0] * Y100 + xy[1] for v in tn[off[s]:off[s + 1]]):
            return s
    raise AssertionError("no subnet at %r" % (xy,))


def test_dag_corners_and_z_shapes(dag_design):
    d = dag_design
    s = _sub_of(d, (20, 30))
    (r, c) = _root_first(d, s, (20, 30), (26, 34))
    assert (r, c) == ((20, 30), (26, 34)), "two-pin root is tree node 0"
    lr, lc = _layer_of_pin(d, s, 20, 30), _layer_of_pin(d, s, 26, 34)
    z = np.zeros(X100 * Y100, np.float32)
    nodes, edges, pts = _dag(d, s, z, z, mode=0)
    # Root R=(20,30), child C=(26,34), slot 0 (database_cuda.hpp:1966-1970).
    # L corners (:1861-1870): (px, cury) = (20,34) and (curx, py) = (26,30).
    # Z shapes along x (pathIndex 0, :1872-1928): len 6, max_z_shape 6, all
    # costs 0 with an empty view, sorted by displace id, select 2 (mode 0):
    # d=1: x = (20*1 + 26*5) / 6 = 25; d=2: x = (40 + 104) / 6 = 24; nodes
    # (x,30) -> (x,34).  Along y (pathIndex 1): len 4, max_z_shape 4: d=1:
    # y = (30 + 34*3) / 4 = 33; d=2: y = (60 + 68) / 4 = 32; nodes (20,y) ->
    # (26,y).  No straight edge R->C (it is diagonal).
    L = 9
    R, C = (lr, 20, 30), (lc, 26, 34)
    want_edges = [(R, (L, 20, 34), 0), ((L, 20, 34), C, 0), (R, (L, 26, 30), 0), ((L, 26, 30), C, 0)]
    for x in (25, 24):
        want_edges += [(R, (L, x, 30), 0), ((L, x, 30), (L, x, 34), 0), ((L, x, 34), C, 0)]
    for y in (33, 32):
        want_edges += [(R, (L, 20, y), 0), ((L, 20, y), (L, 26, y), 0), ((L, 26, y), C, 0)]
    assert edges == sorted(want_edges)
    assert nodes == sorted({e[0] for e in want_edges} | {e[1] for e in want_edges})
    # Scheduler points: both ends of every edge (2-D) plus each Z shape's
    # middle ((x1+x2)/2, (y1+y2)/2): (x,32) for x = 25, 24 and (23,y) for
    # y = 33, 32 (:1921).
    want_pts = {n[1] * Y100 + n[2] for n in nodes} | {25 * Y100 + 32, 24 * Y100 + 32, 23 * Y100 + 33,
                                                      23 * Y100 + 32}
    assert pts == sorted(want_pts)


def test_dag_z_shape_ranking(dag_design):
    # The ranking sums three prefix differences of the y view into an int,
    # truncating after each addition (:1893-1896).  Along y (pathIndex 1) the
    # three terms for displace id d at row yd are ysum(26,yd) - ysum(20,yd),
    # ysum(20,yd) - ysum(20,30) and ysum(26,34) - ysum(26,yd).  With
    # ysum(20, 31..33) = 0.5, 1.5, 2.5, ysum(26, 31..33) = 1, 1, 3 and 0
    # elsewhere:
    #   d=1 (yd=33): trunc(0.5)=0, trunc(0+2.5)=2, trunc(2-3)=-1   -> -1
    #   d=2 (yd=32): trunc(-0.5)=0, trunc(0+1.5)=1, trunc(1-1)=0   ->  0
    #   d=3 (yd=31): trunc(0.5)=0, trunc(0+0.5)=0, trunc(0-1)=-1   -> -1
    # Sorted (cost, d): (-1,1), (-1,3), (0,2); select 2 -> rows 33 and 31.
    # Along x every term is 0, so x = 25, 24 stay.  The view shows no
    # congestion on the tree edge's own columns 20 and 26 between rows 30 and
    # 34 (ysum(20,34) - ysum(20,30) = 0), so no region forms.
    d = dag_design
    s = _sub_of(d, (20, 30))
    ys = np.zeros(X100 * Y100, np.float32)
    for y, v in zip((31, 32, 33), (0.5, 1.5, 2.5)):
        ys[20 * Y100 + y] = v
    for y, v in zip((31, 32, 33), (1.0, 1.0, 3.0)):
        ys[26 * Y100 + y] = v
    nodes, edges, _ = _dag(d, s, np.zeros(X100 * Y100, np.float32), ys, mode=0)
    L = 9
    zy = sorted({n[2] for n in nodes if n[0] == L and n[1] == 20 and n[2] not in (30, 34)})
    assert zy == [31, 33]
    zx = sorted({n[1] for n in nodes if n[0] == L and n[2] == 30 and n[1] not in (20, 26)})
    assert zx == [24, 25]


def test_dag_subdivision(dag_design):
    # connect_node (:1662-1722): Manhattan length 60 (< 150) gets one midpoint
    # ((px+cx)/2, (py+cy)/2); 98 gets one; 60 along x from 0 to 60 gets one at
    # x = 30.  A straight two-pin net with an empty view is one tree edge.
    d = dag_design
    z = np.zeros(X100 * Y100, np.float32)
    L = 9
    for a, b, mid in (((10, 75), (70, 75), [(40, 75)]), ((0, 80), (60, 80), [(30, 80)]),
                      ((1, 95), (99, 95), [(50, 95)])):
        s = _sub_of(d, a)
        r, c = _root_first(d, s, a, b)
        nodes, edges, _ = _dag(d, s, z, z)
        lr, lc = _layer_of_pin(d, s, *r), _layer_of_pin(d, s, *c)
        m = (L,) + mid[0]
        assert edges == sorted([((lr,) + r, m, 0), (m, (lc,) + c, 0)]), (a, b)


def test_dag_subdivision_long_edges():
    # Two midpoints from Manhattan length 150 (:1700-1719):
    # along x at y: (px*2/3 + cx/3, y) and (px/3 + cx*2/3, y) with integer
    # division of each product; (10,40) -> (170,40): 20/3... computed as
    # 10*2/3 = 6, 170/3 = 56 -> 62; 10/3 = 3, 170*2/3 = 113 -> 116.
    import tempfile
    global X100, Y100
    with tempfile.TemporaryDirectory() as tmp:
        X100 = Y100 = 200
        try:
            d = _write_design(tmp, [[(10, 40), (170, 40)], [(50, 3), (50, 190)]])
            z = np.zeros(200 * 200, np.float32)
            L = 9
            s = _sub_of(d, (10, 40))
            r, c = _root_first(d, s, (10, 40), (170, 40))
            lr, lc = _layer_of_pin(d, s, *r), _layer_of_pin(d, s, *c)
            _, edges, _ = _dag(d, s, z, z)
            if r == (10, 40):
                m1, m2 = (L, 62, 40), (L, 116, 40)     # as above
            else:
                # par (170,40), cur (10,40): 170*2/3 = 113, 10/3 = 3 -> 116;
                # 170/3 = 56, 10*2/3 = 6 -> 62
                m1, m2 = (L, 116, 40), (L, 62, 40)
            assert edges == sorted([((lr,) + r, m1, 0), (m1, m2, 0), (m2, (lc,) + c, 0)])
            s = _sub_of(d, (50, 3))
            r, c = _root_first(d, s, (50, 3), (50, 190))
            lr, lc = _layer_of_pin(d, s, *r), _layer_of_pin(d, s, *c)
            _, edges, _ = _dag(d, s, z, z)
            if r == (50, 3):
                # 3*2/3 = 2, 190/3 = 63 -> 65; 3/3 = 1, 190*2/3 = 126 -> 127
                m1, m2 = (L, 50, 65), (L, 50, 127)
            else:
                # 190*2/3 = 126, 3/3 = 1 -> 127; 190/3 = 63, 3*2/3 = 2 -> 65
                m1, m2 = (L, 50, 127), (L, 50, 65)
            assert edges == sorted([((lr,) + r, m1, 0), (m1, m2, 0), (m2, (lc,) + c, 0)])
        finally:
            X100 = Y100 = 100


def test_dag_mirror_trunks(dag_design):
    # Net (20,50)-(40,50), root R = (20,50) (tree node 0), child C = (40,50).
    # xsum(40,50) - xsum(20,50) > 0 marks the edge congested along x
    # (:1540-1547): congestion 1, dir 0; x is the root, so C's region is C
    # (:1570-1580).  Stems of C's region: C's own y (50, :1503-1508) and the
    # root's y appended across the region boundary (:1613-1621) -> [50, 50];
    # range (20, 40).
    # get_mirror_places (:1733-1771): trunk_len 20, max_displace =
    # int(0.1f * 20) = 2, original y 50 with displacement 0.  Down: y=49
    # costs 2 (<= 2), y=48 costs 4 -> low 49.  Up, testing high-1 each time:
    # high 50 -> 51 (tests 49: 2), 52 (tests 50: 0), 53 (tests 51: 2), stops
    # at testing 52 (4) -> high 53.  step 1 ((1 + 3) / 2 < 6 tracks).  Places
    # 49, 51, 52, 53.
    # Each place p (:1823-1845): C is its region's node and not the root, and
    # (40,p) is diagonal to R, so corners (20,p) and (40,50) (layer L) hang
    # from R in slot 0 and both lead to the mirror (40,p).  C is a pin, so a
    # duplicate of C hangs from each mirror in slot deg(C)-1 = 0
    # (:1949-1966).  R -> C stays as the straight edge.
    d = dag_design
    s = _sub_of(d, (20, 50))
    r, c = _root_first(d, s, (20, 50), (40, 50))
    assert (r, c) == ((20, 50), (40, 50))
    lr, lc = _layer_of_pin(d, s, 20, 50), _layer_of_pin(d, s, 40, 50)
    xs = np.zeros(X100 * Y100, np.float32)
    for x in range(30, X100):
        xs[x * Y100 + 50] = 1.0
    nodes, edges, pts = _dag(d, s, xs, np.zeros(X100 * Y100, np.float32), mode=0)
    L = 9
    R, C = (lr, 20, 50), (lc, 40, 50)
    want = [(R, C, 0)]
    for p in (49, 51, 52, 53):
        want += [(R, (L, 20, p), 0), (R, (L, 40, 50), 0), ((L, 20, p), (L, 40, p), 0),
                 ((L, 40, 50), (L, 40, p), 0), ((L, 40, p), C, 0)]
    assert edges == sorted(want)
    assert len(nodes) == 2 + 4 * 4
    assert pts == sorted({20 * Y100 + 50, 40 * Y100 + 50} | {20 * Y100 + p for p in (49, 51, 52, 53)} |
                         {40 * Y100 + p for p in (49, 51, 52, 53)})


@pytest.mark.parametrize("mode,places", [
    # trunk 40: max_displace int(0.1f * 40) = 4 in modes 0/1 (:1457-1463);
    # stems [60, 60] cost 2|p - 60|: low 58 (57 costs 6); high tests high-1:
    # 59:2, 60:0, 61:2, 62:4, 63:6 -> high 64; step 1 ((2 + 4) / 2 < 6).
    (0, [58, 59, 61, 62, 63, 64]),
    # mode 2: ratio 0.15, 9 tracks: max_displace int(0.15f * 40) = 6; low 57
    # (56 costs 8); high tests 59..63 (<= 6), then 64 (8) -> 65; step 1
    # ((3 + 5) / 2 < 9).
    (2, [57, 58, 59, 61, 62, 63, 64, 65]),
])
def test_dag_mirror_places_by_mode(dag_design, mode, places):
    d = dag_design
    s = _sub_of(d, (20, 60))
    r, c = _root_first(d, s, (20, 60), (60, 60))
    assert (r, c) == ((20, 60), (60, 60))
    xs = np.zeros(X100 * Y100, np.float32)
    for x in range(30, X100):
        xs[x * Y100 + 60] = 1.0
    nodes, edges, _ = _dag(d, s, xs, np.zeros(X100 * Y100, np.float32), mode=mode)
    L = 9
    mirrors = sorted(n[2] for n in nodes if n[0] == L and n[1] == 60 and n[2] != 60)
    assert mirrors == places
    assert len(nodes) == 2 + 4 * len(places)
