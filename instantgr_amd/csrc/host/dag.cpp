// Detour DAG construction for rip-up and reroute.
//
// Restates net::generate_detours_reconstruct (src/database_cuda.hpp:1429-2086)
// and select_root_net (:826-853).  From a subnet's Steiner tree and the 2-D
// congestion prefix views it builds a DAG whose source is the tree root:
// congested regions of collinear tree edges get mirror trunks shifted across
// the congestion (with a duplicated pin node under each mirror), every
// L-shaped tree edge gets both L corners and up to 'select' Z shapes ranked by
// the overflow they cross, and long edges are subdivided (one midpoint from 50
// gcells, two from 150).  The nodes are then renumbered in a topological order
// (parents first) with their depth (longest path from the source).
//
// Deterministic semantics where the reference is not deterministic:
//   * cost_road entries the reference leaves unwritten (skipped Z shapes) are
//     left out of the ranking instead of being read uninitialised
//     (src/database_cuda.hpp:1877-1905);
//   * the remaining behaviour, including the displacement scan that tests
//     init_high - 1 (src/database_cuda.hpp:1747), is kept as written.
// Nodes that are neither pins nor tree Steiner points carry layer L, like the
// reference's MAX_LAYER-1 with its 9 routing layers.
//
// Each worker thread reuses one builder (no per-net allocations once warm)
// and writes a chunk of nets into a chunk buffer; the chunks are then copied
// into the final arrays in parallel.  Chunk buffers and final arrays keep
// their memory across builds, so later passes neither zero nor fault it in.
#include "dag.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace igr {
namespace {

struct Builder {
    const Grid& g;
    const float* xsum;
    const float* ysum;
    int mode;
    int X, Y, XYc, L;
    float ratio = 0.1f;
    int num_tracks = 6;
    // tree (CSR borrowed from the subnets)
    int tn = 0;
    const int* pos = nullptr;
    const int64_t* aoff = nullptr;   // relative offsets, tn+1
    const int* tadj = nullptr;
    int root = 0;
    // congestion regions
    std::vector<int> region[2];
    std::vector<std::pair<int, int>> ranges[2];
    std::vector<std::vector<int>> stems[2];
    // tree nodes' layers and coordinates, decoded once per net
    struct XY { int x, y; };
    std::vector<int> tlay;
    std::vector<XY> txy;
    // DAG under construction: per node its encoded cell, its coordinates (so
    // edges never divide the cell index again), its child edges so far and
    // whether its cell is already in points
    struct DNode { int cell, x, y, cpre, listed; };
    std::vector<DNode> dn;
    std::vector<std::pair<int, int>> edges;   // (child, parent*10 + slot) in creation order
    std::vector<int> points;
    // per recursion depth buffers: mirrors created at a node and their places
    std::vector<std::vector<int>> mir_buf[2], place_buf[2];
    // scratch
    std::vector<std::pair<float, int>> road;
    std::vector<int> q, visited, pcnt, poff, plist, visit, seq, st, hset;

    // The scheduler points of the net without duplicates, in order of first
    // occurrence (the scheduler does not depend on their order): an
    // open-addressing set sized to twice the list.
    void dedup_points() {
        size_t cap = 16;
        while (cap < 2 * points.size()) cap <<= 1;
        hset.assign(cap, -1);
        size_t w = 0;
        for (size_t k = 0; k < points.size(); k++) {
            const int p = points[k];
            size_t h = ((uint32_t)p * 2654435761u) & (cap - 1);
            while (hset[h] != -1 && hset[h] != p) h = (h + 1) & (cap - 1);
            if (hset[h] == p) continue;
            hset[h] = p;
            points[w++] = p;
        }
        points.resize(w);
    }

    Builder(const Grid& g_, const float* xs, const float* ys, int mode_)
        : g(g_), xsum(xs), ysum(ys), mode(mode_), X(g_.X), Y(g_.Y), XYc(g_.X * g_.Y), L(g_.L) {
        if (mode == 0 || mode == 1) { ratio = 0.1f; num_tracks = 6; }
        else { ratio = 0.15f; num_tracks = 9; }
    }

    int deg(int x) const { return (int)(aoff[x + 1] - aoff[x]); }

    // src/database_cuda.hpp:826-853: breadth-first from every leaf; the answer
    // is the node of the last queue entry popped.
    int select_root() {
        q.clear();
        visited.assign(tn, 0);
        for (int i = 0; i < tn; i++)
            if (deg(i) == 1) { q.push_back(i); q.push_back(-1); }
        int select = -1;
        for (size_t h = 0; h < q.size(); h += 2) {
            int node = q[h], fa = q[h + 1];
            select = node;
            if (visited[node]) continue;
            visited[node] = 1;
            for (int64_t a = aoff[node]; a < aoff[node + 1]; a++)
                if (tadj[a] != fa) { q.push_back(tadj[a]); q.push_back(node); }
        }
        return select < 0 ? 0 : select;
    }

    int getRegion(int x, int dir) {
        if (x == -1) return -1;
        if (region[dir][x] == -1) return -1;
        if (region[dir][x] != x) {
            int ans = getRegion(region[dir][x], dir);
            region[dir][x] = ans;
            return ans;
        }
        return region[dir][x];
    }

    float ysumd(int x0, int y0, int x1, int y1) const {
        return ysum[(size_t)std::max(x0, x1) * Y + std::max(y0, y1)] - ysum[(size_t)std::min(x0, x1) * Y + std::min(y0, y1)];
    }

    void markCongestion(int x, int par) {
        const int curx = txy[x].x, cury = txy[x].y;
        for (int64_t a = aoff[x]; a < aoff[x + 1]; a++) {
            int e = tadj[a];
            if (e == par) continue;
            const int ex = txy[e].x, ey = txy[e].y;
            int dir = -1, congestion = -1;
            if (ex == curx) {
                if ((ysum[(size_t)curx * Y + std::max(ey, cury)] - ysum[(size_t)curx * Y + std::min(ey, cury)]) > 0) {
                    dir = 1; congestion = 1;
                }
            } else if (ey == cury) {
                if ((xsum[(size_t)std::max(ex, curx) * Y + cury] - xsum[(size_t)std::min(ex, curx) * Y + cury]) > 0) {
                    dir = 0; congestion = 1;
                }
            } else {
                bool cy = (ysum[(size_t)curx * Y + std::max(ey, cury)] - ysum[(size_t)curx * Y + std::min(ey, cury)]) > 0 ||
                          (ysum[(size_t)ex * Y + std::max(ey, cury)] - ysum[(size_t)ex * Y + std::min(ey, cury)]) > 0;
                if (cy) { dir = 1; congestion = 2; }
                if (congestion == -1) {
                    bool cx = (xsum[(size_t)std::max(ex, curx) * Y + cury] - xsum[(size_t)std::min(ex, curx) * Y + cury]) > 0 ||
                              (xsum[(size_t)std::max(ex, curx) * Y + ey] - xsum[(size_t)std::min(ex, curx) * Y + ey]) > 0;
                    if (cx) { dir = 0; congestion = 2; }
                }
            }
            if (congestion == 1) {
                int target;
                if (x != root) {
                    if (region[dir][x] == -1) region[dir][x] = x;
                    target = getRegion(x, dir);
                } else {
                    if (region[dir][e] == -1) region[dir][e] = e;
                    target = getRegion(e, dir);
                }
                region[dir][e] = target;
                if (x != root && target != e) {
                    auto& dst = stems[dir][target];
                    const auto& src = stems[dir][e];
                    dst.insert(dst.end(), src.begin(), src.end());
                }
                ranges[dir][target].first = std::min(ranges[dir][target].first, ranges[dir][e].first);
                ranges[dir][target].second = std::max(ranges[dir][target].second, ranges[dir][e].second);
            } else if (congestion == 2) {
                if (x != root && region[dir][x] == -1) region[dir][x] = x;
                if (region[dir][e] == -1) region[dir][e] = e;
            }
            markCongestion(e, x);
        }
        for (int dir = 0; dir < 2; dir++) {
            int xr = getRegion(x, dir);
            for (int64_t a = aoff[x]; a < aoff[x + 1]; a++) {
                int e = tadj[a];
                if (e == par) continue;
                const int ex = txy[e].x, ey = txy[e].y;
                int er = getRegion(e, dir);
                if (xr != er) {
                    if (xr >= 0) {
                        stems[dir][xr].push_back(dir ? ex : ey);
                        ranges[dir][xr].first = std::min(ranges[dir][xr].first, dir ? ey : ex);
                        ranges[dir][xr].second = std::max(ranges[dir][xr].second, dir ? ey : ex);
                    }
                    if (er >= 0) {
                        stems[dir][er].push_back(dir ? curx : cury);
                        ranges[dir][er].first = std::min(ranges[dir][er].first, dir ? cury : curx);
                        ranges[dir][er].second = std::max(ranges[dir][er].second, dir ? cury : curx);
                    }
                }
            }
        }
    }

    int create_node(int l, int x, int y) {
        int idx = (int)dn.size();
        dn.push_back({l * XYc + x * Y + y, x, y, 0, 0});
        return idx;
    }

    void connect_pre(int par_idx, int cur_idx, int child_id) {
        // Edges carry parent*10 + slot (src/database_cuda.hpp:1655).
        if (child_id < 0 || child_id >= 10) throw std::runtime_error("detour DAG node with more than 10 child slots");
        // A node's cell is listed at its first edge only: a repeat would be
        // dropped by dedup_points anyway, behind the first occurrence.
        DNode& c = dn[cur_idx];
        DNode& p = dn[par_idx];
        if (!c.listed) {
            c.listed = 1;
            points.push_back(c.x * Y + c.y);
        }
        if (!p.listed) {
            p.listed = 1;
            points.push_back(p.x * Y + p.y);
        }
        edges.emplace_back(cur_idx, par_idx * 10 + child_id);
        p.cpre++;
    }

    // src/database_cuda.hpp:1662-1722 (segments are not rebuilt in this pass).
    void connect(int par_idx, int cur_idx, int child_id) {
        const int par_x = dn[par_idx].x, par_y = dn[par_idx].y;
        const int cur_x = dn[cur_idx].x, cur_y = dn[cur_idx].y;
        int dist = std::abs(par_x - cur_x) + std::abs(par_y - cur_y);
        if (dist < 50) {
            connect_pre(par_idx, cur_idx, child_id);
            return;
        }
        if (dist < 150) {
            int mid = create_node(L, (par_x + cur_x) / 2, (par_y + cur_y) / 2);
            connect_pre(par_idx, mid, child_id);
            connect_pre(mid, cur_idx, 0);
            return;
        }
        int m1, m2;
        if (par_y == cur_y) {
            m1 = create_node(L, par_x * 2 / 3 + cur_x / 3, par_y);
            m2 = create_node(L, par_x * 1 / 3 + cur_x * 2 / 3, par_y);
        } else {
            m1 = create_node(L, par_x, par_y * 2 / 3 + cur_y / 3);
            m2 = create_node(L, par_x, par_y * 1 / 3 + cur_y * 2 / 3);
        }
        connect_pre(par_idx, m1, child_id);
        connect_pre(m1, m2, 0);
        connect_pre(m2, cur_idx, 0);
    }

    int calc_displace(int qpos, int dir, int reg) const {
        int ans = 0;
        for (int p : stems[dir][reg]) ans += std::abs(p - qpos);
        return ans;
    }

    // src/database_cuda.hpp:1733-1771.
    void mirror_places(int node, int dir, std::vector<int>& shifts) {
        shifts.clear();
        int reg = region[dir][node];
        const int curx = txy[node].x, cury = txy[node].y;
        int trunk_len = ranges[dir][reg].second - ranges[dir][reg].first;
        int max_displace = (int)(ratio * (float)trunk_len);
        int orig = dir ? curx : cury;
        int orig_disp = calc_displace(orig, dir, reg);
        int lo = orig, hi = orig;
        int bound = dir ? X : Y;
        while (lo - 1 >= 0 && calc_displace(lo - 1, dir, reg) - orig_disp <= max_displace) lo--;
        while (hi + 1 < bound && calc_displace(hi - 1, dir, reg) - orig_disp <= max_displace) hi++;
        int step = 1;
        while ((orig - lo) / (step + 1) + (hi - orig) / (step + 1) >= num_tracks) step++;
        lo = orig - (orig - lo) / step * step;
        hi = orig + (hi - orig) / step * step;
        for (double p = lo; p <= hi; p += step) {
            int shift = (int)(p - orig);
            if (shift == 0) continue;
            shifts.push_back((int)p);
        }
    }

    void ensure_depth(int depth) {
        for (int d = 0; d < 2; d++)
            while ((int)mir_buf[d].size() <= depth) {
                mir_buf[d].emplace_back();
                place_buf[d].emplace_back();
            }
    }

    // mirrors / old places of the parent are mir_buf[*][depth-1] and
    // place_buf[*][depth-1] (absent for the root).
    void dfs(int x, int par, int par_node, int child_idx, int depth, int level) {
        ensure_depth(level);
        const int curl = tlay[x], curx = txy[x].x, cury = txy[x].y;
        int node_idx = create_node(curl, curx, cury);
        std::vector<int>* new_mirrors[2] = {&mir_buf[0][level], &mir_buf[1][level]};
        std::vector<int>* places[2] = {&place_buf[0][level], &place_buf[1][level]};
        for (int d = 0; d < 2; d++) { new_mirrors[d]->clear(); places[d]->clear(); }
        const bool has_parent_data = level > 0;
        if (has_parent_data) {
            for (int dir = 0; dir < 2; dir++) {
                const std::vector<int>& mirrors = mir_buf[dir][level - 1];
                const std::vector<int>& old_places = place_buf[dir][level - 1];
                int rid = getRegion(region[dir][x], dir);
                if (rid < 0) continue;
                if (getRegion(region[dir][par], dir) == par) depth += 2;
                if (x == rid) mirror_places(x, dir, *places[dir]);
                else *places[dir] = old_places;
                if (rid != x && places[dir]->size() != mirrors.size())
                    throw std::runtime_error("detour construction: mirror count mismatch");
                for (size_t m = 0; m < places[dir]->size(); m++) {
                    int pl = (*places[dir])[m];
                    int nx = dir ? pl : curx;
                    int ny = dir ? cury : pl;
                    int nm;
                    if (rid != x) {
                        nm = create_node(L, nx, ny);
                        connect(mirrors[m], nm, child_idx);
                    } else if (x != root) {
                        const int px = dn[par_node].x, py = dn[par_node].y;
                        if (nx != px && ny != py) {
                            int t[2];
                            for (int pi = 0; pi <= 1; pi++) {
                                t[pi] = create_node(L, pi ? nx : px, pi ? py : ny);
                                connect(par_node, t[pi], child_idx);
                            }
                            nm = create_node(L, nx, ny);
                            for (int pi = 0; pi <= 1; pi++) connect(t[pi], nm, 0);
                        } else {
                            nm = create_node(L, nx, ny);
                            connect(par_node, nm, child_idx);
                        }
                    } else {
                        nm = create_node(L, nx, ny);
                    }
                    new_mirrors[dir]->push_back(nm);
                }
            }
        }
        if (par_node != -1) {
            const int px = dn[par_node].x, py = dn[par_node].y;
            if (px != curx && py != cury) {
                for (int pi = 0; pi <= 1; pi++) {
                    int pre = create_node(L, pi ? curx : px, pi ? py : cury);
                    connect(par_node, pre, child_idx);
                    connect(pre, node_idx, 0);
                }
                for (int pi = 0; pi <= 1; pi++) {
                    int len = pi ? std::abs(py - cury) : std::abs(px - curx);
                    int z_shapes = 10, select = 10;
                    if (mode == 0 || mode == 1) { z_shapes = 7; select = 2; }
                    int mz = std::min(z_shapes, len);
                    auto mids = [&](int d, int& x1, int& y1, int& x2, int& y2) {
                        x1 = pi ? px : (px * d + curx * (mz - d)) / mz;
                        y1 = pi ? (py * d + cury * (mz - d)) / mz : py;
                        x2 = pi ? curx : (px * d + curx * (mz - d)) / mz;
                        y2 = pi ? (py * d + cury * (mz - d)) / mz : cury;
                        if (x1 < 0 || y1 < 0 || x1 >= X || y1 >= Y) return false;
                        if (x1 == px && y1 == py) return false;
                        if (x1 == curx && y1 == cury) return false;
                        if (x2 < 0 || y2 < 0 || x2 >= X || y2 >= Y) return false;
                        if (x2 == curx && y2 == cury) return false;
                        if (x2 == px && y2 == py) return false;
                        return true;
                    };
                    road.clear();
                    for (int d = 1; d < mz; d++) {
                        int x1, y1, x2, y2;
                        if (!mids(d, x1, y1, x2, y2)) continue;
                        int of = 0;
                        of += ysumd(x1, y1, x2, y2);
                        of += ysumd(x1, y1, px, py);
                        of += ysumd(curx, cury, x2, y2);
                        road.emplace_back((float)of, d);
                    }
                    std::sort(road.begin(), road.end());
                    for (int ii = 0; ii < select && ii < (int)road.size(); ii++) {
                        int d = road[ii].second, x1, y1, x2, y2;
                        if (!mids(d, x1, y1, x2, y2)) continue;
                        int p1 = create_node(L, x1, y1);
                        connect(par_node, p1, child_idx);
                        points.push_back((x1 + x2) / 2 * Y + (y1 + y2) / 2);
                        int p2 = create_node(L, x2, y2);
                        connect(p1, p2, 0);
                        connect(p2, node_idx, 0);
                    }
                }
            }
            for (int dir = 0; dir < 2; dir++) {
                int rid = getRegion(region[dir][x], dir);
                int prid = getRegion(region[dir][par], dir);
                if (prid >= 0 && rid != prid) {
                    const std::vector<int>& mirrors = mir_buf[dir][level - 1];
                    for (int pm : mirrors) {
                        const int qx = dn[pm].x, qy = dn[pm].y;
                        if (qx != curx && qy != cury) {
                            for (int pi = 0; pi <= 1; pi++) {
                                int ins = create_node(L, pi ? curx : qx, pi ? qy : cury);
                                connect(pm, ins, child_idx);
                                connect(ins, node_idx, 0);
                            }
                        } else {
                            connect(pm, node_idx, child_idx);
                        }
                    }
                }
            }
            if (px == curx || py == cury) connect(par_node, node_idx, child_idx);
        }
        for (int dir = 0; dir < 2; dir++) {
            int rid = getRegion(x, dir);
            if (rid < 0) continue;
            bool is_pin = curl < L;
            if (x != root && is_pin)
                for (int m : *new_mirrors[dir]) {
                    int dup = create_node(curl, curx, cury);
                    connect(m, dup, deg(x) - 1);
                }
        }
        int idx = 0;
        for (int64_t a = aoff[x]; a < aoff[x + 1]; a++) {
            int e = tadj[a];
            if (e == par) continue;
            dfs(e, x, node_idx, idx++, depth + 1, level + 1);
        }
    }

    // The lines of v's next 'bytes' bytes within its capacity, fetched for
    // writing.
    static void prefetch_tail(const RawVec<int>& v, size_t bytes) {
        const char* p = (const char*)(v.data() + v.size());
        const char* e = (const char*)(v.data() + v.capacity());
        for (size_t o = 0; o < bytes && p + o < e; o += 64) __builtin_prefetch(p + o, 1);
    }

    // Build one net's DAG and append it to the chunk output.
    void build(const int* tnode, int tn_, const int64_t* adj_off_abs, const int* tadj_, DagChunk& out,
               int part = DAG_ALL) {
        // The chunk arrays were last written a whole build ago: their next
        // lines (a typical net's share) load while the DAG is being built.
        if (part != DAG_POINTS) {
            for (auto* v : {&out.node, &out.child_num, &out.depth, &out.npar}) prefetch_tail(*v, 384);
            for (auto* v : {&out.par, &out.slot}) prefetch_tail(*v, 512);
        }
        if (part != DAG_STRUCTURE) prefetch_tail(out.points, 256);
        tn = tn_;
        pos = tnode;
        tadj = tadj_;
        rel.assign(adj_off_abs, adj_off_abs + tn + 1);
        for (auto& v : rel) v -= adj_off_abs[0];
        aoff = rel.data();
        tlay.resize(tn);
        txy.resize(tn);
        for (int x = 0; x < tn; x++) {
            tlay[x] = pos[x] / XYc;
            txy[x] = {pos[x] / Y % X, pos[x] % Y};
        }
        root = select_root();
        for (int d = 0; d < 2; d++) {
            region[d].assign(tn, -1);
            ranges[d].resize(tn);
            if ((int)stems[d].size() < tn) stems[d].resize(tn);
            for (int x = 0; x < tn; x++) stems[d][x].clear();
        }
        for (int x = 0; x < tn; x++) {
            const int curl = tlay[x], curx = txy[x].x, cury = txy[x].y;
            for (int dir = 0; dir < 2; dir++) {
                if (curl < L) stems[dir][x].push_back(dir ? curx : cury);
                int t = dir ? cury : curx;
                ranges[dir][x] = {t, t};
            }
        }
        markCongestion(root, -1);
        dn.clear();
        edges.clear();
        points.clear();
        dfs(root, -1, -1, -1, 0, 0);
        if (part != DAG_STRUCTURE) {
            dedup_points();
            out.points.insert(out.points.end(), points.begin(), points.end());
            out.points_per_net.push_back((int)points.size());
            if (part == DAG_POINTS) return;
        }

        // Parent lists in creation order (CSR by child).
        const int n = (int)dn.size();
        poff.assign(n + 1, 0);
        for (auto& e : edges) poff[e.first + 1]++;
        for (int i = 0; i < n; i++) poff[i + 1] += poff[i];
        plist.resize(edges.size());
        pcnt.assign(n, 0);
        for (auto& e : edges) plist[poff[e.first] + pcnt[e.first]++] = e.second;

        // Topological renumbering (src/database_cuda.hpp:2016-2080): a node is
        // pushed once all its child edges are visited; popping the stack gives
        // parents before children.
        // Each node enters the order once, so the order and the walk's stack
        // (node, next parent edge) each hold at most n entries.
        visit.assign(n, 0);
        st.resize(n);
        q.resize(2 * (size_t)n);
        int* order = st.data();
        int* work = q.data();
        const DNode* dnp = dn.data();
        int* vis = visit.data();
        int ns = 0;
        for (int i = 0; i < n; i++) {
            if (dnp[i].cpre != 0) continue;
            // iterative form of the recursive determinSequence
            int top = 0;
            work[0] = i;
            work[1] = poff[i];
            order[ns++] = i;
            while (top >= 0) {
                const int idx = work[2 * top];
                const int k = work[2 * top + 1];
                if (k == poff[idx + 1]) {
                    top--;
                    continue;
                }
                work[2 * top + 1] = k + 1;
                const int p = plist[k] / 10;
                if (++vis[p] == dnp[p].cpre) {
                    order[ns++] = p;
                    top++;
                    work[2 * top] = p;
                    work[2 * top + 1] = poff[p];
                }
            }
        }
        st.resize(ns);
        // Sized once and written through pointers; depth, npar and node are
        // written for every node, child_num accumulates from zero.
        seq.assign(n, 0);
        const size_t nbase = out.node.size(), ebase = out.par.size();
        out.node.resize(nbase + n);
        out.child_num.resize(nbase + n, 0);
        out.depth.resize(nbase + n);
        out.npar.resize(nbase + n);
        out.par.resize(ebase + edges.size());
        out.slot.resize(ebase + edges.size());
        int* o_node = out.node.data() + nbase;
        int* o_cn = out.child_num.data() + nbase;
        int* o_dep = out.depth.data() + nbase;
        int* o_np = out.npar.data() + nbase;
        int* o_par = out.par.data() + ebase;
        int* o_slot = out.slot.data() + ebase;
        int copy = 0, ne = 0;
        for (auto it = st.rbegin(); it != st.rend(); ++it) {
            const int idx = *it;
            int dep = 0;
            for (int k = poff[idx]; k < poff[idx + 1]; k++) {
                const int p = plist[k] / 10, ci = plist[k] % 10;
                const int sp = seq[p];
                o_cn[sp] = std::max(o_cn[sp], ci + 1);
                o_par[ne] = sp;
                o_slot[ne++] = ci;
                dep = std::max(dep, o_dep[sp] + 1);
            }
            o_dep[copy] = dep;
            o_np[copy] = poff[idx + 1] - poff[idx];
            seq[idx] = copy;
            o_node[copy++] = dn[idx].cell;
        }
        if (copy != n) throw std::runtime_error("detour DAG has a cycle");
        out.nodes_per_net.push_back(n);
    }

    std::vector<int64_t> rel;
};

} // namespace

void build_dag(const Grid& g, const float* xsum, const float* ysum, int mode, const int* tnode, int tn,
               const int64_t* adj_off, const int* tadj, DagNet& out) {
    Builder b(g, xsum, ysum, mode);
    DagChunk c;
    b.build(tnode, tn, adj_off, tadj, c);
    out.node.assign(c.node.begin(), c.node.end());
    out.child_num.assign(c.child_num.begin(), c.child_num.end());
    out.depth.assign(c.depth.begin(), c.depth.end());
    out.par.assign(c.par.begin(), c.par.end());
    out.slot.assign(c.slot.begin(), c.slot.end());
    out.points.assign(c.points.begin(), c.points.end());
    out.par_off.assign(c.node.size() + 1, 0);
    for (size_t i = 0; i < c.node.size(); i++) out.par_off[i + 1] = out.par_off[i] + c.npar[i];
}

void build_dags(const Grid& g, const Subnets& s, const int* nets, int64_t n, const float* xsum, const float* ysum,
                int mode, int threads, DagSet& out, int part) {
    if (part != DAG_ALL && part != DAG_POINTS && part != DAG_STRUCTURE) throw std::invalid_argument("unknown DAG part");
    const bool want_pts = part != DAG_STRUCTURE, want_dag = part != DAG_POINTS;
    if (part == DAG_STRUCTURE && (out.n != n || (int64_t)out.pt_off.size() != n + 1))
        throw std::invalid_argument("a structure build follows a points build of the same nets");
    if (threads < 1) threads = 1;
    const int64_t CH = 512;
    const int64_t nch = (n + CH - 1) / CH;
    std::vector<DagChunk>& chunks = out.work;
    if ((int64_t)chunks.size() < nch) chunks.resize(nch);
    std::atomic<int64_t> next(0);
    std::vector<std::string> errs(threads);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t]() {
            try {
                Builder b(g, xsum, ysum, mode);
                for (;;) {
                    int64_t c = next.fetch_add(1);
                    if (c >= nch) break;
                    DagChunk& ck = chunks[c];
                    for (auto* v : {&ck.node, &ck.child_num, &ck.depth, &ck.npar, &ck.par, &ck.slot, &ck.points,
                                    &ck.nodes_per_net, &ck.points_per_net})
                        v->clear();
                    if (want_dag) {
                        for (auto* v : {&ck.node, &ck.child_num, &ck.depth, &ck.npar}) v->reserve(CH * 128);
                        for (auto* v : {&ck.par, &ck.slot}) v->reserve(CH * 160);
                    }
                    if (want_pts) ck.points.reserve(CH * 96);
                    const int64_t i1 = std::min(n, (c + 1) * CH);
                    for (int64_t i = c * CH; i < i1; i++) {
                        // The net list is not in subnet order: fetch the
                        // trees of the next nets ahead, one level of the
                        // offset chain per step.
                        if (i + 3 < i1) __builtin_prefetch(&s.node_off[nets[i + 3]]);
                        if (i + 2 < i1) {
                            const int64_t bn = s.node_off[nets[i + 2]];
                            __builtin_prefetch(&s.node[bn]);
                            __builtin_prefetch(&s.adj_off[bn]);
                        }
                        if (i + 1 < i1) __builtin_prefetch(&s.adj[s.adj_off[s.node_off[nets[i + 1]]]]);
                        int net = nets[i];
                        int64_t b0 = s.node_off[net];
                        int tn = (int)(s.node_off[net + 1] - b0);
                        b.build(&s.node[b0], tn, &s.adj_off[b0], &s.adj[s.adj_off[b0]], ck, part);
                    }
                }
            } catch (const std::exception& e) {
                errs[t] = e.what();
            }
        });
    for (auto& th : pool) th.join();
    for (auto& e : errs)
        if (!e.empty()) throw std::runtime_error(e);
    // Offsets of every chunk in the final arrays.
    std::vector<int64_t> cn(nch + 1, 0), ce(nch + 1, 0), cp(nch + 1, 0);
    for (int64_t c = 0; c < nch; c++) {
        cn[c + 1] = cn[c] + (int64_t)chunks[c].node.size();
        ce[c + 1] = ce[c] + (int64_t)chunks[c].par.size();
        cp[c + 1] = cp[c] + (int64_t)chunks[c].points.size();
    }
    out.n = n;
    // Sized with a quarter to spare when they grow, so the next pass (a few
    // percent larger or smaller) reuses the pages instead of faulting in a
    // fresh multi-GB allocation.
    auto size_to = [](auto& v, int64_t m) {
        if ((int64_t)v.capacity() < m) v.reserve((size_t)(m + m / 4));
        v.resize((size_t)m);
    };
    if (want_dag) {
        size_to(out.node_off, n + 1);
        size_to(out.node, cn[nch]);
        size_to(out.child_num, cn[nch]);
        size_to(out.depth, cn[nch]);
        size_to(out.par_off, cn[nch] + 1);
        size_to(out.par, ce[nch]);
        size_to(out.slot, ce[nch]);
    }
    if (want_pts) {
        size_to(out.pt_off, n + 1);
        size_to(out.pt, cp[nch]);
    }
    next = 0;
    pool.clear();
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            for (;;) {
                int64_t c = next.fetch_add(1);
                if (c >= nch) break;
                const DagChunk& ck = chunks[c];
                if (want_dag) {
                    std::copy(ck.node.begin(), ck.node.end(), out.node.begin() + cn[c]);
                    std::copy(ck.child_num.begin(), ck.child_num.end(), out.child_num.begin() + cn[c]);
                    std::copy(ck.depth.begin(), ck.depth.end(), out.depth.begin() + cn[c]);
                    std::copy(ck.par.begin(), ck.par.end(), out.par.begin() + ce[c]);
                    std::copy(ck.slot.begin(), ck.slot.end(), out.slot.begin() + ce[c]);
                    int64_t e = ce[c];
                    for (size_t k = 0; k < ck.npar.size(); k++) {
                        out.par_off[cn[c] + k] = e;
                        e += ck.npar[k];
                    }
                    int64_t nn = cn[c];
                    for (size_t k = 0; k < ck.nodes_per_net.size(); k++) {
                        out.node_off[c * CH + (int64_t)k] = nn;
                        nn += ck.nodes_per_net[k];
                    }
                }
                if (want_pts) {
                    std::copy(ck.points.begin(), ck.points.end(), out.pt.begin() + cp[c]);
                    int64_t pp = cp[c];
                    for (size_t k = 0; k < ck.points_per_net.size(); k++) {
                        out.pt_off[c * CH + (int64_t)k] = pp;
                        pp += ck.points_per_net[k];
                    }
                }
            }
        });
    for (auto& th : pool) th.join();
    if (want_dag) {
        out.par_off[cn[nch]] = ce[nch];
        out.node_off[n] = cn[nch];
    }
    if (want_pts) out.pt_off[n] = cp[nch];
}

} // namespace igr
