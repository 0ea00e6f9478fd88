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
#include "dag.hpp"

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <queue>
#include <stack>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace igr {
namespace {

// src/database_cuda.hpp:826-853: breadth-first from every leaf; the answer is
// the node of the last queue entry popped.
int select_root(const std::vector<std::vector<int>>& adj) {
    const int n = (int)adj.size();
    std::queue<std::pair<int, int>> q;
    std::vector<char> visited(n, 0);
    int select = -1;
    for (int i = 0; i < n; i++)
        if (adj[i].size() == 1) q.push({i, -1});
    while (!q.empty()) {
        auto f = q.front();
        q.pop();
        select = f.first;
        if (visited[select]) continue;
        visited[select] = 1;
        for (int j : adj[select])
            if (j != f.second) q.push({j, select});
    }
    return select < 0 ? 0 : select;
}

struct Builder {
    const Grid& g;
    const float* xsum;
    const float* ysum;
    int mode;
    int X, Y, XYc, L;
    // tree
    std::vector<std::vector<int>> adj;
    std::vector<int> pos;              // encoded with layer (L = Steiner)
    int root = 0;
    // congestion regions
    std::vector<int> region[2];
    std::vector<std::pair<int, int>> ranges[2];
    std::vector<std::vector<int>> stems[2];
    float ratio = 0.1f;
    int num_tracks = 6;
    // DAG under construction
    std::vector<int> nodes;
    std::vector<std::vector<int>> parents;  // parent*10 + child slot
    std::vector<int> child_num_pre;
    std::vector<int> points;

    Builder(const Grid& g_, const float* xs, const float* ys, int mode_)
        : g(g_), xsum(xs), ysum(ys), mode(mode_), X(g_.X), Y(g_.Y), XYc(g_.X * g_.Y), L(g_.L) {
        if (mode == 0 || mode == 1) { ratio = 0.1f; num_tracks = 6; }
        else { ratio = 0.15f; num_tracks = 9; }
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

    float ysumd(int x0, int y0, int x1, int y1) const {  // ysum[max] - ysum[min] pattern
        return ysum[(size_t)std::max(x0, x1) * Y + std::max(y0, y1)] - ysum[(size_t)std::min(x0, x1) * Y + std::min(y0, y1)];
    }

    void markCongestion(int x, int par) {
        int cur = pos[x];
        int curx = cur / Y % X, cury = cur % Y;
        for (int e : adj[x]) {
            if (e == par) continue;
            int ex = pos[e] / Y % X, ey = pos[e] % Y;
            int dir = -1, congestion = -1;
            if (ex == curx) {
                bool cy = (ysum[(size_t)curx * Y + std::max(ey, cury)] - ysum[(size_t)curx * Y + std::min(ey, cury)]) > 0;
                if (cy) { dir = 1; congestion = 1; }
            } else if (ey == cury) {
                bool cx = (xsum[(size_t)std::max(ex, curx) * Y + cury] - xsum[(size_t)std::min(ex, curx) * Y + cury]) > 0;
                if (cx) { dir = 0; congestion = 1; }
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
                if (x != root)
                    for (int p : std::vector<int>(stems[dir][e])) stems[dir][target].push_back(p);
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
            for (int e : adj[x]) {
                if (e == par) continue;
                int ex = pos[e] / Y % X, ey = pos[e] % Y;
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
        int idx = (int)nodes.size();
        nodes.push_back(l * XYc + x * Y + y);
        parents.emplace_back();
        child_num_pre.push_back(0);
        return idx;
    }

    void connect_pre(int par_idx, int cur_idx, int child_id) {
        // Edges carry parent*10 + slot (src/database_cuda.hpp:1655).
        if (child_id < 0 || child_id >= 10) throw std::runtime_error("detour DAG node with more than 10 child slots");
        int c = nodes[cur_idx] % XYc, p = nodes[par_idx] % XYc;
        points.push_back(c);
        points.push_back(p);
        parents[cur_idx].push_back(par_idx * 10 + child_id);
        child_num_pre[par_idx]++;
    }

    // src/database_cuda.hpp:1662-1722 (segments are not rebuilt in this pass).
    void connect(int par_idx, int cur_idx, int child_id) {
        int par_x = nodes[par_idx] / Y % X, par_y = nodes[par_idx] % Y;
        int cur_x = nodes[cur_idx] / Y % X, cur_y = nodes[cur_idx] % Y;
        int dist = std::abs(par_x - cur_x) + std::abs(par_y - cur_y);
        if (dist < 50) {
            connect_pre(par_idx, cur_idx, child_id);
            return;
        }
        if (dist < 150) {
            int mx = (par_x + cur_x) / 2, my = (par_y + cur_y) / 2;
            int mid = create_node(L, mx, my);
            connect_pre(par_idx, mid, child_id);
            connect_pre(mid, cur_idx, 0);
            return;
        }
        if (par_y == cur_y) {
            int m1 = create_node(L, par_x * 2 / 3 + cur_x / 3, par_y);
            int m2 = create_node(L, par_x * 1 / 3 + cur_x * 2 / 3, par_y);
            connect_pre(par_idx, m1, child_id);
            connect_pre(m1, m2, 0);
            connect_pre(m2, cur_idx, 0);
        } else {
            int m1 = create_node(L, par_x, par_y * 2 / 3 + cur_y / 3);
            int m2 = create_node(L, par_x, par_y * 1 / 3 + cur_y * 2 / 3);
            connect_pre(par_idx, m1, child_id);
            connect_pre(m1, m2, 0);
            connect_pre(m2, cur_idx, 0);
        }
    }

    int calc_displace(int q, int dir, int reg) const {
        int ans = 0;
        for (int p : stems[dir][reg]) ans += std::abs(p - q);
        return ans;
    }

    // src/database_cuda.hpp:1733-1771.
    std::vector<int> mirror_places(int node, int dir) {
        int reg = region[dir][node];
        int curx = pos[node] / Y % X, cury = pos[node] % Y;
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
        std::vector<int> shifts;
        for (double p = lo; p <= hi; p += step) {
            int shift = (int)(p - orig);
            if (shift == 0) continue;
            shifts.push_back((int)p);
        }
        return shifts;
    }

    void dfs(int x, int par, int par_node, int child_idx, int depth, std::vector<std::vector<int>> mirrors,
             const std::vector<std::vector<int>>& old_places) {
        if (mirrors.empty()) mirrors.resize(2);
        int cur = pos[x];
        int curl = cur / XYc, curx = cur / Y % X, cury = cur % Y;
        int node_idx = create_node(curl, curx, cury);
        std::vector<std::vector<int>> new_mirrors(2), places(2);
        if (old_places.size() == 2) {
            for (int dir = 0; dir < 2; dir++) {
                int rid = getRegion(region[dir][x], dir);
                if (rid < 0) continue;
                if (getRegion(region[dir][par], dir) == par) depth += 2;
                if (x == rid) places[dir] = mirror_places(x, dir);
                else places[dir] = old_places[dir];
                if (rid != x && places[dir].size() != mirrors[dir].size())
                    throw std::runtime_error("detour construction: mirror count mismatch");
                for (size_t m = 0; m < places[dir].size(); m++) {
                    int nx = dir ? places[dir][m] : curx;
                    int ny = dir ? cury : places[dir][m];
                    int nm;
                    if (rid != x) {
                        nm = create_node(L, nx, ny);
                        connect(mirrors[dir][m], nm, child_idx);
                    } else if (x != root) {
                        int px = nodes[par_node] / Y % X, py = nodes[par_node] % Y;
                        if (nx != px && ny != py) {
                            int t[2];
                            for (int pi = 0; pi <= 1; pi++) {
                                int mx = pi ? nx : px, my = pi ? py : ny;
                                t[pi] = create_node(L, mx, my);
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
                    new_mirrors[dir].push_back(nm);
                }
            }
        }
        if (par_node != -1) {
            int px = nodes[par_node] / Y % X, py = nodes[par_node] % Y;
            if (px != curx && py != cury) {
                for (int pi = 0; pi <= 1; pi++) {
                    int mx = pi ? curx : px, my = pi ? py : cury;
                    int pre = create_node(L, mx, my);
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
                    std::vector<std::pair<float, int>> road;
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
                    for (int pm : mirrors[dir]) {
                        int qx = nodes[pm] / Y % X, qy = nodes[pm] % Y;
                        if (qx != curx && qy != cury) {
                            for (int pi = 0; pi <= 1; pi++) {
                                int mx = pi ? curx : qx, my = pi ? qy : cury;
                                int ins = create_node(L, mx, my);
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
                for (int m : new_mirrors[dir]) {
                    int dup = create_node(curl, curx, cury);
                    connect(m, dup, (int)adj[x].size() - 1);
                }
        }
        int idx = 0;
        for (int e : adj[x]) {
            if (e == par) continue;
            dfs(e, x, node_idx, idx++, depth + 1, new_mirrors, places);
        }
    }

    void build(const int* tnode, int tn, const int64_t* adj_off, const int* tadj, DagNet& out) {
        adj.assign(tn, {});
        pos.assign(tnode, tnode + tn);
        for (int i = 0; i < tn; i++) adj[i].assign(tadj + adj_off[i], tadj + adj_off[i + 1]);
        root = select_root(adj);
        for (int d = 0; d < 2; d++) {
            region[d].assign(tn, -1);
            ranges[d].assign(tn, {0, 0});
            stems[d].assign(tn, {});
        }
        for (int x = 0; x < tn; x++) {
            int curl = pos[x] / XYc, curx = pos[x] / Y % X, cury = pos[x] % Y;
            for (int dir = 0; dir < 2; dir++) {
                if (curl < L) stems[dir][x].push_back(dir ? curx : cury);
                int t = dir ? cury : curx;
                ranges[dir][x] = {t, t};
            }
        }
        markCongestion(root, -1);
        nodes.clear();
        parents.clear();
        child_num_pre.clear();
        points.clear();
        dfs(root, -1, -1, -1, 0, {}, {});
        std::sort(points.begin(), points.end());
        points.erase(std::unique(points.begin(), points.end()), points.end());

        // Topological renumbering (src/database_cuda.hpp:2016-2080).
        const int n = (int)nodes.size();
        std::vector<int> visit(n, 0), seq(n, 0);
        std::vector<int> st;
        st.reserve(n);
        std::function<void(int)> determine = [&](int idx) {
            st.push_back(idx);
            for (int content : parents[idx]) {
                int p = content / 10;
                if (++visit[p] == child_num_pre[p]) determine(p);
            }
        };
        for (int i = 0; i < n; i++)
            if (child_num_pre[i] == 0) determine(i);
        out.node.assign(n, 0);
        out.par_off.assign(n + 1, 0);
        out.child_num.assign(n, 0);
        out.depth.assign(n, 0);
        out.par.clear();
        out.slot.clear();
        int copy = 0;
        while (!st.empty()) {
            int idx = st.back();
            st.pop_back();
            for (int content : parents[idx]) {
                int p = content / 10, ci = content % 10;
                int sp = seq[p];
                out.child_num[sp] = std::max(out.child_num[sp], ci + 1);
                out.par.push_back(sp);
                out.slot.push_back(ci);
                out.depth[copy] = std::max(out.depth[copy], out.depth[sp] + 1);
            }
            seq[idx] = copy;
            out.par_off[copy + 1] = out.par_off[copy] + (int)parents[idx].size();
            out.node[copy++] = nodes[idx];
        }
        if (copy != n) throw std::runtime_error("detour DAG has a cycle");
        out.points = points;
    }
};

} // namespace

void build_dag(const Grid& g, const float* xsum, const float* ysum, int mode, const int* tnode, int tn,
               const int64_t* adj_off, const int* tadj, DagNet& out) {
    Builder b(g, xsum, ysum, mode);
    b.build(tnode, tn, adj_off, tadj, out);
}

void build_dags(const Grid& g, const Subnets& s, const int* nets, int64_t n, const float* xsum, const float* ysum,
                int mode, int threads, DagSet& out) {
    std::vector<DagNet> per(n);
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    std::vector<std::string> errs(threads);
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t]() {
            try {
                for (int64_t i = t; i < n; i += threads) {
                    int net = nets[i];
                    int64_t b = s.node_off[net];
                    int tn = (int)(s.node_off[net + 1] - b);
                    std::vector<int64_t> aoff(tn + 1);
                    for (int k = 0; k <= tn; k++) aoff[k] = s.adj_off[b + k] - s.adj_off[b];
                    build_dag(g, xsum, ysum, mode, &s.node[b], tn, aoff.data(), &s.adj[s.adj_off[b]], per[i]);
                }
            } catch (const std::exception& e) {
                errs[t] = e.what();
            }
        });
    for (auto& th : pool) th.join();
    for (auto& e : errs)
        if (!e.empty()) throw std::runtime_error(e);
    out = DagSet();
    out.n = n;
    out.node_off.assign(n + 1, 0);
    out.pt_off.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        out.node_off[i + 1] = out.node_off[i] + (int64_t)per[i].node.size();
        out.pt_off[i + 1] = out.pt_off[i] + (int64_t)per[i].points.size();
    }
    const int64_t N = out.node_off[n];
    out.node.resize(N);
    out.child_num.resize(N);
    out.depth.resize(N);
    out.par_off.assign(N + 1, 0);
    out.pt.resize(out.pt_off[n]);
    for (int64_t i = 0; i < n; i++) {
        const DagNet& d = per[i];
        int64_t b = out.node_off[i];
        for (size_t k = 0; k < d.node.size(); k++) {
            out.node[b + k] = d.node[k];
            out.child_num[b + k] = d.child_num[k];
            out.depth[b + k] = d.depth[k];
            out.par_off[b + k + 1] = out.par_off[b + k] + (d.par_off[k + 1] - d.par_off[k]);
        }
        out.par.insert(out.par.end(), d.par.begin(), d.par.end());
        out.slot.insert(out.slot.end(), d.slot.begin(), d.slot.end());
        std::copy(d.points.begin(), d.points.end(), out.pt.begin() + out.pt_off[i]);
    }
}

} // namespace igr
