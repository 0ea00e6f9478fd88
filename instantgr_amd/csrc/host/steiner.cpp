// Rectilinear Steiner trees and the subnet split.
//
// The reference builds trees with FLUTE (src/flute.hpp) whose d>=4 topologies
// come from POST9.dat, a file absent from the reference tree; this builder
// therefore constructs its own trees: a rectilinear minimum spanning tree by
// Prim's algorithm, improved by repeatedly replacing two edges u-v, u-w of a
// node with the three-point Steiner star at the coordinate-wise median of
// u, v, w whenever that shortens the tree.  The result is then normalised
// exactly as the reference normalises FLUTE output in my_flute
// (src/database_cuda.hpp:696-744): node positions sorted and de-duplicated,
// edges sorted by (length, u, v) and re-spanned with Kruskal; adjacency lists
// keep Kruskal's acceptance order.  Lengths are in the cap file's length unit
// (gcell edge lengths), as FLUTE's inputs are.
//
// Nets with more than 12 pins are split into subnets with the union-find pass
// of src/database_cuda.hpp:2111-2228.
#include "design.hpp"
#include "par.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace igr {
namespace {

constexpr int MAX_PIN_SIZE = 12;   // src/database_cuda.hpp:2111

struct Tree2D {
    std::vector<int> pos;                 // sorted distinct x*Y+y
    std::vector<std::vector<int>> adj;    // Kruskal acceptance order
};

struct Coords {
    int X, Y;
    std::vector<long long> drx, dry;      // prefix sums of gcell lengths
    long long dist(int x0, int y0, int x1, int y1) const {
        return std::llabs(drx[x0] - drx[x1]) + std::llabs(dry[y0] - dry[y1]);
    }
};

int median3(int a, int b, int c) { return std::max(std::min(a, b), std::min(std::max(a, b), c)); }

// Normalise a node/edge set the way my_flute does.
void normalise(const Coords& C, const std::vector<int>& nx, const std::vector<int>& ny,
               const std::vector<std::pair<int, int>>& edges, Tree2D& t) {
    std::vector<int> p(nx.size());
    for (size_t i = 0; i < nx.size(); i++) p[i] = nx[i] * C.Y + ny[i];
    t.pos = p;
    std::sort(t.pos.begin(), t.pos.end());
    t.pos.erase(std::unique(t.pos.begin(), t.pos.end()), t.pos.end());
    auto id = [&](int v) { return (int)(std::lower_bound(t.pos.begin(), t.pos.end(), v) - t.pos.begin()); };
    std::vector<std::tuple<long long, int, int>> es;
    es.reserve(edges.size());
    for (auto& e : edges) {
        int u = id(p[e.first]), v = id(p[e.second]);
        if (u == v) continue;
        long long len = C.dist(nx[e.first], ny[e.first], nx[e.second], ny[e.second]);
        es.emplace_back(len, std::min(u, v), std::max(u, v));
    }
    std::sort(es.begin(), es.end());
    std::vector<int> par(t.pos.size());
    std::iota(par.begin(), par.end(), 0);
    std::function<int(int)> find = [&](int x) { return par[x] == x ? x : par[x] = find(par[x]); };
    t.adj.assign(t.pos.size(), {});
    for (auto& e : es) {
        int u = std::get<1>(e), v = std::get<2>(e), pu = find(u), pv = find(v);
        if (pu == pv) continue;
        t.adj[u].push_back(v);
        t.adj[v].push_back(u);
        par[pu] = pv;
    }
}

// Steiner tree over distinct 2-D positions.
void steiner_tree(const Coords& C, const std::vector<int>& pts, Tree2D& t) {
    const int n = (int)pts.size();
    std::vector<int> nx(n), ny(n);
    for (int i = 0; i < n; i++) { nx[i] = pts[i] / C.Y; ny[i] = pts[i] % C.Y; }
    std::vector<std::pair<int, int>> edges;
    if (n >= 2) {
        // Prim, ties to the smallest index.  The nodes outside the tree sit
        // in structure-of-arrays form; one pass per step relaxes them against
        // the node just added and finds the next by the key (distance << sh)
        // | index, whose minimum is the nearest node with the smallest index.
        // The added node leaves by a swap with the last entry.
        int sh = 1;
        while ((1 << sh) <= n) sh++;
        const long long span = C.drx.back() + C.dry.back();
        if (span >= (1LL << (62 - sh))) throw std::runtime_error("Steiner tree: coordinate span too large");
        std::vector<long long> rx(n - 1), ry(n - 1), rk(n - 1);
        std::vector<int> rid(n - 1), rfrom(n - 1, 0);
        const long long x0 = C.drx[nx[0]], y0 = C.dry[ny[0]];
        int m = n - 1, pj = 0;
        for (int r = 0; r < m; r++) {
            const int k = r + 1;
            rx[r] = C.drx[nx[k]];
            ry[r] = C.dry[ny[k]];
            rid[r] = k;
            rk[r] = ((std::llabs(x0 - rx[r]) + std::llabs(y0 - ry[r])) << sh) | k;
            if (rk[r] < rk[pj]) pj = r;
        }
        std::vector<std::vector<int>> adj(n);
        while (m > 0) {
            const int j = rid[pj], fj = rfrom[pj];
            const long long jx = rx[pj], jy = ry[pj];
            adj[fj].push_back(j);
            adj[j].push_back(fj);
            m--;
            rx[pj] = rx[m]; ry[pj] = ry[m]; rk[pj] = rk[m]; rid[pj] = rid[m]; rfrom[pj] = rfrom[m];
            long long nk = LLONG_MAX;
            int nj = 0;
            for (int r = 0; r < m; r++) {
                const long long dk = ((std::llabs(jx - rx[r]) + std::llabs(jy - ry[r])) << sh) | rid[r];
                if (dk < rk[r]) { rk[r] = dk; rfrom[r] = j; }
                if (rk[r] < nk) { nk = rk[r]; nj = r; }
            }
            pj = nj;
        }
        // Steiner improvement: median stars on pairs of incident edges.
        auto D = [&](int a, int b) { return C.dist(nx[a], ny[a], nx[b], ny[b]); };
        auto drop = [&](int a, int b) {
            auto& v = adj[a];
            v.erase(std::find(v.begin(), v.end(), b));
        };
        bool changed = true;
        while (changed) {
            changed = false;
            for (int u = 0; u < (int)adj.size(); u++) {
                for (;;) {
                    long long bestgain = 0;
                    int bv = -1, bw = -1, sx = 0, sy = 0;
                    const auto& au = adj[u];
                    for (size_t a = 0; a < au.size(); a++)
                        for (size_t b = a + 1; b < au.size(); b++) {
                            int v = au[a], w = au[b];
                            int mx = median3(nx[u], nx[v], nx[w]), my = median3(ny[u], ny[v], ny[w]);
                            long long g = D(u, v) + D(u, w)
                                - C.dist(nx[u], ny[u], mx, my) - C.dist(nx[v], ny[v], mx, my)
                                - C.dist(nx[w], ny[w], mx, my);
                            if (g > bestgain) { bestgain = g; bv = v; bw = w; sx = mx; sy = my; }
                        }
                    if (bestgain <= 0) break;
                    drop(u, bv); drop(bv, u); drop(u, bw); drop(bw, u);
                    int s = -1;
                    for (int c : {u, bv, bw})
                        if (nx[c] == sx && ny[c] == sy) { s = c; break; }
                    if (s < 0) {
                        s = (int)nx.size();
                        nx.push_back(sx); ny.push_back(sy);
                        adj.emplace_back();
                    }
                    for (int c : {u, bv, bw})
                        if (c != s) { adj[s].push_back(c); adj[c].push_back(s); }
                    changed = true;
                }
            }
        }
        for (int a = 0; a < (int)adj.size(); a++)
            for (int b : adj[a])
                if (a < b) edges.emplace_back(a, b);
    }
    normalise(C, nx, ny, edges, t);
}

struct SubOut {
    std::vector<int> pins;                // routing-layer pins
    std::vector<int> node;                // with layer (L = Steiner)
    std::vector<std::vector<int>> adj;
};

// Encode tree nodes with their pin layer or L (src/database_cuda.hpp:790-792).
// The pins have distinct 2-D positions, so each node matches at most one.
void encode_layers(const Grid& g, const std::vector<int>& pins, SubOut& s) {
    const int XY = g.X * g.Y;
    if (pins.size() <= 16) {
        for (auto& e : s.node) {
            int pos = e % XY, lay = g.L;
            for (int p : pins)
                if (p % XY == pos) { lay = p / XY; break; }
            e = pos + lay * XY;
        }
        return;
    }
    std::vector<std::pair<int, int>> byPos(pins.size());
    for (size_t k = 0; k < pins.size(); k++) byPos[k] = {pins[k] % XY, pins[k] / XY};
    std::sort(byPos.begin(), byPos.end());
    for (auto& e : s.node) {
        const int pos = e % XY;
        auto it = std::lower_bound(byPos.begin(), byPos.end(), std::make_pair(pos, -1));
        const int lay = it != byPos.end() && it->first == pos ? it->second : g.L;
        e = pos + lay * XY;
    }
}

void split_net(const Coords& C, const Grid& g, const std::vector<int>& pins, std::vector<SubOut>& out) {
    const int XY = g.X * g.Y;
    std::vector<int> pts(pins.size());
    for (size_t i = 0; i < pins.size(); i++) pts[i] = pins[i] % XY;
    std::sort(pts.begin(), pts.end());
    if (std::adjacent_find(pts.begin(), pts.end()) != pts.end())
        throw std::runtime_error("subnet pins must have distinct 2-D positions");
    Tree2D t;
    steiner_tree(C, pts, t);
    if ((int)pins.size() <= MAX_PIN_SIZE) {
        out.emplace_back();
        SubOut& s = out.back();
        s.pins = pins;
        s.node = t.pos;
        s.adj = t.adj;
        encode_layers(g, pins, s);
        return;
    }
    // Union-find split (src/database_cuda.hpp:2133-2228).
    SubOut whole;
    whole.node = t.pos;
    encode_layers(g, pins, whole);
    const int n = (int)t.pos.size();
    std::vector<char> not_steiner(n);
    for (int i = 0; i < n; i++) not_steiner[i] = whole.node[i] < g.L * XY;
    std::vector<int> sz(n, 1), par(n);
    std::iota(par.begin(), par.end(), 0);
    std::function<int(int)> find = [&](int x) { return par[x] == x ? x : par[x] = find(par[x]); };
    std::vector<std::tuple<int, int, int>> edges;
    for (int i = 0; i < n; i++)
        for (int j : t.adj[i])
            if (j < i) edges.emplace_back(j, i, not_steiner[i] + not_steiner[j]);
    std::stable_sort(edges.begin(), edges.end(),
                     [](const auto& a, const auto& b) { return std::get<2>(a) < std::get<2>(b); });
    bool big_steiner = false;
    for (auto& e : edges) {
        int u = find(std::get<0>(e)), v = find(std::get<1>(e));
        if (u == v) continue;
        if (sz[u] + sz[v] > MAX_PIN_SIZE) {
            if (!(not_steiner[std::get<0>(e)] || not_steiner[std::get<1>(e)])) big_steiner = true;
            if (!big_steiner) continue;
        }
        if (sz[u] > sz[v]) std::swap(u, v);
        par[u] = v;
        sz[v] += sz[u];
    }
    std::vector<std::vector<int>> members(n), net_edge(n);
    for (int i = 0; i < n; i++) members[find(i)].push_back(i);
    for (int i = 0; i < (int)edges.size(); i++) {
        int u = std::get<0>(edges[i]), v = std::get<1>(edges[i]);
        int pu = find(u), pv = find(v);
        if (pu == pv) { net_edge[pu].push_back(i); continue; }
        if (sz[u] > sz[v]) { std::swap(u, v); std::swap(pu, pv); }
        if (!not_steiner[v] && not_steiner[u]) { std::swap(u, v); std::swap(pu, pv); }
        sz[pu]++;
        if (sz[pv] == 1) par[pv] = pu;
        net_edge[pu].push_back(i);
        // A boundary node shared by two subnets; appended once per subnet so
        // the subnet tree stays a tree.
        if (std::find(members[pu].begin(), members[pu].end(), v) == members[pu].end())
            members[pu].push_back(v);
    }
    for (int r = 0; r < n; r++) {
        if (find(r) != r || members[r].empty()) continue;
        SubOut s;
        std::vector<int> local(n, -1);
        for (int m : members[r]) {
            local[m] = (int)s.node.size();
            s.node.push_back(whole.node[m]);
            if (whole.node[m] < g.L * XY) s.pins.push_back(whole.node[m]);
        }
        s.adj.assign(s.node.size(), {});
        std::vector<int> up(s.node.size());
        std::iota(up.begin(), up.end(), 0);
        std::function<int(int)> f2 = [&](int x) { return up[x] == x ? x : up[x] = f2(up[x]); };
        for (int ei : net_edge[r]) {
            int a = local[std::get<0>(edges[ei])], b = local[std::get<1>(edges[ei])];
            if (a < 0 || b < 0) continue;
            int ra = f2(a), rb = f2(b);
            if (ra == rb) continue;
            up[ra] = rb;
            s.adj[a].push_back(b);
            s.adj[b].push_back(a);
        }
        if (s.node.size() >= 2) out.push_back(std::move(s));
    }
}

} // namespace

void build_subnets(const Design& d, const NetPins& np, const Grid& g, Subnets& s, int threads) {
    const int64_t N = d.num_nets();
    const int XY = g.X * g.Y;
    Coords C{g.X, g.Y, std::vector<long long>(g.X, 0), std::vector<long long>(g.Y, 0)};
    for (int i = 0; i + 1 < g.X; i++) C.drx[i + 1] = C.drx[i] + g.xlen[i];
    for (int i = 0; i + 1 < g.Y; i++) C.dry[i + 1] = C.dry[i] + g.ylen[i];
    std::vector<std::vector<SubOut>> per(N);
    if (threads < 1) threads = 1;
    std::atomic<int64_t> next(0);
    const int64_t CH = 256;   // nets per work item (heavy nets are spread by the dynamic pull)
    parallel_ranges(threads, threads, [&](int64_t, int64_t, int) {
        std::vector<int> pins;
        for (;;) {
            const int64_t c0 = next.fetch_add(CH);
            if (c0 >= N) break;
            for (int64_t n = c0; n < std::min(N, c0 + CH); n++) {
                int64_t b = np.pin_off[n], e = np.pin_off[n + 1];
                if (e - b <= 1) continue;
                pins.assign(np.pins.begin() + b, np.pins.begin() + e);
                // Metal-1 pins stay on routing layer 0, others drop one layer
                // (src/database_cuda.hpp:2124-2125).
                for (auto& p : pins)
                    if (p >= XY) p -= XY;
                split_net(C, g, pins, per[n]);
            }
        }
    });

    // Sizes per net, offsets, then a parallel fill in net order.
    std::vector<int64_t> c_sub(N + 1, 0), c_pin(N + 1, 0), c_node(N + 1, 0), c_adj(N + 1, 0);
    parallel_ranges(N, threads, [&](int64_t a, int64_t b, int) {
        for (int64_t n = a; n < b; n++) {
            int64_t ps = 0, ns = 0, as = 0;
            for (const auto& so : per[n]) {
                ps += (int64_t)so.pins.size();
                ns += (int64_t)so.node.size();
                for (const auto& v : so.adj) as += (int64_t)v.size();
            }
            c_sub[n + 1] = (int64_t)per[n].size();
            c_pin[n + 1] = ps;
            c_node[n + 1] = ns;
            c_adj[n + 1] = as;
        }
    });
    for (int64_t n = 0; n < N; n++) {
        c_sub[n + 1] += c_sub[n];
        c_pin[n + 1] += c_pin[n];
        c_node[n + 1] += c_node[n];
        c_adj[n + 1] += c_adj[n];
    }
    const int64_t S = c_sub[N], NN = c_node[N];
    s = Subnets();
    s.count = S;
    s.net_sub_off = c_sub;
    s.net_sub.resize(S);
    s.orig.resize(S);
    s.hpwl.resize(S);
    s.pin_off.resize(S + 1);
    s.pins.resize(c_pin[N]);
    s.node_off.resize(S + 1);
    s.node.resize(NN);
    s.adj_off.resize(NN + 1);
    s.adj.resize(c_adj[N]);
    parallel_ranges(N, threads, [&](int64_t a, int64_t b, int) {
        for (int64_t n = a; n < b; n++) {
            int64_t k = c_sub[n], pc = c_pin[n], nc = c_node[n], ac = c_adj[n];
            for (const auto& so : per[n]) {
                s.net_sub[k] = (int)k;
                s.orig[k] = (int)n;
                int minx = g.X, maxx = 0, miny = g.Y, maxy = 0;
                for (int p : so.pins) {
                    int x = p / g.Y % g.X, y = p % g.Y;
                    minx = std::min(minx, x); maxx = std::max(maxx, x);
                    miny = std::min(miny, y); maxy = std::max(maxy, y);
                }
                s.hpwl[k] = maxx - minx + maxy - miny;
                s.pin_off[k] = pc;
                std::copy(so.pins.begin(), so.pins.end(), s.pins.begin() + pc);
                pc += (int64_t)so.pins.size();
                s.node_off[k] = nc;
                for (size_t i = 0; i < so.node.size(); i++) {
                    s.node[nc] = so.node[i];
                    s.adj_off[nc] = ac;
                    std::copy(so.adj[i].begin(), so.adj[i].end(), s.adj.begin() + ac);
                    ac += (int64_t)so.adj[i].size();
                    nc++;
                }
                k++;
            }
            std::vector<SubOut>().swap(per[n]);
        }
    });
    s.pin_off[S] = c_pin[N];
    s.node_off[S] = NN;
    s.adj_off[NN] = c_adj[N];
}

} // namespace igr
