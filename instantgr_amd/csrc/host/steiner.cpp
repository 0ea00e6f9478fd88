// Rectilinear Steiner trees and the subnet split.
//
// The reference builds trees with FLUTE (src/flute.hpp) whose d>=4 topologies
// come from POST9.dat, a file absent from the reference tree; this builder
// therefore constructs its own trees.  Nets of two to four pins get an exact
// rectilinear Steiner minimal tree (the length FLUTE, optimal up to degree 9,
// also reaches): two and three pins by the median star, four pins by
// enumerating Hanan-grid Steiner points.  Larger nets get a rectilinear
// minimum spanning tree by Prim's algorithm, improved by repeatedly replacing
// two edges u-v, u-w of a node with the three-point Steiner star at the
// coordinate-wise median of u, v, w whenever that shortens the tree.  The
// result is then normalised
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
#include <cstdint>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace igr {
namespace {

constexpr int MAX_PIN_SIZE = 12;   // src/database_cuda.hpp:2111

// Adjacency in compressed rows: the neighbours of node i are
// nb[off[i] .. off[i + 1]).
struct Adj {
    std::vector<int> off, nb;
};

// Adjacency of n nodes from an edge list; each node lists its neighbours in
// the order of the edges.
void adj_from_edges(int n, const std::vector<std::pair<int, int>>& edges, Adj& a) {
    a.off.assign(n + 1, 0);
    for (auto& e : edges) {
        a.off[e.first + 1]++;
        a.off[e.second + 1]++;
    }
    for (int i = 0; i < n; i++) a.off[i + 1] += a.off[i];
    a.nb.resize(a.off[n]);
    thread_local std::vector<int> cur;
    cur.assign(a.off.begin(), a.off.end() - 1);
    for (auto& e : edges) {
        a.nb[cur[e.first]++] = e.second;
        a.nb[cur[e.second]++] = e.first;
    }
}

struct Tree2D {
    std::vector<int> pos;    // sorted distinct x*Y+y
    Adj adj;                 // Kruskal acceptance order
};

struct Coords {
    int X, Y;
    std::vector<long long> drx, dry;      // prefix sums of gcell lengths
    long long dist(int x0, int y0, int x1, int y1) const {
        return std::llabs(drx[x0] - drx[x1]) + std::llabs(dry[y0] - dry[y1]);
    }
};

int median3(int a, int b, int c) { return std::max(std::min(a, b), std::min(std::max(a, b), c)); }

// Disjoint sets with path halving (the roots are those of any other find).
struct Dsu {
    std::vector<int> p;
    explicit Dsu(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int find(int x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
};

// Normalise a node/edge set the way my_flute does.
void normalise(const Coords& C, const std::vector<int>& nx, const std::vector<int>& ny,
               const std::vector<std::pair<int, int>>& edges, Tree2D& t) {
    const int n = (int)nx.size();
    // Positions are non-negative ints: (position, index) sorts as one
    // unsigned 64-bit key.
    std::vector<uint64_t> byPos(n);
    for (int i = 0; i < n; i++) byPos[i] = (uint64_t)(uint32_t)(nx[i] * C.Y + ny[i]) << 32 | (uint32_t)i;
    std::sort(byPos.begin(), byPos.end());
    std::vector<int> id(n);
    t.pos.clear();
    for (int k = 0; k < n; k++) {
        if (k == 0 || byPos[k] >> 32 != byPos[k - 1] >> 32) t.pos.push_back((int)(byPos[k] >> 32));
        id[(uint32_t)byPos[k]] = (int)t.pos.size() - 1;
    }
    std::vector<std::pair<int, int>> acc;
    acc.reserve(edges.size());
    Dsu ds((int)t.pos.size());
    auto kruskal = [&](int u, int v) {
        const int pu = ds.find(u), pv = ds.find(v);
        if (pu == pv) return;
        acc.emplace_back(u, v);
        ds.p[pu] = pv;
    };
    // Edges by (length, u, v): one 64-bit key when the node ids fit 16 bits
    // and the lengths 32, else the tuples themselves.
    long long maxlen = 0;
    for (auto& e : edges) maxlen = std::max(maxlen, C.dist(nx[e.first], ny[e.first], nx[e.second], ny[e.second]));
    if (t.pos.size() <= 65536 && maxlen < (1ll << 32)) {
        std::vector<uint64_t> es;
        es.reserve(edges.size());
        for (auto& e : edges) {
            const int u = id[e.first], v = id[e.second];
            if (u == v) continue;
            const uint64_t len = (uint64_t)C.dist(nx[e.first], ny[e.first], nx[e.second], ny[e.second]);
            es.push_back(len << 32 | (uint64_t)std::min(u, v) << 16 | (uint64_t)std::max(u, v));
        }
        std::sort(es.begin(), es.end());
        for (uint64_t e : es) kruskal((int)(e >> 16 & 0xffff), (int)(e & 0xffff));
    } else {
        std::vector<std::tuple<long long, int, int>> es;
        es.reserve(edges.size());
        for (auto& e : edges) {
            int u = id[e.first], v = id[e.second];
            if (u == v) continue;
            long long len = C.dist(nx[e.first], ny[e.first], nx[e.second], ny[e.second]);
            es.emplace_back(len, std::min(u, v), std::max(u, v));
        }
        std::sort(es.begin(), es.end());
        for (auto& e : es) kruskal(std::get<1>(e), std::get<2>(e));
    }
    adj_from_edges((int)t.pos.size(), acc, t.adj);
}

// One Prim step over the nodes outside the tree, kept in ascending index
// order: relax them against the node just added (j at jx, jy) and return the
// position of the nearest, the first on ties, so the smallest index.
template <class T>
__attribute__((target_clones("avx2", "default")))
int prim_step(const T* __restrict rx, const T* __restrict ry, T* __restrict rd, int* __restrict rfrom, int m, T jx,
              T jy, int j) {
    T mn = std::numeric_limits<T>::max();
    for (int r = 0; r < m; r++) {
        const T dx = jx - rx[r], dy = jy - ry[r];
        const T d = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
        const bool lt = d < rd[r];
        rd[r] = lt ? d : rd[r];
        rfrom[r] = lt ? j : rfrom[r];
        mn = rd[r] < mn ? rd[r] : mn;
    }
    int r = 0;
    while (rd[r] != mn) r++;
    return r;
}

// Rectilinear minimum spanning tree by Prim from node 0, ties to the smallest
// index; edges are appended to adj in the order they join the tree.
template <class T>
void prim(const Coords& C, const std::vector<int>& nx, const std::vector<int>& ny, std::vector<std::vector<int>>& adj) {
    const int n = (int)nx.size();
    int m = n - 1;
    std::vector<T> rx(m), ry(m), rd(m, std::numeric_limits<T>::max());
    std::vector<int> rid(m), rfrom(m, 0);
    for (int r = 0; r < m; r++) {
        rx[r] = (T)C.drx[nx[r + 1]];
        ry[r] = (T)C.dry[ny[r + 1]];
        rid[r] = r + 1;
    }
    int j = 0;
    T jx = (T)C.drx[nx[0]], jy = (T)C.dry[ny[0]];
    while (m > 0) {
        const int p = prim_step<T>(rx.data(), ry.data(), rd.data(), rfrom.data(), m, jx, jy, j);
        j = rid[p];
        jx = rx[p];
        jy = ry[p];
        adj[rfrom[p]].push_back(j);
        adj[j].push_back(rfrom[p]);
        m--;
        for (auto* v : {&rx, &ry, &rd}) std::copy(v->begin() + p + 1, v->begin() + m + 1, v->begin() + p);
        for (auto* v : {&rid, &rfrom}) std::copy(v->begin() + p + 1, v->begin() + m + 1, v->begin() + p);
    }
}

// Length of a minimum spanning tree over at most 6 points (O(n^2) Prim).
long long small_mst(const Coords& C, const int* x, const int* y, int n) {
    long long best[6], tot = 0;
    bool in[6] = {false, false, false, false, false, false};
    for (int k = 0; k < n; k++) best[k] = C.dist(x[0], y[0], x[k], y[k]);
    in[0] = true;
    for (int it = 1; it < n; it++) {
        int j = -1;
        for (int k = 0; k < n; k++)
            if (!in[k] && (j < 0 || best[k] < best[j])) j = k;
        in[j] = true;
        tot += best[j];
        for (int k = 0; k < n; k++)
            if (!in[k]) best[k] = std::min(best[k], C.dist(x[j], y[j], x[k], y[k]));
    }
    return tot;
}

// Exact rectilinear Steiner minimal tree of four terminals: the spanning tree
// over the pins plus the best set of at most two Hanan-grid points (Hanan's
// theorem holds for the L1 metric on the gcell-length coordinates, which are
// monotone in the gcell indices).  Fewer Steiner points win ties; among equal
// sets the first in (x, y) order.  Appends the chosen points to nx, ny.
void rsmt4(const Coords& C, std::vector<int>& nx, std::vector<int>& ny) {
    int hx[4], hy[4];
    for (int k = 0; k < 4; k++) { hx[k] = nx[k]; hy[k] = ny[k]; }
    std::sort(hx, hx + 4);
    std::sort(hy, hy + 4);
    const int ux = (int)(std::unique(hx, hx + 4) - hx), uy = (int)(std::unique(hy, hy + 4) - hy);
    int cx[16], cy[16], m = 0;
    for (int a = 0; a < ux; a++)
        for (int b = 0; b < uy; b++) {
            bool pin = false;
            for (int k = 0; k < 4; k++) pin = pin || (nx[k] == hx[a] && ny[k] == hy[b]);
            if (!pin) { cx[m] = hx[a]; cy[m] = hy[b]; m++; }
        }
    int x[6] = {nx[0], nx[1], nx[2], nx[3], 0, 0}, y[6] = {ny[0], ny[1], ny[2], ny[3], 0, 0};
    long long best = small_mst(C, x, y, 4);
    int s1 = -1, s2 = -1;
    for (int i = 0; i < m; i++) {
        x[4] = cx[i]; y[4] = cy[i];
        const long long l1 = small_mst(C, x, y, 5);
        if (l1 < best) { best = l1; s1 = i; s2 = -1; }
    }
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++) {
            x[4] = cx[i]; y[4] = cy[i];
            x[5] = cx[j]; y[5] = cy[j];
            const long long l2 = small_mst(C, x, y, 6);
            if (l2 < best) { best = l2; s1 = i; s2 = j; }
        }
    if (s1 >= 0) { nx.push_back(cx[s1]); ny.push_back(cy[s1]); }
    if (s2 >= 0) { nx.push_back(cx[s2]); ny.push_back(cy[s2]); }
}

// Exact rectilinear Steiner minimal tree of 5..9 terminals by the
// Dreyfus-Wagner recurrence (Dreyfus and Wagner, Networks 1971) over the
// Hanan grid of the terminals (Hanan 1966: some minimal tree uses only grid
// points), with L1 distances in gcell-length coordinates.  The last terminal
// is the root; for every subset D of the others and grid point v,
//   S(D, v) = min_u dist(v, u) + g(D, u),
//   g(D, u) = min over splits D = E + F (E holding D's lowest terminal) of
//             S(E, u) + S(F, u),
// with S({t}, v) = dist(v, t).  The min over u is an L1 distance transform
// on the grid (a forward and a backward sweep per axis), so a subset costs
// O(splits * grid) instead of O(grid^2).  Ties keep the first candidate in
// sweep and split order, so the tree is deterministic.  The tree is read back
// from the argmins as edges between grid points; Steiner points of degree
// two on a shortest path between their neighbours are dropped.  Returns the
// edges over nx/ny (terminals first, Steiner points appended).
// g[v] = min over the splits of 'mask' (the part holding its lowest bit, and
// the rest) of S[e][v] + S[mask ^ e][v].
// Rows are n entries long, n a multiple of 8 (grid points padded).
template <class T>
__attribute__((target_clones("avx2", "default")))
void dw_splits(const T* __restrict S, int n, int mask, T* __restrict g) {
    const int low = mask & -mask, rest = mask ^ low;
    for (int v = 0; v < n; v++) g[v] = std::numeric_limits<T>::max();
    for (int sub = (rest - 1) & rest;; sub = (sub - 1) & rest) {
        const int e = low | sub, f = mask ^ e;
        const T* __restrict se = S + (size_t)e * n;
        const T* __restrict sf = S + (size_t)f * n;
        for (int v0 = 0; v0 < n; v0 += 8)
#pragma GCC unroll 8
            for (int v = v0; v < v0 + 8; v++) {
                const T c = se[v] + sf[v];
                g[v] = c < g[v] ? c : g[v];
            }
        if (sub == 0) break;
    }
}

// Exact rectilinear Steiner minimal tree of 5..9 terminals by the
// Dreyfus-Wagner recurrence (Dreyfus and Wagner, Networks 1971) over the
// Hanan grid of the terminals (Hanan 1966: some minimal tree uses only grid
// points), with L1 distances in gcell-length coordinates.  The last terminal
// is the root; for every subset D of the others and grid point v,
//   S(D, v) = min_u dist(v, u) + g(D, u),
//   g(D, u) = min over splits D = E + F (E holding D's lowest terminal) of
//             S(E, u) + S(F, u),
// with S({t}, v) = dist(v, t).  The min over u is an L1 distance transform
// on the grid (a forward and a backward sweep per axis), so a subset costs
// O(splits * grid) instead of O(grid^2).  The tree is read back top-down:
// at (D, v) the first grid point u (in grid order) with dist(v, u) + g(D, u)
// = S(D, v), then the first split (decreasing subset order) attaining
// g(D, u); so the tree is deterministic.  Steiner points of degree two on a
// shortest path between their neighbours are dropped.  Returns the edges
// over nx/ny (terminals first, Steiner points appended).
template <class T>
void rsmt_dw(const Coords& C, std::vector<int>& nx, std::vector<int>& ny, std::vector<std::pair<int, int>>& edges) {
    const int k = (int)nx.size();
    int hx[9], hy[9];
    for (int i = 0; i < k; i++) { hx[i] = nx[i]; hy[i] = ny[i]; }
    std::sort(hx, hx + k);
    std::sort(hy, hy + k);
    const int HX = (int)(std::unique(hx, hx + k) - hx), HY = (int)(std::unique(hy, hy + k) - hy);
    const int n = HX * HY;
    T cx[9], cy[9];
    for (int i = 0; i < HX; i++) cx[i] = (T)C.drx[hx[i]];
    for (int j = 0; j < HY; j++) cy[j] = (T)C.dry[hy[j]];
    auto dist = [&](int v, int u) {
        const int vi = v / HY, vj = v % HY, ui = u / HY, uj = u % HY;
        return (cx[vi] > cx[ui] ? cx[vi] - cx[ui] : cx[ui] - cx[vi]) +
               (cy[vj] > cy[uj] ? cy[vj] - cy[uj] : cy[uj] - cy[vj]);
    };
    int term[9];
    for (int t = 0; t < k; t++)
        term[t] = (int)(std::lower_bound(hx, hx + HX, nx[t]) - hx) * HY +
                  (int)(std::lower_bound(hy, hy + HY, ny[t]) - hy);
    const int m = k - 1, M = 1 << m;
    const int np = (n + 7) & ~7;   // row stride; padding entries stay 0
    thread_local std::vector<T> S, G;
    S.assign((size_t)M * np, T(0));
    G.resize((size_t)M * np);
    for (int t = 0; t < m; t++) {
        T* st = &S[(size_t)(1 << t) * np];
        for (int v = 0; v < n; v++) st[v] = dist(v, term[t]);
    }
    for (int mask = 3; mask < M; mask++) {
        if ((mask & (mask - 1)) == 0) continue;
        T* g = &G[(size_t)mask * np];
        T* sm = &S[(size_t)mask * np];
        dw_splits<T>(S.data(), np, mask, g);
        // L1 distance transform: along x (j fixed), then along y (i fixed)
        for (int v = 0; v < n; v++) sm[v] = g[v];
        for (int i = 1; i < HX; i++) {
            const T dx = cx[i] - cx[i - 1];
            for (int j = 0; j < HY; j++) sm[i * HY + j] = std::min(sm[i * HY + j], sm[(i - 1) * HY + j] + dx);
        }
        for (int i = HX - 2; i >= 0; i--) {
            const T dx = cx[i + 1] - cx[i];
            for (int j = 0; j < HY; j++) sm[i * HY + j] = std::min(sm[i * HY + j], sm[(i + 1) * HY + j] + dx);
        }
        // (rows i independent: the inner loop carries no dependency)
        for (int j = 1; j < HY; j++) {
            const T dy = cy[j] - cy[j - 1];
            for (int i = 0; i < HX; i++) sm[i * HY + j] = std::min(sm[i * HY + j], sm[i * HY + j - 1] + dy);
        }
        for (int j = HY - 2; j >= 0; j--) {
            const T dy = cy[j + 1] - cy[j];
            for (int i = 0; i < HX; i++) sm[i * HY + j] = std::min(sm[i * HY + j], sm[i * HY + j + 1] + dy);
        }
    }
    // Tree from the recurrence: grid point -> node (terminals keep their ids).
    int node_of[81];
    for (int v = 0; v < n; v++) node_of[v] = -1;
    for (int t = 0; t < k; t++) node_of[term[t]] = t;
    auto node = [&](int v) {
        if (node_of[v] < 0) {
            node_of[v] = (int)nx.size();
            nx.push_back(hx[v / HY]);
            ny.push_back(hy[v % HY]);
        }
        return node_of[v];
    };
    std::vector<std::pair<int, int>> es;
    struct Item { int mask, v; };
    Item stack[64];
    int sp = 0;
    stack[sp++] = {M - 1, term[k - 1]};
    while (sp > 0) {
        const Item it = stack[--sp];
        if ((it.mask & (it.mask - 1)) == 0) {
            const int t = __builtin_ctz(it.mask);
            if (term[t] != it.v) es.emplace_back(node(it.v), t);
            continue;
        }
        const T* g = &G[(size_t)it.mask * np];
        const T target = S[(size_t)it.mask * np + it.v];
        int u = 0;
        while (dist(it.v, u) + g[u] != target) u++;
        if (u != it.v) es.emplace_back(node(it.v), node(u));
        const int low = it.mask & -it.mask, rest = it.mask ^ low;
        int e = low;
        for (int sub = (rest - 1) & rest;; sub = (sub - 1) & rest) {
            e = low | sub;
            if (S[(size_t)e * np + u] + S[(size_t)(it.mask ^ e) * np + u] == g[u]) break;
            if (sub == 0) break;
        }
        stack[sp++] = {it.mask ^ e, u};
        stack[sp++] = {e, u};
    }
    // Drop Steiner points of degree two that lie on a shortest path between
    // their neighbours (and Steiner leaves, which only zero-length edges make).
    const int nn = (int)nx.size();
    std::vector<std::vector<int>> adj(nn);
    for (auto& e : es) {
        adj[e.first].push_back(e.second);
        adj[e.second].push_back(e.first);
    }
    auto D = [&](int a, int b) { return C.dist(nx[a], ny[a], nx[b], ny[b]); };
    bool changed = true;
    while (changed) {
        changed = false;
        for (int s = k; s < nn; s++) {
            auto& a = adj[s];
            if (a.empty() || a.size() > 2) continue;
            if (a.size() == 2 && D(a[0], s) + D(s, a[1]) != D(a[0], a[1])) continue;
            for (int c : a) adj[c].erase(std::find(adj[c].begin(), adj[c].end(), s));
            if (a.size() == 2) {
                adj[a[0]].push_back(a[1]);
                adj[a[1]].push_back(a[0]);
            }
            a.clear();
            changed = true;
        }
    }
    edges.clear();
    for (int a = 0; a < nn; a++)
        for (int b : adj[a])
            if (a < b) edges.emplace_back(a, b);
}

// Steiner tree over distinct 2-D positions.
void steiner_tree(const Coords& C, const std::vector<int>& pts, Tree2D& t) {
    const int n = (int)pts.size();
    std::vector<int> nx(n), ny(n);
    for (int i = 0; i < n; i++) { nx[i] = pts[i] / C.Y; ny[i] = pts[i] % C.Y; }
    std::vector<std::pair<int, int>> edges;
    const long long span = C.drx.back() + C.dry.back();
    if (n >= 5 && n <= 9) {
        if (span < INT_MAX / 32) rsmt_dw<int>(C, nx, ny, edges);
        else rsmt_dw<long long>(C, nx, ny, edges);
    } else if (n == 4) {
        // exact: spanning tree over the pins and the chosen Steiner points
        rsmt4(C, nx, ny);
        std::vector<std::vector<int>> adj(nx.size());
        if (span < INT_MAX / 2) prim<int>(C, nx, ny, adj);
        else prim<long long>(C, nx, ny, adj);
        for (int a = 0; a < (int)adj.size(); a++)
            for (int b : adj[a])
                if (a < b) edges.emplace_back(a, b);
    } else if (n >= 2) {
        std::vector<std::vector<int>> adj(n);
        if (span < INT_MAX / 2) prim<int>(C, nx, ny, adj);
        else prim<long long>(C, nx, ny, adj);
        // Steiner improvement: median stars on pairs of incident edges.
        auto D = [&](int a, int b) { return C.dist(nx[a], ny[a], nx[b], ny[b]); };
        auto drop = [&](int a, int b) {
            auto& v = adj[a];
            v.erase(std::find(v.begin(), v.end(), b));
        };
        // A node whose neighbourhood has not changed since it last found no
        // gain finds none again, so the passes skip it (same result as
        // re-evaluating every node in every pass).
        std::vector<char> dirty(n, 1);
        bool changed = true;
        while (changed) {
            changed = false;
            for (int u = 0; u < (int)adj.size(); u++) {
                if (!dirty[u]) continue;
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
                    if (bestgain <= 0) {
                        dirty[u] = 0;
                        break;
                    }
                    drop(u, bv); drop(bv, u); drop(u, bw); drop(bw, u);
                    int s = -1;
                    for (int c : {u, bv, bw})
                        if (nx[c] == sx && ny[c] == sy) { s = c; break; }
                    if (s < 0) {
                        s = (int)nx.size();
                        nx.push_back(sx); ny.push_back(sy);
                        adj.emplace_back();
                        dirty.push_back(1);
                    }
                    for (int c : {u, bv, bw})
                        if (c != s) { adj[s].push_back(c); adj[c].push_back(s); }
                    for (int c : {u, bv, bw, s}) dirty[c] = 1;
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
    Adj adj;
};

// Encode tree nodes with their pin layer or L (src/database_cuda.hpp:790-792).
// The pins have distinct 2-D positions, so each node matches at most one.
// Tree nodes come from normalise with ascending 2-D positions: then one merge
// with the pins sorted by position finds every match.
void encode_layers(const Grid& g, const std::vector<int>& pins, SubOut& s) {
    const int XY = g.X * g.Y;
    if (std::is_sorted(s.node.begin(), s.node.end()) && (s.node.empty() || (s.node.front() >= 0 && s.node.back() < XY))) {
        thread_local std::vector<std::pair<int, int>> byPos;
        byPos.resize(pins.size());
        for (size_t k = 0; k < pins.size(); k++) byPos[k] = {pins[k] % XY, pins[k] / XY};
        std::sort(byPos.begin(), byPos.end());
        size_t k = 0;
        for (auto& e : s.node) {
            while (k < byPos.size() && byPos[k].first < e) k++;
            const int lay = k < byPos.size() && byPos[k].first == e ? byPos[k].second : g.L;
            e += lay * XY;
        }
        return;
    }
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
        s.node = std::move(t.pos);
        s.adj = std::move(t.adj);
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
    std::vector<int> sz(n, 1);
    Dsu ds(n);
    std::vector<int>& par = ds.p;
    auto find = [&](int x) { return ds.find(x); };
    std::vector<std::tuple<int, int, int>> edges;
    for (int i = 0; i < n; i++)
        for (int k = t.adj.off[i]; k < t.adj.off[i + 1]; k++) {
            const int j = t.adj.nb[k];
            if (j < i) edges.emplace_back(j, i, not_steiner[i] + not_steiner[j]);
        }
    {   // stable order by the pin count of the edge's ends (0, 1 or 2)
        std::vector<std::tuple<int, int, int>> by;
        by.reserve(edges.size());
        for (int key = 0; key <= 2; key++)
            for (auto& e : edges)
                if (std::get<2>(e) == key) by.push_back(e);
        edges.swap(by);
    }
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
    // Member and edge lists per root, kept per thread between nets so the
    // inner lists keep their capacity; emptied here for this net.
    thread_local std::vector<std::vector<int>> members, net_edge;
    if (members.size() < (size_t)n) {
        members.resize(n);
        net_edge.resize(n);
    }
    for (int i = 0; i < n; i++) {
        members[i].clear();
        net_edge[i].clear();
    }
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
    std::vector<int> local(n, -1);
    std::vector<std::pair<int, int>> acc;
    for (int r = 0; r < n; r++) {
        if (find(r) != r || members[r].empty()) continue;
        SubOut s;
        s.node.reserve(members[r].size());
        s.pins.reserve(members[r].size());
        for (int m : members[r]) {
            local[m] = (int)s.node.size();
            s.node.push_back(whole.node[m]);
            if (whole.node[m] < g.L * XY) s.pins.push_back(whole.node[m]);
        }
        thread_local Dsu up(0);
        up.p.resize(s.node.size());
        std::iota(up.p.begin(), up.p.end(), 0);
        acc.clear();
        for (int ei : net_edge[r]) {
            int a = local[std::get<0>(edges[ei])], b = local[std::get<1>(edges[ei])];
            if (a < 0 || b < 0) continue;
            int ra = up.find(a), rb = up.find(b);
            if (ra == rb) continue;
            up.p[ra] = rb;
            acc.emplace_back(a, b);
        }
        adj_from_edges((int)s.node.size(), acc, s.adj);
        for (int m : members[r]) local[m] = -1;
        if (s.node.size() >= 2) out.push_back(std::move(s));
    }
}

} // namespace

void steiner_of_points(const std::vector<int>& x, const std::vector<int>& y, const std::vector<int>& xlen,
                       const std::vector<int>& ylen, std::vector<std::pair<int, int>>& nodes,
                       std::vector<std::pair<int, int>>& edges) {
    const int X = (int)xlen.size() + 1, Y = (int)ylen.size() + 1;
    Coords C{X, Y, std::vector<long long>(X, 0), std::vector<long long>(Y, 0)};
    for (int i = 0; i + 1 < X; i++) C.drx[i + 1] = C.drx[i] + xlen[i];
    for (int i = 0; i + 1 < Y; i++) C.dry[i + 1] = C.dry[i] + ylen[i];
    std::vector<int> pts(x.size());
    for (size_t i = 0; i < x.size(); i++) {
        if (x[i] < 0 || x[i] >= X || y[i] < 0 || y[i] >= Y) throw std::invalid_argument("point outside the grid");
        pts[i] = x[i] * Y + y[i];
    }
    std::sort(pts.begin(), pts.end());
    if (std::adjacent_find(pts.begin(), pts.end()) != pts.end())
        throw std::invalid_argument("points must be distinct");
    Tree2D t;
    steiner_tree(C, pts, t);
    nodes.clear();
    edges.clear();
    for (int p : t.pos) nodes.emplace_back(p / Y, p % Y);
    for (int i = 0; i < (int)t.pos.size(); i++)
        for (int k = t.adj.off[i]; k < t.adj.off[i + 1]; k++)
            if (i < t.adj.nb[k]) edges.emplace_back(i, t.adj.nb[k]);
}

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
                as += (int64_t)so.adj.nb.size();
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
                    const int a0 = so.adj.off[i], a1 = so.adj.off[i + 1];
                    std::copy(so.adj.nb.begin() + a0, so.adj.nb.begin() + a1, s.adj.begin() + ac);
                    ac += a1 - a0;
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
