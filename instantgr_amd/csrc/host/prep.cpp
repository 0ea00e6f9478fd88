// Host preparation of the routing passes and the route file writer.
#include "prep.hpp"
#include "par.hpp"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace igr {

void prepare_pattern_inputs(const Grid& g, const Subnets& s, PatternInputs& pi, int threads) {
    const int64_t S = s.count;
    const int Y = g.Y, XY = g.X * g.Y;
    pi = PatternInputs();
    pi.n = S;
    pi.node_off.assign(S + 1, 0);
    for (int64_t i = 0; i < S; i++) pi.node_off[i + 1] = pi.node_off[i] + (s.node_off[i + 1] - s.node_off[i]);
    pi.node.resize(pi.node_off[S]);
    pi.parent.resize(pi.node_off[S]);

    // Footprints: count first, then fill.
    std::vector<int64_t> npt(S), nh(S), nv(S);
    auto foot = [&](int64_t i, int* pt, int* hs, int* vs, int64_t& cp, int64_t& ch, int64_t& cv) {
        const int64_t b = s.node_off[i], n = s.node_off[i + 1] - b;
        cp = ch = cv = 0;
        for (int64_t k = 0; k < n; k++) {
            int pk = s.node[b + k] % XY;
            if (pt) pt[cp] = pk;
            cp++;
        }
        for (int64_t k = 0; k < n; k++) {
            int pk = s.node[b + k] % XY, xi = pk / Y, yi = pk % Y;
            for (int64_t a = s.adj_off[b + k]; a < s.adj_off[b + k + 1]; a++) {
                int j = s.adj[a];
                if (j >= k) continue;
                int pj = s.node[b + j] % XY, xj = pj / Y, yj = pj % Y;
                int minx = std::min(xi, xj), maxx = std::max(xi, xj), miny = std::min(yi, yj), maxy = std::max(yi, yj);
                if (xi != xj && yi != yj) {
                    if (hs) { hs[2 * ch] = minx * Y + miny; hs[2 * ch + 1] = maxx * Y + miny;
                              hs[2 * ch + 2] = minx * Y + maxy; hs[2 * ch + 3] = maxx * Y + maxy; }
                    ch += 2;
                    if (vs) { vs[2 * cv] = minx * Y + miny; vs[2 * cv + 1] = minx * Y + maxy;
                              vs[2 * cv + 2] = maxx * Y + miny; vs[2 * cv + 3] = maxx * Y + maxy; }
                    cv += 2;
                    if (pt) { pt[cp] = xi * Y + yj; pt[cp + 1] = xj * Y + yi; }
                    cp += 2;
                } else if (xi != xj) {
                    if (hs) { hs[2 * ch] = minx * Y + miny; hs[2 * ch + 1] = maxx * Y + miny; }
                    ch += 1;
                    if (pt) { pt[cp] = minx * Y + miny; pt[cp + 1] = maxx * Y + miny; }
                    cp += 2;
                } else if (yi != yj) {
                    if (vs) { vs[2 * cv] = minx * Y + miny; vs[2 * cv + 1] = minx * Y + maxy; }
                    cv += 1;
                    if (pt) { pt[cp] = minx * Y + miny; pt[cp + 1] = minx * Y + maxy; }
                    cp += 2;
                }
            }
        }
    };
    parallel_ranges(S, threads, [&](int64_t a, int64_t b, int) {
        for (int64_t i = a; i < b; i++) foot(i, nullptr, nullptr, nullptr, npt[i], nh[i], nv[i]);
    });
    pi.pt_off.assign(S + 1, 0);
    pi.hs_off.assign(S + 1, 0);
    pi.vs_off.assign(S + 1, 0);
    for (int64_t i = 0; i < S; i++) {
        pi.pt_off[i + 1] = pi.pt_off[i] + npt[i];
        pi.hs_off[i + 1] = pi.hs_off[i] + nh[i];
        pi.vs_off[i + 1] = pi.vs_off[i] + nv[i];
    }
    pi.pt.resize(pi.pt_off[S]);
    pi.hs.resize(2 * pi.hs_off[S]);
    pi.vs.resize(2 * pi.vs_off[S]);

    if (threads < 1) threads = 1;
    std::vector<int> maxn(threads, 0);
    parallel_ranges(S, threads, [&](int64_t a, int64_t b_, int t) {
            std::vector<std::tuple<int, int, int>> st;   // (node, parent node, parent idx)
            for (int64_t i = a; i < b_; i++) {
                int64_t cp, ch, cv;
                foot(i, &pi.pt[pi.pt_off[i]], &pi.hs[2 * pi.hs_off[i]], &pi.vs[2 * pi.vs_off[i]], cp, ch, cv);
                const int64_t b = s.node_off[i], n = s.node_off[i + 1] - b, o = pi.node_off[i];
                maxn[t] = std::max(maxn[t], (int)n);
                int cnt = 0;
                st.clear();
                st.emplace_back(0, -1, -1);
                while (!st.empty()) {
                    auto [x, par, pidx] = st.back();
                    st.pop_back();
                    int idx = cnt++;
                    pi.node[o + idx] = s.node[b + x];
                    pi.parent[o + idx] = pidx;
                    for (int64_t a = s.adj_off[b + x + 1] - 1; a >= s.adj_off[b + x]; a--) {
                        int e = s.adj[a];
                        if (e != par) st.emplace_back(e, x, idx);
                    }
                }
                if (cnt != n) throw std::runtime_error("subnet tree is not connected");
            }
        });
    pi.max_nodes = *std::max_element(maxn.begin(), maxn.end());
    // Routing order: decreasing hpwl, ties in index order (a stable counting
    // sort; the same order as a stable sort on hpwl).
    int maxh = 0;
    for (int64_t i = 0; i < S; i++) maxh = std::max(maxh, s.hpwl[i]);
    std::vector<int64_t> start(maxh + 2, 0);
    for (int64_t i = 0; i < S; i++) start[maxh - s.hpwl[i] + 1]++;
    for (int h = 0; h <= maxh; h++) start[h + 1] += start[h];
    pi.order.resize(S);
    for (int64_t i = 0; i < S; i++) pi.order[start[maxh - s.hpwl[i]]++] = (int)i;
}

void write_routes(const std::string& path, const Design& d, const NetPins& np, const Grid& g,
                  const Subnets& s, const RouteSet& routes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write '" + path + "'");
    std::vector<char> buf(1 << 24);
    size_t n = 0;
    auto flush = [&]() { std::fwrite(buf.data(), 1, n, f); n = 0; };
    auto put = [&](char c) { if (n == buf.size()) flush(); buf[n++] = c; };
    auto num = [&](int v) {
        char t[12]; int k = 0;
        if (v == 0) { put('0'); return; }
        while (v) { t[k++] = char('0' + v % 10); v /= 10; }
        while (k) put(t[--k]);
    };
    const int X = d.X, Y = d.Y, XY = d.X * d.Y;
    auto route = [&](int a, int b) {
        if (a > b) std::swap(a, b);
        num(a / Y % X); put(' '); num(a % Y); put(' '); num(a / XY); put(' ');
        num(b / Y % X); put(' '); num(b % Y); put(' '); num(b / XY); put('\n');
    };
    for (int64_t net = 0; net < d.num_nets(); net++) {
        for (char ch : d.names[net]) put(ch);
        put('\n'); put('('); put('\n');
        for (int64_t p = np.pin_off[net]; p < np.pin_off[net + 1]; p++)
            if (np.pins[p] < XY) route(np.pins[p], np.pins[p] + XY);
        for (int64_t e = np.extra_off[net]; e < np.extra_off[net + 1]; e += 2) route(np.extra[e], np.extra[e + 1]);
        for (int64_t k = s.net_sub_off[net]; k < s.net_sub_off[net + 1]; k++) {
            int sub = s.net_sub[k];
            const int* r = &routes.data[routes.off[sub]];
            for (int j = 1; j < r[0]; j += 2) route(r[j] + XY, r[j + 1] + XY);
        }
        put(')'); put('\n');
    }
    flush();
    std::fclose(f);
    (void)g;
}

} // namespace igr
