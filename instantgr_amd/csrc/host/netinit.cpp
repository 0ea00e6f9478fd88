// Access-point selection: one access point per pin.
//
// Follows src/database.hpp:38-243.  Nets with at most 4 pins enumerate every
// combination of access points and keep the best by (fewest differing x/y
// coordinates between pin pairs, then smallest bounding box, then smallest
// distance to the net centre, then lowest layer sum); larger nets run the
// greedy overlap clustering.  The reference iterates hash sets whose order is
// implementation-defined; here every set is visited in ascending order, which
// fixes one deterministic member of the reference's possible outcomes.
#include "design.hpp"

#include <algorithm>
#include <cstdlib>
#include <thread>
#include <utility>
#include <vector>

namespace igr {
namespace {

struct PinSel {
    std::vector<int> pins;
    std::vector<int> extra;
};

// Per-thread tables of the greedy clustering, reused across nets.  The maps
// over the grid (position -> index, chosen positions per x and per y) are
// cleared after each net at the entries it touched.
struct Scratch {
    std::vector<int> pidmap;           // 2-D position -> index among the net's positions, or -1
    std::vector<int> xc, yc;           // chosen positions per x / per y
    std::vector<int> pos;              // index -> 2-D position
    std::vector<int> apid;             // access point -> position index
    std::vector<int> pdist, cnt;       // per position: 2x distance to the centre; round count
    std::vector<char> in2d;            // position already chosen
};

void init_net(const Design& d, int64_t n, PinSel& out) {
    const int X = d.X, Y = d.Y, XY = d.X * d.Y;
    const int64_t p0 = d.pin_off[n], num = d.pin_off[n + 1] - p0;
    const int* bb = &d.bbox[4 * n];
    const int c2x = bb[0] + bb[1], c2y = bb[2] + bb[3];
    auto apc = [&](int64_t i) { return (int)(d.ap_off[p0 + i + 1] - d.ap_off[p0 + i]); };
    auto apv = [&](int64_t i, int j) { return d.ap[d.ap_off[p0 + i] + j]; };  // j is 0-based
    // Chosen access points and their 2-D positions (sets kept as vectors,
    // sorted and de-duplicated at the end).
    std::vector<int> p, p2D;

    if (num <= 4) {
        std::vector<int> idx(num, 0), best(num, 0);
        long long min_diff = 1000000000, min_metric = 1000000000, min_metric2 = 1000000000, min_metric0 = 1000000000;
        // Lexicographic enumeration, pin 0 most significant (src/database.hpp:46-90).
        for (;;) {
            int tminx = 1000000000, tmaxx = 0, tminy = 1000000000, tmaxy = 0;
            long long diff = 0, metric = 0, metric2 = 0;
            for (int64_t i = 0; i < num; i++) {
                int v = apv(i, idx[i]);
                int xi = v / Y % X, yi = v % Y, li = v / XY;
                tminx = std::min(tminx, xi); tmaxx = std::max(tmaxx, xi);
                tminy = std::min(tminy, yi); tmaxy = std::max(tmaxy, yi);
                metric += std::abs(2 * xi - c2x) + std::abs(2 * yi - c2y);
                metric2 += li;
                for (int64_t j = 0; j < i; j++) {
                    int u = apv(j, idx[j]);
                    diff += (xi != u / Y % X) + (yi != u % Y);
                }
            }
            long long metric0 = tmaxx - tminx + tmaxy - tminy;
            if (diff < min_diff || (diff == min_diff && metric0 < min_metric0)) {
                best = idx; min_diff = diff; min_metric0 = metric0; min_metric = metric; min_metric2 = metric2;
            } else if (diff == min_diff && metric0 == min_metric0 && metric < min_metric) {
                best = idx; min_diff = diff; min_metric = metric; min_metric2 = metric2;
            } else if (diff == min_diff && metric0 == min_metric0 && metric == min_metric && metric2 < min_metric2) {
                best = idx; min_metric2 = metric2;
            }
            int64_t k = num - 1;
            while (k >= 0 && idx[k] + 1 == apc(k)) idx[k--] = 0;
            if (k < 0) break;
            idx[k]++;
        }
        for (int64_t i = 0; i < num; i++) {
            int v = apv(i, best[i]);
            p.push_back(v);
            p2D.push_back(v % XY);
        }
    } else {
        // Greedy overlap clustering (src/database.hpp:92-197) on local tables:
        // each access point gets the index of its 2-D position and of its x
        // and y coordinate among the net's distinct values, so a round's
        // counting is array work instead of hash and tree lookups.
        thread_local Scratch S;
        if (S.pidmap.size() != (size_t)XY) {
            S.pidmap.assign(XY, -1);
            S.xc.assign(X, 0);
            S.yc.assign(Y, 0);
        }
        const int64_t a0 = d.ap_off[p0], A = d.ap_off[p0 + num] - a0;
        S.pos.clear();
        S.apid.resize(A);
        for (int64_t a = 0; a < A; a++) {
            const int v = d.ap[a0 + a] % XY;
            int& id = S.pidmap[v];
            if (id < 0) {
                id = (int)S.pos.size();
                S.pos.push_back(v);
            }
            S.apid[a] = id;
        }
        const int P = (int)S.pos.size();
        S.pdist.resize(P);
        for (int q = 0; q < P; q++) {
            const int v = S.pos[q];
            S.pdist[q] = std::abs(v / Y * 2 - c2x) + std::abs(v % Y * 2 - c2y);
        }
        S.in2d.assign(P, 0);
        S.cnt.resize(P);
        auto cr_of = [&](int q) { return S.xc[S.pos[q] / Y] + S.yc[S.pos[q] % Y]; };
        auto choose = [&](int q) {
            S.in2d[q] = 1;
            p2D.push_back(S.pos[q]);
            S.xc[S.pos[q] / Y]++;
            S.yc[S.pos[q] % Y]++;
        };
        std::vector<char> visited(num, 0);
        auto ap_first = [&](int64_t i) { return d.ap_off[p0 + i] - a0; };
        for (int64_t round = 0; round < num; round++) {
            int max_q = -1;
            int cnt_repeat = 0, max_coord_repeat = -1, min_dist = 1000000000;
            std::fill(S.cnt.begin(), S.cnt.end(), 0);
            for (int64_t i = 0; i < num; i++) {
                if (visited[i]) continue;
                for (int64_t a = ap_first(i); a < ap_first(i + 1); a++) {
                    const int q = S.apid[a];
                    const int c = ++S.cnt[q];
                    const int dist = S.pdist[q];
                    const int cr = cr_of(q);
                    if (c > cnt_repeat) {
                        max_q = q; cnt_repeat = c; max_coord_repeat = cr; min_dist = dist;
                    } else if (c == cnt_repeat && cr > max_coord_repeat) {
                        max_q = q; max_coord_repeat = cr; min_dist = dist;
                    } else if (c == cnt_repeat && cr == max_coord_repeat && dist < min_dist) {
                        max_q = q; min_dist = dist;
                    }
                }
            }
            if (cnt_repeat <= 1) break;
            if (!S.in2d[max_q]) choose(max_q);
            // Every unvisited pin with an access point at the chosen position
            // takes its first such access point.
            for (int64_t i = 0; i < num; i++) {
                if (visited[i]) continue;
                for (int64_t a = ap_first(i); a < ap_first(i + 1); a++)
                    if (S.apid[a] == max_q) {
                        visited[i] = 1;
                        p.push_back(d.ap[a0 + a]);
                        break;
                    }
            }
        }
        for (int64_t i = 0; i < num; i++) {
            if (visited[i]) continue;
            int min_dist = 1000000000, selected = -1, sel_q = -1, max_overlap = 0;
            for (int64_t a = ap_first(i); a < ap_first(i + 1); a++) {
                const int v = d.ap[a0 + a], q = S.apid[a];
                const int dist = std::abs(v / Y % X * 2 - c2x) + std::abs(v % Y * 2 - c2y);
                const int cr = cr_of(q);
                if (cr > max_overlap) { max_overlap = cr; min_dist = dist; selected = v; sel_q = q; }
                else if (cr == max_overlap && dist < min_dist) { min_dist = dist; selected = v; sel_q = q; }
            }
            p.push_back(selected);
            if (!S.in2d[sel_q]) choose(sel_q);
        }
        for (int q = 0; q < P; q++) {
            S.pidmap[S.pos[q]] = -1;
            if (S.in2d[q]) {
                S.xc[S.pos[q] / Y] = 0;
                S.yc[S.pos[q] % Y] = 0;
            }
        }
    }
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
    std::sort(p2D.begin(), p2D.end());
    p2D.erase(std::unique(p2D.begin(), p2D.end()), p2D.end());
    out.pins.clear();
    out.extra.clear();
    if (p2D.size() < p.size()) {
        // Several layers at one position: the pin takes the highest, the
        // lowest and highest become an extra via stack.  Every chosen access
        // point's position is in p2D, so the groups below follow p2D's order.
        std::vector<std::pair<int, int>> byPos(p.size());
        for (size_t k = 0; k < p.size(); k++) byPos[k] = {p[k] % XY, p[k]};
        std::sort(byPos.begin(), byPos.end());
        for (size_t k = 0; k < byPos.size();) {
            size_t e = k;
            while (e < byPos.size() && byPos[e].first == byPos[k].first) e++;
            const int lo = byPos[k].second, hi = byPos[e - 1].second;
            out.pins.push_back(hi);
            if (lo != hi) { out.extra.push_back(lo); out.extra.push_back(hi); }
            k = e;
        }
    } else {
        out.pins = p;
    }
}

} // namespace

void select_access_points(const Design& d, NetPins& np, int threads) {
    const int64_t N = d.num_nets();
    std::vector<PinSel> sel(N);
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t]() {
            for (int64_t n = t; n < N; n += threads) init_net(d, n, sel[n]);
        });
    for (auto& th : pool) th.join();
    np.pin_off.assign(N + 1, 0);
    np.extra_off.assign(N + 1, 0);
    np.hpwl.resize(N);
    for (int64_t n = 0; n < N; n++) {
        np.pin_off[n + 1] = np.pin_off[n] + (int64_t)sel[n].pins.size();
        np.extra_off[n + 1] = np.extra_off[n] + (int64_t)sel[n].extra.size();
        const int* bb = &d.bbox[4 * n];
        np.hpwl[n] = bb[1] - bb[0] + bb[3] - bb[2];
    }
    np.pins.resize(np.pin_off[N]);
    np.extra.resize(np.extra_off[N]);
    for (int64_t n = 0; n < N; n++) {
        std::copy(sel[n].pins.begin(), sel[n].pins.end(), np.pins.begin() + np.pin_off[n]);
        std::copy(sel[n].extra.begin(), sel[n].extra.end(), np.extra.begin() + np.extra_off[n]);
    }
}

} // namespace igr
