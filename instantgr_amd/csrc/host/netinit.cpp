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
#include <map>
#include <set>
#include <thread>
#include <unordered_map>

namespace igr {
namespace {

struct PinSel {
    std::vector<int> pins;
    std::vector<int> extra;
};

void init_net(const Design& d, int64_t n, PinSel& out) {
    const int X = d.X, Y = d.Y, XY = d.X * d.Y;
    const int64_t p0 = d.pin_off[n], num = d.pin_off[n + 1] - p0;
    const int* bb = &d.bbox[4 * n];
    const int c2x = bb[0] + bb[1], c2y = bb[2] + bb[3];
    auto apc = [&](int64_t i) { return (int)(d.ap_off[p0 + i + 1] - d.ap_off[p0 + i]); };
    auto apv = [&](int64_t i, int j) { return d.ap[d.ap_off[p0 + i] + j]; };  // j is 0-based
    std::set<int> p, p2D;

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
            p.insert(v);
            p2D.insert(v % XY);
        }
    } else {
        std::map<int, int> x_counts, y_counts;
        std::vector<char> visited(num, 0);
        for (int64_t round = 0; round < num; round++) {
            int max_pos = -1;
            size_t cnt_repeat = 0;
            int max_coord_repeat = -1, min_dist = 1000000000;
            std::unordered_map<int, std::vector<int>> cnt;
            for (int64_t i = 0; i < num; i++) {
                if (visited[i]) continue;
                for (int j = 0; j < apc(i); j++) {
                    int pin2d = apv(i, j) % XY;
                    auto& lst = cnt[pin2d];
                    lst.push_back((int)i);
                    int dist = std::abs(pin2d / Y * 2 - c2x) + std::abs(pin2d % Y * 2 - c2y);
                    int cr = x_counts[pin2d / Y] + y_counts[pin2d % Y];
                    if (lst.size() > cnt_repeat) {
                        max_pos = pin2d; cnt_repeat = lst.size(); max_coord_repeat = cr; min_dist = dist;
                    } else if (lst.size() == cnt_repeat && cr > max_coord_repeat) {
                        max_pos = pin2d; max_coord_repeat = cr; min_dist = dist;
                    } else if (lst.size() == cnt_repeat && cr == max_coord_repeat && dist < min_dist) {
                        max_pos = pin2d; min_dist = dist;
                    }
                }
            }
            if (cnt_repeat <= 1) break;
            if (!p2D.count(max_pos)) {
                p2D.insert(max_pos);
                x_counts[max_pos / Y % X]++;
                y_counts[max_pos % Y]++;
            }
            for (int pin_id : cnt[max_pos]) {
                if (visited[pin_id]) continue;
                visited[pin_id] = 1;
                for (int j = 0; j < apc(pin_id); j++)
                    if (apv(pin_id, j) % XY == max_pos) { p.insert(apv(pin_id, j)); break; }
            }
        }
        for (int64_t i = 0; i < num; i++) {
            if (visited[i]) continue;
            int min_dist = 1000000000, selected = -1, max_overlap = 0;
            for (int j = 0; j < apc(i); j++) {
                int v = apv(i, j);
                int dist = std::abs(v / Y % X * 2 - c2x) + std::abs(v % Y * 2 - c2y);
                int pin2d = v % XY;
                int cr = x_counts[pin2d / Y] + y_counts[pin2d % Y];
                if (cr > max_overlap) { max_overlap = cr; min_dist = dist; selected = v; }
                else if (cr == max_overlap && dist < min_dist) { min_dist = dist; selected = v; }
            }
            p.insert(selected);
            if (!p2D.count(selected % XY)) {
                p2D.insert(selected % XY);
                x_counts[selected / Y % X]++;
                y_counts[selected % Y]++;
            }
        }
    }
    out.pins.clear();
    out.extra.clear();
    if (p2D.size() < p.size()) {
        std::map<int, int> pmax, pmin;
        for (int e : p) {
            int pos = e % XY;
            auto it = pmax.find(pos);
            if (it == pmax.end()) { pmax[pos] = e; pmin[pos] = e; }
            else { it->second = std::max(it->second, e); pmin[pos] = std::min(pmin[pos], e); }
        }
        for (int e : p2D) {
            out.pins.push_back(pmax[e]);
            if (pmin[e] != pmax[e]) { out.extra.push_back(pmin[e]); out.extra.push_back(pmax[e]); }
        }
    } else {
        out.pins.assign(p.begin(), p.end());
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
