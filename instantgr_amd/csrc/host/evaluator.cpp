// Route evaluator: connectivity check and FP64 score, following
// run/evaluator.cpp:458-633.
//
// Demand per gcell is 2 * wires + stacked vias in half-track units, held in 16
// bits as the reference's bitfield does (run/evaluator.cpp:132).  A via adds
// demand on its lower layer only where the same net has no wire on that cell
// (run/evaluator.cpp:635-675).  Overflow cost sums over every gcell of every
// routing layer: exp(0.25 * (demand - 2 * capacity)) when capacity > 0.001,
// otherwise exp(0.75 * demand) when demand > 0, weighted per layer.
// Unlike the reference program, an open net is counted instead of ending the
// run.
#include "prep.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace igr {

int evaluate_routes(const Design& d, const std::string& out_path, Score& sc) {
    const int L = d.L, X = d.X, Y = d.Y;
    const size_t XY = (size_t)X * Y, N = (size_t)L * XY;
    auto at = [&](int z, int x, int y) { return (size_t)z * XY + (size_t)x * Y + y; };
    std::vector<int> flag(N, -1), wire(N, 0), svia(N, 0);
    std::vector<long long> vias(L, 0);
    std::unordered_map<std::string, int> by_name;
    by_name.reserve(d.names.size() * 2);
    for (size_t i = 0; i < d.names.size(); i++) by_name[d.names[i]] = (int)i;
    const int nn = (int)d.names.size();

    FILE* f = std::fopen(out_path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open '" + out_path + "'");
    std::vector<char> line(1 << 16);
    int net = -1;
    bool bad = false;
    struct V { int x, y, z; };
    std::vector<V> via_loc;
    std::vector<V> stack;
    sc = Score();
    while (std::fgets(line.data(), (int)line.size(), f)) {
        size_t len = std::strlen(line.data());
        while (len && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
        if (net < 0) {
            auto it = by_name.find(std::string(line.data()));
            if (it == by_name.end()) { std::fclose(f); throw std::runtime_error("unknown net in route file"); }
            net = it->second;
            bad = false;
        } else if (line[0] == '(') {
        } else if (line[0] == ')') {
            // Stacked vias: once per net and location, where the net has no wire.
            for (const V& p : via_loc) {
                size_t k = at(p.z, p.x, p.y);
                if (flag[k] == net) continue;
                flag[k] = net;
                if (d.dir[p.z] == 0) {
                    if (p.x > 0 && p.x < X - 1) { svia[at(p.z, p.x - 1, p.y)]++; svia[k]++; }
                    else if (p.x > 0) svia[at(p.z, p.x - 1, p.y)] += 2;
                    else if (p.x < X - 1) svia[k] += 2;
                } else {
                    if (p.y > 0 && p.y < Y - 1) { svia[at(p.z, p.x, p.y - 1)]++; svia[k]++; }
                    else if (p.y > 0) svia[at(p.z, p.x, p.y - 1)] += 2;
                    else if (p.y < Y - 1) svia[k] += 2;
                }
            }
            for (const V& p : via_loc) { flag[at(p.z, p.x, p.y)] = net; flag[at(p.z + 1, p.x, p.y)] = net; }
            // Connectivity from the first pin's access points.
            bool ok = !bad;
            if (ok) {
                const int mark = net, traced = net + nn;
                int64_t p0 = d.pin_off[net];
                stack.clear();
                for (int64_t a = d.ap_off[p0]; a < d.ap_off[p0 + 1]; a++) {
                    int v = d.ap[a];
                    V q{v / Y % X, v % Y, (int)(v / XY)};
                    size_t k = at(q.z, q.x, q.y);
                    if (flag[k] == mark) { flag[k] = traced; stack.push_back(q); }
                }
                auto visit = [&](int z, int x, int y) {
                    size_t k = at(z, x, y);
                    if (flag[k] == mark) { flag[k] = traced; stack.push_back(V{x, y, z}); }
                };
                while (!stack.empty()) {
                    V q = stack.back();
                    stack.pop_back();
                    if (q.z > 0) {
                        if (d.dir[q.z] == 0) {
                            if (q.x > 0) visit(q.z, q.x - 1, q.y);
                            if (q.x < X - 1) visit(q.z, q.x + 1, q.y);
                        } else {
                            if (q.y > 0) visit(q.z, q.x, q.y - 1);
                            if (q.y < Y - 1) visit(q.z, q.x, q.y + 1);
                        }
                    }
                    if (q.z > 0) visit(q.z - 1, q.x, q.y);
                    if (q.z < L - 1) visit(q.z + 1, q.x, q.y);
                }
                for (int64_t p = p0 + 1; p < d.pin_off[net + 1] && ok; p++) {
                    bool conn = false;
                    for (int64_t a = d.ap_off[p]; a < d.ap_off[p + 1]; a++) {
                        int v = d.ap[a];
                        if (flag[at(v / (int)XY, v / Y % X, v % Y)] == traced) { conn = true; break; }
                    }
                    ok = conn;
                }
            }
            if (!ok) sc.opens++;
            net = -1;
            via_loc.clear();
        } else {
            int xl, yl, zl, xh, yh, zh;
            if (std::sscanf(line.data(), "%d %d %d %d %d %d", &xl, &yl, &zl, &xh, &yh, &zh) != 6) {
                std::fclose(f);
                throw std::runtime_error("malformed route line");
            }
            // A line leaving the grid opens its net and touches no cell (the
            // device evaluator's rule for cells outside the grid).  A via
            // written top-down adds nothing, as in the reference's unsigned
            // layer loop (run/evaluator.cpp:525-530).
            if (xl < 0 || yl < 0 || zl < 0 || xh < 0 || yh < 0 || zh < 0 || xl >= X || xh >= X || yl >= Y ||
                yh >= Y || zl >= L || zh >= L) {
                bad = true;
                continue;
            }
            if (zh != zl) {
                if (xh == xl && yh == yl) {
                    for (int z = zl; z < zh; z++) { vias[z]++; via_loc.push_back(V{xl, yl, z}); }
                } else bad = true;
            } else if (zl == 0) {
                bad = true;
            } else if (d.dir[zl] == 0) {
                if (xh > xl && yh == yl) {
                    for (int x = xl; x < xh; x++) { flag[at(zl, x, yl)] = net; wire[at(zl, x, yl)]++; }
                    flag[at(zl, xh, yl)] = net;
                } else bad = true;
            } else {
                if (yh > yl && xh == xl) {
                    for (int y = yl; y < yh; y++) { flag[at(zl, xl, y)] = net; wire[at(zl, xl, y)]++; }
                    flag[at(zl, xl, yh)] = net;
                } else bad = true;
            }
        }
    }
    std::fclose(f);

    for (int z = 0; z < L; z++) {
        if (z == 0) { sc.via += (double)vias[z] * d.via_cost; continue; }
        unsigned long long total_wl = 0;
        double layer_of = 0;
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++) {
                size_t k = at(z, x, y);
                unsigned demand = (unsigned)(2 * wire[k] + svia[k]) & 0xFFFFu;
                double cap = d.cap[k];
                if (cap > 0.001) {
                    double over = double(demand) - 2 * cap;
                    layer_of += std::exp(over / 2 * 0.5);
                } else if (cap >= 0 && demand > 0) {
                    layer_of += std::exp(1.5 * double(demand) * 0.5);
                }
                if (d.dir[z] == 0) total_wl += (unsigned long long)wire[k] * (x + 1 < X ? d.xlen[x] : 0);
                else total_wl += (unsigned long long)wire[k] * (y + 1 < Y ? d.ylen[y] : 0);
            }
        sc.of += layer_of * d.of_weight[z];
        sc.via += (double)vias[z] * d.via_cost;
        sc.wl += (double)total_wl * d.wire_cost;
    }
    sc.total = sc.of + sc.via + sc.wl;
    return sc.opens == 0 ? 0 : 1;
}

} // namespace igr
