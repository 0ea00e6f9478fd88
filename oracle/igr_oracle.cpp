// CPU oracle for the InstantGR global-routing hot path.
//
// TEST INFRASTRUCTURE ONLY (see igr_oracle.h).  Each function restates the
// reference file:line it cites, sequentially, with the reference's arithmetic
// (float storage, double intermediates, the same summation orders).  Compiled
// with -ffp-contract=off so no multiply-add is fused.
//
// Parity status (what pins each function):
//   orc_expf           own function (replaces CUDA __expf, whose values are
//                      not available): pinned against libm expf (ulp bound),
//                      special values and monotonicity in tests.
//   update_cost        closed forms + invariants (monotone in demand, clamp at
//                      1e12, wcost >= wire cost) in tests.
//   compute_presum     Sklansky tree pinned by exact integer inputs and an FP64
//                      sequential-sum bound; tie to src/graph.hpp:405-430.
//   commits            integer exactness: demand recount from routes.
//   lshape_net         brute force over all L choices and layer assignments
//                      on small trees (tests/test_oracle.py).
//   schedule           invariants (every net batched, priority, no collision
//                      of a batch's point cells).
//   ripup/congestion   closed-form recount in tests.
// The reference ships no tests or golden vectors for any of these.
#include "igr_oracle.h"
#include "orc_expf.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double INF = 1e22;  // src/database_cuda.hpp:53

struct Net {
    std::vector<int> node, par;          // DFS pre-order
    std::vector<int> pts, hs, vs;        // scheduler footprint
};

} // namespace

struct orc_ctx {
    int L, X, Y, DIR, XY;
    double unit_wire, unit_via;
    std::vector<double> w;
    std::vector<int> xlen, ylen;
    std::vector<float> cap, demand, wcost, vcost;
    std::vector<double> presum;
    std::vector<int> pre_demand, timestamp;
    std::vector<Net> nets;
    std::vector<std::vector<int>> routes;
    int stamp = 0;

    int64_t N() const { return (int64_t)L * X * Y; }
    int64_t IDX(int l, int x, int y) const { return (int64_t)l * X * Y + (int64_t)x * Y + y; }
    int dir(int l) const { return (l & 1) ^ DIR; }
};

float orc_expf_scalar(float x) { return orc_expf(x); }
void orc_expf_many(const float* in, float* out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = orc_expf(in[i]);
}

orc_ctx* orc_create(const orc_grid* g) {
    orc_ctx* c = new orc_ctx();
    c->L = g->L; c->X = g->X; c->Y = g->Y; c->DIR = g->DIR;
    c->XY = std::max(g->X, g->Y);
    c->unit_wire = g->unit_wire; c->unit_via = g->unit_via;
    c->w.assign(g->w, g->w + g->L);
    c->xlen.assign(g->xlen, g->xlen + g->X - 1);
    c->ylen.assign(g->ylen, g->ylen + g->Y - 1);
    int64_t n = c->N();
    c->cap.assign(g->cap, g->cap + n);
    c->demand.assign(n, 0.0f);
    c->wcost.assign(n, 0.0f);
    c->vcost.assign(n, 0.0f);
    c->presum.assign(n, 0.0);
    c->pre_demand.assign(n, 0);
    c->timestamp.assign(n, 0);
    return c;
}
void orc_destroy(orc_ctx* c) { delete c; }

void orc_set_demand(orc_ctx* c, const float* d) { std::memcpy(c->demand.data(), d, sizeof(float) * c->N()); }
void orc_get_demand(const orc_ctx* c, float* d) { std::memcpy(d, c->demand.data(), sizeof(float) * c->N()); }
void orc_get_timestamp(const orc_ctx* c, int* t) { std::memcpy(t, c->timestamp.data(), sizeof(int) * c->N()); }
void orc_get_costs(const orc_ctx* c, float* wc, float* vc, double* ps) {
    if (wc) std::memcpy(wc, c->wcost.data(), sizeof(float) * c->N());
    if (vc) std::memcpy(vc, c->vcost.data(), sizeof(float) * c->N());
    if (ps) std::memcpy(ps, c->presum.data(), sizeof(double) * c->N());
}
int orc_stamp(const orc_ctx* c) { return c->stamp; }

// ----------------------------------------------------------------------------
// Logistic congestion cost (src/graph.hpp:23-38, of_cost_scale = 1 by
// src/global.h:32).
// ----------------------------------------------------------------------------
static double of_cost_scaled(float capacity, float demand) {
    if ((double)capacity > 0.001) {
        double a = 0.5 * (double)(float)(demand - capacity);
        double b = 1.0 * 0.5 * (double)(float)(demand - capacity);
        return (double)orc_expf((float)std::fmin(a, b));
    }
    if (demand > 0) {
        double a = 1.5 * (double)demand;
        double b = 1.0 * 1.5 * (double)demand;
        return (double)orc_expf((float)std::fmin(a, b));
    }
    return 0;
}

static double incremental_of_cost(const orc_ctx* c, int l, int x, int y, double inc, double base = 0) {
    int64_t idx = c->IDX(l, x, y);
    float d_inc = (float)((double)c->demand[idx] + inc + base);
    float d_base = (float)((double)c->demand[idx] + base);
    return std::fmin(1e12, (of_cost_scaled(c->cap[idx], d_inc) - of_cost_scaled(c->cap[idx], d_base)) * c->w[l]);
}

// src/graph.hpp:40-47 and :52-71 (cross_points is zero: its allocation is
// only byte-cleared in the reference, src/database_cuda.hpp:2340-2341; this
// restatement fixes it to 0).
void orc_update_cost(orc_ctx* c) {
    for (int l = 0; l < c->L; l++)
        for (int x = 0; x < c->X; x++)
            for (int y = 0; y < c->Y; y++) {
                int64_t idx = c->IDX(l, x, y);
                int d = c->dir(l);
                float wv = (float)incremental_of_cost(c, l, x, y, 1);
                if (d == 0 && x < c->X - 1) wv = (float)((double)wv + c->unit_wire * (double)c->xlen[x]);
                if (d == 1 && y < c->Y - 1) wv = (float)((double)wv + c->unit_wire * (double)c->ylen[y]);
                c->wcost[idx] = wv;
                const double base = 0;
                float vv;
                if (d == 1) {
                    if (y == 0) vv = (float)incremental_of_cost(c, l, x, y, 1, base);
                    else if (y == c->Y - 1) vv = (float)incremental_of_cost(c, l, x, y - 1, 1, base);
                    else vv = (float)(incremental_of_cost(c, l, x, y, 0.5, base) + incremental_of_cost(c, l, x, y - 1, 0.5, base));
                } else {
                    if (x == 0) vv = (float)incremental_of_cost(c, l, x, y, 1, base);
                    else if (x == c->X - 1) vv = (float)incremental_of_cost(c, l, x - 1, y, 1, base);
                    else vv = (float)(incremental_of_cost(c, l, x, y, 0.5, base) + incremental_of_cost(c, l, x - 1, y, 0.5, base));
                }
                c->vcost[idx] = (float)((double)vv + c->unit_via);
            }
}

// Sklansky scan over s[0..n): at level d every index with bit d set adds the
// last element of the preceding 2^d block (src/graph.hpp:412-417).
template <typename T>
static void sklansky(std::vector<T>& s) {
    const int n = (int)s.size();
    for (int d = 0; (1 << d) < n; d++)
        for (int i = 0; i < n; i++)
            if (i >> d & 1) s[i] += s[(i >> d << d) - 1];
}

// Per-track exclusive FP64 prefix of wcost (src/graph.hpp:405-430).
void orc_compute_presum(orc_ctx* c) {
    std::vector<double> s;
    for (int l = 0; l < c->L; l++) {
        if (c->dir(l) == 1) {
            for (int x = 0; x < c->X; x++) {
                s.assign(c->Y, 0.0);
                for (int y = 0; y < c->Y - 1; y++) s[y + 1] = c->wcost[c->IDX(l, x, y)];
                sklansky(s);
                for (int y = 0; y < c->Y; y++) c->presum[c->IDX(l, x, y)] = s[y];
            }
        } else {
            for (int y = 0; y < c->Y; y++) {
                s.assign(c->X, 0.0);
                for (int x = 0; x < c->X - 1; x++) s[x + 1] = c->wcost[c->IDX(l, x, y)];
                sklansky(s);
                for (int x = 0; x < c->X; x++) c->presum[c->IDX(l, x, y)] = s[x];
            }
        }
    }
}

// In-place exclusive int prefix per track (src/graph.hpp:74-99).
static void presum_general(orc_ctx* c, std::vector<int>& a) {
    std::vector<int> s;
    for (int l = 0; l < c->L; l++) {
        if (c->dir(l) == 1) {
            for (int x = 0; x < c->X; x++) {
                s.assign(c->Y, 0);
                for (int y = 0; y < c->Y - 1; y++) s[y + 1] = a[c->IDX(l, x, y)];
                sklansky(s);
                for (int y = 0; y < c->Y; y++) a[c->IDX(l, x, y)] = s[y];
            }
        } else {
            for (int y = 0; y < c->Y; y++) {
                s.assign(c->X, 0);
                for (int x = 0; x < c->X - 1; x++) s[x + 1] = a[c->IDX(l, x, y)];
                sklansky(s);
                for (int x = 0; x < c->X; x++) a[c->IDX(l, x, y)] = s[x];
            }
        }
    }
}

// src/graph.hpp:105-117.
static void add_wire_segment(orc_ctx* c, int l, int minx, int maxx, int miny, int maxy, int stamp, int K = 1) {
    c->pre_demand[c->IDX(l, minx, miny)] += K;
    if (minx == maxx) {
        c->pre_demand[c->IDX(l, minx, maxy)] -= K;
        c->timestamp[c->IDX(l, minx, maxy)] = stamp;
    }
    if (miny == maxy) {
        c->pre_demand[c->IDX(l, maxx, miny)] -= K;
        c->timestamp[c->IDX(l, maxx, miny)] = stamp;
    }
}

// src/graph.hpp:118-137.
static void batch_wire_update(orc_ctx* c, int stamp) {
    presum_general(c, c->pre_demand);
    for (int l = 0; l < c->L; l++)
        for (int x = 0; x < c->X; x++)
            for (int y = 0; y < c->Y; y++) {
                int d = c->dir(l);
                if (d == 0 && x + 1 < c->X && c->pre_demand[c->IDX(l, x + 1, y)] > 0) {
                    c->demand[c->IDX(l, x, y)] += (float)c->pre_demand[c->IDX(l, x + 1, y)];
                    c->timestamp[c->IDX(l, x, y)] = stamp;
                }
                if (d == 1 && y + 1 < c->Y && c->pre_demand[c->IDX(l, x, y + 1)] > 0) {
                    c->demand[c->IDX(l, x, y)] += (float)c->pre_demand[c->IDX(l, x, y + 1)];
                    c->timestamp[c->IDX(l, x, y)] = stamp;
                }
            }
    std::fill(c->pre_demand.begin(), c->pre_demand.end(), 0);
}

// src/graph.hpp:496-543.
static void commit_via_demand(orc_ctx* c, const std::vector<int>& r, int stamp, int K) {
    for (int i = 1; i < r[0]; i += 2) {
        int a = r[i], b = r[i + 1];
        int la = a / c->X / c->Y, lb = b / c->X / c->Y;
        if (la == lb) continue;
        int x = a / c->Y % c->X, y = a % c->Y;
        int minl = std::min(la, lb), maxl = std::max(la, lb);
        for (int l = minl; l < maxl; l++) {
            if (c->timestamp[c->IDX(l, x, y)] >= stamp) continue;
            float k = (float)K, h = 0.5f * (float)K;
            if (c->dir(l) == 1) {
                if (y == 0) c->demand[c->IDX(l, x, y)] += k;
                else if (y == c->Y - 1) c->demand[c->IDX(l, x, y - 1)] += k;
                else { c->demand[c->IDX(l, x, y - 1)] += h; c->demand[c->IDX(l, x, y)] += h; }
            } else {
                if (x == 0) c->demand[c->IDX(l, x, y)] += k;
                else if (x == c->X - 1) c->demand[c->IDX(l, x - 1, y)] += k;
                else { c->demand[c->IDX(l, x - 1, y)] += h; c->demand[c->IDX(l, x, y)] += h; }
            }
        }
    }
}

// src/graph.hpp:468-495 (rip-up with K = -1).
static void commit_wire_demand(orc_ctx* c, const std::vector<int>& r, int stamp, int K) {
    for (int i = 1; i < r[0]; i += 2) {
        int a = r[i], b = r[i + 1];
        if (a / c->X / c->Y != b / c->X / c->Y) continue;
        int l = a / c->X / c->Y;
        int x0 = a / c->Y % c->X, y0 = a % c->Y, x1 = b / c->Y % c->X, y1 = b % c->Y;
        float k = (float)K;
        if (x0 < x1) for (int x = x0; x < x1; x++) { c->demand[c->IDX(l, x, y0)] += k; c->timestamp[c->IDX(l, x, y0)] = stamp; }
        if (x1 < x0) for (int x = x1; x < x0; x++) { c->demand[c->IDX(l, x, y0)] += k; c->timestamp[c->IDX(l, x, y0)] = stamp; }
        if (y0 < y1) for (int y = y0; y < y1; y++) { c->demand[c->IDX(l, x0, y)] += k; c->timestamp[c->IDX(l, x0, y)] = stamp; }
        if (y1 < y0) for (int y = y1; y < y0; y++) { c->demand[c->IDX(l, x0, y)] += k; c->timestamp[c->IDX(l, x0, y)] = stamp; }
        c->timestamp[c->IDX(l, x0, y0)] = stamp;
        c->timestamp[c->IDX(l, x0, y1)] = stamp;
        c->timestamp[c->IDX(l, x1, y0)] = stamp;
        c->timestamp[c->IDX(l, x1, y1)] = stamp;
    }
}

int orc_set_nets(orc_ctx* c, int64_t n, const int64_t* node_off, const int* nodes, const int* parents,
                 const int64_t* pt_off, const int* pts, const int64_t* hs_off, const int* hs,
                 const int64_t* vs_off, const int* vs) {
    c->nets.assign(n, Net());
    c->routes.assign(n, std::vector<int>(1, 1));
    for (int64_t i = 0; i < n; i++) {
        Net& t = c->nets[i];
        t.node.assign(nodes + node_off[i], nodes + node_off[i + 1]);
        t.par.assign(parents + node_off[i], parents + node_off[i + 1]);
        if (pt_off) t.pts.assign(pts + pt_off[i], pts + pt_off[i + 1]);
        if (hs_off) t.hs.assign(hs + 2 * hs_off[i], hs + 2 * hs_off[i + 1]);
        if (vs_off) t.vs.assign(vs + 2 * vs_off[i], vs + 2 * vs_off[i + 1]);
        for (size_t k = 1; k < t.par.size(); k++)
            if (t.par[k] < 0 || t.par[k] >= (int)k) return -1;
    }
    return 0;
}

// ----------------------------------------------------------------------------
// L-shape pattern routing of one net (src/Lshape_route.hpp:9-153).  Returns the
// root cost; writes the route into r and, when commit is set, the wire demand
// differences (pre_demand, timestamp) for batch_wire_update.
// ----------------------------------------------------------------------------
static double lshape_net(orc_ctx* c, int net, std::vector<int>& r, int stamp, bool commit) {
    const Net& t = c->nets[net];
    const int L = c->L, X = c->X, Y = c->Y, LL = L * L;
    const int n = (int)t.node.size();
    std::vector<double> dist((size_t)n * LL, 0.0);
    std::vector<int> from((size_t)n * LL, 0);
    std::vector<int> layer_range(n, 0);
    auto wseg = [&](int l, int xmin, int xmax, int ymin, int ymax) {
        return c->presum[c->IDX(l, xmax, ymax)] - c->presum[c->IDX(l, xmin, ymin)];
    };
    for (int i = 0; i < n; i++) {
        int l = t.node[i] / X / Y, x = t.node[i] / Y % X, y = t.node[i] % Y;
        double* nd = &dist[(size_t)i * LL];
        for (int minl = 0; minl < L; minl++) {
            nd[minl * L + minl] = 0;
            for (int maxl = minl + 1; maxl < L; maxl++)
                nd[minl * L + maxl] = nd[minl * L + maxl - 1] + (double)c->vcost[c->IDX(maxl - 1, x, y)];
            for (int maxl = minl; maxl < L; maxl++)
                if (l < L && (minl > l || maxl < l)) nd[minl * L + maxl] = INF;
        }
    }
    std::vector<double> mcc(L), mcp(LL);
    std::vector<int> cur_from(L);
    for (int i = n - 1; i >= 1; i--) {
        int x = t.node[i] / Y % X, y = t.node[i] % Y;
        int p = t.par[i];
        int px = t.node[p] / Y % X, py = t.node[p] % Y;
        int minx = std::min(x, px), maxx = std::max(x, px), miny = std::min(y, py), maxy = std::max(y, py);
        double* nd = &dist[(size_t)i * LL];
        double* pd = &dist[(size_t)p * LL];
        int* prev = &from[(size_t)i * LL];
        for (int l = 0; l < L; l++) {
            mcc[l] = mcp[l * L + l] = INF;
            cur_from[l] = 0;
            for (int minl = 0; minl <= l; minl++)
                for (int maxl = l; maxl < L; maxl++)
                    if (nd[minl * L + maxl] < mcc[l]) { mcc[l] = nd[minl * L + maxl]; cur_from[l] = minl * L + maxl; }
        }
        if (x == px || y == py) {
            for (int l = 0; l < L; l++) {
                if (c->dir(l) == 0 && y != py) continue;
                if (c->dir(l) == 1 && x != px) continue;
                mcp[l * L + l] = mcc[l] + wseg(l, minx, maxx, miny, maxy);
                prev[l * L + l] = cur_from[l] * LL + l * L + l;
            }
        } else {
            for (int curl = 0; curl < L; curl++) {
                double cost = mcc[curl];
                if (c->dir(curl) == 1) cost += wseg(curl, x, x, miny, maxy);
                else cost += wseg(curl, minx, maxx, y, y);
                for (int parl = (curl & 1) ^ 1; parl < L; parl += 2) {
                    double cost2 = 0;
                    for (int l = std::min(curl, parl) + 1; l < std::max(curl, parl); l++)
                        cost2 += (double)(c->dir(curl) == 1 ? c->vcost[c->IDX(l, x, py)] : c->vcost[c->IDX(l, px, y)]);
                    if (c->dir(parl) == 1) cost2 += wseg(parl, px, px, miny, maxy);
                    else cost2 += wseg(parl, minx, maxx, py, py);
                    if (cost + cost2 < mcp[parl * L + parl]) {
                        mcp[parl * L + parl] = cost + cost2;
                        prev[parl * L + parl] = cur_from[curl] * LL + curl * L + parl;
                    }
                }
            }
        }
        for (int minl = 0; minl < L; minl++) {
            for (int maxl = minl + 1; maxl < L; maxl++) {
                if (mcp[minl * L + maxl - 1] <= mcp[maxl * L + maxl]) {
                    mcp[minl * L + maxl] = mcp[minl * L + maxl - 1];
                    prev[minl * L + maxl] = prev[minl * L + maxl - 1];
                } else {
                    mcp[minl * L + maxl] = mcp[maxl * L + maxl];
                    prev[minl * L + maxl] = prev[maxl * L + maxl];
                }
            }
            for (int maxl = minl; maxl < L; maxl++) pd[minl * L + maxl] += mcp[minl * L + maxl];
        }
    }
    r.assign(1, 1);
    auto push = [&](int a, int b) { r.push_back(a); r.push_back(b); r[0] += 2; };
    layer_range[0] = 0;
    for (int minl = 0; minl < L; minl++)
        for (int maxl = minl; maxl < L; maxl++)
            if (dist[minl * L + maxl] < dist[layer_range[0]]) layer_range[0] = minl * L + maxl;
    double best = dist[layer_range[0]];
    for (int i = 0; i < n; i++) {
        const int* prev = &from[(size_t)i * LL];
        if (i > 0) layer_range[i] = prev[layer_range[t.par[i]]] / L / L;
        int minl = layer_range[i] / L, maxl = layer_range[i] % L;
        int x = t.node[i] / Y % X, y = t.node[i] % Y;
        if (minl < maxl) push((int)c->IDX(minl, x, y), (int)c->IDX(maxl, x, y));
        if (i == 0) continue;
        int p = t.par[i];
        int px = t.node[p] / Y % X, py = t.node[p] % Y;
        int minx = std::min(x, px), maxx = std::max(x, px), miny = std::min(y, py), maxy = std::max(y, py);
        int code = prev[layer_range[p]];
        int curl = code / L % L, parl = code % L;
        if (px == x || py == y) {
            push((int)c->IDX(curl, minx, miny), (int)c->IDX(curl, maxx, maxy));
            if (commit) add_wire_segment(c, curl, minx, maxx, miny, maxy, stamp);
        } else if (c->dir(curl) == 1) {
            push((int)c->IDX(curl, x, y), (int)c->IDX(curl, x, py));
            if (commit) add_wire_segment(c, curl, x, x, miny, maxy, stamp);
            push((int)c->IDX(curl, x, py), (int)c->IDX(parl, x, py));
            push((int)c->IDX(parl, x, py), (int)c->IDX(parl, px, py));
            if (commit) add_wire_segment(c, parl, minx, maxx, py, py, stamp);
        } else {
            push((int)c->IDX(curl, x, y), (int)c->IDX(curl, px, y));
            if (commit) add_wire_segment(c, curl, minx, maxx, y, y, stamp);
            push((int)c->IDX(curl, px, y), (int)c->IDX(parl, px, y));
            push((int)c->IDX(parl, px, y), (int)c->IDX(parl, px, py));
            if (commit) add_wire_segment(c, parl, px, px, miny, maxy, stamp);
        }
    }
    return best;
}

double orc_lshape_probe(orc_ctx* c, int net, int* buf, int cap) {
    std::vector<int> r;
    double best = lshape_net(c, net, r, 0, false);
    for (int i = 0; i < (int)r.size() && i < cap; i++) buf[i] = r[i];
    return best;
}

// One batch of the pattern pass (src/Lshape_route.hpp:235-242).
void orc_lshape_batch(orc_ctx* c, const int* nets, int64_t n) {
    int stamp = ++c->stamp;
    orc_update_cost(c);
    orc_compute_presum(c);
    for (int64_t i = 0; i < n; i++) lshape_net(c, nets[i], c->routes[nets[i]], stamp, true);
    batch_wire_update(c, stamp);
    for (int64_t i = 0; i < n; i++) commit_via_demand(c, c->routes[nets[i]], stamp, 1);
}

// ----------------------------------------------------------------------------
// Batch scheduler (src/database_cuda.hpp:288-434,437-631) with deterministic
// claim semantics: the reference re-reads its own point cell right after the
// atomicMax and abandons the net when a higher net owns it, which races with
// other threads (src/database_cuda.hpp:313-318).  Here a round first claims
// every active net's point cells, then nets owning all their point cells claim
// their neighbour cells and sampled segment cells; nets that lost a point
// claim nothing more.  Commit, revert and clear follow the reference.
// ----------------------------------------------------------------------------
namespace {
int schedule_impl(orc_ctx* c, const int* order, int64_t n, int mode, const int64_t* pt_off, const int* pt,
                  int* batch_of) {
    const int X = c->X, Y = c->Y;
    const int sparsity = (mode == 1 || mode == 2) ? 2 : 15;
    std::vector<int> ids(n);
    for (int64_t i = 0; i < n; i++) ids[i] = order[n - 1 - i];
    std::vector<std::vector<int>> own;
    if (pt_off) {
        own.resize(n);
        for (int64_t i = 0; i < n; i++) own[i].assign(pt + pt_off[n - 1 - i], pt + pt_off[n - i]);
    }
    auto pts_of = [&](int64_t i) -> const std::vector<int>& { return pt_off ? own[i] : c->nets[ids[i]].pts; };
    std::vector<int> batch(n, -1), status(n, 0), map((size_t)X * Y, -1);
    const int xs[4] = {0, 0, -1, 1}, ys[4] = {0, -1, 0, 0};
    int current = 0;
    auto active = [&](int64_t i) { return status[i] != 2 && batch[i] == -1; };
    for (;;) {
        bool have_commit = false;
        std::fill(map.begin(), map.end(), -1);
        std::fill(status.begin(), status.end(), 0);
        for (int round = 0; round < 30; round++) {
            bool have_commit2 = false;
            for (int64_t i = 0; i < n; i++) {
                if (!active(i)) continue;
                status[i] = 1;
                for (int p : pts_of(i)) map[p] = std::max(map[p], (int)i);
            }
            std::vector<int64_t> viable;
            for (int64_t i = 0; i < n; i++) {
                if (!active(i)) continue;
                bool ok = true;
                for (int p : pts_of(i))
                    if (map[p] != (int)i) { ok = false; break; }
                if (ok) viable.push_back(i);
            }
            for (int64_t i : viable) {
                const Net& t = c->nets[ids[i]];
                for (int p : pts_of(i)) {
                    int x = p / Y, y = p % Y;
                    for (int j = 1; j < 4; j++) {
                        int xn = x + xs[j], yn = y + ys[j];
                        if (xn >= 0 && xn < X && yn >= 0 && yn < Y) map[(size_t)xn * Y + yn] = std::max(map[(size_t)xn * Y + yn], (int)i);
                    }
                }
                for (size_t k = 0; k < t.hs.size(); k += 2) {
                    int s = t.hs[k], e = t.hs[k + 1];
                    int y = s % Y;
                    for (int x = s / Y; x <= e / Y; x += sparsity) map[(size_t)x * Y + y] = std::max(map[(size_t)x * Y + y], (int)i);
                }
                for (size_t k = 0; k < t.vs.size(); k += 2) {
                    int s = t.vs[k], e = t.vs[k + 1];
                    int x = s / Y;
                    for (int y = s % Y; y <= e % Y; y += sparsity) map[(size_t)x * Y + y] = std::max(map[(size_t)x * Y + y], (int)i);
                }
            }
            for (int64_t i = 0; i < n; i++) {
                if (!active(i)) continue;
                bool ok = true;
                for (int p : pts_of(i))
                    if (map[p] != (int)i) { ok = false; break; }
                if (ok) { batch[i] = current; have_commit = have_commit2 = true; }
            }
            if (!have_commit2) break;
            for (int64_t i = 0; i < n; i++) {
                if (status[i] != 1 || batch[i] != -1) continue;
                for (int p : pts_of(i)) {
                    int pre = map[p];
                    if (pre != -1 && batch[pre] != -1) { status[i] = 2; break; }
                }
            }
            for (size_t q = 0; q < map.size(); q++)
                if (map[q] >= 0 && status[map[q]] == 2) map[q] = -1;
        }
        if (!have_commit) break;
        current++;
    }
    for (int64_t i = 0; i < n; i++) batch_of[n - 1 - i] = batch[i];
    return current;
}
} // namespace

int orc_schedule(orc_ctx* c, const int* order, int64_t n, int mode, int* batch_of) {
    return schedule_impl(c, order, n, mode, nullptr, nullptr, batch_of);
}

// Pattern pass (src/Lshape_route.hpp:155-257): schedule, reverse the batch
// list, route batches in that order; within a batch nets follow the
// scheduler's index order.
int orc_pattern_route(orc_ctx* c, const int* order, int64_t n) {
    std::vector<int> batch_of(n);
    int nb = orc_schedule(c, order, n, 0, batch_of.data());
    std::vector<std::vector<int>> batches(nb);
    for (int64_t i = n - 1; i >= 0; i--) batches[batch_of[i]].push_back(order[i]);
    for (int b = nb - 1; b >= 0; b--) orc_lshape_batch(c, batches[b].data(), (int64_t)batches[b].size());
    return nb;
}

// Grid state and routes taken over from elsewhere (a routed device context),
// so that a later pass can be replayed from the same state.
int orc_set_state(orc_ctx* c, const float* demand, const int* ts, int stamp, const int64_t* route_off,
                  const int* routes) {
    std::memcpy(c->demand.data(), demand, sizeof(float) * c->N());
    std::memcpy(c->timestamp.data(), ts, sizeof(int) * c->N());
    c->stamp = stamp;
    if (route_off) {
        for (size_t i = 0; i < c->routes.size(); i++) {
            const int* r = routes + route_off[i];
            if (r[0] < 1 || r[0] > route_off[i + 1] - route_off[i]) return -1;
            c->routes[i].assign(r, r + r[0]);
        }
    }
    return 0;
}

// One batch of given routes committed as the reference commits a routed
// batch (src/Lshape_route.hpp:240-241: wire differences, batch_wire_update,
// then commit_via_demand), or ripped up as the DAG pass does
// (src/Lshape_route_detour.hpp:1047-1048: commit_wire_demand, then
// commit_via_demand, both with K = -1).  Each route is buf[0] = next free
// index, then cell pairs; wires run along one layer between the pair's cells.
void orc_commit_routes(orc_ctx* c, int64_t n, const int64_t* route_off, const int* routes, int K) {
    const int stamp = ++c->stamp;
    std::vector<std::vector<int>> rs(n);
    for (int64_t i = 0; i < n; i++) rs[i].assign(routes + route_off[i], routes + route_off[i] + routes[route_off[i]]);
    if (K < 0) {
        for (auto& r : rs) commit_wire_demand(c, r, stamp, -1);
        for (auto& r : rs) commit_via_demand(c, r, stamp, -1);
        return;
    }
    for (auto& r : rs)
        for (size_t k = 1; k + 1 < (size_t)r[0]; k += 2) {
            const int a = std::min(r[k], r[k + 1]), b = std::max(r[k], r[k + 1]);
            const int la = a / c->X / c->Y, lb = b / c->X / c->Y;
            if (la != lb) continue;
            add_wire_segment(c, la, a / c->Y % c->X, b / c->Y % c->X, a % c->Y, b % c->Y, stamp);
        }
    batch_wire_update(c, stamp);
    for (auto& r : rs) commit_via_demand(c, r, stamp, 1);
}

int orc_get_route(const orc_ctx* c, int net, int* buf, int cap) {
    const auto& r = c->routes[net];
    int n = std::min((int)r.size(), cap);
    for (int i = 0; i < n; i++) buf[i] = r[i];
    return (int)r.size();
}

// ----------------------------------------------------------------------------
// Rip-up marking (src/graph.hpp:432-466, :588-601).
// ----------------------------------------------------------------------------
void orc_ripup_mark(orc_ctx* c, int th, uint8_t* is_of) {
    std::vector<int> of(c->N(), 0);
    for (int l = 0; l < c->L; l++)
        for (int x = 0; x < c->X; x++)
            for (int y = 0; y < c->Y; y++) {
                int64_t idx = c->IDX(l, x, y);
                int d = c->dir(l);
                if ((d == 0 && x + 1 < c->X) || (d == 1 && y + 1 < c->Y))
                    of[idx] = (c->cap[idx] + (float)th <= c->demand[idx]) ? 1 : 0;
            }
    presum_general(c, of);
    for (size_t net = 0; net < c->routes.size(); net++) {
        const auto& r = c->routes[net];
        is_of[net] = 0;
        int cnt = 0;
        for (int i = 1; i < r[0]; i += 2) {
            int l = r[i] / c->X / c->Y, x0 = r[i] / c->Y % c->X, y0 = r[i] % c->Y;
            int x1 = r[i + 1] / c->Y % c->X, y1 = r[i + 1] % c->Y;
            if (x0 != x1) cnt += of[c->IDX(l, std::max(x0, x1), y0)] - of[c->IDX(l, std::min(x0, x1), y0)];
            if (y0 != y1) cnt += of[c->IDX(l, x0, std::max(y0, y1))] - of[c->IDX(l, x0, std::min(y0, y1))];
            if (cnt > 10) { is_of[net] = 1; break; }
        }
    }
}

// Congestion view (src/graph.hpp:545-584): max over layers of the overflow
// (demand - capacity where positive), then FP32 Sklansky exclusive prefixes
// along x (per row y) and along y (per column x).
void orc_congestion_view(orc_ctx* c, float* xsum, float* ysum) {
    const int X = c->X, Y = c->Y;
    std::vector<float> v((size_t)X * Y, 0.0f);
    for (int l = 0; l < c->L; l++)
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++) {
                int64_t idx = c->IDX(l, x, y);
                double resource = (double)(float)(c->cap[idx] - c->demand[idx]);
                if (resource < 0) {
                    float& t = v[(size_t)x * Y + y];
                    t = (float)std::fmax((double)t, -resource);
                }
            }
    std::vector<float> s;
    for (int y = 0; y < Y; y++) {
        s.assign(X, 0.0f);
        for (int x = 0; x < X - 1; x++) s[x + 1] = v[(size_t)x * Y + y];
        sklansky(s);
        for (int x = 0; x < X; x++) xsum[(size_t)x * Y + y] = s[x];
    }
    for (int x = 0; x < X; x++) {
        s.assign(Y, 0.0f);
        for (int y = 0; y < Y - 1; y++) s[y + 1] = v[(size_t)x * Y + y];
        sklansky(s);
        for (int y = 0; y < Y; y++) ysum[(size_t)x * Y + y] = s[y];
    }
}

// ----------------------------------------------------------------------------
// Rip-up and reroute over detour DAGs (src/Lshape_route_detour.hpp).
// ----------------------------------------------------------------------------
namespace {

// Ordered-int packing of an FP32 cost and a node id into one 64-bit key
// (src/Lshape_route_detour.hpp:83-104): smaller key = smaller cost, then
// smaller id.
inline uint64_t pack_key(float cost, uint32_t id) {
    int32_t i;
    std::memcpy(&i, &cost, 4);
    int32_t o = i >= 0 ? i : i ^ 0x7FFFFFFF;
    return ((uint64_t)(int64_t)o << 32) | (uint64_t)id;
}
inline void unpack_key(uint64_t k, double* cost, uint32_t* id) {
    int32_t o = (int32_t)(k >> 32);
    int32_t i = o >= 0 ? o : o ^ 0x7FFFFFFF;
    float f;
    std::memcpy(&f, &i, 4);
    *cost = (double)f;
    *id = (uint32_t)(k & 0xFFFFFFFFULL);
}
// Initial slot key (src/Lshape_route_detour.hpp:106).
constexpr uint64_t INIT_KEY = 7208259049189261824ULL;
// Id multiplier and the "straight edge" marker: the reference uses 100 and 82
// with 9 routing layers; any multiplier above every choice code keeps the
// (node, choice) order, and the marker must exceed every layer-pair code.
constexpr uint32_t ID_MUL = 1000;

struct DagView {
    int n;                   // nodes
    const int* node;         // encoded positions
    const int64_t* par_off;  // n+1 (local offsets)
    const int* par;          // local parent index per edge
    const int* slot;         // child slot per edge
    const int* child_num;
    const int* depth;
};

// One net's DAG route (src/Lshape_route_detour.hpp:123-195 precompute_cost_edge,
// :281-406 Lshape_route_node_layer_change_cuda, :520-694
// get_routing_tree_layer_change_cuda).  Writes the route and the wire
// difference array.
void dag_route_net(orc_ctx* c, const DagView& d, std::vector<int>& r, int stamp) {
    const int L = c->L, X = c->X, Y = c->Y, XY = X * Y, ML = L + 1, LL = L * L;
    const int SENT = LL + 1;
    const int n = d.n;
    const int64_t E = d.par_off[n];
    std::vector<int64_t> csum(n + 1, 0);
    for (int i = 0; i < n; i++) csum[i + 1] = csum[i] + d.child_num[i];
    auto pos = [&](int i, int& x, int& y) { x = d.node[i] / Y % X; y = d.node[i] % Y; };
    auto wseg = [&](int l, int xmin, int xmax, int ymin, int ymax) {
        return c->presum[c->IDX(l, xmax, ymax)] - c->presum[c->IDX(l, xmin, ymin)];
    };
    // Edge list: edge k belongs to child node edge_node[k].
    std::vector<int> edge_node(E);
    for (int i = 0; i < n; i++)
        for (int64_t k = d.par_off[i]; k < d.par_off[i + 1]; k++) edge_node[k] = i;
    // precompute_cost_edge
    std::vector<double> cost_edges((size_t)E * LL, INF);
    std::vector<int> best_change((size_t)E * LL, -1);
    for (int64_t e = 0; e < E; e++) {
        int to = d.par[e], from = edge_node[e];
        int tox, toy, fromx, fromy;
        pos(to, tox, toy);
        pos(from, fromx, fromy);
        for (int fl = 0; fl < L; fl++) {
            if ((c->dir(fl) == 0 && fromy != toy) || (c->dir(fl) == 1 && fromx != tox)) {
                for (int tl = 0; tl < L; tl++) cost_edges[e * LL + fl * L + tl] = INF;
                continue;
            }
            int distance = std::abs(tox - fromx) + std::abs(toy - fromy);
            if (distance == 0) continue;
            int dd = distance / std::min(distance, 20);
            bool vert = fromx == tox;
            int s0 = vert ? std::min(fromy, toy) : std::min(fromx, tox);
            int s1 = vert ? std::max(fromy, toy) : std::max(fromx, tox);
            for (int tl = 0; tl < L; tl++) {
                if ((tl % 2) != (fl % 2)) continue;
                double mn = INF;
                int decide = -1;
                for (int sp = s0 + dd; sp < s1; sp += dd) {
                    int sx = vert ? fromx : sp, sy = vert ? sp : fromy;
                    double via = 0;
                    for (int ll = std::min(fl, tl); ll < std::max(fl, tl); ll++) via += (double)c->vcost[c->IDX(ll, sx, sy)];
                    double cost = via;
                    if (vert)
                        cost += (wseg(fl, fromx, fromx, std::min(fromy, sy), std::max(fromy, sy)) +
                                 wseg(tl, tox, tox, std::min(toy, sy), std::max(toy, sy)));
                    else
                        cost += (wseg(fl, std::min(fromx, sx), std::max(fromx, sx), fromy, fromy) +
                                 wseg(tl, std::min(tox, sx), std::max(tox, sx), toy, toy));
                    if (cost < mn) { mn = cost; decide = sp; }
                }
                cost_edges[e * LL + fl * L + tl] = mn;
                best_change[e * LL + fl * L + tl] = decide;
            }
        }
    }
    // Bottom-up DP by decreasing depth.
    std::vector<double> costs((size_t)n * ML, INF);
    std::vector<uint64_t> packed((size_t)csum[n] * ML, INIT_KEY), best((size_t)csum[n] * ML, 0);
    std::vector<int> byd(n);
    for (int i = 0; i < n; i++) byd[i] = i;
    std::stable_sort(byd.begin(), byd.end(), [&](int a, int b) { return d.depth[a] > d.depth[b]; });
    std::vector<double> mcc;
    std::vector<uint64_t> bp;
    for (int node : byd) {
        int x, y;
        pos(node, x, y);
        const int fixed_low = 1 + d.node[node] / XY;
        const int fixed_high = fixed_low == ML ? 0 : fixed_low;
        const int nc = d.child_num[node];
        uint64_t* cur_best = &best[csum[node] * ML];
        const uint64_t* cur_packed = &packed[csum[node] * ML];
        double* cst = &costs[(size_t)node * ML];
        mcc.assign(nc, INF);
        bp.assign(nc, 0);
        for (int low = 1; low <= fixed_low; low++) {
            for (int k = 0; k < nc; k++) mcc[k] = INF;
            double via = 0;
            for (int li = low; li < L + 1; li++) {
                if (li > low) via += (double)c->vcost[c->IDX(li - 2, x, y)];
                for (int k = 0; k < nc; k++) {
                    double cc;
                    uint32_t id;
                    unpack_key(cur_packed[k * ML + li], &cc, &id);
                    if (cc < mcc[k]) { mcc[k] = cc; bp[k] = (uint64_t)id * ML + li; }
                }
                if (li >= fixed_high) {
                    double cost = via;
                    for (int k = 0; k < nc; k++) cost += mcc[k];
                    if (cost < INF && cost < cst[li]) {
                        cst[li] = cost;
                        for (int k = 0; k < nc; k++) cur_best[k * ML + li] = bp[k];
                    }
                }
            }
            for (int li = L + 1 - 2; li >= low; li--)
                if (cst[li + 1] < cst[li]) {
                    cst[li] = cst[li + 1];
                    for (int k = 0; k < nc; k++) cur_best[k * ML + li] = cur_best[k * ML + li + 1];
                }
        }
        for (int64_t e = d.par_off[node]; e < d.par_off[node + 1]; e++) {
            int p = d.par[e], ci = d.slot[e];
            uint64_t* ppk = &packed[csum[p] * ML];
            int px, py;
            pos(p, px, py);
            int minx = std::min(x, px), maxx = std::max(x, px), miny = std::min(y, py), maxy = std::max(y, py);
            std::vector<uint64_t> cache(ML, 0);
            for (int t = 1; t < ML; t++) cache[t] = ppk[ci * ML + t];
            for (int layer = 1; layer < ML; layer++) {
                if (((layer & 1) ^ c->DIR) == 1 && y != py) continue;
                if (((layer & 1) ^ c->DIR) == 0 && x != px) continue;
                double base = cst[layer];
                for (int t = 1 + (1 - layer % 2); t < ML; t += 2) {
                    double cost2;
                    int choice = SENT;
                    if (t == layer) cost2 = base + wseg(layer - 1, minx, maxx, miny, maxy);
                    else if (std::abs(x - px) + std::abs(y - py) > 4) {
                        int q = (layer - 1) * L + (t - 1);
                        double ce = cost_edges[e * LL + q];
                        if (ce < INF) { cost2 = base + ce; choice = q; }
                        else continue;
                    } else continue;
                    uint64_t key = pack_key((float)cost2, (uint32_t)node * ID_MUL + (uint32_t)choice);
                    if (key < cache[t]) cache[t] = key;
                }
            }
            for (int t = 1; t < ML; t++) ppk[ci * ML + t] = std::min(ppk[ci * ML + t], cache[t]);
        }
    }
    // Top-down extraction by increasing depth.
    // Route pieces are collected per node (parent edge first, then the via
    // stack) and concatenated in node order, one fixed order for the
    // reference's racy append.
    std::vector<int> lo(n, -1);
    std::vector<std::vector<int>> em(n);
    int cur_node = 0;
    auto push = [&](int a, int b) { em[cur_node].push_back(a); em[cur_node].push_back(b); };
    for (auto it = byd.rbegin(); it != byd.rend(); ++it) {
        const int node = *it;
        cur_node = node;
        int cx, cy;
        pos(node, cx, cy);
        const int l = d.node[node] / XY;
        const uint64_t* cur_best = &best[csum[node] * ML];
        if (d.par_off[node + 1] == d.par_off[node]) {
            int min_layer = 0;
            double min_cost = costs[(size_t)node * ML];
            for (int layer = 1; layer < ML; layer++)
                if (costs[(size_t)node * ML + layer] < min_cost) { min_cost = costs[(size_t)node * ML + layer]; min_layer = layer; }
            lo[node] = min_layer;
        } else {
            int par_layer = -1, par_idx = -1;
            int64_t par_edge = -1;
            for (int64_t e = d.par_off[node]; e < d.par_off[node + 1]; e++) {
                int pn = d.par[e];
                if (lo[pn] > 0) {
                    par_idx = pn;
                    par_edge = e;
                    par_layer = lo[pn];
                    uint64_t path = best[(csum[pn] + d.slot[e]) * ML + par_layer];
                    uint64_t child = path / ML / ID_MUL, choice = (path / ML) % ID_MUL;
                    if (child == (uint64_t)node) {
                        lo[node] = choice == (uint64_t)SENT ? (int)(path % ML) : (int)(choice / L) + 1;
                        int px, py;
                        pos(pn, px, py);
                        if (px == cx && cy != py) {
                            if (choice == (uint64_t)SENT) {
                                int ly = lo[node] - 1;
                                add_wire_segment(c, ly, px, px, std::min(py, cy), std::max(py, cy), stamp);
                                push((int)c->IDX(ly, px, std::min(py, cy)), (int)c->IDX(ly, px, std::max(py, cy)));
                            } else {
                                int fl = (int)choice / L, tl = (int)choice % L;
                                int ch = best_change[e * LL + fl * L + tl];
                                add_wire_segment(c, fl, cx, cx, std::min(cy, ch), std::max(cy, ch), stamp);
                                add_wire_segment(c, tl, cx, cx, std::min(py, ch), std::max(py, ch), stamp);
                                push((int)c->IDX(fl, cx, std::min(cy, ch)), (int)c->IDX(fl, cx, std::max(cy, ch)));
                                push((int)c->IDX(std::min(fl, tl), cx, ch), (int)c->IDX(std::max(fl, tl), cx, ch));
                                push((int)c->IDX(tl, cx, std::min(py, ch)), (int)c->IDX(tl, cx, std::max(py, ch)));
                            }
                        } else if (py == cy && cx != px) {
                            if (choice == (uint64_t)SENT) {
                                int ly = lo[node] - 1;
                                add_wire_segment(c, ly, std::min(px, cx), std::max(px, cx), py, py, stamp);
                                push((int)c->IDX(ly, std::min(px, cx), py), (int)c->IDX(ly, std::max(px, cx), py));
                            } else {
                                int fl = (int)choice / L, tl = (int)choice % L;
                                int ch = best_change[e * LL + fl * L + tl];
                                add_wire_segment(c, fl, std::min(cx, ch), std::max(cx, ch), cy, cy, stamp);
                                add_wire_segment(c, tl, std::min(px, ch), std::max(px, ch), cy, cy, stamp);
                                push((int)c->IDX(fl, std::min(cx, ch), cy), (int)c->IDX(fl, std::max(cx, ch), cy));
                                push((int)c->IDX(std::min(fl, tl), ch, cy), (int)c->IDX(std::max(fl, tl), ch, cy));
                                push((int)c->IDX(tl, std::min(px, ch), cy), (int)c->IDX(tl, std::max(px, ch), cy));
                            }
                        }
                        break;
                    } else {
                        lo[node] = -1;
                    }
                }
            }
            if (par_layer == -1) { lo[node] = -1; continue; }
            uint64_t path = best[(csum[par_idx] + d.slot[par_edge]) * ML + par_layer];
            if (path / ML / ID_MUL != (uint64_t)node) { lo[node] = -1; continue; }
        }
        // A source without any finite layer has no route (the reference would
        // index layer -1); children without a recorded layer are skipped.
        if (lo[node] < 1) continue;
        int minl = l + 1;
        int maxl = (l + 1) == ML ? 1 : minl;
        minl = std::min(minl, lo[node]);
        maxl = std::max(maxl, lo[node]);
        for (int k = 0; k < d.child_num[node]; k++) {
            int lc = (int)(cur_best[k * ML + lo[node]] % ML);
            if (lc < 1) continue;
            minl = std::min(lc, minl);
            maxl = std::max(lc, maxl);
        }
        if (minl < maxl) push((int)c->IDX(minl - 1, cx, cy), (int)c->IDX(maxl - 1, cx, cy));
    }
    r.assign(1, 1);
    for (int i = 0; i < n; i++) r.insert(r.end(), em[i].begin(), em[i].end());
    r[0] = (int)r.size();
}

// Scheduler core shared by the pattern pass and the DAG pass: points come from
// pt_off/pt per listed net when given, from the pattern footprint otherwise.
int schedule_impl(orc_ctx* c, const int* order, int64_t n, int mode, const int64_t* pt_off, const int* pt,
                  int* batch_of);

} // namespace

int orc_dag_pass(orc_ctx* c, const int* nets, int64_t n, int mode, const int64_t* node_off, const int* node,
                 const int64_t* par_off, const int* par, const int* slot, const int* child_num, const int* depth,
                 const int64_t* pt_off, const int* pt) {
    if (n == 0) return 0;
    std::vector<int> batch_of(n);
    int nb = schedule_impl(c, nets, n, mode, pt_off, pt, batch_of.data());
    std::vector<std::vector<int64_t>> batches(nb);  // list positions
    for (int64_t i = n - 1; i >= 0; i--) batches[batch_of[i]].push_back(i);
    for (int b = nb - 1; b >= 0; b--) {
        const auto& B = batches[b];
        int s1 = ++c->stamp;
        for (int64_t i : B) commit_wire_demand(c, c->routes[nets[i]], s1, -1);
        for (int64_t i : B) commit_via_demand(c, c->routes[nets[i]], s1, -1);
        int s2 = ++c->stamp;
        orc_update_cost(c);
        orc_compute_presum(c);
        for (int64_t i : B) {
            int64_t b0 = node_off[i];
            std::vector<int64_t> poff(node_off[i + 1] - b0 + 1);
            for (size_t k = 0; k < poff.size(); k++) poff[k] = par_off[b0 + k] - par_off[b0];
            DagView v{(int)(node_off[i + 1] - b0), node + b0, poff.data(), par + par_off[b0], slot + par_off[b0],
                      child_num + b0, depth + b0};
            dag_route_net(c, v, c->routes[nets[i]], s2);
        }
        batch_wire_update(c, s2);
        for (int64_t i : B) commit_via_demand(c, c->routes[nets[i]], s2, 1);
    }
    return nb;
}
