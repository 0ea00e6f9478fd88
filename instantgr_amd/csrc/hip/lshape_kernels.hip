// L-shape pattern routing with 3-D layer assignment (src/Lshape_route.hpp:9-153).
//
// Plain form: one thread per net, the reference's launch structure.  The
// per-node DP tables (dist over layer ranges, back-pointers) live in a
// per-batch global scratch slab; each thread walks its tree bottom-up, then
// traces the chosen layer ranges top-down, writes the route and adds its wire
// segments to the demand difference array.  The arithmetic follows the
// reference exactly: FP64 accumulation in the same order (sequential via
// sums, children added in decreasing node order), strict '<' keeping the
// first minimum and '<=' in the layer-range prefix minimum.
#include "igr_ctx.hpp"

namespace igr_dev {

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ double wseg(const GridDev& g, int l, int xmin, int xmax, int ymin, int ymax) {
    return g.presum[gidx(g, l, xmax, ymax)] - g.presum[gidx(g, l, xmin, ymin)];
}

// Difference-array form of one wire segment (src/graph.hpp:105-117).
__device__ __forceinline__ void add_segment(const GridDev& g, int l, int minx, int maxx, int miny, int maxy,
                                            int stamp) {
    atomicAdd(&g.pre[gidx(g, l, minx, miny)], 1);
    if (g.dirty) g.dirty[track_id(g, l, minx, miny)] = 1;
    if (minx == maxx) {
        atomicAdd(&g.pre[gidx(g, l, minx, maxy)], -1);
        g.ts[gidx(g, l, minx, maxy)] = stamp;
    }
    if (miny == maxy) {
        atomicAdd(&g.pre[gidx(g, l, maxx, miny)], -1);
        g.ts[gidx(g, l, maxx, miny)] = stamp;
    }
}

template <int L>
__global__ void __launch_bounds__(TPB) k_lshape_plain(GridDev g, const int* nets, const int64_t* scr_off, int64_t n,
                                                      const int64_t* node_off, const int* node, const int* parent,
                                                      const int64_t* route_off, int* routes, double* dist, int* from,
                                                      int* lrbuf, int stamp) {
    int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    constexpr int LL = L * L;
    const int net = nets[t];
    const int64_t nb = node_off[net];
    const int nn = (int)(node_off[net + 1] - nb);
    const int* nd_pos = node + nb;
    const int* par = parent + nb;
    double* D = dist + scr_off[t] * LL;
    int* F = from + scr_off[t] * LL;
    int* lr = lrbuf + scr_off[t];
    int* r = routes + route_off[net];
    const int X = g.X, Y = g.Y, XY = X * Y;

    for (int i = 0; i < nn; i++) {
        int l = nd_pos[i] / XY, x = nd_pos[i] / Y % X, y = nd_pos[i] % Y;
        double* nd = D + i * LL;
        for (int minl = 0; minl < L; minl++) {
            nd[minl * L + minl] = 0;
            for (int maxl = minl + 1; maxl < L; maxl++)
                nd[minl * L + maxl] = nd[minl * L + maxl - 1] + (double)g.vcost[gidx(g, maxl - 1, x, y)];
            for (int maxl = minl; maxl < L; maxl++)
                if (l < L && (minl > l || maxl < l)) nd[minl * L + maxl] = INF;
        }
    }
    double mcc[L], mcp[LL];
    int cur_from[L];
    for (int i = nn - 1; i >= 1; i--) {
        int x = nd_pos[i] / Y % X, y = nd_pos[i] % Y;
        int p = par[i];
        int px = nd_pos[p] / Y % X, py = nd_pos[p] % Y;
        int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const double* nd = D + i * LL;
        double* pd = D + p * LL;
        int* prev = F + i * LL;
        for (int l = 0; l < L; l++) {
            mcc[l] = mcp[l * L + l] = INF;
            cur_from[l] = 0;
            for (int minl = 0; minl <= l; minl++)
                for (int maxl = l; maxl < L; maxl++)
                    if (nd[minl * L + maxl] < mcc[l]) { mcc[l] = nd[minl * L + maxl]; cur_from[l] = minl * L + maxl; }
        }
        if (x == px || y == py) {
            for (int l = 0; l < L; l++) {
                if (layer_dir(l, g.DIR) == 0 && y != py) continue;
                if (layer_dir(l, g.DIR) == 1 && x != px) continue;
                mcp[l * L + l] = mcc[l] + wseg(g, l, minx, maxx, miny, maxy);
                prev[l * L + l] = cur_from[l] * LL + l * L + l;
            }
        } else {
            for (int curl = 0; curl < L; curl++) {
                double cost = mcc[curl];
                const int cd = layer_dir(curl, g.DIR);
                if (cd == 1) cost += wseg(g, curl, x, x, miny, maxy);
                else cost += wseg(g, curl, minx, maxx, y, y);
                for (int parl = (curl & 1) ^ 1; parl < L; parl += 2) {
                    double cost2 = 0;
                    for (int l = min(curl, parl) + 1; l < max(curl, parl); l++)
                        cost2 += (double)(cd == 1 ? g.vcost[gidx(g, l, x, py)] : g.vcost[gidx(g, l, px, y)]);
                    if (layer_dir(parl, g.DIR) == 1) cost2 += wseg(g, parl, px, px, miny, maxy);
                    else cost2 += wseg(g, parl, minx, maxx, py, py);
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

    int cnt = 1;
    lr[0] = 0;
    for (int minl = 0; minl < L; minl++)
        for (int maxl = minl; maxl < L; maxl++)
            if (D[minl * L + maxl] < D[lr[0]]) lr[0] = minl * L + maxl;
    auto push = [&](int a, int b) { r[cnt] = a; r[cnt + 1] = b; cnt += 2; };
    for (int i = 0; i < nn; i++) {
        const int* prev = F + i * LL;
        if (i > 0) lr[i] = prev[lr[par[i]]] / LL;
        int minl = lr[i] / L, maxl = lr[i] % L;
        int x = nd_pos[i] / Y % X, y = nd_pos[i] % Y;
        if (minl < maxl) {
            push(minl * XY + x * Y + y, maxl * XY + x * Y + y);
            if (g.vc) emit_via(g, minl, maxl, x, y);
        }
        if (i == 0) continue;
        int p = par[i];
        int px = nd_pos[p] / Y % X, py = nd_pos[p] % Y;
        int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        int code = prev[lr[p]];
        int curl = code / L % L, parl = code % L;
        if (px == x || py == y) {
            push(curl * XY + minx * Y + miny, curl * XY + maxx * Y + maxy);
            add_segment(g, curl, minx, maxx, miny, maxy, stamp);
        } else if (layer_dir(curl, g.DIR) == 1) {
            push(curl * XY + x * Y + y, curl * XY + x * Y + py);
            add_segment(g, curl, x, x, miny, maxy, stamp);
            push(curl * XY + x * Y + py, parl * XY + x * Y + py);
            if (g.vc) emit_via(g, curl, parl, x, py);
            push(parl * XY + x * Y + py, parl * XY + px * Y + py);
            add_segment(g, parl, minx, maxx, py, py, stamp);
        } else {
            push(curl * XY + x * Y + y, curl * XY + px * Y + y);
            add_segment(g, curl, minx, maxx, y, y, stamp);
            push(curl * XY + px * Y + y, parl * XY + px * Y + y);
            if (g.vc) emit_via(g, curl, parl, px, y);
            push(parl * XY + px * Y + y, parl * XY + px * Y + py);
            add_segment(g, parl, px, px, miny, maxy, stamp);
        }
    }
    r[0] = cnt;
}

} // namespace

void launch_lshape_plain(igr_ctx* c, const int* d_nets, const int64_t* d_scr_off, int64_t n, int stamp,
                         double bytes) {
    if (n == 0) return;
    KScope ks(c, K_LSHAPE, bytes);
    const unsigned grid = (unsigned)((n + TPB - 1) / TPB);
    NetsDev& N = c->nets;
#define CASE(LV)                                                                                               \
    case LV:                                                                                                   \
        k_lshape_plain<LV><<<grid, TPB, 0, c->stream>>>(c->g, d_nets, d_scr_off, n, N.node_off.p, N.node.p,    \
                                                        N.parent.p, N.route_off.p, N.routes.p,                 \
                                                        c->scratch.dist.p, c->scratch.from.p, c->scratch.lr.p, \
                                                        stamp);                                                \
        break;
    switch (c->g.L) {
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
    default: throw std::runtime_error("unsupported routing layer count");
    }
#undef CASE
    HIP_OK(hipGetLastError());
}

} // namespace igr_dev
