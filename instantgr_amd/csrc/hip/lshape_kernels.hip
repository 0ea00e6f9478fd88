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


// ---------------------------------------------------------------------------
// Batched form: a 16-lane group per net, four nets per wave.  The per-node DP
// tables (cost over layer ranges, back-pointers) live in LDS in triangular
// form (minl <= maxl); lane l owns layer l: it computes the child's best cost
// through layer l, the edge costs on layer l, and the prefix minimum of the
// layer-range row minl = l.  The root's range is the group's min/argmin
// (first minimum in (minl, maxl) order).  Every value is computed with the
// same FP64 operations in the same order as the thread-per-net kernel above,
// so routes and costs are bit-identical.  NC bounds the nodes per net of a
// launch (the host groups a batch's nets by node count); G groups per block.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int tri_row(int L, int minl) { return minl * L - minl * (minl - 1) / 2; }

template <int L, int NC, int G>
__global__ void __launch_bounds__(16 * G) k_lshape_wave(GridDev g, const int* __restrict__ nets, int64_t n,
                                                         const int64_t* __restrict__ node_off,
                                                         const int* __restrict__ node,
                                                         const int* __restrict__ parent,
                                                         const int64_t* __restrict__ route_off, int* routes,
                                                         int stamp) {
    constexpr int LL = L * L, T = L * (L + 1) / 2, LV = L > 1 ? L - 1 : 1;
    __shared__ double sD[G][NC * T];
    __shared__ uint16_t sF[G][NC * T];
    __shared__ int sx[G][NC], sy[G][NC], sl[G][NC], sp[G][NC], slr[G][NC], scode[G][NC];
    __shared__ float sV[G][NC * LV];
    __shared__ double smcc[G][L], swc[G][L], sva[G][L], svb[G][L], sdg[G][L];
    __shared__ int scf[G][L], spv[G][L];
    __shared__ int snn[G];
    const int gi = threadIdx.x >> 4, lane = threadIdx.x & 15;
    const int64_t t = (int64_t)blockIdx.x * G + gi;
    int nn = 0, net = 0;
    int64_t nb = 0;
    if (t < n) {
        net = nets[t];
        nb = node_off[net];
        nn = (int)(node_off[net + 1] - nb);
    }
    if (lane == 0) snn[gi] = nn;
    const int X = g.X, Y = g.Y, XY = X * Y;
    for (int k = lane; k < nn; k += 16) {
        const int p = node[nb + k];
        sl[gi][k] = p / XY;
        sx[gi][k] = p / Y % X;
        sy[gi][k] = p % Y;
        sp[gi][k] = parent[nb + k];
    }
    __syncthreads();
    int nmax = 0;
    for (int k = 0; k < G; k++) nmax = max(nmax, snn[k]);
    for (int k = lane; k < nn * (L - 1); k += 16) {
        const int i = k / (L - 1), l = k - i * (L - 1);
        sV[gi][k] = g.vcost[gidx(g, l, sx[gi][i], sy[gi][i])];
    }
    __syncthreads();
    double* D = sD[gi];
    uint16_t* F = sF[gi];
    // Via-stack cost of every layer range at every node; ranges that miss a
    // pin's layer are INF (src/Lshape_route.hpp:36-48).
    if (lane < L) {
        const int minl = lane, r0 = tri_row(L, minl);
        for (int i = 0; i < nn; i++) {
            double* nd = D + i * T + r0;
            const int pl = sl[gi][i];
            const bool pin = pl < L;
            double acc = 0;
            nd[0] = (pin && minl != pl) ? INF : 0.0;
            for (int maxl = minl + 1; maxl < L; maxl++) {
                acc = acc + (double)sV[gi][i * LV + maxl - 1];
                nd[maxl - minl] = (pin && (minl > pl || maxl < pl)) ? INF : acc;
            }
        }
    }
    __syncthreads();
    // Edges bottom-up, children in decreasing node order.
    for (int i = nmax - 1; i >= 1; i--) {
        const bool act = i < nn && lane < L;
        int x = 0, y = 0, px = 0, py = 0, p = 0;
        if (i < nn) {
            x = sx[gi][i];
            y = sy[gi][i];
            p = sp[gi][i];
            px = sx[gi][p];
            py = sy[gi][p];
        }
        const bool straight = x == px || y == py;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const int l = lane;
        double wc = 0, wp = 0, va = 0, vb = 0;
        if (act) {
            const int d = layer_dir(l, g.DIR);
            if (straight) {
                if ((d == 0 && y == py) || (d == 1 && x == px)) wc = wseg(g, l, minx, maxx, miny, maxy);
            } else {
                wc = d == 1 ? wseg(g, l, x, x, miny, maxy) : wseg(g, l, minx, maxx, y, y);
                wp = d == 1 ? wseg(g, l, px, px, miny, maxy) : wseg(g, l, minx, maxx, py, py);
                va = (double)g.vcost[gidx(g, l, x, py)];
                vb = (double)g.vcost[gidx(g, l, px, y)];
            }
            // Best child range through layer l (first minimum, strict <).
            const double* nd = D + i * T;
            double m = INF;
            int cf = 0;
            for (int minl = 0; minl <= l; minl++) {
                const double* row = nd + tri_row(L, minl) - minl;
                for (int maxl = l; maxl < L; maxl++)
                    if (row[maxl] < m) {
                        m = row[maxl];
                        cf = minl * L + maxl;
                    }
            }
            smcc[gi][l] = m;
            scf[gi][l] = cf;
            swc[gi][l] = wc;
            sva[gi][l] = va;
            svb[gi][l] = vb;
        }
        __syncthreads();
        if (act) {
            double best = INF;
            int bp = 0;
            if (straight) {
                const int d = layer_dir(l, g.DIR);
                if (!((d == 0 && y != py) || (d == 1 && x != px))) {
                    best = smcc[gi][l] + wc;
                    bp = scf[gi][l] * LL + l * L + l;
                }
            } else {
                // Lane l is the parent-side layer; child-side layers of the
                // other parity in increasing order, strict < keeps the first.
                for (int curl = (l & 1) ^ 1; curl < L; curl += 2) {
                    const double cost = smcc[gi][curl] + swc[gi][curl];
                    const double* vv = layer_dir(curl, g.DIR) == 1 ? sva[gi] : svb[gi];
                    double cost2 = 0;
                    for (int k = min(curl, l) + 1; k < max(curl, l); k++) cost2 += vv[k];
                    cost2 += wp;
                    if (cost + cost2 < best) {
                        best = cost + cost2;
                        bp = scf[gi][curl] * LL + curl * L + l;
                    }
                }
            }
            sdg[gi][l] = best;
            spv[gi][l] = bp;
        }
        __syncthreads();
        if (act) {
            // Prefix minimum over maxl ('<=' keeps the lower range) added to
            // the parent's row minl = l; back-pointers of the child.
            const int r0 = tri_row(L, l);
            double* pd = D + p * T + r0;
            uint16_t* fr = F + i * T + r0;
            double m = sdg[gi][l];
            int pv = spv[gi][l];
            pd[0] += m;
            fr[0] = (uint16_t)pv;
            for (int maxl = l + 1; maxl < L; maxl++) {
                const double dm = sdg[gi][maxl];
                if (!(m <= dm)) {
                    m = dm;
                    pv = spv[gi][maxl];
                }
                pd[maxl - l] += m;
                fr[maxl - l] = (uint16_t)pv;
            }
        }
        __syncthreads();
    }
    // Root range: min/argmin over the lanes' rows, first in (minl, maxl) order.
    double m = __builtin_inf();
    int code = 1 << 30;
    if (lane < L && nn > 0) {
        const double* row = D + tri_row(L, lane) - lane;
        m = row[lane];
        code = lane * L + lane;
        for (int maxl = lane + 1; maxl < L; maxl++)
            if (row[maxl] < m) {
                m = row[maxl];
                code = lane * L + maxl;
            }
    }
    for (int off = 8; off >= 1; off >>= 1) {
        const double om = __shfl_xor(m, off, 16);
        const int oc = __shfl_xor(code, off, 16);
        if (om < m || (om == m && oc < code)) {
            m = om;
            code = oc;
        }
    }
    // Traceback top-down (node order is pre-order: parents first).
    if (lane == 0 && nn > 0) {
        slr[gi][0] = code;
        scode[gi][0] = 0;
        for (int i = 1; i < nn; i++) {
            const int lp = slr[gi][sp[gi][i]];
            const int c = F[i * T + tri_row(L, lp / L) + lp % L - lp / L];
            scode[gi][i] = c;
            slr[gi][i] = c / LL;
        }
    }
    __syncthreads();
    if (t >= n) return;
    // Route pairs in node order: the node's via stack, then the edge to its
    // parent (one piece, or wire-via-wire for an L).
    int* r = routes + route_off[net];
    int base = 0;
    for (int c0 = 0; c0 < nn; c0 += 16) {
        const int i = c0 + lane;
        int cntp = 0;
        int minl = 0, maxl = 0, x = 0, y = 0, px = 0, py = 0, cd = 0;
        if (i < nn) {
            minl = slr[gi][i] / L;
            maxl = slr[gi][i] % L;
            x = sx[gi][i];
            y = sy[gi][i];
            cntp = minl < maxl;
            if (i > 0) {
                const int p = sp[gi][i];
                px = sx[gi][p];
                py = sy[gi][p];
                cd = scode[gi][i];
                cntp += (px == x || py == y) ? 1 : 3;
            }
        }
        int incl = cntp;
        for (int off = 1; off < 16; off <<= 1) {
            const int v = __shfl_up(incl, off, 16);
            if (lane >= off) incl += v;
        }
        const int total = __shfl(incl, 15, 16);
        if (i < nn) {
            int k = 1 + 2 * (base + incl - cntp);
            auto push = [&](int a, int b) {
                r[k] = a;
                r[k + 1] = b;
                k += 2;
            };
            if (minl < maxl) {
                push(minl * XY + x * Y + y, maxl * XY + x * Y + y);
                if (g.vc) emit_via(g, minl, maxl, x, y);
            }
            if (i > 0) {
                const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
                const int curl = cd / L % L, parl = cd % L;
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
        }
        base += total;
    }
    if (lane == 0) r[0] = 1 + 2 * base;
}


// LDS bytes of one k_lshape_wave block.
constexpr int lswave_lds(int L, int NC, int G) {
    return G * (NC * (L * (L + 1) / 2) * 10 + NC * 24 + NC * (L > 1 ? L - 1 : 1) * 4 + L * 48 + 4);
}
constexpr int lswave_groups(int L, int NC) {
    return lswave_lds(L, NC, 4) <= 60 * 1024 ? 4 : lswave_lds(L, NC, 2) <= 60 * 1024 ? 2 : 1;
}

template <int L, int NC, int G>
void lswave_launch_g(igr_ctx* c, const int* d_nets, int64_t n, int stamp) {
    NetsDev& N = c->nets;
    const unsigned grid = (unsigned)((n + G - 1) / G);
    k_lshape_wave<L, NC, G><<<grid, 16 * G, 0, c->stream>>>(c->g, d_nets, n, N.node_off.p, N.node.p, N.parent.p,
                                                            N.route_off.p, N.routes.p, stamp);
}

// Nets per workgroup: the tuned count, capped by what the LDS holds, or the
// largest that fits.
template <int L, int NC>
void lswave_launch(igr_ctx* c, const int* d_nets, int64_t n, int stamp) {
    constexpr int GMAX = lswave_groups(L, NC);
    const int G = c->tune.lshape_groups ? std::min(c->tune.lshape_groups, GMAX) : GMAX;
    if (G >= 4 && GMAX >= 4) lswave_launch_g<L, NC, (GMAX >= 4 ? 4 : 1)>(c, d_nets, n, stamp);
    else if (G >= 2 && GMAX >= 2) lswave_launch_g<L, NC, (GMAX >= 2 ? 2 : 1)>(c, d_nets, n, stamp);
    else lswave_launch_g<L, NC, 1>(c, d_nets, n, stamp);
}

template <int L>
void lswave_launch_l(igr_ctx* c, const int* d_nets, int64_t n, int cls, int stamp) {
    switch (cls) {
    case 0: lswave_launch<L, 4>(c, d_nets, n, stamp); break;
    case 1: lswave_launch<L, 8>(c, d_nets, n, stamp); break;
    case 2: lswave_launch<L, 16>(c, d_nets, n, stamp); break;
    default: lswave_launch<L, LSWAVE_MAX_NODES>(c, d_nets, n, stamp); break;
    }
}

} // namespace

int lshape_wave_class(int nodes) {
    return nodes <= 4 ? 0 : nodes <= 8 ? 1 : nodes <= 16 ? 2 : nodes <= LSWAVE_MAX_NODES ? 3 : -1;
}

void launch_lshape_wave(igr_ctx* c, const int* d_nets, int64_t n, int cls, int stamp, double bytes) {
    if (n == 0) return;
    if (cls < 0 || cls > 3) throw std::runtime_error("L-shape node-count class out of range");
    KScope ks(c, K_LSHAPE, bytes);
#define CASE(LV)                                            \
    case LV: lswave_launch_l<LV>(c, d_nets, n, cls, stamp); break;
    switch (c->g.L) {
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
    default: throw std::runtime_error("unsupported routing layer count");
    }
#undef CASE
    HIP_OK(hipGetLastError());
}

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
