// Detour-DAG routing kernels (src/Lshape_route_detour.hpp).
//
// Plain form, the reference's launch structure: per batch one launch computes
// the best single layer change on every straight DAG edge
// (precompute_cost_edge, :123-195), then one DP launch per depth level from
// the deepest up (Lshape_route_node_layer_change_cuda, :281-406) and one
// extraction launch per level from the source down
// (get_routing_tree_layer_change_cuda, :520-694).
//
// DP keys pack the FP32 cost (as an order-preserving int) above a 32-bit id,
// node * 1000 + choice with the node's index inside its own net, so the
// 64-bit atomicMin picks the smallest cost and, among equal costs, the
// smallest (node, choice) pair exactly as the reference's global ids do.
// Each node writes its route pieces (parent edge first, then its via stack)
// into a private 8-int slot; a compaction concatenates them in node order, so
// the route order is deterministic where the reference's atomicAdd append is
// not.
#include "igr_ctx.hpp"

namespace igr_dev {

namespace {

constexpr int TPB = 256;
constexpr uint64_t INIT_KEY = 7208259049189261824ULL;  // :106
constexpr uint32_t ID_MUL = 1000;
constexpr int EMIT = 9;  // count + 8 ints per node

struct DagArgs {
    GridDev g;
    const int* node;
    const int64_t* par_off;
    const int* par;
    const int* slot;
    const int* child_num;
    const int64_t* csum;
    const int* loc;
    const int* enode;
    double* costs;
    uint64_t* packed;
    uint64_t* best;
    double* cost_edges;
    int* best_change;
    int* lo;
    int* emit;
    int64_t node0, edge0, slot0;
    int stamp;
};

__device__ __forceinline__ uint64_t pack_key(float cost, uint32_t id) {
    int32_t i = __float_as_int(cost);
    int32_t o = i >= 0 ? i : i ^ 0x7FFFFFFF;
    return ((uint64_t)(int64_t)o << 32) | (uint64_t)id;
}
__device__ __forceinline__ double unpack_cost(uint64_t k) {
    int32_t o = (int32_t)(k >> 32);
    return (double)__int_as_float(o >= 0 ? o : o ^ 0x7FFFFFFF);
}

__device__ __forceinline__ double wseg(const GridDev& g, int l, int xmin, int xmax, int ymin, int ymax) {
    return g.presum[gidx(g, l, xmax, ymax)] - g.presum[gidx(g, l, xmin, ymin)];
}

__device__ __forceinline__ void add_segment(const GridDev& g, int l, int minx, int maxx, int miny, int maxy,
                                            int stamp) {
    atomicAdd(&g.pre[gidx(g, l, minx, miny)], 1);
    if (minx == maxx) {
        atomicAdd(&g.pre[gidx(g, l, minx, maxy)], -1);
        g.ts[gidx(g, l, minx, maxy)] = stamp;
    }
    if (miny == maxy) {
        atomicAdd(&g.pre[gidx(g, l, maxx, miny)], -1);
        g.ts[gidx(g, l, maxx, miny)] = stamp;
    }
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int* p, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

// precompute_cost_edge: one thread per (edge, from layer).
__global__ void __launch_bounds__(TPB) k_dag_edges(DagArgs a, int64_t edges) {
    int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const GridDev& g = a.g;
    const int L = g.L, LL = L * L;
    if (t >= edges * L) return;
    const int64_t e = a.edge0 + t / L;
    const int fl = (int)(t % L);
    const int64_t er = e - a.edge0;
    const int to = a.par[e], from = a.enode[e];
    const int tox = a.node[to] / g.Y % g.X, toy = a.node[to] % g.Y;
    const int fromx = a.node[from] / g.Y % g.X, fromy = a.node[from] % g.Y;
    if ((layer_dir(fl, g.DIR) == 0 && fromy != toy) || (layer_dir(fl, g.DIR) == 1 && fromx != tox)) {
        for (int tl = 0; tl < L; tl++) a.cost_edges[er * LL + fl * L + tl] = INF;
        return;
    }
    const int distance = abs(tox - fromx) + abs(toy - fromy);
    if (distance == 0) return;
    const int dd = distance / min(distance, 20);
    const bool vert = fromx == tox;
    const int s0 = vert ? min(fromy, toy) : min(fromx, tox);
    const int s1 = vert ? max(fromy, toy) : max(fromx, tox);
    for (int tl = 0; tl < L; tl++) {
        if ((tl % 2) != (fl % 2)) continue;
        double mn = INF;
        int decide = -1;
        for (int sp = s0 + dd; sp < s1; sp += dd) {
            const int sx = vert ? fromx : sp, sy = vert ? sp : fromy;
            double via = 0;
            for (int ll = min(fl, tl); ll < max(fl, tl); ll++) via += (double)g.vcost[gidx(g, ll, sx, sy)];
            double cost = via;
            if (vert)
                cost += (wseg(g, fl, fromx, fromx, min(fromy, sy), max(fromy, sy)) +
                         wseg(g, tl, tox, tox, min(toy, sy), max(toy, sy)));
            else
                cost += (wseg(g, fl, min(fromx, sx), max(fromx, sx), fromy, fromy) +
                         wseg(g, tl, min(tox, sx), max(tox, sx), toy, toy));
            if (cost < mn) { mn = cost; decide = sp; }
        }
        a.cost_edges[er * LL + fl * L + tl] = mn;
        a.best_change[er * LL + fl * L + tl] = decide;
    }
}

// Bottom-up DP of one node (thread) of the current depth level.
__global__ void __launch_bounds__(TPB) k_dag_dp(DagArgs a, const int* lvl_nodes, int64_t lo_i, int64_t hi_i) {
    int64_t t = lo_i + (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= hi_i) return;
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, LL = L * L, SENT = LL + 1, XY = g.X * g.Y;
    const int node = lvl_nodes[t];
    const int nr = (int)(node - a.node0);
    const int x = a.node[node] / g.Y % g.X, y = a.node[node] % g.Y;
    const int fixed_low = 1 + a.node[node] / XY;
    const int fixed_high = fixed_low == ML ? 0 : fixed_low;
    const int nc = a.child_num[node];
    const int64_t cs = a.csum[node] - a.slot0;
    uint64_t* cur_best = a.best + cs * ML;
    const uint64_t* cur_packed = a.packed + cs * ML;
    double* cst = a.costs + (int64_t)nr * ML;
    double mcc[10];
    uint64_t bp[10];
    for (int low = 1; low <= fixed_low; low++) {
        for (int k = 0; k < nc; k++) mcc[k] = INF;
        double via = 0;
        for (int li = low; li < L + 1; li++) {
            if (li > low) via += (double)g.vcost[gidx(g, li - 2, x, y)];
            for (int k = 0; k < nc; k++) {
                uint64_t key = cur_packed[k * ML + li];
                double cc = unpack_cost(key);
                if (cc < mcc[k]) { mcc[k] = cc; bp[k] = (uint64_t)(uint32_t)key * ML + li; }
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
    const uint32_t myid = (uint32_t)a.loc[node] * ID_MUL;
    for (int64_t e = a.par_off[node]; e < a.par_off[node + 1]; e++) {
        const int p = a.par[e], ci = a.slot[e];
        uint64_t* ppk = a.packed + (a.csum[p] - a.slot0 + ci) * ML;
        const int px = a.node[p] / g.Y % g.X, py = a.node[p] % g.Y;
        const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
        const int64_t er = e - a.edge0;
        uint64_t cache[MAXL + 1];
        for (int tt = 1; tt < ML; tt++) cache[tt] = ppk[tt];
        for (int layer = 1; layer < ML; layer++) {
            if (((layer & 1) ^ g.DIR) == 1 && y != py) continue;
            if (((layer & 1) ^ g.DIR) == 0 && x != px) continue;
            const double base = cst[layer];
            for (int tt = 1 + (1 - layer % 2); tt < ML; tt += 2) {
                double cost2;
                int choice = SENT;
                if (tt == layer) cost2 = base + wseg(g, layer - 1, minx, maxx, miny, maxy);
                else if (abs(x - px) + abs(y - py) > 4) {
                    const int q = (layer - 1) * L + (tt - 1);
                    const double ce = a.cost_edges[er * LL + q];
                    if (ce < INF) { cost2 = base + ce; choice = q; }
                    else continue;
                } else continue;
                const uint64_t key = pack_key((float)cost2, myid + (uint32_t)choice);
                if (key < cache[tt]) cache[tt] = key;
            }
        }
        for (int tt = 1; tt < ML; tt++) atomicMin((unsigned long long*)&ppk[tt], (unsigned long long)cache[tt]);
    }
}

// Top-down extraction of one node of the current depth level.
__global__ void __launch_bounds__(TPB) k_dag_extract(DagArgs a, const int* lvl_nodes, int64_t lo_i, int64_t hi_i) {
    int64_t t = lo_i + (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= hi_i) return;
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, LL = L * L, SENT = LL + 1, XY = g.X * g.Y;
    const int node = lvl_nodes[t];
    const int nr = (int)(node - a.node0);
    const int cx = a.node[node] / g.Y % g.X, cy = a.node[node] % g.Y;
    const int l = a.node[node] / XY;
    const uint64_t* cur_best = a.best + (a.csum[node] - a.slot0) * ML;
    int* em = a.emit + (int64_t)nr * EMIT;
    int ne = 0;
    auto push = [&](int u, int v) { em[1 + ne] = u; em[2 + ne] = v; ne += 2; };
    int out;
    if (a.par_off[node + 1] == a.par_off[node]) {
        int min_layer = 0;
        double min_cost = a.costs[(int64_t)nr * ML];
        for (int layer = 1; layer < ML; layer++)
            if (a.costs[(int64_t)nr * ML + layer] < min_cost) { min_cost = a.costs[(int64_t)nr * ML + layer]; min_layer = layer; }
        out = min_layer;
    } else {
        out = -1;
        int par_layer = -1, par_idx = -1;
        int64_t par_edge = -1;
        for (int64_t e = a.par_off[node]; e < a.par_off[node + 1]; e++) {
            const int pn = a.par[e];
            const int plo = a.lo[pn - a.node0];
            if (plo > 0) {
                par_idx = pn;
                par_edge = e;
                par_layer = plo;
                const uint64_t path = a.best[(a.csum[pn] - a.slot0 + a.slot[e]) * ML + par_layer];
                const uint64_t child = path / ML / ID_MUL, choice = (path / ML) % ID_MUL;
                if (child == (uint64_t)a.loc[node]) {
                    out = choice == (uint64_t)SENT ? (int)(path % ML) : (int)(choice / L) + 1;
                    const int px = a.node[pn] / g.Y % g.X, py = a.node[pn] % g.Y;
                    const int64_t er = e - a.edge0;
                    if (px == cx && cy != py) {
                        if (choice == (uint64_t)SENT) {
                            const int ly = out - 1;
                            add_segment(g, ly, px, px, min(py, cy), max(py, cy), a.stamp);
                            push(ly * XY + px * g.Y + min(py, cy), ly * XY + px * g.Y + max(py, cy));
                        } else {
                            const int fl = (int)choice / L, tl = (int)choice % L;
                            const int ch = a.best_change[er * LL + fl * L + tl];
                            add_segment(g, fl, cx, cx, min(cy, ch), max(cy, ch), a.stamp);
                            add_segment(g, tl, cx, cx, min(py, ch), max(py, ch), a.stamp);
                            push(fl * XY + cx * g.Y + min(cy, ch), fl * XY + cx * g.Y + max(cy, ch));
                            push(min(fl, tl) * XY + cx * g.Y + ch, max(fl, tl) * XY + cx * g.Y + ch);
                            push(tl * XY + cx * g.Y + min(py, ch), tl * XY + cx * g.Y + max(py, ch));
                        }
                    } else if (py == cy && cx != px) {
                        if (choice == (uint64_t)SENT) {
                            const int ly = out - 1;
                            add_segment(g, ly, min(px, cx), max(px, cx), py, py, a.stamp);
                            push(ly * XY + min(px, cx) * g.Y + py, ly * XY + max(px, cx) * g.Y + py);
                        } else {
                            const int fl = (int)choice / L, tl = (int)choice % L;
                            const int ch = a.best_change[er * LL + fl * L + tl];
                            add_segment(g, fl, min(cx, ch), max(cx, ch), cy, cy, a.stamp);
                            add_segment(g, tl, min(px, ch), max(px, ch), cy, cy, a.stamp);
                            push(fl * XY + min(cx, ch) * g.Y + cy, fl * XY + max(cx, ch) * g.Y + cy);
                            push(min(fl, tl) * XY + ch * g.Y + cy, max(fl, tl) * XY + ch * g.Y + cy);
                            push(tl * XY + min(px, ch) * g.Y + cy, tl * XY + max(px, ch) * g.Y + cy);
                        }
                    }
                    break;
                } else {
                    out = -1;
                }
            }
        }
        if (par_layer == -1) {
            a.lo[nr] = -1;
            em[0] = 0;
            return;
        }
        const uint64_t path = a.best[(a.csum[par_idx] - a.slot0 + a.slot[par_edge]) * ML + par_layer];
        if (path / ML / ID_MUL != (uint64_t)a.loc[node]) {
            a.lo[nr] = -1;
            em[0] = 0;
            return;
        }
    }
    a.lo[nr] = out;
    // A source without any finite layer has no route (the reference would
    // index layer -1 here); children without a recorded layer are skipped.
    if (out < 1) {
        em[0] = ne;
        return;
    }
    int minl = l + 1;
    int maxl = (l + 1) == ML ? 1 : minl;
    minl = min(minl, out);
    maxl = max(maxl, out);
    for (int k = 0; k < a.child_num[node]; k++) {
        const int lc = (int)(cur_best[k * ML + out] % ML);
        if (lc < 1) continue;
        minl = min(lc, minl);
        maxl = max(lc, maxl);
    }
    if (minl < maxl) push((minl - 1) * XY + cx * g.Y + cy, (maxl - 1) * XY + cx * g.Y + cy);
    em[0] = ne;
}

// Concatenate the node slots of each net in node order into its new route and
// point the net at it.
__global__ void __launch_bounds__(TPB) k_dag_compact(const int* nets, int64_t net0, int64_t net1,
                                                     const int64_t* node_off, const int64_t* route_slot, int* routes,
                                                     const int* emit, int64_t node0, int** rptr) {
    int64_t i = net0 + (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= net1) return;
    int* r = routes + route_slot[i];
    int cnt = 1;
    for (int64_t nd = node_off[i]; nd < node_off[i + 1]; nd++) {
        const int* em = emit + (nd - node0) * EMIT;
        for (int k = 0; k < em[0]; k++) r[cnt + k] = em[1 + k];
        cnt += em[0];
    }
    r[0] = cnt;
    rptr[nets[i]] = r;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

DagArgs make_args(igr_ctx* c, const DagBatch& b, int stamp) {
    DagDev& D = c->dag;
    return DagArgs{c->g, D.node.p, D.par_off.p, D.par.p, D.slot.p, D.child_num.p, D.csum.p, D.loc.p, D.enode.p,
                   D.costs.p, D.packed.p, D.best.p, D.cost_edges.p, D.best_change.p, D.lo.p, D.emit.p,
                   b.node0, b.edge0, b.slot0, stamp};
}

} // namespace

void launch_dag_batch_plain(igr_ctx* c, const DagBatch& b, int stamp, const std::vector<int64_t>& h_lvl_off) {
    DagDev& D = c->dag;
    const int ML = c->g.L + 1;
    const int64_t nn = b.node1 - b.node0, ne = b.edge1 - b.edge0, ns = b.slot1 - b.slot0;
    hipStream_t s = c->stream;
    // Initial values (src/Lshape_route_detour.hpp:75-121).
    k_fill_f64<<<blocks(nn * ML), TPB, 0, s>>>(D.costs.p, nn * ML, INF);
    k_fill_u64<<<blocks(ns * ML), TPB, 0, s>>>(D.packed.p, ns * ML, INIT_KEY);
    k_fill_u64<<<blocks(ns * ML), TPB, 0, s>>>(D.best.p, ns * ML, 0);
    k_fill_i32<<<blocks(nn), TPB, 0, s>>>(D.lo.p, nn, -1);
    HIP_OK(hipGetLastError());
    DagArgs a = make_args(c, b, stamp);
    if (ne > 0) {
        k_dag_edges<<<blocks(ne * c->g.L), TPB, 0, s>>>(a, ne);
        HIP_OK(hipGetLastError());
    }
    for (int d = b.levels - 1; d >= 0; d--) {
        int64_t lo = h_lvl_off[b.lvl0 + d], hi = h_lvl_off[b.lvl0 + d + 1];
        if (hi > lo) k_dag_dp<<<blocks(hi - lo), TPB, 0, s>>>(a, D.lvl_nodes.p, lo, hi);
    }
    for (int d = 0; d < b.levels; d++) {
        int64_t lo = h_lvl_off[b.lvl0 + d], hi = h_lvl_off[b.lvl0 + d + 1];
        if (hi > lo) k_dag_extract<<<blocks(hi - lo), TPB, 0, s>>>(a, D.lvl_nodes.p, lo, hi);
    }
    HIP_OK(hipGetLastError());
}

void launch_dag_compact(igr_ctx* c, const DagBatch& b) {
    DagDev& D = c->dag;
    const int64_t n = b.net1 - b.net0;
    if (n == 0) return;
    k_dag_compact<<<blocks(n), TPB, 0, c->stream>>>(D.nets.p, b.net0, b.net1, D.node_off.p, D.route_slot.p,
                                                      c->nets.dag_routes.back()->p, D.emit.p, b.node0, c->nets.rptr.p);
    HIP_OK(hipGetLastError());
}

} // namespace igr_dev
