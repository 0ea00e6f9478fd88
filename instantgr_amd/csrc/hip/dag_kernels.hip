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

// Parent-slot keys are updated by 64-bit atomics, which complete in L2 and
// leave any L1 copy behind; they are read with device-scope loads, which go
// to L2, so no level has to invalidate the L1 (whose other contents, the DAG
// and grid inputs, stay cached).
__device__ __forceinline__ uint64_t load_key(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double wseg(const GridDev& g, int l, int xmin, int xmax, int ymin, int ymax) {
    return g.presum[gidx(g, l, xmax, ymax)] - g.presum[gidx(g, l, xmin, ymin)];
}

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

// Bottom-up DP of one node by one thread (Lshape_route_node_layer_change_cuda,
// src/Lshape_route_detour.hpp:281-353): for each lowest layer, the running
// minimum of every child slot over the layers above it, the via costs in
// layer order, then the downward sweep.
// Per-lane DP scratch in LDS for k_dag_net, whose workgroup is one wave
// (WTPB lanes) running one of the two node DPs at a time: child keys, best
// child choices, and the current low's choices or running minima.
// Child-major, so the 64 lanes of one child entry hit consecutive banks; kept
// out of registers for more waves per SIMD.
constexpr int WTPB = 64;
__shared__ uint64_t s_key[10][WTPB], s_bp[10][WTPB], s_cb[10][WTPB];

// The sequential DP's running minima (as bits) and choices of one lane: in
// registers, or (LDS) in the LDS scratch above, for a one-wave workgroup only
// (the plain form's k_dag_dp runs 256-thread workgroups).
template <bool LDS>
struct SeqScratch {
    uint64_t m[10], b[10];
    __device__ __forceinline__ uint64_t& mcc(int k) { return m[k]; }
    __device__ __forceinline__ uint64_t& bp(int k) { return b[k]; }
};
template <>
struct SeqScratch<true> {
    __device__ __forceinline__ uint64_t& mcc(int k) { return s_cb[k][threadIdx.x]; }
    __device__ __forceinline__ uint64_t& bp(int k) { return s_bp[k][threadIdx.x]; }
};

// Sequential node DP (a lane per node).
template <bool LDS>
__device__ void dp_node_seq(const DagArgs& a, int node) {
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, XY = g.X * g.Y;
    const int nr = (int)(node - a.node0);
    const int x = a.node[node] / g.Y % g.X, y = a.node[node] % g.Y;
    const int fixed_low = 1 + a.node[node] / XY;
    const int fixed_high = fixed_low == ML ? 0 : fixed_low;
    const int nc = a.child_num[node];
    const int64_t cs = a.csum[node] - a.slot0;
    uint64_t* cur_best = a.best + cs * ML;
    const uint64_t* cur_packed = a.packed + cs * ML;
    double* cst = a.costs + (int64_t)nr * ML;
    SeqScratch<LDS> sc;
#define mcc(k) sc.mcc(k)
#define bp(k) sc.bp(k)
    // LDS: child 0's keys (read at every low) are loaded once into the lane's
    // key slots, which the 16-lane DP alone uses otherwise (up to 10 layers)
    const bool kc = LDS && nc >= 1 && ML <= 10;
    if (kc)
        for (int li = 1; li <= L; li++) s_key[li][threadIdx.x] = load_key(cur_packed + li);
    // low above L visits no layer (the reference's loop runs it empty)
    const int lowtop = min(fixed_low, L);
    for (int low = 1; low <= lowtop; low++) {
        for (int k = 0; k < nc; k++) mcc(k) = __double_as_longlong(INF);
        double via = 0;
        for (int li = low; li < L + 1; li++) {
            if (li > low) via += (double)g.vcost[gidx(g, li - 2, x, y)];
            for (int k = 0; k < nc; k++) {
                const uint64_t key = k == 0 && kc ? s_key[li][threadIdx.x] : load_key(cur_packed + k * ML + li);
                double cc = unpack_cost(key);
                if (cc < __longlong_as_double(mcc(k))) {
                    mcc(k) = __double_as_longlong(cc);
                    bp(k) = (uint64_t)(uint32_t)key * ML + li;
                }
            }
            if (li >= fixed_high) {
                double cost = via;
                for (int k = 0; k < nc; k++) cost += __longlong_as_double(mcc(k));
                if (cost < INF && cost < cst[li]) {
                    cst[li] = cost;
                    for (int k = 0; k < nc; k++) cur_best[k * ML + li] = bp(k);
                }
            }
        }
        for (int li = L + 1 - 2; li >= low; li--)
            if (cst[li + 1] < cst[li]) {
                cst[li] = cst[li + 1];
                for (int k = 0; k < nc; k++) cur_best[k * ML + li] = cur_best[k * ML + li + 1];
            }
    }
#undef mcc
#undef bp
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
    const double* cst = a.costs + (int64_t)nr * ML;
    dp_node_seq<false>(a, node);
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

// Top-down extraction of one node whose parents are all decided.
__device__ void extract_node(const DagArgs& a, int node) {
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, LL = L * L, SENT = LL + 1, XY = g.X * g.Y;
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
                            if (g.vc) emit_via(g, fl, tl, cx, ch);
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
                            if (g.vc) emit_via(g, fl, tl, ch, cy);
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
    if (minl < maxl) {
        push((minl - 1) * XY + cx * g.Y + cy, (maxl - 1) * XY + cx * g.Y + cy);
        if (g.vc) emit_via(g, minl - 1, maxl - 1, cx, cy);
    }
    em[0] = ne;
}

// Plain form: one launch per depth level, one thread per node.
__global__ void __launch_bounds__(TPB) k_dag_extract(DagArgs a, const int* lvl_nodes, int64_t lo_i, int64_t hi_i) {
    int64_t t = lo_i + (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= hi_i) return;
    extract_node(a, lvl_nodes[t]);
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

// ---------------------------------------------------------------------------
// Batched form: one 64-lane workgroup routes one net's whole DAG, walking its
// depth levels with workgroup barriers instead of one launch per level; all
// nets of the batch run concurrently, taken from a list that puts the
// deepest DAGs first (workgroups start in list order, so the longest chains
// start first instead of forming the launch's tail).  Per level the node DP runs one lane
// per node, then the parent-edge candidates run one lane per (edge, layer)
// and go straight to the parent slots with 64-bit atomicMin (the minimum
// over all candidates is the same as the reference's per-thread cache
// followed by one atomicMin).  Extraction walks the levels top-down, then the
// lanes concatenate the node slots into the route with a wave prefix sum.
// ---------------------------------------------------------------------------
struct NetLevels {
    const int64_t* nl_off;     // per net: first entry in lvl_node_off / lvl_edge_off
    const int* nlevels;        // per net: number of depth levels
    const int64_t* lvl_node_off;  // per level: range in dnode
    const int64_t* lvl_edge_off;  // per level: range in dedge
    const int* dnode;          // nodes grouped by depth within each net
    const int64_t* dedge;      // edges of each level's nodes
};

// Node DP with 16 lanes per node, lane s holding layer index s (1..L).  Same
// results as dp_node: for each bottom layer 'low' the running minimum of every
// child slot over layers [low, s] is an inclusive prefix-min across the lanes
// (ties keep the lower layer, as the sequential strict '<' does), the via sum
// is accumulated in layer order inside the lane, and the downward sweep is a
// suffix argmin across the lanes (ties keep the lower layer).  'valid' is
// false on lane groups without a node; every lane of the wave calls this.
__device__ void dp_node16(const DagArgs& a, int node, bool valid, int ncw) {
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, XY = g.X * g.Y;
    const int lane = threadIdx.x & 63, s = lane & 15;
    int nc = 0, fixed_low = 0, fixed_high = 0, x = 0, y = 0, nr = 0;
    int64_t cs = 0;
    if (valid) {
        nr = (int)(node - a.node0);
        x = a.node[node] / g.Y % g.X;
        y = a.node[node] % g.Y;
        fixed_low = 1 + a.node[node] / XY;
        fixed_high = fixed_low == ML ? 0 : fixed_low;
        nc = a.child_num[node];
        cs = a.csum[node] - a.slot0;
    }
    const bool act = valid && s >= 1 && s <= L;
    // per lane in LDS: the child keys (cost above id; padded with INIT_KEY),
    // the best child choices so far (bp) and the current low's (cb), to leave
    // registers for more waves per SIMD
    // (child-major, so the 64 lanes of one child entry hit consecutive banks)
#define key(k) s_key[k][lane]
#define bp(k) s_bp[k][lane]
#define cb(k) s_cb[k][lane]
#pragma unroll
    for (int k = 0; k < 10; k++) key(k) = (k < nc && act) ? load_key(a.packed + (cs + k) * ML + s) : INIT_KEY;
    // lane t of the group holds the via cut cost of layer t at the node
    const float vl = (valid && s < L) ? g.vcost[gidx(g, s, x, y)] : 0.0f;
    double cst = INF;
#pragma unroll
    for (int k = 0; k < 10; k++) bp(k) = 0;
    // Lowest layer: group-uniform bound, the wave runs to its largest.
    // (a low above L visits no layer: the reference's loop runs it empty)
    int lowmax = min(fixed_low, L);
    for (int d = 16; d < 64; d <<= 1) lowmax = max(lowmax, __shfl_xor(lowmax, d, 64));
    for (int low = 1; low <= lowmax; low++) {
        const bool run = act && low <= fixed_low;
        double via = 0;
        // the costs of layers low-1 .. s-2 in ascending order (t runs to L
        // only: layers from L on are past every lane's s)
        for (int t = 2; t <= L; t++) {
            const float vt = __shfl(vl, t - 2, 16);
            if (t > low && t <= s) via += (double)vt;
        }
        double cost = via;
#pragma unroll
        for (int k = 0; k < 10; k++) {
            cb(k) = 0;
            if (k >= ncw) continue;
            const double ck = unpack_cost(key(k));
            double m = (run && s >= low && k < nc && ck < INF) ? ck : INF;
            uint64_t b = (uint64_t)(uint32_t)key(k) * ML + s;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const double om = __shfl_up(m, d, 16);
                const uint64_t ob = __shfl_up(b, d, 16);
                if (s >= d && !(m < om)) { m = om; b = ob; }
            }
            cb(k) = b;
            if (k < nc) cost += m;
        }
        if (run && s >= low && s >= fixed_high && cost < INF && cost < cst) {
            cst = cost;
#pragma unroll
            for (int k = 0; k < 10; k++) bp(k) = cb(k);
        }
        // Suffix argmin over lanes [low, L]: lane s takes the smallest value
        // above it only when strictly smaller.
        double c = (valid && s >= low && s <= L) ? cst : __longlong_as_double(0x7ff0000000000000LL);
        int src = s;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const double oc = __shfl_down(c, d, 16);
            const int os = __shfl_down(src, d, 16);
            if (s + d < 16 && oc < c) { c = oc; src = os; }
        }
        const double nc_cost = __shfl(cst, src, 16);
#pragma unroll
        for (int k = 0; k < 10; k++) {
            if (k >= ncw) continue;
            const uint64_t nb = s_bp[k][(lane & ~15) + src];
            if (run && s >= low && src != s) bp(k) = nb;
        }
        if (run && s >= low && src != s) cst = nc_cost;
    }
    if (valid && s <= L) {
        double* cstp = a.costs + (int64_t)nr * ML;
        cstp[s] = s == 0 ? INF : cst;
        if (s >= 1)
            for (int k = 0; k < nc; k++) a.best[(cs + k) * ML + s] = bp(k);
    }
#undef key
#undef bp
#undef cb
}

// Candidates of one parent edge from one source layer of the child.
__device__ void dp_edge_layer(const DagArgs& a, int64_t e, int layer) {
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1, LL = L * L, SENT = LL + 1;
    const int node = a.enode[e];
    const int nr = (int)(node - a.node0);
    const int x = a.node[node] / g.Y % g.X, y = a.node[node] % g.Y;
    const int p = a.par[e], ci = a.slot[e];
    const int px = a.node[p] / g.Y % g.X, py = a.node[p] % g.Y;
    if (((layer & 1) ^ g.DIR) == 1 && y != py) return;
    if (((layer & 1) ^ g.DIR) == 0 && x != px) return;
    uint64_t* ppk = a.packed + (a.csum[p] - a.slot0 + ci) * ML;
    const int minx = min(x, px), maxx = max(x, px), miny = min(y, py), maxy = max(y, py);
    const int64_t er = e - a.edge0;
    const uint32_t myid = (uint32_t)a.loc[node] * ID_MUL;
    const double base = a.costs[(int64_t)nr * ML + layer];
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
        atomicMin((unsigned long long*)&ppk[tt], (unsigned long long)pack_key((float)cost2, myid + (uint32_t)choice));
    }
}

// Batched form of precompute_cost_edge: one thread per (edge, from layer,
// to layer), computing only the entries the DP reads (to layer of the same
// direction and different from the from layer, from layer along the edge,
// edges longer than 4 gcells).  Each entry is the same strict first minimum
// over the same sampled positions as the reference's.
// Batched form: one wave per DAG edge.  A long straight edge's sample
// positions (at most 40: dd = dist / min(dist, 20)) are read once into LDS:
// the via cut costs of every layer at each sample, and the prefix sums of
// every layer at the samples and at both ends.  The lanes then take the
// (from, to) layer pairs and form each pair's minimum from LDS with the
// arithmetic and order of k_dag_edges_pairs: the via cuts summed from the
// lower layer up, then the two wire pieces, strict '<' keeping the first.
constexpr int EW_SAMPLES = 40;
constexpr int EW_WAVES = TPB / 64;

__global__ void __launch_bounds__(TPB) k_dag_edges_wave(DagArgs a, int64_t edges) {
    __shared__ float s_v[EW_WAVES][EW_SAMPLES][MAXL];      // via cut cost at (layer, sample)
    __shared__ double s_ps[EW_WAVES][EW_SAMPLES][MAXL];    // prefix sum at (layer, sample)
    __shared__ double s_pe[EW_WAVES][2][MAXL];             // prefix sum at the from / to end
    const GridDev& g = a.g;
    const int L = g.L, LL = L * L;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t er = (int64_t)blockIdx.x * EW_WAVES + wv; er < edges; er += (int64_t)gridDim.x * EW_WAVES) {
        const int64_t e = a.edge0 + er;
        const int to = a.par[e], from = a.enode[e];
        const int tox = a.node[to] / g.Y % g.X, toy = a.node[to] % g.Y;
        const int fromx = a.node[from] / g.Y % g.X, fromy = a.node[from] % g.Y;
        const bool vert = fromx == tox;
        if (!vert && fromy != toy) continue;                   // not straight: no layer change
        const int distance = abs(tox - fromx) + abs(toy - fromy);
        if (distance <= 4) continue;
        const int dd = distance / min(distance, 20);
        const int s0 = vert ? min(fromy, toy) : min(fromx, tox);
        const int s1 = vert ? max(fromy, toy) : max(fromx, tox);
        const int ns = (s1 - 1 - s0) / dd;                     // samples s0 + dd, ..., < s1
        const int edir = vert ? 1 : 0;                          // direction the wire pieces run
        // stage: lane = (sample, layer) pairs, then the two ends
        for (int t = lane; t < ns * L; t += 64) {
            const int k = t / L, l = t - k * L;
            const int sp = s0 + (k + 1) * dd;
            const int sx = vert ? fromx : sp, sy = vert ? sp : fromy;
            s_v[wv][k][l] = l < L - 1 ? g.vcost[gidx(g, l, sx, sy)] : 0.0f;
            s_ps[wv][k][l] = layer_dir(l, g.DIR) == edir ? g.presum[gidx(g, l, sx, sy)] : 0.0;
        }
        for (int t = lane; t < 2 * L; t += 64) {
            const int end = t / L, l = t - end * L;
            const int px = end ? tox : fromx, py = end ? toy : fromy;
            s_pe[wv][end][l] = layer_dir(l, g.DIR) == edir ? g.presum[gidx(g, l, px, py)] : 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const int fpos = vert ? fromy : fromx, tpos = vert ? toy : tox;
        for (int q = lane; q < LL; q += 64) {
            const int fl = q / L, tl = q % L;
            if (tl == fl || (tl % 2) != (fl % 2)) continue;
            if (layer_dir(fl, g.DIR) != edir) continue;
            const int lo = min(fl, tl), hi = max(fl, tl);
            double mn = INF;
            int decide = -1;
            for (int k = 0; k < ns; k++) {
                const int sp = s0 + (k + 1) * dd;
                double via = 0;
                for (int ll = lo; ll < hi; ll++) via += (double)s_v[wv][k][ll];
                double cost = via;
                // wire piece on fl between the from end and the sample, on tl
                // between the to end and the sample: presum[max] - presum[min]
                const double wf = sp >= fpos ? s_ps[wv][k][fl] - s_pe[wv][0][fl] : s_pe[wv][0][fl] - s_ps[wv][k][fl];
                const double wt = sp >= tpos ? s_ps[wv][k][tl] - s_pe[wv][1][tl] : s_pe[wv][1][tl] - s_ps[wv][k][tl];
                cost += (wf + wt);
                if (cost < mn) { mn = cost; decide = sp; }
            }
            a.cost_edges[er * LL + q] = mn;
            a.best_change[er * LL + q] = decide;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void __launch_bounds__(TPB) k_dag_edges_pairs(DagArgs a, int64_t edges) {
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const GridDev& g = a.g;
    const int L = g.L, LL = L * L;
    if (t >= edges * LL) return;
    const int64_t er = t / LL;
    const int q = (int)(t % LL), fl = q / L, tl = q % L;
    if (tl == fl || (tl % 2) != (fl % 2)) return;
    const int64_t e = a.edge0 + er;
    const int to = a.par[e], from = a.enode[e];
    const int tox = a.node[to] / g.Y % g.X, toy = a.node[to] % g.Y;
    const int fromx = a.node[from] / g.Y % g.X, fromy = a.node[from] % g.Y;
    if ((layer_dir(fl, g.DIR) == 0 && fromy != toy) || (layer_dir(fl, g.DIR) == 1 && fromx != tox)) return;
    const int distance = abs(tox - fromx) + abs(toy - fromy);
    if (distance <= 4) return;
    const int dd = distance / min(distance, 20);
    const bool vert = fromx == tox;
    const int s0 = vert ? min(fromy, toy) : min(fromx, tox);
    const int s1 = vert ? max(fromy, toy) : max(fromx, tox);
    const int lo = min(fl, tl), hi = max(fl, tl);
    double mn = INF;
    int decide = -1;
    for (int sp = s0 + dd; sp < s1; sp += dd) {
        const int sx = vert ? fromx : sp, sy = vert ? sp : fromy;
        double via = 0;
        for (int ll = lo; ll < hi; ll++) via += (double)g.vcost[gidx(g, ll, sx, sy)];
        double cost = via;
        if (vert)
            cost += (wseg(g, fl, fromx, fromx, min(fromy, sy), max(fromy, sy)) +
                     wseg(g, tl, tox, tox, min(toy, sy), max(toy, sy)));
        else
            cost += (wseg(g, fl, min(fromx, sx), max(fromx, sx), fromy, fromy) +
                     wseg(g, tl, min(tox, sx), max(tox, sx), toy, toy));
        if (cost < mn) { mn = cost; decide = sp; }
    }
    a.cost_edges[er * LL + q] = mn;
    a.best_change[er * LL + q] = decide;
}

// W: waves per SIMD the register allocation must allow (2 leaves the compiler
// free; 4 caps it at 128 registers, spilling the rest).  The LDS scratch
// (15 KB per one-wave workgroup) allows at most 3, so 4 asks for 3.
//
// Persistent form: a grid of at most a few thousand workgroups pulls the
// batch's nets from a device work queue (one counter per batch, zero at the
// pass start) in list order, deepest DAG first, so nets of very uneven size
// (two-pin subnets next to 1000-pin nets' pieces) keep every workgroup busy
// until the queue is empty; each workgroup exits when the counter passes n.
template <int W>
__global__ void __launch_bounds__(WTPB) __attribute__((amdgpu_waves_per_eu(W < 3 ? W : 3, 8)))
k_dag_net(DagArgs a, NetLevels lv, const int* nets, const int* list, int64_t n, int* queue, const int64_t* node_off,
          const int64_t* route_slot, int* routes, int** rptr, int seq_min) {
    const int lane = threadIdx.x;
    for (;;) {
    __syncthreads();   // the previous net's LDS use is over
    int qi = 0;
    if (lane == 0) qi = atomicAdd(queue, 1);
    qi = __shfl(qi, 0, WTPB);
    if (qi >= n) break;
    const int64_t i = list[qi];
    const GridDev& g = a.g;
    const int L = g.L, ML = L + 1;
    const int64_t n0 = node_off[i], n1 = node_off[i + 1];
    const int64_t e0 = a.par_off[n0], e1 = a.par_off[n1];
    const int64_t s0 = a.csum[n0], s1 = a.csum[n1];
    // Initial values (src/Lshape_route_detour.hpp:75-121) and edge layer changes.
    for (int64_t k = lane; k < (n1 - n0) * ML; k += WTPB) a.costs[(n0 - a.node0) * ML + k] = INF;
    for (int64_t k = lane; k < (s1 - s0) * ML; k += WTPB) {
        a.packed[(s0 - a.slot0) * ML + k] = INIT_KEY;
        a.best[(s0 - a.slot0) * ML + k] = 0;
    }
    __syncthreads();
    const int nlev = lv.nlevels[i];
    const int64_t lb = lv.nl_off[i];
    for (int d = nlev - 1; d >= 0; d--) {
        const int64_t a0 = lv.lvl_node_off[lb + d], a1 = lv.lvl_node_off[lb + d + 1];
        if (a1 - a0 >= seq_min) {
            // wide level: a node per lane
            for (int64_t k = a0 + lane; k < a1; k += WTPB) dp_node_seq<true>(a, lv.dnode[k]);
        } else {
            for (int64_t p = a0; p < a1; p += 4) {
                const int64_t k = p + (lane >> 4);
                const bool valid = k < a1;
                const int node = valid ? lv.dnode[k] : 0;
                int ncw = valid ? a.child_num[node] : 0;
                for (int o = 16; o < 64; o <<= 1) ncw = max(ncw, __shfl_xor(ncw, o, 64));
                dp_node16(a, node, valid, ncw);
            }
        }
        __syncthreads();
        const int64_t b0 = lv.lvl_edge_off[lb + d], b1 = lv.lvl_edge_off[lb + d + 1];
        for (int64_t k = lane; k < (b1 - b0) * L; k += WTPB) dp_edge_layer(a, lv.dedge[b0 + k / L], (int)(k % L) + 1);
        __syncthreads();
    }
    for (int d = 0; d < nlev; d++) {
        const int64_t a0 = lv.lvl_node_off[lb + d], a1 = lv.lvl_node_off[lb + d + 1];
        for (int64_t k = a0 + lane; k < a1; k += WTPB) extract_node(a, lv.dnode[k]);
        __syncthreads();
    }
    // Concatenate node slots in node order (wave prefix sum of the counts).
    int* r = routes + route_slot[i];
    int base = 1;
    for (int64_t c0 = n0; c0 < n1; c0 += WTPB) {
        const int64_t nd = c0 + lane;
        const int* em = a.emit + (nd - a.node0) * EMIT;
        const int cnt = nd < n1 ? em[0] : 0;
        int incl = cnt;
        for (int off = 1; off < WTPB; off <<= 1) {
            int v = __shfl_up(incl, off, WTPB);
            if (lane >= off) incl += v;
        }
        const int start = base + incl - cnt;
        for (int k = 0; k < cnt; k++) r[start + k] = em[1 + k];
        base += __shfl(incl, WTPB - 1, WTPB);
    }
    if (lane == 0) {
        r[0] = base;
        rptr[nets[i]] = r;
    }
    }
}

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
        KScope ks(c, K_DAG_EDGES, b.bytes_edges);
        k_dag_edges<<<blocks(ne * c->g.L), TPB, 0, s>>>(a, ne);
        HIP_OK(hipGetLastError());
    }
    {
        KScope ks(c, K_DAG_DP, b.bytes_dp);
        for (int d = b.levels - 1; d >= 0; d--) {
            int64_t lo = h_lvl_off[b.lvl0 + d], hi = h_lvl_off[b.lvl0 + d + 1];
            if (hi > lo) k_dag_dp<<<blocks(hi - lo), TPB, 0, s>>>(a, D.lvl_nodes.p, lo, hi);
        }
    }
    {
        KScope ks(c, K_DAG_EXTRACT);
        for (int d = 0; d < b.levels; d++) {
            int64_t lo = h_lvl_off[b.lvl0 + d], hi = h_lvl_off[b.lvl0 + d + 1];
            if (hi > lo) k_dag_extract<<<blocks(hi - lo), TPB, 0, s>>>(a, D.lvl_nodes.p, lo, hi);
        }
    }
    HIP_OK(hipGetLastError());
}

void launch_dag_batch_wave(igr_ctx* c, const DagBatch& b, int stamp) {
    DagDev& D = c->dag;
    const int64_t n = b.net1 - b.net0;
    if (n == 0) return;
    DagArgs a = make_args(c, b, stamp);
    const int64_t ne = b.edge1 - b.edge0;
    if (ne > 0) {
        KScope ke(c, K_DAG_EDGES, b.bytes_edges);
        const int64_t items = ne * c->g.L * c->g.L;
        if (c->tune.dag_edges == 2)
            k_dag_edges_pairs<<<blocks(items), TPB, 0, c->stream>>>(a, ne);
        else
            k_dag_edges_wave<<<(unsigned)std::min<int64_t>(8192, (ne + EW_WAVES - 1) / EW_WAVES), TPB, 0,
                               c->stream>>>(a, ne);
        HIP_OK(hipGetLastError());
    }
    NetLevels lv{D.nl_off.p, D.nlevels.p, D.lvl_node_off.p, D.lvl_edge_off.p, D.dnode.p, D.dedge.p};
    KScope ks(c, K_DAG_DP, b.bytes_dp);
    // levels with at least seq_min nodes run a node per lane, narrower ones
    // 16 lanes per node
    const int seq_min = c->tune.dag_seq ? c->tune.dag_seq : DEF_DAG_SEQ;
    const unsigned grid = (unsigned)std::min<int64_t>(n, c->tune.dag_grid ? c->tune.dag_grid : DEF_DAG_GRID);
    int* q = D.queue.p + b.index;
    if (c->tune.dag_waves == 4)
        k_dag_net<4><<<grid, WTPB, 0, c->stream>>>(a, lv, D.nets.p, D.glist.p + b.net0, n, q, D.node_off.p,
                                                   D.route_slot.p, c->nets.dag_routes.back()->p, c->nets.rptr.p,
                                                   seq_min);
    else
        k_dag_net<2><<<grid, WTPB, 0, c->stream>>>(a, lv, D.nets.p, D.glist.p + b.net0, n, q, D.node_off.p,
                                                   D.route_slot.p, c->nets.dag_routes.back()->p, c->nets.rptr.p,
                                                   seq_min);
    HIP_OK(hipGetLastError());
}

void launch_dag_queue_reset(igr_ctx* c, int64_t batches) {
    c->dag.queue.alloc(std::max<int64_t>(1, batches));
    k_fill_i32<<<blocks(std::max<int64_t>(1, batches)), TPB, 0, c->stream>>>(c->dag.queue.p, batches, 0);
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
