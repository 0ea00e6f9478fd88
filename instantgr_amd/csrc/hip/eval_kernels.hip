// GPU evaluator: connectivity check and FP64 score of routes, with the
// definitions of run/evaluator.cpp:458-633,635-675 (the same ones the host
// evaluator, csrc/host/evaluator.cpp, applies to a route file), computed from
// the routes in device memory instead of a written and parsed route file.
//
// Layers here are metal layers z = 0..ML-1 (routing layer + 1), cells
// z*X*Y + x*Y + y.  The lines the route file would hold for a net are its
// metal-1 pin vias, its stacked-pin via pairs and its subnets' route pairs
// shifted up one layer (src/graph.hpp:179-235).  Each line becomes one item:
//   * a wire: one layer z >= 1, along the layer's direction, cells lo..hi
//     (hi > lo; wire demand on lo..hi-1, the net present on lo..hi);
//   * a via column: one gcell, layers z0..z1 (a via on each of z0..z1-1, the
//     net present on every layer of the column);
//   * dead: a malformed line, which makes the net open and adds nothing.
// Per net, with its items:
//   * stacked vias: a via's lower cell adds demand (split between the cell and
//     its neighbour along the layer direction) once per net and cell, where
//     none of the net's wires is present;
//   * connectivity: the cells where the net is present are joined along the
//     layer direction on layers above 0 and between adjacent layers of one
//     gcell (the evaluator's search).  Two items are joined when a cell of one
//     is or neighbours a cell of the other in that sense, so union-find over
//     items gives the search's components.  A net is connected when every pin
//     after the first has an access point in a component that holds an access
//     point of the first pin.
// Demand per cell is 2 * wires + stacked vias in 16 bits; the score sums
// wirelength, vias and the overflow cost in FP64 (in another order than the
// host evaluator, which changes the last bits of the sums only).
#include "igr_ctx.hpp"

#include <algorithm>
#include <cmath>
#include <string>

namespace igr_dev {

namespace {

constexpr int ETPB = 256;
constexpr int MAXML = 64;         // metal layers the per-block via counters and via masks hold
constexpr int BIG_NET = 512;      // nets with more items: a workgroup each, items staged in LDS tiles
constexpr int BTILE = 256;        // items per LDS tile (= threads of the workgroup)

struct Item {
    int x0, x1, y0, y1, z0, z1;   // z0 > z1: dead
};

struct EArgs {
    int64_t nn;                   // nets
    int ML, X, Y;
    int64_t XY;
    const int* dir;               // per metal layer
    const int64_t* sub_off;
    const int* subs;
    const int64_t* pin_off;       // selected pins (metal-1 cells get a via up)
    const int* pins;
    const int64_t* extra_off;
    const int* extra;
    const int64_t* dpin_off;      // design pins and their access points
    const int64_t* ap_off;
    const int* ap;
    const int64_t* r_off;         // per subnet: its route, r[0] = ints used
    const int* routes;
    const int64_t* it_off;        // per net: first item
    Item* items;
    int* par;                     // per item: union-find parent (an item of lower index, or itself)
    int* traced;                  // per item: a root reached from the first pin
    int* wire;                    // ML*X*Y
    int* svia;                    // ML*X*Y
    unsigned long long* vias;     // ML
    unsigned long long* opens;
    int* bad;                     // per net
};

__device__ __forceinline__ Item decode(const EArgs& a, int p, int q) {
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
    const Item dead{0, 0, 0, 0, 1, 0};
    if (p < 0 || q >= (int64_t)a.ML * a.XY) return dead;
    const int za = (int)(p / a.XY), xa = p / a.Y % a.X, ya = p % a.Y;
    const int zb = (int)(q / a.XY), xb = q / a.Y % a.X, yb = q % a.Y;
    if (za != zb) return xa == xb && ya == yb ? Item{xa, xa, ya, ya, za, zb} : dead;
    if (za == 0) return dead;
    if (a.dir[za] == 0) return xb > xa && yb == ya ? Item{xa, xb, ya, ya, za, za} : dead;
    return yb > ya && xb == xa ? Item{xa, xa, ya, yb, za, za} : dead;
}

__device__ __forceinline__ bool is_wire(const Item& t) { return t.z0 == t.z1 && (t.x1 > t.x0 || t.y1 > t.y0); }

__device__ __forceinline__ bool contains(const Item& t, int x, int y, int z) {
    return z >= t.z0 && z <= t.z1 && x >= t.x0 && x <= t.x1 && y >= t.y0 && y <= t.y1;
}

// Is a cell of one item the same as, or a search neighbour of, a cell of the other?
__device__ bool joined(const EArgs& a, const Item& s, const Item& t) {
    if (s.z0 > s.z1 || t.z0 > t.z1) return false;
    const bool xo = s.x0 <= t.x1 && t.x0 <= s.x1, yo = s.y0 <= t.y1 && t.y0 <= s.y1;
    const int zl = max(s.z0, t.z0), zh = min(s.z1, t.z1);
    if (zl > zh) {
        // no common layer: adjacent layers of one gcell
        return (s.z1 + 1 == t.z0 || t.z1 + 1 == s.z0) && xo && yo;
    }
    if (xo && yo) return true;   // a common cell
    // neighbours along the direction of a common layer above 0
    const bool xn = s.x0 <= t.x1 + 1 && t.x0 <= s.x1 + 1, yn = s.y0 <= t.y1 + 1 && t.y0 <= s.y1 + 1;
    for (int z = max(zl, 1); z <= zh; z++) {
        if (a.dir[z] == 0 ? (yo && xn) : (xo && yn)) return true;
    }
    return false;
}

__device__ __forceinline__ int uf_load(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Roots have par[x] == x; every other item points to an item of lower index
// on its path to the root, so halving with a minimum keeps that order.
__device__ int uf_find(int* par, int x) {
    for (;;) {
        const int p = uf_load(&par[x]);
        if (p == x) return x;
        const int g = uf_load(&par[p]);
        if (g != p) atomicMin(&par[x], g);
        x = p;
    }
}

__device__ void uf_union(int* par, int x, int y) {
    for (;;) {
        x = uf_find(par, x);
        y = uf_find(par, y);
        if (x == y) return;
        if (x > y) {
            const int t = x;
            x = y;
            y = t;
        }
        if (atomicCAS(&par[y], y, x) == y) return;
    }
}

// A counted via's lower cell (z, x, y): its demand is split between the cell
// and its neighbour along the layer direction (run/evaluator.cpp:635-675).
__device__ void add_svia(const EArgs& a, int x, int y, int z) {
    const int X = a.X, Y = a.Y;
    int* sv = a.svia + z * a.XY;
    if (a.dir[z] == 0) {
        if (x > 0 && x < X - 1) {
            atomicAdd(&sv[(int64_t)(x - 1) * Y + y], 1);
            atomicAdd(&sv[(int64_t)x * Y + y], 1);
        } else if (x > 0) {
            atomicAdd(&sv[(int64_t)(x - 1) * Y + y], 2);
        } else if (x < X - 1) {
            atomicAdd(&sv[(int64_t)x * Y + y], 2);
        }
    } else {
        if (y > 0 && y < Y - 1) {
            atomicAdd(&sv[(int64_t)x * Y + y - 1], 1);
            atomicAdd(&sv[(int64_t)x * Y + y], 1);
        } else if (y > 0) {
            atomicAdd(&sv[(int64_t)x * Y + y - 1], 2);
        } else if (y < Y - 1) {
            atomicAdd(&sv[(int64_t)x * Y + y], 2);
        }
    }
}

// Bits lo..hi (lo <= hi < 64).
__device__ __forceinline__ uint64_t bits(int lo, int hi) {
    const int w = hi - lo + 1;
    return (w >= 64 ? ~0ull : ((1ull << w) - 1)) << lo;
}

// Items per net: metal-1 pin vias, stacked-pin via pairs, route pairs.
__global__ void k_ev_count(EArgs a, int64_t* cnt) {
    const int64_t n = (int64_t)blockIdx.x * ETPB + threadIdx.x;
    if (n >= a.nn) return;
    int64_t c = (a.extra_off[n + 1] - a.extra_off[n]) / 2;
    for (int64_t p = a.pin_off[n]; p < a.pin_off[n + 1]; p++) c += a.pins[p] < a.XY;
    for (int64_t k = a.sub_off[n]; k < a.sub_off[n + 1]; k++) c += (a.routes[a.r_off[a.subs[k]]] - 1) / 2;
    cnt[n] = c;
}

// A thread per net writes its items in line order and marks malformed lines.
__global__ void k_ev_fill(EArgs a) {
    const int64_t n = (int64_t)blockIdx.x * ETPB + threadIdx.x;
    if (n >= a.nn) return;
    Item* it = a.items + a.it_off[n];
    int64_t m = 0;
    bool bad = false;
    const int XY = (int)a.XY;
    for (int64_t p = a.pin_off[n]; p < a.pin_off[n + 1]; p++)
        if (a.pins[p] < XY) it[m++] = decode(a, a.pins[p], a.pins[p] + XY);
    for (int64_t e = a.extra_off[n]; e + 1 < a.extra_off[n + 1]; e += 2) it[m++] = decode(a, a.extra[e], a.extra[e + 1]);
    for (int64_t k = a.sub_off[n]; k < a.sub_off[n + 1]; k++) {
        const int* r = a.routes + a.r_off[a.subs[k]];
        for (int j = 1; j + 1 < r[0]; j += 2) it[m++] = decode(a, r[j] + XY, r[j + 1] + XY);
    }
    for (int64_t i = 0; i < m; i++) {
        bad = bad || it[i].z0 > it[i].z1;
        a.par[a.it_off[n] + i] = (int)i;
        a.traced[a.it_off[n] + i] = 0;
    }
    a.bad[n] = bad;
}

// A thread per item: wire demand and via counts (per block in LDS).
__global__ void k_ev_wires(EArgs a, int64_t total) {
    __shared__ unsigned long long sv[MAXML];
    for (int z = threadIdx.x; z < a.ML; z += ETPB) sv[z] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * ETPB + threadIdx.x; i < total; i += (int64_t)gridDim.x * ETPB) {
        const Item t = a.items[i];
        if (t.z0 > t.z1) continue;
        if (t.z0 != t.z1) {
            for (int z = t.z0; z < t.z1; z++) atomicAdd(&sv[z], 1ull);
        } else if (t.x1 > t.x0) {
            int* w = a.wire + t.z0 * a.XY + t.y0;
            for (int x = t.x0; x < t.x1; x++) atomicAdd(&w[(int64_t)x * a.Y], 1);
        } else {
            int* w = a.wire + t.z0 * a.XY + (int64_t)t.x0 * a.Y;
            for (int y = t.y0; y < t.y1; y++) atomicAdd(&w[y], 1);
        }
    }
    __syncthreads();
    for (int z = threadIdx.x; z < a.ML; z += ETPB)
        if (sv[z]) atomicAdd(&a.vias[z], sv[z]);
}

// A wave per net: stacked vias, components, the connectivity verdict.
__global__ void __launch_bounds__(ETPB) k_ev_net(EArgs a) {
    const int lane = threadIdx.x & 63;
    const int X = a.X, Y = a.Y;
    for (int64_t n = ((int64_t)blockIdx.x * ETPB + threadIdx.x) >> 6; n < a.nn;
         n += ((int64_t)gridDim.x * ETPB) >> 6) {
        const int64_t base = a.it_off[n];
        const int m = (int)(a.it_off[n + 1] - base);
        if (m > BIG_NET) continue;   // k_ev_net_big (the same for the whole wave)
        const Item* it = a.items + base;
        int* par = a.par + base;
        int* tr = a.traced + base;
        // Stacked vias: the lower cell of each via of a column, unless an
        // earlier column holds the same via or a wire of the net is present.
        for (int i = lane; i < m; i += 64) {
            const Item c = it[i];
            if (c.z0 >= c.z1) continue;   // wires and dead lines
            for (int z = max(c.z0, 1); z < c.z1; z++) {
                bool skip = false;
                for (int j = 0; j < m && !skip; j++) {
                    const Item t = it[j];
                    if (j < i && t.z0 < t.z1 && t.x0 == c.x0 && t.y0 == c.y0 && z >= t.z0 && z < t.z1) skip = true;
                    else if (is_wire(t) && t.z0 == z && contains(t, c.x0, c.y0, z)) skip = true;
                }
                if (skip) continue;
                add_svia(a, c.x0, c.y0, z);
            }
        }
        const int64_t p0 = a.dpin_off[n], p1 = a.dpin_off[n + 1];
        if (p1 - p0 < 2) {
            if (lane == 0 && a.bad[n]) atomicAdd(a.opens, 1ull);
            continue;   // a single pin is connected unless a line is malformed
        }
        // Components.
        for (int i = 0; i < m; i++) {
            const Item s = it[i];
            for (int j = i + 1 + lane; j < m; j += 64)
                if (joined(a, s, it[j])) uf_union(par, i, j);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        // Roots of the components holding an access point of the first pin.
        for (int64_t k = a.ap_off[p0]; k < a.ap_off[p0 + 1]; k++) {
            const int v = a.ap[k];
            const int z = (int)(v / a.XY), x = v / Y % X, y = v % Y;
            for (int j = lane; j < m; j += 64)
                if (contains(it[j], x, y, z)) atomicExch(&tr[uf_find(par, j)], 1);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        bool ok = !a.bad[n];
        for (int64_t p = p0 + 1; p < p1 && ok; p++) {
            bool hit = false;
            for (int64_t k = a.ap_off[p]; k < a.ap_off[p + 1]; k++) {
                const int v = a.ap[k];
                const int z = (int)(v / a.XY), x = v / Y % X, y = v % Y;
                for (int j = lane; j < m && !hit; j += 64)
                    hit = contains(it[j], x, y, z) && uf_load(&tr[uf_find(par, j)]) != 0;
            }
            ok = __ballot(hit) != 0;
        }
        if (lane == 0 && !ok) atomicAdd(a.opens, 1ull);
    }
}

// A workgroup per net with more than BIG_NET items: the same work as
// k_ev_net, with the item pairs compared through LDS tiles.  A thread keeps
// one item (or one via column's counted-via mask) in registers and compares
// it with each tile, so global loads grow with m * m / BTILE, not m * m.
__global__ void __launch_bounds__(BTILE) k_ev_net_big(EArgs a, const int* big, int nbig) {
    __shared__ Item tile[BTILE];
    __shared__ int s_hit;
    const int tid = threadIdx.x;
    const int X = a.X, Y = a.Y;
    const Item dead{0, 0, 0, 0, 1, 0};
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {   // the same for the whole workgroup
        const int n = big[b];
        const int64_t base = a.it_off[n];
        const int m = (int)(a.it_off[n + 1] - base);
        const Item* it = a.items + base;
        int* par = a.par + base;
        int* tr = a.traced + base;
        // Stacked vias: bit z of 'skip' drops the via on layer z of column i.
        for (int c0 = 0; c0 < m; c0 += BTILE) {
            const int i = c0 + tid;
            const Item c = i < m ? it[i] : dead;
            const bool col = c.z0 < c.z1;
            uint64_t skip = 0;
            for (int t0 = 0; t0 < m; t0 += BTILE) {
                __syncthreads();
                if (t0 + tid < m) tile[tid] = it[t0 + tid];
                __syncthreads();
                const int tn = min(BTILE, m - t0);
                if (!col) continue;
                for (int k = 0; k < tn; k++) {
                    const Item t = tile[k];
                    if (t.x0 > c.x0 || t.x1 < c.x0 || t.y0 > c.y0 || t.y1 < c.y0) continue;
                    if (t.z0 < t.z1) {
                        // an earlier column over the same gcell holds the vias t.z0..t.z1-1
                        if (t0 + k < i) {
                            const int lo = max(t.z0, c.z0), hi = min(t.z1, c.z1) - 1;
                            if (lo <= hi) skip |= bits(lo, hi);
                        }
                    } else if (is_wire(t) && t.z0 >= c.z0 && t.z0 < c.z1) {
                        skip |= 1ull << t.z0;   // a wire of the net on that layer
                    }
                }
            }
            if (col)
                for (int z = max(c.z0, 1); z < c.z1; z++)
                    if (!((skip >> z) & 1)) add_svia(a, c.x0, c.y0, z);
        }
        const int64_t p0 = a.dpin_off[n], p1 = a.dpin_off[n + 1];
        if (p1 - p0 < 2) {
            if (tid == 0 && a.bad[n]) atomicAdd(a.opens, 1ull);
            continue;
        }
        // Components: item i against every later item, a tile at a time.
        for (int c0 = 0; c0 < m; c0 += BTILE) {
            const int i = c0 + tid;
            const Item s = i < m ? it[i] : dead;
            for (int t0 = c0; t0 < m; t0 += BTILE) {
                __syncthreads();
                if (t0 + tid < m) tile[tid] = it[t0 + tid];
                __syncthreads();
                const int tn = min(BTILE, m - t0);
                if (i >= m) continue;
                for (int k = 0; k < tn; k++)
                    if (t0 + k > i && joined(a, s, tile[k])) uf_union(par, i, t0 + k);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        __syncthreads();
        for (int64_t k = a.ap_off[p0]; k < a.ap_off[p0 + 1]; k++) {
            const int v = a.ap[k];
            const int z = (int)(v / a.XY), x = v / Y % X, y = v % Y;
            for (int j = tid; j < m; j += BTILE)
                if (contains(it[j], x, y, z)) atomicExch(&tr[uf_find(par, j)], 1);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        __syncthreads();
        bool ok = !a.bad[n];
        for (int64_t p = p0 + 1; p < p1 && ok; p++) {
            if (tid == 0) s_hit = 0;
            __syncthreads();
            bool hit = false;
            for (int64_t k = a.ap_off[p]; k < a.ap_off[p + 1] && !hit; k++) {
                const int v = a.ap[k];
                const int z = (int)(v / a.XY), x = v / Y % X, y = v % Y;
                for (int j = tid; j < m && !hit; j += BTILE)
                    hit = contains(it[j], x, y, z) && uf_load(&tr[uf_find(par, j)]) != 0;
            }
            if (hit) s_hit = 1;
            __syncthreads();
            ok = s_hit != 0;
            __syncthreads();
        }
        if (tid == 0 && !ok) atomicAdd(a.opens, 1ull);
    }
}

// Score partials per block: overflow cost (FP64) and wirelength (integer
// length units), over the cells of metal layers 1..ML-1.
__global__ void k_ev_score(EArgs a, const double* cap, const double* weight, const int* xlen, const int* ylen,
                           double* part_of, unsigned long long* part_wl) {
    __shared__ double s_of[ETPB];
    __shared__ unsigned long long s_wl[ETPB];
    double of = 0;
    unsigned long long wl = 0;
    const int64_t cells = (int64_t)(a.ML - 1) * a.XY;
    for (int64_t i = (int64_t)blockIdx.x * ETPB + threadIdx.x; i < cells; i += (int64_t)gridDim.x * ETPB) {
        const int64_t k = a.XY + i;
        const int z = (int)(k / a.XY);
        const int x = (int)(k / a.Y % a.X), y = (int)(k % a.Y);
        const unsigned demand = (unsigned)(2 * a.wire[k] + a.svia[k]) & 0xFFFFu;
        const double c = cap[k];
        double t = 0;
        if (c > 0.001) t = exp((double(demand) - 2 * c) / 2 * 0.5);
        else if (c >= 0 && demand > 0) t = exp(1.5 * double(demand) * 0.5);
        of += t * weight[z];
        const int len = a.dir[z] == 0 ? (x + 1 < a.X ? xlen[x] : 0) : (y + 1 < a.Y ? ylen[y] : 0);
        wl += (unsigned long long)a.wire[k] * (unsigned long long)len;
    }
    s_of[threadIdx.x] = of;
    s_wl[threadIdx.x] = wl;
    __syncthreads();
    for (int s = ETPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_of[threadIdx.x] += s_of[threadIdx.x + s];
            s_wl[threadIdx.x] += s_wl[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_of[blockIdx.x] = s_of[0];
        part_wl[blockIdx.x] = s_wl[0];
    }
}

__global__ void k_ev_zero(int* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * ETPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ETPB) p[i] = 0;
}

inline unsigned eblocks(int64_t n, int64_t cap = 65535) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, (n + ETPB - 1) / ETPB));
}

} // namespace

void evaluate_routes_dev(igr_ctx* c, const igr_eval_nets* e, const int64_t* h_route_off, const int* h_routes,
                         igr_score* out) {
    const int ML = c->g.L + 1, X = c->g.X, Y = c->g.Y;
    const int64_t XY = (int64_t)X * Y, NN = e->nets, S = c->nets.n;
    if (ML > MAXML) throw std::invalid_argument("too many metal layers for the evaluator");
    if (NN <= 0) throw std::invalid_argument("no nets");
    for (int z = 0; z < ML; z++)
        if (e->dir[z] != 0 && e->dir[z] != 1) throw std::invalid_argument("layer direction must be 0 or 1");
    auto monotone = [](const int64_t* off, int64_t n, const char* what) {
        if (off[0] != 0) throw std::invalid_argument(std::string(what) + " offsets must start at 0");
        for (int64_t i = 0; i < n; i++)
            if (off[i + 1] < off[i]) throw std::invalid_argument(std::string(what) + " offsets must not decrease");
    };
    monotone(e->sub_off, NN, "subnet");
    monotone(e->pin_off, NN, "pin");
    monotone(e->extra_off, NN, "stacked-via");
    monotone(e->dpin_off, NN, "design-pin");
    monotone(e->ap_off, e->dpin_off[NN], "access-point");
    const int64_t nsub = e->sub_off[NN];
    for (int64_t k = 0; k < nsub; k++)
        if (e->subs[k] < 0 || e->subs[k] >= S) throw std::invalid_argument("subnet index out of range");
    const int64_t npin = e->pin_off[NN], nextra = e->extra_off[NN], ndp = e->dpin_off[NN], nap = e->ap_off[ndp];
    for (int64_t p = 0; p < npin; p++)
        if (e->pins[p] < 0 || e->pins[p] >= ML * XY) throw std::invalid_argument("pin cell out of range");
    for (int64_t p = 0; p < nextra; p++)
        if (e->extra[p] < 0 || e->extra[p] >= ML * XY) throw std::invalid_argument("extra via cell out of range");
    for (int64_t p = 0; p < nap; p++)
        if (e->ap[p] < 0 || e->ap[p] >= ML * XY) throw std::invalid_argument("access point out of range");
    hipStream_t st = c->stream;
    // Routes: the caller's, or the context's gathered into one array.
    DevBuf<int64_t> r_off;
    DevBuf<int> routes;
    if (h_routes) {
        std::vector<int64_t> o(h_route_off, h_route_off + S + 1);
        for (int64_t i = 0; i < S; i++) {
            const int used = h_routes[o[i]];
            if (o[i + 1] < o[i] || used < 1 || o[i] + used > o[i + 1] || (used - 1) % 2)
                throw std::invalid_argument("malformed route slice");
            for (int j = 1; j < used; j++)
                if (h_routes[o[i] + j] < 0 || h_routes[o[i] + j] >= (ML - 1) * XY)
                    throw std::invalid_argument("route cell out of range");
        }
        r_off.upload(o.data(), S + 1, st);
        routes.upload(h_routes, o[S], st);
    } else {
        std::vector<int64_t> o(S + 1, 0);
        for (int64_t i = 0; i < S; i++) o[i + 1] = o[i] + c->nets.h_rcap[i];
        r_off.upload(o.data(), S + 1, st);
        routes.alloc(std::max<int64_t>(1, o[S]));
        gather_routes(c, r_off.p, routes.p);
    }
    DevBuf<int> dir, subs, pins, extra, ap, wire, svia, bad, par, traced;
    DevBuf<int64_t> sub_off, pin_off, extra_off, dpin_off, ap_off, it_off, cnt;
    DevBuf<double> cap, weight, part_of;
    DevBuf<unsigned long long> vias, opens, part_wl;
    DevBuf<Item> items;
    dir.upload(e->dir, ML, st);
    sub_off.upload(e->sub_off, NN + 1, st);
    subs.upload(e->subs, std::max<int64_t>(1, nsub), st);
    pin_off.upload(e->pin_off, NN + 1, st);
    pins.upload(e->pins, std::max<int64_t>(1, npin), st);
    extra_off.upload(e->extra_off, NN + 1, st);
    extra.upload(e->extra, std::max<int64_t>(1, nextra), st);
    dpin_off.upload(e->dpin_off, NN + 1, st);
    ap_off.upload(e->ap_off, ndp + 1, st);
    ap.upload(e->ap, std::max<int64_t>(1, nap), st);
    cap.upload(e->capacity, ML * XY, st);
    weight.upload(e->of_weight, ML, st);
    EArgs a{NN, ML, X, Y, XY, dir.p, sub_off.p, subs.p, pin_off.p, pins.p, extra_off.p, extra.p, dpin_off.p,
            ap_off.p, ap.p, r_off.p, routes.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, nullptr};
    cnt.alloc(NN);
    k_ev_count<<<(unsigned)((NN + ETPB - 1) / ETPB), ETPB, 0, st>>>(a, cnt.p);
    HIP_OK(hipGetLastError());
    std::vector<int64_t> hc(NN), ho(NN + 1, 0);
    HIP_OK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int64_t) * NN, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int64_t n = 0; n < NN; n++) {
        if (hc[n] >= 0x7fffffff) throw std::invalid_argument("a net has too many route lines for the evaluator");
        ho[n + 1] = ho[n] + hc[n];
    }
    const int64_t total = ho[NN];
    it_off.upload(ho.data(), NN + 1, st);
    items.alloc(std::max<int64_t>(1, total));
    par.alloc(std::max<int64_t>(1, total));
    traced.alloc(std::max<int64_t>(1, total));
    wire.alloc(ML * XY);
    svia.alloc(ML * XY);
    bad.alloc(NN);
    vias.zero(ML, st);
    opens.zero(1, st);
    k_ev_zero<<<eblocks(ML * XY, 4096), ETPB, 0, st>>>(wire.p, ML * XY);
    k_ev_zero<<<eblocks(ML * XY, 4096), ETPB, 0, st>>>(svia.p, ML * XY);
    a.it_off = it_off.p;
    a.items = items.p;
    a.par = par.p;
    a.traced = traced.p;
    a.wire = wire.p;
    a.svia = svia.p;
    a.vias = vias.p;
    a.opens = opens.p;
    a.bad = bad.p;
    k_ev_fill<<<(unsigned)((NN + ETPB - 1) / ETPB), ETPB, 0, st>>>(a);
    k_ev_wires<<<eblocks(total, 4096), ETPB, 0, st>>>(a, total);
    k_ev_net<<<eblocks(NN * 64, 8192), ETPB, 0, st>>>(a);
    std::vector<int> hbig;
    for (int64_t n = 0; n < NN; n++)
        if (hc[n] > BIG_NET) hbig.push_back((int)n);
    DevBuf<int> big;
    if (!hbig.empty()) {
        big.upload(hbig.data(), (int64_t)hbig.size(), st);
        k_ev_net_big<<<(unsigned)std::min<size_t>(hbig.size(), 4096), BTILE, 0, st>>>(a, big.p, (int)hbig.size());
    }
    const unsigned sb = eblocks((ML - 1) * XY, 2048);
    part_of.alloc(sb);
    part_wl.alloc(sb);
    k_ev_score<<<sb, ETPB, 0, st>>>(a, cap.p, weight.p, c->xlen.p, c->ylen.p, part_of.p, part_wl.p);
    HIP_OK(hipGetLastError());
    std::vector<double> hof(sb);
    std::vector<unsigned long long> hwl(sb), hv(ML), ho1(1);
    HIP_OK(hipMemcpyAsync(hof.data(), part_of.p, sizeof(double) * sb, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hwl.data(), part_wl.p, sizeof(unsigned long long) * sb, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hv.data(), vias.p, sizeof(unsigned long long) * ML, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(ho1.data(), opens.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double of = 0, via = 0;
    unsigned long long wl = 0;
    for (unsigned b = 0; b < sb; b++) {
        of += hof[b];
        wl += hwl[b];
    }
    for (int z = 0; z < ML; z++) via += (double)hv[z] * c->g.unit_via;
    out->wirelength = (double)wl * c->g.unit_wire;
    out->via = via;
    out->overflow = of;
    out->total = out->overflow + out->via + out->wirelength;
    out->opens = (int64_t)ho1[0];
}

} // namespace igr_dev
