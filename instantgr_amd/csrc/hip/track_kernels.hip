// Batched form of the grid updates: one fused pass per dirty track.
//
// Between two routed batches the reference runs, over the whole grid, the
// wire-demand scan and commit (src/graph.hpp:104-137), the via commit
// (:496-543), for rip-up batches the wire and via rip-up (:468-543 with
// K = -1), then the logistic cost update (:40-71) and the FP64 prefix sums
// (:405-430).  Each of these only touches cells of one track (a via cut at
// (l, x, y) adds demand to the two edges of that cell's track on layer l), so
// one workgroup per track does them all in registers, in the reference's
// order, and streams each array once:
//   1. wire commit: inclusive scan of the difference array, demand += count,
//      timestamp = commit stamp on covered edges;
//   2. via commit: cuts whose cell timestamp is older than the commit stamp
//      add half a track to each adjacent edge (a whole one at track ends);
//   3. rip-up of the next batch: demand -= covering wire count, timestamp =
//      rip stamp on covered cells, then the gated via cuts subtract;
//   4. wcost / vcost from the new demand, FP64 prefix sums of wcost.
// Demand changes by integers and halves, so the result equals the sequential
// one.  A track no scatter touched keeps its demand and so its costs: the
// routing kernels and the rip-up scatter flag every track they write into.
#include "igr_ctx.hpp"
#include "scan.h"

namespace igr_dev {

namespace {

constexpr int SB = 256;   // block size of the scatter and fill kernels

// Value of the first element of the next thread (0 past the last thread).
template <typename T, int B>
__device__ __forceinline__ T next_first(T first, T* sh) {
    __syncthreads();
    sh[threadIdx.x] = first;
    __syncthreads();
    return threadIdx.x + 1 < B ? sh[threadIdx.x + 1] : T(0);
}
// Value of the last element of the previous thread (0 for thread 0).
template <typename T>
__device__ __forceinline__ T prev_last(T last, T* sh) {
    __syncthreads();
    sh[threadIdx.x] = last;
    __syncthreads();
    return threadIdx.x > 0 ? sh[threadIdx.x - 1] : T(0);
}

// Via cuts gv at the cells apply sign * (1 or 1/2) to the adjacent edges.
template <int K>
__device__ __forceinline__ void apply_vias(float (&dem)[K], const int (&gv)[K], int gnext, int len, float sign) {
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int e = threadIdx.x * K + j;
        if (e > len - 2) continue;
        const int g0 = gv[j];
        const int g1 = j + 1 < K ? gv[j + 1] : gnext;
        if (g0) dem[j] += sign * (e == 0 ? (float)g0 : 0.5f * (float)g0);
        if (g1) dem[j] += sign * (e + 1 == len - 1 ? (float)g1 : 0.5f * (float)g1);
    }
}

// s_c: commit stamp (0 = nothing to commit); s_r: rip-up stamp (0 = no
// rip-up); cost: recompute costs and prefix sums (and clear the dirty flag).
// stats[0] += cells of the tracks processed, stats[1] += those with rip-up.
template <int K, int B>
__global__ void __launch_bounds__(B) k_track_fused(GridDev g, int s_c, int s_r, int cost,
                                                   unsigned long long* stats) {
    constexpr int NW = B / 64;
    __shared__ int sh_i[B];
    __shared__ float sh_f[B];
    __shared__ double sh_d[B];
    __shared__ int lds_i[NW];
    __shared__ double lds_d[NW];
    int t = blockIdx.x;
    if (!g.dirty[t]) return;
    int l = 0;
    while (l < g.L - 1 && t >= g.tbase[l + 1]) l++;
    t -= g.tbase[l];
    const int dir = layer_dir(l, g.DIR);
    const int len = dir ? g.Y : g.X;
    const int64_t base = (int64_t)l * g.XYc + (int64_t)t * len;
    float dem[K], cap[K];
    int ts[K];
    uint32_t wr_dem = 0, wr_ts = 0, clr = 0;   // per-element bit masks
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = threadIdx.x * K + j;
        const bool in = p < len;
        dem[j] = in ? g.demand[base + p] : 0.0f;
        cap[j] = in ? g.cap[base + p] : 0.0f;
        ts[j] = in ? g.ts[base + p] : 0;
    }
    if (s_c) {
        int inc[K], gv[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int p = threadIdx.x * K + j;
            inc[j] = p < len ? g.pre[base + p] : 0;
            gv[j] = p < len ? g.vc[base + p] : 0;
            if (inc[j] | gv[j]) clr |= 1u << j;
        }
        block_sklansky<int, K, NW>(inc, lds_i);
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int p = threadIdx.x * K + j;
            if (p <= len - 2 && inc[j] > 0) {
                dem[j] += (float)inc[j];
                wr_dem |= 1u << j;
                if (ts[j] != s_c) { ts[j] = s_c; wr_ts |= 1u << j; }
            }
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (ts[j] >= s_c) gv[j] = 0;
            if (gv[j]) wr_dem |= (3u << j) >> 1;  // edges j-1 and j of this thread
        }
        const int gn = next_first<int, B>(gv[0], sh_i);
        if (gn) wr_dem |= 1u << (K - 1);
        apply_vias<K>(dem, gv, gn, len, 1.0f);
    }
    if (s_r) {
        int dec[K], cov[K], gr[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int p = threadIdx.x * K + j;
            const bool in = p < len;
            dec[j] = in ? g.rpre[base + p] : 0;
            cov[j] = in ? g.rcov[base + p] : 0;
            gr[j] = in ? g.rvc[base + p] : 0;
            if (dec[j] | cov[j] | gr[j]) clr |= 1u << j;
        }
        block_sklansky<int, K, NW>(dec, lds_i);
        block_sklansky<int, K, NW>(cov, lds_i);
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int p = threadIdx.x * K + j;
            if (p <= len - 2 && dec[j] > 0) { dem[j] -= (float)dec[j]; wr_dem |= 1u << j; }
            if (p < len && cov[j] > 0 && ts[j] != s_r) { ts[j] = s_r; wr_ts |= 1u << j; }
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (ts[j] >= s_r) gr[j] = 0;
            if (gr[j]) wr_dem |= (3u << j) >> 1;
        }
        const int gn = next_first<int, B>(gr[0], sh_i);
        if (gn) wr_dem |= 1u << (K - 1);
        apply_vias<K>(dem, gr, gn, len, -1.0f);
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = threadIdx.x * K + j;
        if (p >= len) continue;
        if (wr_dem >> j & 1) g.demand[base + p] = dem[j];
        if (wr_ts >> j & 1) g.ts[base + p] = ts[j];
        if (clr >> j & 1) {
            if (s_c) { g.pre[base + p] = 0; g.vc[base + p] = 0; }
            if (s_r) { g.rpre[base + p] = 0; g.rcov[base + p] = 0; g.rvc[base + p] = 0; }
        }
    }
    if (threadIdx.x == 0) {
        atomicAdd(&stats[0], (unsigned long long)len);
        if (s_r) atomicAdd(&stats[1], (unsigned long long)len);
    }
    if (!cost) return;
    // Costs from the new demand (src/graph.hpp:40-71).
    const double w = g.w[l];
    const int* elen = dir ? g.ylen : g.xlen;
    const float dprev = prev_last(dem[K - 1], sh_f);
    const float cprev = prev_last(cap[K - 1], sh_f);
    float wc[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = threadIdx.x * K + j;
        if (p >= len) { wc[j] = 0.0f; continue; }
        float x = (float)inc_cost(cap[j], dem[j], 1.0, w);
        if (p < len - 1) x = (float)((double)x + g.unit_wire * (double)elen[p]);
        wc[j] = x;
        const float dp = j > 0 ? dem[j - 1] : dprev, cp = j > 0 ? cap[j - 1] : cprev;
        float v;
        if (p == 0) v = (float)inc_cost(cap[j], dem[j], 1.0, w);
        else if (p == len - 1) v = (float)inc_cost(cp, dp, 1.0, w);
        else v = (float)(inc_cost(cap[j], dem[j], 0.5, w) + inc_cost(cp, dp, 0.5, w));
        g.wcost[base + p] = wc[j];
        g.vcost[base + p] = (float)((double)v + g.unit_via);
    }
    // Exclusive FP64 Sklansky prefix of wcost (src/graph.hpp:405-430).
    const double wprev = prev_last((double)wc[K - 1], sh_d);
    double s[K];
#pragma unroll
    for (int j = K - 1; j >= 0; j--) {
        const int p = threadIdx.x * K + j;
        s[j] = (p >= 1 && p < len) ? (j > 0 ? (double)wc[j - 1] : wprev) : 0.0;
    }
    block_sklansky<double, K, NW>(s, lds_d);
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = threadIdx.x * K + j;
        if (p < len) g.presum[base + p] = s[j];
    }
    if (threadIdx.x == 0) g.dirty[blockIdx.x] = 0;
}

// Rip-up footprint of routed nets: +1/-1 wire differences at the two ends of
// each wire, +1/-1 covered-cell differences at its first cell and one past
// its last, one count per via cut, and the dirty flag of every track touched.
// Wires that do not run along their layer's direction set *bad.
__global__ void __launch_bounds__(SB) k_rip_scatter(GridDev g, const int* nets, int64_t n, int* const* rptr,
                                                    int* bad) {
    const int64_t t = (int64_t)blockIdx.x * SB + threadIdx.x;
    if (t >= n) return;
    const int* r = rptr[nets[t]];
    const int XY = g.X * g.Y;
    for (int i = 1; i < r[0]; i += 2) {
        const int a = r[i], b = r[i + 1];
        const int la = a / XY, lb = b / XY;
        const int x0 = a / g.Y % g.X, y0 = a % g.Y, x1 = b / g.Y % g.X, y1 = b % g.Y;
        if (la == lb) {
            const int d = layer_dir(la, g.DIR);
            if ((d == 0 && y0 != y1) || (d == 1 && x0 != x1)) { atomicOr(bad, 1); continue; }
            const int lo = d ? min(y0, y1) : min(x0, x1), hi = d ? max(y0, y1) : max(x0, x1);
            const int len = d ? g.Y : g.X;
            const int64_t c0 = d ? gidx(g, la, x0, lo) : gidx(g, la, lo, y0);
            const int64_t c1 = d ? gidx(g, la, x0, hi) : gidx(g, la, hi, y0);
            if (lo < hi) {
                atomicAdd(&g.rpre[c0], 1);
                atomicAdd(&g.rpre[c1], -1);
            }
            atomicAdd(&g.rcov[c0], 1);
            if (hi + 1 < len) atomicAdd(&g.rcov[c1 + 1], -1);
            g.dirty[track_id(g, la, x0, y0)] = 1;
        } else {
            const int lo = min(la, lb), hi = max(la, lb);
            for (int l = lo; l < hi; l++) {
                atomicAdd(&g.rvc[gidx(g, l, x0, y0)], 1);
                g.dirty[track_id(g, l, x0, y0)] = 1;
            }
        }
    }
}

__global__ void k_fill_i(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + SB - 1) / SB); }

} // namespace

BatchedState::BatchedState(igr_ctx* c_) : c(c_) {
    const int64_t n = c->cells();
    const int T = c->g.tbase[c->g.L];
    c->vc.zero(n, c->stream);
    c->rpre.zero(n, c->stream);
    c->rcov.zero(n, c->stream);
    c->rvc.zero(n, c->stream);
    c->bad.zero(1, c->stream);
    c->dirty.alloc(T);
    k_fill_i<<<blocks(T), SB, 0, c->stream>>>(c->dirty.p, T, 1);
    HIP_OK(hipGetLastError());
    GridDev& g = c->g;
    g.vc = c->vc.p;
    g.rpre = c->rpre.p;
    g.rcov = c->rcov.p;
    g.rvc = c->rvc.p;
    g.dirty = c->dirty.p;
}

BatchedState::~BatchedState() {
    GridDev& g = c->g;
    g.vc = g.rpre = g.rcov = g.rvc = g.dirty = nullptr;
}

void BatchedState::check() {
    int bad = 0;
    HIP_OK(hipMemcpyAsync(&bad, c->bad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (bad) throw std::runtime_error("a ripped route has a wire across its layer's direction");
}

void launch_track_fused(igr_ctx* c, int s_commit, int s_rip, bool cost) {
    KScope ks(c, K_TRACK_FUSED);
    const int T = c->g.tbase[c->g.L];
    const int ml = std::max(c->g.X, c->g.Y);
    const int cs = cost ? 1 : 0;
    unsigned long long* st = c->track_stats.p;
    hipStream_t s = c->stream;
    if (ml <= 256) k_track_fused<1, 256><<<T, 256, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    else if (ml <= 512) k_track_fused<2, 256><<<T, 256, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    else if (ml <= 1024) k_track_fused<4, 256><<<T, 256, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    else if (ml <= 2048) k_track_fused<2, 1024><<<T, 1024, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    else if (ml <= 4096) k_track_fused<4, 1024><<<T, 1024, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    else k_track_fused<8, 1024><<<T, 1024, 0, s>>>(c->g, s_commit, s_rip, cs, st);
    HIP_OK(hipGetLastError());
}

void launch_rip_scatter(igr_ctx* c, const int* d_nets, int64_t n) {
    if (n == 0) return;
    KScope ks(c, K_RIP);
    k_rip_scatter<<<blocks(n), SB, 0, c->stream>>>(c->g, d_nets, n, c->nets.rptr.p, c->bad.p);
    HIP_OK(hipGetLastError());
}

} // namespace igr_dev
