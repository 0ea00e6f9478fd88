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
template <typename T>
__device__ __forceinline__ T next_first(T first, T* sh) {
    __syncthreads();
    sh[threadIdx.x] = first;
    __syncthreads();
    return threadIdx.x + 1 < blockDim.x ? sh[threadIdx.x + 1] : T(0);
}
// Value of the last element of the previous thread (0 for thread 0).
template <typename T>
__device__ __forceinline__ T prev_last(T last, T* sh) {
    __syncthreads();
    sh[threadIdx.x] = last;
    __syncthreads();
    return threadIdx.x > 0 ? sh[threadIdx.x - 1] : T(0);
}

// K consecutive elements of a track per thread.  With K a multiple of 4 and
// 16-byte aligned tracks (v4), four elements move in one 16-byte access, so a
// wave reads or writes one contiguous span.
template <int K, typename T, typename V4>
__device__ __forceinline__ void load_k(const T* __restrict__ src, int64_t base, int p0, int len, bool v4,
                                       T (&dst)[K]) {
    if constexpr (K % 4 == 0) {
        if (v4) {
#pragma unroll
            for (int j = 0; j < K; j += 4) {
                if (p0 + j < len) {
                    const V4 q = *reinterpret_cast<const V4*>(src + base + p0 + j);
                    dst[j] = q.x; dst[j + 1] = q.y; dst[j + 2] = q.z; dst[j + 3] = q.w;
                } else {
                    dst[j] = dst[j + 1] = dst[j + 2] = dst[j + 3] = T(0);
                }
            }
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < K; j++) dst[j] = p0 + j < len ? src[base + p0 + j] : T(0);
}

template <int K>
__device__ __forceinline__ void store_k(float* __restrict__ dst, int64_t base, int p0, int len, bool v4,
                                        const float (&v)[K]) {
    if constexpr (K % 4 == 0) {
        if (v4) {
#pragma unroll
            for (int j = 0; j < K; j += 4)
                if (p0 + j < len)
                    *reinterpret_cast<float4*>(dst + base + p0 + j) = make_float4(v[j], v[j + 1], v[j + 2], v[j + 3]);
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (p0 + j < len) dst[base + p0 + j] = v[j];
}

template <int K>
__device__ __forceinline__ void store_k(double* __restrict__ dst, int64_t base, int p0, int len, bool v4,
                                        const double (&v)[K]) {
    if constexpr (K % 4 == 0) {
        if (v4) {
#pragma unroll
            for (int j = 0; j < K; j += 2)
                if (p0 + j < len) *reinterpret_cast<double2*>(dst + base + p0 + j) = make_double2(v[j], v[j + 1]);
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (p0 + j < len) dst[base + p0 + j] = v[j];
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

// Wire cost, via cost and the exclusive prefix input (the predecessor's wire
// cost) of a thread's K cells from their demand and capacity, with pdc the
// predecessor cell's (demand, capacity).
template <int K, bool TAB>
__device__ __forceinline__ void track_costs(const GridDev& g, int l, int dir, int len, int p0, const float (&dem)[K],
                                            const float (&cap)[K], float2 pdc, const float* tab, float (&wc)[K],
                                            float (&vc)[K], double (&s)[K]) {
    const double w = g.w[l];
    const int* elen = dir ? g.ylen : g.xlen;
    // edge lengths of the cells p0 - 1 .. p0 + K - 1 (0 from the last cell on)
    int el[K + 1];
    el[0] = p0 >= 1 && p0 - 1 < len - 1 ? elen[p0 - 1] : 0;
    if (K % 4 == 0 && p0 + K <= len - 1) {
#pragma unroll
        for (int j = 0; j < K; j += 4) {
            const int4 q = *reinterpret_cast<const int4*>(elen + p0 + j);
            el[j + 1] = q.x; el[j + 2] = q.y; el[j + 3] = q.z; el[j + 4] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < K; j++) el[j + 1] = p0 + j < len - 1 ? elen[p0 + j] : 0;
    }
    auto wire = [&](double a1, int p, int e) {
        const float x = (float)a1;
        const float y = (float)((double)x + g.unit_wire * (double)e);
        return p < len - 1 ? y : x;
    };
    double a1p = 0.0, ahp = 0.0;
    float wcp = 0.0f;
    if (p0 >= 1 && p0 - 1 < len) {
        inc_pair<TAB>(pdc.y, pdc.x, w, tab, a1p, ahp);
        wcp = wire(a1p, p0 - 1, el[0]);
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = p0 + j;
        if (p >= len) { wc[j] = vc[j] = 0.0f; s[j] = 0.0; continue; }
        double a1, ah;
        inc_pair<TAB>(cap[j], dem[j], w, tab, a1, ah);
        wc[j] = wire(a1, p, el[j + 1]);
        float v;
        if (p == 0) v = (float)a1;
        else if (p == len - 1) v = (float)a1p;
        else v = (float)(ah + ahp);
        vc[j] = (float)((double)v + g.unit_via);
        s[j] = p >= 1 ? (double)wcp : 0.0;   // exclusive prefix input
        a1p = a1;
        ahp = ah;
        wcp = wc[j];
    }
}

// s_c: commit stamp (0 = nothing to commit); s_r: rip-up stamp (0 = no
// rip-up); cost: recompute costs and prefix sums (and clear the dirty flag).
// stats[0] += cells of the tracks processed, stats[1] += those with rip-up.
// The workgroup is the fewest waves (at most 16) whose threads cover a track
// at K cells each.
template <int K>
__global__ void __launch_bounds__(1024) k_track_fused(GridDev g, int s_c, int s_r, int cost,
                                                      unsigned long long* stats) {
    constexpr int NWMAX = 16;
    const int nw = blockDim.x >> 6;
    __shared__ int sh_i[NWMAX * 64];
    __shared__ float2 sh_f2[NWMAX * 64];
    __shared__ int lds_i[NWMAX];
    __shared__ double lds_d[NWMAX];
    __shared__ float s_exp[EXPTAB_N];
    int t = blockIdx.x;
    if (!g.dirty[t]) return;
    int l = 0;
    while (l < g.L - 1 && t >= g.tbase[l + 1]) l++;
    t -= g.tbase[l];
    const int dir = layer_dir(l, g.DIR);
    const int len = dir ? g.Y : g.X;
    const int64_t base = (int64_t)l * g.XYc + (int64_t)t * len;
    const bool v4 = (len & 3) == 0;
    const int p0 = threadIdx.x * K;
    float dem[K], cap[K];
    int ts[K];
    uint32_t wr_dem = 0, wr_ts = 0, clr = 0;   // per-element bit masks
    load_k<K, float, float4>(g.demand, base, p0, len, v4, dem);
    load_k<K, float, float4>(g.cap, base, p0, len, v4, cap);
    if (s_c || s_r) load_k<K, int, int4>(g.ts, base, p0, len, v4, ts);
    else
#pragma unroll
        for (int j = 0; j < K; j++) ts[j] = 0;
    if (s_c) {
        int inc[K], gv[K];
        load_k<K, int, int4>(g.pre, base, p0, len, v4, inc);
        load_k<K, int, int4>(g.vc, base, p0, len, v4, gv);
#pragma unroll
        for (int j = 0; j < K; j++)
            if (inc[j] | gv[j]) clr |= 1u << j;
        block_sklansky_rt<int, K, NWMAX>(inc, lds_i, nw);
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
        const int gn = next_first<int>(gv[0], sh_i);
        if (gn) wr_dem |= 1u << (K - 1);
        apply_vias<K>(dem, gv, gn, len, 1.0f);
    }
    if (s_r) {
        int dec[K], cov[K], gr[K];
        load_k<K, int, int4>(g.rpre, base, p0, len, v4, dec);
        load_k<K, int, int4>(g.rcov, base, p0, len, v4, cov);
        load_k<K, int, int4>(g.rvc, base, p0, len, v4, gr);
#pragma unroll
        for (int j = 0; j < K; j++)
            if (dec[j] | cov[j] | gr[j]) clr |= 1u << j;
        block_sklansky_rt<int, K, NWMAX>(dec, lds_i, nw);
        block_sklansky_rt<int, K, NWMAX>(cov, lds_i, nw);
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
        const int gn = next_first<int>(gr[0], sh_i);
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
    // Per element the increments of one and of half a track (inc_pair, the
    // shared f(d) term); the via cost takes the element's and its
    // predecessor's, the prefix input is the predecessor's wire cost.  The
    // predecessor of a thread's first element is recomputed from the previous
    // thread's last demand and capacity (one exchange, whose barriers also
    // publish the LDS exp table).  The exponents come from the table when
    // every element of the wave allows it, else they are computed; the
    // values are the same either way.
    for (int i = threadIdx.x; i < EXPTAB_N; i += blockDim.x) s_exp[i] = g.exptab[i];
    const float2 pdc = prev_last(make_float2(dem[K - 1], cap[K - 1]), sh_f2);
    float wc[K], vc[K];
    double s[K];
    bool q = quarter_grid(pdc.y, pdc.x);
#pragma unroll
    for (int j = 0; j < K; j++) q = q && quarter_grid(cap[j], dem[j]);
    if (__all(q)) track_costs<K, true>(g, l, dir, len, p0, dem, cap, pdc, s_exp, wc, vc, s);
    else track_costs<K, false>(g, l, dir, len, p0, dem, cap, pdc, s_exp, wc, vc, s);
    store_k<K>(g.wcost, base, p0, len, v4, wc);
    store_k<K>(g.vcost, base, p0, len, v4, vc);
    // Exclusive FP64 Sklansky prefix of wcost (src/graph.hpp:405-430).
    block_sklansky_rt<double, K, NWMAX>(s, lds_d, nw);
    store_k<K>(g.presum, base, p0, len, v4, s);
    if (threadIdx.x == 0) g.dirty[blockIdx.x] = 0;
}

// Cost update and prefix sums alone (no commit, no rip-up: the full sweep at
// a pass start and the timed sweep), as a persistent kernel: a grid sized to
// the workgroups the GPU holds walks the tracks, each workgroup loading its
// next dirty track's demand and capacity before it computes the current one,
// so loads stay in flight while the costs, the scan and the stores run.  The
// exp table is copied to LDS once per workgroup.  The values are those of
// k_track_fused with s_c = s_r = 0.
template <int K>
__global__ void __launch_bounds__(1024) k_track_cost(GridDev g, unsigned long long* stats) {
    constexpr int NWMAX = 16;
    const int nw = blockDim.x >> 6;
    __shared__ float2 sh_f2[NWMAX * 64];
    __shared__ double lds_d[NWMAX];
    __shared__ float s_exp[EXPTAB_N];
    for (int i = threadIdx.x; i < EXPTAB_N; i += blockDim.x) s_exp[i] = g.exptab[i];
    const int T = g.tbase[g.L];
    const int p0 = threadIdx.x * K;
    auto next_dirty = [&](int t) {
        while (t < T && !g.dirty[t]) t += gridDim.x;
        return t;
    };
    auto geometry = [&](int t, int& l, int& len, int64_t& base) {
        l = 0;
        while (l < g.L - 1 && t >= g.tbase[l + 1]) l++;
        len = layer_dir(l, g.DIR) ? g.Y : g.X;
        base = (int64_t)l * g.XYc + (int64_t)(t - g.tbase[l]) * len;
    };
    int t = next_dirty((int)blockIdx.x);
    float dem[K], cap[K];
    int l = 0, len = 0;
    int64_t base = 0;
    if (t < T) {
        geometry(t, l, len, base);
        load_k<K, float, float4>(g.demand, base, p0, len, (len & 3) == 0, dem);
        load_k<K, float, float4>(g.cap, base, p0, len, (len & 3) == 0, cap);
    }
    while (t < T) {
        // the next track's loads first
        const int tn = next_dirty(t + (int)gridDim.x);
        int ln = 0, lenn = 0;
        int64_t basen = 0;
        float demn[K], capn[K];
        if (tn < T) {
            geometry(tn, ln, lenn, basen);
            load_k<K, float, float4>(g.demand, basen, p0, lenn, (lenn & 3) == 0, demn);
            load_k<K, float, float4>(g.cap, basen, p0, lenn, (lenn & 3) == 0, capn);
        }
        const bool v4 = (len & 3) == 0;
        const float2 pdc = prev_last(make_float2(dem[K - 1], cap[K - 1]), sh_f2);
        float wc[K], vc[K];
        double sv[K];
        bool q = quarter_grid(pdc.y, pdc.x);
#pragma unroll
        for (int j = 0; j < K; j++) q = q && quarter_grid(cap[j], dem[j]);
        if (__all(q)) track_costs<K, true>(g, l, layer_dir(l, g.DIR), len, p0, dem, cap, pdc, s_exp, wc, vc, sv);
        else track_costs<K, false>(g, l, layer_dir(l, g.DIR), len, p0, dem, cap, pdc, s_exp, wc, vc, sv);
        store_k<K>(g.wcost, base, p0, len, v4, wc);
        store_k<K>(g.vcost, base, p0, len, v4, vc);
        block_sklansky_rt<double, K, NWMAX>(sv, lds_d, nw);
        store_k<K>(g.presum, base, p0, len, v4, sv);
        if (threadIdx.x == 0) {
            g.dirty[t] = 0;
            atomicAdd(&stats[0], (unsigned long long)len);
        }
        t = tn;
        l = ln;
        len = lenn;
        base = basen;
#pragma unroll
        for (int j = 0; j < K; j++) {
            dem[j] = demn[j];
            cap[j] = capn[j];
        }
    }
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

// Full-grid sweeps of the cost update and prefix sums, timed with events on
// the context stream: the plain form's kernels (update_wcost, update_vcost,
// presum) or the fused track kernel with every track dirty and nothing to
// commit.  Average milliseconds per sweep over reps sweeps after one warm-up.
double time_cost_sweep(igr_ctx* c, int form, int reps) {
    hipStream_t s = c->stream;
    const int T = c->g.tbase[c->g.L];
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    double ms_total = 0;
    try {
        if (form == IGR_FORM_PLAIN) {
            for (int r = 0; r <= reps; r++) {
                HIP_OK(hipEventRecord(e0, s));
                launch_update_cost(c);
                launch_presum(c);
                HIP_OK(hipEventRecord(e1, s));
                HIP_OK(hipEventSynchronize(e1));
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) ms_total += ms;
            }
        } else {
            BatchedState bs(c);
            for (int r = 0; r <= reps; r++) {
                k_fill_i<<<blocks(T), SB, 0, s>>>(c->dirty.p, T, 1);
                HIP_OK(hipEventRecord(e0, s));
                launch_track_fused(c, 0, 0, true);
                HIP_OK(hipEventRecord(e1, s));
                HIP_OK(hipEventSynchronize(e1));
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) ms_total += ms;
            }
            bs.check();
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    HIP_OK(hipEventDestroy(e0));
    HIP_OK(hipEventDestroy(e1));
    return ms_total / std::max(1, reps);
}

void launch_track_fused(igr_ctx* c, int s_commit, int s_rip, bool cost) {
    KScope ks(c, K_TRACK_FUSED);
    const int T = c->g.tbase[c->g.L];
    const int ml = std::max(c->g.X, c->g.Y);
    const int cs = cost ? 1 : 0;
    unsigned long long* st = c->track_stats.p;
    hipStream_t s = c->stream;
    // Elements per thread: tuned, or 4 up to 4096-cell tracks and 8 above
    // (4 measured fastest on 800-, 1500- and 2500-cell tracks); the workgroup
    // is the fewest waves covering the longest track.
    const int K = c->tune.track_k ? c->tune.track_k : (ml <= 4096 ? 4 : 8);
    const int B = ((ml + K - 1) / K + 63) / 64 * 64;
    if (B > 1024) throw std::invalid_argument("track longer than the fused kernel's tile");
    if (!s_commit && !s_rip && cost) {
        // cost update and prefix sums only: persistent, prefetching kernel
        int cus = 0;
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        auto go = [&](auto kern) {
            int per = 0;
            HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, B, 0));
            const int grid = std::max(1, std::min(T, std::max(1, per) * cus));
            kern<<<grid, B, 0, s>>>(c->g, st);
        };
        switch (K) {
        case 4: go(k_track_cost<4>); break;
        case 8: go(k_track_cost<8>); break;
        case 16: go(k_track_cost<16>); break;
        default: throw std::invalid_argument("track_k must be 4, 8 or 16");
        }
        HIP_OK(hipGetLastError());
        return;
    }
    switch (K) {
    case 4: k_track_fused<4><<<T, B, 0, s>>>(c->g, s_commit, s_rip, cs, st); break;
    case 8: k_track_fused<8><<<T, B, 0, s>>>(c->g, s_commit, s_rip, cs, st); break;
    case 16: k_track_fused<16><<<T, B, 0, s>>>(c->g, s_commit, s_rip, cs, st); break;
    default: throw std::invalid_argument("track_k must be 4, 8 or 16");
    }
    HIP_OK(hipGetLastError());
}

void launch_rip_scatter(igr_ctx* c, const int* d_nets, int64_t n, hipStream_t s) {
    if (n == 0) return;
    if (s == nullptr || s == c->stream) {
        KScope ks(c, K_RIP);
        k_rip_scatter<<<blocks(n), SB, 0, c->stream>>>(c->g, d_nets, n, c->nets.rptr.p, c->bad.p);
    } else {
        k_rip_scatter<<<blocks(n), SB, 0, s>>>(c->g, d_nets, n, c->nets.rptr.p, c->bad.p);
    }
    HIP_OK(hipGetLastError());
}

} // namespace igr_dev
