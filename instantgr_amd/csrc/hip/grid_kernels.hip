// Grid kernels of the plain (reference launch structure) form: logistic cost
// update, FP64 per-track prefix sums, wire and via demand commits, rip-up
// marking, congestion view, and layout conversion for host transfers.
// All are HBM-bandwidth bound streaming kernels over the track-major grid.
#include "igr_ctx.hpp"
#include "scan.h"

namespace igr_dev {

namespace {

constexpr int TPB = 256;

// Track of a two-dimensional launch (blockIdx.y = track in layer order):
// its layer, length and first cell.  blockIdx.y is uniform, so this is scalar
// work; no per-cell division.
__device__ __forceinline__ void track_at(const GridDev& g, int t, int& l, int& len, int64_t& base) {
    l = 0;
    while (l < g.L - 1 && t >= g.tbase[l + 1]) l++;
    len = layer_dir(l, g.DIR) ? g.Y : g.X;
    base = (int64_t)l * g.XYc + (int64_t)(t - g.tbase[l]) * len;
}

// N consecutive cells of a track per thread (N a multiple of 4), with 16-byte
// accesses when the track length is a multiple of 4, which keeps every track
// 16-byte aligned.  All of a thread's loads are issued before any use, so a
// wave keeps N * 256 bytes per array in flight.
template <int N, typename T, typename V4>
__device__ __forceinline__ void ldn(const T* __restrict__ a, int64_t base, int p0, int len, T (&v)[N]) {
    if ((len & 3) == 0) {
#pragma unroll
        for (int j = 0; j < N; j += 4) {
            if (p0 + j < len) {
                const V4 q = *reinterpret_cast<const V4*>(a + base + p0 + j);
                v[j] = q.x; v[j + 1] = q.y; v[j + 2] = q.z; v[j + 3] = q.w;
            } else {
                v[j] = v[j + 1] = v[j + 2] = v[j + 3] = T(0);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = p0 + j < len ? a[base + p0 + j] : T(0);
    }
}
template <int N>
__device__ __forceinline__ void stn(float* __restrict__ a, int64_t base, int p0, int len, const float (&v)[N]) {
    if ((len & 3) == 0) {
#pragma unroll
        for (int j = 0; j < N; j += 4)
            if (p0 + j < len) *reinterpret_cast<float4*>(a + base + p0 + j) = make_float4(v[j], v[j + 1], v[j + 2], v[j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (p0 + j < len) a[base + p0 + j] = v[j];
    }
}

constexpr int CPT = 16;   // cells per thread of the cost kernels
constexpr int CTPB = 64;  // their workgroup: one wave, 1024 cells of a track

// Copies the exp table into LDS (all threads of the block take part; the
// caller synchronises before use).
__device__ __forceinline__ void load_exptab(const GridDev& g, float* s_exp) {
    for (int i = threadIdx.x; i < EXPTAB_N; i += blockDim.x) s_exp[i] = g.exptab[i];
}

// Gcell lengths el[p0 .. p0 + CPT) of a track's edges (0 past its last edge,
// which has no length: the track has len - 1 edges).
__device__ __forceinline__ void load_elen(const int* __restrict__ el, int p0, int len, int (&e)[CPT]) {
    if (p0 + CPT <= len - 1) {
#pragma unroll
        for (int j = 0; j < CPT; j += 4) {
            const int4 q = *reinterpret_cast<const int4*>(el + p0 + j);
            e[j] = q.x; e[j + 1] = q.y; e[j + 2] = q.z; e[j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < CPT; j++) e[j] = p0 + j < len - 1 ? el[p0 + j] : 0;
    }
}

template <bool TAB>
__device__ __forceinline__ void wcost_cells(const GridDev& g, const float (&c)[CPT], const float (&d)[CPT], int p0,
                                            int len, double w, const int (&e)[CPT], const float* tab,
                                            float (&o)[CPT]) {
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const float a1 = (float)inc_one<TAB>(c[j], d[j], 1.0f, w, tab);
        const float wl = (float)((double)a1 + g.unit_wire * (double)e[j]);
        o[j] = p0 + j < len - 1 ? wl : a1;
    }
}

// wcost = f32(inc_cost(+1)) then f32(wcost + unit_wire * edge length) for all
// but the last cell of a track (src/graph.hpp:40-47).  The exponents come
// from the LDS table when every cell of the wave allows it (quarter_grid),
// else they are computed; the values are the same either way.
__global__ void __launch_bounds__(CTPB) k_update_wcost(GridDev g) {
    __shared__ float s_exp[EXPTAB_N];
    int l, len;
    int64_t base;
    track_at(g, blockIdx.y, l, len, base);
    const int p0 = (blockIdx.x * CTPB + threadIdx.x) * CPT;
    float c[CPT], d[CPT], o[CPT];
    int e[CPT];
    if (p0 < len) {
        ldn<CPT, float, float4>(g.cap, base, p0, len, c);
        ldn<CPT, float, float4>(g.demand, base, p0, len, d);
        load_elen(layer_dir(l, g.DIR) ? g.ylen : g.xlen, p0, len, e);
    }
    load_exptab(g, s_exp);
    __syncthreads();
    if (p0 >= len) return;
    const double w = g.w[l];
    bool q = true;
#pragma unroll
    for (int j = 0; j < CPT; j++) q = q && quarter_grid(c[j], d[j]);
    if (__all(q)) wcost_cells<true>(g, c, d, p0, len, w, e, s_exp, o);
    else wcost_cells<false>(g, c, d, p0, len, w, e, s_exp, o);
    stn<CPT>(g.wcost, base, p0, len, o);
}

// Interior cells take the half increments of the cell and the one before;
// the track's first cell the whole increment of its edge, the last cell the
// whole increment of the edge before it (only the threads holding a track
// end take that branch).
template <bool TAB>
__device__ __forceinline__ void vcost_cells(const GridDev& g, const float (&c)[CPT], const float (&d)[CPT], float cp,
                                            float dp, int p0, int len, double w, const float* tab, float (&o)[CPT]) {
    double hp = inc_one<TAB>(cp, dp, 0.5f, w, tab);   // half increment of the cell before
    float vv[CPT];
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const double h = inc_one<TAB>(c[j], d[j], 0.5f, w, tab);
        vv[j] = (float)(h + hp);
        hp = h;
    }
    if (p0 == 0 || p0 + CPT > len - 1) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int p = p0 + j;
            if (p == 0) vv[j] = (float)inc_one<TAB>(c[j], d[j], 1.0f, w, tab);
            else if (p == len - 1) vv[j] = (float)inc_one<TAB>(j ? c[j - 1] : cp, j ? d[j - 1] : dp, 1.0f, w, tab);
        }
    }
#pragma unroll
    for (int j = 0; j < CPT; j++) o[j] = (float)((double)vv[j] + g.unit_via);
}

// Via cost: half the cost increment of the two track edges meeting at the
// cell, the whole increment at a track end, plus unit_via
// (src/graph.hpp:52-71; cross_points is identically zero).  A thread's cells
// and the one before them give the half increments it needs.
__global__ void __launch_bounds__(CTPB) k_update_vcost(GridDev g) {
    __shared__ float s_exp[EXPTAB_N];
    int l, len;
    int64_t base;
    track_at(g, blockIdx.y, l, len, base);
    const int p0 = (blockIdx.x * CTPB + threadIdx.x) * CPT;
    float c[CPT], d[CPT], o[CPT];
    float cp = 0.0f, dp = 0.0f;
    if (p0 < len) {
        if (p0 > 0) {
            cp = g.cap[base + p0 - 1];
            dp = g.demand[base + p0 - 1];
        }
        ldn<CPT, float, float4>(g.cap, base, p0, len, c);
        ldn<CPT, float, float4>(g.demand, base, p0, len, d);
    }
    load_exptab(g, s_exp);
    __syncthreads();
    if (p0 >= len) return;
    const double w = g.w[l];
    bool q = quarter_grid(cp, dp);
#pragma unroll
    for (int j = 0; j < CPT; j++) q = q && quarter_grid(c[j], d[j]);
    if (__all(q)) vcost_cells<true>(g, c, d, cp, dp, p0, len, w, s_exp, o);
    else vcost_cells<false>(g, c, d, cp, dp, p0, len, w, s_exp, o);
    stn<CPT>(g.vcost, base, p0, len, o);
}

// Track index -> (layer, first cell, length).
__device__ __forceinline__ void track_of(const GridDev& g, int t, int64_t& base, int& len) {
    int l = 0;
    for (;; l++) {
        int cnt = layer_dir(l, g.DIR) ? g.X : g.Y;
        if (t < cnt) break;
        t -= cnt;
    }
    len = layer_dir(l, g.DIR) ? g.Y : g.X;
    base = (int64_t)l * g.XYc + (int64_t)t * len;
}

// presum[p] = Sklansky-exclusive prefix of wcost along the track, FP64
// (src/graph.hpp:405-430).  One workgroup per track (blockIdx.x = track) of
// just enough waves for its cells, K cells per thread: 16-byte loads of
// wcost[p0 .. p0+3], the element before p0 from the previous lane.
template <int K>
__global__ void __launch_bounds__(1024) k_presum(GridDev g) {
    __shared__ double lds[16];
    int l, len;
    int64_t base;
    track_at(g, blockIdx.x, l, len, base);
    const int nw = blockDim.x >> 6;
    const int p0 = threadIdx.x * K;
    float w[K];
    if constexpr (K % 4 == 0) {
        if ((len & 3) == 0) {
#pragma unroll
            for (int j = 0; j < K; j += 4) {
                if (p0 + j < len) {
                    const float4 q = *reinterpret_cast<const float4*>(g.wcost + base + p0 + j);
                    w[j] = q.x; w[j + 1] = q.y; w[j + 2] = q.z; w[j + 3] = q.w;
                } else {
                    w[j] = w[j + 1] = w[j + 2] = w[j + 3] = 0.0f;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < K; j++) w[j] = p0 + j < len ? g.wcost[base + p0 + j] : 0.0f;
        }
    } else {
#pragma unroll
        for (int j = 0; j < K; j++) w[j] = p0 + j < len ? g.wcost[base + p0 + j] : 0.0f;
    }
    // element p0 - 1: the previous lane's last, or a load at a wave's first lane
    float before = __shfl_up(w[K - 1], 1, 64);
    if ((threadIdx.x & 63) == 0) before = (p0 >= 1 && p0 - 1 < len) ? g.wcost[base + p0 - 1] : 0.0f;
    double s[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = p0 + j;
        s[j] = (p >= 1 && p < len) ? (double)(j ? w[j - 1] : before) : 0.0;
    }
    block_sklansky_rt<double, K, 16>(s, lds, nw);
    if constexpr (K % 2 == 0) {
        if ((len & 1) == 0) {
#pragma unroll
            for (int j = 0; j < K; j += 2)
                if (p0 + j < len) *reinterpret_cast<double2*>(g.presum + base + p0 + j) = make_double2(s[j], s[j + 1]);
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (p0 + j < len) g.presum[base + p0 + j] = s[j];
}

// In-place exclusive int prefix of the difference array along each track
// (src/graph.hpp:74-99).
template <int K>
__global__ void __launch_bounds__(TPB) k_scan_int(GridDev g, int* a) {
    __shared__ int lds[TPB / 64];
    int64_t base;
    int len;
    track_of(g, blockIdx.x, base, len);
    int s[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        s[j] = (p >= 1 && p < len) ? a[base + p - 1] : 0;
    }
    block_sklansky<int, K, TPB / 64>(s, lds);
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        if (p < len) a[base + p] = s[j];
    }
}

// demand += covered count, timestamp = stamp where the count is positive
// (src/graph.hpp:118-132).  Two-dimensional launch: blockIdx.y = track.
__global__ void __launch_bounds__(TPB) k_commit_edges(GridDev g, int stamp) {
    int l, len;
    int64_t base;
    track_at(g, blockIdx.y, l, len, base);
    const int p = blockIdx.x * TPB + threadIdx.x;
    if (p + 1 < len) {
        const int64_t i = base + p;
        int v = g.pre[i + 1];
        if (v > 0) {
            g.demand[i] += (float)v;
            g.ts[i] = stamp;
        }
    }
}

// Via demand of routed nets, gated by the commit timestamps
// (src/graph.hpp:496-543).
__global__ void __launch_bounds__(TPB) k_via_commit(GridDev g, const int* nets, int64_t n, int* const* rptr, int stamp,
                                                    int K) {
    int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    const int* r = rptr[nets[t]];
    const int XY = g.X * g.Y;
    const float k = (float)K, h = 0.5f * (float)K;
    for (int i = 1; i < r[0]; i += 2) {
        int a = r[i], b = r[i + 1];
        int la = a / XY, lb = b / XY;
        if (la == lb) continue;
        int x = a / g.Y % g.X, y = a % g.Y;
        int minl = min(la, lb), maxl = max(la, lb);
        for (int l = minl; l < maxl; l++) {
            if (g.ts[gidx(g, l, x, y)] >= stamp) continue;
            if (layer_dir(l, g.DIR) == 1) {
                if (y == 0) atomicAdd(&g.demand[gidx(g, l, x, y)], k);
                else if (y == g.Y - 1) atomicAdd(&g.demand[gidx(g, l, x, y - 1)], k);
                else { atomicAdd(&g.demand[gidx(g, l, x, y - 1)], h); atomicAdd(&g.demand[gidx(g, l, x, y)], h); }
            } else {
                if (x == 0) atomicAdd(&g.demand[gidx(g, l, x, y)], k);
                else if (x == g.X - 1) atomicAdd(&g.demand[gidx(g, l, x - 1, y)], k);
                else { atomicAdd(&g.demand[gidx(g, l, x - 1, y)], h); atomicAdd(&g.demand[gidx(g, l, x, y)], h); }
            }
        }
    }
}

// of = 1 on track edges whose demand reaches capacity + threshold
// (src/graph.hpp:432-438).  Two-dimensional launch: blockIdx.y = track.
__global__ void __launch_bounds__(TPB) k_mark_edges(GridDev g, int th, int* of) {
    int l, len;
    int64_t base;
    track_at(g, blockIdx.y, l, len, base);
    const int p = blockIdx.x * TPB + threadIdx.x;
    if (p >= len) return;
    const int64_t i = base + p;
    of[i] = (p + 1 < len && g.cap[i] + (float)th <= g.demand[i]) ? 1 : 0;
}

// A net is ripped when more than 10 full edges lie under its wires
// (src/graph.hpp:441-466).
__global__ void __launch_bounds__(TPB) k_mark_nets(GridDev g, int64_t n, int* const* rptr, const int* of,
                                                   uint8_t* is_of) {
    int64_t net = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (net >= n) return;
    const int* r = rptr[net];
    const int XY = g.X * g.Y;
    int cnt = 0;
    uint8_t res = 0;
    for (int i = 1; i < r[0]; i += 2) {
        int l = r[i] / XY, x0 = r[i] / g.Y % g.X, y0 = r[i] % g.Y;
        int x1 = r[i + 1] / g.Y % g.X, y1 = r[i + 1] % g.Y;
        if (x0 != x1) cnt += of[gidx(g, l, max(x0, x1), y0)] - of[gidx(g, l, min(x0, x1), y0)];
        if (y0 != y1) cnt += of[gidx(g, l, x0, max(y0, y1))] - of[gidx(g, l, x0, min(y0, y1))];
        if (cnt > 10) { res = 1; break; }
    }
    is_of[net] = res;
}

// 2-D congestion view: max over layers of (demand - capacity) where positive
// (src/graph.hpp:545-556, with a zeroed start and an exact max).
__global__ void __launch_bounds__(TPB) k_congestion(GridDev g, float* v) {
    int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= g.XYc) return;
    int x = (int)(q / g.Y), y = (int)(q % g.Y);
    float m = 0.0f;
    for (int l = 0; l < g.L; l++) {
        int64_t i = gidx(g, l, x, y);
        double resource = (double)(g.cap[i] - g.demand[i]);
        if (resource < 0) m = (float)fmax((double)m, -resource);
    }
    v[q] = m;
}

// FP32 Sklansky exclusive prefix of v along x (per y, stride Y) or along y
// (per x, stride 1) (src/graph.hpp:558-584).
template <int K>
__global__ void __launch_bounds__(TPB) k_view_scan(const float* v, float* out, int len, int64_t stride_pos,
                                                   int64_t stride_line) {
    __shared__ float lds[TPB / 64];
    const int64_t base = (int64_t)blockIdx.x * stride_line;
    float s[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        s[j] = (p >= 1 && p < len) ? v[base + (int64_t)(p - 1) * stride_pos] : 0.0f;
    }
    block_sklansky<float, K, TPB / 64>(s, lds);
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        if (p < len) out[base + (int64_t)p * stride_pos] = s[j];
    }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_to_ref(GridDev g, const T* src, T* dst) {
    int64_t r = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (r >= (int64_t)g.L * g.XYc) return;
    int l = (int)(r / g.XYc);
    int64_t q = r - (int64_t)l * g.XYc;
    int x = (int)(q / g.Y), y = (int)(q % g.Y);
    dst[r] = src[gidx(g, l, x, y)];
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_from_ref(GridDev g, const T* src, T* dst) {
    int64_t r = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (r >= (int64_t)g.L * g.XYc) return;
    int l = (int)(r / g.XYc);
    int64_t q = r - (int64_t)l * g.XYc;
    int x = (int)(q / g.Y), y = (int)(q % g.Y);
    dst[gidx(g, l, x, y)] = src[r];
}

__global__ void k_exptab(float* t) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i < EXPTAB_N) t[i] = igr_expf((float)(i + EXPTAB_LO) * 0.25f);
}

__global__ void k_expf(const float* in, float* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = igr_expf(in[i]);
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

int tracks(const GridDev& g) {
    int t = 0;
    for (int l = 0; l < g.L; l++) t += layer_dir(l, g.DIR) ? g.X : g.Y;
    return t;
}

} // namespace

#define DISPATCH_K(maxlen, KERNEL, GRID, STREAM, ...)                                                   \
    do {                                                                                               \
        int ml_ = (maxlen);                                                                            \
        if (ml_ <= TPB) KERNEL<1><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                              \
        else if (ml_ <= 2 * TPB) KERNEL<2><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                     \
        else if (ml_ <= 4 * TPB) KERNEL<4><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                     \
        else if (ml_ <= 8 * TPB) KERNEL<8><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                     \
        else if (ml_ <= 16 * TPB) KERNEL<16><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                   \
        else if (ml_ <= 32 * TPB) KERNEL<32><<<GRID, TPB, 0, STREAM>>>(__VA_ARGS__);                   \
        else throw std::runtime_error("track longer than 8192 gcells is not supported");               \
        HIP_OK(hipGetLastError());                                                                     \
    } while (0)

// Grid of a two-dimensional launch over every track, 'per' cells per thread.
dim3 track_grid(const GridDev& g, int tpb, int per) {
    const int ml = std::max(g.X, g.Y);
    if (g.tbase[g.L] > 65535) throw std::runtime_error("more than 65535 tracks");
    return dim3((unsigned)((ml + tpb * per - 1) / (tpb * per)), (unsigned)g.tbase[g.L]);
}

void launch_update_cost(igr_ctx* c) {
    int64_t n = c->cells();
    KScope ks(c, K_COST, 16.0 * n);   // read cap, demand; write wcost, vcost
    k_update_wcost<<<track_grid(c->g, CTPB, CPT), CTPB, 0, c->stream>>>(c->g);
    HIP_OK(hipGetLastError());
    k_update_vcost<<<track_grid(c->g, CTPB, CPT), CTPB, 0, c->stream>>>(c->g);
    HIP_OK(hipGetLastError());
}

void launch_presum(igr_ctx* c) {
    KScope ks(c, K_PRESUM, 12.0 * c->cells());   // read wcost f32, write presum f64
    // 16 cells per thread (tunable); the workgroup is the fewest waves
    // covering the longest track.
    const int ml = std::max(c->g.X, c->g.Y);
    const int K = c->tune.presum_k ? c->tune.presum_k : 16;
    const int B = ((ml + K - 1) / K + 63) / 64 * 64;
    if (B > 1024) throw std::runtime_error("track longer than the prefix-sum kernel's tile");
    switch (K) {
    case 4: k_presum<4><<<tracks(c->g), B, 0, c->stream>>>(c->g); break;
    case 8: k_presum<8><<<tracks(c->g), B, 0, c->stream>>>(c->g); break;
    case 16: k_presum<16><<<tracks(c->g), B, 0, c->stream>>>(c->g); break;
    default: throw std::invalid_argument("presum_k must be 4, 8 or 16");
    }
    HIP_OK(hipGetLastError());
}

void launch_wire_commit(igr_ctx* c, int stamp) {
    // scan: read/write the difference array; commit: read it, read/write
    // demand, write timestamps where covered (not counted); clear it
    KScope ks(c, K_WIRE_COMMIT, 16.0 * c->cells());
    DISPATCH_K(std::max(c->g.X, c->g.Y), k_scan_int, tracks(c->g), c->stream, c->g, c->g.pre);
    int64_t n = c->cells();
    k_commit_edges<<<track_grid(c->g, TPB, 1), TPB, 0, c->stream>>>(c->g, stamp);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemsetAsync(c->g.pre, 0, sizeof(int) * n, c->stream));
}

void launch_via_commit(igr_ctx* c, const int* d_nets, int64_t n, int stamp, int K) {
    if (n == 0) return;
    KScope ks(c, K_VIA_COMMIT);
    k_via_commit<<<blocks(n), TPB, 0, c->stream>>>(c->g, d_nets, n, c->nets.rptr.p, stamp, K);
    HIP_OK(hipGetLastError());
}

// Rip-up of the wires of routed nets: demand -= 1 on every wire edge and the
// timestamp set on every wire cell (src/graph.hpp:468-495 with K = -1).
__global__ void __launch_bounds__(TPB) k_rip_wire(GridDev g, const int* nets, int64_t n, int* const* rptr, int stamp) {
    int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    const int* r = rptr[nets[t]];
    const int XY = g.X * g.Y;
    for (int i = 1; i < r[0]; i += 2) {
        int a = r[i], b = r[i + 1];
        if (a / XY != b / XY) continue;
        int l = a / XY;
        int x0 = a / g.Y % g.X, y0 = a % g.Y, x1 = b / g.Y % g.X, y1 = b % g.Y;
        if (x0 < x1) for (int x = x0; x < x1; x++) { atomicAdd(&g.demand[gidx(g, l, x, y0)], -1.0f); g.ts[gidx(g, l, x, y0)] = stamp; }
        if (x1 < x0) for (int x = x1; x < x0; x++) { atomicAdd(&g.demand[gidx(g, l, x, y0)], -1.0f); g.ts[gidx(g, l, x, y0)] = stamp; }
        if (y0 < y1) for (int y = y0; y < y1; y++) { atomicAdd(&g.demand[gidx(g, l, x0, y)], -1.0f); g.ts[gidx(g, l, x0, y)] = stamp; }
        if (y1 < y0) for (int y = y1; y < y0; y++) { atomicAdd(&g.demand[gidx(g, l, x0, y)], -1.0f); g.ts[gidx(g, l, x0, y)] = stamp; }
        g.ts[gidx(g, l, x0, y0)] = stamp;
        g.ts[gidx(g, l, x0, y1)] = stamp;
        g.ts[gidx(g, l, x1, y0)] = stamp;
        g.ts[gidx(g, l, x1, y1)] = stamp;
    }
}

void launch_rip_wire(igr_ctx* c, const int* d_nets, int64_t n, int stamp) {
    if (n == 0) return;
    KScope ks(c, K_RIP);
    k_rip_wire<<<blocks(n), TPB, 0, c->stream>>>(c->g, d_nets, n, c->nets.rptr.p, stamp);
    HIP_OK(hipGetLastError());
}

void launch_to_ref(igr_ctx* c, const void* src, void* dst, int elem_bytes) {
    int64_t n = c->cells();
    if (elem_bytes == 4) k_to_ref<int><<<blocks(n), TPB, 0, c->stream>>>(c->g, (const int*)src, (int*)dst);
    else k_to_ref<double><<<blocks(n), TPB, 0, c->stream>>>(c->g, (const double*)src, (double*)dst);
    HIP_OK(hipGetLastError());
}

void launch_from_ref(igr_ctx* c, const void* src, void* dst, int elem_bytes) {
    int64_t n = c->cells();
    if (elem_bytes == 4) k_from_ref<int><<<blocks(n), TPB, 0, c->stream>>>(c->g, (const int*)src, (int*)dst);
    else k_from_ref<double><<<blocks(n), TPB, 0, c->stream>>>(c->g, (const double*)src, (double*)dst);
    HIP_OK(hipGetLastError());
}

void launch_ripup_mark(igr_ctx* c, int th, uint8_t* d_is_of, int* d_of) {
    k_mark_edges<<<track_grid(c->g, TPB, 1), TPB, 0, c->stream>>>(c->g, th, d_of);
    HIP_OK(hipGetLastError());
    DISPATCH_K(std::max(c->g.X, c->g.Y), k_scan_int, tracks(c->g), c->stream, c->g, d_of);
    k_mark_nets<<<blocks(c->nets.n), TPB, 0, c->stream>>>(c->g, c->nets.n, c->nets.rptr.p, d_of, d_is_of);
    HIP_OK(hipGetLastError());
}

void launch_congestion_map(igr_ctx* c, float* d_v) {
    const GridDev& g = c->g;
    k_congestion<<<blocks(g.XYc), TPB, 0, c->stream>>>(g, d_v);
    HIP_OK(hipGetLastError());
}

void launch_view_scan(igr_ctx* c, const float* d_v, float* d_out, int axis) {
    const GridDev& g = c->g;
    if (axis == 0)   // along x for every y: elements x*Y + y, lines y
        DISPATCH_K(g.X, k_view_scan, g.Y, c->stream, d_v, d_out, g.X, (int64_t)g.Y, (int64_t)1);
    else             // along y for every x: elements x*Y + y, lines x
        DISPATCH_K(g.Y, k_view_scan, g.X, c->stream, d_v, d_out, g.Y, (int64_t)1, (int64_t)g.Y);
}

void launch_congestion(igr_ctx* c, float* d_xsum, float* d_ysum) {
    c->scratch.view.alloc(c->g.XYc);
    launch_congestion_map(c, c->scratch.view.p);
    launch_view_scan(c, c->scratch.view.p, d_xsum, 0);
    launch_view_scan(c, c->scratch.view.p, d_ysum, 1);
    HIP_OK(hipStreamSynchronize(c->stream));
}

void launch_exptab(float* t, hipStream_t s) {
    k_exptab<<<blocks(EXPTAB_N), TPB, 0, s>>>(t);
    HIP_OK(hipGetLastError());
}

void launch_expf(const float* in, float* out, int64_t n, hipStream_t s) {
    k_expf<<<blocks(n), TPB, 0, s>>>(in, out, n);
    HIP_OK(hipGetLastError());
}

namespace {
// Plain 16-byte-vector copy: the bandwidth reference and the calibration
// kernel for the memory counters (bytes read = bytes written = n * 16).  Each
// thread moves four 16-byte vectors, a wave's loads covering 4 KiB, all four
// loads issued before the stores.
__global__ void __launch_bounds__(TPB) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    int64_t n) {
    const int64_t i0 = (int64_t)blockIdx.x * TPB * 4 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t i = i0 + (int64_t)k * TPB;
        if (i < n) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t i = i0 + (int64_t)k * TPB;
        if (i < n) dst[i] = v[k];
    }
}
} // namespace

double copy_probe(int64_t bytes, int reps, hipStream_t s) {
    const int64_t n = bytes / 16;
    if (n <= 0 || reps <= 0) throw std::invalid_argument("copy probe needs bytes >= 16 and reps >= 1");
    DevBuf<float4> a, b;
    a.zero(n, s);
    b.alloc(n);
    const unsigned grid = (unsigned)((n + 4 * TPB - 1) / (4 * TPB));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    k_copy_probe<<<grid, TPB, 0, s>>>(a.p, b.p, n);   // warm-up
    HIP_OK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) k_copy_probe<<<grid, TPB, 0, s>>>(a.p, b.p, n);
    HIP_OK(hipEventRecord(e1, s));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_OK(hipGetLastError());
    return 2.0 * 16.0 * (double)n * reps / (ms * 1e-3) / 1e9;
}

} // namespace igr_dev
