// Grid kernels of the plain (reference launch structure) form: logistic cost
// update, FP64 per-track prefix sums, wire and via demand commits, rip-up
// marking, congestion view, and layout conversion for host transfers.
// All are HBM-bandwidth bound streaming kernels over the track-major grid.
#include "igr_ctx.hpp"
#include "scan.h"

namespace igr_dev {

namespace {

constexpr int TPB = 256;

// Decode a device cell offset into (layer, track, position, track length).
struct CellPos {
    int l, track, pos, len;
};
__device__ __forceinline__ CellPos decode(const GridDev& g, int64_t i) {
    CellPos c;
    c.l = (int)(i / g.XYc);
    int64_t r = i - (int64_t)c.l * g.XYc;
    c.len = layer_dir(c.l, g.DIR) ? g.Y : g.X;
    c.track = (int)(r / c.len);
    c.pos = (int)(r - (int64_t)c.track * c.len);
    return c;
}

// wcost = f32(inc_cost(+1)) then f32(wcost + unit_wire * edge length) for all
// but the last cell of a track (src/graph.hpp:40-47).
__global__ void __launch_bounds__(TPB) k_update_wcost(GridDev g) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)g.L * g.XYc) return;
    CellPos c = decode(g, i);
    float wv = (float)inc_cost(g.cap[i], g.demand[i], 1.0, g.w[c.l]);
    if (c.pos < c.len - 1) {
        int el = layer_dir(c.l, g.DIR) ? g.ylen[c.pos] : g.xlen[c.pos];
        wv = (float)((double)wv + g.unit_wire * (double)el);
    }
    g.wcost[i] = wv;
}

// Via cost: half the cost increment of the two track edges meeting at the
// cell, the whole increment at a track end, plus unit_via
// (src/graph.hpp:52-71; cross_points is identically zero).
__global__ void __launch_bounds__(TPB) k_update_vcost(GridDev g) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)g.L * g.XYc) return;
    CellPos c = decode(g, i);
    double wl = g.w[c.l];
    float vv;
    if (c.pos == 0) vv = (float)inc_cost(g.cap[i], g.demand[i], 1.0, wl);
    else if (c.pos == c.len - 1) vv = (float)inc_cost(g.cap[i - 1], g.demand[i - 1], 1.0, wl);
    else vv = (float)(inc_cost(g.cap[i], g.demand[i], 0.5, wl) + inc_cost(g.cap[i - 1], g.demand[i - 1], 0.5, wl));
    g.vcost[i] = (float)((double)vv + g.unit_via);
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
// (src/graph.hpp:405-430).  One workgroup per track.
template <int K>
__global__ void __launch_bounds__(TPB) k_presum(GridDev g) {
    __shared__ double lds[TPB / 64];
    int64_t base;
    int len;
    track_of(g, blockIdx.x, base, len);
    double s[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        s[j] = (p >= 1 && p < len) ? (double)g.wcost[base + p - 1] : 0.0;
    }
    block_sklansky<double, K, TPB / 64>(s, lds);
#pragma unroll
    for (int j = 0; j < K; j++) {
        int p = threadIdx.x * K + j;
        if (p < len) g.presum[base + p] = s[j];
    }
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
// (src/graph.hpp:118-132).
__global__ void __launch_bounds__(TPB) k_commit_edges(GridDev g, int stamp) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)g.L * g.XYc) return;
    CellPos c = decode(g, i);
    if (c.pos + 1 < c.len) {
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
// (src/graph.hpp:432-438).
__global__ void __launch_bounds__(TPB) k_mark_edges(GridDev g, int th, int* of) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)g.L * g.XYc) return;
    CellPos c = decode(g, i);
    of[i] = (c.pos + 1 < c.len && g.cap[i] + (float)th <= g.demand[i]) ? 1 : 0;
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

void launch_update_cost(igr_ctx* c) {
    int64_t n = c->cells();
    KScope ks(c, K_COST, 16.0 * n);   // read cap, demand; write wcost, vcost
    k_update_wcost<<<blocks(n), TPB, 0, c->stream>>>(c->g);
    HIP_OK(hipGetLastError());
    k_update_vcost<<<blocks(n), TPB, 0, c->stream>>>(c->g);
    HIP_OK(hipGetLastError());
}

void launch_presum(igr_ctx* c) {
    KScope ks(c, K_PRESUM, 12.0 * c->cells());   // read wcost f32, write presum f64
    DISPATCH_K(std::max(c->g.X, c->g.Y), k_presum, tracks(c->g), c->stream, c->g);
}

void launch_wire_commit(igr_ctx* c, int stamp) {
    // scan: read/write the difference array; commit: read it, read/write
    // demand, write timestamps where covered (not counted); clear it
    KScope ks(c, K_WIRE_COMMIT, 16.0 * c->cells());
    DISPATCH_K(std::max(c->g.X, c->g.Y), k_scan_int, tracks(c->g), c->stream, c->g, c->g.pre);
    int64_t n = c->cells();
    k_commit_edges<<<blocks(n), TPB, 0, c->stream>>>(c->g, stamp);
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
    int64_t n = c->cells();
    k_mark_edges<<<blocks(n), TPB, 0, c->stream>>>(c->g, th, d_of);
    HIP_OK(hipGetLastError());
    DISPATCH_K(std::max(c->g.X, c->g.Y), k_scan_int, tracks(c->g), c->stream, c->g, d_of);
    k_mark_nets<<<blocks(c->nets.n), TPB, 0, c->stream>>>(c->g, c->nets.n, c->nets.rptr.p, d_of, d_is_of);
    HIP_OK(hipGetLastError());
}

void launch_congestion(igr_ctx* c, float* d_xsum, float* d_ysum) {
    const GridDev& g = c->g;
    DevBuf<float> v;
    v.alloc(g.XYc);
    k_congestion<<<blocks(g.XYc), TPB, 0, c->stream>>>(g, v.p);
    HIP_OK(hipGetLastError());
    // along x for every y: elements x*Y + y, lines y
    DISPATCH_K(g.X, k_view_scan, g.Y, c->stream, v.p, d_xsum, g.X, (int64_t)g.Y, (int64_t)1);
    // along y for every x: elements x*Y + y, lines x
    DISPATCH_K(g.Y, k_view_scan, g.X, c->stream, v.p, d_ysum, g.Y, (int64_t)1, (int64_t)g.Y);
    HIP_OK(hipStreamSynchronize(c->stream));
}

void launch_expf(const float* in, float* out, int64_t n, hipStream_t s) {
    k_expf<<<blocks(n), TPB, 0, s>>>(in, out, n);
    HIP_OK(hipGetLastError());
}

namespace {
// Plain 16-byte-vector copy: the bandwidth reference and the calibration
// kernel for the memory counters (bytes read = bytes written = n * 16).
__global__ void __launch_bounds__(TPB) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) dst[i] = src[i];
}
} // namespace

double copy_probe(int64_t bytes, int reps, hipStream_t s) {
    const int64_t n = bytes / 16;
    if (n <= 0 || reps <= 0) throw std::invalid_argument("copy probe needs bytes >= 16 and reps >= 1");
    DevBuf<float4> a, b;
    a.zero(n, s);
    b.alloc(n);
    const unsigned grid = (unsigned)std::min<int64_t>((n + TPB - 1) / TPB, 256 * 64);
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
