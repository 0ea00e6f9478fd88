// Device-side definitions shared by the routing kernels: grid layout, the
// routing exp, the logistic cost and small helpers.
//
// Grid layout in HBM (structure of arrays, one array per quantity):
// routing layer l occupies [l*X*Y, (l+1)*X*Y).  Inside a layer the cells are
// stored track-major: a vertical layer (dir 1, tracks along y) keeps
// x*Y + y, a horizontal layer (dir 0, tracks along x) keeps y*X + x, so every
// track (row of a horizontal layer, column of a vertical layer) is contiguous
// and the per-track scans and commits read and write whole cache lines.
// Routes and every host-visible array keep the reference encoding
// l*X*Y + x*Y + y (src/database_cuda.hpp:49).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igr_dev {

constexpr double INF = 1e22;      // src/database_cuda.hpp:53
constexpr int MAXL = 16;          // largest supported routing-layer count

struct GridDev {
    int L, X, Y, DIR;
    int64_t XYc;                  // X*Y
    double unit_wire, unit_via;
    const double* w;              // per layer overflow weight (device)
    const int* xlen;
    const int* ylen;
    const float* cap;
    float* demand;
    float* wcost;
    float* vcost;
    double* presum;
    int* pre;                     // wire demand differences
    int* ts;                      // commit timestamps
    // Batched form only (null in the plain form): per-cell via-cut counts of
    // the routed batch, rip-up differences of the next batch, and a dirty flag
    // per track (tracks ordered by layer, then x or y; tbase[l] = first track
    // of layer l).
    int* vc;
    int* rpre;
    int* rcov;
    int* rvc;
    int* dirty;
    // igr_expf at the multiples of 1/4 from EXPTAB_LO/4 to EXPTAB_HI/4 (the
    // cost kernels copy it to LDS; see exp_tab)
    const float* exptab;
    int tbase[MAXL + 1];
};

constexpr int EXPTAB_LO = -348, EXPTAB_HI = 355, EXPTAB_N = EXPTAB_HI - EXPTAB_LO + 1;

__host__ __device__ __forceinline__ int layer_dir(int l, int DIR) { return (l & 1) ^ DIR; }

// Device offset of cell (l, x, y).
__host__ __device__ __forceinline__ int64_t gidx(const GridDev& g, int l, int x, int y) {
    return (int64_t)l * g.XYc + (layer_dir(l, g.DIR) ? (int64_t)x * g.Y + y : (int64_t)y * g.X + x);
}

// Track holding cell (l, x, y).
__device__ __forceinline__ int track_id(const GridDev& g, int l, int x, int y) {
    return g.tbase[l] + (layer_dir(l, g.DIR) ? x : y);
}

// Batched form: record the via cuts of a via between layers l0 and l1 at
// (x, y) for the fused track update (each cut l in [min, max) adds demand on
// layer l, src/graph.hpp:511-541).
__device__ __forceinline__ void emit_via(const GridDev& g, int l0, int l1, int x, int y) {
    const int lo = l0 < l1 ? l0 : l1, hi = l0 < l1 ? l1 : l0;
    for (int l = lo; l < hi; l++) {
        atomicAdd(&g.vc[gidx(g, l, x, y)], 1);
        g.dirty[track_id(g, l, x, y)] = 1;
    }
}

// ---------------------------------------------------------------------------
// Routing exp: Cody-Waite reduction and a degree-7 Taylor polynomial with fused
// multiply-adds; +inf above 88.72, 0 below -87, NaN propagates.  The CPU
// oracle evaluates the identical operation sequence, so both sides agree bit
// for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow2i(int k) { return __int_as_float((k + 127) << 23); }

// (Branch-free: the polynomial runs on x clamped to [-87, 88.72], which
// leaves every in-range x unchanged; the special cases are selected after.)
__device__ __forceinline__ float igr_expf(float x0) {
    const float x = __builtin_fminf(__builtin_fmaxf(x0, -87.0f), 88.72f);
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018704e-06f;
    float n = __builtin_rintf(x * log2e);
    float r = __builtin_fmaf(-n, ln2_hi, x);
    r = __builtin_fmaf(-n, ln2_lo, r);
    float p = 1.98412698412698413e-04f;
    p = __builtin_fmaf(p, r, 1.38888888888888889e-03f);
    p = __builtin_fmaf(p, r, 8.33333333333333333e-03f);
    p = __builtin_fmaf(p, r, 4.16666666666666667e-02f);
    p = __builtin_fmaf(p, r, 1.66666666666666667e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    int k = (int)n;
    int k1 = k / 2, k2 = k - k1;
    float v = (p * pow2i(k1)) * pow2i(k2);
    v = x0 > 88.72f ? __int_as_float(0x7f800000) : v;
    v = x0 < -87.0f ? 0.0f : v;
    return x0 != x0 ? x0 : v;
}

// igr_expf(a) for a multiple of 1/4 from the table: igr_expf is +inf above
// 88.72 (so from 355/4 on) and 0 below -87 (so below -348/4), and the table
// holds igr_expf of exactly the floats in between.
// (Branch-free: a clamped index, then selects.)
__device__ __forceinline__ float exp_tab(float a, const float* tab) {
    const float q = a * 4.0f;   // exact: scaling by a power of two
    const int i = (int)__builtin_fminf(__builtin_fmaxf(q, (float)EXPTAB_LO), (float)EXPTAB_HI) - EXPTAB_LO;
    float v = tab[i];
    v = q > (float)EXPTAB_HI ? __int_as_float(0x7f800000) : v;
    return q < (float)EXPTAB_LO ? 0.0f : v;
}

// True when the table may serve every exponent of an edge's cost terms: an
// integer capacity and a half-integer demand make 0.5 (d + inc - c) and
// 1.5 (d + inc) multiples of 1/4 for inc in {0, 0.5, 1}.
__device__ __forceinline__ bool quarter_grid(float c, float d) {
    const float d2 = 2.0f * d;
    return c == __builtin_rintf(c) && d2 == __builtin_rintf(d2) && __builtin_fabsf(c) < 4.0e6f &&
           __builtin_fabsf(d) < 4.0e6f;
}

// The logistic term of src/graph.hpp:23-32 as a float, before the widening
// to double: (double)c > 0.001 is c >= 0.001f (the float nearest 0.001 lies
// above it), and the products 0.5 (d - c), 1.5 d and the sum d + inc are
// exact in double, so rounding them once to float gives the float operation's
// value.  TAB: exponent from the table (the caller has checked quarter_grid).
// Branch-free: both exponent arguments are formed, one exp is taken.
template <bool TAB>
__device__ __forceinline__ float of_term(float c, float d, const float* tab) {
    const bool with_cap = c >= 0.001f;
    const float a = with_cap ? 0.5f * (d - c) : 1.5f * d;
    const float e = TAB ? exp_tab(a, tab) : igr_expf(a);
    return (with_cap || d > 0.0f) ? e : 0.0f;
}

// Cost increments of one edge for one track (a1) and half a track (ah),
// sharing the f(d) term: inc_cost(c, d, 1.0, w) and inc_cost(c, d, 0.5, w)
// of below, bit for bit.
template <bool TAB>
__device__ __forceinline__ void inc_pair(float c, float d, double w, const float* tab, double& a1, double& ah) {
    const double f0 = (double)of_term<TAB>(c, d, tab);
    a1 = fmin(1e12, ((double)of_term<TAB>(c, d + 1.0f, tab) - f0) * w);
    ah = fmin(1e12, ((double)of_term<TAB>(c, d + 0.5f, tab) - f0) * w);
}
template <bool TAB>
__device__ __forceinline__ double inc_one(float c, float d, float inc, double w, const float* tab) {
    return fmin(1e12, ((double)of_term<TAB>(c, d + inc, tab) - (double)of_term<TAB>(c, d, tab)) * w);
}

// Logistic overflow cost of one edge (src/graph.hpp:23-32 with
// of_cost_scale = 1): e^(0.5 (d - c)) for c > 0.001, e^(1.5 d) for d > 0.
__device__ __forceinline__ double of_cost(float c, float d) {
    if ((double)c > 0.001) {
        double a = 0.5 * (double)(d - c);
        return (double)igr_expf((float)a);
    }
    if (d > 0) {
        double a = 1.5 * (double)d;
        return (double)igr_expf((float)a);
    }
    return 0;
}

// Cost of adding 'inc' tracks of demand to an edge (src/graph.hpp:34-38):
// min(1e12, (f(d + inc) - f(d)) * w) with the min absorbing inf - inf.
__device__ __forceinline__ double inc_cost(float c, float d, double inc, double w) {
    float d1 = (float)((double)d + inc + 0.0);
    double v = (of_cost(c, d1) - of_cost(c, d)) * w;
    return fmin(1e12, v);
}

} // namespace igr_dev
