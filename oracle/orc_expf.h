/* Oracle copy of the single-precision exponential shared by the routing path.
 *
 * TEST INFRASTRUCTURE ONLY: part of the CPU oracle; never linked into the
 * product.
 *
 * The reference computes the logistic congestion cost with CUDA's __expf
 * (src/graph.hpp:24-30), whose values are defined by the CUDA toolkit and
 * cannot be reproduced here.  The routing path instead fixes its own exp,
 * which the GPU kernels and this oracle both implement with the same
 * operations so their results are bit-identical:
 *   Cody-Waite argument reduction (Cody & Waite 1980): n = rint(x*log2(e)),
 *   r = x - n*ln2_hi - n*ln2_lo with ln2_hi = 0x1.62e4p-1 (low bits zero so
 *   n*ln2_hi is exact) and ln2_lo = ln2 - ln2_hi; exp(r) by the degree-7
 *   Taylor polynomial in Horner form with fused multiply-adds (truncation
 *   error |r|^8/8! < 6e-9 on |r| <= ln2/2); the result scaled by 2^n through
 *   two exact power-of-two multiplies.
 * Inputs above 88.72 give +inf, inputs below -87 give 0 (no subnormals are
 * produced on either side), NaN propagates.
 */
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline float orc_pow2i(int k) { /* 2^k for k in [-126, 127] */
    uint32_t b = (uint32_t)(k + 127) << 23;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static inline float orc_expf(float x) {
    if (x != x) return x;
    if (x > 88.72f) return INFINITY;
    if (x < -87.0f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018704e-06f;
    float n = rintf(x * log2e);
    float r = fmaf(-n, ln2_hi, x);
    r = fmaf(-n, ln2_lo, r);
    float p = 1.98412698412698413e-04f;      /* 1/5040 */
    p = fmaf(p, r, 1.38888888888888889e-03f); /* 1/720 */
    p = fmaf(p, r, 8.33333333333333333e-03f); /* 1/120 */
    p = fmaf(p, r, 4.16666666666666667e-02f); /* 1/24 */
    p = fmaf(p, r, 1.66666666666666667e-01f); /* 1/6 */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)n;
    int k1 = k / 2, k2 = k - k1;
    return (p * orc_pow2i(k1)) * orc_pow2i(k2);
}
