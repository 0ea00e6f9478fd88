// Workgroup Sklansky scan reproducing the reference's association order
// (src/graph.hpp:412-417): over a virtual array s[0..n), level d adds to every
// element whose index has bit d set the last element of the preceding 2^d
// block.  Thread t of the workgroup holds elements t*K .. t*K+K-1 in
// registers (K a power of two), so levels below log2(K) run inside a thread,
// the next six levels across the lanes of a wave (64-wide shuffles) and the
// rest across waves through LDS.  Each element receives the same additions in
// the same level order as the sequential loop, so floating-point results are
// bit-identical to it.  Elements at or beyond n never influence those below n.
#pragma once
#include <hip/hip_runtime.h>

namespace igr_dev {

template <typename T, int K, int NW>
__device__ __forceinline__ void block_sklansky(T (&s)[K], T* lds_tot /* NW entries */) {
    // Levels inside the thread.
#pragma unroll
    for (int d = 0; (1 << d) < K; d++) {
#pragma unroll
        for (int j = 0; j < K; j++)
            if (j >> d & 1) s[j] += s[(j >> d << d) - 1];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Levels across the lanes of a wave.
#pragma unroll
    for (int D = 0; D < 6; D++) {
        int partner = ((lane >> D) << D) - 1;
        T v = __shfl(s[K - 1], partner < 0 ? 0 : partner, 64);
        if (lane >> D & 1) {
#pragma unroll
            for (int j = 0; j < K; j++) s[j] += v;
        }
    }
    // Levels across waves.
    if (NW > 1) {
        if (lane == 63) lds_tot[wave] = s[K - 1];
        __syncthreads();
        T W[NW];
#pragma unroll
        for (int i = 0; i < NW; i++) W[i] = lds_tot[i];
#pragma unroll
        for (int b = 0; (1 << b) < NW; b++) {
            T add = W[(((wave >> b) << b) - 1) < 0 ? 0 : (((wave >> b) << b) - 1)];
            if (wave >> b & 1) {
#pragma unroll
                for (int j = 0; j < K; j++) s[j] += add;
            }
            T Wn[NW];
#pragma unroll
            for (int i = 0; i < NW; i++) Wn[i] = (i >> b & 1) ? W[i] + W[((i >> b) << b) - 1] : W[i];
#pragma unroll
            for (int i = 0; i < NW; i++) W[i] = Wn[i];
        }
        __syncthreads();
    }
}

} // namespace igr_dev

namespace igr_dev {

// The same scan for a workgroup of nw waves known only at run time
// (nw <= NWMAX, blockDim.x == 64 * nw).  Waves past nw hold nothing, so the
// association order over the virtual array is the one above.
template <typename T, int K, int NWMAX>
__device__ __forceinline__ void block_sklansky_rt(T (&s)[K], T* lds_tot /* NWMAX entries */, int nw) {
#pragma unroll
    for (int d = 0; (1 << d) < K; d++) {
#pragma unroll
        for (int j = 0; j < K; j++)
            if (j >> d & 1) s[j] += s[(j >> d << d) - 1];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int D = 0; D < 6; D++) {
        int partner = ((lane >> D) << D) - 1;
        T v = __shfl(s[K - 1], partner < 0 ? 0 : partner, 64);
        if (lane >> D & 1) {
#pragma unroll
            for (int j = 0; j < K; j++) s[j] += v;
        }
    }
    if (nw > 1) {
        if (lane == 63) lds_tot[wave] = s[K - 1];
        __syncthreads();
        // Wave totals after each cross-wave level, kept in registers: W[i]
        // for i < nw.  Level b adds the last total of the preceding 2^b block.
        T W[NWMAX];
#pragma unroll
        for (int i = 0; i < NWMAX; i++) W[i] = i < nw ? lds_tot[i] : T(0);
#pragma unroll
        for (int b = 0; (1 << b) < NWMAX; b++) {
            if ((1 << b) >= nw) continue;   // uniform: no further level
            const int src = ((wave >> b) << b) - 1;
            T add = T(0);
#pragma unroll
            for (int i = 0; i < NWMAX; i++)
                if (i == src) add = W[i];
            if (wave >> b & 1) {
#pragma unroll
                for (int j = 0; j < K; j++) s[j] += add;
            }
            T Wn[NWMAX];
#pragma unroll
            for (int i = 0; i < NWMAX; i++) Wn[i] = (i >> b & 1) ? W[i] + W[((i >> b) << b) - 1] : W[i];
#pragma unroll
            for (int i = 0; i < NWMAX; i++) W[i] = Wn[i];
        }
        __syncthreads();
    }
}

} // namespace igr_dev
