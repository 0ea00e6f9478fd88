// Batched form of the conflict-free net batch scheduler: the same batches as
// the round-by-round scheduler (src/database_cuda.hpp:288-434,437-631, with
// the deterministic claim semantics stated in sched_kernels.hip), with less
// work and no host synchronisation per round.
//
//  * Round 0 of every batch starts from the map every unassigned net claims
//    with its point cells: per cell, the highest-priority unassigned net that
//    lists the cell.  It is kept across batches from a per-cell list of nets
//    sorted by priority (built once per pass) and a head index per cell that
//    moves past assigned nets at the end of each batch, so round 0 claims
//    nothing.
//  * Within a batch a cell's value only grows (claims are maxima) except where
//    the end of a round releases the cells of blocked nets, so a later point
//    claim can only change a released cell.
//  * Two ways to run the later rounds give the same batches at different
//    costs per round:
//    - active-list rounds (kernels k_a_*): the nets still active in the batch
//      form a compacted list rebuilt after each round that committed; they
//      claim their points where the map holds a lower value, and the revert
//      tests each of them against the owners of its points.  Work per round
//      follows the active nets' point counts.
//    - release-driven rounds (kernels k_r_*): the end of a round lists the
//      cells it released; each goes to its highest-priority active lister (the
//      first active entry of the cell's list, scanned by a lane group), and
//      only those new owners are tested for viability (a net can only turn
//      viable by taking a released cell).  A net is blocked when one of its
//      points is owned by a net committed in this batch; owners committed in
//      earlier rounds already blocked every active net they could (their
//      cells only shrink), so the revert lists the cells of the nets committed
//      in this round and blocks the active nets in those cells' lists.  Work
//      per round follows the released cells and committed nets.
//    On a 640k-cell grid with 0.5M nets the active-list rounds cost less; on a
//    2.25M-cell grid with 3M-10M nets the release-driven ones do.
//  * List orders are free: every decision is an atomic maximum or a test
//    against the map.
//  * The round and batch control lives in a device word block; the host
//    enqueues rounds in chunks and reads back one flag per chunk.  Kernels of
//    rounds after the end return at once.
#include "igr_ctx.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace igr_dev {

namespace {

constexpr int TPB = 256;
constexpr int ROUNDS = 30;        // rounds per batch (src/database_cuda.hpp:612)

// control state words (two slots of C_WORDS, then the placed-net counter).
// C_NU / C_UL: length and buffer of the nets round 0 tests.  Active-list
// rounds: the unassigned list, and C_NA / C_AL the same for the batch's
// active list (C_NA < 0: the unassigned list itself).  Release-driven rounds:
// the round-0 candidates (below); C_NV0, the nets round 0 found viable (list
// v0); C_NR, cells the previous round released (list rel); C_NK, new owners
// of those cells (list cand); C_NC, nets committed in this round (list comm);
// C_NH, cells those nets own (list hits).
// C_LASTC, nets placed when the lists were last compacted.
enum { C_DONE = 0, C_CUR, C_ROUND, C_COMMIT, C_HAVE, C_NVIABLE, C_NU, C_UL, C_NA, C_AL, C_NR, C_NK, C_NC, C_NH,
       C_NV0, C_LASTC, C_WORDS };
constexpr int NLISTS = 3;         // net list buffers of n entries each
constexpr int COMPACT_DIV = 64;   // release-driven rounds: compact the cell lists every n/64 placed nets

struct BArgs {
    int64_t n;
    const int* ids;               // scheduler index -> net id (segments)
    const int* pids;              // scheduler index -> row of the point arrays
    const int64_t* pt_off;
    const int* pt;
    const int64_t* hs_off;
    const int* hs;
    const int64_t* vs_off;
    const int* vs;
    int* map;                     // X*Y owner or -1
    int* base;                    // X*Y highest unassigned net listing the cell, or -1
    const int64_t* c_off;         // X*Y+1 per-cell list offsets
    int64_t* c_end;               // X*Y end of each list (below c_off[q+1] once compacted)
    int* c_net;                   // per-cell lists, highest index first
    int64_t* head;                // X*Y first unassigned entry of each list
    int* batch;
    int* ctrl;
    int* lists;                   // NLISTS buffers of n net indices
    // release-driven rounds only
    int* rel;                     // X*Y cells released by the previous round
    int* cand;                    // nets that took a released cell this round
    int* cstamp;                  // per net: 1 + the round that listed it in cand, -(1 + g) when the
                                  // end of round g listed it as a round-0 candidate
    int* v0;                      // nets viable in round 0 of the batch
    int* comm;                    // nets committed this round
    int* hits;                    // X*Y cells owned by the nets committed this round
    int* hstamp;                  // per cell: 1 + the round that listed it in hits
    int64_t* cur;                 // X*Y: no active net before this entry of the cell's list in this batch
    unsigned long long* stat;     // rounds run, nets tested for viability, unassigned at batch starts
    // per net two bits, 16 nets a word (a few MB even at 10M nets, so the
    // random lookups of the list scans stay in L2): bit 0 assigned, bit 1
    // blocked in this batch; a net is active while both are clear
    uint32_t* nst;
    int64_t nwords;
    int X, Y, sparsity;
    // optional per-round record (IGR_SCHED_TRACE): TR_WORDS per round
    long long* trace;
    int64_t trace_cap;
    // kernel timing only: bytes the release-driven round kernels read and
    // write, one slot per workgroup index modulo BYTE_SLOTS
    unsigned long long* bytes;
};
constexpr int BYTE_SLOTS = 4096;

// Adds a wave's byte count to its workgroup's slot (profiling only; every
// lane of the wave calls it together, at the kernel's end).
__device__ __forceinline__ void flush_bytes(const BArgs& a, unsigned long long nb) {
    if (!a.bytes) return;
    for (int off = 32; off > 0; off >>= 1) nb += __shfl_xor(nb, off);
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(&a.bytes[blockIdx.x % BYTE_SLOTS], nb);
}

// Per-round record: batch, round in the batch, nets tested, viable,
// committed, cells those own, cells released by the previous round, new
// owners, device wall clock at the round's last kernel.
enum { TR_CUR = 0, TR_ROUND, TR_TESTED, TR_VIABLE, TR_COMM, TR_HITS, TR_REL, TR_NEWOWN, TR_CLOCK, TR_WORDS };

__device__ __forceinline__ uint32_t nstate(const BArgs& a, int64_t i) {
    return (a.nst[i >> 4] >> (2 * (i & 15))) & 3u;
}
__device__ __forceinline__ bool assigned(const BArgs& a, int64_t i) { return nstate(a, i) & 1u; }
__device__ __forceinline__ bool blocked(const BArgs& a, int64_t i) { return nstate(a, i) & 2u; }
__device__ __forceinline__ void set_assigned(const BArgs& a, int64_t i) {
    atomicOr(&a.nst[i >> 4], 1u << (2 * (i & 15)));
}
__device__ __forceinline__ void set_blocked(const BArgs& a, int64_t i) {
    atomicOr(&a.nst[i >> 4], 2u << (2 * (i & 15)));
}
// batch end: every blocked net is active again (assigned bits kept)
constexpr uint32_t NST_ASSIGNED = 0x55555555u;

// Active-list rounds: the batch's active list of the round whose state is st.
__device__ __forceinline__ const int* a_list(const BArgs& a, const int* st, int64_t& na) {
    na = st[C_NA] < 0 ? st[C_NU] : st[C_NA];
    return a.lists + (int64_t)st[C_AL] * a.n;
}

// Release-driven rounds: the nets whose viability the round whose state is st
// tests: the round-0 candidates in round 0 (every net in the first batch), the
// new owners of released cells later.
__device__ __forceinline__ const int* r_list(const BArgs& a, const int* st, int64_t& na) {
    if (st[C_ROUND] == 0) {
        na = st[C_NU];
        return a.lists + (int64_t)st[C_UL] * a.n;
    }
    na = st[C_NK];
    return a.cand;
}

// not assigned and not blocked in this batch
__device__ __forceinline__ bool active(const BArgs& a, int64_t i) { return nstate(a, i) == 0; }

// ---- per-cell lists -------------------------------------------------------
__global__ void k_cell_count(BArgs a, int* cnt) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) atomicAdd(&cnt[a.pt[k]], 1);
}

__global__ void k_cell_fill(BArgs a, int* fillc) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
        const int p = a.pt[k];
        const int pos = atomicAdd(&fillc[p], 1);
        a.c_net[a.c_off[p] + pos] = (int)i;
    }
}

// Heap sort of one cell's list into decreasing order (min-heap), then the
// initial head and base.
__device__ void sift_min(int* v, int64_t root, int64_t n) {
    for (;;) {
        int64_t c = 2 * root + 1;
        if (c >= n) return;
        if (c + 1 < n && v[c + 1] < v[c]) c++;
        if (v[c] >= v[root]) return;
        const int t = v[c];
        v[c] = v[root];
        v[root] = t;
        root = c;
    }
}

__global__ void k_cell_sort(BArgs a, int64_t cells) {
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= cells) return;
    const int64_t b = a.c_off[q], m = a.c_off[q + 1] - b;
    int* v = a.c_net + b;
    for (int64_t r = m / 2 - 1; r >= 0; r--) sift_min(v, r, m);
    for (int64_t e = m - 1; e > 0; e--) {
        const int t = v[0];
        v[0] = v[e];
        v[e] = t;
        sift_min(v, 0, e);
    }
    a.head[q] = b;
    a.c_end[q] = b + m;
    if (a.cur) a.cur[q] = b;
    const int top = m > 0 ? v[0] : -1;
    a.base[q] = top;
    a.map[q] = top;
}

// ---- exclusive scan of the per-cell counts into int64 offsets ------------
constexpr int SCAN_PER = 16;
constexpr int SCAN_BLOCK = TPB * SCAN_PER;

__device__ int64_t block_excl_scan(int64_t v, int64_t* sh, int64_t* total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < TPB; d <<= 1) {
        const int64_t t = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int64_t incl = sh[threadIdx.x];
    *total = sh[TPB - 1];
    __syncthreads();
    return incl - v;
}

__global__ void k_scan_reduce(const int* cnt, int64_t n, int64_t* bsum) {
    __shared__ int64_t sh[TPB];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_PER;
    int64_t s = 0;
    for (int j = 0; j < SCAN_PER; j++)
        if (b0 + j < n) s += cnt[b0 + j];
    int64_t tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_bsums(int64_t* bsum, int64_t nb) {
    __shared__ int64_t sh[TPB];
    const int64_t per = (nb + TPB - 1) / TPB;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int64_t s = 0;
    for (int64_t j = 0; j < per; j++)
        if (b0 + j < nb) s += bsum[b0 + j];
    int64_t tot;
    int64_t run = block_excl_scan(s, sh, &tot);
    for (int64_t j = 0; j < per; j++)
        if (b0 + j < nb) {
            const int64_t v = bsum[b0 + j];
            bsum[b0 + j] = run;
            run += v;
        }
}

__global__ void k_scan_apply(const int* cnt, int64_t n, const int64_t* bsum, int64_t* off) {
    __shared__ int64_t sh[TPB];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_PER;
    int64_t s = 0;
    for (int j = 0; j < SCAN_PER; j++)
        if (b0 + j < n) s += cnt[b0 + j];
    int64_t tot;
    int64_t run = block_excl_scan(s, sh, &tot) + bsum[blockIdx.x];
    for (int j = 0; j < SCAN_PER; j++)
        if (b0 + j < n) {
            off[b0 + j] = run;
            run += cnt[b0 + j];
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == TPB - 1) off[n] = run;
}

// ---- rounds -----------------------------------------------------------------
// Round g reads the control state in slot g & 1; its last kernel writes the
// successor state into the other slot, so no kernel reads a word another
// kernel of the same round rewrites.  A net's point list is walked by a group
// of G lanes (one step covers G points), with wave ballots for the early
// exits; a viable net's neighbour and segment claims are spread over a whole
// wave.
__device__ __forceinline__ uint64_t group_bits(uint64_t m, int G) {
    const int base = (int)(threadIdx.x & 63) & ~(G - 1);
    const uint64_t mask = G == 64 ? ~0ull : ((1ull << G) - 1);
    return (m >> base) & mask;
}

// Appends v to list for every thread with want, one counter atomic per
// workgroup (a returning atomic per wave on one counter serialises thousands
// of waves).  Every thread of the workgroup calls it; s holds TPB/64+1 ints.
__device__ __forceinline__ void block_append(bool want, int v, int* counter, int* list, int* s) {
    constexpr int NW = TPB / 64;
    const uint64_t m = __ballot(want);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s[wv] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < NW; w++) {
            const int c = s[w];
            s[w] = tot;
            tot += c;
        }
        s[NW] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    if (want) list[s[NW] + s[wv] + __popcll(m & ((1ull << lane) - 1))] = v;
    __syncthreads();   // s is reused by the next call
}

// A workgroup's appends to one list, gathered in LDS over the whole kernel
// and written with one counter atomic at the end: counter atomics per wave or
// per pass serialise thousands of waves on the one address each round.  add()
// may be called by diverged lanes (the ballot covers the active lanes, and the
// lane that takes the offsets is one of them); entries beyond the LDS
// capacity go out at once.  Every thread of the workgroup calls init() and
// flush().
constexpr int GL_CAP = 2048;
struct GroupList {
    int* buf;                     // LDS, GL_CAP entries
    int* hdr;                     // LDS: count, then the list offset
    int* counter;
    int* list;
    __device__ __forceinline__ void init() {
        if (threadIdx.x == 0) hdr[0] = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void add(bool want, int v) {
        const uint64_t m = __ballot(want);
        if (m == 0) return;
        const int lane = threadIdx.x & 63;
        const int leader = __builtin_ctzll(m);
        int off = 0;
        if (lane == leader) off = atomicAdd(&hdr[0], __popcll(m));
        off = __shfl(off, leader);
        const int pos = off + __popcll(m & ((1ull << lane) - 1));
        if (want) {
            if (pos < GL_CAP) buf[pos] = v;
            else list[atomicAdd(counter, 1)] = v;
        }
    }
    __device__ __forceinline__ void flush() {
        __syncthreads();
        const int nb = hdr[0] < GL_CAP ? hdr[0] : GL_CAP;
        if (threadIdx.x == 0) hdr[1] = nb ? atomicAdd(counter, nb) : 0;
        __syncthreads();
        for (int i = threadIdx.x; i < nb; i += blockDim.x) list[hdr[1] + i] = buf[i];
    }
};

// Calls f(valid, q) for the sampled cells q of a net's segments (every
// sparsity-th cell from each segment's start, src/database_cuda.hpp:320-352),
// the cells of up to 64 segments at a time spread evenly over the wave's
// lanes (a lane per segment would walk a long segment alone).  Every lane of
// the wave calls it together; f sees valid = false on the lanes past the last
// cell and may use ballots.
template <class F>
__device__ __forceinline__ void wave_segment_cells(const BArgs& a, int net, F f) {
    const int lane = threadIdx.x & 63;
    const int Y = a.Y, sp = a.sparsity;
    for (int dir = 0; dir < 2; dir++) {
        const int64_t s0 = dir ? a.vs_off[net] : a.hs_off[net];
        const int64_t s1 = dir ? a.vs_off[net + 1] : a.hs_off[net + 1];
        const int* seg = dir ? a.vs : a.hs;
        for (int64_t sb = s0; sb < s1; sb += 64) {
            const int64_t k = sb + lane;
            int cnt = 0, first = 0, step = 0;
            if (k < s1) {
                const int b = seg[2 * k], e = seg[2 * k + 1];
                if (dir == 0) {          // along x at y = b % Y
                    const int x0 = b / Y, x1 = e / Y;
                    cnt = x1 >= x0 ? (x1 - x0) / sp + 1 : 0;
                    first = x0 * Y + b % Y;
                    step = sp * Y;
                } else {                 // along y at x = b / Y
                    const int y0 = b % Y, y1 = e % Y;
                    cnt = y1 >= y0 ? (y1 - y0) / sp + 1 : 0;
                    first = (b / Y) * Y + y0;
                    step = sp;
                }
            }
            int incl = cnt;              // inclusive scan of the cell counts
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(incl, d);
                if (lane >= d) incl += v;
            }
            const int total = __shfl(incl, 63);
            const int excl = incl - cnt;
            for (int cb = 0; cb < total; cb += 64) {
                const int c = cb + lane;
                int sg = 0;              // the first lane whose inclusive count exceeds c
                for (int b2 = 32; b2 > 0; b2 >>= 1)
                    if (__shfl(incl, sg + b2 - 1) <= c) sg += b2;
                const int f0 = __shfl(first, sg), st0 = __shfl(step, sg), ex = __shfl(excl, sg);
                const bool valid = c < total;
                f(valid, valid ? f0 + (c - ex) * st0 : 0);
            }
        }
    }
}

// 3. (both) a wave per viable net claims the neighbour cells (0,-1), (-1,0),
//    (1,0) of its points and every sparsity-th cell of its segments
__global__ void __launch_bounds__(TPB) k_b_claim_rest(BArgs a, int64_t g, const int* vlist) {
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE]) return;
    const int nv = st[C_NVIABLE];
    const int lane = threadIdx.x & 63;
    const int X = a.X, Y = a.Y;
    unsigned long long nb = 0;
    for (int w = (blockIdx.x * TPB + threadIdx.x) >> 6; w < nv; w += (gridDim.x * TPB) >> 6) {
        const int i = vlist[w];
        const int net = a.ids[i];
        if (lane == 0) nb += 4 + 4 + 4 + 16 + 32;   // list entry, id, row, point and segment offsets
        for (int64_t k = a.pt_off[a.pids[i]] + lane; k < a.pt_off[a.pids[i] + 1]; k += 64) {
            const int p = a.pt[k], x = p / Y, y = p % Y;
            nb += 4 + 12;   // point, three neighbour claims
            if (y - 1 >= 0) atomicMax(&a.map[x * Y + y - 1], i);
            if (x - 1 >= 0) atomicMax(&a.map[(x - 1) * Y + y], i);
            if (x + 1 < X) atomicMax(&a.map[(x + 1) * Y + y], i);
        }
        wave_segment_cells(a, net, [&](bool valid, int q) {
            if (valid) {
                atomicMax(&a.map[q], i);
                nb += 4;
            }
        });
    }
    flush_bytes(a, nb);
}

// ---- active-list rounds -----------------------------------------------------
// 1. point claims of the active nets (from round 1; round 0 starts from the
//    kept map): only cells holding a lower value can change
template <int G>
__global__ void __launch_bounds__(TPB) k_a_claim(BArgs a, int64_t g) {
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE] || st[C_ROUND] == 0) return;
    int64_t na;
    const int* al = a_list(a, st, na);
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    for (int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G; t < na; t += nth) {
        const int64_t i = al[t];
        if (!active(a, i)) continue;
        const int64_t k1 = a.pt_off[a.pids[i] + 1];
        for (int64_t k = a.pt_off[a.pids[i]] + lg; k < k1; k += G) {
            int* m = &a.map[a.pt[k]];
            if (*m < (int)i) atomicMax(m, (int)i);
        }
    }
}

// 2. active nets owning every point cell go to the viable list
template <int G>
__global__ void __launch_bounds__(TPB) k_a_viable(BArgs a, int64_t g, int* vlist) {
    int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the successor lists' counters, filled by this round's last kernel
        int* nx = a.ctrl + ((g + 1) & 1) * C_WORDS;
        nx[C_NA] = 0;
        nx[C_NU] = 0;
    }
    int64_t na;
    const int* al = a_list(a, st, na);
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    const int64_t rounds = (na + nth - 1) / nth;   // the same for every lane of the wave
    for (int64_t r = 0; r < rounds; r++) {
        const int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G + r * nth;
        const int64_t i = t < na ? al[t] : 0;
        const bool live = t < na && active(a, i);
        const int64_t k0 = live ? a.pt_off[a.pids[i]] : 0, k1 = live ? a.pt_off[a.pids[i] + 1] : 0;
        bool fail = !live;
        for (int64_t k = k0 + lg;; k += G) {
            const bool f = !fail && k < k1 && a.map[a.pt[k]] != (int)i;
            fail = fail || group_bits(__ballot(f), G) != 0;
            if (__ballot(!fail && k + G < k1) == 0) break;
        }
        if (lg == 0 && !fail) vlist[atomicAdd(&st[C_NVIABLE], 1)] = (int)i;
    }
}

// 4. a wave per viable net: commit it if it still owns every point cell
__global__ void __launch_bounds__(TPB) k_a_commit(BArgs a, int64_t g, const int* vlist) {
    int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE]) return;
    const int nv = st[C_NVIABLE];
    const int lane = threadIdx.x & 63;
    for (int w = (blockIdx.x * TPB + threadIdx.x) >> 6; w < nv; w += (gridDim.x * TPB) >> 6) {
        const int i = vlist[w];
        bool lost = false;
        for (int64_t k = a.pt_off[a.pids[i]] + lane; k < a.pt_off[a.pids[i] + 1]; k += 64)
            lost = lost || a.map[a.pt[k]] != i;
        if (__ballot(lost) == 0 && lane == 0) {
            a.batch[i] = st[C_CUR];
            set_assigned(a, i);
            atomicAdd(&a.ctrl[2 * C_WORDS], 1);   // nets placed so far
            st[C_COMMIT] = 1;
        }
    }
}

// 5. after a commit: active nets with a point cell owned by a committed net
//    sit out the batch
template <int G>
__global__ void __launch_bounds__(TPB) k_a_revert(BArgs a, int64_t g) {
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE] || !st[C_COMMIT]) return;
    int64_t na;
    const int* al = a_list(a, st, na);
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    const int64_t rounds = (na + nth - 1) / nth;
    for (int64_t r = 0; r < rounds; r++) {
        const int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G + r * nth;
        const int64_t i = t < na ? al[t] : 0;
        const bool live = t < na && active(a, i);
        const int64_t k0 = live ? a.pt_off[a.pids[i]] : 0, k1 = live ? a.pt_off[a.pids[i] + 1] : 0;
        bool hit = false;
        for (int64_t k = k0 + lg;; k += G) {
            bool h = false;
            if (live && !hit && k < k1) {
                const int pre = a.map[a.pt[k]];
                h = pre != -1 && assigned(a, pre);
            }
            hit = hit || group_bits(__ballot(h), G) != 0;
            if (__ballot(live && !hit && k + G < k1) == 0) break;
        }
        if (lg == 0 && live && hit) {
            set_blocked(a, i);
        }
    }
}

// 6. end of the round.  The batch goes on after a commit below the round
//    limit: cells of blocked nets are released.  Otherwise the batch ends:
//    each cell's list moves past assigned nets, the next batch starts from
//    the kept map with every unassigned net active.  The successor lists go
//    to a buffer neither current list uses.  Thread 0 writes the successor
//    control state.
__global__ void __launch_bounds__(TPB) k_a_clear_end(BArgs a, int64_t g, int64_t cells) {
    __shared__ int s_app[TPB / 64 + 1];
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    int* nx = a.ctrl + ((g + 1) & 1) * C_WORDS;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * TPB;   // grid-stride over cells and list entries
    const bool done = st[C_DONE] != 0;
    const bool commit = st[C_COMMIT] != 0;
    const int round = st[C_ROUND];
    const bool end = !commit || round + 1 >= ROUNDS;
    int fresh = 0;   // the first list buffer neither current list uses
    while (fresh == st[C_AL] || fresh == st[C_UL]) fresh++;
    int* nl = a.lists + (int64_t)fresh * a.n;
    if (!done) {
        if (end) {
            for (int64_t q = t; q < cells; q += stride) {
                int64_t h = a.head[q];
                const int64_t e = a.c_end[q];
                while (h < e && assigned(a, a.c_net[h])) h++;
                a.head[q] = h;
                const int b = h < e ? a.c_net[h] : -1;
                a.base[q] = b;
                a.map[q] = b;
            }
            const int* ul = a.lists + (int64_t)st[C_UL] * a.n;
            const int64_t nu = st[C_NU], passes = (nu + stride - 1) / stride;   // the same for every thread
            for (int64_t ps = 0; ps < passes; ps++) {
                const int64_t u = ps * stride + t;
                const int i = u < nu ? ul[u] : 0;
                const bool left = u < nu && !assigned(a, i);
                block_append(left, i, &nx[C_NU], nl, s_app);
            }
            // every unassigned net is active again in the next batch
            for (int64_t w = t; w < a.nwords; w += stride) a.nst[w] &= NST_ASSIGNED;
        } else {
            for (int64_t q = t; q < cells; q += stride) {
                const int m = a.map[q];
                if (m >= 0 && blocked(a, m)) a.map[q] = -1;
            }
            int64_t na;
            const int* al = a_list(a, st, na);
            const int64_t passes = (na + stride - 1) / stride;   // the same for every thread
            for (int64_t ps = 0; ps < passes; ps++) {
                const int64_t u = ps * stride + t;
                const int i = u < na ? al[u] : 0;
                block_append(u < na && active(a, i), i, &nx[C_NA], nl, s_app);
            }
        }
    }
    if (t == 0) {
        if (!done) {
            int64_t na;
            (void)a_list(a, st, na);
            if (a.trace && g < a.trace_cap) {
                long long* rec = a.trace + g * TR_WORDS;
                rec[TR_CUR] = st[C_CUR];
                rec[TR_ROUND] = round;
                rec[TR_TESTED] = na;
                rec[TR_VIABLE] = st[C_NVIABLE];
                rec[TR_COMM] = st[C_NC];
                rec[TR_HITS] = st[C_NH];
                rec[TR_REL] = st[C_NR];
                rec[TR_NEWOWN] = st[C_NK];
                rec[TR_CLOCK] = (long long)wall_clock64();
            }
            a.stat[0] += 1;
            a.stat[1] += (unsigned long long)na;
            if (round == 0) a.stat[2] += (unsigned long long)st[C_NU];
        }
        int d = st[C_DONE], cur = st[C_CUR], r = st[C_ROUND], have = st[C_HAVE] | (commit ? 1 : 0);
        if (!d) {
            if (!end) {
                r = round + 1;
            } else if (!have) {
                d = 1;                       // a batch without commits: nothing left to place
            } else {
                cur += 1;
                r = 0;
                have = 0;
                if ((int64_t)a.ctrl[2 * C_WORDS] == a.n) d = 1;
            }
        }
        nx[C_DONE] = d;
        nx[C_CUR] = cur;
        nx[C_ROUND] = r;
        nx[C_HAVE] = have;
        nx[C_COMMIT] = 0;
        nx[C_NVIABLE] = 0;
        if (done) {
            nx[C_NA] = st[C_NA];
            nx[C_AL] = st[C_AL];
            nx[C_NU] = st[C_NU];
            nx[C_UL] = st[C_UL];
        } else if (end) {
            nx[C_NA] = -1;           // the new unassigned list is the active list
            nx[C_AL] = fresh;
            nx[C_UL] = fresh;
        } else {
            nx[C_AL] = fresh;        // C_NA counted above
            nx[C_NU] = st[C_NU];
            nx[C_UL] = st[C_UL];
        }
    }
}

// ---- release-driven rounds --------------------------------------------------
// 1. from round 1: each cell the previous round released goes to its
//    highest-priority active lister (the first active entry of its list),
//    which joins the viability test.  A group of G lanes scans a cell's list
//    G entries a step.
template <int G>
__global__ void __launch_bounds__(TPB) k_r_claim(BArgs a, int64_t g) {
    __shared__ int s_buf[GL_CAP], s_hdr[2];
    int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE] || st[C_ROUND] == 0) return;
    GroupList gl{s_buf, s_hdr, &st[C_NK], a.cand};
    gl.init();
    const int64_t nr = st[C_NR];
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    const int64_t rounds = (nr + nth - 1) / nth;   // the same for every thread
    unsigned long long nb = 0;
    for (int64_t r = 0; r < rounds; r++) {
        const int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G + r * nth;
        const bool live = t < nr;
        const int q = live ? a.rel[t] : 0;
        int64_t h = live ? a.cur[q] : 0;
        const int64_t e = live ? a.c_end[q] : 0;
        if (live && lg == 0) nb += 4 + 16;   // cell, its cursor and list end
        bool found = !live;
        bool fresh = false;   // this lane took the cell for a net not yet listed this round
        int owner = 0;
        for (;;) {
            const int64_t k = h + lg;
            int j = -1;
            bool act = false;
            if (!found && k < e) {
                j = a.c_net[k];
                act = active(a, j);
                nb += 8;   // list entry, state word
            }
            const uint64_t m = group_bits(__ballot(act), G);
            if (m != 0 && lg == __builtin_ctzll(m)) {
                a.map[q] = j;
                a.cur[q] = k;   // the entries before k stay inactive for the rest of the batch
                owner = j;
                fresh = atomicExch(&a.cstamp[j], (int)g + 1) != (int)g + 1;
                nb += 4 + 8 + 8 + (fresh ? 4 : 0);   // map, cursor, stamp, new-owner entry
            }
            found = found || m != 0;
            h += G;
            if (__ballot(!found && h < e) == 0) break;
        }
        gl.add(fresh, owner);
    }
    gl.flush();
    flush_bytes(a, nb);
}

// 2. the nets under test (every unassigned net in round 0, the new owners of
//    released cells later) that own every point cell go to the viable list
template <int G>
__global__ void __launch_bounds__(TPB) k_r_viable(BArgs a, int64_t g, int* vlist) {
    int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the successor lists' counters, filled by this round's last kernel
        int* nx = a.ctrl + ((g + 1) & 1) * C_WORDS;
        nx[C_NU] = 0;
        nx[C_NR] = 0;
    }
    int64_t na;
    const int* al = r_list(a, st, na);
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    const int64_t rounds = (na + nth - 1) / nth;   // the same for every lane of the wave
    unsigned long long nb = 0;
    for (int64_t r = 0; r < rounds; r++) {
        const int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G + r * nth;
        const int64_t i = t < na ? al[t] : 0;
        const bool live = t < na && active(a, i);
        if (t < na && lg == 0) nb += 4 + 4 + (live ? 4 + 16 : 0);   // entry, state, row, point offsets
        const int64_t k0 = live ? a.pt_off[a.pids[i]] : 0, k1 = live ? a.pt_off[a.pids[i] + 1] : 0;
        bool fail = !live;
        for (int64_t k = k0 + lg;; k += G) {
            const bool chk = !fail && k < k1;
            const bool f = chk && a.map[a.pt[k]] != (int)i;
            if (chk) nb += 8;   // point, map cell
            fail = fail || group_bits(__ballot(f), G) != 0;
            if (__ballot(!fail && k + G < k1) == 0) break;
        }
        if (lg == 0 && !fail) {
            vlist[atomicAdd(&st[C_NVIABLE], 1)] = (int)i;
            if (st[C_ROUND] == 0) a.v0[atomicAdd(&st[C_NV0], 1)] = (int)i;
            nb += st[C_ROUND] == 0 ? 8 : 4;
        }
    }
    flush_bytes(a, nb);
}

// 4. a wave per viable net: commit it if it still owns every point cell;
//    a committed net's wave then lists the cells it owns (points, their
//    neighbours, sampled segment cells), once each.  Commits do not change
//    the map, so a net's cells are the same whichever nets of the round
//    commit around it.
__global__ void __launch_bounds__(TPB) k_r_commit(BArgs a, int64_t g, const int* vlist) {
    __shared__ int s_buf[GL_CAP], s_hdr[2];
    int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE]) return;
    GroupList gl{s_buf, s_hdr, &st[C_NH], a.hits};
    gl.init();
    const int nv = st[C_NVIABLE];
    const int lane = threadIdx.x & 63;
    const int X = a.X, Y = a.Y;
    const int stamp = (int)g + 1;
    unsigned long long nb = 0;
    for (int w = (blockIdx.x * TPB + threadIdx.x) >> 6; w < nv; w += (gridDim.x * TPB) >> 6) {
        const int i = vlist[w];
        const int64_t k0 = a.pt_off[a.pids[i]], k1 = a.pt_off[a.pids[i] + 1];
        if (lane == 0) nb += 4 + 4 + 16;   // list entry, row, point offsets
        bool lost = false;
        for (int64_t k = k0 + lane; k < k1; k += 64) {
            lost = lost || a.map[a.pt[k]] != i;
            nb += 8;   // point, map cell
        }
        if (__ballot(lost) != 0) continue;
        if (lane == 0) {
            a.comm[atomicAdd(&st[C_NC], 1)] = i;
            a.batch[i] = st[C_CUR];
            set_assigned(a, i);
            atomicAdd(&a.ctrl[2 * C_WORDS], 1);   // nets placed so far
            st[C_COMMIT] = 1;
            nb += 16 + 16;   // committed entry, batch, state word; segment offsets
        }
        auto hit = [&](int q) {
            const bool own = a.map[q] == i;
            const bool h = own && atomicExch(&a.hstamp[q], stamp) != stamp;
            nb += 4 + (own ? 8 : 0) + (h ? 4 : 0);   // map cell, stamp, hit entry
            gl.add(h, q);
        };
        for (int64_t k = k0 + lane; k < k1; k += 64) {
            const int p = a.pt[k], x = p / Y, y = p % Y;
            hit(p);
            if (y - 1 >= 0) hit(x * Y + y - 1);
            if (x - 1 >= 0) hit((x - 1) * Y + y);
            if (x + 1 < X) hit((x + 1) * Y + y);
        }
        wave_segment_cells(a, a.ids[i], [&](bool valid, int q) {
            const bool own = valid && a.map[q] == i;
            const bool h = own && atomicExch(&a.hstamp[q], stamp) != stamp;
            if (valid) nb += 4 + (own ? 8 : 0) + (h ? 4 : 0);
            gl.add(h, q);
        });
    }
    gl.flush();
    flush_bytes(a, nb);
}

// 5. every active net listing one of those cells sits out the batch (a net
//    is blocked when one of its point cells is owned by a net committed in
//    this batch; owners committed in earlier rounds blocked their listers
//    then, and their cells only shrink).  A group of G lanes per cell (a
//    wave: the lists of cells in dense regions run to thousands of nets).
template <int G>
__global__ void __launch_bounds__(TPB) k_r_block(BArgs a, int64_t g) {
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    if (st[C_DONE] || !st[C_COMMIT]) return;
    const int64_t nh = st[C_NH];
    const int64_t nth = (int64_t)gridDim.x * TPB / G;
    const int lg = threadIdx.x & (G - 1);
    unsigned long long nb = 0;
    for (int64_t t = ((int64_t)blockIdx.x * TPB + threadIdx.x) / G; t < nh; t += nth) {
        const int q = a.hits[t];
        const int64_t e = a.c_end[q];
        for (int64_t k = a.cur[q] + lg; k < e; k += G) {
            const int j = a.c_net[k];
            nb += 8;   // list entry, state word
            if (active(a, j)) {
                set_blocked(a, j);
                nb += 4;
            }
        }
        // every entry from the cursor on is now inactive for the rest of the
        // batch (a later release or hit of the cell finds no active lister)
        if (lg == 0) {
            a.cur[q] = e;
            nb += 4 + 24;   // hit entry, list end, cursor read and write
        }
    }
    flush_bytes(a, nb);
}

// 6. end of the round.  The batch goes on after a commit below the round
//    limit: cells of blocked nets are released.  Otherwise the batch ends:
//    each cell's list moves past assigned nets, the next batch starts from
//    the kept map with every unassigned net active.  The successor lists go
//    to a buffer neither current list uses.  Thread 0 writes the successor
//    control state.
__global__ void __launch_bounds__(TPB) k_r_clear_end(BArgs a, int64_t g, int64_t cells) {
    __shared__ int s_buf[GL_CAP], s_hdr[2];
    const int* st = a.ctrl + (g & 1) * C_WORDS;
    int* nx = a.ctrl + ((g + 1) & 1) * C_WORDS;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * TPB;   // grid-stride over cells and list entries
    const bool done = st[C_DONE] != 0;
    const bool commit = st[C_COMMIT] != 0;
    const int round = st[C_ROUND];
    const bool end = !commit || round + 1 >= ROUNDS;
    const int fresh = 1 - st[C_UL];   // the candidate-list buffer not in use
    int* nl = a.lists + (int64_t)fresh * a.n;
    unsigned long long nb = 0;
    if (!done) {
        if (end) {
            // Round 0 of the next batch tests only nets that may own every
            // point cell of the kept map: a net viable in round 0 of this
            // batch and still unassigned keeps its cells (a list's head moves
            // past assigned nets only), and any other net can only become
            // viable by taking a cell whose head moved.
            // Every COMPACT_DIV-th of the nets placed since the last time,
            // the lists drop their assigned entries (order kept): the scans
            // of the later rounds then skip no dead entries.
            GroupList gl{s_buf, s_hdr, &nx[C_NU], nl};
            gl.init();
            const int stamp = -(int)(g + 1);
            const bool compact = (int64_t)(a.ctrl[2 * C_WORDS] - st[C_LASTC]) * COMPACT_DIV >= a.n;
            if (compact) {
                // a group of CG lanes per cell compacts its list in place:
                // every lane reads its entry before the group's ballot, and
                // an entry moves to a slot at or below its own
                constexpr int CG = 16;
                const int lg = threadIdx.x & (CG - 1);
                const int64_t ng = stride / CG;
                const int64_t gpasses = (cells + ng - 1) / ng;
                for (int64_t ps = 0; ps < gpasses; ps++) {
                    const int64_t q = ps * ng + t / CG;
                    const bool live = q < cells;
                    const int64_t s0 = live ? a.c_off[q] : 0, e = live ? a.c_end[q] : 0;
                    int64_t w = s0;
                    for (int64_t k = live ? a.head[q] : 0;; k += CG) {
                        const int64_t kk = k + lg;
                        int j = -1;
                        bool keep = false;
                        if (kk < e) nb += 12;   // entry read, state word, entry written
                        if (kk < e) {
                            j = a.c_net[kk];
                            keep = !assigned(a, j);
                        }
                        const uint64_t m = group_bits(__ballot(keep), CG);
                        if (keep) a.c_net[w + __popcll(m & ((1ull << lg) - 1))] = j;
                        w += __popcll(m);
                        if (__ballot(k + CG < e) == 0) break;
                    }
                    int b = -1;
                    bool want = false;
                    if (live && lg == 0) {
                        a.c_end[q] = w;
                        a.head[q] = s0;
                        a.cur[q] = s0;
                        b = w > s0 ? a.c_net[s0] : -1;
                        const bool moved = b != a.base[q];
                        if (moved) a.base[q] = b;
                        a.map[q] = b;
                        want = moved && b >= 0 && atomicExch(&a.cstamp[b], stamp) != stamp;
                    }
                    gl.add(want, b);
                }
            } else {
                const int64_t passes = (cells + stride - 1) / stride;
                for (int64_t ps = 0; ps < passes; ps++) {
                    const int64_t q = ps * stride + t;
                    bool want = false;
                    int b = -1;
                    if (q < cells) {
                        const int64_t h0 = a.head[q], e = a.c_end[q];
                        int64_t h = h0;
                        while (h < e && assigned(a, a.c_net[h])) h++;
                        nb += 16 + 8 * (h - h0 + (h < e ? 1 : 0)) + 8 + 12;   // head, end; entries; cur, base, map
                        if (h != h0) a.head[q] = h;
                        b = h < e ? a.c_net[h] : -1;
                        // the first unassigned lister changed exactly where
                        // the previous one was assigned
                        const bool moved = b != a.base[q];
                        a.cur[q] = h;
                        if (moved) a.base[q] = b;
                        a.map[q] = b;
                        want = moved && b >= 0 && atomicExch(&a.cstamp[b], stamp) != stamp;
                    }
                    gl.add(want, b);
                }
            }
            const int64_t nv0 = st[C_NV0], vpasses = (nv0 + stride - 1) / stride;
            for (int64_t ps = 0; ps < vpasses; ps++) {
                const int64_t u = ps * stride + t;
                const int i = u < nv0 ? a.v0[u] : 0;
                const bool want = u < nv0 && !assigned(a, i) && atomicExch(&a.cstamp[i], stamp) != stamp;
                gl.add(want, i);
            }
            gl.flush();
            // every unassigned net is active again in the next batch
            for (int64_t w = t; w < a.nwords; w += stride) {
                a.nst[w] &= NST_ASSIGNED;
                nb += 8;
            }
        } else {
            // UNR independent cells per lane and pass
            GroupList gl{s_buf, s_hdr, &nx[C_NR], a.rel};
            gl.init();
            constexpr int UNR = 4;
            const int64_t per = stride * UNR;
            for (int64_t ps = 0; ps * per < cells; ps++) {
                int m[UNR];
                for (int u = 0; u < UNR; u++) {
                    const int64_t q = ps * per + u * stride + t;
                    m[u] = q < cells ? a.map[q] : -1;
                }
                bool rl[UNR];
                for (int u = 0; u < UNR; u++) rl[u] = m[u] >= 0 && blocked(a, m[u]);
                for (int u = 0; u < UNR; u++)
                    nb += (ps * per + u * stride + t < cells ? 4 : 0) + (m[u] >= 0 ? 4 : 0) + (rl[u] ? 8 : 0);
                for (int u = 0; u < UNR; u++) {
                    const int64_t q = ps * per + u * stride + t;
                    if (rl[u]) a.map[q] = -1;
                    gl.add(rl[u], (int)q);
                }
            }
            gl.flush();
        }
    }
    if (t == 0) {
        if (!done) {
            int64_t na;
            (void)r_list(a, st, na);
            if (a.trace && g < a.trace_cap) {
                long long* rec = a.trace + g * TR_WORDS;
                rec[TR_CUR] = st[C_CUR];
                rec[TR_ROUND] = round;
                rec[TR_TESTED] = na;
                rec[TR_VIABLE] = st[C_NVIABLE];
                rec[TR_COMM] = st[C_NC];
                rec[TR_HITS] = st[C_NH];
                rec[TR_REL] = st[C_NR];
                rec[TR_NEWOWN] = st[C_NK];
                rec[TR_CLOCK] = (long long)wall_clock64();
            }
            a.stat[0] += 1;
            a.stat[1] += (unsigned long long)na;
            // unassigned nets at the batch start (placed now, less round 0's commits)
            if (round == 0) a.stat[2] += (unsigned long long)(a.n - (a.ctrl[2 * C_WORDS] - st[C_NC]));
        }
        int d = st[C_DONE], cur = st[C_CUR], r = st[C_ROUND], have = st[C_HAVE] | (commit ? 1 : 0);
        if (!d) {
            if (!end) {
                r = round + 1;
            } else if (!have) {
                d = 1;                       // a batch without commits: nothing left to place
            } else {
                cur += 1;
                r = 0;
                have = 0;
                if ((int64_t)a.ctrl[2 * C_WORDS] == a.n) d = 1;
            }
        }
        nx[C_DONE] = d;
        nx[C_CUR] = cur;
        nx[C_ROUND] = r;
        nx[C_HAVE] = have;
        nx[C_COMMIT] = 0;
        nx[C_NVIABLE] = 0;
        nx[C_NV0] = !done && end ? 0 : st[C_NV0];
        nx[C_LASTC] = !done && end && (int64_t)(a.ctrl[2 * C_WORDS] - st[C_LASTC]) * COMPACT_DIV >= a.n
                          ? a.ctrl[2 * C_WORDS]
                          : st[C_LASTC];
        nx[C_NK] = 0;
        nx[C_NC] = 0;
        nx[C_NH] = 0;
        if (done) {
            nx[C_NU] = st[C_NU];
            nx[C_UL] = st[C_UL];
            nx[C_NR] = st[C_NR];
        } else if (end) {
            nx[C_UL] = fresh;        // C_NU counted above
        } else {
            nx[C_NU] = st[C_NU];     // C_NR counted above
            nx[C_UL] = st[C_UL];
        }
    }
    flush_bytes(a, nb);
}

__global__ void k_fill_b(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

// Start state: every net unassigned and active, listed in buffer 0.
__global__ void k_start(BArgs a) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < a.n) a.lists[i] = (int)i;
    if (i == 0) {
        a.ctrl[C_NU] = (int)a.n;
        a.ctrl[C_UL] = 0;
        a.ctrl[C_NA] = -1;   // active-list rounds: the unassigned list is the active list
        a.ctrl[C_AL] = 0;
    }
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

} // namespace

int run_schedule_batched(igr_ctx* c, const int* h_order, int64_t n, int mode, int* h_batch_of,
                         const int64_t* d_pt_off, const int* d_pt) {
    if (n == 0) return 0;
    if (n >= 0x7fffffff) throw std::invalid_argument("too many nets for the scheduler");
    KScope ks(c, K_SCHED);
    Sched& S = c->sched;
    NetsDev& N = c->nets;
    hipStream_t st = c->stream;
    std::vector<int> ids(n);
    for (int64_t i = 0; i < n; i++) ids[i] = h_order[n - 1 - i];
    S.ids.upload(ids.data(), n, st);
    if (d_pt_off) {
        std::vector<int> pids(n);
        for (int64_t i = 0; i < n; i++) pids[i] = (int)(n - 1 - i);
        S.pids.upload(pids.data(), n, st);
    }
    const int64_t XYc = (int64_t)c->g.X * c->g.Y;
    S.map.alloc(XYc);
    S.base.alloc(XYc);
    S.head.alloc(XYc);
    S.c_off.alloc(XYc + 1);
    S.c_end.alloc(XYc);
    S.cnt.zero(XYc, st);
    S.batch.alloc(n);
    S.viable.alloc(n);
    S.lists.alloc(NLISTS * n);
    // Release-driven rounds from this many nets (active-list rounds below).
    const bool release = c->tune.sched_kind ? c->tune.sched_kind == 2 : n >= DEF_SCHED_RELEASE_NETS;
    if (release) {
        S.rel.alloc(XYc);
        S.cand.alloc(n);
        S.cstamp.zero(n, st);
        S.comm.alloc(n);
        S.hits.alloc(XYc);
        S.hstamp.zero(XYc, st);
        S.cur.alloc(XYc);
    }
    S.stat.zero(3, st);
    S.ctrl.zero(2 * C_WORDS + 1, st);
    const int64_t nwords = (n + 15) / 16;
    S.bits.zero(nwords, st);
    BArgs a{n, S.ids.p, d_pt_off ? S.pids.p : S.ids.p, d_pt_off ? d_pt_off : N.pt_off.p, d_pt_off ? d_pt : N.pt.p,
            N.hs_off.p, N.hs.p, N.vs_off.p, N.vs.p, S.map.p, S.base.p, S.c_off.p, S.c_end.p, nullptr, S.head.p,
            S.batch.p, S.ctrl.p, S.lists.p,
            release ? S.rel.p : nullptr, release ? S.cand.p : nullptr, release ? S.cstamp.p : nullptr,
            release ? S.lists.p + 2 * n : nullptr,
            release ? S.comm.p : nullptr, release ? S.hits.p : nullptr, release ? S.hstamp.p : nullptr,
            release ? S.cur.p : nullptr, S.stat.p,
            (uint32_t*)S.bits.p, nwords,
            c->g.X, c->g.Y,
            (mode == 1 || mode == 2) ? 2 : 15, nullptr, 0, nullptr};
    // kernel timing on: the round kernels count the bytes they move
    DevBuf<unsigned long long> byte_slots;
    if (c->prof.on && release) {
        byte_slots.zero(BYTE_SLOTS, st);
        a.bytes = byte_slots.p;
    }
    // IGR_SCHED_TRACE=<file>: append one line per round (the TR_* record)
    const char* trace_path = std::getenv("IGR_SCHED_TRACE");
    DevBuf<long long> trace;
    if (trace_path && *trace_path) {
        a.trace_cap = 1 << 20;
        trace.zero(a.trace_cap * TR_WORDS, st);
        a.trace = trace.p;
    }
    // Per-cell lists of the nets' point cells, highest priority first.
    k_cell_count<<<blocks(n), TPB, 0, st>>>(a, S.cnt.p);
    const int64_t nb = (XYc + SCAN_BLOCK - 1) / SCAN_BLOCK;
    S.bsum.alloc(nb);
    k_scan_reduce<<<(unsigned)nb, TPB, 0, st>>>(S.cnt.p, XYc, S.bsum.p);
    k_scan_bsums<<<1, TPB, 0, st>>>(S.bsum.p, nb);
    k_scan_apply<<<(unsigned)nb, TPB, 0, st>>>(S.cnt.p, XYc, S.bsum.p, S.c_off.p);
    HIP_OK(hipGetLastError());
    int64_t entries = 0;
    HIP_OK(hipMemcpyAsync(&entries, S.c_off.p + XYc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    S.c_net.alloc(std::max<int64_t>(1, entries));
    a.c_net = S.c_net.p;
    HIP_OK(hipMemsetAsync(S.cnt.p, 0, sizeof(int) * XYc, st));
    k_cell_fill<<<blocks(n), TPB, 0, st>>>(a, S.cnt.p);
    k_cell_sort<<<blocks(XYc), TPB, 0, st>>>(a, XYc);
    k_fill_b<<<blocks(n), TPB, 0, st>>>(S.batch.p, n, -1);
    k_start<<<blocks(n), TPB, 0, st>>>(a);
    HIP_OK(hipGetLastError());
    // Lanes per net: about half the average point count, at most 16.
    const int64_t entries_total = entries;
    const double avg = (double)entries_total / (double)n;
    const int G = c->tune.sched_group ? c->tune.sched_group : avg <= 6 ? 4 : avg <= 16 ? 8 : 16;
    const int CHUNK = c->tune.sched_chunk ? c->tune.sched_chunk : DEF_SCHED_CHUNK;
    const unsigned ggrid = (unsigned)std::min<int64_t>(2048, blocks(n * G));
    const unsigned wgrid = 256;   // one wave per viable net
    const unsigned gm = (unsigned)std::min<int64_t>(2048, blocks(std::max(XYc, n)));   // grid-stride
    int* h_ctrl = nullptr;
    HIP_OK(hipHostMalloc(&h_ctrl, sizeof(int) * (2 * C_WORDS + 1)));
    const int64_t max_rounds = (int64_t)ROUNDS * (n + 1) + CHUNK;
    int64_t g = 0;
    const auto t_start = std::chrono::steady_clock::now();
    auto t_note = t_start;
    try {
        for (;;) {
            if (g > max_rounds) throw std::runtime_error("scheduler did not finish");
            const auto t_now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double>(t_now - t_note).count() > 20) {
                t_note = t_now;
                progress("scheduler: %d batches after %.1f s", h_ctrl[(g & 1) * C_WORDS + C_CUR],
                         std::chrono::duration<double>(t_now - t_start).count());
            }
            for (int r = 0; r < CHUNK; r++, g++) {
                if (release) {
                    switch (G) {
                    case 4:
                        k_r_claim<4><<<ggrid, TPB, 0, st>>>(a, g);
                        k_r_viable<4><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    case 8:
                        k_r_claim<8><<<ggrid, TPB, 0, st>>>(a, g);
                        k_r_viable<8><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    default:
                        k_r_claim<16><<<ggrid, TPB, 0, st>>>(a, g);
                        k_r_viable<16><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    }
                    k_b_claim_rest<<<wgrid, TPB, 0, st>>>(a, g, S.viable.p);
                    k_r_commit<<<wgrid, TPB, 0, st>>>(a, g, S.viable.p);
                    k_r_block<64><<<ggrid, TPB, 0, st>>>(a, g);
                    k_r_clear_end<<<gm, TPB, 0, st>>>(a, g, XYc);
                } else {
                    switch (G) {
                    case 4:
                        k_a_claim<4><<<ggrid, TPB, 0, st>>>(a, g);
                        k_a_viable<4><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    case 8:
                        k_a_claim<8><<<ggrid, TPB, 0, st>>>(a, g);
                        k_a_viable<8><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    default:
                        k_a_claim<16><<<ggrid, TPB, 0, st>>>(a, g);
                        k_a_viable<16><<<ggrid, TPB, 0, st>>>(a, g, S.viable.p);
                        break;
                    }
                    k_b_claim_rest<<<wgrid, TPB, 0, st>>>(a, g, S.viable.p);
                    k_a_commit<<<wgrid, TPB, 0, st>>>(a, g, S.viable.p);
                    switch (G) {
                    case 4: k_a_revert<4><<<ggrid, TPB, 0, st>>>(a, g); break;
                    case 8: k_a_revert<8><<<ggrid, TPB, 0, st>>>(a, g); break;
                    default: k_a_revert<16><<<ggrid, TPB, 0, st>>>(a, g); break;
                    }
                    k_a_clear_end<<<gm, TPB, 0, st>>>(a, g, XYc);
                }
            }
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(h_ctrl, S.ctrl.p, sizeof(int) * (2 * C_WORDS + 1), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (h_ctrl[(g & 1) * C_WORDS + C_DONE]) break;
        }
    } catch (...) {
        (void)hipHostFree(h_ctrl);
        throw;
    }
    const int nbatch = h_ctrl[(g & 1) * C_WORDS + C_CUR];
    HIP_OK(hipHostFree(h_ctrl));
    if (a.bytes) {
        std::vector<unsigned long long> slots(BYTE_SLOTS);
        HIP_OK(hipMemcpyAsync(slots.data(), a.bytes, sizeof(unsigned long long) * BYTE_SLOTS, hipMemcpyDeviceToHost,
                              st));
        HIP_OK(hipStreamSynchronize(st));
        double tot = 0;
        for (unsigned long long v : slots) tot += (double)v;
        c->prof.bytes[K_SCHED] += tot;
    }
    if (a.trace) {
        const int64_t nrec = std::min<int64_t>(g, a.trace_cap);
        std::vector<long long> rec((size_t)(nrec * TR_WORDS));
        HIP_OK(hipMemcpyAsync(rec.data(), a.trace, sizeof(long long) * rec.size(), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        int khz = 0;
        HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        if (FILE* f = std::fopen(trace_path, "a")) {
            std::fprintf(f, "# nets %lld rounds %s lanes %d clock_khz %d\n", (long long)n,
                         release ? "release-driven" : "active-list", G, khz);
            std::fprintf(f, "g batch round tested viable committed hits released new_owners clock\n");
            for (int64_t r = 0; r < nrec; r++) {
                const long long* q = rec.data() + r * TR_WORDS;
                if (r > 0 && q[TR_CLOCK] == 0) break;   // rounds after the last batch record nothing
                std::fprintf(f, "%lld %lld %lld %lld %lld %lld %lld %lld %lld %lld\n", (long long)r, q[0], q[1], q[2],
                             q[3], q[4], q[5], q[6], q[7], q[8]);
            }
            std::fclose(f);
        }
    }
    {
        unsigned long long sv[3] = {0, 0, 0};
        HIP_OK(hipMemcpyAsync(sv, S.stat.p, sizeof(sv), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        progress("scheduler (%s rounds): %d batches, %llu rounds, %.0f nets tested per round, %.0f unassigned per "
                 "batch", release ? "release-driven" : "active-list", nbatch, sv[0],
                 sv[0] ? (double)sv[1] / (double)sv[0] : 0.0, nbatch ? (double)sv[2] / nbatch : 0.0);
    }
    std::vector<int> b(n);
    HIP_OK(hipMemcpyAsync(b.data(), S.batch.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; i++) {
        if (b[i] < 0) throw std::runtime_error("scheduler left a net without a batch");
        h_batch_of[n - 1 - i] = b[i];
    }
    return nbatch;
}

} // namespace igr_dev
