// Batched form of the conflict-free net batch scheduler: the same batches as
// the round-by-round scheduler (src/database_cuda.hpp:288-434,437-631, with
// the deterministic claim semantics stated in sched_kernels.hip), with less
// work and no host synchronisation per round.
//
//  * Round 0 of every batch starts from the map every unassigned net claims
//    with its point cells: per cell, the highest-priority unassigned net that
//    lists the cell.  It is kept across batches from a per-cell list of nets
//    sorted by priority (built once per pass) and a head index per cell that
//    skips assigned nets; only the point cells of the nets a batch committed
//    change, so round 0 claims nothing.
//  * In later rounds a point claim can only change a cell the previous
//    round's clear emptied (values otherwise only grow within a batch), so a
//    net claims a point only where the map holds a lower value.
//  * The round and batch control lives in a device word block; the host
//    enqueues rounds in chunks and reads back one flag per chunk.  Kernels of
//    rounds after the end return at once.
#include "igr_ctx.hpp"

#include <algorithm>

namespace igr_dev {

namespace {

constexpr int TPB = 256;
constexpr int ROUNDS = 30;        // rounds per batch (src/database_cuda.hpp:612)
constexpr int CHUNK = 16;         // rounds enqueued per host check

enum { C_DONE = 0, C_CUR, C_ROUND, C_COMMIT, C_REVERT, C_END, C_HAVE, C_ASSIGNED, C_NCOMMITTED, C_WORDS };

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
    int* c_net;                   // per-cell lists, highest index first
    int64_t* head;                // X*Y first unassigned entry of each list
    int* batch;
    int* status;                  // 1 active in this batch, 2 blocked, 0 assigned earlier
    int* viable;
    int* committed;               // nets committed in the current batch
    int* ctrl;
    int X, Y, sparsity;
};

__device__ __forceinline__ bool active(const BArgs& a, int64_t i) { return a.status[i] == 1 && a.batch[i] == -1; }

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
__device__ __forceinline__ void claim(int* m, int i) {
    if (*m < i) atomicMax(m, i);
}

__device__ __forceinline__ bool owns_points(const BArgs& a, int64_t i) {
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++)
        if (a.map[a.pt[k]] != (int)i) return false;
    return true;
}

// 1. point claims (round 0 starts from the kept map)
__global__ void __launch_bounds__(TPB) k_b_claim(BArgs a) {
    if (a.ctrl[C_DONE] || a.ctrl[C_ROUND] == 0) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !active(a, i)) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) claim(&a.map[a.pt[k]], (int)i);
}

// 2. nets owning every point cell
__global__ void __launch_bounds__(TPB) k_b_viable(BArgs a) {
    if (a.ctrl[C_DONE]) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n) return;
    a.viable[i] = active(a, i) && owns_points(a, i);
}

// 3. their neighbour cells and sampled segment cells
__global__ void __launch_bounds__(TPB) k_b_claim_rest(BArgs a) {
    if (a.ctrl[C_DONE]) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !a.viable[i]) return;
    const int net = a.ids[i];
    const int X = a.X, Y = a.Y;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
        const int p = a.pt[k], x = p / Y, y = p % Y;
        if (y - 1 >= 0) claim(&a.map[x * Y + y - 1], (int)i);
        if (x - 1 >= 0) claim(&a.map[(x - 1) * Y + y], (int)i);
        if (x + 1 < X) claim(&a.map[(x + 1) * Y + y], (int)i);
    }
    for (int64_t k = a.hs_off[net]; k < a.hs_off[net + 1]; k++) {
        const int s = a.hs[2 * k], e = a.hs[2 * k + 1], y = s % Y;
        for (int x = s / Y; x <= e / Y; x += a.sparsity) claim(&a.map[x * Y + y], (int)i);
    }
    for (int64_t k = a.vs_off[net]; k < a.vs_off[net + 1]; k++) {
        const int s = a.vs[2 * k], e = a.vs[2 * k + 1], x = s / Y;
        for (int y = s % Y; y <= e % Y; y += a.sparsity) claim(&a.map[x * Y + y], (int)i);
    }
}

// 4. commit the viable nets still owning every point cell
__global__ void __launch_bounds__(TPB) k_b_commit(BArgs a) {
    if (a.ctrl[C_DONE]) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !a.viable[i] || !owns_points(a, i)) return;
    a.batch[i] = a.ctrl[C_CUR];
    a.committed[atomicAdd(&a.ctrl[C_NCOMMITTED], 1)] = (int)i;
    atomicAdd(&a.ctrl[C_ASSIGNED], 1);
    a.ctrl[C_COMMIT] = 1;
}

// Single thread: end of the round's commit.
__global__ void k_b_control(int* ctrl) {
    if (ctrl[C_DONE]) return;
    const int round = ctrl[C_ROUND];
    if (ctrl[C_COMMIT]) {
        ctrl[C_COMMIT] = 0;
        ctrl[C_HAVE] = 1;
        ctrl[C_REVERT] = 1;
        ctrl[C_END] = round + 1 >= ROUNDS ? 1 : 0;
        ctrl[C_ROUND] = round + 1;
    } else {
        ctrl[C_REVERT] = 0;
        ctrl[C_END] = 1;
    }
}

// 5. nets with a point cell owned by a committed net sit out the batch
__global__ void __launch_bounds__(TPB) k_b_revert(BArgs a) {
    if (a.ctrl[C_DONE] || !a.ctrl[C_REVERT]) return;
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !active(a, i)) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
        const int pre = a.map[a.pt[k]];
        if (pre != -1 && a.batch[pre] != -1) { a.status[i] = 2; return; }
    }
}

// 6. their cells are released
__global__ void __launch_bounds__(TPB) k_b_clear(BArgs a, int64_t cells) {
    if (a.ctrl[C_DONE] || !a.ctrl[C_REVERT]) return;
    const int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= cells) return;
    const int m = a.map[q];
    if (m >= 0 && a.status[m] == 2) a.map[q] = -1;
}

// 7. end of a batch: the lists of the committed nets' point cells (disjoint
//    between committed nets) move to their first unassigned entry
__global__ void __launch_bounds__(TPB) k_b_base(BArgs a) {
    if (a.ctrl[C_DONE] || !a.ctrl[C_END]) return;
    const int nc = a.ctrl[C_NCOMMITTED];
    for (int t = blockIdx.x * TPB + threadIdx.x; t < nc; t += gridDim.x * TPB) {
        const int i = a.committed[t];
        for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
            const int p = a.pt[k];
            int64_t h = a.head[p];
            const int64_t e = a.c_off[p + 1];
            while (h < e && a.batch[a.c_net[h]] != -1) h++;
            a.head[p] = h;
            a.base[p] = h < e ? a.c_net[h] : -1;
        }
    }
}

// 8. the next batch starts from the kept map with every unassigned net
__global__ void __launch_bounds__(TPB) k_b_next(BArgs a, int64_t cells) {
    if (a.ctrl[C_DONE] || !a.ctrl[C_END]) return;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t < cells) a.map[t] = a.base[t];
    if (t < a.n) a.status[t] = a.batch[t] == -1 ? 1 : 0;
}

// Single thread: start the next batch, or finish.
__global__ void k_b_advance(int* ctrl, int64_t n) {
    if (ctrl[C_DONE] || !ctrl[C_END]) return;
    if (!ctrl[C_HAVE]) {
        ctrl[C_DONE] = 1;     // a batch without commits: nothing left to place
        return;
    }
    ctrl[C_CUR] += 1;
    ctrl[C_ROUND] = 0;
    ctrl[C_END] = 0;
    ctrl[C_HAVE] = 0;
    ctrl[C_NCOMMITTED] = 0;
    if ((int64_t)ctrl[C_ASSIGNED] == n) ctrl[C_DONE] = 1;
}

__global__ void k_fill_b(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
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
    S.cnt.zero(XYc, st);
    S.batch.alloc(n);
    S.status.alloc(n);
    S.viable.alloc(n);
    S.committed.alloc(n);
    S.ctrl.zero(C_WORDS, st);
    BArgs a{n, S.ids.p, d_pt_off ? S.pids.p : S.ids.p, d_pt_off ? d_pt_off : N.pt_off.p, d_pt_off ? d_pt : N.pt.p,
            N.hs_off.p, N.hs.p, N.vs_off.p, N.vs.p, S.map.p, S.base.p, S.c_off.p, nullptr, S.head.p,
            S.batch.p, S.status.p, S.viable.p, S.committed.p, S.ctrl.p, c->g.X, c->g.Y,
            (mode == 1 || mode == 2) ? 2 : 15};
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
    k_fill_b<<<blocks(n), TPB, 0, st>>>(S.status.p, n, 1);
    HIP_OK(hipGetLastError());
    int* h_ctrl = nullptr;
    HIP_OK(hipHostMalloc(&h_ctrl, sizeof(int) * C_WORDS));
    const int64_t max_rounds = (int64_t)ROUNDS * (n + 1) + CHUNK;
    const unsigned gn = blocks(n), gc = blocks(XYc), gm = blocks(std::max(XYc, n));
    try {
        for (int64_t rounds = 0;; rounds += CHUNK) {
            if (rounds > max_rounds) throw std::runtime_error("scheduler did not finish");
            for (int r = 0; r < CHUNK; r++) {
                k_b_claim<<<gn, TPB, 0, st>>>(a);
                k_b_viable<<<gn, TPB, 0, st>>>(a);
                k_b_claim_rest<<<gn, TPB, 0, st>>>(a);
                k_b_commit<<<gn, TPB, 0, st>>>(a);
                k_b_control<<<1, 1, 0, st>>>(S.ctrl.p);
                k_b_revert<<<gn, TPB, 0, st>>>(a);
                k_b_clear<<<gc, TPB, 0, st>>>(a, XYc);
                k_b_base<<<256, TPB, 0, st>>>(a);
                k_b_next<<<gm, TPB, 0, st>>>(a, XYc);
                k_b_advance<<<1, 1, 0, st>>>(S.ctrl.p, n);
            }
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(h_ctrl, S.ctrl.p, sizeof(int) * C_WORDS, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (h_ctrl[C_DONE]) break;
        }
    } catch (...) {
        (void)hipHostFree(h_ctrl);
        throw;
    }
    const int nbatch = h_ctrl[C_CUR];
    HIP_OK(hipHostFree(h_ctrl));
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
