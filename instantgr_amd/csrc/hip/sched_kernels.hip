// Conflict-free net batch scheduler (src/database_cuda.hpp:288-434,437-631).
//
// Nets compete for the cells of an X*Y ownership map by atomicMax of their
// priority index (larger hpwl = larger index wins); a net joins the current
// batch when it owns every one of its point cells.  Up to 30 rounds per batch;
// nets blocked by a committed net sit out the rest of the batch.
//
// The reference re-reads a point cell right after its own atomicMax and
// abandons the net if a higher one owns it, racing with concurrent claims
// (src/database_cuda.hpp:313-318).  Here that read is given barrier
// semantics: a round first claims every active net's point cells (kernel 1),
// then decides which nets still own all their point cells (kernel 2), and only
// those claim their neighbour cells and sampled segment cells (kernel 3).
// Every result is a pure function of the input order, identical to the CPU
// oracle's.
#include "igr_ctx.hpp"

#include <chrono>

#include <algorithm>

namespace igr_dev {

namespace {

constexpr int TPB = 256;

struct SchedArgs {
    int64_t n;
    const int* ids;              // scheduler index -> net id (segments)
    const int* pids;             // scheduler index -> row of the point arrays
    const int64_t* pt_off;
    const int* pt;
    const int64_t* hs_off;
    const int* hs;
    const int64_t* vs_off;
    const int* vs;
    int* map;                    // X*Y owner index or -1
    int* batch;                  // per scheduler index
    int* status;                 // 0 normal, 1 competing, 2 blocked for this batch
    int* viable;
    int* flags;                  // [0] = committed this round
    int X, Y, sparsity;
};

__device__ __forceinline__ bool active(const SchedArgs& a, int64_t i) { return a.status[i] != 2 && a.batch[i] == -1; }

__global__ void __launch_bounds__(TPB) k_claim_points(SchedArgs a) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !active(a, i)) return;
    a.status[i] = 1;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) atomicMax(&a.map[a.pt[k]], (int)i);
}

__global__ void __launch_bounds__(TPB) k_viable(SchedArgs a) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n) return;
    int ok = 0;
    if (active(a, i)) {
        ok = 1;
        for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++)
            if (a.map[a.pt[k]] != (int)i) { ok = 0; break; }
    }
    a.viable[i] = ok;
}

__global__ void __launch_bounds__(TPB) k_claim_rest(SchedArgs a) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !a.viable[i]) return;
    const int net = a.ids[i];
    const int X = a.X, Y = a.Y;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
        int p = a.pt[k], x = p / Y, y = p % Y;
        if (y - 1 >= 0) atomicMax(&a.map[x * Y + y - 1], (int)i);
        if (x - 1 >= 0) atomicMax(&a.map[(x - 1) * Y + y], (int)i);
        if (x + 1 < X) atomicMax(&a.map[(x + 1) * Y + y], (int)i);
    }
    for (int64_t k = a.hs_off[net]; k < a.hs_off[net + 1]; k++) {
        int s = a.hs[2 * k], e = a.hs[2 * k + 1], y = s % Y;
        for (int x = s / Y; x <= e / Y; x += a.sparsity) atomicMax(&a.map[x * Y + y], (int)i);
    }
    for (int64_t k = a.vs_off[net]; k < a.vs_off[net + 1]; k++) {
        int s = a.vs[2 * k], e = a.vs[2 * k + 1], x = s / Y;
        for (int y = s % Y; y <= e % Y; y += a.sparsity) atomicMax(&a.map[x * Y + y], (int)i);
    }
}

__global__ void __launch_bounds__(TPB) k_commit(SchedArgs a, int cur) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || !active(a, i)) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++)
        if (a.map[a.pt[k]] != (int)i) return;
    a.batch[i] = cur;
    a.flags[0] = 1;
}

__global__ void __launch_bounds__(TPB) k_revert(SchedArgs a) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= a.n || a.status[i] != 1 || a.batch[i] != -1) return;
    for (int64_t k = a.pt_off[a.pids[i]]; k < a.pt_off[a.pids[i] + 1]; k++) {
        int pre = a.map[a.pt[k]];
        if (pre != -1 && a.batch[pre] != -1) { a.status[i] = 2; return; }
    }
}

__global__ void __launch_bounds__(TPB) k_clear(SchedArgs a) {
    int64_t q = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= (int64_t)a.X * a.Y) return;
    int m = a.map[q];
    if (m >= 0 && a.status[m] == 2) a.map[q] = -1;
}

__global__ void k_fill(int* p, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

} // namespace

int run_schedule(igr_ctx* c, const int* h_order, int64_t n, int mode, int* h_batch_of, const int64_t* d_pt_off,
                 const int* d_pt) {
    if (n == 0) return 0;
    Sched& S = c->sched;
    NetsDev& N = c->nets;
    hipStream_t st = c->stream;
    std::vector<int> ids(n);
    for (int64_t i = 0; i < n; i++) ids[i] = h_order[n - 1 - i];
    S.ids.upload(ids.data(), n, st);
    std::vector<int> pids;
    if (d_pt_off) {
        pids.resize(n);
        for (int64_t i = 0; i < n; i++) pids[i] = (int)(n - 1 - i);
        S.pids.upload(pids.data(), n, st);
    }
    const int64_t XYc = (int64_t)c->g.X * c->g.Y;
    S.map.alloc(XYc);
    S.batch.alloc(n);
    S.status.alloc(n);
    S.viable.alloc(n);
    S.flags.alloc(4);
    SchedArgs a{n, S.ids.p, d_pt_off ? S.pids.p : S.ids.p, d_pt_off ? d_pt_off : N.pt_off.p, d_pt_off ? d_pt : N.pt.p, N.hs_off.p, N.hs.p, N.vs_off.p, N.vs.p,
                S.map.p, S.batch.p, S.status.p, S.viable.p, S.flags.p, c->g.X, c->g.Y,
                (mode == 1 || mode == 2) ? 2 : 15};
    k_fill<<<blocks(n), TPB, 0, st>>>(S.batch.p, n, -1);
    int cur = 0;
    int* h_flag = nullptr;
    HIP_OK(hipHostMalloc(&h_flag, sizeof(int)));
    const auto t_start = std::chrono::steady_clock::now();
    auto t_note = t_start;
    try {
        for (;;) {
            const auto t_now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double>(t_now - t_note).count() > 20) {
                t_note = t_now;
                progress("scheduler: %d batches after %.1f s", cur,
                         std::chrono::duration<double>(t_now - t_start).count());
            }
            k_fill<<<blocks(XYc), TPB, 0, st>>>(S.map.p, XYc, -1);
            k_fill<<<blocks(n), TPB, 0, st>>>(S.status.p, n, 0);
            bool have_commit = false;
            for (int round = 0; round < 30; round++) {
                k_fill<<<1, TPB, 0, st>>>(S.flags.p, 1, 0);
                k_claim_points<<<blocks(n), TPB, 0, st>>>(a);
                k_viable<<<blocks(n), TPB, 0, st>>>(a);
                k_claim_rest<<<blocks(n), TPB, 0, st>>>(a);
                k_commit<<<blocks(n), TPB, 0, st>>>(a, cur);
                HIP_OK(hipGetLastError());
                HIP_OK(hipMemcpyAsync(h_flag, S.flags.p, sizeof(int), hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                if (!*h_flag) break;
                have_commit = true;
                k_revert<<<blocks(n), TPB, 0, st>>>(a);
                k_clear<<<blocks(XYc), TPB, 0, st>>>(a);
            }
            if (!have_commit) break;
            cur++;
        }
    } catch (...) {
        (void)hipHostFree(h_flag);
        throw;
    }
    HIP_OK(hipHostFree(h_flag));
    std::vector<int> b(n);
    HIP_OK(hipMemcpyAsync(b.data(), S.batch.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; i++) {
        if (b[i] < 0) throw std::runtime_error("scheduler left a net without a batch");
        h_batch_of[n - 1 - i] = b[i];
    }
    return cur;
}

} // namespace igr_dev
