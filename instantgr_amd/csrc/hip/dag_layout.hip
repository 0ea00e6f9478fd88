// Device layout of one rip-up-and-reroute pass's DAGs (batched form).
//
// The host uploads the DAGs as the builder produced them (net list order,
// local parent indices) and, once the scheduler has fixed the routing order,
// the per-net bases of that order.  One 64-lane workgroup per routed net then
// writes the net's nodes, parent edges and child-slot offsets at their global
// positions, groups its nodes and parent edges by depth level (the order inside
// a level is free: the DAG kernels combine a level's nodes with atomic minima
// over keys that carry the node id, and emit routes per node), and checks the
// DAG's structure (node range, slot counts, a single source, parents before
// children and at a smaller depth, valid slots), raising *err on a violation.
#include "igr_ctx.hpp"

namespace igr_dev {

namespace {

constexpr int WG = 64;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__global__ void __launch_bounds__(WG) k_dag_layout(DagLayoutArgs a) {
    const int64_t r = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t i = a.rorder[r];
    const int64_t b0 = a.node_off[i], nn = a.node_off[i + 1] - b0;
    const int64_t e0 = a.par_off[b0];
    const int64_t g0 = a.nbase[r], ge0 = a.ebase[r], gs0 = a.sbase[r];
    const int nl = (int)(a.lbase[r + 1] - a.lbase[r]);
    const int64_t lo = a.lbase[r] + r;   // nl + 1 level boundaries per net
    for (int d = lane; d <= nl; d += WG) {
        a.cur_n[lo + d] = 0;
        a.cur_e[lo + d] = 0;
    }
    __syncthreads();
    bool bad = false;
    int64_t carry = 0;
    for (int64_t base = 0; base < nn; base += WG) {
        const int64_t k = base + lane;
        const bool valid = k < nn;
        const int cn = valid ? a.child_num[b0 + k] : 0;
        const int64_t incl = wave_incl_scan(cn);
        if (valid) {
            const int v = a.nodes[b0 + k];
            bad |= v < 0 || v >= a.node_limit || cn < 0 || cn > 10;
            a.node[g0 + k] = v;
            a.o_child_num[g0 + k] = cn;
            a.loc[g0 + k] = (int)k;
            a.csum[g0 + k] = gs0 + carry + incl - cn;
            const int64_t pb = a.par_off[b0 + k], pe = a.par_off[b0 + k + 1];
            bad |= (k == 0) != (pe == pb);
            a.o_par_off[g0 + k] = ge0 + (pb - e0);
            for (int64_t q = pb; q < pe; q++) {
                const int p = a.par[q];
                const int sl = a.slot[q];
                if (p < 0 || p >= k) {
                    bad = true;
                } else if (sl < 0 || sl >= a.child_num[b0 + p] || a.depth[b0 + p] >= a.depth[b0 + k]) {
                    bad = true;
                }
                a.o_par[ge0 + (q - e0)] = (int)(g0 + p);
                a.o_slot[ge0 + (q - e0)] = sl;
                a.enode[ge0 + (q - e0)] = (int)(g0 + k);
            }
            const int d = a.depth[b0 + k];
            if (d < 0 || d >= nl) {
                bad = true;
            } else {
                atomicAdd((unsigned long long*)&a.cur_n[lo + d], 1ull);
                atomicAdd((unsigned long long*)&a.cur_e[lo + d], (unsigned long long)(pe - pb));
            }
        }
        carry += __shfl(incl, 63, 64);
    }
    __syncthreads();
    // level boundaries (exclusive prefix of the counts); the counters become
    // the fill cursors
    int64_t cn_carry = 0, ce_carry = 0;
    for (int base = 0; base <= nl; base += WG) {
        const int d = base + lane;
        const int64_t cn = d < nl ? a.cur_n[lo + d] : 0, ce = d < nl ? a.cur_e[lo + d] : 0;
        const int64_t in = wave_incl_scan(cn), ie = wave_incl_scan(ce);
        if (d <= nl) {
            const int64_t pn = g0 + cn_carry + in - cn, pe = ge0 + ce_carry + ie - ce;
            a.lvl_node_off[lo + d] = pn;
            a.lvl_edge_off[lo + d] = pe;
            a.cur_n[lo + d] = pn;
            a.cur_e[lo + d] = pe;
        }
        cn_carry += __shfl(in, 63, 64);
        ce_carry += __shfl(ie, 63, 64);
    }
    __syncthreads();
    for (int64_t k = lane; k < nn; k += WG) {
        const int d = a.depth[b0 + k];
        if (d < 0 || d >= nl) continue;
        const int64_t pos = (int64_t)atomicAdd((unsigned long long*)&a.cur_n[lo + d], 1ull);
        a.dnode[pos] = (int)(g0 + k);
        const int64_t pb = a.par_off[b0 + k], pe = a.par_off[b0 + k + 1];
        const int64_t ep = (int64_t)atomicAdd((unsigned long long*)&a.cur_e[lo + d], (unsigned long long)(pe - pb));
        for (int64_t t = 0; t < pe - pb; t++) a.dedge[ep + t] = ge0 + (pb - e0) + t;
    }
    if (lane == 0) {
        a.nl_off[r] = lo;
        a.nlevels[r] = nl;
    }
    if (bad) atomicOr(a.err, 1);
}

} // namespace

void launch_dag_layout(igr_ctx* c, const DagLayoutArgs& a, int64_t n) {
    if (n == 0) return;
    k_dag_layout<<<(unsigned)n, WG, 0, c->stream>>>(a);
    HIP_OK(hipGetLastError());
}

} // namespace igr_dev
