// Routing context: device grid state, net data and the launch helpers shared
// by the kernel translation units.
#pragma once
#include "igr_dev.h"
#include "igr.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#define HIP_OK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

namespace igr_dev {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) count = 1;
        HIP_OK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIP_OK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(size_t count, hipStream_t s) {
        alloc(count);
        HIP_OK(hipMemsetAsync(p, 0, count * sizeof(T), s));
    }
};

// Net trees and footprints on the device (pattern pass).  Every net's current
// route is reached through rptr[net]: pattern routes live in 'routes' at
// route_off[net]; a rip-up-and-reroute pass moves the nets it reroutes to a
// buffer of its own (dag_routes).
struct NetsDev {
    int64_t n = 0;
    DevBuf<int64_t> node_off, route_off, pt_off, hs_off, vs_off;
    DevBuf<int> node, parent, routes, pt, hs, vs;
    DevBuf<int*> rptr;
    std::vector<int64_t> h_node_off, h_route_off;
    std::vector<int> h_node;          // host copies used by the host-side batch logic
    std::vector<int*> h_rptr;
    std::vector<int64_t> h_rcap;      // capacity of each net's current route slot
    std::vector<double> h_lbytes;     // algorithmic bytes of each net's L-shape DP
    std::vector<DevBuf<int>*> dag_routes;   // freed (null) once no net's route lives there
    std::vector<int64_t> dag_live;          // nets whose route lives in dag_routes[k]
    std::vector<int> h_rbuf;                // per net: its dag_routes buffer, -1 = pattern slot
    int max_nodes = 0;
    ~NetsDev() { clear_dag_routes(); }
    void clear_dag_routes() {
        for (auto* b : dag_routes) delete b;
        dag_routes.clear();
        dag_live.clear();
        std::fill(h_rbuf.begin(), h_rbuf.end(), -1);
    }
    // Net 'net' leaves its rip-up-and-reroute buffer; an emptied buffer is
    // freed (the caller has synchronised with every kernel that used it).
    void release(int net) {
        const int k = h_rbuf[net];
        if (k < 0) return;
        h_rbuf[net] = -1;
        if (--dag_live[k] == 0) {
            delete dag_routes[k];
            dag_routes[k] = nullptr;
        }
    }
    // The listed nets' routes now live in the newest buffer at rbase + slot[r].
    void adopt(const int* rnets, const int64_t* slot, int64_t n, int* rbase) {
        const int k = (int)dag_routes.size() - 1;
        dag_live.resize(dag_routes.size(), 0);
        for (int64_t r = 0; r < n; r++) {
            const int net = rnets[r];
            release(net);
            h_rptr[net] = rbase + slot[r];
            h_rcap[net] = slot[r + 1] - slot[r];
            h_rbuf[net] = k;
            dag_live[k]++;
        }
        if (dag_live[k] == 0) {
            delete dag_routes[k];
            dag_routes[k] = nullptr;
        }
    }
};

struct Sched {
    DevBuf<int> map, batch, status, viable, ids, pids, flags;
    // batched form: per-cell net lists, their heads and the round-0 map; the
    // compacted unassigned / active net lists; release-driven rounds: the
    // released cells, new owners, committed nets and their cells, with stamps
    DevBuf<int> base, cnt, c_net, ctrl, lists, rel, cand, cstamp, comm, hits, hstamp;
    DevBuf<int64_t> head, c_off, c_end, bsum, cur;
    DevBuf<unsigned long long> stat;
    DevBuf<int> bits;             // per-net state bits (assigned, blocked in this batch)
};

// One rip-up-and-reroute pass on the device: DAGs of the rerouted nets in
// routing order (batch by batch), with global node/edge/slot numbering.
struct DagDev {
    DevBuf<int> node, par, slot, child_num, loc, enode, node_net, lvl_nodes, nets;
    DevBuf<int64_t> node_off, par_off, csum, route_slot;
    DevBuf<int64_t> lvl_off;
    DevBuf<double> costs, cost_edges;
    DevBuf<uint64_t> packed, best;
    DevBuf<int> lo, best_change, emit;
    DevBuf<int> pt;
    DevBuf<int64_t> pt_off;
    // Batched form: depth levels of every net (nodes and their parent edges).
    DevBuf<int64_t> nl_off, lvl_node_off, lvl_edge_off, dedge;
    DevBuf<int> nlevels, dnode;
    // Batched form: the DAGs as uploaded (net list order, local parents), the
    // per-routed-net bases and the level cursors of the device layout.
    DevBuf<int64_t> r_node_off, r_par_off, ebase, sbase, lbase, cur_n, cur_e;
    DevBuf<int> r_nodes, r_par, r_slot, r_child_num, r_depth, rorder, err;
    // Batched form: the routed nets of each batch in launch order, and the
    // per-batch work-queue counters.
    DevBuf<int> glist, queue;
};

// Arguments of the device DAG layout (one workgroup per routed net).
struct DagLayoutArgs {
    const int64_t* node_off;      // uploaded DAGs, net list order
    const int* nodes;
    const int64_t* par_off;
    const int* par;
    const int* slot;
    const int* child_num;
    const int* depth;
    const int* rorder;            // routed net -> list position
    const int64_t* nbase;         // per routed net: first node, edge, slot, level
    const int64_t* ebase;
    const int64_t* sbase;
    const int64_t* lbase;
    int node_limit;               // (L + 1) * X * Y
    int* node;
    int* o_child_num;
    int* loc;
    int64_t* csum;
    int64_t* o_par_off;
    int* o_par;
    int* o_slot;
    int* enode;
    int64_t* nl_off;
    int* nlevels;
    int64_t* lvl_node_off;
    int64_t* lvl_edge_off;
    int* dnode;
    int64_t* dedge;
    int64_t* cur_n;
    int64_t* cur_e;
    int* err;
};

struct Scratch {
    DevBuf<double> dist;
    DevBuf<int> from, lr;
    DevBuf<int> net_list;
    DevBuf<int64_t> scr_off;
    DevBuf<int> of;
    DevBuf<uint8_t> is_of;
    DevBuf<float> xsum, ysum, view;
};

// Per-kernel-class device time, measured with events on the context stream
// when enabled (igr_set_profiling).
enum KClass {
    K_COST = 0, K_PRESUM, K_LSHAPE, K_WIRE_COMMIT, K_VIA_COMMIT, K_DAG_EDGES, K_DAG_DP, K_DAG_EXTRACT, K_SCHED,
    K_TRACK_FUSED, K_RIP, K_NCLASS
};

struct KProf {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<int> cls;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> rec;
    double bytes[16] = {};         // algorithmic HBM bytes per kernel class
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void reset() {
        used = 0;
        cls.clear();
        rec.clear();
        for (double& b : bytes) b = 0;
    }
    ~KProf() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

} // namespace igr_dev

struct igr_ctx {
    igr_dev::KProf prof;
    igr_dev::GridDev g{};
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // uploads that overlap the compute stream
    hipStream_t aux_stream = nullptr;    // a DAG pass's rip-up scatters beside the DAG kernels
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int device = 0;
    igr_dev::DevBuf<double> w;
    igr_dev::DevBuf<int> xlen, ylen;
    igr_dev::DevBuf<float> cap, demand, wcost, vcost, exptab;
    igr_dev::DevBuf<double> presum;
    igr_dev::DevBuf<int> pre, ts;
    // Batched form: via-cut counts, rip-up differences, dirty-track flags, the
    // bad-route flag and the fused-pass cell counters (processed, ripped).
    igr_dev::DevBuf<int> vc, rpre, rcov, rvc, dirty, bad;
    igr_dev::DevBuf<unsigned long long> track_stats;
    igr_dev::NetsDev nets;
    igr_dev::Sched sched;
    igr_dev::Scratch scratch;
    igr_dev::DagDev dag;
    int stamp = 0;
    igr_tuning tune{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // launch-shape overrides (0 = built-in)
    double last_dag_prep_s = 0;   // host preparation time of the last DAG pass
    double last_dag_prep_parts[4] = {0, 0, 0, 0};   // fill, batch metadata, upload, allocation
    int64_t cells() const { return (int64_t)g.L * g.X * g.Y; }
};

namespace igr_dev {
// Records the device time of the launches made during its lifetime and adds
// their algorithmic HBM bytes to the class total.
struct KScope {
    igr_ctx* c;
    int k;
    hipEvent_t a = nullptr;
    KScope(igr_ctx* c_, int k_, double bytes = 0) : c(c_), k(k_) {
        if (c->prof.on) {
            a = c->prof.get();
            HIP_OK(hipEventRecord(a, c->stream));
            c->prof.bytes[k] += bytes;
        }
    }
    ~KScope() {
        if (c->prof.on && a) {
            hipEvent_t b = c->prof.get();
            if (hipEventRecord(b, c->stream) == hipSuccess) {
                c->prof.cls.push_back(k);
                c->prof.rec.emplace_back(a, b);
            }
        }
    }
};

// Records msg as the calling thread's igr_last_error() and returns status.
int fail(int status, const std::string& msg);

// IGR_PROGRESS set in the environment: one line per stage on stderr, for
// long runs watched from outside.
void progress(const char* fmt, ...);

// Built-in launch-shape choices (igr_tuning fields left at 0).
constexpr int DEF_GRAPH = 1;
constexpr int DEF_SCHED_CHUNK = 16;
constexpr int DEF_DAG_ORDER = 1;
constexpr int DEF_DAG_WAVES = 2;
constexpr int DEF_DAG_SEQ = 16;
constexpr int DEF_DAG_GRID = 2048;
// Batched scheduler: release-driven rounds from this many scheduled nets,
// active-list rounds below.
constexpr int64_t DEF_SCHED_RELEASE_NETS = 1500000;

// Records the work enqueued on a stream between construction and launch()
// as one graph, then launches it; inactive, the work runs as it is enqueued.
// A capture still open when the scope unwinds is ended and dropped.
struct StreamGraph {
    hipStream_t s;
    bool on;
    explicit StreamGraph(hipStream_t s_, bool on_) : s(s_), on(on_) {
        if (on) HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    }
    void launch() {
        if (!on) return;
        on = false;
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamEndCapture(s, &g));
        hipGraphExec_t x = nullptr;
        hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(x, s);
        if (x) (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        HIP_OK(e);
    }
    ~StreamGraph() {
        if (on) {
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        }
    }
};

// Kernel launchers (defined in the .hip files).
void launch_update_cost(igr_ctx* c);
void launch_presum(igr_ctx* c);
void launch_wire_commit(igr_ctx* c, int stamp);
void launch_via_commit(igr_ctx* c, const int* d_nets, int64_t n, int stamp, int K);
void launch_lshape_plain(igr_ctx* c, const int* d_nets, const int64_t* d_scr_off, int64_t n, int stamp,
                         double bytes = 0);
// Batched form: the L-shape DP with a 16-lane group per net and the DP tables
// in LDS.  Nets of one launch have node counts in one class (lshape_wave_class:
// up to 4, 8, 16 or LSWAVE_MAX_NODES nodes; -1 above).
constexpr int LSWAVE_MAX_NODES = 24;
int lshape_wave_class(int nodes);
void launch_lshape_wave(igr_ctx* c, const int* d_nets, int64_t n, int cls, int stamp, double bytes = 0);
void launch_to_ref(igr_ctx* c, const void* src, void* dst, int elem_bytes);
void launch_from_ref(igr_ctx* c, const void* src, void* dst, int elem_bytes);
void launch_ripup_mark(igr_ctx* c, int th, uint8_t* d_is_of, int* d_of);
void launch_congestion(igr_ctx* c, float* d_xsum, float* d_ysum);
// Its parts, on caller buffers: the per-gcell view (X*Y floats), and one
// prefix view of it (axis 0 along x, 1 along y).  No allocation, no sync.
void launch_congestion_map(igr_ctx* c, float* d_v);
void launch_view_scan(igr_ctx* c, const float* d_v, float* d_out, int axis);
// Static scratch of the no-allocation entry points: looked up once per
// device (call with that device current); the error of the lookup, if any,
// is kept for those entry points to report.
void noalloc_init(int device);
// Scheduler; with d_pt_off/d_pt given, the points of order[i] are list entry i
// of those arrays instead of the pattern footprint.
int run_schedule(igr_ctx* c, const int* h_order, int64_t n, int mode, int* h_batch_of,
                 const int64_t* d_pt_off = nullptr, const int* d_pt = nullptr);
// The same schedule computed incrementally with device-side round control
// (batched form).
int run_schedule_batched(igr_ctx* c, const int* h_order, int64_t n, int mode, int* h_batch_of,
                         const int64_t* d_pt_off = nullptr, const int* d_pt = nullptr);
void launch_rip_wire(igr_ctx* c, const int* d_nets, int64_t n, int stamp);
// Plain (reference launch structure) device work of one DAG batch, after the
// rip-up and the cost/prefix update.
struct DagBatch {
    int64_t net0, net1, node0, node1, edge0, edge1, slot0, slot1;
    int levels;
    int index = 0;              // batch number in the pass (work-queue counter)
    int64_t lvl0;               // first entry of this batch in lvl_off
    double bytes_edges = 0, bytes_dp = 0;   // algorithmic HBM bytes
};
void launch_dag_batch_plain(igr_ctx* c, const DagBatch& b, int stamp, const std::vector<int64_t>& h_lvl_off);
void launch_dag_compact(igr_ctx* c, const DagBatch& b);
// Batched form: persistent workgroups pulling the batch's nets from a work
// queue, all depth levels of a net in one workgroup; the pass's per-batch
// queue counters are zeroed first.
void launch_dag_batch_wave(igr_ctx* c, const DagBatch& b, int stamp);
void launch_dag_queue_reset(igr_ctx* c, int64_t batches);
// Batched form: the device layout of a pass's DAGs.
void launch_dag_layout(igr_ctx* c, const DagLayoutArgs& a, int64_t n);
void launch_expf(const float* in, float* out, int64_t n, hipStream_t s);
// Fills the cost kernels' exp table (EXPTAB_N floats, GridDev::exptab).
void launch_exptab(float* t, hipStream_t s);
// Device-to-device copy of 'bytes' repeated 'reps' times; returns GB/s
// (bytes read + written).
double copy_probe(int64_t bytes, int reps, hipStream_t s);

// Batched-form grid state for the lifetime of one pass: clears the scatter
// arrays, flags every track dirty and points the grid at them; the plain
// kernels see null pointers outside it.
struct BatchedState {
    igr_ctx* c;
    explicit BatchedState(igr_ctx* c_);
    ~BatchedState();
    BatchedState(const BatchedState&) = delete;
    BatchedState& operator=(const BatchedState&) = delete;
    void check();   // throws if a ripped route could not be scattered
};
// Fused per-track pass: commit of the routed batch (s_commit > 0), rip-up of
// the next one (s_rip > 0), then costs and prefix sums when 'cost' is set.
void launch_track_fused(igr_ctx* c, int s_commit, int s_rip, bool cost);
double time_cost_sweep(igr_ctx* c, int form, int reps);
// Every net's route into d_out at d_off[i] (slices of the nets' route capacities).
void gather_routes(igr_ctx* c, const int64_t* d_off, int* d_out);
// The GPU evaluator (eval_kernels.hip).
void evaluate_routes_dev(igr_ctx* c, const igr_eval_nets* e, const int64_t* h_route_off, const int* h_routes,
                         igr_score* out);
void launch_rip_scatter(igr_ctx* c, const int* d_nets, int64_t n, hipStream_t s = nullptr);
} // namespace igr_dev
