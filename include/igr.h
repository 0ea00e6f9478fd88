/* C ABI of the MI355X global-routing path (libigr.so).
 *
 * The reference exposes no library interface: its routing stage is a set of
 * C++ functions over process-global __managed__ state, called from route()
 * (src/main.cpp:22-73).  This ABI replaces that internal boundary with an
 * explicit context on one GPU.  Conventions: inputs are host pointers that are
 * copied (the caller keeps ownership), outputs go to caller-allocated host
 * buffers, every call returns IGR_OK (0) or a negative status and never exits
 * the process, igr_last_error() describes the last failure of the calling
 * thread, and one host thread drives a context at a time.  Net lists
 * (order, nets) name each uploaded net at most once; a list with a repeated
 * or out-of-range id is rejected with IGR_EINVAL before any device work.
 *
 * Grid arrays use the reference cell index l*X*Y + x*Y + y over routing layers
 * (metal layer - 1), see src/database_cuda.hpp:49 and :2106.
 */
#ifndef IGR_H
#define IGR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { IGR_OK = 0, IGR_EINVAL = -1, IGR_EHIP = -2, IGR_ESTATE = -3 };
enum { IGR_FORM_PLAIN = 0, IGR_FORM_BATCHED = 1 };

typedef struct igr_ctx igr_ctx;

typedef struct igr_grid_desc {
    int L, X, Y, DIR;            /* routing layers, gcells, direction of routing layer 0 (0 = horizontal) */
    double unit_wire, unit_via;
    const double* of_weight;     /* L overflow weights */
    const int* xlen;             /* X-1 gcell edge lengths */
    const int* ylen;             /* Y-1 */
    const float* capacity;       /* L*X*Y */
} igr_grid_desc;

/* Subnet trees in DFS pre-order (node 0 is the root, a parent precedes its
 * children); node = l*X*Y + x*Y + y with l = L for Steiner nodes.  Scheduler
 * footprints: points as x*Y+y, horizontal/vertical segments as pairs of
 * x*Y+y.  Offsets have n+1 entries; segment offsets count segments. */
typedef struct igr_net_desc {
    int64_t n;
    const int64_t* node_off;
    const int* nodes;
    const int* parents;
    const int64_t* pt_off;
    const int* pts;
    const int64_t* hs_off;
    const int* hs;
    const int64_t* vs_off;
    const int* vs;
} igr_net_desc;

typedef struct igr_pass_stats {
    int batches;
    double schedule_s;            /* scheduler wall time */
    double route_s;               /* batch routing wall time (device work included) */
    int64_t nets;
} igr_pass_stats;

/* Context lifetime.  Replaces build_cuda_database's grid part
 * (src/database_cuda.hpp:2258-2361). */
int igr_create(const igr_grid_desc* desc, igr_ctx** out);
void igr_destroy(igr_ctx* ctx);
const char* igr_last_error(void);

/* Grid state (reference layout on the host side). */
int igr_set_demand(igr_ctx* ctx, const float* demand);
int igr_get_demand(igr_ctx* ctx, float* demand);
int igr_get_timestamp(igr_ctx* ctx, int* ts);
int igr_get_costs(igr_ctx* ctx, float* wcost, float* vcost, double* presum);  /* any may be NULL */

/* graph::update_cost (src/graph.hpp:140-146) */
int igr_update_cost(igr_ctx* ctx);
/* graph::compute_presum (src/graph.hpp:405-430) */
int igr_compute_presum(igr_ctx* ctx);

/* Nets of the pattern pass (the tree/DFS upload of src/Lshape_route.hpp:184-229). */
int igr_set_nets(igr_ctx* ctx, const igr_net_desc* nets);

/* generate_batches_rsmt_gpu (src/database_cuda.hpp:437-631): order is the
 * routing priority (highest first); batch_of[i] receives the batch of
 * order[i]; *nbatches the batch count.  Both forms give the same batches
 * (plain: the reference's rounds with a host check per round; batched:
 * incremental rounds with device-side control). */
int igr_schedule(igr_ctx* ctx, const int* order, int64_t n, int mode, int form, int* batch_of, int* nbatches);

/* One pattern-routing batch (src/Lshape_route.hpp:235-242): cost update,
 * prefix sums, L-shape DP of every listed net, wire and via commit. */
int igr_route_batch(igr_ctx* ctx, const int* nets, int64_t n, int form);

/* Lshape_route::Lshape_route (src/Lshape_route.hpp:155-257) after the trees:
 * schedule, then route the batches in the reference order. */
int igr_pattern_route(igr_ctx* ctx, const int* order, int64_t n, int form, igr_pass_stats* stats);

/* Detour DAGs of the nets of one rip-up-and-reroute pass, one per listed net
 * in list order, nodes in topological order (parents first): node i of a net
 * has parents par[par_off[i] .. par_off[i+1]) (indices local to the net,
 * par_off global over all nodes) and fills child slot slot[k] of parent
 * par[k]; child_num[i] is node i's slot count and depth[i] its longest-path
 * depth.  pts are the scheduler points of each net.  node_off, par_off and
 * pt_off start at 0. */
typedef struct igr_dag_desc {
    int64_t n;
    const int64_t* node_off;
    const int* nodes;
    const int64_t* par_off;
    const int* par;
    const int* slot;
    const int* child_num;
    const int* depth;
    const int64_t* pt_off;
    const int* pts;
} igr_dag_desc;

/* One rip-up-and-reroute pass (Lshape_route_detour::Lshape_route_detour_wrap,
 * src/Lshape_route_detour.hpp:726-1112): 'nets' in routing priority order
 * (highest first); schedules with the DAG points, then per batch (last
 * first) rips up the nets' current routes, updates costs and prefix sums,
 * routes the DAGs and commits the new routes, which replace the old ones. */
int igr_dag_pass(igr_ctx* ctx, const int* nets, int64_t n, int mode, int form, const igr_dag_desc* dags,
                 igr_pass_stats* stats);

/* Fills the DAG structure of a description whose n, pt_off and pts are set
 * (node_off, nodes, par_off, par, slot, child_num, depth); returns IGR_OK or
 * a non-zero status.  The arrays must stay valid until the pass returns. */
typedef int (*igr_dag_complete_fn)(void* user, igr_dag_desc* dags);

/* igr_dag_pass in the batched form with the DAGs' structure built while the
 * scheduler runs: 'points' holds n, pt_off and pts (the scheduler's input);
 * complete(user, d) is called once, on a helper thread, with a copy of it and
 * fills the rest.  Routes are those of igr_dag_pass with the full
 * description.  No counterpart in the reference, whose pass
 * (src/Lshape_route_detour.hpp:726-1112) starts once the DAGs are built. */
int igr_dag_pass_overlapped(igr_ctx* ctx, const int* nets, int64_t n, int mode, const igr_dag_desc* points,
                            igr_dag_complete_fn complete, void* user, igr_pass_stats* stats);

#define IGR_MAX_RRR 8

typedef struct igr_run_opts {
    int form;                     /* IGR_FORM_PLAIN or IGR_FORM_BATCHED */
    int rrr_iterations;           /* rip-up-and-reroute passes after the pattern pass */
    int threads;                  /* host threads for the detour DAGs */
} igr_run_opts;

typedef struct igr_run_stats {
    double upload_s, pattern_s, pattern_schedule_s, total_s;
    int pattern_batches, iterations;
    int64_t subnets;
    int64_t ripped[IGR_MAX_RRR];
    int rrr_batches[IGR_MAX_RRR];
    int64_t rrr_dag_nodes[IGR_MAX_RRR];
    double rrr_view_s[IGR_MAX_RRR], rrr_dag_build_s[IGR_MAX_RRR], rrr_route_s[IGR_MAX_RRR];
    double rrr_schedule_s[IGR_MAX_RRR], rrr_prep_s[IGR_MAX_RRR];
    double rrr_prep_parts[IGR_MAX_RRR][4];   /* host fill, batch metadata, upload, allocation */
    /* batched form: the DAG structure, built on the pass's helper thread
     * while the scheduler runs (rrr_dag_build_s then times the points) */
    double rrr_dag_rest_s[IGR_MAX_RRR];
} igr_run_stats;

struct igrh_design;
/* The routing stage end to end (src/main.cpp:22-73) on a design whose trees
 * igrh_build has prepared (include/igr_host.h): nets upload, pattern pass,
 * then rip-up and reroute passes until 'rrr_iterations' or no net overflows.
 * Routes stay in the context (igr_get_routes). */
int igr_run(igr_ctx* ctx, struct igrh_design* design, const igr_run_opts* opts, igr_run_stats* stats);

/* Routes: r[0] = next free slot (1 = empty), then pairs of cell indices. */
int igr_get_route(igr_ctx* ctx, int net, int* buf, int cap, int* len);
/* All routes packed: off[n+1] (host) where off[i+1]-off[i] is net i's slot
 * capacity, data with data_cap >= igr_routes_capacity() entries. */
int igr_get_routes(igr_ctx* ctx, int64_t* off, int* data, int64_t data_cap);

/* GPU evaluator: connectivity check and FP64 score (wirelength, vias,
 * overflow) with the definitions of run/evaluator.cpp:458-633,635-675, computed
 * from routes in device memory instead of a written and parsed route file
 * (the reference scores with the separate evaluator program,
 * run/evaluator.cpp:279-301).  The route file's lines are rebuilt as the
 * writer makes them (src/graph.hpp:179-235): per net its metal-1 pin vias,
 * its stacked-pin via pairs and its subnets' routes one layer up.  Unlike the
 * reference program, an open net is counted instead of ending the run.
 * route_off / routes: routes in igr_get_routes' layout, or NULL for the
 * routes the context holds.  Metal layers are routing layers + 1; cells
 * z*X*Y + x*Y + y. */
typedef struct igr_eval_nets {
    int64_t nets;
    const int64_t* sub_off;      /* nets+1: each net's subnets (context net indices) */
    const int* subs;
    const int64_t* pin_off;      /* nets+1: selected pin cells; a cell on metal 1 gets a via to metal 2 */
    const int* pins;
    const int64_t* extra_off;    /* nets+1: stacked-pin via pairs (metal-layer cells) */
    const int* extra;
    const int64_t* dpin_off;     /* nets+1: design pins */
    const int64_t* ap_off;       /* design pins+1: access points (metal-layer cells) */
    const int* ap;
    const int* dir;              /* per metal layer: 0 horizontal, 1 vertical */
    const double* of_weight;     /* per metal layer */
    const double* capacity;      /* metal layers x X x Y */
} igr_eval_nets;
typedef struct igr_score {
    double wirelength, via, overflow, total;
    int64_t opens;
} igr_score;
int igr_evaluate(igr_ctx* ctx, const igr_eval_nets* nets, const int64_t* route_off, const int* routes,
                 igr_score* out);
int64_t igr_routes_capacity(igr_ctx* ctx);

/* graph::ripup's marking (src/graph.hpp:588-601): is_of[net] = 1 when more
 * than 10 of the net's wire edges have demand >= capacity + threshold. */
int igr_ripup_mark(igr_ctx* ctx, int threshold, uint8_t* is_of);
/* graph::extract_congestionView{,_xsum,_ysum} (src/graph.hpp:545-584):
 * X*Y prefix views indexed x*Y + y. */
int igr_congestion_view(igr_ctx* ctx, float* xsum, float* ysum);
/* No counterpart in the reference: igr_congestion_view for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same outputs; its device work
 * uses a static scratch array of 2 * IGR_NOALLOC_MAX_GCELLS floats (64 MiB)
 * per GPU instead of allocating, so grids of more than IGR_NOALLOC_MAX_GCELLS
 * gcells (X*Y) are rejected with IGR_EINVAL.  All calls on one GPU share that
 * array: calls that overlap on one GPU race. */
#define IGR_NOALLOC_MAX_GCELLS (1 << 23)
int igr_congestion_view_noalloc(igr_ctx* ctx, float* xsum, float* ysum);

/* Device time per kernel class, measured with events on the context's stream
 * while enabled (enabling clears earlier records).  Classes: 0 cost update,
 * 1 prefix sums, 2 L-shape DP, 3 wire commit, 4 via commit, 5 DAG edge
 * layer changes, 6 DAG DP, 7 DAG extraction, 8 scheduler, 9 fused track
 * update, 10 rip-up.  ms and launches have IGR_KCLASSES entries. */
#define IGR_KCLASSES 11
int igr_set_profiling(igr_ctx* ctx, int on);
int igr_kernel_times(igr_ctx* ctx, double* ms, int64_t* launches);
/* Algorithmic HBM bytes moved by the launches recorded since profiling was
 * enabled, per kernel class (IGR_KCLASSES entries; 0 where no byte model is
 * defined).  The per-unit figures are stated beside each launcher; the
 * scheduler's release-driven rounds count the bytes their kernels read and
 * write on the device. */
int igr_kernel_bytes(igr_ctx* ctx, double* bytes);

/* Device memory of the context's GPU: bytes free and total (hipMemGetInfo on
 * that device).  No counterpart in the reference; for sizing designs. */
int igr_device_memory(igr_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* Launch-shape choices of the batched form; 0 in a field keeps the built-in
 * choice.  No counterpart in the reference (its launch shapes are fixed);
 * the autotuner (instantgr_amd/autotune.py) measures and sets them.  Results
 * do not depend on any of them. */
typedef struct igr_tuning {
    int graph;          /* 1: each pass's batch sequence captured and launched as one hipGraph; 2: stream launches */
    int sched_group;    /* lanes per net in the scheduler's claim and viability kernels: 4, 8 or 16 */
    int sched_chunk;    /* scheduler rounds enqueued between host checks, 1..256 */
    int dag_order;      /* workgroup order of a DAG batch: 1 deepest DAG first, 2 routing order */
    int dag_waves;      /* DAG kernel register budget: 2 (compiler's choice) or 4 waves per SIMD */
    int track_k;        /* fused track kernel: cells per thread, 4, 8 or 16 */
    int dag_edges;      /* layer-change costs: 1 a wave per edge with LDS staging, 2 a thread per layer pair */
    int dag_seq;        /* DAG levels with at least this many nodes run a node per lane, narrower ones 16
                           lanes per node (1 << 30: always 16 lanes) */
    int sched_kind;     /* scheduler rounds after the first of a batch: 1 active-list (work follows the
                           active nets), 2 release-driven (work follows the released cells); built-in
                           choice 2 from 1.5M scheduled nets */
    int presum_k;       /* plain form's prefix-sum kernel: cells per thread, 4, 8 or 16 */
    int dag_grid;       /* DAG kernel: persistent workgroups pulling a batch's nets from its work queue */
    int lshape_groups;  /* pattern DP: nets (16-lane groups) per workgroup, 1, 2 or 4; the workgroup's LDS
                           slab holds their layer-range tables (0: 4, fewer where the tables need more room) */
} igr_tuning;
int igr_set_tuning(igr_ctx* ctx, const igr_tuning* t);
/* The values in force (built-in choices resolved, per-pass choices as 0). */
int igr_get_tuning(igr_ctx* ctx, igr_tuning* t);

/* Serving loop: a routed design kept resident on the device takes reroute
 * requests (lists of subnets) as they arrive.  No counterpart in the
 * reference, whose route() runs once per process (src/main.cpp:22-73); each
 * request is a rip-up-and-reroute pass (Lshape_route_detour_wrap,
 * src/Lshape_route_detour.hpp:726-1112) over its subnets, with the detour
 * DAGs built from the congestion view at the start of its tick.
 *
 * A worker thread owns the context while the server is open (the caller
 * makes no other call on that context until igr_server_close).  Each tick it
 * admits waiting requests in arrival order within the budgets below, passing
 * over any request whose footprint (current routes, DAG nodes and points of
 * its subnets, on 16x16-gcell tiles) meets an earlier waiting or admitted
 * request's footprint, and routes the admitted requests' subnets in one
 * batched pass.  Requests admitted together therefore touch disjoint cells. */
typedef struct igr_serve_opts {
    int threads;              /* host threads for the DAG builder (0: 8) */
    int64_t max_tick_nets;    /* subnets per tick (0: no limit; the first waiting request is always admitted) */
    int64_t scratch_bytes;    /* estimated device scratch of the DAG tables per tick (0: no limit) */
    int hold_ms;              /* a tick starts this long after its first request arrived, gathering
                                 later arrivals into it (0: at once) */
    int overtake;             /* 0: admission stops at the first request that cannot join the tick (the
                                 ticks keep arrival order); 1: later requests with disjoint footprints
                                 may pass a waiting one */
} igr_serve_opts;

typedef struct igr_request_result {
    int64_t id, nets;
    double submit_s, start_s, done_s;   /* steady-clock seconds */
    double latency_s;                   /* done - submit */
    int tick;                           /* tick that routed the request */
    int64_t tick_nets;                  /* subnets routed in that tick, all requests together */
} igr_request_result;

typedef struct igr_server_stats {
    int64_t ticks, requests, nets, max_tick_nets;
    double prep_s, route_s;             /* host preparation (view, DAGs, admission) and pass time, summed */
} igr_server_stats;

typedef struct igr_server igr_server;
int igr_server_open(igr_ctx* ctx, struct igrh_design* design, const igr_serve_opts* opts, igr_server** out);
/* Thread-safe; copies the list. */
int igr_server_submit(igr_server* srv, const int* nets, int64_t n, int64_t* id);
/* Blocks until the request is routed; a request is waited for once. */
int igr_server_wait(igr_server* srv, int64_t id, igr_request_result* out);
int igr_server_stats_get(igr_server* srv, igr_server_stats* out);
/* Routes every waiting request, then stops the worker. */
int igr_server_close(igr_server* srv);
/* Message of the last failed igr_server_* call on the calling thread. */
const char* igr_server_error(void);

/* Current commit timestamp counter. */
int igr_stamp(igr_ctx* ctx);

/* Milliseconds per full-grid cost update plus prefix sums (average of reps
 * sweeps after a warm-up), on the context's current demand: the plain form's
 * kernels (update_wcost, update_vcost, compute_presum; src/graph.hpp:40-104,
 * 405-430) or the batched form's fused track kernel with every track dirty.
 * The algorithmic bytes of one sweep are 28 per cell (cost update 16: read
 * capacity and demand, write wire and via cost; prefix sums 12: read wire
 * cost, write the FP64 sums).  No counterpart in the reference; a
 * measurement entry point. */
int igr_cost_sweep(igr_ctx* ctx, int form, int reps, double* ms);

/* Bandwidth reference: a plain 16-byte-vector device copy of 'bytes',
 * repeated 'reps' times; *gbs = (bytes read + written) / time.  Also the
 * calibration kernel for the memory traffic counters. */
int igr_copy_probe(int64_t bytes, int reps, double* gbs);

/* The routing exp evaluated on the device (diagnostic). */
int igr_expf(const float* in, float* out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
