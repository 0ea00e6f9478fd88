/* CPU oracle for the InstantGR global-routing hot path.
 *
 * TEST INFRASTRUCTURE: a sequential C++ restatement of the reference's GPU
 * routing logic, used only by tests/, the smoke check and the CPU baseline leg
 * of bench.py as the checker.  It is never part of the product path.
 *
 * Parity status of each function is listed at the top of igr_oracle.cpp.
 * Conventions: cells are indexed l*X*Y + x*Y + y (src/database_cuda.hpp:49);
 * routes are stored per net as the reference does: buf[0] = next free index
 * (starts at 1), then pairs of cell indices (src/Lshape_route.hpp:99-151).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int L, X, Y, DIR;          /* routing layers, gcells, direction of layer 0 */
    double unit_wire, unit_via;
    const double* w;           /* overflow weight per routing layer (L) */
    const int* xlen;           /* X-1 */
    const int* ylen;           /* Y-1 */
    const float* cap;          /* L*X*Y */
} orc_grid;

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const orc_grid* g);
void orc_destroy(orc_ctx* c);

float orc_expf_scalar(float x);
void orc_expf_many(const float* in, float* out, int64_t n);

/* Grid state. */
void orc_set_demand(orc_ctx* c, const float* demand);
void orc_get_demand(const orc_ctx* c, float* demand);
void orc_get_timestamp(const orc_ctx* c, int* ts);
void orc_update_cost(orc_ctx* c);                 /* src/graph.hpp:140-146 */
void orc_compute_presum(orc_ctx* c);              /* src/graph.hpp:405-430 */
void orc_get_costs(const orc_ctx* c, float* wcost, float* vcost, double* presum);

/* Nets: trees in DFS pre-order (node 0 = root, parents precede children),
 * node = l*X*Y + x*Y + y with l = L for Steiner nodes; scheduler footprints
 * (points as x*Y+y, h/v segments as pairs of x*Y+y). */
int orc_set_nets(orc_ctx* c, int64_t n,
                 const int64_t* node_off, const int* nodes, const int* parents,
                 const int64_t* pt_off, const int* pts,
                 const int64_t* hs_off, const int* hs,
                 const int64_t* vs_off, const int* vs);

/* Deterministic batch scheduler (src/database_cuda.hpp:437-631).  'order' is
 * the routing priority order (highest first); on return batch_of[i] is the
 * batch index of order[i]. Returns the number of batches. */
int orc_schedule(orc_ctx* c, const int* order, int64_t n, int mode, int* batch_of);

/* One pattern-routing batch (src/Lshape_route.hpp:235-242). */
void orc_lshape_batch(orc_ctx* c, const int* nets, int64_t n);
/* Whole pattern pass: schedule + batches in the reference order. */
int orc_pattern_route(orc_ctx* c, const int* order, int64_t n);

int orc_get_route(const orc_ctx* c, int net, int* buf, int cap);

/* Replaces demand, timestamps, the stamp counter and (route_off non-NULL)
 * every net's route (route_off[i] indexes routes; routes[route_off[i]] is the
 * route's next free index).  Returns -1 on a route longer than its slot. */
int orc_set_state(orc_ctx* c, const float* demand, const int* ts, int stamp, const int64_t* route_off,
                  const int* routes);
/* Commits n given routes as one batch (K = 1: wire differences,
 * batch_wire_update, via commit; src/Lshape_route.hpp:240-241) or rips them
 * up (K = -1: commit_wire_demand then commit_via_demand;
 * src/Lshape_route_detour.hpp:1047-1048), with a new stamp. */
void orc_commit_routes(orc_ctx* c, int64_t n, const int64_t* route_off, const int* routes, int K);
int orc_stamp(const orc_ctx* c);

/* Rip-up marking (src/graph.hpp:432-466,588-601): is_of[net] = 1 when more
 * than 10 full edges lie under the net's wires. */
void orc_ripup_mark(orc_ctx* c, int threshold, uint8_t* is_of);
/* Congestion view (src/graph.hpp:545-584) with zeroed start. */
void orc_congestion_view(orc_ctx* c, float* xsum, float* ysum);

/* One rip-up-and-reroute pass over detour DAGs (src/Lshape_route_detour.hpp:
 * 726-1112): 'nets' in routing priority order (highest first); DAG i belongs
 * to nets[i] (node_off/pt_off per listed net, par_off global over nodes, par
 * local).  Schedules with the DAG points, then per batch (last first) rips up
 * the nets' current routes, recomputes costs, routes the DAGs and commits.
 * Returns the number of batches. */
int orc_dag_pass(orc_ctx* c, const int* nets, int64_t n, int mode, const int64_t* node_off, const int* node,
                 const int64_t* par_off, const int* par, const int* slot, const int* child_num, const int* depth,
                 const int64_t* pt_off, const int* pt);

/* Single-net L-shape DP exposed for brute-force checks: routes the tree of
 * net 'net' against the current costs without committing; returns the optimal
 * root cost and writes the route. */
double orc_lshape_probe(orc_ctx* c, int net, int* buf, int cap);

#ifdef __cplusplus
}
#endif
