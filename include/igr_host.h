/* Host library of the MI355X global-routing path (libigr_host.so).
 *
 * Pure host code: design generation and the .cap/.net formats, access-point
 * selection, subnets and Steiner trees, routing inputs, the route file writer
 * and the evaluator.  Every entry point returns 0 (or a non-null handle) on
 * success; on failure igrh_error() describes the error of the calling thread.
 *
 * Reference counterparts:
 *   igrh_load      db::read (src/database.hpp:247-415) and db::net::init
 *                  (src/database.hpp:38-243, run through igrh_build)
 *   igrh_build     build_cuda_database's host part (src/database_cuda.hpp:
 *                  2096-2256) and the tree/DFS preparation of
 *                  Lshape_route (src/Lshape_route.hpp:155-208)
 *   igrh_write_routes  graph::output_nets (src/graph.hpp:179-235)
 *   igrh_evaluate  the evaluator program (run/evaluator.cpp:458-633)
 */
#ifndef IGR_HOST_H
#define IGR_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct igrh_design igrh_design;

typedef struct igrh_gen_params {
    int layers, xsize, ysize;
    int64_t nets;
    int pin_cap;
    double geom_p, sigma, blockage_frac, heavy_frac;
    int heavy_min, heavy_max, heavy_window, hotspots;
    double hotspot_scale;
    uint64_t seed;
} igrh_gen_params;

typedef struct igrh_grid_info {
    int L, X, Y, DIR;              /* routing layers (metal layers - 1) */
    double unit_wire, unit_via;
    const double* w;               /* per routing layer overflow weight */
    const int* xlen;
    const int* ylen;
    const float* cap;              /* L*X*Y, index l*X*Y + x*Y + y */
    int metal_layers;
    int64_t nets, subnets;
    int max_nodes;
} igrh_grid_info;

typedef struct igrh_score {
    double wirelength, via, overflow, total;
    int64_t opens;
} igrh_score;

const char* igrh_error(void);

/* Benchmark configurations 0..4 (nets_override/seed_override 0 = keep). */
igrh_design* igrh_generate(int preset, int64_t nets_override, uint64_t seed_override);
igrh_design* igrh_generate_custom(const igrh_gen_params* p);
igrh_design* igrh_load(const char* cap_path, const char* net_path);
int igrh_save(const igrh_design* h, const char* cap_path, const char* net_path);
void igrh_free(igrh_design* h);

/* Access points, routing grid, subnets, trees and routing inputs. */
int igrh_build(igrh_design* h, int threads);
double igrh_build_seconds(const igrh_design* h);
int igrh_grid(const igrh_design* h, igrh_grid_info* out);

/* Named host arrays (node_off, node, parent, pt_off, pt, hs_off, hs, vs_off,
 * vs, order, hpwl, ...): returns the element count and sets *ptr. */
int64_t igrh_array(const igrh_design* h, const char* name, const void** ptr);

/* Detour DAGs of the listed subnets for rip-up and reroute
 * (net::generate_detours_reconstruct, src/database_cuda.hpp:1429-2086) from
 * the X*Y congestion prefix views; results in the dag_* arrays, one DAG per
 * listed net in list order (dag_node_off, dag_node, dag_par_off (global over
 * nodes), dag_par (local indices), dag_slot, dag_child_num, dag_depth,
 * dag_pt_off, dag_pt). */
int igrh_build_dags(igrh_design* h, const int* nets, int64_t n, const float* xsum, const float* ysum, int mode,
                    int threads);

/* igrh_build_dags in two steps, so that the scheduler can start on the points
 * while the rest is built (no counterpart in the reference, which builds the
 * DAGs whole before its pass).  IGRH_DAG_POINTS writes only dag_pt_off and
 * dag_pt; IGRH_DAG_STRUCTURE, called next with the same nets and views,
 * writes the other dag_* arrays and leaves the points as they are (it fails
 * unless the points of as many nets precede it); IGRH_DAG_ALL is
 * igrh_build_dags.  The two steps give the arrays of one full build. */
#define IGRH_DAG_ALL 0
#define IGRH_DAG_POINTS 1
#define IGRH_DAG_STRUCTURE 2
int igrh_build_dags_part(igrh_design* h, const int* nets, int64_t n, const float* xsum, const float* ysum, int mode,
                         int threads, int part);

/* Rectilinear Steiner tree of n distinct gcells (x[i], y[i]) on a grid of
 * xsize*ysize gcells with edge lengths xlen (xsize-1 entries) and ylen
 * (ysize-1): the tree the subnet builder gives a net with these pins, in
 * place of flute::flute (src/flute.hpp:1239-1350) as my_flute calls it
 * (src/database_cuda.hpp:696-744).  Writes the nodes as (x, y) pairs into
 * nodes_xy (room for node_cap nodes; a tree of n pins has fewer than 2n) and
 * the *n_nodes - 1 edges as node index pairs into edges. */
int igrh_steiner(int n, const int* x, const int* y, int xsize, int ysize, const int* xlen, const int* ylen,
                 int node_cap, int* nodes_xy, int* edges, int* n_nodes);

/* Routes per subnet: off[s] indexes data; data[off[s]] = next free slot,
 * then cell index pairs (routing layers). */
int igrh_write_routes(const igrh_design* h, const char* out_path, const int64_t* off, const int* data,
                      int64_t data_len);
int igrh_evaluate(const igrh_design* h, const char* out_path, igrh_score* out);

#ifdef __cplusplus
}
namespace igr {
struct Design;
struct NetPins;
struct Grid;
struct Subnets;
struct PatternInputs;
}
const igr::Design* igrh_design_ref(const igrh_design* h);
const igr::NetPins* igrh_netpins_ref(const igrh_design* h);
const igr::Grid* igrh_grid_ref(const igrh_design* h);
const igr::Subnets* igrh_subnets_ref(const igrh_design* h);
const igr::PatternInputs* igrh_inputs_ref(const igrh_design* h);
namespace igr { struct DagSet; }
const igr::DagSet* igrh_dags_ref(const igrh_design* h);
#endif
#endif
