// Detour DAGs for rip-up and reroute (host construction).
#pragma once
#include "design.hpp"

#include <memory>
#include <utility>
#include <vector>

namespace igr {

// One net's DAG in topological order (parents before children): node i has
// parents par[par_off[i] .. par_off[i+1]) (local indices) and occupies child
// slot slot[k] of parent par[k]; child_num[i] is its number of child slots,
// depth[i] its longest-path depth from the source (node 0).  Nodes are
// l*X*Y + x*Y + y with l = L for nodes that are not pins.
struct DagNet {
    std::vector<int> node, par_off, par, slot, child_num, depth, points;
};

// Consecutive nets' DAGs as one builder thread produces them (par local,
// npar = parent count per node).
struct DagChunk {
    RawVec<int> node, child_num, depth, npar, par, slot, points, nodes_per_net, points_per_net;
};

// DAGs of several nets concatenated; par_off is global over all nodes and
// par holds local (per-net) node indices.  The arrays and the per-chunk work
// buffers keep their capacity from one build to the next.
struct DagSet {
    int64_t n = 0;
    RawVec<int64_t> node_off, par_off, pt_off;
    RawVec<int> node, par, slot, child_num, depth, pt;
    std::vector<DagChunk> work;
};

void build_dag(const Grid& g, const float* xsum, const float* ysum, int mode, const int* tnode, int tn,
               const int64_t* adj_off, const int* tadj, DagNet& out);

// What a build writes into the DAG set: every array, only the scheduler
// points (pt_off, pt), or every array but the points, which it leaves as they
// are.  A points build followed by a structure build over the same nets and
// views gives the arrays of a full build.
enum DagPart { DAG_ALL = 0, DAG_POINTS = 1, DAG_STRUCTURE = 2 };

void build_dags(const Grid& g, const Subnets& s, const int* nets, int64_t n, const float* xsum, const float* ysum,
                int mode, int threads, DagSet& out, int part = DAG_ALL);

} // namespace igr
