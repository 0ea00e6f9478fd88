// Inputs of the GPU routing passes prepared on the host: trees flattened in
// DFS pre-order, scheduler footprints and the routing priority order.
#pragma once
#include "design.hpp"

namespace igr {

struct PatternInputs {
    int64_t n = 0;
    std::vector<int64_t> node_off;   // per subnet
    // written in full after sizing (every node by the DFS, every footprint
    // entry by the fill pass, every position by the counting sort)
    RawVec<int> node;                // l*X*Y + x*Y + y (l = L for Steiner), DFS pre-order
    RawVec<int> parent;              // local index of the parent (-1 for the root)
    std::vector<int64_t> pt_off;     // scheduler points (x*Y+y)
    RawVec<int> pt;
    std::vector<int64_t> hs_off;     // horizontal segments, pairs of x*Y+y
    RawVec<int> hs;
    std::vector<int64_t> vs_off;     // vertical segments
    RawVec<int> vs;
    RawVec<int> order;               // subnets by hpwl, largest first (ties: lower id first)
    int max_nodes = 0;
};

// DFS pre-order of each subnet tree from its node 0 (src/Lshape_route.hpp:192-208)
// and the footprints of src/database_cuda.hpp:768-789.
void prepare_pattern_inputs(const Grid& g, const Subnets& s, PatternInputs& pi, int threads);

// Routes in the reference's per-net format: r[0] = next free slot (starts at
// 1) followed by pairs of cell indices l*X*Y + x*Y + y.
struct RouteSet {
    std::vector<int64_t> off;        // per subnet, capacity slice
    std::vector<int> data;
};

// Route file writer (src/graph.hpp:179-235): per original net its name, "(",
// metal-1 pin vias, the stacked-pin vias, the subnet routes with layers
// shifted up by one, ")".
void write_routes(const std::string& path, const Design& d, const NetPins& np, const Grid& g,
                  const Subnets& s, const RouteSet& routes);

// Evaluator (run/evaluator.cpp:458-633): connectivity check and the score
// wirelength + via + overflow in FP64.  Returns 0 when every net is connected.
struct Score {
    double wl = 0, via = 0, of = 0, total = 0;
    int64_t opens = 0;
};
int evaluate_routes(const Design& d, const std::string& out_path, Score& sc);

} // namespace igr
