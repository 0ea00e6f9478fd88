// Host-side design model for the global-routing path: the parsed .cap/.net
// contents (metal-layer coordinates), pin access-point selection, the routing
// grid handed to the GPU (routing-layer coordinates) and the subnets with their
// Steiner trees.  Pure C++17; no GPU code here.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace igr {

// Allocator that leaves elements default-initialised on resize (no zero
// fill), for arrays that are written in full right after they are sized (the
// pattern inputs, the detour DAGs): their pages are then first touched by the
// parallel fill instead of a serial zeroing pass.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using RawVec = std::vector<T, NoInitAlloc<T>>;

// ---------------------------------------------------------------------------
// Parsed design in metal-layer coordinates.  A 3-D position is encoded as
// l*X*Y + x*Y + y, as the reference does (src/database.hpp:358).
// ---------------------------------------------------------------------------
struct Design {
    int L = 0, X = 0, Y = 0;               // metal layers, gcells in x and y
    double wire_cost = 0, via_cost = 0;    // unit wire / via cost
    std::vector<double> of_weight;         // per metal layer overflow weight
    std::vector<int> xlen, ylen;           // gcell edge lengths (X-1, Y-1)
    std::vector<int> dir;                  // per metal layer: 0 = horizontal, 1 = vertical
    std::vector<double> min_len;           // per metal layer (carried, unused by routing)
    std::vector<double> cap;               // capacity[l*X*Y + x*Y + y]

    // Nets: names, pins grouped as access points (CSR over pins, then APs).
    std::vector<std::string> names;
    std::vector<int64_t> pin_off;          // net -> first pin (size nets+1)
    std::vector<int64_t> ap_off;           // pin -> first access point (size pins+1)
    std::vector<int> ap;                   // access point positions
    std::vector<int> bbox;                 // per net: minx,maxx,miny,maxy over all APs

    int64_t num_nets() const { return (int64_t)names.size(); }
};

// One access point chosen per pin (src/database.hpp:38-243).  'pins' holds one
// position per distinct 2-D location; 'extra' holds via pairs that join pins
// stacked on the same 2-D location.
struct NetPins {
    std::vector<int64_t> pin_off;   // net -> selected pins
    std::vector<int> pins;
    std::vector<int64_t> extra_off; // net -> extra route ints (pairs)
    std::vector<int> extra;
    std::vector<int> hpwl;          // per net, from the AP bounding box
};

// ---------------------------------------------------------------------------
// Routing grid in routing-layer coordinates: routing layer l = metal l+1
// (src/database_cuda.hpp:2106,2285-2290).  Direction of routing layer l is
// (l & 1) ^ DIR with DIR = direction of metal 2 (src/database_cuda.hpp:2258).
// ---------------------------------------------------------------------------
struct Grid {
    int L = 0, X = 0, Y = 0, DIR = 0;
    double wire_cost = 0, via_cost = 0;
    std::vector<double> w;          // overflow weight per routing layer
    std::vector<int> xlen, ylen;
    std::vector<float> cap;         // [l*X*Y + x*Y + y]
    int64_t cells() const { return (int64_t)L * X * Y; }
    int layer_dir(int l) const { return (l & 1) ^ DIR; }
};

// ---------------------------------------------------------------------------
// Subnets (nets of at most 12 pins) with their trees.  Tree nodes are stored as
// l*X*Y + x*Y + y with l = L for Steiner nodes (src/database_cuda.hpp:790-792).
// Adjacency lists keep the insertion order the tree builder produced, which
// fixes the DFS child order used by the pattern router.
// ---------------------------------------------------------------------------
struct Subnets {
    int64_t count = 0;
    std::vector<int> orig;           // original net id
    std::vector<int> hpwl;           // 2-D half perimeter of the pins
    std::vector<int64_t> pin_off;    // routing-layer pins
    std::vector<int> pins;
    std::vector<int64_t> node_off;   // tree nodes
    std::vector<int> node;           // encoded position with layer (L = Steiner)
    std::vector<int64_t> adj_off;    // per global node index, CSR into adj
    std::vector<int> adj;            // local node indices
    // Subnets of every original net and the count of unfinished subnets.
    std::vector<int64_t> net_sub_off;
    std::vector<int> net_sub;
};

// ----------------------------- I/O ----------------------------------------
// Parse the reference .cap / .net formats (src/database.hpp:247-415).  Throws
// std::runtime_error on malformed input.
void read_cap(const std::string& path, Design& d);
void read_net(const std::string& path, Design& d);
void write_cap(const std::string& path, const Design& d);
void write_net(const std::string& path, const Design& d);

// Access-point choice per pin (src/database.hpp:38-243), deterministic.
void select_access_points(const Design& d, NetPins& np, int threads);

// Routing grid from the design (src/database_cuda.hpp:2258-2290).
void build_grid(const Design& d, Grid& g);

// Subnets and Steiner trees (src/database_cuda.hpp:2109-2256, :696-820).
void build_subnets(const Design& d, const NetPins& np, const Grid& g, Subnets& s, int threads);
// Steiner tree of distinct gcells (x[i], y[i]) with the given gcell edge
// lengths, as the subnet builder makes it: node positions (x, y) and edges.
void steiner_of_points(const std::vector<int>& x, const std::vector<int>& y, const std::vector<int>& xlen,
                       const std::vector<int>& ylen, std::vector<std::pair<int, int>>& nodes,
                       std::vector<std::pair<int, int>>& edges);

// Synthetic design generator (deterministic, own PRNG).
struct GenConfig {
    int L = 10, X = 200, Y = 200;
    int64_t nets = 20000;
    int pin_cap = 30;             // max pins per net
    double geom_p = 0.35;         // pins = 2 + geometric(p)
    double sigma = 8.0;           // cluster spread in gcells
    double blockage_frac = 0.05;  // fraction of edges zeroed by rectangles
    double heavy_frac = 0.0;      // fraction of heavy nets
    int heavy_min = 100, heavy_max = 1000, heavy_window = 300;
    int hotspots = 0;             // rectangles with capacity * hotspot_scale
    double hotspot_scale = 0.7;
    uint64_t seed = 1;
};
void generate_design(const GenConfig& c, Design& d);
GenConfig preset_config(int idx);   // the five benchmark configurations

} // namespace igr
