// C interface of the host library (libigr_host.so): design generation and
// I/O, access points, subnets and trees, routing inputs, route writing and the
// evaluator.  Used by the C++ driver, the GPU library and the Python tests.
#include "igr_host.h"
#include "prep.hpp"
#include "dag.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

using namespace igr;

struct igrh_design {
    Design d;
    NetPins np;
    Grid g;
    Subnets s;
    PatternInputs pi;
    DagSet dags;
    bool built = false;
    double t_tree = 0;
};

static thread_local std::string g_err;
const char* igrh_error(void) { return g_err.c_str(); }

template <typename F>
static int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

igrh_design* igrh_generate(int preset, int64_t nets_override, uint64_t seed_override) {
    igrh_design* h = new igrh_design();
    if (guard([&]() {
            GenConfig c = preset_config(preset);
            if (nets_override > 0) c.nets = nets_override;
            if (seed_override) c.seed = seed_override;
            generate_design(c, h->d);
        })) { delete h; return nullptr; }
    return h;
}

igrh_design* igrh_generate_custom(const igrh_gen_params* p) {
    igrh_design* h = new igrh_design();
    if (guard([&]() {
            GenConfig c;
            c.L = p->layers; c.X = p->xsize; c.Y = p->ysize; c.nets = p->nets;
            c.pin_cap = p->pin_cap; c.geom_p = p->geom_p; c.sigma = p->sigma;
            c.blockage_frac = p->blockage_frac; c.heavy_frac = p->heavy_frac;
            c.heavy_min = p->heavy_min; c.heavy_max = p->heavy_max; c.heavy_window = p->heavy_window;
            c.hotspots = p->hotspots; c.hotspot_scale = p->hotspot_scale; c.seed = p->seed;
            generate_design(c, h->d);
        })) { delete h; return nullptr; }
    return h;
}

igrh_design* igrh_load(const char* cap_path, const char* net_path) {
    igrh_design* h = new igrh_design();
    if (guard([&]() {
            read_cap(cap_path, h->d);
            read_net(net_path, h->d);
        })) { delete h; return nullptr; }
    return h;
}

int igrh_save(const igrh_design* h, const char* cap_path, const char* net_path) {
    return guard([&]() {
        write_cap(cap_path, h->d);
        write_net(net_path, h->d);
    });
}

void igrh_free(igrh_design* h) { delete h; }

int igrh_build(igrh_design* h, int threads) {
    return guard([&]() {
        using clk = std::chrono::steady_clock;
        auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto t0 = clk::now();
        select_access_points(h->d, h->np, threads);
        build_grid(h->d, h->g);
        const auto t1 = clk::now();
        build_subnets(h->d, h->np, h->g, h->s, threads);
        const auto t2 = clk::now();
        prepare_pattern_inputs(h->g, h->s, h->pi, threads);
        const auto t3 = clk::now();
        h->t_tree = sec(t0, t3);
        if (std::getenv("IGR_PROGRESS"))
            std::fprintf(stderr, "[igr] trees and pattern inputs: access points and grid %.3f s, trees and subnets "
                                 "%.3f s, pattern inputs %.3f s\n", sec(t0, t1), sec(t1, t2), sec(t2, t3));
        h->built = true;
    });
}

double igrh_build_seconds(const igrh_design* h) { return h->t_tree; }

int igrh_grid(const igrh_design* h, igrh_grid_info* out) {
    if (!h->built) { g_err = "design not built"; return -1; }
    const Grid& g = h->g;
    out->L = g.L; out->X = g.X; out->Y = g.Y; out->DIR = g.DIR;
    out->unit_wire = g.wire_cost; out->unit_via = g.via_cost;
    out->w = g.w.data(); out->xlen = g.xlen.data(); out->ylen = g.ylen.data(); out->cap = g.cap.data();
    out->metal_layers = h->d.L;
    out->nets = h->d.num_nets();
    out->subnets = h->s.count;
    out->max_nodes = h->pi.max_nodes;
    return 0;
}

int64_t igrh_array(const igrh_design* h, const char* name, const void** ptr) {
    const PatternInputs& pi = h->pi;
    const Subnets& s = h->s;
    std::string n(name);
#define A(key, vec) if (n == key) { *ptr = (vec).data(); return (int64_t)(vec).size(); }
    A("node_off", pi.node_off) A("node", pi.node) A("parent", pi.parent)
    A("pt_off", pi.pt_off) A("pt", pi.pt) A("hs_off", pi.hs_off) A("hs", pi.hs)
    A("vs_off", pi.vs_off) A("vs", pi.vs) A("order", pi.order)
    A("hpwl", s.hpwl) A("orig", s.orig) A("sub_pin_off", s.pin_off) A("sub_pins", s.pins)
    A("tree_node_off", s.node_off) A("tree_node", s.node) A("tree_adj_off", s.adj_off) A("tree_adj", s.adj)
    A("net_sub_off", s.net_sub_off) A("net_sub", s.net_sub)
    A("pin_off", h->np.pin_off) A("pins", h->np.pins) A("extra_off", h->np.extra_off) A("extra", h->np.extra)
    A("ap_off", h->d.ap_off) A("ap", h->d.ap) A("net_pin_off", h->d.pin_off)
    A("metal_cap", h->d.cap) A("metal_dir", h->d.dir) A("metal_of_weight", h->d.of_weight)
    A("dag_node_off", h->dags.node_off) A("dag_node", h->dags.node) A("dag_par_off", h->dags.par_off)
    A("dag_par", h->dags.par) A("dag_slot", h->dags.slot) A("dag_child_num", h->dags.child_num)
    A("dag_depth", h->dags.depth) A("dag_pt_off", h->dags.pt_off) A("dag_pt", h->dags.pt)
#undef A
    g_err = "unknown array " + n;
    *ptr = nullptr;
    return -1;
}

int igrh_build_dags_part(igrh_design* h, const int* nets, int64_t n, const float* xsum, const float* ysum, int mode,
                         int threads, int part) {
    return guard([&]() {
        if (!h->built) throw std::runtime_error("design not built");
        for (int64_t i = 0; i < n; i++)
            if (nets[i] < 0 || nets[i] >= h->s.count) throw std::runtime_error("net id out of range");
        build_dags(h->g, h->s, nets, n, xsum, ysum, mode, threads, h->dags, part);
    });
}

int igrh_build_dags(igrh_design* h, const int* nets, int64_t n, const float* xsum, const float* ysum, int mode,
                     int threads) {
    return igrh_build_dags_part(h, nets, n, xsum, ysum, mode, threads, IGRH_DAG_ALL);
}

int igrh_write_routes(const igrh_design* h, const char* out_path, const int64_t* off, const int* data,
                      int64_t data_len) {
    return guard([&]() {
        RouteSet rs;
        rs.off.assign(off, off + h->s.count + 1);
        rs.data.assign(data, data + data_len);
        write_routes(out_path, h->d, h->np, h->g, h->s, rs);
    });
}

int igrh_evaluate(const igrh_design* h, const char* out_path, igrh_score* out) {
    return guard([&]() {
        Score sc;
        evaluate_routes(h->d, out_path, sc);
        out->wirelength = sc.wl; out->via = sc.via; out->overflow = sc.of; out->total = sc.total;
        out->opens = sc.opens;
    });
}

const igr::Design* igrh_design_ref(const igrh_design* h) { return &h->d; }
const igr::NetPins* igrh_netpins_ref(const igrh_design* h) { return &h->np; }
const igr::Grid* igrh_grid_ref(const igrh_design* h) { return &h->g; }
const igr::Subnets* igrh_subnets_ref(const igrh_design* h) { return &h->s; }
const igr::PatternInputs* igrh_inputs_ref(const igrh_design* h) { return &h->pi; }
const igr::DagSet* igrh_dags_ref(const igrh_design* h) { return &h->dags; }

int igrh_steiner(int n, const int* x, const int* y, int xsize, int ysize, const int* xlen, const int* ylen,
                 int node_cap, int* nodes_xy, int* edges, int* n_nodes) {
    return guard([&]() {
        if (n < 1 || !x || !y || !nodes_xy || !edges || !n_nodes || xsize < 1 || ysize < 1 ||
            (xsize > 1 && !xlen) || (ysize > 1 && !ylen))
            throw std::invalid_argument("igrh_steiner: bad arguments");
        std::vector<int> vx(x, x + n), vy(y, y + n);
        std::vector<int> lx(xlen, xlen + (xsize - 1)), ly(ylen, ylen + (ysize - 1));
        for (int v : lx)
            if (v < 0) throw std::invalid_argument("negative gcell length");
        for (int v : ly)
            if (v < 0) throw std::invalid_argument("negative gcell length");
        std::vector<std::pair<int, int>> nodes, es;
        steiner_of_points(vx, vy, lx, ly, nodes, es);
        if ((int)nodes.size() > node_cap) throw std::invalid_argument("igrh_steiner: node_cap too small");
        for (size_t i = 0; i < nodes.size(); i++) {
            nodes_xy[2 * i] = nodes[i].first;
            nodes_xy[2 * i + 1] = nodes[i].second;
        }
        for (size_t i = 0; i < es.size(); i++) {
            edges[2 * i] = es[i].first;
            edges[2 * i + 1] = es[i].second;
        }
        *n_nodes = (int)nodes.size();
    });
}
