// C ABI of libigr.so: context lifetime, grid state transfers, nets, the
// scheduler and the pattern pass.
#include "igr_ctx.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <string>

using namespace igr_dev;

static thread_local std::string g_err;

const char* igr_last_error(void) { return g_err.c_str(); }

template <typename F>
static int guard(F&& f) {
    try {
        f();
        return IGR_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return IGR_EINVAL;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return IGR_ESTATE;
    } catch (const std::exception& e) {
        g_err = e.what();
        return IGR_EHIP;
    }
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int igr_create(const igr_grid_desc* d, igr_ctx** out) {
    *out = nullptr;
    igr_ctx* c = new igr_ctx();
    int rc = guard([&]() {
        if (!d || d->L < 1 || d->L > 12 || d->X < 2 || d->Y < 2)
            throw std::invalid_argument("grid must have 1..12 routing layers and at least 2x2 gcells");
        if ((int64_t)d->L * d->X * d->Y > 0x7fffffffLL)
            throw std::invalid_argument("grid too large for 32-bit cell indices");
        if (std::max(d->X, d->Y) > 8192) throw std::invalid_argument("tracks longer than 8192 gcells");
        int ndev = 0;
        HIP_OK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw std::runtime_error("no GPU visible");
        HIP_OK(hipGetDevice(&c->device));
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        GridDev& g = c->g;
        g.L = d->L; g.X = d->X; g.Y = d->Y; g.DIR = d->DIR & 1;
        g.XYc = (int64_t)d->X * d->Y;
        g.unit_wire = d->unit_wire; g.unit_via = d->unit_via;
        const int64_t n = c->cells();
        c->w.upload(d->of_weight, d->L, c->stream);
        c->xlen.upload(d->xlen, d->X - 1, c->stream);
        c->ylen.upload(d->ylen, d->Y - 1, c->stream);
        DevBuf<float> tmp;
        tmp.upload(d->capacity, n, c->stream);
        c->cap.alloc(n);
        c->demand.zero(n, c->stream);
        c->wcost.zero(n, c->stream);
        c->vcost.zero(n, c->stream);
        c->presum.zero(n, c->stream);
        c->pre.zero(n, c->stream);
        c->ts.zero(n, c->stream);
        g.w = c->w.p; g.xlen = c->xlen.p; g.ylen = c->ylen.p; g.cap = c->cap.p;
        g.demand = c->demand.p; g.wcost = c->wcost.p; g.vcost = c->vcost.p; g.presum = c->presum.p;
        g.pre = c->pre.p; g.ts = c->ts.p;
        launch_from_ref(c, tmp.p, c->cap.p, 4);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
    if (rc != IGR_OK) {
        igr_destroy(c);
        return rc;
    }
    *out = c;
    return IGR_OK;
}

void igr_destroy(igr_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    hipStream_t s = c->stream;
    delete c;
    if (s) (void)hipStreamDestroy(s);
}

static void to_host_ref(igr_ctx* c, const void* dev, void* host, int eb) {
    const int64_t n = c->cells();
    if (eb == 4) {
        DevBuf<int> t;
        t.alloc(n);
        launch_to_ref(c, dev, t.p, 4);
        HIP_OK(hipMemcpyAsync(host, t.p, 4 * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    } else {
        DevBuf<double> t;
        t.alloc(n);
        launch_to_ref(c, dev, t.p, 8);
        HIP_OK(hipMemcpyAsync(host, t.p, 8 * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
}

int igr_set_demand(igr_ctx* c, const float* demand) {
    return guard([&]() {
        DevBuf<float> t;
        t.upload(demand, c->cells(), c->stream);
        launch_from_ref(c, t.p, c->demand.p, 4);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}
int igr_get_demand(igr_ctx* c, float* demand) { return guard([&]() { to_host_ref(c, c->demand.p, demand, 4); }); }
int igr_get_timestamp(igr_ctx* c, int* ts) { return guard([&]() { to_host_ref(c, c->ts.p, ts, 4); }); }
int igr_get_costs(igr_ctx* c, float* wcost, float* vcost, double* presum) {
    return guard([&]() {
        if (wcost) to_host_ref(c, c->wcost.p, wcost, 4);
        if (vcost) to_host_ref(c, c->vcost.p, vcost, 4);
        if (presum) to_host_ref(c, c->presum.p, presum, 8);
    });
}

int igr_update_cost(igr_ctx* c) {
    return guard([&]() {
        launch_update_cost(c);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}
int igr_compute_presum(igr_ctx* c) {
    return guard([&]() {
        launch_presum(c);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int igr_set_nets(igr_ctx* c, const igr_net_desc* d) {
    return guard([&]() {
        if (!d || d->n < 0 || !d->node_off || !d->nodes || !d->parents)
            throw std::invalid_argument("net descriptor needs node_off, nodes and parents");
        NetsDev& N = c->nets;
        const int64_t n = d->n;
        const int XY = c->g.X * c->g.Y;
        N.n = n;
        N.h_node_off.assign(d->node_off, d->node_off + n + 1);
        N.h_node.assign(d->nodes, d->nodes + N.h_node_off[n]);
        N.h_route_off.assign(n + 1, 0);
        N.max_nodes = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t k = N.h_node_off[i + 1] - N.h_node_off[i];
            if (k < 1) throw std::invalid_argument("every net needs at least one node");
            for (int64_t j = 0; j < k; j++) {
                int p = d->parents[N.h_node_off[i] + j];
                if (j == 0 ? p != -1 : (p < 0 || p >= j)) throw std::invalid_argument("nodes must be in DFS pre-order");
                int v = d->nodes[N.h_node_off[i] + j];
                if (v < 0 || v >= (c->g.L + 1) * XY) throw std::invalid_argument("node outside the grid");
            }
            N.max_nodes = std::max<int>(N.max_nodes, (int)k);
            N.h_route_off[i + 1] = N.h_route_off[i] + std::max<int64_t>(3, 8 * k - 5);
        }
        hipStream_t s = c->stream;
        N.node_off.upload(d->node_off, n + 1, s);
        N.node.upload(d->nodes, N.h_node_off[n], s);
        N.parent.upload(d->parents, N.h_node_off[n], s);
        N.route_off.upload(N.h_route_off.data(), n + 1, s);
        // Every route starts empty: r[0] = 1.
        std::vector<int> init(N.h_route_off[n], 0);
        for (int64_t i = 0; i < n; i++) init[N.h_route_off[i]] = 1;
        N.routes.upload(init.data(), init.size(), s);
        std::vector<int64_t> z(n + 1, 0);
        N.pt_off.upload(d->pt_off ? d->pt_off : z.data(), n + 1, s);
        N.hs_off.upload(d->hs_off ? d->hs_off : z.data(), n + 1, s);
        N.vs_off.upload(d->vs_off ? d->vs_off : z.data(), n + 1, s);
        int64_t npt = d->pt_off ? d->pt_off[n] : 0, nhs = d->hs_off ? d->hs_off[n] : 0,
                nvs = d->vs_off ? d->vs_off[n] : 0;
        for (int64_t i = 0; i < npt; i++)
            if (d->pts[i] < 0 || d->pts[i] >= XY) throw std::invalid_argument("scheduler point outside the grid");
        for (int64_t i = 0; i < 2 * nhs; i++)
            if (d->hs[i] < 0 || d->hs[i] >= XY) throw std::invalid_argument("segment outside the grid");
        for (int64_t i = 0; i < 2 * nvs; i++)
            if (d->vs[i] < 0 || d->vs[i] >= XY) throw std::invalid_argument("segment outside the grid");
        N.pt.upload(d->pts, npt, s);
        N.hs.upload(d->hs, 2 * nhs, s);
        N.vs.upload(d->vs, 2 * nvs, s);
        HIP_OK(hipStreamSynchronize(s));
    });
}

int igr_schedule(igr_ctx* c, const int* order, int64_t n, int mode, int* batch_of, int* nbatches) {
    return guard([&]() {
        if (n > c->nets.n) throw std::invalid_argument("more nets than uploaded");
        for (int64_t i = 0; i < n; i++)
            if (order[i] < 0 || order[i] >= c->nets.n) throw std::invalid_argument("net id out of range");
        *nbatches = run_schedule(c, order, n, mode, batch_of);
    });
}

namespace {

// Batch-local DP scratch offsets and the scratch size for a list of nets.
int64_t scratch_offsets(const igr_ctx* c, const int* nets, int64_t n, int64_t* off) {
    int64_t acc = 0;
    for (int64_t i = 0; i < n; i++) {
        off[i] = acc;
        acc += c->nets.h_node_off[nets[i] + 1] - c->nets.h_node_off[nets[i]];
    }
    return acc;
}

void ensure_scratch(igr_ctx* c, int64_t nodes) {
    const int64_t LL = (int64_t)c->g.L * c->g.L;
    c->scratch.dist.alloc(nodes * LL);
    c->scratch.from.alloc(nodes * LL);
    c->scratch.lr.alloc(nodes);
}

// Device work of one plain batch: the reference sequence of launches.
void plain_batch(igr_ctx* c, const int* d_nets, const int64_t* d_off, int64_t n) {
    int stamp = ++c->stamp;
    launch_update_cost(c);
    launch_presum(c);
    launch_lshape_plain(c, d_nets, d_off, n, stamp);
    launch_wire_commit(c, stamp);
    launch_via_commit(c, d_nets, n, stamp, 1);
}

void check_nets(const igr_ctx* c, const int* nets, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (nets[i] < 0 || nets[i] >= c->nets.n) throw std::invalid_argument("net id out of range");
}

} // namespace

int igr_route_batch(igr_ctx* c, const int* nets, int64_t n, int form) {
    return guard([&]() {
        check_nets(c, nets, n);
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        std::vector<int64_t> off(n);
        int64_t tot = scratch_offsets(c, nets, n, off.data());
        ensure_scratch(c, tot);
        c->scratch.net_list.upload(nets, n, c->stream);
        c->scratch.scr_off.upload(off.data(), n, c->stream);
        plain_batch(c, c->scratch.net_list.p, c->scratch.scr_off.p, n);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int igr_pattern_route(igr_ctx* c, const int* order, int64_t n, int form, igr_pass_stats* st) {
    return guard([&]() {
        check_nets(c, order, n);
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        double t0 = now_s();
        std::vector<int> batch_of(n);
        int nb = run_schedule(c, order, n, 0, batch_of.data());
        double t1 = now_s();
        // Batches in routing order: the scheduler's last batch first; inside a
        // batch, the scheduler's index order (src/Lshape_route.hpp:175-183).
        std::vector<int64_t> cnt(nb + 1, 0);
        for (int64_t i = 0; i < n; i++) cnt[nb - 1 - batch_of[i] + 1]++;
        for (int b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
        std::vector<int> list(n);
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t i = n - 1; i >= 0; i--) list[pos[nb - 1 - batch_of[i]]++] = order[i];
        std::vector<int64_t> off(n);
        int64_t max_nodes = 0;
        for (int b = 0; b < nb; b++)
            max_nodes = std::max(max_nodes, scratch_offsets(c, list.data() + cnt[b], cnt[b + 1] - cnt[b],
                                                            off.data() + cnt[b]));
        ensure_scratch(c, max_nodes);
        c->scratch.net_list.upload(list.data(), n, c->stream);
        c->scratch.scr_off.upload(off.data(), n, c->stream);
        for (int b = 0; b < nb; b++)
            plain_batch(c, c->scratch.net_list.p + cnt[b], c->scratch.scr_off.p + cnt[b], cnt[b + 1] - cnt[b]);
        HIP_OK(hipStreamSynchronize(c->stream));
        double t2 = now_s();
        if (st) {
            st->batches = nb;
            st->schedule_s = t1 - t0;
            st->route_s = t2 - t1;
            st->nets = n;
        }
    });
}

int igr_get_route(igr_ctx* c, int net, int* buf, int cap, int* len) {
    return guard([&]() {
        if (net < 0 || net >= c->nets.n) throw std::invalid_argument("net id out of range");
        int64_t b = c->nets.h_route_off[net], e = c->nets.h_route_off[net + 1];
        std::vector<int> tmp(e - b);
        HIP_OK(hipMemcpyAsync(tmp.data(), c->nets.routes.p + b, sizeof(int) * (e - b), hipMemcpyDeviceToHost,
                              c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        int used = tmp[0];
        if (len) *len = used;
        for (int i = 0; i < used && i < cap; i++) buf[i] = tmp[i];
    });
}

int64_t igr_routes_capacity(igr_ctx* c) { return c->nets.h_route_off.empty() ? 0 : c->nets.h_route_off.back(); }

int igr_get_routes(igr_ctx* c, int64_t* off, int* data, int64_t cap) {
    return guard([&]() {
        const int64_t n = c->nets.n, tot = c->nets.h_route_off[n];
        if (cap < tot) throw std::invalid_argument("route buffer too small");
        HIP_OK(hipMemcpyAsync(data, c->nets.routes.p, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        std::memcpy(off, c->nets.h_route_off.data(), sizeof(int64_t) * (n + 1));
    });
}

int igr_ripup_mark(igr_ctx* c, int th, uint8_t* is_of) {
    return guard([&]() {
        c->scratch.of.alloc(c->cells());
        c->scratch.is_of.alloc(c->nets.n);
        launch_ripup_mark(c, th, c->scratch.is_of.p, c->scratch.of.p);
        HIP_OK(hipMemcpyAsync(is_of, c->scratch.is_of.p, c->nets.n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int igr_congestion_view(igr_ctx* c, float* xsum, float* ysum) {
    return guard([&]() {
        const int64_t q = c->g.XYc;
        c->scratch.xsum.alloc(q);
        c->scratch.ysum.alloc(q);
        launch_congestion(c, c->scratch.xsum.p, c->scratch.ysum.p);
        HIP_OK(hipMemcpyAsync(xsum, c->scratch.xsum.p, 4 * q, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(ysum, c->scratch.ysum.p, 4 * q, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int igr_stamp(igr_ctx* c) { return c->stamp; }

int igr_expf(const float* in, float* out, int64_t n) {
    return guard([&]() {
        hipStream_t s;
        HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DevBuf<float> a, b;
        a.upload(in, n, s);
        b.alloc(n);
        launch_expf(a.p, b.p, n, s);
        HIP_OK(hipMemcpyAsync(out, b.p, 4 * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipStreamDestroy(s));
    });
}
