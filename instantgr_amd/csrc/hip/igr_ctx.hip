// C ABI of libigr.so: context lifetime, grid state transfers, nets, the
// scheduler and the pattern pass.
#include "igr_ctx.hpp"
#include "par.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <thread>

using namespace igr_dev;

static thread_local std::string g_err;

const char* igr_last_error(void) { return g_err.c_str(); }

int igr_dev::fail(int status, const std::string& msg) {
    g_err = msg;
    return status;
}

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

void igr_dev::progress(const char* fmt, ...) {
    static const bool on = std::getenv("IGR_PROGRESS") != nullptr;
    if (!on) return;
    va_list ap;
    va_start(ap, fmt);
    std::fputs("[igr] ", stderr);
    std::vfprintf(stderr, fmt, ap);
    std::fputc('\n', stderr);
    std::fflush(stderr);
    va_end(ap);
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
        noalloc_init(c->device);
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
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
        g.tbase[0] = 0;
        for (int l = 0; l < g.L; l++) g.tbase[l + 1] = g.tbase[l] + (layer_dir(l, g.DIR) ? g.X : g.Y);
        c->track_stats.zero(2, c->stream);
        c->exptab.alloc(EXPTAB_N);
        launch_exptab(c->exptab.p, c->stream);
        g.exptab = c->exptab.p;
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
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->aux_stream) (void)hipStreamSynchronize(c->aux_stream);
    hipStream_t s = c->stream, cs = c->copy_stream, as = c->aux_stream;
    hipEvent_t ef = c->ev_fork, ej = c->ev_join;
    delete c;
    if (s) (void)hipStreamDestroy(s);
    if (cs) (void)hipStreamDestroy(cs);
    if (as) (void)hipStreamDestroy(as);
    if (ef) (void)hipEventDestroy(ef);
    if (ej) (void)hipEventDestroy(ej);
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
int igr_cost_sweep(igr_ctx* c, int form, int reps, double* ms) {
    return guard([&]() {
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        if (reps < 1 || !ms) throw std::invalid_argument("reps >= 1 and an output");
        *ms = time_cost_sweep(c, form, reps);
    });
}

int igr_compute_presum(igr_ctx* c) {
    return guard([&]() {
        launch_presum(c);
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

namespace {
// Every route starts empty: r[0] = 1 at each net's slot (the rest is zero).
__global__ void k_route_starts(const int64_t* route_off, int64_t n, int* routes) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) routes[route_off[i]] = 1;
}

int host_threads() { return std::max(1, std::min(16, (int)std::thread::hardware_concurrency())); }
} // namespace

int igr_set_nets(igr_ctx* c, const igr_net_desc* d) {
    return guard([&]() {
        if (!d || d->n < 0 || !d->node_off || !d->nodes || !d->parents)
            throw std::invalid_argument("net descriptor needs node_off, nodes and parents");
        NetsDev& N = c->nets;
        const int64_t n = d->n;
        const int XY = c->g.X * c->g.Y, Y = c->g.Y, Lr = c->g.L;
        const int64_t ncells = (int64_t)(c->g.L + 1) * XY;
        const int nth = host_threads();
        N.n = n;
        N.h_node_off.assign(d->node_off, d->node_off + n + 1);
        N.h_node.assign(d->nodes, d->nodes + N.h_node_off[n]);
        N.h_route_off.assign(n + 1, 0);
        N.h_lbytes.assign(n, 0.0);
        // Per net, in parallel: the DFS pre-order and grid checks, its route
        // slot size (prefix-summed below) and the algorithmic bytes of its
        // L-shape DP: per node its position, parent, via-stack costs (4 B per
        // layer) and route output; per straight edge two prefix sums per
        // layer; per L-shaped edge four prefix sums per layer plus the corner
        // via costs.
        std::vector<int> mxn(nth, 0);
        igr::parallel_ranges(n, nth, [&](int64_t a, int64_t b, int t) {
            for (int64_t i = a; i < b; i++) {
                const int64_t o = N.h_node_off[i], k = N.h_node_off[i + 1] - o;
                if (k < 1) throw std::invalid_argument("every net needs at least one node");
                int64_t straight = 0;
                for (int64_t j = 0; j < k; j++) {
                    const int p = d->parents[o + j];
                    if (j == 0 ? p != -1 : (p < 0 || p >= j)) throw std::invalid_argument("nodes must be in DFS pre-order");
                    const int v = d->nodes[o + j];
                    if (v < 0 || v >= ncells) throw std::invalid_argument("node outside the grid");
                    if (j > 0) {
                        const int u = d->nodes[o + p];
                        if ((v % XY) / Y == (u % XY) / Y || v % Y == u % Y) straight++;
                    }
                }
                mxn[t] = std::max<int>(mxn[t], (int)k);
                N.h_route_off[i + 1] = std::max<int64_t>(3, 8 * k - 5);
                N.h_lbytes[i] = (double)k * (8 + 4 * Lr + 32) + (double)straight * 8 * Lr +
                                (double)(k - 1 - straight) * 40 * Lr;
            }
        });
        N.max_nodes = *std::max_element(mxn.begin(), mxn.end());
        for (int64_t i = 0; i < n; i++) N.h_route_off[i + 1] += N.h_route_off[i];
        const int64_t npt = d->pt_off ? d->pt_off[n] : 0, nhs = d->hs_off ? d->hs_off[n] : 0,
                      nvs = d->vs_off ? d->vs_off[n] : 0;
        auto check_cells = [&](const int* a, int64_t cnt, const char* what) {
            igr::parallel_ranges(cnt, nth, [&](int64_t a0, int64_t a1, int) {
                for (int64_t i = a0; i < a1; i++)
                    if (a[i] < 0 || a[i] >= XY) throw std::invalid_argument(what);
            });
        };
        check_cells(d->pts, npt, "scheduler point outside the grid");
        check_cells(d->hs, 2 * nhs, "segment outside the grid");
        check_cells(d->vs, 2 * nvs, "segment outside the grid");
        hipStream_t s = c->stream;
        N.node_off.upload(d->node_off, n + 1, s);
        N.node.upload(d->nodes, N.h_node_off[n], s);
        N.parent.upload(d->parents, N.h_node_off[n], s);
        N.route_off.upload(N.h_route_off.data(), n + 1, s);
        N.routes.zero(N.h_route_off[n], s);
        if (n > 0) hipLaunchKernelGGL(k_route_starts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N.route_off.p, n,
                                      N.routes.p);
        HIP_OK(hipGetLastError());
        N.h_rbuf.assign(n, -1);
        N.clear_dag_routes();
        N.h_rptr.resize(n);
        N.h_rcap.resize(n);
        for (int64_t i = 0; i < n; i++) {
            N.h_rptr[i] = N.routes.p + N.h_route_off[i];
            N.h_rcap[i] = N.h_route_off[i + 1] - N.h_route_off[i];
        }
        N.rptr.upload(N.h_rptr.data(), n, s);
        std::vector<int64_t> z(n + 1, 0);
        N.pt_off.upload(d->pt_off ? d->pt_off : z.data(), n + 1, s);
        N.hs_off.upload(d->hs_off ? d->hs_off : z.data(), n + 1, s);
        N.vs_off.upload(d->vs_off ? d->vs_off : z.data(), n + 1, s);
        N.pt.upload(d->pts, npt, s);
        N.hs.upload(d->hs, 2 * nhs, s);
        N.vs.upload(d->vs, 2 * nvs, s);
        HIP_OK(hipStreamSynchronize(s));
    });
}

namespace {
// Net lists name each net at most once: a repeated net would be ripped up
// twice in one pass and its route slot released while still in use.
void check_nets(const igr_ctx* c, const int* nets, int64_t n) {
    if (n < 0 || (n > 0 && !nets)) throw std::invalid_argument("bad net list");
    std::vector<uint8_t> seen((size_t)c->nets.n, 0);
    for (int64_t i = 0; i < n; i++) {
        if (nets[i] < 0 || nets[i] >= c->nets.n) throw std::invalid_argument("net id out of range");
        if (seen[nets[i]]) throw std::invalid_argument("net listed twice");
        seen[nets[i]] = 1;
    }
}
} // namespace

int igr_schedule(igr_ctx* c, const int* order, int64_t n, int mode, int form, int* batch_of, int* nbatches) {
    return guard([&]() {
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        check_nets(c, order, n);
        *nbatches = form == IGR_FORM_PLAIN ? run_schedule(c, order, n, mode, batch_of)
                                           : run_schedule_batched(c, order, n, mode, batch_of);
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
void plain_batch(igr_ctx* c, const int* d_nets, const int64_t* d_off, int64_t n, double lbytes = 0) {
    int stamp = ++c->stamp;
    launch_update_cost(c);
    launch_presum(c);
    launch_lshape_plain(c, d_nets, d_off, n, stamp, lbytes);
    launch_wire_commit(c, stamp);
    launch_via_commit(c, d_nets, n, stamp, 1);
}


// Pattern routing writes the pattern slot: point nets that a rip-up pass
// moved elsewhere back at it.
void reset_pattern_ptrs(igr_ctx* c, const int* nets, int64_t n) {
    NetsDev& N = c->nets;
    bool dirty = false;
    for (int64_t i = 0; i < n; i++) {
        int* p = N.routes.p + N.h_route_off[nets[i]];
        if (N.h_rptr[nets[i]] != p) {
            N.release(nets[i]);
            N.h_rptr[nets[i]] = p;
            N.h_rcap[nets[i]] = N.h_route_off[nets[i] + 1] - N.h_route_off[nets[i]];
            dirty = true;
        }
    }
    if (dirty) N.rptr.upload(N.h_rptr.data(), N.n, c->stream);
}

__global__ void k_gather_routes(int* const* rptr, const int64_t* out_off, int* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int* r = rptr[i];
    int* o = out + out_off[i];
    const int used = r[0];
    for (int k = 0; k < used; k++) o[k] = r[k];
}

} // namespace

int igr_route_batch(igr_ctx* c, const int* nets, int64_t n, int form) {
    return guard([&]() {
        check_nets(c, nets, n);
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        reset_pattern_ptrs(c, nets, n);
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
        reset_pattern_ptrs(c, order, n);
        double t0 = now_s();
        std::vector<int> batch_of(n);
        int nb = form == IGR_FORM_PLAIN ? run_schedule(c, order, n, 0, batch_of.data())
                                        : run_schedule_batched(c, order, n, 0, batch_of.data());
        double t1 = now_s();
        progress("pattern pass: %d batches scheduled in %.3f s", nb, t1 - t0);
        // Batches in routing order: the scheduler's last batch first; inside a
        // batch, the scheduler's index order (src/Lshape_route.hpp:175-183).
        std::vector<int64_t> cnt(nb + 1, 0);
        for (int64_t i = 0; i < n; i++) cnt[nb - 1 - batch_of[i] + 1]++;
        for (int b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
        std::vector<int> list(n);
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t i = n - 1; i >= 0; i--) list[pos[nb - 1 - batch_of[i]]++] = order[i];
        std::vector<int64_t> off(n);
        if (form == IGR_FORM_PLAIN) {
            int64_t max_nodes = 0;
            for (int b = 0; b < nb; b++)
                max_nodes = std::max(max_nodes, scratch_offsets(c, list.data() + cnt[b], cnt[b + 1] - cnt[b],
                                                                off.data() + cnt[b]));
            ensure_scratch(c, max_nodes);
            c->scratch.net_list.upload(list.data(), n, c->stream);
            c->scratch.scr_off.upload(off.data(), n, c->stream);
            std::vector<double> lbytes(nb, 0.0);
            for (int b = 0; b < nb; b++)
                for (int64_t i = cnt[b]; i < cnt[b + 1]; i++) lbytes[b] += c->nets.h_lbytes[list[i]];
            for (int b = 0; b < nb; b++)
                plain_batch(c, c->scratch.net_list.p + cnt[b], c->scratch.scr_off.p + cnt[b], cnt[b + 1] - cnt[b],
                            lbytes[b]);
        } else {
            // Each batch's nets grouped by node-count class, one launch per
            // class: the order of nets inside a batch changes no result (their
            // footprints are disjoint and their demand updates integer adds).
            // Nets above the largest class (none under the 12-pin subnet cap)
            // take the thread-per-net kernel.
            constexpr int NCLS = 5;
            std::vector<int> wl(n);
            std::vector<int64_t> co((size_t)nb * NCLS + 1, 0);
            std::vector<double> cbytes((size_t)nb * NCLS, 0.0);
            {
                std::vector<int> cls(n);
                for (int64_t i = 0; i < n; i++) {
                    const int k = lshape_wave_class(
                        (int)(c->nets.h_node_off[list[i] + 1] - c->nets.h_node_off[list[i]]));
                    cls[i] = k < 0 ? NCLS - 1 : k;
                }
                for (int b = 0; b < nb; b++)
                    for (int64_t i = cnt[b]; i < cnt[b + 1]; i++) {
                        co[(size_t)b * NCLS + cls[i] + 1]++;
                        cbytes[(size_t)b * NCLS + cls[i]] += c->nets.h_lbytes[list[i]];
                    }
                for (size_t k = 0; k + 1 < co.size(); k++) co[k + 1] += co[k];
                std::vector<int64_t> pos(co.begin(), co.end() - 1);
                for (int b = 0; b < nb; b++)
                    for (int64_t i = cnt[b]; i < cnt[b + 1]; i++) wl[pos[(size_t)b * NCLS + cls[i]]++] = list[i];
            }
            int64_t max_nodes = 0;
            for (int b = 0; b < nb; b++) {
                const int64_t a = co[(size_t)b * NCLS + NCLS - 1], e = co[(size_t)b * NCLS + NCLS];
                max_nodes = std::max(max_nodes, scratch_offsets(c, wl.data() + a, e - a, off.data() + a));
            }
            if (max_nodes) ensure_scratch(c, max_nodes);
            c->scratch.net_list.upload(wl.data(), n, c->stream);
            c->scratch.scr_off.upload(off.data(), n, c->stream);
            // Costs of every track once, then per batch: DP (scattering wires,
            // vias and dirty flags), then the fused commit + cost pass over the
            // touched tracks.  The last pass leaves the costs as the plain
            // form does (computed before the last batch).
            BatchedState bs(c);
            launch_track_fused(c, 0, 0, true);
            for (int b = 0; b < nb; b++) {
                const int stamp = ++c->stamp;
                for (int k = 0; k < NCLS; k++) {
                    const int64_t a = co[(size_t)b * NCLS + k], e = co[(size_t)b * NCLS + k + 1];
                    if (a == e) continue;
                    if (k < NCLS - 1)
                        launch_lshape_wave(c, c->scratch.net_list.p + a, e - a, k, stamp, cbytes[(size_t)b * NCLS + k]);
                    else
                        launch_lshape_plain(c, c->scratch.net_list.p + a, c->scratch.scr_off.p + a, e - a, stamp,
                                            cbytes[(size_t)b * NCLS + k]);
                }
                launch_track_fused(c, stamp, 0, b + 1 < nb);
            }
            HIP_OK(hipStreamSynchronize(c->stream));
        }
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
        const int64_t rc = c->nets.h_rcap[net];
        std::vector<int> tmp(rc);
        HIP_OK(hipMemcpyAsync(tmp.data(), c->nets.h_rptr[net], sizeof(int) * rc, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        int used = tmp[0];
        if (len) *len = used;
        for (int i = 0; i < used && i < cap; i++) buf[i] = tmp[i];
    });
}

int64_t igr_routes_capacity(igr_ctx* c) {
    int64_t t = 0;
    for (int64_t v : c->nets.h_rcap) t += v;
    return t;
}

namespace igr_dev {
void gather_routes(igr_ctx* c, const int64_t* d_off, int* d_out) {
    const int64_t n = c->nets.n;
    if (n) k_gather_routes<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->nets.rptr.p, d_off, d_out, n);
    HIP_OK(hipGetLastError());
}
} // namespace igr_dev

int igr_evaluate(igr_ctx* c, const igr_eval_nets* e, const int64_t* route_off, const int* routes, igr_score* out) {
    return guard([&]() {
        if (!e || !out) throw std::invalid_argument("null argument");
        if ((route_off == nullptr) != (routes == nullptr))
            throw std::invalid_argument("route_off and routes go together");
        if (!route_off && c->nets.n == 0) throw std::invalid_argument("the context holds no routes");
        evaluate_routes_dev(c, e, route_off, routes, out);
    });
}

int igr_get_routes(igr_ctx* c, int64_t* off, int* data, int64_t cap) {
    return guard([&]() {
        const int64_t n = c->nets.n;
        std::vector<int64_t> o(n + 1, 0);
        for (int64_t i = 0; i < n; i++) o[i + 1] = o[i] + c->nets.h_rcap[i];
        if (cap < o[n]) throw std::invalid_argument("route buffer too small");
        DevBuf<int64_t> doff;
        DevBuf<int> out;
        doff.upload(o.data(), n + 1, c->stream);
        out.alloc(o[n]);
        gather_routes(c, doff.p, out.p);
        HIP_OK(hipMemcpyAsync(data, out.p, sizeof(int) * o[n], hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        std::memcpy(off, o.data(), sizeof(int64_t) * (n + 1));
    });
}

namespace {

void pfor(int64_t cnt, const std::function<void(int64_t, int64_t)>& fn) {
    const int nthreads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const int64_t chunk = (cnt + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        const int64_t a0 = t * chunk, a1 = std::min(cnt, a0 + chunk);
        if (a0 < a1) th.emplace_back(fn, a0, a1);
    }
    for (auto& x : th) x.join();
}

// Batched form of one rip-up-and-reroute pass.  While the scheduler runs on
// the compute stream, a helper thread uploads the DAGs as built (copy stream)
// and measures every net (nodes, edges, slots, levels, algorithmic bytes).
// With 'complete' set, dg holds only the points on entry and the helper first
// has the rest of it built, so that building overlaps the scheduler too.
// Once the routing order is known the host only computes per-net bases and
// batch bounds; the device lays the DAGs out (k_dag_layout) and checks them.
void dag_pass_batched(igr_ctx* c, const int* nets, int64_t n, int mode, const igr_dag_desc* dg_in,
                      const std::function<void(igr_dag_desc&)>& complete, igr_pass_stats* st) {
    const int XY = c->g.X * c->g.Y, L = c->g.L, LL = L * L, ML = L + 1;
    igr_dag_desc dgv = *dg_in;
    const igr_dag_desc* dg = &dgv;
    // The scheduler's points are checked here, the DAG structure on the device.
    if (dg->pt_off[0] != 0) throw std::invalid_argument("DAG point offsets must start at 0");
    std::atomic<int> bad_pt{0};
    pfor(n, [&](int64_t a0, int64_t a1) {
        int bad = 0;
        for (int64_t i = a0; i < a1; i++) {
            bad |= dg->pt_off[i + 1] < dg->pt_off[i];
            for (int64_t q = dg->pt_off[i]; q < dg->pt_off[i + 1]; q++) bad |= dg->pts[q] < 0 || dg->pts[q] >= XY;
        }
        if (bad) bad_pt.store(1);
    });
    if (bad_pt.load()) throw std::invalid_argument("DAG point outside the grid");
    DagDev& D = c->dag;
    hipStream_t s = c->stream;
    const double t0 = now_s();
    D.pt_off.upload(dg->pt_off, n + 1, s);
    D.pt.upload(dg->pts, dg->pt_off[n], s);

    double pairs[2] = {0, 0}, pvia[2] = {0, 0};
    for (int fl = 0; fl < L; fl++)
        for (int tl = fl & 1; tl < L; tl += 2) {
            pairs[layer_dir(fl, c->g.DIR)] += 1;
            pvia[layer_dir(fl, c->g.DIR)] += std::abs(fl - tl);
        }
    int64_t NN = 0, NE = 0;
    std::vector<int64_t> l_ns(n);
    std::vector<int> l_nl(n);
    std::vector<double> l_eb(n), l_db(n);
    std::atomic<int> bad_dag{0};
    std::exception_ptr herr;
    double t_helper = 0;
    std::thread helper([&]() {
        try {
            const double h0 = now_s();
            if (complete) complete(dgv);
            if (dg->node_off[0] != 0 || dg->par_off[0] != 0) throw std::invalid_argument("DAG offsets must start at 0");
            std::atomic<int> empty{0};
            pfor(n, [&](int64_t a0, int64_t a1) {
                for (int64_t i = a0; i < a1; i++)
                    if (dg->node_off[i + 1] <= dg->node_off[i]) empty.store(1);
            });
            if (empty.load()) throw std::invalid_argument("empty DAG");
            NN = dg->node_off[n];
            NE = dg->par_off[NN];
            hipStream_t cs = c->copy_stream;
            D.r_node_off.upload(dg->node_off, n + 1, cs);
            D.r_nodes.upload(dg->nodes, NN, cs);
            D.r_par_off.upload(dg->par_off, NN + 1, cs);
            D.r_par.upload(dg->par, NE, cs);
            D.r_slot.upload(dg->slot, NE, cs);
            D.r_child_num.upload(dg->child_num, NN, cs);
            D.r_depth.upload(dg->depth, NN, cs);
            // Per net (list order): slots, levels and the algorithmic bytes
            // of the DAG kernels (the models are stated in the plain path).
            pfor(n, [&](int64_t a0, int64_t a1) {
                for (int64_t i = a0; i < a1; i++) {
                    const int64_t b0 = dg->node_off[i], b1 = dg->node_off[i + 1];
                    int64_t slots = 0;
                    int maxd = 0;
                    double eb = 0, longp = 0;
                    for (int64_t k = b0; k < b1; k++) {
                        // the layout kernel trusts the edge offsets and depths
                        if (dg->par_off[k + 1] < dg->par_off[k] || dg->par_off[k + 1] > NE || dg->depth[k] < 0 ||
                            dg->depth[k] >= b1 - b0 || dg->child_num[k] < 0 || dg->child_num[k] > 10) {
                            bad_dag.store(1);
                            break;
                        }
                        slots += dg->child_num[k];
                        maxd = std::max(maxd, dg->depth[k]);
                        const int v = dg->nodes[k];
                        for (int64_t q = dg->par_off[k]; q < dg->par_off[k + 1]; q++) {
                            const int64_t pl = dg->par[q];
                            if (pl < 0 || pl >= b1 - b0) continue;   // rejected by the layout
                            const int u = dg->nodes[b0 + pl];
                            const int vx = v % XY / c->g.Y, vy = v % c->g.Y, ux = u % XY / c->g.Y, uy = u % c->g.Y;
                            const int dist = std::abs(vx - ux) + std::abs(vy - uy);
                            if (dist <= 4 || (vx != ux && vy != uy)) continue;
                            const int dir = vx == ux ? 1 : 0;
                            const int dd = dist / std::min(dist, 20);
                            eb += ((dist - 1) / dd) * (16 * pairs[dir] + 4 * pvia[dir]) + 12 * pairs[dir];
                            longp += pairs[dir];
                        }
                    }
                    const int64_t nn = b1 - b0, ne = dg->par_off[b1] - dg->par_off[b0];
                    l_ns[i] = slots;
                    l_nl[i] = maxd + 1;
                    l_eb[i] = eb;
                    l_db[i] = nn * (4.0 + 4 * L + 16 * ML + 32) + ne * (8.0 + 8 * L + 8 * ML) + longp * 8 +
                              slots * 16.0 * ML;
                }
            });
            HIP_OK(hipStreamSynchronize(cs));
            t_helper = now_s() - h0;
        } catch (...) {
            herr = std::current_exception();
        }
    });
    std::vector<int> batch_of(n);
    int nb = 0;
    try {
        nb = run_schedule_batched(c, nets, n, mode, batch_of.data(), D.pt_off.p, D.pt.p);
    } catch (...) {
        helper.join();
        throw;
    }
    helper.join();
    if (herr) std::rethrow_exception(herr);
    if (bad_dag.load()) throw std::invalid_argument("malformed DAG (edge offsets, depths or slot counts)");
    progress("DAG pass: %d batches scheduled in %.3f s", nb, now_s() - t0);
    const double t1 = now_s();

    // Routing order: batch nb-1 first, inside a batch decreasing list
    // position (the scheduler's index order).
    std::vector<int64_t> cntb(nb + 1, 0);
    for (int64_t i = 0; i < n; i++) cntb[nb - 1 - batch_of[i] + 1]++;
    for (int b = 0; b < nb; b++) cntb[b + 1] += cntb[b];
    std::vector<int> rorder(n), rnets(n);
    {
        std::vector<int64_t> pos(cntb.begin(), cntb.end() - 1);
        for (int64_t i = n - 1; i >= 0; i--) rorder[pos[nb - 1 - batch_of[i]]++] = (int)i;
    }
    std::vector<int64_t> nbase(n + 1, 0), ebase(n + 1, 0), sbase(n + 1, 0), lbase(n + 1, 0), route_slot(n + 1, 0);
    for (int64_t r = 0; r < n; r++) {
        const int64_t i = rorder[r];
        const int64_t nn = dg->node_off[i + 1] - dg->node_off[i];
        rnets[r] = nets[i];
        nbase[r + 1] = nbase[r] + nn;
        ebase[r + 1] = ebase[r] + (dg->par_off[dg->node_off[i + 1]] - dg->par_off[dg->node_off[i]]);
        sbase[r + 1] = sbase[r] + l_ns[i];
        lbase[r + 1] = lbase[r] + l_nl[i];
        route_slot[r + 1] = route_slot[r] + 8 * nn + 1;
    }
    const int64_t NS = sbase[n], NL = lbase[n];
    std::vector<DagBatch> batches(nb);
    int64_t mx_nodes = 0, mx_edges = 0, mx_slots = 0;
    // Per batch, the routed nets in launch order of the workgroup-per-net
    // kernel: deepest DAG first, then most nodes (dag_order 1), or routing
    // order (2).
    const bool deep_first = (c->tune.dag_order ? c->tune.dag_order : DEF_DAG_ORDER) == 1;
    std::vector<int> glist(n);
    std::vector<std::pair<int64_t, int64_t>> gl;   // (sort key, routed net)
    for (int b = 0; b < nb; b++) {
        DagBatch& B = batches[b];
        gl.clear();
        for (int64_t r = cntb[b]; r < cntb[b + 1]; r++)
            gl.emplace_back(deep_first ? -((lbase[r + 1] - lbase[r]) << 32 | (nbase[r + 1] - nbase[r])) : 0, r);
        if (deep_first) std::sort(gl.begin(), gl.end());
        for (size_t k = 0; k < gl.size(); k++) glist[cntb[b] + k] = (int)gl[k].second;
        B.net0 = cntb[b];
        B.net1 = cntb[b + 1];
        B.node0 = nbase[B.net0];
        B.node1 = nbase[B.net1];
        B.edge0 = ebase[B.net0];
        B.edge1 = ebase[B.net1];
        B.slot0 = sbase[B.net0];
        B.slot1 = sbase[B.net1];
        int maxl = 1;
        for (int64_t r = B.net0; r < B.net1; r++) {
            maxl = std::max(maxl, l_nl[rorder[r]]);
            B.bytes_edges += l_eb[rorder[r]];
            B.bytes_dp += l_db[rorder[r]];
        }
        B.levels = maxl;
        B.index = b;
        B.lvl0 = 0;
        mx_nodes = std::max(mx_nodes, B.node1 - B.node0);
        mx_edges = std::max(mx_edges, B.edge1 - B.edge0);
        mx_slots = std::max(mx_slots, B.slot1 - B.slot0);
    }
    // IGR_DAG_TRACE=<file>: append one line per batch (nets, nodes, edges,
    // slots, deepest DAG's levels, largest DAG's nodes).
    if (const char* tp = std::getenv("IGR_DAG_TRACE")) {
        if (FILE* f = std::fopen(tp, "a")) {
            std::fprintf(f, "# DAG pass: %lld nets, %d batches\nbatch nets nodes edges slots levels max_nodes\n",
                         (long long)n, nb);
            for (int b = 0; b < nb; b++) {
                const DagBatch& B = batches[b];
                int64_t mxn = 0;
                for (int64_t r = B.net0; r < B.net1; r++) mxn = std::max(mxn, nbase[r + 1] - nbase[r]);
                std::fprintf(f, "%d %lld %lld %lld %lld %d %lld\n", b, (long long)(B.net1 - B.net0),
                             (long long)(B.node1 - B.node0), (long long)(B.edge1 - B.edge0),
                             (long long)(B.slot1 - B.slot0), B.levels, (long long)mxn);
            }
            std::fclose(f);
        }
    }
    D.glist.upload(glist.data(), n, s);
    D.nets.upload(rnets.data(), n, s);
    D.rorder.upload(rorder.data(), n, s);
    D.node_off.upload(nbase.data(), n + 1, s);
    D.ebase.upload(ebase.data(), n + 1, s);
    D.sbase.upload(sbase.data(), n + 1, s);
    D.lbase.upload(lbase.data(), n + 1, s);
    D.route_slot.upload(route_slot.data(), n + 1, s);
    D.node.alloc(NN);
    D.child_num.alloc(NN);
    D.loc.alloc(NN);
    D.csum.alloc(NN + 1);
    D.par_off.alloc(NN + 1);
    D.par.alloc(NE);
    D.slot.alloc(NE);
    D.enode.alloc(NE);
    D.nl_off.alloc(n);
    D.nlevels.alloc(n);
    D.lvl_node_off.alloc(NL + n);
    D.lvl_edge_off.alloc(NL + n);
    D.cur_n.alloc(NL + n);
    D.cur_e.alloc(NL + n);
    D.dnode.alloc(NN);
    D.dedge.alloc(NE);
    D.err.zero(1, s);
    const int64_t tails[2] = {NE, NS};
    HIP_OK(hipMemcpyAsync(D.par_off.p + NN, &tails[0], sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(D.csum.p + NN, &tails[1], sizeof(int64_t), hipMemcpyHostToDevice, s));
    const double t_meta = now_s();
    DagLayoutArgs la{D.r_node_off.p, D.r_nodes.p, D.r_par_off.p, D.r_par.p, D.r_slot.p, D.r_child_num.p, D.r_depth.p,
                     D.rorder.p, D.node_off.p, D.ebase.p, D.sbase.p, D.lbase.p, (L + 1) * XY,
                     D.node.p, D.child_num.p, D.loc.p, D.csum.p, D.par_off.p, D.par.p, D.slot.p, D.enode.p,
                     D.nl_off.p, D.nlevels.p, D.lvl_node_off.p, D.lvl_edge_off.p, D.dnode.p, D.dedge.p,
                     D.cur_n.p, D.cur_e.p, D.err.p};
    launch_dag_layout(c, la, n);
    int err = 0;
    HIP_OK(hipMemcpyAsync(&err, D.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (err) throw std::invalid_argument("malformed DAG (node range, slot counts, source, parent order or slots)");
    const double t_up = now_s();
    D.costs.alloc(mx_nodes * ML);
    D.lo.alloc(mx_nodes);
    D.emit.alloc(mx_nodes * 9);
    D.packed.alloc(std::max<int64_t>(1, mx_slots) * ML);
    D.best.alloc(std::max<int64_t>(1, mx_slots) * ML);
    D.cost_edges.alloc(std::max<int64_t>(1, mx_edges) * LL);
    D.best_change.alloc(std::max<int64_t>(1, mx_edges) * LL);
    NetsDev& N = c->nets;
    N.dag_routes.push_back(new DevBuf<int>());
    N.dag_routes.back()->alloc(route_slot[n]);
    int* rbase = N.dag_routes.back()->p;
    const double tp = now_s();

    // Batch b: DAG routing (scattering its wires, vias and dirty flags), the
    // rip-up scatter of batch b+1 from its current routes, then one fused
    // pass: commit b, rip up b+1, costs.  Stamps follow the plain sequence
    // (rip-up, then commit, per batch).
    {
        BatchedState bs(c);
        launch_dag_queue_reset(c, (int64_t)batches.size());
        // The whole sequence is recorded as one graph unless kernel timing
        // (events on the stream) is on.
        StreamGraph sg(s, !c->prof.on && (c->tune.graph ? c->tune.graph : DEF_GRAPH) == 1);
        // The rip-up scatter of batch b+1 (the old routes of other nets, into
        // the rip-up difference arrays) runs on the side stream beside batch
        // b's DAG kernels (the new routes, into the commit arrays) and joins
        // before the fused update that consumes both.  Kernel timing keeps
        // one stream.
        const bool side = !c->prof.on;
        int s1 = ++c->stamp;
        launch_rip_scatter(c, D.nets.p + batches[0].net0, batches[0].net1 - batches[0].net0);
        launch_track_fused(c, 0, s1, true);
        for (size_t b = 0; b < batches.size(); b++) {
            if (b % 512 == 511) progress("DAG pass: %zu of %zu batches enqueued", b + 1, batches.size());
            const bool more = b + 1 < batches.size();
            if (side && more) {
                const DagBatch& Bn = batches[b + 1];
                HIP_OK(hipEventRecord(c->ev_fork, s));
                HIP_OK(hipStreamWaitEvent(c->aux_stream, c->ev_fork, 0));
                launch_rip_scatter(c, D.nets.p + Bn.net0, Bn.net1 - Bn.net0, c->aux_stream);
                HIP_OK(hipEventRecord(c->ev_join, c->aux_stream));
            }
            const int s2 = ++c->stamp;
            launch_dag_batch_wave(c, batches[b], s2);
            if (more) {
                const DagBatch& Bn = batches[b + 1];
                s1 = ++c->stamp;
                if (side) HIP_OK(hipStreamWaitEvent(s, c->ev_join, 0));
                else launch_rip_scatter(c, D.nets.p + Bn.net0, Bn.net1 - Bn.net0);
                launch_track_fused(c, s2, s1, true);
            } else {
                launch_track_fused(c, s2, 0, false);
            }
        }
        sg.launch();
        bs.check();
    }
    HIP_OK(hipStreamSynchronize(s));
    N.adopt(rnets.data(), route_slot.data(), n, rbase);
    const double t2 = now_s();
    c->last_dag_prep_s = tp - t1;
    c->last_dag_prep_parts[0] = t_meta - t1;     // routing-order bases, batch bounds, small uploads
    c->last_dag_prep_parts[1] = t_up - t_meta;   // device layout
    c->last_dag_prep_parts[2] = t_helper;        // DAG upload and measuring (overlaps the scheduler)
    c->last_dag_prep_parts[3] = tp - t_up;       // scratch allocation
    if (st) *st = igr_pass_stats{nb, t1 - t0, t2 - t1, n};
}

} // namespace

int igr_dag_pass(igr_ctx* c, const int* nets, int64_t n, int mode, int form, const igr_dag_desc* dg,
                 igr_pass_stats* st) {
    return guard([&]() {
        check_nets(c, nets, n);
        if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) throw std::invalid_argument("unknown form");
        if (!dg || dg->n != n) throw std::invalid_argument("one DAG per listed net");
        if (n == 0) {
            if (st) *st = igr_pass_stats{0, 0, 0, 0};
            return;
        }
        if (form == IGR_FORM_BATCHED) {
            dag_pass_batched(c, nets, n, mode, dg, nullptr, st);
            return;
        }
        const int XY = c->g.X * c->g.Y, L = c->g.L, LL = L * L, ML = L + 1;
        // Validate the DAGs before anything reaches the device.
        if (dg->node_off[0] != 0 || dg->par_off[0] != 0 || dg->pt_off[0] != 0)
            throw std::invalid_argument("DAG offsets must start at 0");
        for (int64_t i = 0; i < n; i++) {
            const int64_t b = dg->node_off[i], e = dg->node_off[i + 1];
            if (e <= b) throw std::invalid_argument("empty DAG");
            for (int64_t k = b; k < e; k++) {
                if (dg->nodes[k] < 0 || dg->nodes[k] >= (L + 1) * XY) throw std::invalid_argument("DAG node outside the grid");
                if (dg->child_num[k] < 0 || dg->child_num[k] > 10) throw std::invalid_argument("DAG node with > 10 slots");
                if ((k == b) != (dg->par_off[k + 1] == dg->par_off[k]))
                    throw std::invalid_argument("DAG node 0 must be the only source");
                for (int64_t q = dg->par_off[k]; q < dg->par_off[k + 1]; q++) {
                    int p = dg->par[q];
                    if (p < 0 || p >= k - b) throw std::invalid_argument("DAG parents must precede children");
                    if (dg->slot[q] < 0 || dg->slot[q] >= dg->child_num[b + p]) throw std::invalid_argument("bad DAG slot");
                }
            }
            for (int64_t q = dg->pt_off[i]; q < dg->pt_off[i + 1]; q++)
                if (dg->pts[q] < 0 || dg->pts[q] >= XY) throw std::invalid_argument("DAG point outside the grid");
        }
        DagDev& D = c->dag;
        hipStream_t s = c->stream;
        double t0 = now_s();
        D.pt_off.upload(dg->pt_off, n + 1, s);
        D.pt.upload(dg->pts, dg->pt_off[n], s);
        std::vector<int> batch_of(n);
        int nb = run_schedule(c, nets, n, mode, batch_of.data(), D.pt_off.p, D.pt.p);
        double t1 = now_s();
        progress("DAG pass: %d batches scheduled in %.3f s", nb, t1 - t0);

        // Routing order: batch nb-1 first, inside a batch decreasing list
        // position (the scheduler's index order).
        std::vector<int64_t> cntb(nb + 1, 0);
        for (int64_t i = 0; i < n; i++) cntb[nb - 1 - batch_of[i] + 1]++;
        for (int b = 0; b < nb; b++) cntb[b + 1] += cntb[b];
        std::vector<int64_t> rorder(n);
        {
            std::vector<int64_t> pos(cntb.begin(), cntb.end() - 1);
            for (int64_t i = n - 1; i >= 0; i--) rorder[pos[nb - 1 - batch_of[i]]++] = i;
        }
        // Per routed net sizes, then bases, then a parallel fill.
        std::vector<int64_t> nbase(n + 1, 0), ebase(n + 1, 0), sbase(n + 1, 0), lbase(n + 1, 0), route_slot(n + 1, 0);
        const int nthreads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
        auto pfor = [&](int64_t cnt, const std::function<void(int64_t, int64_t)>& fn) {
            std::vector<std::thread> th;
            int64_t chunk = (cnt + nthreads - 1) / nthreads;
            for (int t = 0; t < nthreads; t++) {
                int64_t a0 = t * chunk, a1 = std::min(cnt, a0 + chunk);
                if (a0 < a1) th.emplace_back(fn, a0, a1);
            }
            for (auto& x : th) x.join();
        };
        // Algorithmic bytes (HBM) of the DAG kernels per routed net.
        //  edge layer changes, per edge longer than 4 and per (from, to) layer
        //  pair of its direction: at each sampled switch point two prefix
        //  sums (16 B) and the via costs between the layers (4 B each), plus
        //  the result (12 B);
        //  DP and extraction: per node position, via-stack costs, layer costs
        //  written and read back, route output (32 B); per edge parent and
        //  slot, same-layer prefix pair per layer of its direction, the packed
        //  key atomics per layer; per long edge the layer-change costs read;
        //  per child slot the keys read and the back-pointers written.
        double pairs[2] = {0, 0}, pvia[2] = {0, 0};
        for (int fl = 0; fl < L; fl++)
            for (int tl = fl & 1; tl < L; tl += 2) {
                pairs[layer_dir(fl, c->g.DIR)] += 1;
                pvia[layer_dir(fl, c->g.DIR)] += std::abs(fl - tl);
            }
        std::vector<double> ebytes(n, 0.0), dbytes(n, 0.0);
        pfor(n, [&](int64_t a0, int64_t a1) {
            for (int64_t r = a0; r < a1; r++) {
                const int64_t i = rorder[r], b0 = dg->node_off[i], b1 = dg->node_off[i + 1];
                int64_t slots = 0;
                int maxd = 0;
                for (int64_t k = b0; k < b1; k++) { slots += dg->child_num[k]; maxd = std::max(maxd, dg->depth[k]); }
                double eb = 0, longp = 0;
                for (int64_t k = b0; k < b1; k++) {
                    const int v = dg->nodes[k];
                    for (int64_t q = dg->par_off[k]; q < dg->par_off[k + 1]; q++) {
                        const int u = dg->nodes[b0 + dg->par[q]];
                        const int vx = v % XY / c->g.Y, vy = v % c->g.Y, ux = u % XY / c->g.Y, uy = u % c->g.Y;
                        const int dist = std::abs(vx - ux) + std::abs(vy - uy);
                        if (dist <= 4 || (vx != ux && vy != uy)) continue;
                        const int dir = vx == ux ? 1 : 0;
                        const int dd = dist / std::min(dist, 20);
                        const int ns = (dist - 1) / dd;
                        eb += ns * (16 * pairs[dir] + 4 * pvia[dir]) + 12 * pairs[dir];
                        longp += pairs[dir];
                    }
                }
                const int64_t nn = b1 - b0, ne = dg->par_off[b1] - dg->par_off[b0];
                ebytes[r] = eb;
                dbytes[r] = nn * (4.0 + 4 * L + 16 * ML + 32) + ne * (8.0 + 8 * L + 8 * ML) + longp * 8 + slots * 16.0 * ML;
                nbase[r + 1] = b1 - b0;
                ebase[r + 1] = dg->par_off[b1] - dg->par_off[b0];
                sbase[r + 1] = slots;
                lbase[r + 1] = maxd + 1;
                route_slot[r + 1] = 8 * (b1 - b0) + 1;
            }
        });
        for (int64_t r = 0; r < n; r++) {
            nbase[r + 1] += nbase[r]; ebase[r + 1] += ebase[r]; sbase[r + 1] += sbase[r];
            lbase[r + 1] += lbase[r]; route_slot[r + 1] += route_slot[r];
        }
        const int64_t NN = nbase[n], NE = ebase[n], NS = sbase[n];
        std::vector<int> rnets(n), node(NN), par(NE), slot(NE), child_num(NN), loc(NN), enode(NE);
        std::vector<int64_t> par_off(NN + 1), csum(NN + 1);
        pfor(n, [&](int64_t a0, int64_t a1) {
            for (int64_t r = a0; r < a1; r++) {
                const int64_t i = rorder[r], b0 = dg->node_off[i], b1 = dg->node_off[i + 1];
                const int64_t g0 = nbase[r];
                int64_t e = ebase[r], s = sbase[r];
                rnets[r] = nets[i];
                for (int64_t k = b0; k < b1; k++) {
                    const int64_t gk = g0 + (k - b0);
                    node[gk] = dg->nodes[k];
                    child_num[gk] = dg->child_num[k];
                    loc[gk] = (int)(k - b0);
                    csum[gk] = s;
                    s += dg->child_num[k];
                    par_off[gk] = e;
                    for (int64_t q = dg->par_off[k]; q < dg->par_off[k + 1]; q++, e++) {
                        par[e] = (int)(g0 + dg->par[q]);
                        slot[e] = dg->slot[q];
                        enode[e] = (int)gk;
                    }
                }
            }
        });
        par_off[NN] = NE;
        csum[NN] = NS;
        const double t_fill = now_s();
        std::vector<DagBatch> batches;
        std::vector<int64_t> lvl_off(1, 0);
        std::vector<int> lvl_nodes;
        int64_t mx_nodes = 0, mx_edges = 0, mx_slots = 0;
        for (int b = 0; b < nb; b++) {
            DagBatch B{};
            B.net0 = cntb[b];
            B.net1 = cntb[b + 1];
            B.node0 = nbase[B.net0];
            B.node1 = nbase[B.net1];
            B.edge0 = ebase[B.net0];
            B.edge1 = ebase[B.net1];
            B.slot0 = sbase[B.net0];
            B.slot1 = sbase[B.net1];
            int maxd = 0;
            for (int64_t r = B.net0; r < B.net1; r++) {
                maxd = std::max<int>(maxd, (int)(lbase[r + 1] - lbase[r]) - 1);
                B.bytes_edges += ebytes[r];
                B.bytes_dp += dbytes[r];
            }
            B.levels = maxd + 1;
            B.lvl0 = (int64_t)lvl_off.size() - 1;
            {
                // All nodes of the batch bucketed by depth (stable in node order).
                std::vector<int64_t> cnt(B.levels + 1, 0);
                for (int64_t r = B.net0; r < B.net1; r++) {
                    const int64_t i = rorder[r];
                    for (int64_t q = dg->node_off[i]; q < dg->node_off[i + 1]; q++) cnt[dg->depth[q] + 1]++;
                }
                for (int d = 0; d < B.levels; d++) cnt[d + 1] += cnt[d];
                const int64_t base = (int64_t)lvl_nodes.size();
                lvl_nodes.resize(base + (B.node1 - B.node0));
                std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
                for (int64_t r = B.net0; r < B.net1; r++) {
                    const int64_t i = rorder[r];
                    for (int64_t q = dg->node_off[i]; q < dg->node_off[i + 1]; q++)
                        lvl_nodes[base + pos[dg->depth[q]]++] = (int)(nbase[r] + (q - dg->node_off[i]));
                }
                for (int d = 0; d < B.levels; d++) lvl_off.push_back(base + cnt[d + 1]);
            }
            mx_nodes = std::max(mx_nodes, B.node1 - B.node0);
            mx_edges = std::max(mx_edges, B.edge1 - B.edge0);
            mx_slots = std::max(mx_slots, B.slot1 - B.slot0);
            batches.push_back(B);
        }
        const double t_meta = now_s();
        D.nets.upload(rnets.data(), rnets.size(), s);
        D.node.upload(node.data(), node.size(), s);
        D.par.upload(par.data(), par.size(), s);
        D.slot.upload(slot.data(), slot.size(), s);
        D.child_num.upload(child_num.data(), child_num.size(), s);
        D.loc.upload(loc.data(), loc.size(), s);
        D.enode.upload(enode.data(), enode.size(), s);
        D.node_off.upload(nbase.data(), nbase.size(), s);
        D.par_off.upload(par_off.data(), par_off.size(), s);
        D.csum.upload(csum.data(), csum.size(), s);
        D.route_slot.upload(route_slot.data(), route_slot.size(), s);
        D.lvl_nodes.upload(lvl_nodes.data(), lvl_nodes.size(), s);
        HIP_OK(hipStreamSynchronize(s));
        const double t_up = now_s();
        D.costs.alloc(mx_nodes * ML);
        D.lo.alloc(mx_nodes);
        D.emit.alloc(mx_nodes * 9);
        D.packed.alloc(std::max<int64_t>(1, mx_slots) * ML);
        D.best.alloc(std::max<int64_t>(1, mx_slots) * ML);
        D.cost_edges.alloc(std::max<int64_t>(1, mx_edges) * LL);
        D.best_change.alloc(std::max<int64_t>(1, mx_edges) * LL);
        NetsDev& N = c->nets;
        N.dag_routes.push_back(new DevBuf<int>());
        N.dag_routes.back()->alloc(route_slot[n]);
        int* rbase = N.dag_routes.back()->p;
        double tp = now_s();

        double t_note = tp;
        for (const DagBatch& B : batches) {
            if (now_s() - t_note > 20) {
                t_note = now_s();
                progress("DAG pass: batch %lld of %d routed", (long long)(&B - batches.data()), nb);
            }
            const int* bn = D.nets.p + B.net0;
            const int64_t cnt = B.net1 - B.net0;
            const int s1 = ++c->stamp;
            launch_rip_wire(c, bn, cnt, s1);
            launch_via_commit(c, bn, cnt, s1, -1);
            const int s2 = ++c->stamp;
            launch_update_cost(c);
            launch_presum(c);
            launch_dag_batch_plain(c, B, s2, lvl_off);
            launch_dag_compact(c, B);
            launch_wire_commit(c, s2);
            launch_via_commit(c, bn, cnt, s2, 1);
            HIP_OK(hipStreamSynchronize(s));
        }
        HIP_OK(hipStreamSynchronize(s));
        N.adopt(rnets.data(), route_slot.data(), n, rbase);
        double t2 = now_s();
        c->last_dag_prep_s = tp - t1;
        c->last_dag_prep_parts[0] = t_fill - t1;
        c->last_dag_prep_parts[1] = t_meta - t_fill;
        c->last_dag_prep_parts[2] = t_up - t_meta;
        c->last_dag_prep_parts[3] = tp - t_up;
        if (st) *st = igr_pass_stats{nb, t1 - t0, t2 - t1, n};
    });
}

int igr_dag_pass_overlapped(igr_ctx* c, const int* nets, int64_t n, int mode, const igr_dag_desc* points,
                            igr_dag_complete_fn complete, void* user, igr_pass_stats* st) {
    return guard([&]() {
        check_nets(c, nets, n);
        if (!points || points->n != n || !points->pt_off || !points->pts)
            throw std::invalid_argument("one point list per listed net");
        if (!complete) throw std::invalid_argument("no completion callback");
        if (n == 0) {
            if (st) *st = igr_pass_stats{0, 0, 0, 0};
            return;
        }
        dag_pass_batched(c, nets, n, mode, points, [&](igr_dag_desc& d) {
            const int rc = complete(user, &d);
            if (rc != IGR_OK) throw std::invalid_argument("DAG completion failed with status " + std::to_string(rc));
            // the points are the ones the scheduler already has
            d.n = n;
            d.pt_off = points->pt_off;
            d.pts = points->pts;
            if (!d.node_off || !d.nodes || !d.par_off || !d.par || !d.slot || !d.child_num || !d.depth)
                throw std::invalid_argument("DAG completion left arrays unset");
        }, st);
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

int igr_set_tuning(igr_ctx* c, const igr_tuning* t) {
    return guard([&]() {
        if (!t) throw std::invalid_argument("null tuning");
        if (t->graph < 0 || t->graph > 2) throw std::invalid_argument("graph: 0, 1 or 2");
        if (t->sched_group != 0 && t->sched_group != 4 && t->sched_group != 8 && t->sched_group != 16)
            throw std::invalid_argument("sched_group: 0, 4, 8 or 16");
        if (t->sched_chunk < 0 || t->sched_chunk > 256) throw std::invalid_argument("sched_chunk: 0..256");
        if (t->dag_order < 0 || t->dag_order > 2) throw std::invalid_argument("dag_order: 0, 1 or 2");
        if (t->dag_waves != 0 && t->dag_waves != 2 && t->dag_waves != 4)
            throw std::invalid_argument("dag_waves: 0, 2 or 4");
        if (t->dag_edges < 0 || t->dag_edges > 2) throw std::invalid_argument("dag_edges: 0, 1 or 2");
        if (t->dag_seq < 0) throw std::invalid_argument("dag_seq: >= 0");
        if (t->sched_kind < 0 || t->sched_kind > 2) throw std::invalid_argument("sched_kind: 0, 1 or 2");
        if (t->track_k != 0 && t->track_k != 4 && t->track_k != 8 && t->track_k != 16)
            throw std::invalid_argument("track_k: 0, 4, 8 or 16");
        if (t->track_k && t->track_k * 1024 < std::max(c->g.X, c->g.Y))
            throw std::invalid_argument("track_k too small for this grid's tracks");
        if (t->presum_k != 0 && t->presum_k != 4 && t->presum_k != 8 && t->presum_k != 16)
            throw std::invalid_argument("presum_k: 0, 4, 8 or 16");
        if (t->presum_k && t->presum_k * 1024 < std::max(c->g.X, c->g.Y))
            throw std::invalid_argument("presum_k too small for this grid's tracks");
        if (t->dag_grid < 0) throw std::invalid_argument("dag_grid: >= 0");
        if (t->lshape_groups != 0 && t->lshape_groups != 1 && t->lshape_groups != 2 && t->lshape_groups != 4)
            throw std::invalid_argument("lshape_groups: 0, 1, 2 or 4");
        c->tune = *t;
    });
}

int igr_get_tuning(igr_ctx* c, igr_tuning* t) {
    return guard([&]() {
        if (!t) throw std::invalid_argument("null tuning");
        *t = c->tune;
        if (!t->graph) t->graph = DEF_GRAPH;
        if (!t->sched_chunk) t->sched_chunk = DEF_SCHED_CHUNK;
        if (!t->dag_order) t->dag_order = DEF_DAG_ORDER;
        if (!t->dag_waves) t->dag_waves = DEF_DAG_WAVES;
        if (!t->dag_edges) t->dag_edges = 1;
        if (!t->dag_seq) t->dag_seq = DEF_DAG_SEQ;
        if (!t->track_k) t->track_k = std::max(c->g.X, c->g.Y) <= 4096 ? 4 : 8;
        if (!t->presum_k) t->presum_k = 16;
        if (!t->dag_grid) t->dag_grid = DEF_DAG_GRID;
    });
}

int igr_set_profiling(igr_ctx* c, int on) {
    return guard([&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        c->prof.reset();
        c->prof.on = on != 0;
        HIP_OK(hipMemset(c->track_stats.p, 0, 2 * sizeof(unsigned long long)));
    });
}

int igr_device_memory(igr_ctx* c, size_t* free_bytes, size_t* total_bytes) {
    if (!c || !free_bytes || !total_bytes) return fail(IGR_EINVAL, "null argument");
    return guard([&]() {
        int prev = 0;
        HIP_OK(hipGetDevice(&prev));
        HIP_OK(hipSetDevice(c->device));
        const hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
        HIP_OK(hipSetDevice(prev));
        HIP_OK(e);
    });
}

int igr_kernel_bytes(igr_ctx* c, double* bytes) {
    return guard([&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        for (int k = 0; k < K_NCLASS; k++) bytes[k] = c->prof.bytes[k];
        // fused track pass: 36 B per processed cell (read demand, capacity,
        // timestamp, wire differences, via cuts; write wcost, vcost, presum),
        // 12 B more where a rip-up is applied (its three difference arrays)
        unsigned long long st[2] = {0, 0};
        HIP_OK(hipMemcpy(st, c->track_stats.p, sizeof(st), hipMemcpyDeviceToHost));
        bytes[K_TRACK_FUSED] = 36.0 * (double)st[0] + 12.0 * (double)st[1];
    });
}

int igr_kernel_times(igr_ctx* c, double* ms, int64_t* launches) {
    static_assert(K_NCLASS == IGR_KCLASSES, "kernel class count");
    return guard([&]() {
        HIP_OK(hipStreamSynchronize(c->stream));
        for (int k = 0; k < K_NCLASS; k++) { ms[k] = 0; launches[k] = 0; }
        for (size_t i = 0; i < c->prof.rec.size(); i++) {
            float t = 0;
            HIP_OK(hipEventElapsedTime(&t, c->prof.rec[i].first, c->prof.rec[i].second));
            ms[c->prof.cls[i]] += t;
            launches[c->prof.cls[i]]++;
        }
    });
}

int igr_copy_probe(int64_t bytes, int reps, double* gbs) {
    return guard([&]() {
        hipStream_t s;
        HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            *gbs = copy_probe(bytes, reps, s);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        HIP_OK(hipStreamDestroy(s));
    });
}

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
