// Serving loop: a routed design stays resident on the device and takes
// reroute requests (lists of subnets) as they arrive.
//
// One worker thread owns the context.  Each tick it
//   1. takes the congestion view of the current state,
//   2. builds the detour DAGs of the pending requests at the head of the queue
//      (up to twice the tick budget in nets),
//   3. admits requests in arrival order while the tick's net and scratch
//      budgets allow and each new request's footprint is disjoint from the
//      admitted ones'; by default admission stops at the first request that
//      cannot join, so the ticks route the requests in arrival order and the
//      result is that of routing them one at a time.  With 'overtake', a
//      later request may pass a waiting one whose footprint it does not meet
//      (the waiting request's DAGs are rebuilt next tick and may then reach
//      further, so this trades the one-at-a-time equivalence for admission),
//   4. runs one batched rip-up-and-reroute pass over the admitted requests'
//      subnets together: their nets share the scheduler's batches and the
//      routing launches.
// A footprint is the set of 16x16-gcell tiles covered by the request's
// subnets: their current routes, their DAG nodes and scheduler points, one
// gcell around.  Requests admitted together have disjoint footprints, so no
// subnet of one reads a cell another writes, and each request's batches and
// routes are the ones it gets alone from the same state (up to the rounding
// of the FP64 track prefix sums, which span the whole track).
// Requests that arrive while a tick runs wait for the next one; their DAGs
// are built from the state that tick leaves.
#include "igr_ctx.hpp"
#include "igr_host.h"
#include "dag.hpp"
#include "design.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>

using namespace igr_dev;

namespace {

constexpr int TILE = 16;
constexpr int TPB = 256;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Bounding box (x0, x1, y0, y1) of each listed net's current route; an empty
// route gives x0 > x1.
__global__ void k_route_bbox(const int* nets, int64_t n, int* const* rptr, int XY, int Y, int* box) {
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    const int* r = rptr[nets[t]];
    int x0 = 0x7fffffff, x1 = -1, y0 = 0x7fffffff, y1 = -1;
    for (int i = 1; i < r[0]; i++) {
        const int p = r[i] % XY, x = p / Y, y = p % Y;
        x0 = min(x0, x); x1 = max(x1, x);
        y0 = min(y0, y); y1 = max(y1, y);
    }
    box[4 * t] = x0;
    box[4 * t + 1] = x1;
    box[4 * t + 2] = y0;
    box[4 * t + 3] = y1;
}

struct Request {
    int64_t id = 0;
    std::vector<int> nets;
    double t_submit = 0, t_start = 0, t_done = 0;
    int tick = -1;
    int64_t tick_nets = 0;
    bool done = false;
    int status = IGR_OK;
    std::string err;
    // per tick: the request's DAGs inside the candidate DagSet and its tiles
    int64_t first = 0;
    std::vector<int> tiles;
    int64_t scratch = 0;
};

} // namespace

struct igr_server {
    igr_ctx* c = nullptr;
    igrh_design* h = nullptr;
    igr_serve_opts o{};
    int mode = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
    int64_t next_id = 1;
    bool stop = false;
    int failed = 0;
    std::string fail_msg;
    igr_server_stats st{};
    std::thread worker;

    void run();
    void tick(std::vector<std::shared_ptr<Request>>& cand);
};

void igr_server::tick(std::vector<std::shared_ptr<Request>>& cand) {
    const igr::Grid& g = *igrh_grid_ref(h);
    const igr::Subnets& sn = *igrh_subnets_ref(h);
    const int X = g.X, Y = g.Y, XY = X * Y;
    const int TX = (X + TILE - 1) / TILE, TY = (Y + TILE - 1) / TILE;
    const double t0 = now_s();
    // 1-2. view and DAGs of every candidate's nets, in candidate order
    std::vector<float> xs((size_t)XY), ys((size_t)XY);
    int rc = igr_congestion_view(c, xs.data(), ys.data());
    if (rc) throw std::runtime_error(std::string("congestion view: ") + igr_last_error());
    std::vector<int> all;
    for (auto& r : cand) {
        r->first = (int64_t)all.size();
        all.insert(all.end(), r->nets.begin(), r->nets.end());
    }
    const int64_t na = (int64_t)all.size();
    if (igrh_build_dags(h, all.data(), na, xs.data(), ys.data(), mode, std::max(1, o.threads)))
        throw std::runtime_error(std::string("DAG build: ") + igrh_error());
    const igr::DagSet& D = *igrh_dags_ref(h);
    // current route boxes
    std::vector<int> box((size_t)(4 * std::max<int64_t>(1, na)));
    {
        DevBuf<int> d_nets, d_box;
        d_nets.upload(all.data(), na, c->stream);
        d_box.alloc(4 * std::max<int64_t>(1, na));
        k_route_bbox<<<(unsigned)((na + TPB - 1) / TPB), TPB, 0, c->stream>>>(d_nets.p, na, c->nets.rptr.p, XY, Y,
                                                                               d_box.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(box.data(), d_box.p, sizeof(int) * 4 * na, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    // footprints (tiles) and scratch estimates per candidate
    const int ML = g.L + 1;
    for (auto& r : cand) {
        std::vector<char> mark((size_t)TX * TY, 0);
        r->tiles.clear();
        r->scratch = 0;
        for (size_t k = 0; k < r->nets.size(); k++) {
            const int64_t q = r->first + (int64_t)k;
            int x0 = box[4 * q], x1 = box[4 * q + 1], y0 = box[4 * q + 2], y1 = box[4 * q + 3];
            for (int64_t v = D.node_off[q]; v < D.node_off[q + 1]; v++) {
                const int p = D.node[v] % XY, x = p / Y, y = p % Y;
                x0 = std::min(x0, x); x1 = std::max(x1, x);
                y0 = std::min(y0, y); y1 = std::max(y1, y);
            }
            for (int64_t v = D.pt_off[q]; v < D.pt_off[q + 1]; v++) {
                const int p = D.pt[v], x = p / Y, y = p % Y;
                x0 = std::min(x0, x); x1 = std::max(x1, x);
                y0 = std::min(y0, y); y1 = std::max(y1, y);
            }
            if (x0 > x1) continue;
            x0 = std::max(0, x0 - 1); x1 = std::min(X - 1, x1 + 1);
            y0 = std::max(0, y0 - 1); y1 = std::min(Y - 1, y1 + 1);
            for (int tx = x0 / TILE; tx <= x1 / TILE; tx++)
                for (int ty = y0 / TILE; ty <= y1 / TILE; ty++) {
                    char& m = mark[(size_t)tx * TY + ty];
                    if (!m) { m = 1; r->tiles.push_back(tx * TY + ty); }
                }
            // device tables per DAG node and slot (layer costs, keys,
            // back-pointers, route slot) and per edge (layer-change costs)
            const int64_t nn = D.node_off[q + 1] - D.node_off[q];
            const int64_t ne = D.par_off[D.node_off[q + 1]] - D.par_off[D.node_off[q]];
            r->scratch += nn * (8 * ML + 16 * ML * 2 + 40) + ne * (12 * (ML - 1) * (ML - 1) + 24);
        }
    }
    // 3. admission
    std::vector<char> claimed((size_t)TX * TY, 0);
    std::vector<std::shared_ptr<Request>> adm;
    int64_t nets_in = 0, scratch_in = 0;
    for (auto& r : cand) {
        bool meets = false;
        for (int t : r->tiles) meets = meets || claimed[t];
        const int64_t nn = (int64_t)r->nets.size();
        const bool fits = adm.empty() || ((o.max_tick_nets <= 0 || nets_in + nn <= o.max_tick_nets) &&
                                          (o.scratch_bytes <= 0 || scratch_in + r->scratch <= o.scratch_bytes));
        if (!meets && fits) {
            adm.push_back(r);
            nets_in += nn;
            scratch_in += r->scratch;
        } else if (!o.overtake) {
            break;   // arrival order: nothing passes a waiting request
        }
        for (int t : r->tiles) claimed[t] = 1;
    }
    // 4. one pass over the admitted subnets, by decreasing hpwl (stable in
    // arrival order), with their DAGs
    std::vector<std::pair<int, int64_t>> order;   // (net, index in 'all')
    for (auto& r : adm)
        for (size_t k = 0; k < r->nets.size(); k++) order.emplace_back(r->nets[k], r->first + (int64_t)k);
    std::stable_sort(order.begin(), order.end(),
                     [&](const std::pair<int, int64_t>& a, const std::pair<int, int64_t>& b) {
                         return sn.hpwl[a.first] > sn.hpwl[b.first];
                     });
    const int64_t n = (int64_t)order.size();
    std::vector<int> nets(n);
    std::vector<int64_t> node_off(n + 1, 0), par_off(1, 0), pt_off(n + 1, 0);
    std::vector<int> node, par, slot, cnum, depth, pt;
    for (int64_t i = 0; i < n; i++) {
        const int64_t q = order[i].second;
        nets[i] = order[i].first;
        const int64_t v0 = D.node_off[q], v1 = D.node_off[q + 1];
        for (int64_t v = v0; v < v1; v++) {
            node.push_back(D.node[v]);
            cnum.push_back(D.child_num[v]);
            depth.push_back(D.depth[v]);
            for (int64_t e = D.par_off[v]; e < D.par_off[v + 1]; e++) {
                par.push_back(D.par[e]);
                slot.push_back(D.slot[e]);
            }
            par_off.push_back((int64_t)par.size());
        }
        node_off[i + 1] = (int64_t)node.size();
        pt.insert(pt.end(), D.pt.begin() + D.pt_off[q], D.pt.begin() + D.pt_off[q + 1]);
        pt_off[i + 1] = (int64_t)pt.size();
    }
    igr_dag_desc dd{n, node_off.data(), node.data(), par_off.data(), par.data(), slot.data(),
                    cnum.data(), depth.data(), pt_off.data(), pt.data()};
    igr_pass_stats ps{};
    const double t1 = now_s();
    rc = igr_dag_pass(c, nets.data(), n, mode, IGR_FORM_BATCHED, &dd, &ps);
    if (rc) throw std::runtime_error(std::string("reroute pass: ") + igr_last_error());
    const double t2 = now_s();
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& r : adm) {
            r->t_start = t0;
            r->t_done = t2;
            r->tick = (int)st.ticks;
            r->tick_nets = n;
            r->done = true;
            pending.erase(std::find(pending.begin(), pending.end(), r));
        }
        st.ticks++;
        st.requests += (int64_t)adm.size();
        st.nets += n;
        st.prep_s += t1 - t0;
        st.route_s += t2 - t1;
        st.max_tick_nets = std::max(st.max_tick_nets, n);
    }
    cv_done.notify_all();
}

void igr_server::run() {
    if (hipSetDevice(c->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(mu);
        failed = 1;
        fail_msg = "cannot select the context's device";
        return;
    }
    for (;;) {
        std::vector<std::shared_ptr<Request>> cand;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            if (pending.empty()) return;   // stop with nothing left
            if (o.hold_ms > 0 && !stop) {
                const double until = pending.front()->t_submit + 1e-3 * o.hold_ms;
                const double wait = until - now_s();
                if (wait > 0) cv_work.wait_for(lk, std::chrono::duration<double>(wait), [&] { return stop; });
            }
            int64_t nets = 0;
            const int64_t cap = o.max_tick_nets > 0 ? 2 * o.max_tick_nets : INT64_MAX;
            for (auto& r : pending) {
                if (!cand.empty() && nets + (int64_t)r->nets.size() > cap) break;
                cand.push_back(r);
                nets += (int64_t)r->nets.size();
            }
        }
        try {
            tick(cand);
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            failed = 1;
            fail_msg = e.what();
            for (auto& r : pending) {
                r->done = true;
                r->status = IGR_EHIP;
                r->err = fail_msg;
            }
            pending.clear();
            cv_done.notify_all();
            return;
        }
    }
}

namespace {
thread_local std::string s_err;

template <typename F>
int sguard(F&& f) {
    try {
        f();
        return IGR_OK;
    } catch (const std::invalid_argument& e) {
        s_err = e.what();
        return IGR_EINVAL;
    } catch (const std::logic_error& e) {
        s_err = e.what();
        return IGR_ESTATE;
    } catch (const std::exception& e) {
        s_err = e.what();
        return IGR_EHIP;
    }
}
} // namespace

const char* igr_server_error(void) { return s_err.c_str(); }

int igr_server_open(igr_ctx* c, igrh_design* h, const igr_serve_opts* o, igr_server** out) {
    return sguard([&]() {
        if (!c || !h || !out) throw std::invalid_argument("null argument");
        const igr::Subnets* sn = igrh_subnets_ref(h);
        if (!sn || sn->count != c->nets.n) throw std::invalid_argument("design and context hold different subnets");
        auto* s = new igr_server();
        s->c = c;
        s->h = h;
        if (o) s->o = *o;
        if (s->o.threads <= 0) s->o.threads = 8;
        s->mode = sn->count > 20000000 ? 2 : (sn->count > 5000000 ? 1 : 0);
        s->worker = std::thread([s]() { s->run(); });
        *out = s;
    });
}

int igr_server_submit(igr_server* s, const int* nets, int64_t n, int64_t* id) {
    return sguard([&]() {
        if (!s || (!nets && n > 0) || n < 0) throw std::invalid_argument("bad request");
        if (n == 0) throw std::invalid_argument("empty request");
        std::vector<int> v(nets, nets + n);
        for (int x : v)
            if (x < 0 || x >= s->c->nets.n) throw std::invalid_argument("net id out of range");
        std::sort(v.begin(), v.end());
        if (std::adjacent_find(v.begin(), v.end()) != v.end()) throw std::invalid_argument("repeated net in a request");
        auto r = std::make_shared<Request>();
        r->nets.assign(nets, nets + n);
        r->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->failed) throw std::runtime_error("server failed: " + s->fail_msg);
            if (s->stop) throw std::logic_error("server closing");
            r->id = s->next_id++;
            s->reqs[r->id] = r;
            s->pending.push_back(r);
        }
        s->cv_work.notify_one();
        if (id) *id = r->id;
    });
}

int igr_server_wait(igr_server* s, int64_t id, igr_request_result* out) {
    std::shared_ptr<Request> r;
    int rc = sguard([&]() {
        if (!s) throw std::invalid_argument("null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) throw std::invalid_argument("unknown request id");
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->reqs.erase(id);
    });
    if (rc) return rc;
    if (out) {
        out->id = r->id;
        out->nets = (int64_t)r->nets.size();
        out->submit_s = r->t_submit;
        out->start_s = r->t_start;
        out->done_s = r->t_done;
        out->latency_s = r->t_done - r->t_submit;
        out->tick = r->tick;
        out->tick_nets = r->tick_nets;
    }
    if (r->status) s_err = r->err;
    return r->status;
}

int igr_server_stats_get(igr_server* s, igr_server_stats* out) {
    return sguard([&]() {
        if (!s || !out) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->st;
    });
}

int igr_server_close(igr_server* s) {
    if (!s) return IGR_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    const int rc = s->failed ? IGR_EHIP : IGR_OK;
    if (s->failed) s_err = s->fail_msg;
    delete s;
    return rc;
}
