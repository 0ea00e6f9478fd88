// End-to-end routing stage on one GPU: the reference's route()
// (src/main.cpp:22-73) with the rip-up-and-reroute step repeated.
//
//   trees (host, already built) -> pattern pass -> for each RRR iteration:
//   congestion view -> rip-up marking -> detour DAGs (host threads) ->
//   DAG pass.
//
// The reference runs one RRR iteration and asserts on a second rip-up
// (src/graph.hpp:446-452); here every iteration marks nets over their current
// routes, whichever pass produced them.  'mode' follows src/main.cpp:39-40:
// 1 above 5M subnets, 2 above 20M, set after the pattern pass.
#include "igr_ctx.hpp"
#include "igr_host.h"
#include "dag.hpp"
#include "prep.hpp"

#include <algorithm>
#include <chrono>
#include <numeric>

using namespace igr_dev;

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
} // namespace

int igr_run(igr_ctx* c, igrh_design* h, const igr_run_opts* o, igr_run_stats* st) {
    igr_run_stats S{};
    const int form = o ? o->form : IGR_FORM_PLAIN;
    const int iters = o ? o->rrr_iterations : 1;
    const int threads = o && o->threads > 0 ? o->threads : 8;
    if (!c || !h) return fail(IGR_EINVAL, "null context or design");
    if (form != IGR_FORM_PLAIN && form != IGR_FORM_BATCHED) return fail(IGR_EINVAL, "unknown form");
    if (iters < 0 || iters > IGR_MAX_RRR) return fail(IGR_EINVAL, "rip-up-and-reroute passes outside 0..IGR_MAX_RRR");
    double t0 = now_s();
    const igr::PatternInputs& pi = *igrh_inputs_ref(h);
    const igr::Subnets& sn = *igrh_subnets_ref(h);
    igr_net_desc nd{pi.n, pi.node_off.data(), pi.node.data(), pi.parent.data(), pi.pt_off.data(), pi.pt.data(),
                    pi.hs_off.data(), pi.hs.data(), pi.vs_off.data(), pi.vs.data()};
    int rc = igr_set_nets(c, &nd);
    if (rc) return rc;
    double t1 = now_s();
    S.upload_s = t1 - t0;
    igr_pass_stats ps{};
    rc = igr_pattern_route(c, pi.order.data(), pi.n, form, &ps);
    if (rc) return rc;
    double t2 = now_s();
    S.pattern_s = t2 - t1;
    S.pattern_batches = ps.batches;
    S.pattern_schedule_s = ps.schedule_s;
    progress("pattern pass: %.3f s", S.pattern_s);
    const int mode = pi.n > 20000000 ? 2 : (pi.n > 5000000 ? 1 : 0);
    const int64_t XYc = (int64_t)c->g.X * c->g.Y;
    std::vector<float> xs(XYc), ys(XYc);
    std::vector<uint8_t> is_of(pi.n);
    for (int it = 0; it < iters; it++) {
        double a = now_s();
        if ((rc = igr_congestion_view(c, xs.data(), ys.data()))) return rc;
        if ((rc = igr_ripup_mark(c, 0, is_of.data()))) return rc;
        // The marked subnets by decreasing HPWL, ties in index order: a
        // counting sort over the HPWL values (non-negative ints, at most
        // X + Y), the stable order std::stable_sort would give.
        std::vector<int> nets;
        {
            int hmax = 0;
            int64_t cnt = 0;
            for (int64_t i = 0; i < pi.n; i++)
                if (is_of[i]) {
                    hmax = std::max(hmax, sn.hpwl[i]);
                    cnt++;
                }
            std::vector<int64_t> start((size_t)hmax + 2, 0);
            for (int64_t i = 0; i < pi.n; i++)
                if (is_of[i]) start[(size_t)(hmax - sn.hpwl[i]) + 1]++;
            for (size_t h = 1; h < start.size(); h++) start[h] += start[h - 1];
            nets.resize((size_t)cnt);
            for (int64_t i = 0; i < pi.n; i++)
                if (is_of[i]) nets[(size_t)start[(size_t)(hmax - sn.hpwl[i])]++] = (int)i;
        }
        S.ripped[it] = (int64_t)nets.size();
        if (nets.empty()) break;
        double b = now_s();
        const igr::DagSet& D = *igrh_dags_ref(h);
        double cc;
        if (form == IGR_FORM_BATCHED) {
            // The scheduler needs only the DAGs' points: build those, then
            // the rest of each DAG on the pass's helper thread while the
            // scheduler runs, with three host threads left to the scheduler's
            // launches and the uploads.
            if ((rc = igrh_build_dags_part(h, nets.data(), (int64_t)nets.size(), xs.data(), ys.data(), mode, threads,
                                           IGRH_DAG_POINTS))) {
                return fail(IGR_EINVAL, std::string("detour DAG points: ") + igrh_error());
            }
            cc = now_s();
            struct Completion {
                igrh_design* h;
                const std::vector<int>* nets;
                const float *xs, *ys;
                int mode, threads;
                double secs;
                std::string err;
            } cp{h, &nets, xs.data(), ys.data(), mode, std::max(1, threads - 3), 0.0, std::string()};
            auto complete = [](void* user, igr_dag_desc* d) -> int {
                Completion& x = *static_cast<Completion*>(user);
                const double t = now_s();
                if (igrh_build_dags_part(x.h, x.nets->data(), (int64_t)x.nets->size(), x.xs, x.ys, x.mode, x.threads,
                                         IGRH_DAG_STRUCTURE)) {
                    x.err = igrh_error();
                    return IGR_EINVAL;
                }
                const igr::DagSet& ds = *igrh_dags_ref(x.h);
                d->node_off = ds.node_off.data();
                d->nodes = ds.node.data();
                d->par_off = ds.par_off.data();
                d->par = ds.par.data();
                d->slot = ds.slot.data();
                d->child_num = ds.child_num.data();
                d->depth = ds.depth.data();
                x.secs = now_s() - t;
                return IGR_OK;
            };
            igr_dag_desc pd{D.n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, D.pt_off.data(),
                            D.pt.data()};
            if ((rc = igr_dag_pass_overlapped(c, nets.data(), (int64_t)nets.size(), mode, &pd, complete, &cp, &ps))) {
                if (!cp.err.empty()) return fail(rc, std::string("detour DAGs: ") + cp.err);
                return rc;
            }
            S.rrr_dag_rest_s[it] = cp.secs;
        } else {
            if ((rc = igrh_build_dags(h, nets.data(), (int64_t)nets.size(), xs.data(), ys.data(), mode, threads))) {
                return fail(IGR_EINVAL, std::string("detour DAGs: ") + igrh_error());
            }
            cc = now_s();
            igr_dag_desc dd{D.n, D.node_off.data(), D.node.data(), D.par_off.data(), D.par.data(), D.slot.data(),
                            D.child_num.data(), D.depth.data(), D.pt_off.data(), D.pt.data()};
            if ((rc = igr_dag_pass(c, nets.data(), (int64_t)nets.size(), mode, form, &dd, &ps))) return rc;
        }
        double e = now_s();
        S.rrr_view_s[it] = b - a;
        S.rrr_dag_build_s[it] = cc - b;
        S.rrr_route_s[it] = e - cc;
        S.rrr_schedule_s[it] = ps.schedule_s;
        S.rrr_prep_s[it] = c->last_dag_prep_s;
        for (int k = 0; k < 4; k++) S.rrr_prep_parts[it][k] = c->last_dag_prep_parts[k];
        S.rrr_batches[it] = ps.batches;
        S.rrr_dag_nodes[it] = D.node_off.empty() ? 0 : D.node_off.back();
        S.iterations = it + 1;
        progress("rip-up-and-reroute pass %d: %.3f s (view %.3f, DAG build %.3f, behind the scheduler %.3f, %lld nets)",
                 it, e - a, b - a, cc - b, S.rrr_dag_rest_s[it], (long long)nets.size());
    }
    S.total_s = now_s() - t0;
    S.subnets = pi.n;
    if (st) *st = S;
    return IGR_OK;
}
