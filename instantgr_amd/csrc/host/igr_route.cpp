// Command-line routing stage, a drop-in for the reference binary
// (README.md:20-22, src/main.cpp:91-102):
//   igr_route -cap <file> -net <file> -out <file> [-form plain|batched]
//             [-rrr N] [-threads T]
// Reads the .cap/.net files, builds the Steiner trees, routes on the GPU
// (pattern pass plus N rip-up-and-reroute passes; the default 1 is what the
// reference's route() runs, src/main.cpp:33-71) and writes the route file in
// the reference format.  Exits non-zero on any error, 2 on a bad command line.
#include "igr.h"
#include "igr_host.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    std::string cap, net, out;
    igr_run_opts opts{IGR_FORM_BATCHED, 1, 8};
    auto number = [](const std::string& v, int lo, int hi, int* out) {
        char* end = nullptr;
        const long x = std::strtol(v.c_str(), &end, 10);
        if (v.empty() || *end || x < lo || x > hi) return false;
        *out = (int)x;
        return true;
    };
    const char* usage = "usage: %s -cap <file> -net <file> -out <file> [-form plain|batched] [-rrr 0..%d] "
                        "[-threads 1..256]\n";
    if (argc % 2 != 1) {
        std::fprintf(stderr, usage, argv[0], IGR_MAX_RRR);
        return 2;
    }
    for (int i = 1; i + 1 < argc; i += 2) {
        std::string k = argv[i], v = argv[i + 1];
        bool ok = true;
        if (k == "-cap") cap = v;
        else if (k == "-net") net = v;
        else if (k == "-out") out = v;
        else if (k == "-form") {
            ok = v == "plain" || v == "batched";
            opts.form = v == "plain" ? IGR_FORM_PLAIN : IGR_FORM_BATCHED;
        } else if (k == "-rrr") ok = number(v, 0, IGR_MAX_RRR, &opts.rrr_iterations);
        else if (k == "-threads") ok = number(v, 1, 256, &opts.threads);
        else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
        if (!ok) {
            std::fprintf(stderr, "bad value '%s' for %s\n", v.c_str(), k.c_str());
            return 2;
        }
    }
    if (cap.empty() || net.empty() || out.empty()) {
        std::fprintf(stderr, usage, argv[0], IGR_MAX_RRR);
        return 2;
    }
    double t0 = now_s();
    igrh_design* h = igrh_load(cap.c_str(), net.c_str());
    if (!h) {
        std::fprintf(stderr, "input: %s\n", igrh_error());
        return 1;
    }
    double t1 = now_s();
    if (igrh_build(h, opts.threads)) {
        std::fprintf(stderr, "trees: %s\n", igrh_error());
        return 1;
    }
    double t2 = now_s();
    igrh_grid_info gi;
    igrh_grid(h, &gi);
    igr_grid_desc gd{gi.L, gi.X, gi.Y, gi.DIR, gi.unit_wire, gi.unit_via, gi.w, gi.xlen, gi.ylen, gi.cap};
    igr_ctx* ctx = nullptr;
    if (igr_create(&gd, &ctx)) {
        std::fprintf(stderr, "gpu: %s\n", igr_last_error());
        return 1;
    }
    igr_run_stats st;
    if (igr_run(ctx, h, &opts, &st)) {
        std::fprintf(stderr, "route: %s\n", igr_last_error());
        return 1;
    }
    double t3 = now_s();
    int64_t cap_total = igr_routes_capacity(ctx);
    std::vector<int64_t> off(gi.subnets + 1);
    std::vector<int> data(cap_total > 0 ? cap_total : 1);
    if (igr_get_routes(ctx, off.data(), data.data(), (int64_t)data.size()) ||
        igrh_write_routes(h, out.c_str(), off.data(), data.data(), (int64_t)data.size())) {
        std::fprintf(stderr, "output: %s %s\n", igr_last_error(), igrh_error());
        return 1;
    }
    double t4 = now_s();
    std::printf("nets %lld subnets %lld | input %.2fs trees %.2fs route %.2fs (pattern %.2fs, %d batches",
                (long long)gi.nets, (long long)gi.subnets, t1 - t0, t2 - t1, t3 - t2, st.pattern_s, st.pattern_batches);
    for (int i = 0; i < st.iterations; i++)
        std::printf("; rrr%d %lld nets %.2fs", i + 1, (long long)st.ripped[i],
                    st.rrr_view_s[i] + st.rrr_dag_build_s[i] + st.rrr_route_s[i]);
    std::printf(") output %.2fs total %.2fs\n", t4 - t3, t4 - t0);
    igr_destroy(ctx);
    igrh_free(h);
    return 0;
}
