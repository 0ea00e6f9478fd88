// Drives the host library under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by tests/test_host_sanitize.py): generation, a .cap/.net save and
// reload, trees and subnets, detour DAGs in all three modes over a synthetic
// congestion view, the route writer and the host evaluator, on the
// configuration-0 generator and on the configuration-2 generator's heavy nets.
#include "igr_host.h"

#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const std::string cap = dir + "/s.cap", net = dir + "/s.net", out = dir + "/s.out";
    for (int preset : {0, 2}) {
        igrh_design* g = igrh_generate(preset, preset == 2 ? 3000 : 4000, 0);
        if (!g || igrh_save(g, cap.c_str(), net.c_str())) { printf("generate/save: %s\n", igrh_error()); return 1; }
        igrh_free(g);
        igrh_design* h = igrh_load(cap.c_str(), net.c_str());
        if (!h || igrh_build(h, 4)) { printf("load/build: %s\n", igrh_error()); return 1; }
        igrh_grid_info gi;
        igrh_grid(h, &gi);
        const int X = gi.X, Y = gi.Y;
        std::vector<float> xs((size_t)X * Y), ys((size_t)X * Y);
        std::mt19937 rng(1);
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++) {
                const size_t k = (size_t)x * Y + y;
                const float v = (float)(rng() % 1000) / 500.0f - 0.5f;
                xs[k] = v + (x ? xs[k - Y] : 0.0f);
                ys[k] = v + (y ? ys[k - 1] : 0.0f);
            }
        std::vector<int> nets(gi.subnets);
        for (int64_t i = 0; i < gi.subnets; i++) nets[i] = (int)i;
        for (int mode = 0; mode < 3; mode++) {
            if (igrh_build_dags(h, nets.data(), (int64_t)nets.size(), xs.data(), ys.data(), mode, 4)) {
                printf("dags: %s\n", igrh_error());
                return 1;
            }
            // the same DAGs in two steps, points then structure
            if (igrh_build_dags_part(h, nets.data(), (int64_t)nets.size(), xs.data(), ys.data(), mode, 3,
                                     IGRH_DAG_POINTS) ||
                igrh_build_dags_part(h, nets.data(), (int64_t)nets.size(), xs.data(), ys.data(), mode, 2,
                                     IGRH_DAG_STRUCTURE)) {
                printf("dags in two steps: %s\n", igrh_error());
                return 1;
            }
        }
        // a structure build that does not follow the points of as many nets
        if (!igrh_build_dags_part(h, nets.data(), (int64_t)nets.size() - 1, xs.data(), ys.data(), 0, 2,
                                  IGRH_DAG_STRUCTURE)) {
            printf("structure build without its points accepted\n");
            return 1;
        }
        std::vector<int64_t> off(gi.subnets + 1);
        std::vector<int> data;
        for (int64_t i = 0; i < gi.subnets; i++) {
            off[i] = (int64_t)data.size();
            data.push_back(1);   // no route lines: every multi-pin net is open
        }
        off[gi.subnets] = (int64_t)data.size();
        igrh_score sc;
        if (igrh_write_routes(h, out.c_str(), off.data(), data.data(), (int64_t)data.size())) {
            printf("write: %s\n", igrh_error());
            return 1;
        }
        igrh_evaluate(h, out.c_str(), &sc);
        printf("preset %d: %lld subnets, %lld open nets\n", preset, (long long)gi.subnets, (long long)sc.opens);
        // Malformed lines inside the first net: outside the grid on every
        // axis, negative, a via past the top layer, a via written top-down.
        {
            std::string text;
            if (FILE* f = std::fopen(out.c_str(), "rb")) {
                char buf[1 << 16];
                size_t k;
                while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
                std::fclose(f);
            }
            const size_t at = text.find("(\n");
            if (at == std::string::npos) { printf("route file without a net block\n"); return 1; }
            char bad[512];
            std::snprintf(bad, sizeof bad,
                          "%d 0 1 %d 0 1\n0 %d 2 0 %d 2\n-1 0 1 3 0 1\n0 0 -2 0 0 1\n0 0 1 0 0 %d\n3 3 4 3 3 2\n"
                          "0 0 %d 0 0 %d\n",
                          X, X + 5, Y, Y + 3, gi.metal_layers + 2, gi.metal_layers, gi.metal_layers + 1);
            text.insert(at + 2, bad);
            const std::string out2 = dir + "/s_bad.out";
            FILE* f = std::fopen(out2.c_str(), "wb");
            std::fwrite(text.data(), 1, text.size(), f);
            std::fclose(f);
            igrh_score sb;
            igrh_evaluate(h, out2.c_str(), &sb);
            if (sb.opens < 1) { printf("malformed lines did not open the net\n"); return 1; }
        }
        igrh_free(h);
    }
    printf("clean\n");
    return 0;
}
