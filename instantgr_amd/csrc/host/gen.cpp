// Seeded synthetic design generator for the five benchmark configurations.
//
// Own PRNG (splitmix64 seeding a xoshiro256** stream) and own distributions, so
// the same seed gives the same design on every machine: geometric(p) by
// Bernoulli trials, an approximately normal variate as the Irwin-Hall sum of
// twelve uniforms minus six, uniform integers by 128-bit multiply-shift.
#include "design.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace igr {
namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& v : s) {  // splitmix64
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
            v = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // Uniform integer in [a, b].
    int range(int a, int b) {
        uint64_t span = (uint64_t)(b - a) + 1;
        return a + (int)(((unsigned __int128)next() * span) >> 64);
    }
    int geometric(double p) {  // failures before the first success
        int k = 0;
        while (uniform() >= p) k++;
        return k;
    }
    double normal() {
        double acc = 0;
        for (int i = 0; i < 12; i++) acc += uniform();
        return acc - 6.0;
    }
};

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

} // namespace

GenConfig preset_config(int idx) {
    GenConfig c;
    switch (idx) {
    case 0: c.L = 10; c.X = c.Y = 200; c.nets = 20000; c.pin_cap = 30; c.seed = 1; break;
    case 1: case 4: c.L = 10; c.X = c.Y = 800; c.nets = 500000; c.pin_cap = 64; c.seed = 2; break;
    case 2:
        c.L = 10; c.X = c.Y = 1500; c.nets = 3000000; c.pin_cap = 64; c.seed = 3;
        c.heavy_frac = 0.03; c.heavy_min = 100; c.heavy_max = 1000; c.heavy_window = 300;
        c.hotspots = 20; c.hotspot_scale = 0.7;
        break;
    case 3: c.L = 12; c.X = c.Y = 2500; c.nets = 8000000; c.pin_cap = 64; c.seed = 4; break;
    default: throw std::runtime_error("configuration index must be 0..4");
    }
    return c;
}

void generate_design(const GenConfig& c, Design& d) {
    if (c.L < 3 || c.X < 2 || c.Y < 2) throw std::runtime_error("generator: grid too small");
    Rng rng(c.seed);
    d = Design();
    d.L = c.L; d.X = c.X; d.Y = c.Y;
    d.wire_cost = 1.0;
    d.via_cost = 4.0;
    d.of_weight.assign(c.L, 1.0);
    d.xlen.assign(c.X - 1, 1);
    d.ylen.assign(c.Y - 1, 1);
    d.dir.resize(c.L);
    for (int l = 0; l < c.L; l++) d.dir[l] = l & 1;  // metal1 horizontal, alternating
    d.min_len.assign(c.L, 1.0);
    const size_t XY = (size_t)c.X * c.Y;
    d.cap.assign((size_t)c.L * XY, 0.0);
    for (int l = 1; l < c.L; l++)
        for (size_t i = 0; i < XY; i++) d.cap[l * XY + i] = rng.range(2, 6);

    // Blockages: rectangles on random routing layers zeroing capacity until the
    // requested fraction of edges is blocked.
    const size_t target = (size_t)(c.blockage_frac * (double)((c.L - 1) * XY));
    size_t zeroed = 0;
    const int rmax = std::max(4, std::min(c.X, c.Y) / 10);
    while (zeroed < target) {
        int l = rng.range(1, c.L - 1);
        int w = rng.range(4, rmax), h = rng.range(4, rmax);
        int x0 = rng.range(0, c.X - 1), y0 = rng.range(0, c.Y - 1);
        for (int x = x0; x < std::min(c.X, x0 + w) && zeroed < target; x++)
            for (int y = y0; y < std::min(c.Y, y0 + h) && zeroed < target; y++) {
                double& v = d.cap[l * XY + (size_t)x * c.Y + y];
                if (v != 0.0) { v = 0.0; zeroed++; }
            }
    }
    // Hotspots: capacity scaled down (floored) on every routing layer.
    for (int k = 0; k < c.hotspots; k++) {
        int w = rng.range(std::max(2, c.X / 20), std::max(3, c.X / 8));
        int h = rng.range(std::max(2, c.Y / 20), std::max(3, c.Y / 8));
        int x0 = rng.range(0, c.X - 1), y0 = rng.range(0, c.Y - 1);
        for (int l = 1; l < c.L; l++)
            for (int x = x0; x < std::min(c.X, x0 + w); x++)
                for (int y = y0; y < std::min(c.Y, y0 + h); y++) {
                    double& v = d.cap[l * XY + (size_t)x * c.Y + y];
                    v = std::floor(v * c.hotspot_scale);
                }
    }

    // Nets.
    d.names.resize(c.nets);
    d.pin_off.assign(1, 0);
    d.ap_off.assign(1, 0);
    d.ap.clear();
    d.ap.reserve((size_t)c.nets * 4);
    d.bbox.reserve((size_t)c.nets * 4);
    for (int64_t n = 0; n < c.nets; n++) {
        d.names[n] = "net" + std::to_string(n);
        bool heavy = c.heavy_frac > 0 && rng.uniform() < c.heavy_frac;
        int k;
        int minx = c.X, maxx = 0, miny = c.Y, maxy = 0;
        if (heavy) {
            k = rng.range(c.heavy_min, c.heavy_max);
            int wx = std::min(c.heavy_window, c.X), wy = std::min(c.heavy_window, c.Y);
            int x0 = rng.range(0, c.X - wx), y0 = rng.range(0, c.Y - wy);
            for (int i = 0; i < k; i++) {
                int x, y, l, v;
                do {  // distinct 3-D positions within a net
                    x = x0 + rng.range(0, wx - 1); y = y0 + rng.range(0, wy - 1); l = rng.range(0, 1);
                    v = (int)(l * XY + (size_t)x * c.Y + y);
                } while (std::find(d.ap.end() - i, d.ap.end(), v) != d.ap.end());
                d.ap.push_back(v);
                d.ap_off.push_back((int64_t)d.ap.size());
                minx = std::min(minx, x); maxx = std::max(maxx, x);
                miny = std::min(miny, y); maxy = std::max(maxy, y);
            }
        } else {
            k = std::min(c.pin_cap, 2 + rng.geometric(c.geom_p));
            double cx = rng.uniform() * (c.X - 1), cy = rng.uniform() * (c.Y - 1);
            for (int i = 0; i < k; i++) {
                int x, y, l, v, tries = 0;
                do {  // distinct 3-D positions within a net
                    x = clampi((int)std::floor(cx + c.sigma * rng.normal() + 0.5), 0, c.X - 1);
                    y = clampi((int)std::floor(cy + c.sigma * rng.normal() + 0.5), 0, c.Y - 1);
                    l = rng.range(0, 1);
                    v = (int)(l * XY + (size_t)x * c.Y + y);
                } while (std::find(d.ap.end() - i, d.ap.end(), v) != d.ap.end() && ++tries < 1000);
                d.ap.push_back(v);
                d.ap_off.push_back((int64_t)d.ap.size());
                minx = std::min(minx, x); maxx = std::max(maxx, x);
                miny = std::min(miny, y); maxy = std::max(maxy, y);
            }
        }
        d.pin_off.push_back((int64_t)d.ap_off.size() - 1);
        d.bbox.insert(d.bbox.end(), {minx, maxx, miny, maxy});
    }
}

} // namespace igr
