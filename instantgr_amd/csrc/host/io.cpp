// .cap / .net parsers and writers.
//
// Formats follow the reference reader (src/database.hpp:247-415) and the
// evaluator reader (run/evaluator.cpp:311-456): the .cap file holds
// "L X Y", unit wire and via costs, L overflow weights, X-1 and Y-1 gcell edge
// lengths, then per layer "metalN dir min_len" followed by Y rows of X
// capacities.  The .net file holds, per net, its name, "(", one line
// "[(l, x, y), ...]" per pin listing its access points, and ")".
// Numbers are unsigned decimal (the reference scanner has no sign or exponent).
#include "design.hpp"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

namespace igr {
namespace {

std::vector<char> slurp(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open '" + path + "'");
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)n + 1);
    if (n > 0 && std::fread(buf.data(), 1, (size_t)n, f) != (size_t)n) {
        std::fclose(f);
        throw std::runtime_error("short read on '" + path + "'");
    }
    std::fclose(f);
    buf[(size_t)n] = 0;
    return buf;
}

struct Scanner {
    const char* p;
    const char* end;
    std::string what;
    void fail(const char* msg) const { throw std::runtime_error(what + ": " + msg); }
    void skip_to_digit() {
        while (p < end && !std::isdigit((unsigned char)*p)) ++p;
        if (p >= end) fail("unexpected end of file");
    }
    // Unsigned decimal field; values above INT_MAX are malformed input.
    int read_int() {
        skip_to_digit();
        long long v = 0;
        while (p < end && std::isdigit((unsigned char)*p)) {
            v = v * 10 + (*p++ - '0');
            if (v > 0x7fffffffLL) fail("integer field out of range");
        }
        return (int)v;
    }
    // Digits with an optional '.', as the reference scanner reads them.
    double read_double() {
        skip_to_digit();
        double ans = 0, scale = 0;
        while (p < end && (std::isdigit((unsigned char)*p) || *p == '.')) {
            if (*p == '.') scale = 1;
            else { scale *= 10; ans = ans * 10 + (*p - '0'); }
            ++p;
        }
        return scale == 0 ? ans : ans / scale;
    }
    std::string read_token() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
        if (p >= end) fail("unexpected end of file");
        const char* b = p;
        while (p < end && *p != ' ' && *p != '\n' && *p != '\r' && *p != '\t') ++p;
        return std::string(b, p);
    }
};

} // namespace

void read_cap(const std::string& path, Design& d) {
    std::vector<char> buf = slurp(path);
    Scanner s{buf.data(), buf.data() + buf.size() - 1, "cap file '" + path + "'"};
    d.L = s.read_int();
    d.X = s.read_int();
    d.Y = s.read_int();
    if (d.L < 2 || d.X < 2 || d.Y < 2) s.fail("grid must have at least 2 layers and 2x2 gcells");
    if ((long long)d.L * d.X * d.Y > 0x7fffffffLL) s.fail("grid too large for 32-bit cell indices");
    d.wire_cost = s.read_double();
    d.via_cost = s.read_double();
    d.of_weight.resize(d.L);
    for (int l = 0; l < d.L; l++) d.of_weight[l] = s.read_double();
    d.xlen.resize(d.X - 1);
    d.ylen.resize(d.Y - 1);
    for (int i = 0; i < d.X - 1; i++) d.xlen[i] = s.read_int();
    for (int i = 0; i < d.Y - 1; i++) d.ylen[i] = s.read_int();
    d.dir.resize(d.L);
    d.min_len.resize(d.L);
    d.cap.assign((size_t)d.L * d.X * d.Y, 0.0);
    for (int l = 0; l < d.L; l++) {
        std::string name = s.read_token();
        if (name != "metal" + std::to_string(l + 1)) s.fail(("unexpected layer name " + name).c_str());
        d.dir[l] = s.read_int();
        if (d.dir[l] != 0 && d.dir[l] != 1) s.fail("layer direction must be 0 or 1");
        d.min_len[l] = s.read_double();
        for (int y = 0; y < d.Y; y++)
            for (int x = 0; x < d.X; x++)
                d.cap[(size_t)l * d.X * d.Y + (size_t)x * d.Y + y] = s.read_double();
        if (l && d.dir[l] == d.dir[l - 1]) s.fail("layer directions must alternate");
    }
}

void read_net(const std::string& path, Design& d) {
    if (d.X == 0) throw std::runtime_error("read the cap file before the net file");
    std::vector<char> buf = slurp(path);
    Scanner s{buf.data(), buf.data() + buf.size() - 1, "net file '" + path + "'"};
    const int XY = d.X * d.Y;
    d.names.clear();
    d.pin_off.assign(1, 0);
    d.ap_off.assign(1, 0);
    d.ap.clear();
    d.bbox.clear();
    auto at_space = [&]() {
        while (s.p < s.end && (*s.p == ' ' || *s.p == '\n' || *s.p == '\r' || *s.p == '\t')) ++s.p;
    };
    at_space();
    while (s.p < s.end) {
        d.names.push_back(s.read_token());
        int minx = d.X, maxx = 0, miny = d.Y, maxy = 0;
        for (;;) {                       // pins
            for (;;) {                   // access points of one pin
                int l = s.read_int(), x = s.read_int(), y = s.read_int();
                if (l < 0 || x < 0 || y < 0 || l >= d.L || x >= d.X || y >= d.Y)
                    s.fail("access point outside the grid");
                minx = std::min(minx, x); maxx = std::max(maxx, x);
                miny = std::min(miny, y); maxy = std::max(maxy, y);
                d.ap.push_back(l * XY + x * d.Y + y);
                bool pin_end = false;
                while (s.p < s.end && !std::isdigit((unsigned char)*s.p))
                    if (*s.p++ == ']') { pin_end = true; break; }
                if (pin_end) break;
                if (s.p >= s.end) s.fail("unterminated pin");
            }
            d.ap_off.push_back((int64_t)d.ap.size());
            bool net_end = false;
            while (s.p < s.end && !std::isdigit((unsigned char)*s.p))
                if (*s.p++ == ')') { net_end = true; break; }
            if (net_end) break;
            if (s.p >= s.end) s.fail("unterminated net");
        }
        d.pin_off.push_back((int64_t)d.ap_off.size() - 1);
        d.bbox.insert(d.bbox.end(), {minx, maxx, miny, maxy});
        at_space();
    }
}

namespace {
struct Out {
    FILE* f;
    std::vector<char> buf;
    size_t n = 0;
    explicit Out(const std::string& path) : f(std::fopen(path.c_str(), "wb")), buf(1 << 22) {
        if (!f) throw std::runtime_error("cannot write '" + path + "'");
    }
    ~Out() { flush(); std::fclose(f); }
    void flush() { if (n) std::fwrite(buf.data(), 1, n, f); n = 0; }
    void put(char c) { if (n == buf.size()) flush(); buf[n++] = c; }
    void str(const char* s) { while (*s) put(*s++); }
    void num(long long v) {
        char t[24]; int k = 0;
        if (v == 0) { put('0'); return; }
        while (v) { t[k++] = char('0' + v % 10); v /= 10; }
        while (k) put(t[--k]);
    }
    void dbl(double v) {
        char t[64];
        std::snprintf(t, sizeof t, "%.6g", v);
        str(t);
    }
};
} // namespace

void write_cap(const std::string& path, const Design& d) {
    Out o(path);
    o.num(d.L); o.put(' '); o.num(d.X); o.put(' '); o.num(d.Y); o.put('\n');
    o.dbl(d.wire_cost); o.put(' '); o.dbl(d.via_cost);
    for (int l = 0; l < d.L; l++) { o.put(' '); o.dbl(d.of_weight[l]); }
    o.put('\n');
    for (int i = 0; i < d.X - 1; i++) { if (i) o.put(' '); o.num(d.xlen[i]); }
    o.put('\n');
    for (int i = 0; i < d.Y - 1; i++) { if (i) o.put(' '); o.num(d.ylen[i]); }
    o.put('\n');
    for (int l = 0; l < d.L; l++) {
        o.str("metal"); o.num(l + 1); o.put(' '); o.num(d.dir[l]); o.put(' '); o.dbl(d.min_len[l]); o.put('\n');
        for (int y = 0; y < d.Y; y++) {
            for (int x = 0; x < d.X; x++) {
                if (x) o.put(' ');
                double c = d.cap[(size_t)l * d.X * d.Y + (size_t)x * d.Y + y];
                if (c == (double)(long long)c) o.num((long long)c); else o.dbl(c);
            }
            o.put('\n');
        }
    }
}

void write_net(const std::string& path, const Design& d) {
    Out o(path);
    const int XY = d.X * d.Y;
    for (int64_t n = 0; n < d.num_nets(); n++) {
        o.str(d.names[n].c_str()); o.str("\n(\n");
        for (int64_t p = d.pin_off[n]; p < d.pin_off[n + 1]; p++) {
            o.put('[');
            for (int64_t a = d.ap_off[p]; a < d.ap_off[p + 1]; a++) {
                int v = d.ap[a];
                if (a > d.ap_off[p]) o.str(", ");
                o.put('('); o.num(v / XY); o.str(", "); o.num(v / d.Y % d.X); o.str(", "); o.num(v % d.Y); o.put(')');
            }
            o.str("]\n");
        }
        o.str(")\n");
    }
}

void build_grid(const Design& d, Grid& g) {
    g.L = d.L - 1;
    g.X = d.X;
    g.Y = d.Y;
    g.DIR = d.dir[1];
    g.wire_cost = d.wire_cost;
    g.via_cost = d.via_cost;
    g.w.assign(d.of_weight.begin() + 1, d.of_weight.end());
    g.xlen = d.xlen;
    g.ylen = d.ylen;
    const size_t XY = (size_t)d.X * d.Y;
    g.cap.resize((size_t)g.L * XY);
    for (int l = 0; l < g.L; l++)
        for (size_t i = 0; i < XY; i++) g.cap[l * XY + i] = (float)d.cap[(l + 1) * XY + i];
}

} // namespace igr
