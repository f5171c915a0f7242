// Host-only checks of the library's CPU-side code, built by tests/test_host_sanitizers.py with AddressSanitizer
// and UndefinedBehaviorSanitizer (g++, no HIP): the tile-table parser (csrc/tune_parse.hpp) and the serving
// loop's arena allocator and batch-latency model (csrc/serve_host.hpp).
//
//   host_checks <committed tune_table.txt> <scratch directory>
//
// Prints one line per check group and exits 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "serve_host.hpp"
#include "tune_parse.hpp"

namespace {

int g_fail = 0;

void check(bool c, const char *what, int line) {
    if (!c) {
        std::fprintf(stderr, "FAILED (line %d): %s\n", line, what);
        ++g_fail;
    }
}
#define CHECK(c) check((c), #c, __LINE__)

int parse_text(const std::string &dir, const std::string &text, pe::tune::Table &t) {
    const std::string path = dir + "/table.txt";
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return -100;
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
    f = std::fopen(path.c_str(), "r");
    const int n = pe::tune::parse(f, t);
    std::fclose(f);
    return n;
}

void tune_checks(const char *committed, const std::string &dir) {
    pe::tune::Table t;
    // the committed table: every coef / mmse / solve line is an entry
    FILE *f = std::fopen(committed, "r");
    CHECK(f != nullptr);
    if (f) {
        int lines = 0;
        char buf[512];
        while (std::fgets(buf, sizeof(buf), f))
            if (!std::strncmp(buf, "coef ", 5) || !std::strncmp(buf, "mmse ", 5) || !std::strncmp(buf, "solve ", 6))
                ++lines;
        std::rewind(f);
        CHECK(pe::tune::parse(f, t) == lines);
        std::fclose(f);
    }
    // fields, defaults and malformed lines
    CHECK(parse_text(dir, "# comment\ncoef 8 11 64 1 16 4 2\nmmse 8 64 4096 16 8 2 1\nsolve 8 64 -1 64 4\n"
                          "slot 32 4 1\n", t) == 3);
    CHECK(t.coef.size() == 1 && t.coef[0].f32_kc == 16 && t.coef[0].mfma_waves == 4 && t.coef[0].mfma_spw == 2);
    CHECK(t.mmse.size() == 1 && t.mmse[0].lanes == 16 && t.mmse[0].tile == 8 && t.mmse[0].waves == 2 &&
          t.mmse[0].keep == 1);
    CHECK(t.solve.size() == 1 && t.solve[0].tile == 64 && t.solve[0].waves == 4);
    CHECK(t.slot.mmse_tile == 32 && t.slot.mmse_waves == 4 && t.slot.mmse_keep == 1);
    CHECK(parse_text(dir, "coef 1 2 3\nmmse 8 64\nsolve 8 64 -1 64\nsolve 8 64 -1 300 4\nbogus\n\n", t) == 0);
    CHECK(parse_text(dir, "coef 8 11 64 1 7 3 99\nmmse 8 64 -1 5 999 3 7\n", t) == 2);
    CHECK(t.coef[0].f32_kc == 8 && t.coef[0].mfma_waves == 1 && t.coef[0].mfma_spw == 1);
    CHECK(t.mmse[0].lanes == 8 && t.mmse[0].tile == 0 && t.mmse[0].waves == 1 && t.mmse[0].keep == 0);
    // the last line without a newline still counts
    CHECK(parse_text(dir, "mmse 8 64 -1 16", t) == 1);
    // an over-long line is skipped whole: its tail must not be read as a line of its own
    std::string longline = "# " + std::string(400, 'x') + "\n";
    std::string sneaky = "# " + std::string(253, 'x') + "coef 1 1 1 1 8 1 1\n";  // 'coef' at the buffer's end
    CHECK(parse_text(dir, longline + sneaky + "mmse 4 32 -1 8\n", t) == 1);
    CHECK(t.coef.empty() && t.mmse.size() == 1 && t.mmse[0].L == 4);
    CHECK(parse_text(dir, std::string(5000, 'a'), t) == 0);  // one huge line without a newline
    CHECK(parse_text(dir, "", t) == 0);
    std::printf("tune table parser: ok\n");
}

// Random alloc / release against a shadow list of live blocks: blocks stay inside the arena, never overlap,
// free bytes always equal capacity minus live bytes, and releasing everything coalesces back to one block.
void arena_checks() {
    for (int trial = 0; trial < 20; ++trial) {
        std::mt19937_64 rng(1234 + trial);
        const size_t cap = (size_t)(1 + rng() % 64) << 20;
        pe::serve_host::Arena a(cap);
        struct Live {
            size_t off, bytes;
        };
        std::vector<Live> live;
        size_t used = 0;
        for (int step = 0; step < 4000; ++step) {
            if (live.empty() || rng() % 3) {
                const size_t n = 1 + rng() % (cap / 8);
                size_t off = 0;
                if (a.alloc(n, off)) {
                    const size_t r = (n + 255) & ~size_t(255);
                    CHECK(off % 256 == 0 && off + r <= cap);
                    for (const Live &l : live) CHECK(off + r <= l.off || l.off + l.bytes <= off);
                    live.push_back({off, r});
                    used += r;
                } else {
                    CHECK(a.free_bytes() < ((n + 255) & ~size_t(255)) || a.free_blocks() > 1);
                }
            } else {
                const size_t i = rng() % live.size();
                a.release(live[i].off, live[i].bytes);
                used -= live[i].bytes;
                live[i] = live.back();
                live.pop_back();
            }
            CHECK(a.free_bytes() == cap - used);
        }
        for (const Live &l : live) a.release(l.off, l.bytes);
        CHECK(a.free_bytes() == cap && a.free_blocks() == 1);
        size_t off = 1;
        CHECK(a.alloc(cap, off) && off == 0);  // the whole arena again
    }
    std::printf("serving arena: ok\n");
}

void latency_checks() {
    pe::serve_host::LatencyFit f;
    for (int i = 0; i < 7; ++i) f.observe(20e-6 + 1e-9 * (1000 + 500 * i), 1000 + 500 * i);
    CHECK(!f.ready);  // fewer than kMinPoints batches
    for (int i = 7; i < 16; ++i) f.observe(20e-6 + 1e-9 * (1000 + 500 * i), 1000 + 500 * i);
    CHECK(f.ready && std::fabs(f.b - 1e-9) < 1e-15 && std::fabs(f.a - 20e-6) < 1e-12);
    f.observe(5e-3, 4000);  // one batch delayed by the host: the robust fit barely moves
    CHECK(std::fabs(f.b - 1e-9) < 1e-10 && std::fabs(f.a - 20e-6) < 2e-6);
    for (int i = 0; i < 100; ++i) f.observe(1e-4, 3000);  // identical sizes: no slope, no division by zero
    CHECK(std::isfinite(f.a) && std::isfinite(f.b));
    pe::serve_host::LatencyModel m;
    CHECK(m.predict(1000, 0) == 0.0);  // nothing predicted before a fit exists
    for (int i = 0; i < 16; ++i) m.observe(10e-6 + 2e-9 * 100 * (i + 1), 100 * (i + 1), 0);
    CHECK(std::fabs(m.predict(1000, 0) - (10e-6 + 2e-6)) < 1e-9);
    for (int i = 0; i < 1000; ++i) m.outcome(true, 300e-6);
    CHECK(m.margin == 300e-6);  // capped
    for (int i = 0; i < 100000; ++i) m.outcome(false, 300e-6);
    CHECK(m.margin == 0.0);  // floor
    std::printf("latency model: ok\n");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: host_checks <tune_table.txt> <scratch dir>\n");
        return 2;
    }
    tune_checks(argv[1], argv[2]);
    arena_checks();
    latency_checks();
    if (g_fail) std::fprintf(stderr, "%d check(s) failed\n", g_fail);
    return g_fail ? 1 : 0;
}
