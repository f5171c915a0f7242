// tune_parse.hpp -- the text format of the tile table, parsed on the host (no HIP): used by pe_api.cpp and by the
// CPU-only checks under AddressSanitizer.
//
// Lines (first match wins, -1 = any), '#' starts a comment line:
//   coef <L> <S> <A> <prec> <f32_kc> <mfma_wpb> <mfma_spw>
//   mmse <L> <A> <nsc_max> <lanes_per_subcarrier> [<subcarriers_per_tile> [<waves_per_workgroup> [<keep>]]]
//        (tile <= 64 x waves and <= 128; keep 1: rows kept in registers where the tile allows it)
//   solve <L> <A> <nsc_max> <subcarriers_per_tile> <waves_per_workgroup>   (the one-pass chain's solve kernel)
//   slot <mmse_subcarriers_per_tile> [<mmse_waves_per_workgroup> [<mmse_keep>]]
// Out-of-range fields fall back to defaults (coef, mmse, slot) or drop the line (solve). A line longer than the
// read buffer is skipped whole, so its tail is never read as a line of its own.
#pragma once
#include <cstdio>
#include <cstring>
#include <vector>

namespace pe {
namespace tune {

struct Coef {
    int L, S, A, prec;
    int f32_kc, mfma_waves, mfma_spw;
};
struct Mmse {
    int L, A, nsc_max, lanes, tile, waves, keep;  // tile 0: default for the size
};
struct Solve {
    int L, A, nsc_max, tile, waves;
};
struct Slot {
    int mmse_tile = 0, mmse_waves = 0, mmse_keep = 0;  // zero: executor defaults
};
struct Table {
    std::vector<Coef> coef;
    std::vector<Mmse> mmse;
    std::vector<Solve> solve;
    Slot slot;
    int entries() const { return (int)(coef.size() + mmse.size() + solve.size()); }
};

// One line (NUL-terminated, newline optional) into t.
inline void parse_line(const char *line, Table &t) {
    if (line[0] == '#' || line[0] == '\n' || line[0] == '\0') return;
    Coef c{};
    Mmse m{};
    Solve s{};
    int kc = 0, wpb = 0, spw = 0, a1 = 0, a2 = 0, a3 = 0;
    if (sscanf(line, "coef %d %d %d %d %d %d %d", &c.L, &c.S, &c.A, &c.prec, &kc, &wpb, &spw) == 7) {
        c.f32_kc = (kc == 16 || kc == 4) ? kc : 8;
        c.mfma_waves = (wpb == 2 || wpb == 4) ? wpb : 1;
        c.mfma_spw = spw > 1 && spw <= 8 ? spw : 1;
        t.coef.push_back(c);
    } else if (sscanf(line, "mmse %d %d %d %d %d %d %d", &m.L, &m.A, &m.nsc_max, &m.lanes, &a1, &a2, &a3) >= 4) {
        m.keep = a3 == 1 ? 1 : 0;
        m.lanes = (m.lanes == 4 || m.lanes == 16) ? m.lanes : 8;
        m.waves = (a2 == 2 || a2 == 4 || a2 == 8) ? a2 : 1;
        m.tile = (a1 >= 1 && a1 <= 64 * m.waves && a1 <= 128) ? a1 : 0;  // 128 x 84 floats: 43 KB of LDS
        t.mmse.push_back(m);
    } else if (sscanf(line, "solve %d %d %d %d %d", &s.L, &s.A, &s.nsc_max, &s.tile, &s.waves) == 5) {
        s.waves = (s.waves == 2 || s.waves == 4 || s.waves == 8) ? s.waves : 1;
        if (s.tile >= 1 && s.tile <= 64 * s.waves && s.tile <= 128) t.solve.push_back(s);
    } else if (sscanf(line, "slot %d %d %d", &a1, &a2, &a3) >= 1) {
        t.slot.mmse_keep = a3 == 1 ? 1 : 0;
        t.slot.mmse_waves = (a2 == 2 || a2 == 4 || a2 == 8) ? a2 : 1;
        t.slot.mmse_tile = (a1 >= 1 && a1 <= 64 * t.slot.mmse_waves && a1 <= 128) ? a1 : 0;
    }
}

// Every line of f into t (which is cleared first); returns t.entries().
inline int parse(FILE *f, Table &t) {
    t = Table();
    char line[256];
    bool tail = false;  // inside the rest of an over-long line
    while (fgets(line, sizeof(line), f)) {
        const size_t n = strlen(line);
        const bool whole = n > 0 && line[n - 1] == '\n';
        const bool last = !whole && feof(f);  // final line without a newline
        if (!tail && (whole || last)) parse_line(line, t);
        tail = !whole && !last;
    }
    return t.entries();
}

}  // namespace tune
}  // namespace pe
