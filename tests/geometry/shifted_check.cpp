// Host-only check of the shifted-stripe schedule (sw_host.h's shifted_* helpers) against the
// tile geometry the sweep kernels use (make_tile in sw_kernels.hip, restated here): for each
// region shape, every diagonal's stripes match a brute-force scan, every lane covers every
// column exactly once, no two tiles of a launch touch one bottom-row or lane-slot location
// when one of them writes it, and every read finds the value the right tile wrote in an
// earlier launch.  Prints "ok" or the first failure; exit status 1 on failure.
#include <cstdio>
#include <vector>

#include "sw_host.h"

using swh_host::shifted_first_diag;
using swh_host::shifted_last_diag;
using swh_host::shifted_stripes;

namespace {

struct Tile {
    bool exists, first, last;
    int64_t C0, w;  // lane 0's columns [C0, C0 + w), relative to the region
};

Tile tile(int64_t s, int64_t g, int64_t cols, int64_t W) {
    Tile t{};
    const int64_t x0 = g * W - 64 * s;
    t.exists = g >= 0 && x0 + W > 0 && x0 < cols;
    if (!t.exists) return t;
    t.C0 = x0 > 0 ? x0 : 0;
    t.w = (x0 + W < cols ? x0 + W : cols) - t.C0;
    t.first = x0 <= 0;
    t.last = x0 + W >= cols;
    return t;
}

// Columns [a, b) (relative to the region) that lane l of tile t computes: its tile columns
// x = step - l are x >= 0 in a first segment, x >= -l in the others, and x < w in a last
// segment, x < w - l in the others (TileCtx in sw_kernels.hip).
void lane_cols(const Tile& t, int l, int64_t& a, int64_t& b) {
    a = t.C0 + (t.first ? 0 : -l);
    b = t.C0 + (t.last ? t.w : t.w - l);
}

int fail(const char* what, int64_t W, int64_t cols, int64_t S, int64_t d) {
    printf("FAIL %s: W %lld cols %lld S %lld d %lld\n", what, (long long)W, (long long)cols, (long long)S,
           (long long)d);
    return 1;
}

int check(int64_t W, int64_t cols, int64_t S) {
    const int64_t D = shifted_last_diag(S - 1, cols, W) + 1;
    // bottom-row slots [parity][col]: stripe that last wrote the value (-1 none); lane slots
    // [parity][stripe]: segment that last wrote them
    std::vector<int64_t> row_writer(2 * cols, -1), lane_writer(2 * S, -1);
    std::vector<int64_t> covered(3 * cols, 0);  // lanes 0, 31, 63
    std::vector<int> rd(2 * cols), wr(2 * cols);
    for (int64_t d = 0; d < D + 2; ++d) {
        int64_t lo, hi;
        shifted_stripes(d, S, cols, W, lo, hi);
        int64_t blo = -1, bhi = -2;
        for (int64_t s = 0; s < S; ++s)
            if (tile(s, d - s, cols, W).exists) {
                if (blo < 0) blo = s;
                bhi = s;
            }
        if ((lo > hi) != (blo < 0) || (blo >= 0 && (lo != blo || hi != bhi))) return fail("stripes", W, cols, S, d);
        for (int64_t s = blo; s >= 0 && s <= bhi; ++s)
            if (!tile(s, d - s, cols, W).exists) return fail("contiguous", W, cols, S, d);
        if (d < D && lo > hi) return fail("empty diagonal", W, cols, S, d);
        std::fill(rd.begin(), rd.end(), 0);
        std::fill(wr.begin(), wr.end(), 0);
        for (int64_t s = lo; s <= hi; ++s) {
            const int64_t g = d - s;
            const Tile t = tile(s, g, cols, W);
            if (t.w % 64 && !t.last) return fail("width", W, cols, S, d);
            for (int l : {0, 31, 63}) {
                int64_t a, b;
                lane_cols(t, l, a, b);
                for (int64_t c = a; c < b; ++c) {
                    if (c < 0 || c >= cols) return fail("lane bounds", W, cols, S, d);
                    covered[(size_t)((l == 0 ? 0 : l == 31 ? 1 : 2) * cols + c)]++;
                }
            }
            // lane 0 reads the stripe above's bottom row at its columns and the corner one
            // column left of them (outside a first segment)
            if (s > 0) {
                int64_t a0, b0;
                lane_cols(t, 0, a0, b0);
                for (int64_t c = a0 - (t.first ? 0 : 1); c < b0; ++c) {
                    if (c < 0 || c >= cols) return fail("read bounds", W, cols, S, d);
                    rd[(size_t)((s & 1) * cols + c)]++;
                    if (row_writer[(size_t)((s & 1) * cols + c)] != s - 1) return fail("read source", W, cols, S, d);
                }
            }
            // lane 63 writes the bottom row at its columns
            if (s + 1 < S) {
                int64_t a, b;
                lane_cols(t, 63, a, b);
                for (int64_t c = a; c < b; ++c) {
                    if (c < 0 || c >= cols) return fail("write bounds", W, cols, S, d);
                    wr[(size_t)(((s + 1) & 1) * cols + c)]++;
                }
            }
            if (!t.first && lane_writer[(size_t)((g & 1) * S + s)] != g - 1) return fail("lane slot source", W, cols, S, d);
        }
        for (size_t k = 0; k < rd.size(); ++k)
            if (wr[k] > 1 || (wr[k] && rd[k])) return fail("race", W, cols, S, d);
        for (int64_t s = lo; s <= hi; ++s) {
            const int64_t g = d - s;
            const Tile t = tile(s, g, cols, W);
            if (s + 1 < S) {
                int64_t a, b;
                lane_cols(t, 63, a, b);
                for (int64_t c = a; c < b; ++c) row_writer[(size_t)(((s + 1) & 1) * cols + c)] = s;
            }
            if (!t.last) lane_writer[(size_t)(((g + 1) & 1) * S + s)] = g;
        }
    }
    for (size_t k = 0; k < covered.size(); ++k)
        if (covered[k] != S) return fail("coverage", W, cols, S, (int64_t)k);
    return 0;
}

}  // namespace

int main() {
    int bad = 0;
    long shapes = 0;
    for (int64_t W : {64, 128, 192, 256, 320, 448, 512, 1024})
        for (int64_t cols : {W + 1, 2 * W, 2 * W + 5, (int64_t)1000, (int64_t)4097, (int64_t)20000})
            for (int64_t S : {1, 2, 3, 7, 64, 300}) {
                if (cols <= W) continue;
                ++shapes;
                bad += check(W, cols, S);
                if (bad) return 1;
            }
    printf("ok: %ld shapes\n", shapes);
    return 0;
}
