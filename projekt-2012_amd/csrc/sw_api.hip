// Host side of the C-ABI (include/swalign_hip.h): context, device buffers and the stage
// driver.  The stage structure replaces Sterbic/Projekt-2012's host loop
// (src/main.cu:158-248) and its Crosspointer (src/Crosspointer.cpp:104-338):
//   1. forward sweep over the whole matrix, keeping checkpoint rows {H,F} every K rows
//      and checkpoint columns {H,E} every Kc columns in HBM, plus per-row maxima;
//      a sweep is a sequence of launches, one per anti-diagonal of its tile grid;
//   2. argmax over the row maxima (smallest row holding the score);
//   3. end-cell column search: the band of blocks above that row is recomputed exactly
//      from its checkpoints, all column blocks in parallel;
//   4. traceback: the block holding the current path cell is recomputed exactly from its
//      top checkpoint row and left checkpoint column, writing 4-bit direction codes, and
//      a walker follows the canonical greedy rule until it leaves the block.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "sw_host.h"

using namespace swk;

namespace swh_host {

int map_hip(swh_ctx* c, hipError_t e) {
    if (c) c->last = e;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SWH_ERR_OUT_OF_MEMORY;
    return SWH_ERR_HIP;
}


hipError_t ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 64;
    if (b.fixed) return b.cap >= bytes ? hipSuccess : hipErrorOutOfMemory;  // a slice: never allocates
    if (b.p && b.cap >= bytes) return hipSuccess;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.cap = bytes;
    else (void)hipGetLastError();
    return e;
}

hipError_t ensure_host(HostBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 64;
    if (b.h && b.cap >= bytes) return hipSuccess;
    if (b.h) (void)hipHostFree(b.h);
    b = HostBuf{};
    hipError_t e = hipHostMalloc(&b.h, bytes, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&b.d, b.h, 0);
    if (e == hipSuccess) {
        b.cap = bytes;
    } else {
        (void)hipGetLastError();
        if (b.h) (void)hipHostFree(b.h);
        b = HostBuf{};
    }
    return e;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Direction codes kept per traceback round (bounds the lookahead).
constexpr int64_t kTracebackDirsCap = INT64_C(2) << 30;

int64_t next_pow2(int64_t v) {
    int64_t p = 64;
    while (p < v) p <<= 1;
    return p;
}

bool scoring_ok(const swh_scoring* s) {
    // src/SWutils.cpp:36: match >= 1, every penalty <= -1
    return s && s->match >= 1 && s->mismatch <= -1 && s->gap_open <= -1 && s->gap_extend <= -1;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Columns per tile: `want` if given, else one tile per resident sweep wave (wps per SIMD)
// across `cols` columns, at least 256 (a tile spends 63 extra steps filling and draining
// its lanes).
int64_t tile_cols(swh_ctx* c, int64_t cols, int64_t want, int wps) {
    if (want > 0) return want;
    const int64_t target = (int64_t)kSweepWaves * wps * std::max(1, c->props.multiProcessorCount);
    return std::max<int64_t>(256, ceil_div(cols, target));
}

// Relative time of a wave that shares its SIMD with a second sweep wave: two waves per
// SIMD retire 1.31x the steps of one (measured on MI355X on the 1 Mbp pair: 91.9 us per
// launch of 489-column tiles at two waves against 113.5 us for 977-column tiles at one; the
// second wave issues while the first waits on dependencies, LDS and branches).  With it the
// model predicts the measured ratio of the two sweeps (0.978 against 0.977).
constexpr double kTwoWaveCost = 2.0 / 1.31;

// Stage-1 segment width of a single pair: one tile per resident sweep wave across the
// columns, or with skewed stripes (whose launches hold every other diagonal's tiles) two,
// rounded up to a multiple of 64.
int64_t pair_tile_cols(swh_ctx* c, int64_t cols, int wps, bool skew) {
    return skew ? ceil_div(tile_cols(c, cols, 0, 2 * wps), 64) * 64 : tile_cols(c, cols, 0, wps);
}

// Segment width of a single pair's int32 stage 1 in shifted stripes: the widest diagonal
// holds min(S, cols / (W + 64)) tiles (stripe s starts on diagonal s + floor(64s / W)).  The
// narrowest W, a multiple of 64, at which that is at most min(S, four tiles per SIMD): fewer
// launches than narrower segments, one round of four waves per SIMD, and no fewer tiles than
// the stripes allow; at least 192 (below it a launch's fixed cost outweighs its steps).
// Measured stage 1 (ms) by width:
//   300 kbp:  64 97.8, 128 86.3, *192* 80.8, 256 81.1
//   500 kbp:  128 161.7, *192* 157.2, 256 165.8, 320 172.9
//   1 Mbp:    *192* 450.5, 384 489.9, 448 451.1, 512 483.7 (skewed stripes of 256: 474)
//   1.2 Mbp:  *256* 612.7, 448 658.4, 576 631.6, 640 667.8
int64_t shifted_pair_cols(swh_ctx* c, int64_t rows, int64_t cols) {
    const int64_t simds = (int64_t)kSweepWaves * std::max(1, c->props.multiProcessorCount);
    const int64_t S = ceil_div(rows, kStripeRows);
    const int64_t per_tile = std::max(ceil_div(cols, S), ceil_div(cols, 4 * simds));  // W + 64
    return std::max<int64_t>(192, ceil_div(per_tile - 64, 64) * 64);
}

// Modelled duration (in tile steps) of a sweep over rows x cols with wps waves per SIMD:
// one launch per tile anti-diagonal, each lasting ceil(tiles / resident waves) tile times.  A
// rectangular tile takes W + 63 steps; a skewed stripe's segments W (tile (s, g) on diagonal
// 2s + g; the lanes' fill and drain fall only in a stripe's first and last segments).
double sweep_model(swh_ctx* c, int64_t rows, int64_t cols, int wps, bool skew) {
    const int64_t W = pair_tile_cols(c, cols, wps, skew);
    const int64_t S = ceil_div(rows, kStripeRows), G = ceil_div(cols, W);
    const int64_t P = (int64_t)kSweepWaves * wps * std::max(1, c->props.multiProcessorCount);
    skew = skew && G >= 2;
    const int64_t D = skew ? 2 * (S - 1) + G : S + G - 1;
    double t = 0;
    for (int64_t d = 0; d < D; ++d) {
        const int64_t tiles = skew ? std::min(S - 1, d / 2) - std::max<int64_t>(0, (d - G + 2) / 2) + 1
                                   : std::min(d, S - 1) - std::max<int64_t>(0, d - (G - 1)) + 1;
        t += (double)ceil_div(tiles, P) * (double)(skew ? W : W + 63);
    }
    return t * (wps == 2 ? kTwoWaveCost : 1.0);
}

// Occupancy model of a single pair's stage 1 in microseconds, used to choose between the
// int32 sweep and the packed one.  Each launch (tile anti-diagonal) spreads its waves (one
// per int32 tile, one per two packed tiles) over the SIMDs; a SIMD holding q waves runs
// their steps at kStepCost[q] times one int32 wave's step, and each round of waves on a
// SIMD adds a fixed kRoundUs (loads before the first step, stores after the last).  Fitted
// on MI355X to 38 measured stage-1 times of square pairs of 1, 2, 3 and 10 Mbp (tile widths
// 64 to 4883, one to four waves per SIMD, both precisions): within 10% for packed widths of
// 128 and more.
constexpr double kStepUs = 0.091;
constexpr double kRoundUs = 5.2;
constexpr double kStepCostI32[kMaxSweepWps + 1] = {0.0, 1.0, 1.772, 2.753, 3.215};
// (packed waves hold 139 VGPRs: a fourth on a SIMD waits for one of three to end)
constexpr double kStepCostPk[kMaxSweepWps + 1] = {0.0, 1.625, 2.550, 3.653, 3.653 * 4 / 3};
// the packed sweep replaces the int32 one for a single pair only when the model (of int32
// skewed stripes) expects it to take less than this fraction of the time, and for pairs of
// at least kPackedPairMinLen on each side.  Measured against the int32 sweep in shifted
// stripes (with the first and last diagonals in int32, packed_launch): 457 against 453 ms at
// 1 Mbp, 913 against 911 ms at 1.5 Mbp, 1308 against 1527 ms at 2 Mbp
constexpr double kPackedPairMargin = 0.95;
constexpr int64_t kPackedPairMinLen = INT64_C(1500000);

double occupancy_model(swh_ctx* c, int64_t rows, int64_t cols, int64_t W, int wps, bool skew, bool packed) {
    const int64_t S = ceil_div(rows, kStripeRows), G = ceil_div(cols, W);
    const int64_t simds = (int64_t)kSweepWaves * std::max(1, c->props.multiProcessorCount);
    skew = skew && G >= 2;
    const int64_t D = skew ? 2 * (S - 1) + G : S + G - 1;
    const double st = (double)(skew ? W : W + 63) * kStepUs;
    const double* T = packed ? kStepCostPk : kStepCostI32;
    double t = 0;
    for (int64_t d = 0; d < D; ++d) {
        const int64_t tiles = skew ? std::min(S - 1, d / 2) - std::max<int64_t>(0, (d - G + 2) / 2) + 1
                                   : std::min(d, S - 1) - std::max<int64_t>(0, d - (G - 1)) + 1;
        const bool pk = packed && tiles > simds * 3 / 2;  // as packed_launch
        const double* Td = pk ? T : kStepCostI32;
        const int64_t q = ceil_div(pk ? ceil_div(tiles, 2) : tiles, simds);
        const int64_t full = q / wps, rest = q % wps;
        t += (double)full * (st * Td[wps] + kRoundUs) + (rest ? st * Td[rest] + kRoundUs : 0.0);
    }
    return t;
}

int usable_wps(const swh_ctx* c, int w) { return w >= 1 && w <= kMaxSweepWps && c->lds_pad[w] ? w : 1; }

// A packed sweep's launch runs the packed kernel only when its tiles give the SIMDs two waves
// or more (1.5 tiles per SIMD); fewer tiles (a wavefront's first and last diagonals) run
// faster one per wave in int32.  Both kernels read and write the same int32 hand-off
// buffers, lane slots, checkpoints and codes.
bool packed_launch(const swh_ctx* c, int64_t ntiles) {
    if (c->packed_min_tiles >= 0) return ntiles >= c->packed_min_tiles;
    return ntiles > (int64_t)kSweepWaves * std::max(1, c->props.multiProcessorCount) * 3 / 2;
}

// Packed skewed stripes for a single pair: the fastest modelled of segment widths that put
// about one, one half or two packed waves per SIMD slot on the widest diagonal (two or three
// waves per SIMD, widths of whole chunks from 128 up to the int16 bound).  Returns the
// modelled time, or a negative value when no width fits.
double packed_pair_choice(swh_ctx* c, int64_t rows, int64_t cols, int64_t w16, int wps_req, int64_t* W_out,
                          int* wps_out) {
    const int64_t simds = (int64_t)kSweepWaves * std::max(1, c->props.multiProcessorCount);
    const int64_t wmax = (w16 - 63) / 64 * 64;
    double best = -1.0;
    for (int wps = wps_req ? wps_req : 2; wps <= (wps_req ? wps_req : 3); ++wps) {
        const int wu = usable_wps(c, wps);
        // the widest skewed diagonal holds G / 2 tiles: G = 4 * simds * wps fills every slot
        const int64_t base = ceil_div(cols, 4 * simds * wu);
        for (int64_t f2 : {1, 2, 4}) {  // half, one and two times the filling width
            const int64_t W = std::min(wmax, std::max<int64_t>(128, ceil_div(base * f2 / 2, 64) * 64));
            if (W < 128 || ceil_div(cols, W) < 2) continue;
            const double t = occupancy_model(c, rows, cols, W, wu, true, true);
            if (best < 0 || t < best) {
                best = t;
                *W_out = W;
                *wps_out = wu;
            }
        }
    }
    // The width is rounded down to a power of two (from 256; more tiles per diagonal than the
    // model credits run well when they oversubscribe the resident waves): the 10 Mbp pair's
    // stage 1 took 21.6 s at 512 columns against 22.4 s at the model's 832 (22.0 to 22.4 s at
    // 384 to 640).
    if (best >= 0) {
        int64_t W2 = 256;
        while (W2 * 2 <= *W_out) W2 *= 2;
        if (*W_out >= 256 && W2 * 2 > *W_out) *W_out = W2;
    }
    return best;
}

bool many_walks(const swh_ctx* c, int64_t njobs) {
    const int64_t per_cu = (int64_t)c->props.maxSharedMemoryPerMultiProcessor / (int64_t)walk_lds_bytes();
    return njobs > std::max<int64_t>(1, per_cu) * std::max(1, c->props.multiProcessorCount);
}

// Sweep waves per SIMD for a rows x cols sweep: whichever of one or two the model prefers.
int choose_wps(swh_ctx* c, int64_t rows, int64_t cols, bool skew) {
    if (c->wps_fixed) return c->wps_fixed;
    if (!c->lds_pad[2]) return 1;
    return sweep_model(c, rows, cols, 2, skew) < sweep_model(c, rows, cols, 1, skew) ? 2 : 1;
}

// Runs a sweep over `regions` with tiles of kStripeRows x W: one launch per anti-diagonal
// of the tile grid, all regions side by side.  The grid has G segments per stripe (the
// widest region's, or G_fix when given: kDirs lays codes out by it).  Optional events are
// recorded immediately before the first and after the last launch.
int run_sweep(swh_ctx* c, int mode, SweepParams p, std::vector<Region>& regions, int64_t W, int64_t G_fix,
              hipEvent_t ev_before, hipEvent_t ev_after, int wps, int64_t* launches, int64_t tstride, bool packed,
              bool sync, int skew) {
    const int64_t nr = (int64_t)regions.size();
    std::vector<int64_t> Sr((size_t)nr), Gr((size_t)nr), Lr((size_t)nr);
    int64_t G = 0, D = 0, cb_elems = 0, rb_elems = 0, cor_elems = 0;
    // skewed stripes (stage 1, W a multiple of 64, regions of two or more segments): tile
    // (s, g) on diagonal 2s + g, or s + g in shifted stripes, lane slots as the segments'
    // hand-off
    if (mode != kStage1 || W % 64 != 0) skew = 0;
    if (packed && skew == 2) skew = 1;  // the packed kernel takes skewed stripes only (make_tile)
    for (int64_t q = 0; q < nr; ++q) {
        Region& r = regions[(size_t)q];
        const int64_t rows = r.r1 - r.r0 + 1, cols = r.c1 - r.c0 + 1;
        if (rows <= 0 || cols <= 0) return SWH_ERR_INTERNAL;
        Sr[(size_t)q] = ceil_div(rows, kStripeRows);
        Gr[(size_t)q] = ceil_div(cols, W);
        r.skew = Gr[(size_t)q] >= 2 ? skew : 0;
        Lr[(size_t)q] = r.skew == 2 ? shifted_last_diag(Sr[(size_t)q] - 1, cols, W) + 1
                        : r.skew    ? 2 * (Sr[(size_t)q] - 1) + Gr[(size_t)q]
                                    : Sr[(size_t)q] + Gr[(size_t)q] - 1;
        G = std::max(G, Gr[(size_t)q]);
        D = std::max(D, Lr[(size_t)q]);
        cb_elems += r.skew ? 2 * Sr[(size_t)q] * 64 * (kRowsPerLane + 1) : 2 * rows;
        rb_elems += 2 * cols;
        cor_elems += 2 * Sr[(size_t)q];
    }
    if (G_fix > 0) {
        if (G > G_fix) return SWH_ERR_INTERNAL;
        G = G_fix;
    }
    HCK(c, ensure(c->cb, sizeof(int2) * std::max<int64_t>(cb_elems, 1)));
    HCK(c, ensure(c->cor, sizeof(int32_t) * std::max<int64_t>(cor_elems, 1)));
    HCK(c, ensure(c->rbuf, sizeof(int2) * std::max<int64_t>(rb_elems, 1)));
    int64_t cbo = 0, rbo = 0, coro = 0;
    for (int64_t q = 0; q < nr; ++q) {
        Region& r = regions[(size_t)q];
        r.W = W;
        r.G = G;
        r.tstride = tstride;
        r.cb = (int2*)c->cb.p + cbo;
        r.rb = (int2*)c->rbuf.p + rbo;
        r.cor = (int32_t*)c->cor.p + coro;
        cbo += r.skew ? 2 * Sr[(size_t)q] * 64 * (kRowsPerLane + 1) : 2 * (r.r1 - r.r0 + 1);
        rbo += 2 * (r.c1 - r.c0 + 1);
        coro += 2 * Sr[(size_t)q];
    }
    // Spans of every launch: for diagonal d, each region with tiles on it contributes its
    // stripes [max(0, d - G_r + 1), min(S_r - 1, d)].  Regions are visited longest-lived
    // first (a counting sort by the number of diagonals) so that each diagonal stops at the
    // first region that has ended.
    const auto t_spans = std::chrono::steady_clock::now();
    std::vector<int64_t> order((size_t)nr);
    int64_t total = 0;
    {
        std::vector<int64_t> at((size_t)D + 2, 0);  // regions living longer than each count
        for (int64_t q = 0; q < nr; ++q) ++at[(size_t)(D - Lr[(size_t)q])];
        int64_t acc = 0;
        for (auto& v : at) {
            const int64_t n_here = v;
            v = acc;
            acc += n_here;
        }
        for (int64_t q = 0; q < nr; ++q) {
            order[(size_t)at[(size_t)(D - Lr[(size_t)q])]++] = q;
            total += Lr[(size_t)q];
        }
    }
    HCK(c, ensure(c->regions, sizeof(Region) * std::max<int64_t>(nr, 1)));
    HCK(c, hipMemcpyAsync(c->regions.p, regions.data(), sizeof(Region) * nr, hipMemcpyHostToDevice, c->stream));
    HCK(c, ensure(c->spans, sizeof(TileSpan) * (size_t)std::max<int64_t>(total, 1)));
    // pinned, grow-only: the uploads are asynchronous (the caller synchronises the stream
    // before the next sweep, as for `regions`); a context that may not allocate pinned
    // memory stages them in host memory (its uploads complete before the call returns)
    TileSpan* hs;
    if (c->fixed_memory) {
        c->host_spans.resize((size_t)std::max<int64_t>(total, 1));
        hs = c->host_spans.data();
    } else {
        if ((size_t)total > c->pinned_spans_cap) {
            if (c->pinned_spans) (void)hipHostFree(c->pinned_spans);
            c->pinned_spans = nullptr;
            c->pinned_spans_cap = 0;
            HCK(c, hipHostMalloc((void**)&c->pinned_spans, sizeof(TileSpan) * (size_t)std::max<int64_t>(total, 1), 0));
            c->pinned_spans_cap = (size_t)std::max<int64_t>(total, 1);
        }
        hs = c->pinned_spans;
    }
    TileSpan* ds = (TileSpan*)c->spans.p;
    p.regions = (const Region*)c->regions.p;
    // Large span lists (many regions) are built a few diagonals ahead of the launches that
    // use them, each part uploaded before its launches, so that the device starts at once
    // and the host builds the rest while it sweeps; small ones are built and uploaded whole.
    const bool pipelined = total > (int64_t)1 << 16;
    int64_t pos = 0, d_built = 0;
    std::vector<int64_t> span_at((size_t)D + 1, 0), tiles_at((size_t)D, 0);
    // shifted regions: the stripes [lo, hi] on the diagonal last built (stripe 0 alone on
    // diagonal 0); each region is visited on every diagonal it lives on, in order
    std::vector<int64_t> sh_lo, sh_hi;
    if (skew == 2) {
        sh_lo.assign((size_t)nr, 0);
        sh_hi.assign((size_t)nr, 0);
    }
    auto build_to = [&](int64_t d_end) {
        const int64_t from = pos;
        for (int64_t d = d_built; d < d_end; ++d) {
            span_at[(size_t)d] = pos;
            int64_t t = 0;
            for (int64_t q : order) {
                const int64_t Sq = Sr[(size_t)q], Gq = Gr[(size_t)q];
                if (d >= Lr[(size_t)q]) break;
                int64_t s_lo, s_hi;
                if (regions[(size_t)q].skew == 2) {
                    // from diagonal d - 1's stripes: each bound moves by at most one
                    int64_t& lo = sh_lo[(size_t)q];
                    int64_t& hi = sh_hi[(size_t)q];
                    if (d > 0) {
                        const int64_t cols = regions[(size_t)q].c1 - regions[(size_t)q].c0 + 1;
                        while (hi + 1 < Sq && shifted_first_diag(hi + 1, W) <= d) ++hi;
                        while (lo < hi && shifted_last_diag(lo, cols, W) < d) ++lo;
                    }
                    s_lo = lo;
                    s_hi = hi;
                } else if (regions[(size_t)q].skew) {  // g = d - 2s in [0, Gq)
                    s_lo = std::max<int64_t>(0, (d - Gq + 2) / 2);
                    s_hi = std::min(Sq - 1, d / 2);
                } else {
                    s_lo = std::max<int64_t>(0, d - (Gq - 1));
                    s_hi = std::min(Sq - 1, d);
                }
                hs[pos++] = TileSpan{t, (int32_t)q, (int32_t)s_lo, d};
                t += s_hi - s_lo + 1;
            }
            tiles_at[(size_t)d] = t;
        }
        span_at[(size_t)d_end] = pos;
        d_built = d_end;
        return hipMemcpyAsync(ds + from, hs + from, sizeof(TileSpan) * (size_t)(pos - from), hipMemcpyHostToDevice,
                              c->stream);
    };
    if (!pipelined) HCK(c, build_to(D));
    if (c->trace && nr > 1000) {
        fprintf(stderr, "swalign: sweep set-up: %lld regions, %lld spans%s, %.3f ms before the first launch\n",
                (long long)nr, (long long)total, pipelined ? " (built along the launches)" : "",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_spans).count());
    }
    if (ev_before) HCK(c, hipEventRecord(ev_before, c->stream));
    // diagonals built per part: doubling from 2, up to about 65,536 spans a part (16 diagonals
    // of the 65,536-pair batch; a single pair's whole list in a few parts)
    const int64_t chunk_max = std::max<int64_t>(16, ((int64_t)1 << 16) / std::max<int64_t>(1, total / std::max<int64_t>(1, D)));
    int64_t chunk = 2;
    for (int64_t d = 0; d < D; ++d) {
        if (d == d_built) {
            HCK(c, build_to(std::min(D, d + chunk)));
            chunk = std::min<int64_t>(2 * chunk, chunk_max);
        }
        p.spans = ds + span_at[(size_t)d];
        p.nspans = (int32_t)(span_at[(size_t)d + 1] - span_at[(size_t)d]);
        p.ntiles = tiles_at[(size_t)d];
        if (packed && (mode == kStage1 || mode == kDirs) && packed_launch(c, p.ntiles))
            HCK(c, launch_sweep_pk(mode, p, c->lds_pad[usable_wps(c, wps)], c->stream));
        else HCK(c, launch_sweep(mode, p, c->lds_pad[usable_wps(c, wps)], c->stream));
    }
    if (ev_after) HCK(c, hipEventRecord(ev_after, c->stream));
    if (launches) *launches = D;
    // `regions` and the context's span list fed asynchronous copies: either wait for them
    // here, or (sync == false) the caller synchronises the stream before it changes
    // `regions` or sweeps again
    if (sync) HCK(c, hipStreamSynchronize(c->stream));
    return SWH_OK;
}

int code_shift1_for(const uint32_t mask[8], const swh_scoring* sc) {
    if ((int64_t)sc->match - sc->mismatch > 255) return 0;
    for (int sh = 0; sh <= 5; ++sh) {
        unsigned used = 0;
        bool ok = true;
        for (int x = 0; x < 256 && ok; ++x) {
            if (!(mask[x >> 5] >> (x & 31) & 1u)) continue;
            const unsigned code = 1u << ((x >> sh) & 7);
            ok = !(used & code);
            used |= code;
        }
        if (ok) return sh + 1;
    }
    return 0;
}

int code2_shift1_for(const uint32_t mask[8], const swh_scoring* sc) {
    const int64_t mo = (int64_t)sc->match - sc->gap_open, xo = (int64_t)sc->mismatch - sc->gap_open;
    if (mo > 255 || xo < 0 || xo > 255) return 0;
    for (int sh = 0; sh <= 6; ++sh) {
        unsigned used = 0;
        bool ok = true;
        for (int x = 0; x < 256 && ok; ++x) {
            if (!(mask[x >> 5] >> (x & 31) & 1u)) continue;
            const unsigned code = 1u << ((x >> sh) & 3);
            ok = !(used & code);
            used |= code;
        }
        if (ok) return sh + 1;
    }
    return 0;
}

int code_shift1_device(swh_ctx* c, const std::vector<ByteRange>& ranges, const swh_scoring* sc, int* shift1,
                       int* shift2) {
    *shift1 = 0;
    if (shift2) *shift2 = 0;
    if (c->codes_off || ranges.empty()) return SWH_OK;
    HCK(c, ensure(c->ranges, sizeof(ByteRange) * ranges.size()));
    HCK(c, ensure(c->alpha, sizeof(uint32_t) * 8));
    HCK(c, hipMemcpyAsync(c->ranges.p, ranges.data(), sizeof(ByteRange) * ranges.size(), hipMemcpyHostToDevice,
                          c->stream));
    HCK(c, hipMemsetAsync(c->alpha.p, 0, sizeof(uint32_t) * 8, c->stream));
    HCK(c, launch_alphabet((const ByteRange*)c->ranges.p, (int64_t)ranges.size(), (uint32_t*)c->alpha.p, c->stream));
    uint32_t mask[8] = {};
    HCK(c, hipMemcpyAsync(mask, c->alpha.p, sizeof mask, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    *shift1 = code_shift1_for(mask, sc);
    if (shift2) *shift2 = code2_shift1_for(mask, sc);
    return SWH_OK;
}

int64_t int16_tile_cols(const swh_scoring* sc) {
    const int64_t M = sc->match, X = -(int64_t)sc->mismatch, O = -(int64_t)sc->gap_open, E = -(int64_t)sc->gap_extend;
    if (M > 1000 || X > 1000 || O > 1000 || E > 1000) return 0;
    // H moves by at most M + O between neighbouring cells: a tile of 256 rows by W columns and
    // its inputs lie within (256 + W + 2) (M + O) of the tile's base, E and F up to O + E below
    const int64_t w = (32000 - (M + O + X + E)) / (M + O) - (kStripeRows + 2);
    return w > 0 ? w : 0;
}

int64_t plan_mem(int64_t m, int64_t n, int64_t K, int64_t Kc) {
    return ((m / K) * n + (n / Kc) * m) * (int64_t)sizeof(int2);
}

// Checkpoint spacing of one pair: K rows (multiple of 256) and Kc columns (power of two),
// from the options or doubled from max(m, n)/64 (1024 to 8192) until the checkpoints fit
// `budget` bytes.
int choose_plan(int64_t m, int64_t n, const swh_opts* o, int64_t budget, Plan* pl) {
    int64_t K = o && o->checkpoint_rows > 0 ? ceil_div(o->checkpoint_rows, kStripeRows) * kStripeRows : 0;
    // the sweeps take Kc as a power of two >= 64 (a lane meets one checkpoint column per chunk)
    int64_t Kc = o && o->checkpoint_cols > 0 ? std::max<int64_t>(64, next_pow2(o->checkpoint_cols)) : 0;
    if (o && o->score_only) {
        K = INT64_C(1) << 62;
        Kc = INT64_C(1) << 62;
    } else if (K == 0) {
        // max(m, n)/64 rows, 1024 to 8192: a short pair's end search and traceback blocks stay
        // small, a long pair keeps few checkpoints
        K = std::min<int64_t>(8192, std::max<int64_t>(1024, next_pow2(std::max(m, n) / 64)));
        if (Kc == 0) Kc = next_pow2(K);
        while (plan_mem(m, n, K, Kc) > budget && K < (INT64_C(1) << 40)) {
            K *= 2;
            Kc *= 2;
        }
    } else if (Kc == 0) {
        Kc = next_pow2(K);
        while (plan_mem(m, n, K, Kc) > budget && Kc < (INT64_C(1) << 40)) Kc *= 2;
    }
    if (!(o && o->score_only) && plan_mem(m, n, K, Kc) > budget) return SWH_ERR_BUDGET;
    pl->K = K;
    pl->Kc = Kc;
    pl->nsr = m / K;
    pl->nsc = n / Kc;
    return SWH_OK;
}

// Checkpoint budget when the options give none: half of the memory available to the call --
// what is still free plus the checkpoint buffers the context already holds (a later call
// reuses them) -- and never less than those buffers, so that repeated calls keep one plan.
// HBM, or host memory for host-resident checkpoints.
int64_t default_budget(const swh_ctx* c, const swh_opts* o) {
    if (o && o->hbm_budget_bytes > 0) return o->hbm_budget_bytes;
    int64_t held, freeb;
    if (o && o->host_checkpoints) {
        const long pages = sysconf(_SC_AVPHYS_PAGES), psz = sysconf(_SC_PAGESIZE);
        held = (int64_t)(c->hsrow.cap + c->hscol.cap);
        freeb = pages > 0 && psz > 0 ? (int64_t)pages * psz : (INT64_C(16) << 30);
    } else {
        size_t f = 0, t = 0;
        (void)hipMemGetInfo(&f, &t);
        held = (int64_t)(c->srow.cap + c->scol.cap);
        freeb = (int64_t)f;
    }
    return std::max(held, (freeb + held) / 2);
}

// Columns per tile of a batch sweep: wide enough that the 63 fill steps of a tile cost
// about a tenth, narrow enough that short pairs still cut into several tiles.
constexpr int64_t kBatchTileCols = 512;
// Checkpoint blocks of a batch, rows by columns: the traceback recomputes about the block
// height per cell of the path while stage 1 writes a checkpoint row per stripe, and a column
// spacing that is a multiple of the tile width costs stage 1 nothing per step (65,536-pair
// batch, W = 512: 256 x 512 723 ms a batch, 512 x 512 744, 256 x 256 751, 1024 x 1024 892).
constexpr int64_t kBatchCkptRows = 256, kBatchCkptCols = 512;
// Sweep waves per SIMD of a batch: a batch has tiles to spare, and several waves per SIMD hide
// the per-tile set-up and memory waits that one or two leave exposed (65,536-pair batch,
// forward sweep: 1210 / 763 / 668 / 650 ms at 1 / 2 / 3 / 4 waves per SIMD, rectangular
// tiles).  The packed stage 1 of skewed stripes holds 139 VGPRs: three waves fit a SIMD.
constexpr int kBatchWps = 3;
// A tuned candidate replaces the batch defaults only when it is faster than this fraction of
// the defaults' time on the sample (on the 65,536-pair batch, 256-column tiles timed 2.2%
// faster than the default 512 on the sample, and 4% slower on the whole batch).
constexpr double kTuneMargin = 0.95;

// Runs stage 1, the end search and (for jobs that want it) the traceback of every job,
// all jobs side by side in each sweep.  Fills out[k] (result_init'ed by the caller) and
// jobs[k].status.  Stage times are those of the whole batch.
constexpr int64_t kAlphabetRange = INT64_C(1) << 16;  // bytes per workgroup of k_alphabet

// Host wall time between the stages of run_jobs, printed to stderr when the context traces.
struct StageClock {
    bool on;
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void mark(const char* stage, int64_t n) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "swalign: %-10s %9.3f ms (%lld)\n", stage,
                std::chrono::duration<double, std::milli>(now - last).count(), (long long)n);
        last = now;
    }
};

// Traceback rounds of a batch planned on the device, one block per walking pair and round
// (no lookahead), while the walking pairs' tiles fill the device: the plan kernel builds each
// pair's region and walk arguments from its walk state in device memory, the codes sweep and
// the walk run as in host-planned rounds, and a compaction kernel lists the pairs that still
// walk.  The host reads the number of walking pairs back once every kRoundsPerSync rounds
// (the launches shrink to it) and stops when the rest would be better served by host-planned
// rounds with lookahead.  On return `active` and `states` hold the pairs still walking.
constexpr int kRoundsPerSync = 4;
// Alignments from which their CIGARs are built on the device rather than on the host.
constexpr int64_t kGpuCigarMinJobs = 256;

int device_rounds(swh_ctx* c, std::vector<Job>& jobs, std::vector<int64_t>& active, std::vector<WalkState>& states,
                  SweepParams p, int2* srow, int2* scol, int64_t Wtb, int64_t tstride, int64_t Gblk, bool tb_packed,
                  int tb_wps, int64_t dirs_cap, int64_t& guard, int64_t guard_max) {
    const int64_t nj = (int64_t)jobs.size();
    const int64_t N0 = (int64_t)active.size();
    int64_t Kb = 0, Wb = 0;  // rows and columns of the largest block
    for (int64_t k : active) {
        Kb = std::max(Kb, std::min(jobs[k].m, jobs[k].pl.K));
        Wb = std::max(Wb, std::min(jobs[k].n, jobs[k].pl.Kc));
    }
    const int64_t Sb = ceil_div(Kb, kStripeRows);
    const int64_t D = Sb + Gblk - 1;
    const int64_t slot_dirs = Sb * Gblk * tstride * 64;  // code elements per slot
    const int64_t slots = (int64_t)kSweepWaves * tb_wps * (tb_packed ? 2 : 1) * std::max(1, c->props.multiProcessorCount);
    // the host-planned rounds would look ahead (tiles of all walking pairs at one block
    // ahead fit the resident waves): leave the rest to them
    auto fills = [&](int64_t nact) { return lookahead_blocks(1) * nact * Sb * Gblk > slots * D; };
    // every walking pair holds a code slot: up to half of the free memory (the code buffer
    // already held counts as free), at least the host-planned rounds' share
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    const int64_t cap = std::max<int64_t>(dirs_cap, ((int64_t)freeb + (int64_t)c->dirs.cap) / 2);
    if (N0 == 0 || !fills(N0) || N0 >= INT32_MAX || N0 * slot_dirs * (int64_t)sizeof(uint16_t) > cap)
        return SWH_OK;

    std::vector<TbPair> tp((size_t)nj);
    for (int64_t k : active) {
        const Job& j = jobs[k];
        TbPair& q = tp[(size_t)k];
        q.A = j.A;
        q.B = j.B;
        q.m = j.m;
        q.n = j.n;
        q.K = j.pl.K;
        q.Kc = j.pl.Kc;
        q.srow = j.pl.nsr ? srow + j.srow_off : nullptr;
        q.scol = j.pl.nsc ? scol + j.scol_off : nullptr;
        q.ops = (uint8_t*)c->ops.p + j.ops_off;
        q.xps = (XPoint*)c->xps.p + j.xps_off;
        q.xcap = j.xcap;
    }
    std::vector<int32_t> act0((size_t)N0 + 2);  // two counters, then the first list
    act0[0] = (int32_t)N0;
    act0[1] = 0;
    for (int64_t x = 0; x < N0; ++x) act0[(size_t)x + 2] = (int32_t)active[(size_t)x];
    std::vector<TileSpan> spans((size_t)(D * N0));
    std::vector<int64_t> per_diag((size_t)D);
    for (int64_t d = 0; d < D; ++d) {
        const int64_t s_lo = std::max<int64_t>(0, d - (Gblk - 1)), s_hi = std::min(Sb - 1, d);
        per_diag[(size_t)d] = s_hi - s_lo + 1;
        for (int64_t x = 0; x < N0; ++x)
            spans[(size_t)(d * N0 + x)] = TileSpan{x * per_diag[(size_t)d], (int32_t)x, (int32_t)s_lo, d};
    }
    HCK(c, ensure(c->tb_pairs, sizeof(TbPair) * (size_t)nj));
    HCK(c, ensure(c->tb_states, sizeof(WalkState) * (size_t)nj));
    HCK(c, ensure(c->tb_act, sizeof(int32_t) * (size_t)(2 * N0 + 2)));
    HCK(c, ensure(c->tb_pre, sizeof(TbPre) * (size_t)N0));
    HCK(c, ensure(c->regions, sizeof(Region) * (size_t)N0));
    HCK(c, ensure(c->wregs, sizeof(WalkRegion) * (size_t)N0));
    HCK(c, ensure(c->wargs, sizeof(WalkArgs) * (size_t)N0));
    HCK(c, ensure(c->spans, sizeof(TileSpan) * spans.size()));
    HCK(c, ensure(c->cb, sizeof(int2) * (size_t)(N0 * 2 * Kb)));
    HCK(c, ensure(c->rbuf, sizeof(int2) * (size_t)(N0 * 2 * Wb)));
    HCK(c, ensure(c->cor, sizeof(int32_t) * (size_t)(N0 * 2 * Sb)));
    HCK(c, ensure(c->dirs, sizeof(uint16_t) * (size_t)(N0 * slot_dirs)));
    HCK(c, hipMemcpyAsync(c->tb_pairs.p, tp.data(), sizeof(TbPair) * tp.size(), hipMemcpyHostToDevice, c->stream));
    HCK(c, hipMemcpyAsync(c->tb_states.p, states.data(), sizeof(WalkState) * (size_t)nj, hipMemcpyHostToDevice,
                          c->stream));
    HCK(c, hipMemcpyAsync(c->tb_act.p, act0.data(), sizeof(int32_t) * act0.size(), hipMemcpyHostToDevice, c->stream));
    HCK(c, hipMemcpyAsync(c->spans.p, spans.data(), sizeof(TileSpan) * spans.size(), hipMemcpyHostToDevice,
                          c->stream));
    int32_t* cnt = (int32_t*)c->tb_act.p;
    int32_t* lists[2] = {cnt + 2, cnt + 2 + N0};
    TbRoundArgs ra{};
    ra.pairs = (const TbPair*)c->tb_pairs.p;
    ra.states = (WalkState*)c->tb_states.p;
    ra.regions = (Region*)c->regions.p;
    ra.wregs = (WalkRegion*)c->wregs.p;
    ra.wargs = (WalkArgs*)c->wargs.p;
    ra.pre = (TbPre*)c->tb_pre.p;
    ra.W = Wtb;
    ra.G = Gblk;
    ra.tstride = tstride;
    ra.cb = (int2*)c->cb.p;
    ra.cb_stride = 2 * Kb;
    ra.rb = (int2*)c->rbuf.p;
    ra.rb_stride = 2 * Wb;
    ra.cor = (int32_t*)c->cor.p;
    ra.cor_stride = 2 * Sb;
    ra.dirs = (uint16_t*)c->dirs.p;
    ra.dirs_stride = slot_dirs;
    ra.sc = p.sc;
    p.regions = (const Region*)c->regions.p;
    const size_t lds = c->lds_pad[usable_wps(c, tb_wps)];
    int cur = 0;
    int64_t nb = N0;
    int32_t walking = (int32_t)N0;
    const auto t_start = std::chrono::steady_clock::now();
    while (walking > 0 && fills(walking)) {
        for (int r = 0; r < kRoundsPerSync; ++r) {
            if (++guard > guard_max) return SWH_ERR_INTERNAL;
            ra.act_in = lists[cur];
            ra.nact_in = cnt + cur;
            ra.act_out = lists[1 - cur];
            ra.nact_out = cnt + 1 - cur;
            ra.nslots = nb;
            HCK(c, launch_tb_plan(ra, c->stream));
            for (int64_t d = 0; d < D; ++d) {
                p.spans = (const TileSpan*)c->spans.p + d * N0;
                p.nspans = (int32_t)nb;
                p.ntiles = nb * per_diag[(size_t)d];
                if (tb_packed && packed_launch(c, p.ntiles)) HCK(c, launch_sweep_pk(kDirs, p, lds, c->stream));
                else HCK(c, launch_sweep(kDirs, p, lds, c->stream));
            }
            HCK(c, launch_walk((const WalkArgs*)c->wargs.p, (int)nb, many_walks(c, nb), c->stream));
            HCK(c, launch_tb_compact(ra, c->stream));
            cur = 1 - cur;
        }
        HCK(c, hipMemcpyAsync(&walking, cnt + cur, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        if (c->trace)
            fprintf(stderr, "swalign: device rounds to %lld: %d pairs walking (%lld slots, %lld launches per round), "
                    "%.3f ms\n", (long long)guard, walking, (long long)nb, (long long)(D + 3),
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count());
        nb = walking;
    }
    HCK(c, hipMemcpyAsync(states.data(), c->tb_states.p, sizeof(WalkState) * (size_t)nj, hipMemcpyDeviceToHost,
                          c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    std::vector<int64_t> next;
    for (int64_t k : active) {
        WalkState& st = states[(size_t)k];
        if (st.done == -1) {
            jobs[k].status = SWH_ERR_INTERNAL;  // the walk stopped moving
            st.done = 1;
        } else if (!st.done) {
            next.push_back(k);
        }
    }
    active.swap(next);
    return SWH_OK;
}

int run_jobs(swh_ctx* c, std::vector<Job>& jobs, const swh_scoring* sc, const EngineOpts& eo, swh_result* out) {
    {
        const char* pm = getenv("SWALIGN_PACKED_MIN_TILES");
        c->packed_min_tiles = pm && *pm ? std::max<int64_t>(0, atoll(pm)) : -1;
    }
    const int64_t nj = (int64_t)jobs.size();
    StageClock clk{c->trace};
    int64_t rm_elems = 0, sr_elems = 0, scl_elems = 0, ops_elems = 0, xps_elems = 0;
    int64_t max_m = 0, max_n = 0;
    for (Job& j : jobs) {
        j.rowmax_off = rm_elems;
        j.srow_off = sr_elems;
        j.scol_off = scl_elems;
        j.ops_off = ops_elems;
        j.xps_off = xps_elems;
        j.xcap = j.pl.nsr + 2;
        rm_elems += j.m;
        sr_elems += j.pl.nsr * j.n;
        scl_elems += j.pl.nsc * j.m;
        ops_elems += j.m + j.n + 16;
        xps_elems += j.xcap;
        max_m = std::max(max_m, j.m);
        max_n = std::max(max_n, j.n);
    }
    HCK(c, ensure(c->rowmax, sizeof(int32_t) * std::max<int64_t>(rm_elems, 1)));
    HCK(c, ensure(c->rowseg, sizeof(int32_t) * std::max<int64_t>(rm_elems, 1)));
    HCK(c, ensure(c->segs, sizeof(int32_t) * std::max<int64_t>(nj, 1)));
    int2* srow;
    int2* scol;
    if (eo.host_ckpt) {
        HCK(c, ensure_host(c->hsrow, sizeof(int2) * std::max<int64_t>(sr_elems, 1)));
        HCK(c, ensure_host(c->hscol, sizeof(int2) * std::max<int64_t>(scl_elems, 1)));
        srow = (int2*)c->hsrow.d;
        scol = (int2*)c->hscol.d;
    } else {
        HCK(c, ensure(c->srow, sizeof(int2) * std::max<int64_t>(sr_elems, 1)));
        HCK(c, ensure(c->scol, sizeof(int2) * std::max<int64_t>(scl_elems, 1)));
        srow = (int2*)c->srow.p;
        scol = (int2*)c->scol.p;
    }
    HCK(c, ensure(c->key, sizeof(unsigned long long) * std::max<int64_t>(nj, 1)));
    HCK(c, ensure(c->found, sizeof(int32_t) * std::max<int64_t>(nj, 1)));
    int32_t* rowmax = (int32_t*)c->rowmax.p;
    int32_t* rowseg = (int32_t*)c->rowseg.p;  // first stage-1 segment holding each row maximum
    int32_t* segs = (int32_t*)c->segs.p;
    const Score S{sc->match, sc->mismatch, sc->gap_open, sc->gap_extend};
    SweepParams p{};
    p.sc = S;
    for (int64_t k = 0; k < nj; ++k) {
        out[k].m = jobs[k].m;
        out[k].n = jobs[k].n;
        out[k].swapped = jobs[k].swapped ? 1 : 0;
        out[k].score = 0;
        out[k].end_row = out[k].end_col = out[k].start_row = out[k].start_col = -1;
        out[k].precision_used = 32;
        out[k].checkpoint_rows = jobs[k].pl.nsr ? jobs[k].pl.K : 0;
        out[k].checkpoint_cols = jobs[k].pl.nsc ? jobs[k].pl.Kc : 0;
    }

    // ---- stage 1: all pairs, one region each --------------------------------------
    std::vector<Region> regions;
    std::vector<int64_t> live;  // jobs with cells
    for (int64_t k = 0; k < nj; ++k) {
        const Job& j = jobs[k];
        if (j.m == 0 || j.n == 0) continue;
        Region r{};
        r.A = j.A;
        r.B = j.B;
        r.m = j.m;
        r.n = j.n;
        r.r0 = 0;
        r.r1 = j.m - 1;
        r.c0 = 0;
        r.c1 = j.n - 1;
        r.rowmax = rowmax + j.rowmax_off;
        r.rowseg = rowseg + j.rowmax_off;
        r.srow = j.pl.nsr ? srow + j.srow_off : nullptr;
        r.K = j.pl.K;
        r.scol = j.pl.nsc ? scol + j.scol_off : nullptr;
        r.Kc = j.pl.Kc;
        regions.push_back(r);
        live.push_back(k);
    }
    if (live.empty()) return SWH_OK;
    // packed 16-bit stage 1 (two tiles per wave) when every tile's values fit int16 around
    // the tile's base: by default for batches, for single pairs on request (precision 16) or
    // where the occupancy model prefers it (below)
    const int64_t w16 = int16_tile_cols(sc);
    bool packed = (eo.precision == 16 || (eo.precision == 0 && eo.batch)) && w16 > 0;
    int wps = 1;
    int64_t W = 0;
    // skewed stage-1 stripes (no lane fill and drain between segments) need W a multiple of
    // 64 and, packed, 63 more columns of int16 headroom (a skewed tile spans W + 63 columns)
    bool skew = eo.skew;
    bool shift = eo.layout == 2;  // shifted stripes: by default for int32 single pairs (below)
    if (eo.batch) {
        wps = usable_wps(c, eo.wps ? eo.wps : (c->wps_fixed ? c->wps_fixed : kBatchWps));
        W = eo.tile_cols > 0 ? eo.tile_cols : kBatchTileCols;
        skew = skew && W % 64 == 0 && (!packed || W + 63 <= w16);
        packed = packed && (skew ? W + 63 : W) <= w16;
    } else {
        // single pairs: packed skewed stripes on request (precision 16), or by default where
        // the occupancy model expects them to beat the int32 sweep by a margin (long pairs)
        int64_t Wp = 0;
        int wp = 1;
        bool pk_skew = false;
        if (skew && w16 > 0 && eo.precision != 32) {
            if (eo.tile_cols > 0) {
                pk_skew = packed && eo.tile_cols % 64 == 0 && eo.tile_cols + 63 <= w16;
                Wp = eo.tile_cols;
                wp = usable_wps(c, eo.wps ? eo.wps : 2);
            } else if (packed || std::min(max_m, max_n) >= kPackedPairMinLen) {
                const double tp = packed_pair_choice(c, max_m, max_n, w16, eo.wps, &Wp, &wp);
                if (tp >= 0 && packed) {
                    pk_skew = true;
                } else if (tp >= 0) {
                    const int ws = eo.wps ? usable_wps(c, eo.wps) : choose_wps(c, max_m, max_n, true);
                    const int wr = eo.wps ? usable_wps(c, eo.wps) : choose_wps(c, max_m, max_n, false);
                    const bool s32 = sweep_model(c, max_m, max_n, ws, true) < sweep_model(c, max_m, max_n, wr, false);
                    const double ti =
                        s32 ? occupancy_model(c, max_m, max_n, pair_tile_cols(c, max_n, ws, true), ws, true, false)
                            : occupancy_model(c, max_m, max_n, pair_tile_cols(c, max_n, wr, false), wr, false, false);
                    pk_skew = tp < kPackedPairMargin * ti;
                }
            }
        }
        if (pk_skew) {
            // two tiles per wave in skewed stripes (10 Mbp pair: 22.4 s against 32.0 s for the
            // int32 rectangular sweep; 3 Mbp: 2.58 against 3.31 s; 1 Mbp stays int32: 470 ms
            // against 489 ms)
            packed = true;
            W = Wp;
            wps = wp;
        } else if (skew && !packed &&
                   (eo.layout != 1 || eo.skew_forced || eo.tile_cols > 0 ||
                    sweep_model(c, max_m, max_n, choose_wps(c, max_m, max_n, true), true) <
                        sweep_model(c, max_m, max_n, choose_wps(c, max_m, max_n, false), false))) {
            // int32 single pairs run shifted stripes unless skewed ones are asked for, which
            // follow the tile-step model: shifted stage 1 beat rectangular tiles at every size
            // measured (ms): 3 kbp 0.75 against 0.87, 10 kbp 2.47 / 3.00, 50 kbp 12.4 / 15.1,
            // 300 kbp 80.8 / 103.1, 1 Mbp x 300 kbp 216 / 233, 1 Mbp 451 / 532
            // a skewed stripe's launches hold tiles of every other diagonal: segments of half
            // the width keep the resident waves busy.  Skewed when the model says so: the 1 Mbp
            // pair (4.58 M against 5.02 M modelled tile steps; 470 against 531 ms measured)
            // Shifted stripes (stripe s's segments 64s columns left) unless skewed ones are
            // asked for: stripe s starts on diagonal s + floor(64s / W) instead of 2s, so the
            // wavefront fills the device sooner and takes fewer launches (1 Mbp pair: 10417
            // launches of 192 columns against 11719 of 256, stage 1 451 against 474 ms)
            shift = eo.layout != 1;
            wps = eo.wps          ? usable_wps(c, eo.wps)
                  : eo.tile_cols > 0 ? 1
                  : shift         ? usable_wps(c, kMaxSweepWps)
                                  : choose_wps(c, max_m, max_n, true);
            W = eo.tile_cols > 0 ? eo.tile_cols : shift ? shifted_pair_cols(c, max_m, max_n) : pair_tile_cols(c, max_n, wps, true);
            skew = W % 64 == 0;
        } else {
            skew = false;
            if (packed) {  // two tiles per wave, one wave per SIMD
                W = tile_cols(c, max_n, eo.tile_cols, 2);
                if (W > w16 && eo.tile_cols > 0) {
                    packed = false;  // the requested width does not fit int16
                } else if (W > w16) {
                    W = w16;  // narrower tiles: more of them per launch, two waves per SIMD if needed
                    const int64_t slots = (int64_t)kSweepWaves * std::max(1, c->props.multiProcessorCount);
                    if (ceil_div(max_n, W) > 2 * slots && c->lds_pad[2]) wps = 2;
                }
                if (eo.wps) wps = usable_wps(c, eo.wps);
            }
            if (!packed) {
                wps = eo.wps ? usable_wps(c, eo.wps) : eo.tile_cols > 0 ? 1 : choose_wps(c, max_m, max_n, false);
                W = tile_cols(c, max_n, eo.tile_cols, wps);
            }
        }
    }
    {  // symbol codes for the int32 sweeps, from the bytes of the pairs
        std::vector<ByteRange> br;
        for (int64_t k : live) {
            const Job& j = jobs[k];
            for (int64_t o = 0; o < j.m; o += kAlphabetRange) br.push_back(ByteRange{j.A + o, std::min(kAlphabetRange, j.m - o)});
            for (int64_t o = 0; o < j.n; o += kAlphabetRange) br.push_back(ByteRange{j.B + o, std::min(kAlphabetRange, j.n - o)});
        }
        const int rc = code_shift1_device(c, br, sc, &p.code_shift1, &p.code2_shift1);
        if (rc) return rc;
    }
    clk.mark("setup", (int64_t)live.size());
    if (packed) shift = false;  // as run_sweep: packed sweeps run skewed stripes
    int64_t launches = 0;
    int rc = run_sweep(c, kStage1, p, regions, W, 0, c->ev[0], c->ev[1], wps, &launches, 0, packed, true,
                       skew ? (shift ? 2 : 1) : 0);
    if (rc) return rc;
    clk.mark("stage1", launches);
    for (int64_t k : live) {
        out[k].precision_used = packed ? 16 : 32;
        out[k].score_codes = p.code_shift1 != 0;
    }

    // ---- argmax of the row maxima, per pair ---------------------------------------
    std::vector<ArgmaxJob> aj;  // source of an asynchronous copy: lives until the sync below
    {
        for (int64_t k : live)
            aj.push_back(ArgmaxJob{rowmax + jobs[k].rowmax_off, jobs[k].m, (unsigned long long*)c->key.p + k,
                                   rowseg + jobs[k].rowmax_off, segs + k});
        HCK(c, ensure(c->argjobs, sizeof(ArgmaxJob) * aj.size()));
        HCK(c, hipMemcpyAsync(c->argjobs.p, aj.data(), sizeof(ArgmaxJob) * aj.size(), hipMemcpyHostToDevice, c->stream));
        HCK(c, hipMemsetAsync(c->key.p, 0, sizeof(unsigned long long) * nj, c->stream));
        HCK(c, hipMemsetAsync(segs, 0xFF, sizeof(int32_t) * nj, c->stream));  // -1: search every block
        const int bpj = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(max_m, 16384)));
        HCK(c, launch_argmax_rows((const ArgmaxJob*)c->argjobs.p, (int)aj.size(), bpj, c->stream));
        HCK(c, launch_end_segment((const ArgmaxJob*)c->argjobs.p, (int)aj.size(), c->stream));
    }
    std::vector<unsigned long long> keys((size_t)nj, 0);
    std::vector<int32_t> endseg((size_t)nj, -1);
    HCK(c, hipMemcpyAsync(keys.data(), c->key.p, sizeof(unsigned long long) * nj, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipMemcpyAsync(endseg.data(), segs, sizeof(int32_t) * nj, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    const float ms_fwd = elapsed(c->ev[0], c->ev[1]);
    clk.mark("argmax", (int64_t)aj.size());
    std::vector<int64_t> scored;
    for (int64_t k : live) {
        Job& j = jobs[k];
        j.best = (int32_t)(keys[k] >> 32);
        j.bi = keys[k] ? (int64_t)(0xFFFFFFFFu - (uint32_t)(keys[k] & 0xFFFFFFFFu)) : -1;
        out[k].score = j.best > 0 ? j.best : 0;
        out[k].forward_launches = (int32_t)launches;
        out[k].forward_tile_cols = (int32_t)W;
        out[k].ms_forward = ms_fwd;
        out[k].ms_total = ms_fwd;
        if (j.best > 0) scored.push_back(k);
    }
    if (scored.empty()) return SWH_OK;

    // ---- end-cell search: per pair, the band above the end row, one region per
    // checkpoint-column block.  Only the blocks over the first stage-1 segment (tile
    // columns [g W, (g+1) W)) holding the end row's maximum are searched: no column left
    // of it holds the score in that row, and the first one that does lies inside it. ------
    {
        regions.clear();
        int64_t wmax = 0;
        for (int64_t k : scored) {
            const Job& j = jobs[k];
            const int64_t br0 = (j.bi / j.pl.K) * j.pl.K;
            const int64_t nq = ceil_div(j.n, std::min(j.pl.Kc, j.n));
            const int64_t bw = std::min(j.pl.Kc, j.n);
            wmax = std::max(wmax, bw);
            int64_t q0 = 0, q1 = nq - 1;
            const int64_t g = endseg[(size_t)k];
            // the segment's lane-0 columns: [gW, (g+1)W), in shifted stripes 64 columns left
            // per stripe above the end row's (see TileCtx in sw_kernels.hip)
            const bool sk = skew && ceil_div(j.n, W) >= 2;
            const int64_t x0 = g * W - ((sk && shift) ? 64 * (j.bi / kStripeRows) : 0);
            if (g >= 0 && x0 < j.n) {
                // the end row's columns in it, in a skewed stripe shifted left by the row's lane
                const int64_t l = sk ? (j.bi % kStripeRows) / kRowsPerLane : 0;
                const int64_t lo = std::max<int64_t>(0, x0 - l);
                const int64_t hi = (x0 + W >= j.n) ? j.n - 1 : std::min(j.n, x0 + W - l) - 1;
                q0 = lo / bw;
                q1 = hi / bw;
            }
            for (int64_t q = q0; q <= q1; ++q) {
                Region r{};
                r.A = j.A;
                r.B = j.B;
                r.m = j.m;
                r.n = j.n;
                r.r0 = br0;
                r.r1 = j.bi;
                r.c0 = q * bw;
                r.c1 = std::min(j.n, r.c0 + bw) - 1;
                r.top = br0 > 0 ? srow + j.srow_off + (br0 / j.pl.K - 1) * j.n + r.c0 : nullptr;
                r.left = r.c0 > 0 ? scol + j.scol_off + (r.c0 / j.pl.Kc - 1) * j.m : nullptr;
                r.K = j.pl.K;
                r.Kc = j.pl.Kc;
                r.target_row = j.bi;
                r.target = j.best;
                r.found = (int32_t*)c->found.p + k;
                regions.push_back(r);
            }
        }
        std::vector<int32_t> big((size_t)nj, INT32_MAX);  // copied before run_sweep's sync
        HCK(c, hipMemcpyAsync(c->found.p, big.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, c->stream));
        const int64_t Ws = eo.batch ? W : tile_cols(c, wmax, eo.tile_cols);
        rc = run_sweep(c, kSearch, p, regions, Ws);
        if (rc) return rc;
        std::vector<int32_t> found((size_t)nj, -1);
        HCK(c, hipMemcpyAsync(found.data(), c->found.p, sizeof(int32_t) * nj, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipEventRecord(c->ev[3], c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        const float ms_search = elapsed(c->ev[1], c->ev[3]);
        clk.mark("search", (int64_t)regions.size());
        for (int64_t k : scored) {
            Job& j = jobs[k];
            if (found[k] == INT32_MAX || found[k] < 0) {
                j.status = SWH_ERR_INTERNAL;
                continue;
            }
            j.bj = found[k];
            out[k].end_row = j.bi;
            out[k].end_col = j.bj;
            out[k].ms_end_search = ms_search;
            out[k].ms_total = ms_fwd + ms_search;
        }
    }

    // ---- traceback rounds ----------------------------------------------------------
    // Each round recomputes, for every walking pair, the checkpoint block of its current
    // cell (cut at the cell) and up to L blocks above and to the left, all in one sweep
    // (independent regions, each from its own checkpoints); then one walk per pair
    // follows the path through its blocks until it leaves them.
    std::vector<int64_t> active;
    for (int64_t k : scored)
        if (jobs[k].traceback && jobs[k].status == SWH_OK) active.push_back(k);
    if (active.empty()) return SWH_OK;
    HCK(c, hipEventRecord(c->ev[4], c->stream));
    HCK(c, ensure(c->ops, (size_t)ops_elems));
    HCK(c, ensure(c->xps, sizeof(XPoint) * (size_t)xps_elems));
    std::vector<WalkState> states((size_t)nj);
    for (int64_t k : active) {
        WalkState& st = states[(size_t)k];
        st = WalkState{};
        st.i = jobs[k].bi;
        st.j = jobs[k].bj;
        st.start_i = st.i;
        st.start_j = st.j;
        st.last_row = st.i;
        st.v = jobs[k].best;
    }
    int64_t wmax = 0;
    for (int64_t k : active) wmax = std::max(wmax, std::min(jobs[k].n, jobs[k].pl.Kc));
    const int64_t Wtb = eo.batch ? W : tile_cols(c, wmax, eo.tile_cols);
    const int64_t tstride = ceil_div(Wtb + 63, 64) * 64;
    const int64_t Gblk = ceil_div(wmax, Wtb);
    const bool tb_packed = eo.batch && packed && Wtb <= w16;
    const int tb_wps = eo.batch ? wps : 1;  // a batch's codes sweeps run as many waves per SIMD as its stage 1
    auto blk_dirs = [&](const Job& j) {
        return ceil_div(std::min(j.m, j.pl.K), kStripeRows) * Gblk * tstride * 64 * (int64_t)sizeof(uint16_t);
    };
    // direction codes per round: a quarter of the free memory, between 2 and 32 GiB
    int64_t dirs_cap = kTracebackDirsCap;
    if (c->dirs_cap_fixed > 0) {
        dirs_cap = c->dirs_cap_fixed;  // the codes' share of a fixed scratch area
    } else {
        size_t freeb = 0, totalb = 0;
        (void)hipMemGetInfo(&freeb, &totalb);
        dirs_cap = std::max<int64_t>(kTracebackDirsCap, std::min<int64_t>((int64_t)(freeb / 4), INT64_C(32) << 30));
    }
    int64_t guard = 0;
    int64_t max_blocks = 0, max_blk_rows = 0;
    for (int64_t k : active) max_blk_rows = std::max(max_blk_rows, jobs[k].pl.K);
    for (int64_t k : active)
        max_blocks = std::max(max_blocks, jobs[k].m / jobs[k].pl.K + jobs[k].n / jobs[k].pl.Kc + 8);
    const int64_t guard_max = 4 * (max_blocks + (int64_t)active.size());  // rounds, at most
    if (eo.batch && c->dirs_cap_fixed == 0) {
        rc = device_rounds(c, jobs, active, states, p, srow, scol, Wtb, tstride, Gblk, tb_packed, tb_wps, dirs_cap,
                           guard, guard_max);
        if (rc) return rc;
    }
    const auto t_rounds = std::chrono::steady_clock::now();
    while (!active.empty()) {
        if (++guard > guard_max) return SWH_ERR_INTERNAL;
        // lookahead and the pairs walked this round, within the direction-code budget
        int64_t L = std::max(0, eo.lookahead);
        int64_t need = 0;
        for (int64_t k : active) need += blk_dirs(jobs[k]);
        if (eo.lookahead_auto) {
            // batches: blocks ahead of the walk while the codes sweep's launches still fit the
            // resident waves (the rounds of the last, long pairs are latency-bound: their extra
            // blocks cost no time and save rounds); none while many pairs walk
            int64_t tiles = 0;  // tiles of one block of every walking pair
            for (int64_t k : active) tiles += ceil_div(std::min(jobs[k].m, jobs[k].pl.K), kStripeRows) * Gblk;
            const int64_t diags = ceil_div(std::min(max_m, max_blk_rows), kStripeRows) + Gblk - 1;
            const int64_t slots = (int64_t)kSweepWaves * tb_wps * (tb_packed ? 2 : 1) *
                                  std::max(1, c->props.multiProcessorCount);
            L = 0;
            for (int64_t l : {1, 2, 3, 5, 7})
                if (lookahead_blocks(l) * tiles <= slots * diags) L = l;
        }
        while (L > 0 && lookahead_blocks(L) * need > dirs_cap) --L;
        std::vector<int64_t> round;
        int64_t used = 0;
        for (int64_t k : active) {
            const int64_t b = lookahead_blocks(L) * blk_dirs(jobs[k]);
            if (!round.empty() && used + b > dirs_cap) break;
            round.push_back(k);
            used += b;
        }
        const int64_t nr = (int64_t)round.size();
        std::vector<WalkRegion> wregs((size_t)(nr * (L + 1) * (L + 1)));
        std::vector<WalkArgs> wargs((size_t)nr);
        regions.clear();
        std::vector<std::pair<int64_t, int64_t>> slots;  // (wregs index, dirs offset)
        int64_t dirs_elems = 0;
        for (int64_t x = 0; x < nr; ++x) {
            const Job& j = jobs[round[x]];
            const WalkState& st = states[(size_t)round[x]];
            const int64_t brow = st.i / j.pl.K, bcol = st.j / j.pl.Kc;
            for (int64_t da = 0; da <= L; ++da) {
                for (int64_t db = 0; db <= L; ++db) {
                    const int64_t wi = x * (L + 1) * (L + 1) + da * (L + 1) + db;
                    WalkRegion& w = wregs[(size_t)wi];
                    w = WalkRegion{};
                    if (brow - da < 0 || bcol - db < 0 || !lookahead_block(L, da, db)) continue;
                    Region r{};
                    r.A = j.A;
                    r.B = j.B;
                    r.m = j.m;
                    r.n = j.n;
                    r.r0 = (brow - da) * j.pl.K;
                    r.r1 = da == 0 ? st.i : std::min(j.m, r.r0 + j.pl.K) - 1;
                    r.c0 = (bcol - db) * j.pl.Kc;
                    const int64_t c_end = std::min(j.n, r.c0 + j.pl.Kc) - 1;  // the block's last column
                    r.c1 = db == 0 ? st.j : c_end;
                    w.cols = r.c1 - r.c0 + 1;
                    // The block cut at the current cell would end in a ragged tile; the columns
                    // right of the cell are computed too, up to whole tiles within the block, so
                    // that the packed sweep's two tiles per wave have equal widths (a wave whose
                    // tiles differ in width takes per-value selects every step).  The walk
                    // never visits them.
                    r.c1 = std::min(c_end, r.c0 + ceil_div(w.cols, Wtb) * Wtb - 1);
                    r.top = r.r0 > 0 ? srow + j.srow_off + (r.r0 / j.pl.K - 1) * j.n + r.c0 : nullptr;
                    r.left = r.c0 > 0 ? scol + j.scol_off + (r.c0 / j.pl.Kc - 1) * j.m : nullptr;
                    r.K = j.pl.K;
                    r.Kc = j.pl.Kc;
                    w.r0 = r.r0;
                    w.c0 = r.c0;
                    w.rows = r.r1 - r.r0 + 1;
                    slots.emplace_back(wi, dirs_elems);
                    dirs_elems += ceil_div(w.rows, kStripeRows) * Gblk * tstride * 64;
                    regions.push_back(r);
                }
            }
        }
        HCK(c, ensure(c->dirs, sizeof(uint16_t) * (size_t)std::max<int64_t>(dirs_elems, 1)));
        for (size_t q = 0; q < slots.size(); ++q) {
            uint16_t* d = (uint16_t*)c->dirs.p + slots[q].second;
            regions[q].dirs = d;
            wregs[(size_t)slots[q].first].dirs = d;
        }
        // batches: packed 16-bit codes sweep (two tiles per wave) when stage 1 ran packed and the
        // tiles fit, two waves per SIMD; a single pair's rounds are latency-bound: int32, one wave
        // no synchronisation: the walk's arguments are built while the codes sweep runs (the
        // round ends with one synchronisation for the walk states)
        int64_t round_launches = 0;
        rc = run_sweep(c, kDirs, p, regions, Wtb, Gblk, nullptr, nullptr, tb_wps, &round_launches, tstride, tb_packed,
                       false);
        if (rc) return rc;
        if (c->trace) {
            int64_t cells = 0, tiles = 0;
            for (const Region& r : regions) {
                cells += (r.r1 - r.r0 + 1) * (r.c1 - r.c0 + 1);
                tiles += ceil_div(r.r1 - r.r0 + 1, kStripeRows) * ceil_div(r.c1 - r.c0 + 1, Wtb);
            }
            fprintf(stderr, "swalign: round %lld: %lld pairs, lookahead %lld, %zu blocks, %lld tiles, %.3g cells, "
                    "%lld launches (from %.3f ms)\n", (long long)guard, (long long)nr, (long long)L, regions.size(),
                    (long long)tiles, (double)cells, (long long)round_launches,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_rounds).count());
        }
        // device arrays of the round; allocated before their addresses go into the walk args
        HCK(c, ensure(c->wregs, sizeof(WalkRegion) * wregs.size()));
        HCK(c, ensure(c->wargs, sizeof(WalkArgs) * wargs.size()));
        HCK(c, ensure(c->walk, sizeof(WalkState) * (size_t)nr));
        for (int64_t x = 0; x < nr; ++x) {
            const Job& j = jobs[round[x]];
            const WalkState& st = states[(size_t)round[x]];
            WalkArgs& wa = wargs[(size_t)x];
            wa = WalkArgs{};
            wa.A = j.A;
            wa.B = j.B;
            wa.regs = (const WalkRegion*)c->wregs.p + x * (L + 1) * (L + 1);
            wa.L = L;
            wa.brow = st.i / j.pl.K;
            wa.bcol = st.j / j.pl.Kc;
            wa.tstride = tstride;
            wa.G = Gblk;
            wa.W = Wtb;
            wa.K = j.pl.K;
            wa.Kc = j.pl.Kc;
            wa.srow = j.pl.nsr ? srow + j.srow_off : nullptr;
            wa.n = j.n;
            wa.st = (WalkState*)c->walk.p + x;
            wa.ops = (uint8_t*)c->ops.p + j.ops_off;
            wa.xps = (XPoint*)c->xps.p + j.xps_off;
            wa.xcap = j.xcap;
            wa.sc = p.sc;
        }
        // the round's walk states, compacted
        std::vector<WalkState> rst((size_t)nr);
        for (int64_t x = 0; x < nr; ++x) rst[(size_t)x] = states[(size_t)round[x]];
        HCK(c, hipMemcpyAsync(c->walk.p, rst.data(), sizeof(WalkState) * nr, hipMemcpyHostToDevice, c->stream));
        HCK(c, hipMemcpyAsync(c->wregs.p, wregs.data(), sizeof(WalkRegion) * wregs.size(), hipMemcpyHostToDevice,
                              c->stream));
        HCK(c, hipMemcpyAsync(c->wargs.p, wargs.data(), sizeof(WalkArgs) * wargs.size(), hipMemcpyHostToDevice,
                              c->stream));
        HCK(c, launch_walk((const WalkArgs*)c->wargs.p, (int)nr, many_walks(c, nr), c->stream));
        std::vector<WalkState> after((size_t)nr);
        HCK(c, hipMemcpyAsync(after.data(), c->walk.p, sizeof(WalkState) * nr, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        std::vector<char> finished((size_t)nj, 0);
        for (int64_t x = 0; x < nr; ++x) {
            const int64_t k = round[x];
            const WalkState& pv = rst[(size_t)x];
            const WalkState& st = after[(size_t)x];
            states[(size_t)k] = st;
            if (st.done) {
                finished[(size_t)k] = 1;
            } else if (st.i == pv.i && st.j == pv.j && st.layer == pv.layer) {
                jobs[k].status = SWH_ERR_INTERNAL;  // no progress
                finished[(size_t)k] = 1;
            }
        }
        std::vector<int64_t> next;
        next.reserve(active.size());
        for (int64_t k : active)
            if (!finished[(size_t)k]) next.push_back(k);
        active.swap(next);
    }
    HCK(c, hipEventRecord(c->ev[5], c->stream));
    clk.mark("traceback", guard);

    // ---- results: ops (reversed) -> run-length CIGAR, crosspoints start -> end ---------
    // Many alignments (a batch): the CIGAR text is built on the device (k_cigar, one wave
    // per alignment: a count pass, offsets on the host, a write pass) and only the text is
    // read back, a fraction of the ops' bytes (65,536-pair batch: 105 MB instead of 492 MB,
    // and 8 ms instead of 38 ms).  Few alignments: each wave would walk a long alignment
    // alone (2 ms per pass for the 1 Mbp pair's 204k ops), so the ops are read back and
    // encoded on the host threads (0.2 ms).  The allocation-free entry point, whose buffers
    // are fixed slices, reads the ops back into the caller's memory.
    int64_t n_tb = 0;
    for (int64_t k : scored) n_tb += (jobs[k].traceback && jobs[k].status == SWH_OK) ? 1 : 0;
    if (n_tb == 0) return SWH_OK;
    const bool gpu_cigar = !c->fixed_memory && !c->ops_dst && n_tb >= kGpuCigarMinJobs;
    uint8_t* ops_h = c->ops_dst;  // the caller's memory in the allocation-free entry point
    std::vector<CigarJob> cj;
    std::vector<int64_t> cj_at((size_t)nj, -1);  // job -> index in cj
    const char* text_h = nullptr;
    std::vector<XPoint> xps_h(eo.want_xp ? (size_t)xps_elems : 0);
    if (gpu_cigar) {
        for (int64_t k : scored) {
            const Job& j = jobs[k];
            if (!j.traceback || j.status != SWH_OK) continue;
            cj_at[(size_t)k] = (int64_t)cj.size();
            cj.push_back(CigarJob{(const uint8_t*)c->ops.p + j.ops_off, states[(size_t)k].nops, 0, 0});
        }
        const int64_t ncj = (int64_t)cj.size();
        HCK(c, ensure(c->cigjobs, sizeof(CigarJob) * (size_t)std::max<int64_t>(ncj, 1)));
        HCK(c, hipMemcpyAsync(c->cigjobs.p, cj.data(), sizeof(CigarJob) * ncj, hipMemcpyHostToDevice, c->stream));
        HCK(c, launch_cigar((CigarJob*)c->cigjobs.p, ncj, nullptr, c->stream));
        HCK(c, hipMemcpyAsync(cj.data(), c->cigjobs.p, sizeof(CigarJob) * ncj, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        int64_t total = 0;
        for (CigarJob& q : cj) {
            q.at = total;
            total += q.len;
        }
        HCK(c, ensure(c->cigtext, (size_t)std::max<int64_t>(total, 1)));
        HCK(c, hipMemcpyAsync(c->cigjobs.p, cj.data(), sizeof(CigarJob) * ncj, hipMemcpyHostToDevice, c->stream));
        HCK(c, launch_cigar((CigarJob*)c->cigjobs.p, ncj, (char*)c->cigtext.p, c->stream));
        if ((size_t)total > c->host_ops_cap) {  // pinned, grow-only: the copy runs at full link speed
            if (c->host_ops) (void)hipHostFree(c->host_ops);
            c->host_ops = nullptr;
            c->host_ops_cap = 0;
            HCK(c, hipHostMalloc((void**)&c->host_ops, (size_t)total, 0));
            c->host_ops_cap = (size_t)total;
        }
        HCK(c, hipMemcpyAsync(c->host_ops, c->cigtext.p, (size_t)total, hipMemcpyDeviceToHost, c->stream));
        text_h = (const char*)c->host_ops;
        clk.mark("cigar_gpu", total);
    } else {
        if (!ops_h) {  // pinned, grow-only: the copy runs at full link speed
            if ((size_t)ops_elems > c->host_ops_cap) {
                if (c->host_ops) (void)hipHostFree(c->host_ops);
                c->host_ops = nullptr;
                c->host_ops_cap = 0;
                HCK(c, hipHostMalloc((void**)&c->host_ops, (size_t)ops_elems, 0));
                c->host_ops_cap = (size_t)ops_elems;
            }
            ops_h = c->host_ops;
        } else if ((size_t)ops_elems > c->ops_dst_cap) {
            return SWH_ERR_INTERNAL;
        }
        HCK(c, hipMemcpyAsync(ops_h, c->ops.p, (size_t)ops_elems, hipMemcpyDeviceToHost, c->stream));
    }
    if (eo.want_xp)
        HCK(c, hipMemcpyAsync(xps_h.data(), c->xps.p, sizeof(XPoint) * xps_elems, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    const float ms_tb = elapsed(c->ev[4], c->ev[5]);
    clk.mark("readback", gpu_cigar ? (int64_t)c->host_ops_cap : ops_elems);
    // results on the host threads, pairs split between them
    std::vector<int> fail(16, 0);
    auto build = [&](int tid, int nth) {
        for (size_t x = (size_t)tid; x < scored.size(); x += (size_t)nth) {
            const int64_t k = scored[x];
            Job& j = jobs[k];
            swh_result& r = out[k];
            if (!j.traceback || j.status != SWH_OK) continue;
            r.ms_traceback = ms_tb;
            r.ms_total = r.ms_forward + r.ms_end_search + ms_tb;
            const CigarJob* q = gpu_cigar ? &cj[(size_t)cj_at[(size_t)k]] : nullptr;
            if (fill_alignment(r, states[(size_t)k], gpu_cigar ? nullptr : ops_h + j.ops_off,
                               q ? text_h + q->at : nullptr, q ? q->len : 0,
                               eo.want_xp ? xps_h.data() + j.xps_off : nullptr, j.xcap, j.best, eo.want_xp,
                               c->cigar_dst, c->cigar_dst_cap) != SWH_OK) {
                fail[(size_t)tid] = 1;
                return;
            }
        }
    };
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(),
                                                                (int64_t)scored.size() / 64 + 1}));
    if (nth == 1) {
        build(0, 1);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t) pool.emplace_back(build, t, nth);
        for (auto& t : pool) t.join();
    }
    for (int f : fail)
        if (f) return c->cigar_dst ? SWH_ERR_INVALID_ARG : SWH_ERR_OUT_OF_MEMORY;
    clk.mark("cigars", (int64_t)scored.size());
    return SWH_OK;
}

void result_init(swh_result* r) { memset(r, 0, sizeof *r); }

// The result's start cell, CIGAR and crosspoints.  The CIGAR is either the text the device
// built (text, text_len) or the run-length encoding of the ops read back (ops, written
// end -> start).
int fill_alignment(swh_result& r, const WalkState& st, const uint8_t* ops, const char* text, int64_t text_len,
                   const XPoint* xs, int64_t xcap, int32_t best, bool want_xp, char* dst, int64_t dst_cap) {
    r.start_row = st.start_i;
    r.start_col = st.start_j;
    r.n_ops = st.nops;
    if (text) {
        char* out = (char*)malloc((size_t)text_len + 1);
        if (!out) return SWH_ERR_OUT_OF_MEMORY;
        memcpy(out, text, (size_t)text_len);
        out[text_len] = 0;
        r.cigar = out;
        r.cigar_len = text_len;
    } else {
        // a run of n ops takes at most (digits of n) + 1 <= 2n characters
        if (dst && dst_cap < 2 * st.nops + 1) return SWH_ERR_INVALID_ARG;
        char* out = dst ? dst : (char*)malloc((size_t)(2 * st.nops + 1));
        if (!out) return SWH_ERR_OUT_OF_MEMORY;
        size_t len = 0;
        for (int64_t b = st.nops - 1; b >= 0;) {  // ops were written end -> start
            int64_t e = b;
            while (e >= 0 && ops[e] == ops[b]) --e;
            char digits[24];
            int nd = 0;
            for (uint64_t v = (uint64_t)(b - e); v; v /= 10) digits[nd++] = (char)('0' + v % 10);
            while (nd) out[len++] = digits[--nd];
            out[len++] = (char)ops[b];
            b = e;
        }
        out[len] = 0;
        r.cigar_len = (int64_t)len;
        if (dst) return SWH_OK;  // the caller's buffer: r.cigar stays NULL, nothing to free
        char* shrunk = (char*)realloc(out, len + 1);
        r.cigar = shrunk ? shrunk : out;
    }
    if (want_xp) {
        const int64_t nx = std::min<int64_t>(st.nxp, xcap);
        r.xpoints = (swh_xpoint*)malloc(sizeof(swh_xpoint) * (size_t)(nx + 1));
        if (!r.xpoints) return SWH_ERR_OUT_OF_MEMORY;
        for (int64_t q = 0; q < nx; ++q) {
            const XPoint& xp = xs[nx - 1 - q];
            r.xpoints[q] = swh_xpoint{xp.row, xp.col, xp.score, xp.layer};
        }
        r.xpoints[nx] = swh_xpoint{r.end_row, r.end_col, best, 0};
        r.n_xpoints = nx + 1;
    }
    return SWH_OK;
}

EngineOpts engine_opts(const swh_opts* o, bool batch) {
    EngineOpts eo;
    eo.tile_cols = o ? o->tile_cols : 0;
    eo.precision = o ? o->precision : 0;
    eo.wps = o && o->waves_per_simd >= 1 && o->waves_per_simd <= kMaxSweepWps ? o->waves_per_simd : 0;
    eo.want_xp = o && o->want_xpoints;
    eo.batch = batch;
    eo.host_ckpt = o && o->host_checkpoints;
    eo.lookahead = batch ? 0 : 7;
    eo.lookahead_auto = batch && !(o && o->tb_lookahead != 0);
    if (o && o->tb_lookahead != 0) eo.lookahead = std::max(0, o->tb_lookahead);
    // skewed stage-1 stripes by default (single pairs: where the wavefront model prefers them);
    // SWALIGN_SKEW=0 / 1 / 2 forces rectangular tiles / skewed stripes / shifted stripes
    const char* sk = getenv("SWALIGN_SKEW");
    eo.skew = sk ? atoi(sk) != 0 : true;
    eo.skew_forced = sk && atoi(sk) != 0;
    eo.layout = sk && (atoi(sk) == 1 || atoi(sk) == 2) ? atoi(sk) : 0;
    return eo;
}

// A job for device-resident sequences (orientation of src/SWutils.cpp:84-88).
Job make_job(const uint8_t* d1, int64_t len1, const uint8_t* d2, int64_t len2, bool traceback) {
    Job j;
    j.swapped = len1 < len2;
    j.A = j.swapped ? d2 : d1;
    j.B = j.swapped ? d1 : d2;
    j.m = j.swapped ? len2 : len1;
    j.n = j.swapped ? len1 : len2;
    j.traceback = traceback;
    return j;
}

}  // namespace swh_host

using namespace swh_host;

extern "C" {

const char* swh_strerror(int status) {
    switch (status) {
        case SWH_OK: return "OK";
        case SWH_ERR_INVALID_ARG: return "Invalid argument.";
        case SWH_ERR_SCORING: return "One or more scoring values were not usable!";
        case SWH_ERR_NO_DEVICE: return "No HIP capable device detected.";
        case SWH_ERR_HIP: return "HIP runtime call failed.";
        case SWH_ERR_OUT_OF_MEMORY: return "Device memory allocation failed.";
        case SWH_ERR_BUDGET: return "Request exceeds the HBM budget.";
        case SWH_ERR_INTERNAL: return "Internal traceback inconsistency.";
        case SWH_ERR_NOT_FOUND: return "Not found.";
        case SWH_PENDING: return "Pending.";
        case SWH_ERR_SCRATCH:
            return "The pair does not fit the fixed scratch memory of the allocation-free entry point.";
        default: return "Unknown status.";
    }
}

int swh_ctx_create(int device, swh_ctx** out) {
    if (!out) return SWH_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        (void)hipGetLastError();
        return SWH_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) return SWH_ERR_INVALID_ARG;
    swh_ctx* c = new swh_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&c->props, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < 8 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
    if (e != hipSuccess) {
        delete c;
        return SWH_ERR_HIP;
    }
    // Reserve LDS per sweep workgroup so that exactly w of them fit a CU (w = 1..4): the
    // waves of a launch then spread evenly, w per SIMD (an uneven spread leaves some SIMDs
    // with more tiles than others, and a launch lasts as long as its busiest SIMD).
    const char* env = getenv("SWALIGN_SWEEP_WAVES_PER_SIMD");
    if (env && atoi(env) >= 1 && atoi(env) <= kMaxSweepWps) c->wps_fixed = atoi(env);
    // SWALIGN_SYMBOL_CODES=0 keeps the byte-compare score in every sweep
    const char* codes = getenv("SWALIGN_SYMBOL_CODES");
    c->codes_off = codes && atoi(codes) == 0;
    const char* trace = getenv("SWALIGN_TRACE");
    c->trace = trace && atoi(trace) != 0;
    const size_t per_cu = (size_t)c->props.maxSharedMemoryPerMultiProcessor, stat = sweep_static_lds();
    for (int w = 1; w <= kMaxSweepWps; ++w) {
        const size_t want = per_cu / (size_t)(w + 1) + 1024;
        if (want > stat && want <= (size_t)c->props.sharedMemPerBlock) c->lds_pad[w] = want - stat;
    }
    if (c->lds_pad[1] && sweep_allow_lds(c->lds_pad[1]) != hipSuccess) {
        (void)hipGetLastError();
        for (size_t& pad : c->lds_pad) pad = 0;
    }
    // the allocation-free entry point's scratch array on this device (resolved once per
    // device); a failure here only disables that entry point
    c->scratch_status = scratch_resolve(device, &c->scratch);
    *out = c;
    return SWH_OK;
}

int swh_ctx_destroy(swh_ctx* c) {
    if (!c) return SWH_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    DevBuf* bufs[] = {&c->seqA, &c->seqB, &c->rbuf, &c->cb, &c->cor, &c->rowmax, &c->srow, &c->scol,
                      &c->dirs, &c->regions, &c->key, &c->found, &c->walk, &c->ops, &c->xps, &c->wregs,
                      &c->argjobs, &c->wargs, &c->seqs, &c->spans, &c->alpha, &c->ranges, &c->rowseg,
                      &c->segs, &c->tb_pairs, &c->tb_states, &c->tb_act, &c->tb_pre, &c->cigjobs,
                      &c->cigtext};
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->staging) (void)hipHostFree(c->staging);
    if (c->host_ops) (void)hipHostFree(c->host_ops);
    if (c->pinned_spans) (void)hipHostFree(c->pinned_spans);
    for (HostBuf* b : {&c->hsrow, &c->hscol})
        if (b->h) (void)hipHostFree(b->h);
    delete c;
    return SWH_OK;
}

int swh_device_info(swh_ctx* c, char* name, int64_t name_cap, int32_t* cus, int32_t* clock_khz, int64_t* hbm) {
    if (!c) return SWH_ERR_INVALID_ARG;
    if (name && name_cap > 0) {
        snprintf(name, (size_t)name_cap, "%s", c->props.name);
    }
    if (cus) *cus = c->props.multiProcessorCount;
    if (clock_khz) *clock_khz = c->props.clockRate;
    if (hbm) *hbm = (int64_t)c->props.totalGlobalMem;
    return SWH_OK;
}

int swh_align_pair_device(swh_ctx* c, const uint8_t* d1, int64_t len1, const uint8_t* d2, int64_t len2,
                          const swh_scoring* sc, const swh_opts* o, swh_result* out) {
    if (!c || !out || len1 < 0 || len2 < 0 || (len1 && !d1) || (len2 && !d2)) return SWH_ERR_INVALID_ARG;
    result_init(out);
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    HCK(c, hipSetDevice(c->device));
    std::vector<Job> jobs{make_job(d1, len1, d2, len2, !(o && o->score_only))};
    int rc = choose_plan(jobs[0].m, jobs[0].n, o, default_budget(c, o), &jobs[0].pl);
    if (rc) return rc;
    rc = run_jobs(c, jobs, sc, engine_opts(o, false), out);
    return rc ? rc : jobs[0].status;
}

int swh_align_pair(swh_ctx* c, const uint8_t* s1, int64_t len1, const uint8_t* s2, int64_t len2,
                   const swh_scoring* sc, const swh_opts* o, swh_result* out) {
    if (!c || !out || len1 < 0 || len2 < 0 || (len1 && !s1) || (len2 && !s2)) return SWH_ERR_INVALID_ARG;
    result_init(out);
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    HCK(c, hipSetDevice(c->device));
    HCK(c, ensure(c->seqA, (size_t)len1 + 64));
    HCK(c, ensure(c->seqB, (size_t)len2 + 64));
    if (len1) HCK(c, hipMemcpyAsync(c->seqA.p, s1, (size_t)len1, hipMemcpyHostToDevice, c->stream));
    if (len2) HCK(c, hipMemcpyAsync(c->seqB.p, s2, (size_t)len2, hipMemcpyHostToDevice, c->stream));
    return swh_align_pair_device(c, (const uint8_t*)c->seqA.p, len1, (const uint8_t*)c->seqB.p, len2, sc, o, out);
}

void swh_result_free(swh_result* r) {
    if (!r) return;
    free(r->cigar);
    free(r->xpoints);
    r->cigar = nullptr;
    r->xpoints = nullptr;
}

int swh_forward(swh_ctx* c, const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const swh_scoring* sc,
                int64_t K, int64_t Kc, int32_t* score, int64_t* end_row, int64_t* end_col, int32_t* rowmax,
                int32_t* srow_out, int32_t* scol_out) {
    if (!c || m < 0 || n < 0 || (m && !A) || (n && !B) || !score || !end_row || !end_col)
        return SWH_ERR_INVALID_ARG;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    if (K <= 0 || K % kStripeRows != 0 || Kc < 64 || (Kc & (Kc - 1)) != 0) return SWH_ERR_INVALID_ARG;
    HCK(c, hipSetDevice(c->device));
    HCK(c, ensure(c->seqA, (size_t)m + 64));
    HCK(c, ensure(c->seqB, (size_t)n + 64));
    if (m) HCK(c, hipMemcpyAsync(c->seqA.p, A, (size_t)m, hipMemcpyHostToDevice, c->stream));
    if (n) HCK(c, hipMemcpyAsync(c->seqB.p, B, (size_t)n, hipMemcpyHostToDevice, c->stream));
    // A and B are already oriented: one job without the swap rule
    Job j;
    j.A = (const uint8_t*)c->seqA.p;
    j.B = (const uint8_t*)c->seqB.p;
    j.m = m;
    j.n = n;
    j.traceback = false;
    j.pl = Plan{K, Kc, m / K, n / Kc};
    std::vector<Job> jobs{j};
    swh_result r;
    result_init(&r);
    const int rc = run_jobs(c, jobs, sc, engine_opts(nullptr, false), &r);
    swh_result_free(&r);
    if (rc) return rc;
    if (jobs[0].status) return jobs[0].status;
    *score = r.score;
    *end_row = r.end_row;
    *end_col = r.end_col;
    if (rowmax && m) HCK(c, hipMemcpy(rowmax, c->rowmax.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
    if (srow_out && m / K > 0)
        HCK(c, hipMemcpy(srow_out, c->srow.p, sizeof(int2) * (m / K) * n, hipMemcpyDeviceToHost));
    if (scol_out && n / Kc > 0)
        HCK(c, hipMemcpy(scol_out, c->scol.p, sizeof(int2) * (n / Kc) * m, hipMemcpyDeviceToHost));
    return SWH_OK;
}

}  // extern "C"

namespace {

int tune_jobs(swh_ctx* c, const std::vector<Job>& jobs, const swh_scoring* sc, const swh_opts* o, int64_t sample,
              swh_batch_tuning* out);

// Shared by the batch entry points: jobs over device-resident pairs, one checkpoint
// spacing for the batch (kBatchCkptRows x kBatchCkptCols, or the options, doubled to fit
// the budget); with
// opts->autotune the settings are first chosen by timing candidates on a sample.
int batch_run(swh_ctx* c, std::vector<Job>& jobs, const swh_scoring* sc, const swh_opts* o, swh_result* out,
              int32_t* statuses) {
    const int64_t budget = default_budget(c, o);
    swh_opts ob = o ? *o : swh_opts{};
    if (ob.autotune) {
        swh_batch_tuning t{};
        const int rc = tune_jobs(c, jobs, sc, &ob, 2048, &t);
        if (rc) return rc;
        ob.autotune = 0;
        ob.tile_cols = t.tile_cols;
        ob.waves_per_simd = t.waves_per_simd;
        ob.checkpoint_rows = t.checkpoint_rows;
        ob.checkpoint_cols = t.checkpoint_cols;
    }
    if (!ob.score_only) {
        int64_t K = ob.checkpoint_rows > 0 ? ceil_div(ob.checkpoint_rows, kStripeRows) * kStripeRows : kBatchCkptRows;
        int64_t Kc = ob.checkpoint_cols > 0   ? std::max<int64_t>(64, next_pow2(ob.checkpoint_cols))
                     : ob.checkpoint_rows > 0 ? next_pow2(K)
                                              : kBatchCkptCols;
        int64_t ckpt = 0;
        while (true) {
            ckpt = 0;
            for (const Job& j : jobs) ckpt += plan_mem(j.m, j.n, K, Kc);
            if (ckpt <= budget || K >= (INT64_C(1) << 40)) break;
            K *= 2;
            Kc *= 2;
        }
        if (ckpt > budget) return SWH_ERR_BUDGET;
        ob.checkpoint_rows = K;
        ob.checkpoint_cols = Kc;
    }
    for (Job& j : jobs) {
        const int rc = choose_plan(j.m, j.n, &ob, INT64_MAX / 4, &j.pl);
        if (rc) return rc;
    }
    const int rc = run_jobs(c, jobs, sc, engine_opts(&ob, true), out);
    if (statuses)
        for (size_t k = 0; k < jobs.size(); ++k) statuses[k] = rc ? rc : jobs[k].status;
    return rc;
}

// Times each candidate setting on an evenly spaced sample of the jobs (full alignments,
// results discarded) and keeps the fastest.
int tune_jobs(swh_ctx* c, const std::vector<Job>& jobs, const swh_scoring* sc, const swh_opts* o, int64_t sample,
              swh_batch_tuning* out) {
    const int64_t n = (int64_t)jobs.size();
    sample = std::max<int64_t>(1, std::min(sample > 0 ? sample : 2048, n));
    std::vector<Job> base;
    for (int64_t k = 0; k < sample; ++k) base.push_back(jobs[(size_t)(k * n / sample)]);
    const int Ws[] = {256, 384, 512, 768, 1024};
    const int Ps[] = {2, 3, 4};
    const int64_t Ks[][2] = {{256, 512}, {512, 512}, {1024, 1024}, {2048, 2048}};  // rows x columns
    std::vector<swh_result> res((size_t)sample);
    double best = 1e300;
    *out = swh_batch_tuning{};
    {  // one untimed run first, so that buffer growth is not charged to the first candidate
        std::vector<Job> js = base;
        swh_opts ot = o ? *o : swh_opts{};
        ot.autotune = 0;
        ot.checkpoint_rows = ot.checkpoint_cols = 2048;
        for (Job& j : js) {
            const int rc = choose_plan(j.m, j.n, &ot, INT64_MAX / 4, &j.pl);
            if (rc) return rc;
        }
        for (auto& r : res) result_init(&r);
        const int rc = run_jobs(c, js, sc, engine_opts(&ot, true), res.data());
        for (auto& r : res) swh_result_free(&r);
        if (rc) return rc;
    }
    int cand = 0;
    const int64_t budget = default_budget(c, o);
    for (const auto& KK : Ks) {
        const int64_t K = KK[0], Kc = KK[1];
        // a spacing whose checkpoints for the whole batch exceed the budget is not a candidate
        int64_t need = 0;
        if (!(o && o->score_only))
            for (const Job& j : jobs) need += plan_mem(j.m, j.n, K, Kc);
        if (need > budget) continue;
        for (int P : Ps) {
            const int Pu = usable_wps(c, P);  // 1 where the LDS reservation is unavailable
            if (Pu != P && P != Ps[0]) continue;
            for (int W : Ws) {
                std::vector<Job> js = base;
                swh_opts ot = o ? *o : swh_opts{};
                ot.autotune = 0;
                ot.tile_cols = W;
                ot.waves_per_simd = Pu;
                ot.checkpoint_rows = K;
                ot.checkpoint_cols = Kc;
                for (Job& j : js) {
                    const int rc = choose_plan(j.m, j.n, &ot, INT64_MAX / 4, &j.pl);
                    if (rc) return rc;
                }
                for (auto& r : res) result_init(&r);
                HCK(c, hipStreamSynchronize(c->stream));
                const auto t0 = std::chrono::steady_clock::now();
                const int rc = run_jobs(c, js, sc, engine_opts(&ot, true), res.data());
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                for (auto& r : res) swh_result_free(&r);
                if (rc == SWH_ERR_OUT_OF_MEMORY) continue;  // does not fit beside other allocations
                if (rc) return rc;
                ++cand;
                if (W == kBatchTileCols && K == kBatchCkptRows && Kc == kBatchCkptCols && Pu == usable_wps(c, kBatchWps))
                    out->ms_default = ms;
                if (ms < best) {
                    best = ms;
                    out->tile_cols = W;
                    out->waves_per_simd = Pu;
                    out->checkpoint_rows = K;
                    out->checkpoint_cols = Kc;
                    out->ms_best = ms;
                }
            }
        }
    }
    // A sample times candidates within ~1% of each other that differ by several percent on
    // the whole batch (its launches hold fewer tiles, so the longest pairs' critical path
    // weighs more): the default stays unless a candidate beats it clearly.
    if (out->ms_default > 0 && best > kTuneMargin * out->ms_default) {
        out->tile_cols = kBatchTileCols;
        out->waves_per_simd = usable_wps(c, kBatchWps);
        out->checkpoint_rows = kBatchCkptRows;
        out->checkpoint_cols = kBatchCkptCols;
        out->ms_best = out->ms_default;
    }
    out->candidates = cand;
    return cand ? SWH_OK : SWH_ERR_BUDGET;
}

}  // namespace

extern "C" {

int swh_batch_tune(swh_ctx* c, const swh_pair* pairs, int64_t npairs, int32_t device, const swh_scoring* sc,
                   const swh_opts* o, int64_t sample, swh_batch_tuning* out) {
    if (!c || !out || npairs <= 0 || !pairs) return SWH_ERR_INVALID_ARG;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    HCK(c, hipSetDevice(c->device));
    const bool tb = !(o && o->score_only);
    std::vector<Job> jobs;
    if (device) {
        for (int64_t k = 0; k < npairs; ++k)
            jobs.push_back(make_job(pairs[k].seq1, pairs[k].len1, pairs[k].seq2, pairs[k].len2, tb));
        return tune_jobs(c, jobs, sc, o, sample, out);
    }
    // host pairs: the sample is uploaded once
    const int64_t ns = std::max<int64_t>(1, std::min(sample > 0 ? sample : 2048, npairs));
    int64_t total = 0;
    for (int64_t k = 0; k < ns; ++k) {
        const swh_pair& q = pairs[k * npairs / ns];
        if (q.len1 < 0 || q.len2 < 0 || (q.len1 && !q.seq1) || (q.len2 && !q.seq2)) return SWH_ERR_INVALID_ARG;
        total += q.len1 + q.len2;
    }
    DevBuf buf;
    HCK(c, ensure(buf, (size_t)total + 64));
    std::vector<uint8_t> host((size_t)total + 1);
    int64_t off = 0;
    for (int64_t k = 0; k < ns; ++k) {
        const swh_pair& q = pairs[k * npairs / ns];
        if (q.len1) memcpy(host.data() + off, q.seq1, (size_t)q.len1);
        if (q.len2) memcpy(host.data() + off + q.len1, q.seq2, (size_t)q.len2);
        const uint8_t* d = (const uint8_t*)buf.p + off;
        jobs.push_back(make_job(d, q.len1, d + q.len1, q.len2, tb));
        off += q.len1 + q.len2;
    }
    int rc = hipMemcpy(buf.p, host.data(), (size_t)total, hipMemcpyHostToDevice) == hipSuccess ? SWH_OK : SWH_ERR_HIP;
    if (!rc) rc = tune_jobs(c, jobs, sc, o, ns, out);
    (void)hipFree(buf.p);
    return rc;
}

int swh_batch_align_device(swh_ctx* c, const swh_pair* pairs, int64_t npairs, const swh_scoring* sc,
                           const swh_opts* o, swh_result* out, int32_t* statuses) {
    if (!c || (npairs > 0 && (!pairs || !out)) || npairs < 0) return SWH_ERR_INVALID_ARG;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    HCK(c, hipSetDevice(c->device));
    for (int64_t k = 0; k < npairs; ++k) result_init(&out[k]);
    if (npairs == 0) return SWH_OK;
    std::vector<Job> jobs((size_t)npairs);
    const bool tb = !(o && o->score_only);
    for (int64_t k = 0; k < npairs; ++k) {
        const swh_pair& q = pairs[k];
        if (q.len1 < 0 || q.len2 < 0 || (q.len1 && !q.seq1) || (q.len2 && !q.seq2)) return SWH_ERR_INVALID_ARG;
        jobs[(size_t)k] = make_job(q.seq1, q.len1, q.seq2, q.len2, tb);
    }
    return batch_run(c, jobs, sc, o, out, statuses);
}

int swh_batch_align(swh_ctx* c, const swh_pair* pairs, int64_t npairs, const swh_scoring* sc,
                    const swh_opts* o, swh_result* out, int32_t* statuses) {
    if (!c || (npairs > 0 && (!pairs || !out)) || npairs < 0) return SWH_ERR_INVALID_ARG;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    HCK(c, hipSetDevice(c->device));
    for (int64_t k = 0; k < npairs; ++k) result_init(&out[k]);
    if (npairs == 0) return SWH_OK;
    // inputs packed into one pinned staging buffer and copied with one transfer
    int64_t total = 0;
    for (int64_t k = 0; k < npairs; ++k) {
        const swh_pair& q = pairs[k];
        if (q.len1 < 0 || q.len2 < 0 || (q.len1 && !q.seq1) || (q.len2 && !q.seq2)) return SWH_ERR_INVALID_ARG;
        total += q.len1 + q.len2;
    }
    HCK(c, hipStreamSynchronize(c->stream));  // the staging buffer may still feed a copy
    if ((size_t)total + 64 > c->staging_cap) {
        if (c->staging) (void)hipHostFree(c->staging);
        c->staging = nullptr;
        c->staging_cap = 0;
        HCK(c, hipHostMalloc((void**)&c->staging, (size_t)total + 64, 0));
        c->staging_cap = (size_t)total + 64;
    }
    HCK(c, ensure(c->seqs, (size_t)total + 64));
    const uint8_t* dev = (const uint8_t*)c->seqs.p;
    std::vector<Job> jobs((size_t)npairs);
    const bool tb = !(o && o->score_only);
    int64_t off = 0;
    for (int64_t k = 0; k < npairs; ++k) {
        const swh_pair& q = pairs[k];
        if (q.len1) memcpy(c->staging + off, q.seq1, (size_t)q.len1);
        if (q.len2) memcpy(c->staging + off + q.len1, q.seq2, (size_t)q.len2);
        jobs[(size_t)k] = make_job(dev + off, q.len1, dev + off + q.len1, q.len2, tb);
        off += q.len1 + q.len2;
    }
    if (total) HCK(c, hipMemcpyAsync(c->seqs.p, c->staging, (size_t)total, hipMemcpyHostToDevice, c->stream));
    return batch_run(c, jobs, sc, o, out, statuses);
}

void swh_free(void* p) { free(p); }

}  // extern "C"
