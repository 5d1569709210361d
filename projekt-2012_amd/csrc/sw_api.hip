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
#include <vector>

#include "../../include/swalign_hip.h"
#include "sw_internal.h"

using namespace swk;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct swh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    DevBuf seqA, seqB, rbuf, cb, cor, rowmax, srow, scol, dirs, regions, key, found, walk, ops, xps, wregs;
    hipEvent_t ev[8] = {};
    size_t lds_pad[3] = {0, 0, 0};  // dynamic LDS per sweep workgroup for w workgroups per CU
    int wps_fixed = 0;              // sweep waves per SIMD forced by the environment, 0 = model
    hipError_t last = hipSuccess;
};

namespace {

int map_hip(swh_ctx* c, hipError_t e) {
    if (c) c->last = e;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SWH_ERR_OUT_OF_MEMORY;
    return SWH_ERR_HIP;
}

#define HCK(ctx, x)                                   \
    do {                                              \
        hipError_t e__ = (x);                         \
        if (e__ != hipSuccess) return map_hip(ctx, e__); \
    } while (0)

hipError_t ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 64;
    if (b.p && b.cap >= bytes) return hipSuccess;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.cap = bytes;
    else (void)hipGetLastError();
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
int64_t tile_cols(swh_ctx* c, int64_t cols, int64_t want, int wps = 1) {
    if (want > 0) return want;
    const int64_t target = (int64_t)kSweepWaves * wps * std::max(1, c->props.multiProcessorCount);
    return std::max<int64_t>(256, ceil_div(cols, target));
}

// Relative time of a wave that shares its SIMD with a second sweep wave: two waves per
// SIMD retire 1.28x the steps of one (measured on MI355X: the second wave issues while the
// first waits on scalar, LDS and branch instructions).
constexpr double kTwoWaveCost = 2.0 / 1.28;

// Modelled duration (in tile steps) of a sweep over rows x cols with wps waves per SIMD:
// one launch per tile anti-diagonal, each lasting ceil(tiles / resident waves) tile times.
double sweep_model(swh_ctx* c, int64_t rows, int64_t cols, int wps) {
    const int64_t W = tile_cols(c, cols, 0, wps);
    const int64_t S = ceil_div(rows, kStripeRows), G = ceil_div(cols, W);
    const int64_t P = (int64_t)kSweepWaves * wps * std::max(1, c->props.multiProcessorCount);
    double t = 0;
    for (int64_t d = 0; d < S + G - 1; ++d) {
        const int64_t tiles = std::min(d, S - 1) - std::max<int64_t>(0, d - (G - 1)) + 1;
        t += (double)ceil_div(tiles, P) * (double)(W + 63);
    }
    return t * (wps == 2 ? kTwoWaveCost : 1.0);
}

// Sweep waves per SIMD for a rows x cols sweep: whichever of one or two the model prefers.
int choose_wps(swh_ctx* c, int64_t rows, int64_t cols) {
    if (c->wps_fixed) return c->wps_fixed;
    if (!c->lds_pad[2]) return 1;
    return sweep_model(c, rows, cols, 2) < sweep_model(c, rows, cols, 1) ? 2 : 1;
}

// Runs a sweep over `regions` with tiles of kStripeRows x W: one launch per anti-diagonal
// of the tile grid, all regions side by side.  The grid has G segments per stripe (the
// widest region's, or G_fix when given: kDirs lays codes out by it).  Optional events are
// recorded immediately before the first and after the last launch.
int run_sweep(swh_ctx* c, int mode, SweepParams p, std::vector<Region>& regions, int64_t W, int64_t G_fix = 0,
              hipEvent_t ev_before = nullptr, hipEvent_t ev_after = nullptr, int wps = 1,
              int64_t* launches = nullptr) {
    int64_t S = 0, G = 0, cb_elems = 0, rb_elems = 0, cor_elems = 0;
    for (const Region& r : regions) {
        const int64_t rows = r.r1 - r.r0 + 1, cols = r.c1 - r.c0 + 1;
        if (rows <= 0 || cols <= 0) return SWH_ERR_INTERNAL;
        S = std::max(S, ceil_div(rows, kStripeRows));
        G = std::max(G, ceil_div(cols, W));
        cb_elems += 2 * rows;
        rb_elems += 2 * cols;
    }
    if (G_fix > 0) {
        if (G > G_fix) return SWH_ERR_INTERNAL;
        G = G_fix;
    }
    cor_elems = 2 * S * (int64_t)regions.size();
    HCK(c, ensure(c->cb, sizeof(int2) * cb_elems));
    HCK(c, ensure(c->cor, sizeof(int32_t) * cor_elems));
    HCK(c, ensure(c->rbuf, sizeof(int2) * rb_elems));
    int64_t cbo = 0, rbo = 0;
    for (size_t q = 0; q < regions.size(); ++q) {
        regions[q].cb = (int2*)c->cb.p + cbo;
        regions[q].rb = (int2*)c->rbuf.p + rbo;
        regions[q].cor = (int32_t*)c->cor.p + 2 * S * (int64_t)q;
        cbo += 2 * (regions[q].r1 - regions[q].r0 + 1);
        rbo += 2 * (regions[q].c1 - regions[q].c0 + 1);
    }
    HCK(c, ensure(c->regions, sizeof(Region) * regions.size()));
    HCK(c, hipMemcpyAsync(c->regions.p, regions.data(), sizeof(Region) * regions.size(), hipMemcpyHostToDevice,
                          c->stream));
    p.regions = (const Region*)c->regions.p;
    p.nregions = (int32_t)regions.size();
    p.S = S;
    p.G = G;
    p.W = W;
    if (ev_before) HCK(c, hipEventRecord(ev_before, c->stream));
    for (int64_t d = 0; d < S + G - 1; ++d) {
        const int64_t s_lo = std::max<int64_t>(0, d - (G - 1)), s_hi = std::min(S - 1, d);
        p.d = d;
        p.s_lo = s_lo;
        p.per = (int32_t)(s_hi - s_lo + 1);
        HCK(c, launch_sweep(mode, p, c->lds_pad[wps == 2 ? 2 : 1], c->stream));
    }
    if (ev_after) HCK(c, hipEventRecord(ev_after, c->stream));
    if (launches) *launches = S + G - 1;
    return SWH_OK;
}

struct Plan {
    int64_t K, Kc, nsr, nsc;
};

int choose_plan(swh_ctx* c, int64_t m, int64_t n, const swh_opts* o, Plan* pl) {
    int64_t K = o && o->checkpoint_rows > 0 ? ceil_div(o->checkpoint_rows, kStripeRows) * kStripeRows : 0;
    int64_t Kc = o && o->checkpoint_cols > 0 ? next_pow2(o->checkpoint_cols) : 0;
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    int64_t budget = o && o->hbm_budget_bytes > 0 ? o->hbm_budget_bytes : (int64_t)(freeb / 2);
    auto mem = [&](int64_t k, int64_t kc) { return ((m / k) * n + (n / kc) * m) * (int64_t)sizeof(int2); };
    if (o && o->score_only) {
        K = INT64_C(1) << 62;
        Kc = INT64_C(1) << 62;
    } else if (K == 0) {
        K = 8192;
        if (Kc == 0) Kc = next_pow2(K);
        while (mem(K, Kc) > budget && K < (INT64_C(1) << 40)) {
            K *= 2;
            Kc *= 2;
        }
    } else if (Kc == 0) {
        Kc = next_pow2(K);
        while (mem(K, Kc) > budget && Kc < (INT64_C(1) << 40)) Kc *= 2;
    }
    if (!(o && o->score_only) && mem(K, Kc) > budget) return SWH_ERR_BUDGET;
    pl->K = K;
    pl->Kc = Kc;
    pl->nsr = m / K;
    pl->nsc = n / Kc;
    return SWH_OK;
}

// Core: A (m rows) and B (n columns) resident on the device.
int align_core(swh_ctx* c, const uint8_t* dA, int64_t m, const uint8_t* dB, int64_t n, const swh_scoring* sc,
               const swh_opts* o, swh_result* out, bool traceback = true) {
    const bool want_tb = traceback && !(o && o->score_only);
    const bool want_xp = o && o->want_xpoints;
    out->m = m;
    out->n = n;
    out->score = 0;
    out->end_row = out->end_col = out->start_row = out->start_col = -1;
    out->precision_used = 32;
    if (m == 0 || n == 0) return SWH_OK;
    Plan pl;
    int rc = choose_plan(c, m, n, o, &pl);
    if (rc) return rc;
    out->checkpoint_rows = pl.nsr ? pl.K : 0;
    out->checkpoint_cols = pl.nsc ? pl.Kc : 0;

    const int64_t want_w = o ? o->tile_cols : 0;
    HCK(c, ensure(c->rowmax, sizeof(int32_t) * m));
    HCK(c, ensure(c->srow, sizeof(int2) * std::max<int64_t>(pl.nsr * n, 1)));
    HCK(c, ensure(c->scol, sizeof(int2) * std::max<int64_t>(pl.nsc * m, 1)));
    HCK(c, ensure(c->key, 64));
    HCK(c, ensure(c->found, 64));
    int2* srow = (int2*)c->srow.p;
    int2* scol = (int2*)c->scol.p;

    Score S{sc->match, sc->mismatch, sc->gap_open, sc->gap_extend};
    SweepParams p{};
    p.A = dA;
    p.m = m;
    p.B = dB;
    p.n = n;
    p.sc = S;
    p.rowmax = (int32_t*)c->rowmax.p;
    p.srow = pl.nsr ? srow : nullptr;
    p.K = pl.K;
    p.scol = pl.nsc ? scol : nullptr;
    p.Kc = pl.Kc;

    // ---- stage 1 ------------------------------------------------------------------
    {
        Region r{};
        r.r0 = 0;
        r.r1 = m - 1;
        r.c0 = 0;
        r.c1 = n - 1;
        std::vector<Region> regions{r};
        const int wps = want_w > 0 ? 1 : choose_wps(c, m, n);
        int64_t launches = 0;
        rc = run_sweep(c, kStage1, p, regions, tile_cols(c, n, want_w, wps), 0, c->ev[0], c->ev[1], wps,
                       &launches);
        out->forward_launches = (int32_t)launches;
        if (rc) return rc;
    }
    HCK(c, hipMemsetAsync(c->key.p, 0, 8, c->stream));
    HCK(c, launch_argmax_rows((const int32_t*)c->rowmax.p, m, (unsigned long long*)c->key.p, c->stream));
    unsigned long long key = 0;
    HCK(c, hipMemcpyAsync(&key, c->key.p, 8, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    out->ms_forward = elapsed(c->ev[0], c->ev[1]);
    const int32_t best = (int32_t)(key >> 32);
    const int64_t bi = (int64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu));
    out->score = best;
    if (best <= 0) {
        out->ms_total = out->ms_forward;
        return SWH_OK;
    }

    // ---- end-cell column search (the argmax above is counted here) ---------------------
    const hipEvent_t ev_search0 = c->ev[1];
    const int64_t br0 = (bi / pl.K) * pl.K;
    const int64_t nq = ceil_div(n, pl.Kc);
    {
        // one region per checkpoint-column block: blocks start from their own checkpoints
        std::vector<Region> regions;
        regions.reserve(nq);
        for (int64_t q = 0; q < nq; ++q) {
            Region r{};
            r.r0 = br0;
            r.r1 = bi;
            r.c0 = q * pl.Kc;
            r.c1 = std::min(n, r.c0 + pl.Kc) - 1;
            r.top = br0 > 0 ? srow + (br0 / pl.K - 1) * n + r.c0 : nullptr;
            r.left = r.c0 > 0 ? scol + (r.c0 / pl.Kc - 1) * m : nullptr;
            r.found = (int32_t*)c->found.p;
            regions.push_back(r);
        }
        const int32_t big = INT32_MAX;
        HCK(c, hipMemcpyAsync(c->found.p, &big, 4, hipMemcpyHostToDevice, c->stream));
        SweepParams ps = p;
        ps.target_row = bi;
        ps.target = best;
        rc = run_sweep(c, kSearch, ps, regions, tile_cols(c, std::min(n, pl.Kc), want_w));
        if (rc) return rc;
        int32_t bj = -1;
        HCK(c, hipMemcpyAsync(&bj, c->found.p, 4, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipEventRecord(c->ev[3], c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        if (bj == INT32_MAX || bj < 0) return SWH_ERR_INTERNAL;
        out->end_row = bi;
        out->end_col = bj;
    }
    out->ms_end_search = elapsed(ev_search0, c->ev[3]);
    if (!want_tb) {
        out->ms_total = out->ms_forward + out->ms_end_search;
        return SWH_OK;
    }

    // ---- traceback ----------------------------------------------------------------
    HCK(c, hipEventRecord(c->ev[4], c->stream));
    const int64_t xcap = pl.nsr + 2;
    HCK(c, ensure(c->ops, (size_t)(m + n + 16)));
    HCK(c, ensure(c->xps, sizeof(XPoint) * xcap));
    HCK(c, ensure(c->walk, sizeof(WalkState)));
    WalkState st{};
    st.i = out->end_row;
    st.j = out->end_col;
    st.start_i = st.i;
    st.start_j = st.j;
    st.last_row = st.i;
    st.layer = 0;
    st.done = 0;
    HCK(c, hipMemcpyAsync(c->walk.p, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    // Rounds: the checkpoint block holding the current cell (cut at that cell) and the
    // full blocks up to L block-rows above and L block-columns to the left are
    // recomputed in one sweep (independent regions, each from its own checkpoints), then
    // the walk follows the path through them until it leaves them.
    const int64_t Wtb = tile_cols(c, std::min(n, pl.Kc), want_w);
    const int64_t tstride = ceil_div(Wtb + 63, 64) * 64;
    const int64_t Gblk = ceil_div(std::min(n, pl.Kc), Wtb);
    const int64_t blk_dirs = ceil_div(std::min(m, pl.K), kStripeRows) * Gblk * tstride * 64 * (int64_t)sizeof(uint16_t);
    int64_t L = o && o->tb_lookahead != 0 ? std::max(0, o->tb_lookahead) : 3;
    while (L > 0 && (L + 1) * (L + 1) * blk_dirs > kTracebackDirsCap) --L;
    std::vector<WalkRegion> wregs((size_t)((L + 1) * (L + 1)));
    HCK(c, ensure(c->wregs, sizeof(WalkRegion) * wregs.size()));
    int64_t guard = 0;
    while (!st.done) {
        if (++guard > m / std::max<int64_t>(pl.K, 1) + n / std::max<int64_t>(pl.Kc, 1) + 8) return SWH_ERR_INTERNAL;
        const int64_t brow = st.i / pl.K, bcol = st.j / pl.Kc;
        std::vector<Region> regions;
        std::vector<int64_t> slot;
        int64_t dirs_elems = 0;
        for (int64_t da = 0; da <= L; ++da) {
            for (int64_t db = 0; db <= L; ++db) {
                WalkRegion& w = wregs[(size_t)(da * (L + 1) + db)];
                w = WalkRegion{};
                if (brow - da < 0 || bcol - db < 0) continue;
                Region r{};
                r.r0 = (brow - da) * pl.K;
                r.r1 = da == 0 ? st.i : std::min(m, r.r0 + pl.K) - 1;
                r.c0 = (bcol - db) * pl.Kc;
                r.c1 = db == 0 ? st.j : std::min(n, r.c0 + pl.Kc) - 1;
                r.top = r.r0 > 0 ? srow + (r.r0 / pl.K - 1) * n + r.c0 : nullptr;
                r.left = r.c0 > 0 ? scol + (r.c0 / pl.Kc - 1) * m : nullptr;
                w.r0 = r.r0;
                w.c0 = r.c0;
                w.rows = r.r1 - r.r0 + 1;
                w.cols = r.c1 - r.c0 + 1;
                slot.push_back(dirs_elems);
                dirs_elems += ceil_div(w.rows, kStripeRows) * Gblk * tstride * 64;
                regions.push_back(r);
            }
        }
        HCK(c, ensure(c->dirs, sizeof(uint16_t) * (size_t)dirs_elems));
        for (size_t k = 0, q = 0; k < wregs.size(); ++k) {
            if (wregs[k].rows <= 0) continue;
            uint16_t* d = (uint16_t*)c->dirs.p + slot[q];
            regions[q].dirs = d;
            wregs[k].dirs = d;
            ++q;
        }
        SweepParams pd = p;
        pd.tstride = tstride;
        rc = run_sweep(c, kDirs, pd, regions, Wtb, Gblk);
        if (rc) return rc;
        HCK(c, hipMemcpyAsync(c->wregs.p, wregs.data(), sizeof(WalkRegion) * wregs.size(), hipMemcpyHostToDevice,
                              c->stream));
        WalkArgs wa{};
        wa.A = dA;
        wa.B = dB;
        wa.regs = (const WalkRegion*)c->wregs.p;
        wa.L = L;
        wa.brow = brow;
        wa.bcol = bcol;
        wa.tstride = tstride;
        wa.G = Gblk;
        wa.W = Wtb;
        wa.K = pl.K;
        wa.Kc = pl.Kc;
        wa.srow = pl.nsr ? srow : nullptr;
        wa.n = n;
        wa.st = (WalkState*)c->walk.p;
        wa.ops = (uint8_t*)c->ops.p;
        wa.xps = (XPoint*)c->xps.p;
        wa.xcap = xcap;
        HCK(c, launch_walk(wa, c->stream));
        const int64_t pi = st.i, pj = st.j;
        HCK(c, hipMemcpyAsync(&st, c->walk.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        if (!st.done && st.i == pi && st.j == pj) return SWH_ERR_INTERNAL;  // no progress
    }
    HCK(c, hipEventRecord(c->ev[5], c->stream));
    std::string ops((size_t)st.nops, '\0');
    if (st.nops) HCK(c, hipMemcpyAsync(&ops[0], c->ops.p, (size_t)st.nops, hipMemcpyDeviceToHost, c->stream));
    std::vector<XPoint> xps((size_t)st.nxp);
    if (want_xp && st.nxp)
        HCK(c, hipMemcpyAsync(xps.data(), c->xps.p, sizeof(XPoint) * st.nxp, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    out->ms_traceback = elapsed(c->ev[4], c->ev[5]);
    out->ms_total = out->ms_forward + out->ms_end_search + out->ms_traceback;
    out->start_row = st.start_i;
    out->start_col = st.start_j;
    std::reverse(ops.begin(), ops.end());
    std::string rle;
    rle.reserve(ops.size() / 4 + 16);
    char num[32];
    for (size_t a = 0; a < ops.size();) {
        size_t b = a;
        while (b < ops.size() && ops[b] == ops[a]) ++b;
        int k = snprintf(num, sizeof num, "%zu%c", b - a, ops[a]);
        rle.append(num, (size_t)k);
        a = b;
    }
    out->n_ops = (int64_t)ops.size();
    out->cigar = (char*)malloc(rle.size() + 1);
    if (!out->cigar) return SWH_ERR_OUT_OF_MEMORY;
    memcpy(out->cigar, rle.data(), rle.size());
    out->cigar[rle.size()] = 0;
    out->cigar_len = (int64_t)rle.size();
    if (want_xp) {
        const int64_t nx = std::min<int64_t>(st.nxp, xcap);
        out->xpoints = (swh_xpoint*)malloc(sizeof(swh_xpoint) * (nx + 1));
        if (!out->xpoints) return SWH_ERR_OUT_OF_MEMORY;
        for (int64_t k = 0; k < nx; ++k) {
            const XPoint& x = xps[(size_t)(nx - 1 - k)];
            out->xpoints[k] = swh_xpoint{x.row, x.col, x.score, x.layer};
        }
        out->xpoints[nx] = swh_xpoint{out->end_row, out->end_col, best, 0};
        out->n_xpoints = nx + 1;
    }
    return SWH_OK;
}

void result_init(swh_result* r) { memset(r, 0, sizeof *r); }

}  // namespace

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
    // Reserve LDS per sweep workgroup so that exactly w of them fit a CU (w = 1, 2): the
    // waves of a launch then spread evenly, w per SIMD (an uneven spread leaves some SIMDs
    // with more tiles than others, and a launch lasts as long as its busiest SIMD).
    const char* env = getenv("SWALIGN_SWEEP_WAVES_PER_SIMD");
    if (env && (atoi(env) == 1 || atoi(env) == 2)) c->wps_fixed = atoi(env);
    const size_t per_cu = (size_t)c->props.maxSharedMemoryPerMultiProcessor, stat = sweep_static_lds();
    for (int w = 1; w <= 2; ++w) {
        const size_t want = per_cu / (size_t)(w + 1) + 1024;
        if (want > stat && want <= (size_t)c->props.sharedMemPerBlock) c->lds_pad[w] = want - stat;
    }
    if (c->lds_pad[1] && sweep_allow_lds(c->lds_pad[1]) != hipSuccess) {
        (void)hipGetLastError();
        c->lds_pad[1] = c->lds_pad[2] = 0;
    }
    *out = c;
    return SWH_OK;
}

int swh_ctx_destroy(swh_ctx* c) {
    if (!c) return SWH_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    DevBuf* bufs[] = {&c->seqA, &c->seqB, &c->rbuf, &c->cb, &c->cor, &c->rowmax, &c->srow, &c->scol,
                      &c->dirs, &c->regions, &c->key, &c->found, &c->walk, &c->ops, &c->xps, &c->wregs};
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
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
    const bool swap = len1 < len2;  // src/SWutils.cpp:84-88
    out->swapped = swap ? 1 : 0;
    const uint8_t* A = swap ? d2 : d1;
    const uint8_t* B = swap ? d1 : d2;
    const int64_t m = swap ? len2 : len1, n = swap ? len1 : len2;
    return align_core(c, A, m, B, n, sc, o, out);
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
    swh_opts o{};
    o.checkpoint_rows = K;
    o.checkpoint_cols = Kc;
    o.hbm_budget_bytes = INT64_MAX / 4;
    swh_result r;
    result_init(&r);
    const int rc = align_core(c, (const uint8_t*)c->seqA.p, m, (const uint8_t*)c->seqB.p, n, sc, &o, &r,
                              /*traceback=*/false);
    if (rc) return rc;
    *score = r.score;
    *end_row = r.end_row;
    *end_col = r.end_col;
    if (rowmax && m) HCK(c, hipMemcpy(rowmax, c->rowmax.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
    if (srow_out && m / K > 0)
        HCK(c, hipMemcpy(srow_out, c->srow.p, sizeof(int2) * (m / K) * n, hipMemcpyDeviceToHost));
    if (scol_out && n / Kc > 0)
        HCK(c, hipMemcpy(scol_out, c->scol.p, sizeof(int2) * (n / Kc) * m, hipMemcpyDeviceToHost));
    swh_result_free(&r);
    return SWH_OK;
}

int swh_batch_align(swh_ctx* c, const swh_pair* pairs, int64_t npairs, const swh_scoring* sc,
                    const swh_opts* o, swh_result* out, int32_t* statuses) {
    if (!c || (npairs > 0 && (!pairs || !out)) || npairs < 0) return SWH_ERR_INVALID_ARG;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    for (int64_t k = 0; k < npairs; ++k) {
        const int rc = swh_align_pair(c, pairs[k].seq1, pairs[k].len1, pairs[k].seq2, pairs[k].len2, sc, o, &out[k]);
        if (statuses) statuses[k] = rc;
    }
    return SWH_OK;
}

void swh_free(void* p) { free(p); }

}  // extern "C"
