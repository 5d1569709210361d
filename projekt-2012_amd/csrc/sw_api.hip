// Host side of the C-ABI (include/swalign_hip.h): context, device buffers and the stage
// driver.  The stage structure replaces Sterbic/Projekt-2012's host loop
// (src/main.cu:158-248) and its Crosspointer (src/Crosspointer.cpp:104-338):
//   1. forward sweep over the whole matrix, keeping checkpoint rows {H,F} every K rows
//      and checkpoint columns {H,E} every Kc columns in HBM, plus per-row maxima;
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
    DevBuf seqA, seqB, bnd, prog, rowmax, srow, scol, dirs, tiles, key, found, walk, ops, xps, ticket, err;
    int occ_stage1 = 0, occ_dirs = 0, occ_search = 0;
    hipEvent_t ev[8] = {};
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
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.cap = bytes;
    else (void)hipGetLastError();
    return e;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int64_t next_pow2(int64_t v) {
    int64_t p = 64;
    while (p < v) p <<= 1;
    return p;
}

bool scoring_ok(const swh_scoring* s) {
    // src/SWutils.cpp:36: match >= 1, every penalty <= -1
    return s && s->match >= 1 && s->mismatch <= -1 && s->gap_open <= -1 && s->gap_extend <= -1;
}

int grid_for(swh_ctx* c, int occ, int64_t stripes) {
    int64_t cap = (int64_t)std::max(occ, 1) * c->props.multiProcessorCount;
    return (int)std::max<int64_t>(1, std::min<int64_t>(stripes, cap));
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Runs the sweep over `tiles` (already in host memory).  Optional events are recorded
// immediately before and after the kernel launch.
int run_sweep(swh_ctx* c, int mode, SweepParams p, const std::vector<Tile>& tiles, hipEvent_t ev_before = nullptr,
              hipEvent_t ev_after = nullptr) {
    HCK(c, ensure(c->tiles, sizeof(Tile) * tiles.size()));
    HCK(c, hipMemcpyAsync(c->tiles.p, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice,
                          c->stream));
    HCK(c, ensure(c->ticket, 64));
    HCK(c, hipMemsetAsync(c->ticket.p, 0, 8, c->stream));
    HCK(c, ensure(c->err, 64));
    HCK(c, hipMemsetAsync(c->err.p, 0, 32, c->stream));
    int64_t total = 0;
    for (auto& t : tiles) total += t.nstripes;
    p.tiles = (const Tile*)c->tiles.p;
    p.ntiles = (int32_t)tiles.size();
    p.total_stripes = total;
    p.ticket = (unsigned long long*)c->ticket.p;
    p.err = (int32_t*)c->err.p;
    if (p.spin_limit == 0) p.spin_limit = kDefaultSpinLimit;
    const int occ = mode == kStage1 ? c->occ_stage1 : (mode == kDirs ? c->occ_dirs : c->occ_search);
    if (ev_before) HCK(c, hipEventRecord(ev_before, c->stream));
    HCK(c, launch_sweep(mode, p, grid_for(c, occ, total), c->stream));
    if (ev_after) HCK(c, hipEventRecord(ev_after, c->stream));
    return SWH_OK;
}

// Reads the sweep's failure record (after the stream has been synchronised).
int check_sweep(swh_ctx* c) {
    int32_t e[5] = {0, 0, 0, 0, 0};
    HCK(c, hipMemcpy(e, c->err.p, sizeof e, hipMemcpyDeviceToHost));
    if (e[0] != 0) {
        fprintf(stderr, "swalign: stripe wait exceeded its bound (tile %d, stripe %d, need %d, seen %d)\n", e[1],
                e[2], e[3], e[4]);
        return SWH_ERR_INTERNAL;
    }
    return SWH_OK;
}

struct Plan {
    int64_t K, Kc, nsr, nsc;
};

int choose_plan(swh_ctx* c, int64_t m, int64_t n, const swh_opts* o, Plan* pl) {
    int64_t K = o && o->checkpoint_rows > 0 ? ceil_div(o->checkpoint_rows, kStripeRows) * kStripeRows : 0;
    int64_t Kc = o && o->checkpoint_cols > 0 ? next_pow2(o->checkpoint_cols) : 0;
    size_t freeb = 0, totalb = 0;
    hipMemGetInfo(&freeb, &totalb);
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

    const int64_t nstripes = ceil_div(m, kStripeRows);
    HCK(c, ensure(c->bnd, sizeof(int2) * n));
    HCK(c, ensure(c->prog, sizeof(int32_t) * std::max<int64_t>(nstripes, 1) + 64));
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
    p.spin_limit = (o && o->wait_spin_limit > 0) ? (uint32_t)o->wait_spin_limit : kDefaultSpinLimit;

    // ---- stage 1 ------------------------------------------------------------------
    HCK(c, hipMemsetAsync(c->prog.p, 0, sizeof(int32_t) * nstripes, c->stream));
    {
        Tile t{};
        t.r0 = 0;
        t.r1 = m - 1;
        t.c0 = 0;
        t.c1 = n - 1;
        t.stripe0 = 0;
        t.nstripes = nstripes;
        t.bnd = (int2*)c->bnd.p;
        t.prog = (int32_t*)c->prog.p;
        std::vector<Tile> tiles{t};
        rc = run_sweep(c, kStage1, p, tiles, c->ev[0], c->ev[1]);
        if (rc) return rc;
    }
    HCK(c, hipMemsetAsync(c->key.p, 0, 8, c->stream));
    HCK(c, launch_argmax_rows((const int32_t*)c->rowmax.p, m, (unsigned long long*)c->key.p, c->stream));
    unsigned long long key = 0;
    HCK(c, hipMemcpyAsync(&key, c->key.p, 8, hipMemcpyDeviceToHost, c->stream));
    HCK(c, hipStreamSynchronize(c->stream));
    if ((rc = check_sweep(c))) return rc;
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
    const int64_t band_rows = bi - br0 + 1;
    const int64_t sp = ceil_div(band_rows, kStripeRows);
    const int64_t nq = ceil_div(n, pl.Kc);
    HCK(c, ensure(c->prog, sizeof(int32_t) * std::max<int64_t>(nstripes, sp * nq) + 64));
    {
        std::vector<Tile> tiles;
        tiles.reserve(nq);
        for (int64_t q = 0; q < nq; ++q) {
            Tile t{};
            t.r0 = br0;
            t.r1 = bi;
            t.c0 = q * pl.Kc;
            t.c1 = std::min(n, t.c0 + pl.Kc) - 1;
            t.stripe0 = q * sp;
            t.nstripes = sp;
            t.bnd = (int2*)c->bnd.p + t.c0;
            t.prog = (int32_t*)c->prog.p + q * sp;
            t.top = br0 > 0 ? srow + (br0 / pl.K - 1) * n + t.c0 : nullptr;
            t.left = t.c0 > 0 ? scol + (t.c0 / pl.Kc - 1) * m : nullptr;
            t.found = (int32_t*)c->found.p;
            tiles.push_back(t);
        }
        const int32_t big = INT32_MAX;
        HCK(c, hipMemcpyAsync(c->found.p, &big, 4, hipMemcpyHostToDevice, c->stream));
        HCK(c, hipMemsetAsync(c->prog.p, 0, sizeof(int32_t) * sp * nq, c->stream));
        SweepParams ps = p;
        ps.target_row = bi;
        ps.target = best;
        rc = run_sweep(c, kSearch, ps, tiles);
        if (rc) return rc;
        int32_t bj = -1;
        HCK(c, hipMemcpyAsync(&bj, c->found.p, 4, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipEventRecord(c->ev[3], c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        if ((rc = check_sweep(c))) return rc;
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
    int64_t guard = 0;
    while (!st.done) {
        if (++guard > (m / std::max<int64_t>(pl.K, 1) + n / std::max<int64_t>(pl.Kc, 1) + 8) * 4)
            return SWH_ERR_INTERNAL;
        const int64_t r0 = (st.i / pl.K) * pl.K, c0 = (st.j / pl.Kc) * pl.Kc;
        const int64_t rows = st.i - r0 + 1, cols = st.j - c0 + 1;
        const int64_t ns = ceil_div(rows, kStripeRows);
        const int64_t tstride = ceil_div(cols + 63, 64) * 64;
        HCK(c, ensure(c->dirs, sizeof(uint16_t) * (size_t)(ns * tstride * 64)));
        HCK(c, ensure(c->prog, sizeof(int32_t) * std::max<int64_t>({nstripes, sp * nq, ns}) + 64));
        HCK(c, hipMemsetAsync(c->prog.p, 0, sizeof(int32_t) * ns, c->stream));
        Tile t{};
        t.r0 = r0;
        t.r1 = st.i;
        t.c0 = c0;
        t.c1 = st.j;
        t.stripe0 = 0;
        t.nstripes = ns;
        t.bnd = (int2*)c->bnd.p;
        t.prog = (int32_t*)c->prog.p;
        t.top = r0 > 0 ? srow + (r0 / pl.K - 1) * n + c0 : nullptr;
        t.left = c0 > 0 ? scol + (c0 / pl.Kc - 1) * m : nullptr;
        t.dirs = (uint16_t*)c->dirs.p;
        t.tstride = tstride;
        std::vector<Tile> tiles{t};
        rc = run_sweep(c, kDirs, p, tiles);
        if (rc) return rc;
        WalkArgs wa{};
        wa.A = dA;
        wa.B = dB;
        wa.dirs = (const uint16_t*)c->dirs.p;
        wa.tstride = tstride;
        wa.r0 = r0;
        wa.c0 = c0;
        wa.rows = rows;
        wa.cols = cols;
        wa.K = pl.K;
        wa.srow = pl.nsr ? srow : nullptr;
        wa.n = n;
        wa.st = (WalkState*)c->walk.p;
        wa.ops = (uint8_t*)c->ops.p;
        wa.xps = (XPoint*)c->xps.p;
        wa.xcap = xcap;
        HCK(c, launch_walk(wa, c->stream));
        HCK(c, hipMemcpyAsync(&st, c->walk.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
        HCK(c, hipStreamSynchronize(c->stream));
        if ((rc = check_sweep(c))) return rc;
        if (!st.done && (st.i >= r0 && st.j >= c0)) return SWH_ERR_INTERNAL;  // no progress
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
    c->occ_stage1 = sweep_occupancy(kStage1);
    c->occ_dirs = sweep_occupancy(kDirs);
    c->occ_search = sweep_occupancy(kSearch);
    *out = c;
    return SWH_OK;
}

int swh_ctx_destroy(swh_ctx* c) {
    if (!c) return SWH_ERR_INVALID_ARG;
    hipSetDevice(c->device);
    DevBuf* bufs[] = {&c->seqA, &c->seqB, &c->bnd, &c->prog, &c->rowmax, &c->srow, &c->scol, &c->dirs,
                      &c->tiles, &c->key, &c->found, &c->walk, &c->ops, &c->xps, &c->ticket, &c->err};
    if (c->stream) hipStreamSynchronize(c->stream);
    for (DevBuf* b : bufs)
        if (b->p) hipFree(b->p);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
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
