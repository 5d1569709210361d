// Host-side internals shared by the C-ABI (sw_api.hip) and the request server
// (sw_serve.hip): the context, device buffers, checkpoint plans and the sweep driver.
// Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/swalign_hip.h"
#include "sw_internal.h"

// A grow-only device allocation, or (fixed) a slice of memory the context does not own:
// ensure() then only checks the capacity and never allocates.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool fixed = false;
};

// A grow-only pinned host allocation mapped into the device's address space.
struct HostBuf {
    void* h = nullptr;  // host address
    void* d = nullptr;  // device address of the same memory
    size_t cap = 0;
};

// Sweep waves resident per SIMD that an LDS reservation can enforce (the packed sweep's
// register use allows four).
constexpr int kMaxSweepWps = 4;

struct swh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    DevBuf seqA, seqB, rbuf, cb, cor, rowmax, srow, scol, dirs, regions, key, found, walk, ops, xps, wregs,
        argjobs, wargs, seqs, spans, alpha, ranges, rowseg, segs;
    DevBuf tb_pairs, tb_states, tb_act, tb_pre;  // traceback rounds planned on the device
    DevBuf cigjobs, cigtext;                      // CIGAR text built on the device
    uint8_t* staging = nullptr;  // pinned host staging of batch inputs
    size_t staging_cap = 0;
    uint8_t* host_ops = nullptr;  // pinned host copy of the traceback ops
    size_t host_ops_cap = 0;
    HostBuf hsrow, hscol;         // host-resident checkpoint rows / columns (swh_opts.host_checkpoints)
    hipEvent_t ev[8] = {};
    size_t lds_pad[kMaxSweepWps + 1] = {};  // dynamic LDS per sweep workgroup for w workgroups per CU
    int wps_fixed = 0;              // sweep waves per SIMD forced by the environment, 0 = model
    bool codes_off = false;         // symbol-code scores disabled by the environment
    bool trace = false;             // SWALIGN_TRACE=1: host wall time of each stage to stderr
    // tiles a packed sweep's launch needs to run packed (packed_launch); -1 = 1.5 per SIMD.
    // SWALIGN_PACKED_MIN_TILES=n, read at the start of each alignment call, sets it (tests use
    // 0 to run every launch of a packed sweep packed)
    int64_t packed_min_tiles = -1;
    hipError_t last = hipSuccess;
    // The allocation-free entry point runs on a copy of the context whose DevBufs are fixed
    // slices of the device's static scratch array; these then replace the allocations of
    // the traceback's host buffers (nullptr: the grow-only pinned buffers above).
    void* scratch = nullptr;        // this device's address of the scratch array
    int scratch_status = 0;         // SWH_OK, or why it could not be resolved
    uint8_t* ops_dst = nullptr;     // host destination of the traceback ops
    size_t ops_dst_cap = 0;
    char* cigar_dst = nullptr;      // caller's CIGAR buffer (the result's cigar stays NULL)
    int64_t cigar_dst_cap = 0;
    int64_t dirs_cap_fixed = 0;     // direction-code bytes per traceback round, 0 = from free HBM
    bool fixed_memory = false;      // every device buffer is a fixed slice: nothing may be allocated
    std::vector<swk::TileSpan> host_spans;  // launch spans of the latest sweep, fixed-memory contexts
    swk::TileSpan* pinned_spans = nullptr;  // launch spans of the latest sweep (pinned, grow-only)
    size_t pinned_spans_cap = 0;
};

namespace swh_host {

int map_hip(swh_ctx* c, hipError_t e);

#define HCK(ctx, x)                                      \
    do {                                                 \
        hipError_t e__ = (x);                            \
        if (e__ != hipSuccess) return swh_host::map_hip(ctx, e__); \
    } while (0)

hipError_t ensure(DevBuf& b, size_t bytes);
hipError_t ensure_host(HostBuf& b, size_t bytes);
int64_t ceil_div(int64_t a, int64_t b);
int64_t next_pow2(int64_t v);
bool scoring_ok(const swh_scoring* s);
float elapsed(hipEvent_t a, hipEvent_t b);
void result_init(swh_result* r);

// Columns per tile: `want` if given, else one tile per resident sweep wave.
int64_t tile_cols(swh_ctx* c, int64_t cols, int64_t want, int wps = 1);
int choose_wps(swh_ctx* c, int64_t rows, int64_t cols, bool skew);
// w if the context can hold the sweep to w waves per SIMD (1 <= w <= kMaxSweepWps), else 1.
int usable_wps(const swh_ctx* c, int w);
// True if a packed sweep's launch of ntiles tiles should run the packed kernel (two waves
// per SIMD or more), false if the int32 kernel (one tile per wave) is faster for it.
bool packed_launch(const swh_ctx* c, int64_t ntiles);
// True if njobs walks exceed what the device's LDS holds at once with large walk windows.
bool many_walks(const swh_ctx* c, int64_t njobs);

// One sweep over `regions`, one launch per tile anti-diagonal (see sw_api.hip).  skew: 0
// rectangular tiles, 1 skewed stripes, 2 shifted stripes (stage 1; see Region::skew).
int run_sweep(swh_ctx* c, int mode, swk::SweepParams p, std::vector<swk::Region>& regions, int64_t W,
              int64_t G_fix = 0, hipEvent_t ev_before = nullptr, hipEvent_t ev_after = nullptr, int wps = 1,
              int64_t* launches = nullptr, int64_t tstride = 0, bool packed = false,
              bool sync = true, int skew = 0);

// Shifted stripes (W a multiple of 64, at least two segments): stripe s's segments g run from
// floor(64s / W), the first whose lane-0 columns [gW - 64s, gW - 64s + W) reach column 0, to
// floor((cols - 1 + 64s) / W), the one holding column cols - 1 for lane 0; tile (s, g) lies on
// diagonal s + g.
inline int64_t shifted_first_diag(int64_t s, int64_t W) { return s + 64 * s / W; }
inline int64_t shifted_last_diag(int64_t s, int64_t cols, int64_t W) { return s + (cols - 1 + 64 * s) / W; }
// Stripes [lo, hi] of a region of S shifted stripes with a tile on diagonal d (none if lo > hi).
inline void shifted_stripes(int64_t d, int64_t S, int64_t cols, int64_t W, int64_t& lo, int64_t& hi) {
    const int64_t k = W / 64;  // the first diagonal grows by k + 1 every k stripes
    int64_t h = std::min(S - 1, d * k / (k + 1));
    while (h + 1 <= S - 1 && shifted_first_diag(h + 1, W) <= d) ++h;
    while (h >= 0 && shifted_first_diag(h, W) > d) --h;
    int64_t l = std::max<int64_t>(0, std::min(h, (d * W - (cols - 1)) * k / ((k + 1) * W)));
    while (l > 0 && shifted_last_diag(l - 1, cols, W) >= d) --l;
    while (l <= h && shifted_last_diag(l, cols, W) < d) ++l;
    lo = l;
    hi = h;
}

// Widest stage-1 tile (columns) whose values stay in int16 relative to the tile's base under
// scoring sc, for the packed stage 1; 0 if none.
int64_t int16_tile_cols(const swh_scoring* sc);

// Symbol codes for the sweeps' scores (SweepParams::code_shift1): the smallest shift s in
// [0, 5] under which (byte >> s) & 7 is one-to-one on the bytes present (mask: bit x of
// word x >> 5), plus one; 0 if there is none or match - mismatch exceeds a byte.
int code_shift1_for(const uint32_t mask[8], const swh_scoring* sc);
// Same, for the bytes of device ranges (one presence kernel and a read-back).
int code_shift1_device(swh_ctx* c, const std::vector<swk::ByteRange>& ranges, const swh_scoring* sc,
                       int* shift1, int* shift2 = nullptr);
// Packed stage 1 (SweepParams::code2_shift1): the smallest shift s in [0, 6] under which
// (byte >> s) & 3 is one-to-one on the bytes present, plus one; 0 if there is none or a
// profile byte (match - gap_open, mismatch - gap_open) leaves [0, 255].
int code2_shift1_for(const uint32_t mask[8], const swh_scoring* sc);

// Traceback lookahead: block (da, db) above / left of the current cell's checkpoint block is
// recomputed in a round with lookahead L if da, db <= L and, beyond L = 3, |da - db| <= 2 (a
// band along the diagonal, where paths run; a path that leaves it is picked up next round).
inline bool lookahead_block(int64_t L, int64_t da, int64_t db) {
    return da <= L && db <= L && (L <= 3 || (da > db ? da - db : db - da) <= 2);
}
// Blocks per round with lookahead L.
inline int64_t lookahead_blocks(int64_t L) {
    int64_t n = 0;
    for (int64_t da = 0; da <= L; ++da)
        for (int64_t db = 0; db <= L; ++db) n += lookahead_block(L, da, db);
    return n;
}

struct Plan {
    int64_t K, Kc, nsr, nsc;
};
int64_t plan_mem(int64_t m, int64_t n, int64_t K, int64_t Kc);
int choose_plan(int64_t m, int64_t n, const swh_opts* o, int64_t budget, Plan* pl);

// Fills r's start, CIGAR and crosspoints.  The CIGAR is the text built on the device (text,
// text_len characters) when text is given, else the run-length encoding of ops (written
// end -> start).  With dst the encoded CIGAR is written there (dst_cap bytes, NUL included)
// and r.cigar stays NULL.
int fill_alignment(swh_result& r, const swk::WalkState& st, const uint8_t* ops, const char* text, int64_t text_len,
                   const swk::XPoint* xps, int64_t xcap, int32_t best, bool want_xp, char* dst = nullptr,
                   int64_t dst_cap = 0);

// The device's static scratch array of the allocation-free entry point (sw_noalloc.hip):
// resolves its address on `device` (which must be current) once per device and process.
int scratch_resolve(int device, void** out);

// The stage driver (sw_api.hip), shared with the allocation-free entry point.
// One pair of an alignment batch (a single alignment is a batch of one).
struct Job {
    const uint8_t* A = nullptr;  // device, oriented: A = rows (the longer), B = columns
    const uint8_t* B = nullptr;
    int64_t m = 0, n = 0;
    bool swapped = false;
    bool traceback = true;
    Plan pl{};
    int64_t rowmax_off = 0, srow_off = 0, scol_off = 0, ops_off = 0, xps_off = 0, xcap = 0;
    int32_t best = 0;
    int64_t bi = -1, bj = -1;
    swk::WalkState st{};
    int status = SWH_OK;
};

struct EngineOpts {
    int64_t tile_cols = 0;   // 0 = auto
    int precision = 0;       // 0 = packed 16-bit stage 1 where it cannot overflow, 32 = int32 only
    int wps = 0;             // sweep waves per SIMD, 0 = auto
    int lookahead = 7;       // traceback blocks ahead per direction (0 = none)
    bool lookahead_auto = false;  // batches: lookahead chosen per round from the walking pairs
    bool want_xp = false;
    bool batch = false;      // many pairs: fixed tile width, two waves per SIMD
    bool host_ckpt = false;  // checkpoints in mapped pinned host memory
    bool skew = false;       // skewed stage-1 stripes (see run_sweep)
    bool skew_forced = false;  // single pairs: skewed even where the model prefers rectangles
    int layout = 0;          // skewed stripes: 0 default, 1 skewed, 2 shifted (Region::skew)
};

// Stage 1, end search and traceback of every job, all jobs side by side in each sweep.
int run_jobs(swh_ctx* c, std::vector<Job>& jobs, const swh_scoring* sc, const EngineOpts& eo, swh_result* out);
EngineOpts engine_opts(const swh_opts* o, bool batch);
// A job for device-resident sequences, oriented as the reference does (longer = rows).
Job make_job(const uint8_t* d1, int64_t len1, const uint8_t* d2, int64_t len2, bool traceback);

}  // namespace swh_host
