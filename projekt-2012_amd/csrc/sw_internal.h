// Internal types shared by the HIP kernels (sw_kernels.hip) and the host side of the
// C-ABI (sw_api.hip).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swk {

// Rows per lane of the forward sweep; a wave owns a stripe of 64*kRowsPerLane rows.
constexpr int kRowsPerLane = 4;
constexpr int kStripeRows = 64 * kRowsPerLane;
// Waves per sweep workgroup, one tile each.  Sweeps reserve enough LDS that one
// workgroup fits a CU, so the waves of a launch spread one per SIMD.
constexpr int kSweepWaves = 4;

// Sweep modes.
enum SweepMode : int {
    kStage1 = 0,   // whole matrix: row maxima, checkpoint rows and columns
    kDirs = 1,     // block recomputation writing 4-bit direction codes per cell
    kSearch = 2    // block recomputation looking for the first column holding a value
};

struct Score {
    int32_t M, X, O, E;  // match, mismatch, gap open ("first"), gap extension
};

// One independent rectangle of one pair's DP matrix, rows [r0, r1] x columns [c0, c1]
// (absolute within the pair, inclusive).  It is cut into tiles of kStripeRows rows
// (stripe s) by W columns (segment g); tile (s, g) needs the bottom row of (s-1, g), the
// right column of (s, g-1) and the bottom-right cell of (s-1, g-1).  All tiles with
// s + g = d are independent and form one launch, together with the same diagonal of every
// other region of the sweep (other blocks, other pairs); launches run in order of d, so
// every hand-off between tiles crosses a kernel boundary and no wave waits on another.
struct Region {
    const uint8_t* A;  // the pair's rows (m bytes)
    const uint8_t* B;  // the pair's columns (n bytes)
    int64_t m, n;
    int64_t r0, r1, c0, c1;
    int64_t W;         // columns per tile
    int64_t G;         // kDirs: segments per stripe in the code layout (>= tiles across the region)
    int64_t tstride;   // kDirs: steps stored per tile
    const int2* top;   // {H,F} of row r0-1 at columns c0..c1 (index col-c0); nullptr = zero row
    const int2* left;  // {H,E} of column c0-1 indexed by row; nullptr = zero column
    int2* rb;          // 2 x (c1-c0+1) {H,F}: bottom row of stripe s, kept in slot (s+1)&1
    int2* cb;          // 2 x (r1-r0+1) {H,E}: right column of segment g, kept in slot (g+1)&1
    int32_t* cor;      // 2 x stripes: H(R0-1, C0-1) of the next segment, kept in slot (g+1)&1
    // kStage1 outputs (nullable checkpoints)
    int32_t* rowmax;   // m row maxima
    int32_t* rowseg;   // nullable: m first segments (tile columns) holding each row maximum
    int2* srow;        // (m/K) x n {H,F}: rows K-1, 2K-1, ...
    int64_t K;         // multiple of kStripeRows
    int2* scol;        // (n/Kc) x m {H,E}: columns Kc-1, 2Kc-1, ...
    int64_t Kc;        // power of two >= 64
    // kDirs
    uint16_t* dirs;    // [tile s*G+g][step][lane] packed direction codes
    // kSearch
    int64_t target_row;
    int32_t target;
    int32_t skew;      // kStage1: skewed segments (TileCtx in sw_kernels.hip), W a multiple of 64, cb
                       // holding lane slots instead of a column: 1 = tiles (s, g) at diagonal d = 2s + g,
                       // 2 = shifted stripes (stripe s's segment grid 64s columns left), d = s + g
    int32_t* found;    // smallest column of target_row holding the target value
};

// The tiles of one region on one of its anti-diagonals d: stripes s_lo, s_lo+1, ... are
// tiles first_tile, first_tile+1, ... of the launch.  Regions of one launch may be at
// different diagonals (pairs admitted at different times).
struct TileSpan {
    int64_t first_tile;
    int32_t region;
    int32_t s_lo;
    int64_t d;
};

struct SweepParams {
    Score sc;
    const Region* regions;
    const TileSpan* spans;  // this launch's spans, ordered by first_tile
    int32_t nspans;
    int32_t code_shift1;    // 0: compare bytes; s + 1: symbol code (byte >> s) & 7 selects a row profile byte
    int64_t ntiles;         // tiles of this launch (grid = ceil(ntiles / kSweepWaves) workgroups)
    int32_t code2_shift1;   // packed stage 1: 0 = compare bytes; s + 1: 2-bit codes (byte >> s) & 3
    int32_t _pad2;
};

struct WalkState {
    int64_t i, j;              // current cell (absolute)
    int64_t start_i, start_j;  // first aligned pair seen so far
    int64_t nops;              // ops written (reverse order)
    int64_t nxp;               // crosspoints written
    int64_t last_row;          // last row visited (crosspoint detection)
    int32_t layer;             // 0 = H, 1 = E, 2 = F
    int32_t done;              // 1 when the start of the alignment was reached
    int32_t v;                 // value of the current layer's matrix at (i, j) (the score at the end cell)
    int32_t _pad;
};

struct XPoint {
    int64_t row, col;
    int32_t score;
    int32_t layer;
};

// A recomputed region of direction codes: rows [r0, r0+rows), cols [c0, c0+cols).
struct WalkRegion {
    int64_t r0, c0, rows, cols;
    const uint16_t* dirs;        // Region::dirs layout
};

// One pair's walk in a traceback round: it visits the pair's regions of the round.  Region
// a*(L+1)+b covers (part of) the checkpoint block (brow-a, bcol-b); rows == 0 marks a
// block not recomputed.
struct WalkArgs {
    const uint8_t* A;
    const uint8_t* B;
    const WalkRegion* regs;
    int64_t L, brow, bcol;       // lookahead and the block of the round's first region
    int64_t tstride, G, W;       // steps per tile, segments per stripe, columns per segment
    int64_t K, Kc;               // checkpoint row / column spacing
    const int2* srow;            // checkpoint rows (crosspoint values), may be null
    int64_t n;
    WalkState* st;
    uint8_t* ops;
    XPoint* xps;
    int64_t xcap;
    Score sc;                    // the walk follows the values of the cells it visits
};

// Launchers (sw_kernels.hip).
hipError_t launch_sweep(int mode, const SweepParams& p, size_t lds_pad, hipStream_t s);
// Stage 1 in packed 16-bit arithmetic, two tiles per wave (the caller guarantees no value
// leaves int16).
hipError_t launch_sweep_pk(int mode, const SweepParams& p, size_t lds_pad, hipStream_t s);  // kStage1 or kDirs
size_t sweep_static_lds();                     // static LDS of a sweep workgroup
hipError_t sweep_allow_lds(size_t dyn_bytes);  // permits dyn_bytes of dynamic LDS
// Row argmax of many pairs: job k reduces rowmax[off[k] .. off[k] + len[k]) into key[k]
// (score << 32 | ~row; key must be zeroed).
struct ArgmaxJob {
    const int32_t* rowmax;
    int64_t m;
    unsigned long long* key;
    const int32_t* rowseg;  // nullable: row -> first segment holding the row maximum
    int32_t* seg;           // receives rowseg[end row] (with rowseg)
};
hipError_t launch_argmax_rows(const ArgmaxJob* jobs, int njobs, int blocks_per_job, hipStream_t s);
// After launch_argmax_rows on the same stream: seg = rowseg[end row] for jobs with rowseg.
hipError_t launch_end_segment(const ArgmaxJob* jobs, int njobs, hipStream_t s);
// A byte range in device memory.
struct ByteRange {
    const uint8_t* p;
    int64_t n;
};
// ORs into mask[8] (zeroed by the caller) bit x for every byte x present in the ranges.
hipError_t launch_alphabet(const ByteRange* ranges, int64_t nranges, uint32_t* mask, hipStream_t s);

// Run-length CIGAR text of many alignments, one wave per alignment: ops[0 .. nops) hold the
// alignment's '=', 'X', 'I', 'D' ops end -> start, the text reads start -> end.  The count
// pass (out == nullptr) writes each job's text length to len; the write pass writes the
// text (no terminator) at out + at.
struct CigarJob {
    const uint8_t* ops;
    int64_t nops;
    int64_t at;   // write pass: offset of the job's text in out
    int64_t len;  // count pass: characters of the job's text
};
hipError_t launch_cigar(CigarJob* jobs, int64_t njobs, char* out, hipStream_t s);

// One workgroup per job.  many: more walks than the device holds at once with the large
// window (walk_lds_bytes() of LDS each), which then run with small windows.
hipError_t launch_walk(const WalkArgs* jobs, int njobs, bool many, hipStream_t s);
size_t walk_lds_bytes();

// ---- traceback rounds planned on the device ------------------------------------------
// While many pairs walk, each round's regions, walk arguments and the next round's list of
// walking pairs are built by kernels from the walk states in device memory, so that rounds
// follow each other without the host.  Round slot x (the x-th walking pair) owns fixed
// slices of the hand-off buffers and of the direction codes.

// What the round kernels need of pair k; constant over the traceback.
struct TbPair {
    const uint8_t* A;
    const uint8_t* B;
    int64_t m, n, K, Kc;
    const int2* srow;  // the pair's checkpoint rows, or nullptr (none)
    const int2* scol;  // the pair's checkpoint columns, or nullptr (none)
    uint8_t* ops;
    XPoint* xps;
    int64_t xcap;
};

// Position of slot x's pair when its round was planned (progress check).
struct TbPre {
    int64_t i, j;
    int32_t layer, _pad;
};

struct TbRoundArgs {
    const TbPair* pairs;
    WalkState* states;        // per pair; done = -1 marks a pair whose walk stopped moving
    const int32_t* act_in;    // walking pairs of this round, *nact_in of them
    const int32_t* nact_in;
    int32_t* act_out;         // the next round's, built by the compaction
    int32_t* nact_out;        // zeroed by the plan kernel
    Region* regions;          // [slot]
    WalkRegion* wregs;        // [slot]
    WalkArgs* wargs;          // [slot]; st == nullptr for an empty slot
    TbPre* pre;               // [slot]
    int64_t nslots;           // slots of the launch (at least *nact_in)
    int64_t W, G, tstride;    // tile columns, segments per stripe of the code layout, steps per tile
    int2* cb;                 // + slot * cb_stride
    int64_t cb_stride;
    int2* rb;                 // + slot * rb_stride
    int64_t rb_stride;
    int32_t* cor;             // + slot * cor_stride
    int64_t cor_stride;
    uint16_t* dirs;           // + slot * dirs_stride
    int64_t dirs_stride;
    Score sc;
};
// Regions, walk arguments and progress records of one round (block of each walking pair's
// current cell, cut at the cell); slots past *nact_in get empty regions.
hipError_t launch_tb_plan(const TbRoundArgs& a, hipStream_t s);
// After the round's walk: the pairs still walking go to act_out (any order), and a pair
// whose walk did not move is marked done = -1 and dropped.
hipError_t launch_tb_compact(const TbRoundArgs& a, hipStream_t s);

}  // namespace swk
