// HIP kernels (gfx950) for the SWalign hot path.
//
// Forward sweep.  The DP matrix (rows = A, columns = B) is cut into tiles of
// kStripeRows = 64*R rows by W columns; one 64-lane wavefront computes one tile, left
// to right.  Lane l holds rows [l*R, l*R+R) of the tile and, at step t, updates column
// t-l: the tile advances along anti-diagonals, so lane l needs the bottom-row {H,F} of
// lane l-1 from the previous step, which it receives with one DPP wave_shr:1 per value.
// Lane 0 takes the tile's top boundary (the bottom row of the tile above) from register
// queues filled by coalesced loads once per 64 steps; lane 63's bottom row is staged in
// LDS and written once per 64 steps with a coalesced store.  Tiles are scheduled by
// anti-diagonal of the tile grid, one kernel launch per diagonal, so a tile's inputs were
// all written by earlier launches: no wave waits on another wave.
//
// Cell update (Sterbic/Projekt-2012 src/FindAlignment.cuh:50-56, zero boundaries as in
// src/SWutils.cpp:131-136), kept in "plus open" form Ho = H + O so that one add serves
// both gap recurrences:
//     E  = max(E + ext, Ho_left)               F = max(F_up + ext, Ho_up)
//     H  = max(0, Hdiag + s, E, F)             Ho = H + O
// with Hdiag + s computed as Ho_diag + (s - O).
//
// Substitution score.  In general s = (a == b) ? match : mismatch, one compare and one
// select per cell.  When the pair's bytes map one-to-one onto 3-bit symbol codes
// ((byte >> shift) & 7, chosen by the host from the bytes present), each row instead holds
// an 8-byte profile whose byte c is (match - mismatch) if c is the row's code, else 0, and
// each column carries a v_perm_b32 selector for its code: one v_perm_b32 yields the bonus
// and Hdiag + s = Ho_diag + (mismatch - O) + bonus is one three-operand add.
#include "sw_internal.h"

namespace swk {

namespace {

// Steps unrolled per loop iteration: the int32 sweeps (one wave per SIMD for single pairs)
// gain from the longer blocks the scheduler sees; the packed sweep's larger body runs best
// at 8 (its 16-step loop no longer fits the instruction cache well).
#ifndef UNROLL_STEPS
#define UNROLL_STEPS 16
#endif
#ifndef UNROLL_STEPS_PK
#define UNROLL_STEPS_PK 8
#endif
#define SWK_DO_PRAGMA(x) _Pragma(#x)
#define SWK_UNROLL(n) SWK_DO_PRAGMA(unroll n)

constexpr int kDppWaveShl1 = 0x130;  // lane l <- lane l+1
constexpr int kDppWaveShr1 = 0x138;  // lane l <- lane l-1, lane 0 keeps `old`

__device__ __forceinline__ int shr1(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShr1, 0xF, 0xF, false);
}
// Queue shift: lane 63 receives 0 (bound_ctrl), which is never read, so the shift needs
// no copy of the queue to preserve an `old` value and shr1 can consume the queue in place.
__device__ __forceinline__ int shl1(int v) { return __builtin_amdgcn_mov_dpp(v, kDppWaveShl1, 0xF, 0xF, true); }

// Copies a wave-uniform value into a VGPR the compiler cannot see through (a DPP
// row_mirror of a uniform value is the value itself), so that per-step selects keep
// both operands in registers instead of re-materialising them every step.
constexpr int kDppRowMirror = 0x140;
__device__ __forceinline__ int to_vgpr(int v) {
    return __builtin_amdgcn_mov_dpp(v, kDppRowMirror, 0xF, 0xF, false);
}

// Lane within the wavefront (a sweep workgroup holds several independent waves).
__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63u); }

// One tile of a sweep: rows [R0, R0 + 64R) clipped to r1.  Lane l takes column C0 + t - l at
// step t (every tile), over the tile columns x = t - l with, per lane,
//   first: x >= 0 (the lane starts at step l), else x >= -l (it starts at step 0);
//   last:  x <  w (it ends at step w - 1 + l), else x < w - l (it ends at step w - 1).
// A rectangular tile (columns [C0, C0 + w) for every lane) is first and last.  A skewed
// stage-1 stripe cuts its columns into segments of W columns for lane 0 and W columns shifted
// left by l for lane l: only its first segment starts lanes late and only its last ends them
// late, so its other tiles spend no steps filling or draining lanes.  In shifted stripes the
// segment grid of stripe s starts 64s columns left of column 0 (lane 0 of segment g takes
// columns [gW - 64s, gW - 64s + W), clipped to the region): the bottom row of stripe s - 1's
// segment g (lane 63's columns [gW - 64s + 1, ...)) and g - 1 then hold all the top boundary
// segment g of stripe s reads, so tile (s, g) follows both one diagonal later.
struct TileCtx {
    const uint8_t* A;   // the pair's rows and columns
    const uint8_t* B;
    int64_t m;          // the pair's rows (stride of the checkpoint columns)
    int64_t R0, r1, C0, w;
    bool first, last;
    int2* scol;         // stage 1: checkpoint columns (nullable), every Kc columns
    int64_t Kc;
    int32_t* rowmax;    // stage 1: row maxima
    int32_t* rowseg;    // stage 1 (nullable): first segment holding each row maximum
    int32_t seg;        // segment of the tile within its region
    int64_t target_row; // kSearch
    int32_t target;
    const int2* top;    // {H,F} of row R0-1 at columns C0.. (index col-C0); nullptr = zero row
    int2* bot;          // receives {H,F} of the tile's last row (index col-C0); nullptr = none
    int2* srow;         // stage 1: checkpoint row receiving the last row (index col-C0) or nullptr
    uint16_t* dirs;     // kDirs: [step][lane] codes of this tile
};

// Per-wave LDS of the sweep: lane 63's bottom-row outputs of the current chunk.  Every lane
// stores its bottom row each step, with no EXEC mask: lane 63 into out[step], lane l < 63
// into spare[l + step], which nothing reads (a masked store for lane 63 alone cost three
// more instructions per step).
struct SweepLds {
    int h[64], f[64];               // lane 63's {H+O, F+E} of steps 0..63
    int spare_h[128], spare_f[128];
};
// Where a lane's per-step bottom-row stores of a chunk go (see SweepLds): H and F go to
// separate arrays, so that the compiler pairs the stores of two steps without first moving
// the values into adjacent registers.
struct BottomSlots {
    int* h;
    int* f;
};
__device__ __forceinline__ BottomSlots bottom_slots(SweepLds& lds) {
    const int l = wave_lane();
    return l == 63 ? BottomSlots{lds.h, lds.f} : BottomSlots{lds.spare_h + l, lds.spare_f + l};
}

// Column symbol carried down the lanes: the byte itself, or with symbol codes the
// v_perm_b32 selector of its code (byte 0 picks profile byte `code`, bytes 1-3 read 0).
// Pad columns get a value that matches nothing (0x100 / all-zero selector).
constexpr int kPadByte = 0x100;
constexpr int kPadSel = 0x0c0c0c0c;
template <bool PROF>
__device__ __forceinline__ int col_symbol(const SweepParams& p, uint32_t byte) {
    if (!PROF) return (int)byte;
    return (int)(((byte >> (p.code_shift1 - 1)) & 7u) | 0x0c0c0c00u);
}

// Loads the inputs of chunk cc (steps [64cc, 64cc+64)) one column per lane: the top
// boundary {H+O, F+E} and the column symbol of column 64cc + lane.
template <bool PROF>
__device__ __forceinline__ void load_chunk_inputs(const SweepParams& p, const TileCtx& T, int64_t cc,
                                                  int2& in, int& b1) {
    const int lane = wave_lane();
    const int64_t col = 64 * cc + lane;
    int H = 0, F = 0;
    if (col < T.w && T.top) {
        const int2 v = T.top[col];
        H = v.x;
        F = v.y;
    }
    in = make_int2(H + p.sc.O, F + p.sc.E);
    b1 = (col < T.w) ? col_symbol<PROF>(p, T.B[T.C0 + col]) : (PROF ? kPadSel : kPadByte);
}

// Row symbols of a lane: lo[k] is row k's byte, or with symbol codes lo/hi[k] hold row k's
// 8-byte profile (bytes 0-3 / 4-7).  Rows past the region match nothing.
template <int R>
struct RowSyms {
    uint32_t lo[R], hi[R];
};
template <bool PROF>
__device__ __forceinline__ void row_symbol(const SweepParams& p, bool in, uint32_t byte, uint32_t& lo,
                                           uint32_t& hi) {
    if (!PROF) {
        lo = in ? byte : 0x200u;
        hi = 0;
        return;
    }
    const uint32_t code = (byte >> (p.code_shift1 - 1)) & 7u;
    const uint64_t prof = in ? (uint64_t)(uint32_t)(p.sc.M - p.sc.X) << (8 * code) : 0;
    lo = (uint32_t)prof;
    hi = (uint32_t)(prof >> 32);
}

// Checkpoint column of a stage-1 chunk: the step s in [0, 64) at which this lane's column
// C0 + 64c + s - lane (tile columns [0, w)) is one (column + 1 a multiple of Kc), or -1; dst
// receives the lane's first row of that checkpoint column.  Kc is a power of two >= 64, so a
// lane meets at most one checkpoint column in the 64 steps of a chunk.  The lane's tile
// columns are [xlo, xlo + xlen) (see TileCtx).
__device__ __forceinline__ int dump_step(int64_t C0, int64_t xlo, int64_t xlen, int64_t Kc, int2* scol, int64_t m,
                                         int64_t c, int64_t row0, int2*& dst) {
    dst = nullptr;
    if (!scol) return -1;
    const int64_t x0 = 64 * c - wave_lane();  // the lane's tile column at step 0
    const int64_t X1 = C0 + x0 + 1;
    const int64_t st = (-X1) & (Kc - 1);      // steps to the next column whose index + 1 is a multiple of Kc
    const int64_t x = x0 + st;
    if (st >= 64 || x < xlo || x >= xlo + xlen) return -1;
    dst = scol + (((X1 + st) >> __builtin_ctzll((unsigned long long)Kc)) - 1) * m + row0;
    return (int)st;
}

// 64 steps of one tile.  The chunk's inputs wait in register queues, one column per
// lane: Qh/Qf hold the top boundary, Qb the column symbols; each step lane 0 takes the head
// of the queues (DPP wave_shr with the queue as `old`) and the queues shift down one lane
// (wave_shl), so the inner loop touches no memory but lane 63's output.  bl is the
// column byte of the lane's current column, passed down the lanes with the diagonal.
// EDGE: some (lane, step) pairs fall outside [0, w) and must leave the lane state
// untouched.  DUMP: stage-1 chunk containing a checkpoint column.  START: the first chunk of
// a tile at least 64 columns wide, where the only steps outside [0, w) are those before a
// lane's first column (step < lane): those lanes sit the step out under the EXEC mask, which
// leaves their state untouched without per-value selects; the queue shifts above the branch
// run on every lane.  END: a later chunk, where the only steps outside [0, w) are those past
// a lane's last column; those lanes sit out the same way.
template <int R, int MODE, bool PROF, bool EDGE, bool DUMP, bool START = false, bool END = false>
__device__ __forceinline__ void chunk_steps(const SweepParams& p, const TileCtx& T, int64_t c, int64_t row0,
                                            const RowSyms<R>& a, int (&Ho)[R], int (&E)[R], int (&rmax)[R],
                                            int& Ho_up_prev, int& Ho_bot, int& Fe_bot, int& Qh, int& Qf, int& Qb,
                                            int& bl, BottomSlots s_out, int& found_col) {
    const int lane = wave_lane();
    const int O = p.sc.O, Ev = p.sc.E;
    const int MO = to_vgpr(p.sc.M - O), XO = to_vgpr(p.sc.X - O);
    const int64_t w = T.w;
    uint16_t* dcol = nullptr;
    if (MODE == kDirs) dcol = T.dirs + (64 * c) * 64 + lane;
    // the lane's tile columns: [xlo, xlo + xlen) (see TileCtx)
    const int xlo = T.first ? 0 : -lane;
    const int xlen = (int)(T.last ? w : w - lane) - xlo;
    int2* ddst = nullptr;
    const int sdump = DUMP ? dump_step(T.C0, xlo, xlen, T.Kc, T.scol, T.m, c, row0, ddst) : -1;
    int svH[R], svE[R];  // DUMP: the state at the lane's checkpoint column, stored after the chunk
    if (DUMP) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            svH[k] = Ho[k];
            svE[k] = E[k];
        }
    }
    const uint32_t col0 = (uint32_t)(64 * c) - (uint32_t)lane;  // the lane's tile column at step 0
    const int s_end = END ? (int)(w - (64 * c - lane)) : 64;  // first step past column w - 1
    SWK_UNROLL(UNROLL_STEPS)
    for (int s = 0; s < 64; ++s) {
        int Ho_up = shr1(Qh, Ho_bot);
        int Fe_up = shr1(Qf, Fe_bot);
        bl = shr1(Qb, bl);
        Qh = shl1(Qh);
        Qf = shl1(Qf);
        Qb = shl1(Qb);
        if (START && lane > s) continue;  // before the lane's first column
        if (END && s >= s_end) continue;  // past the lane's last column
        const uint32_t b = (uint32_t)bl;
        bool valid = true;
        int col = 0;
        if (EDGE || MODE == kSearch) {
            // tile column 64c + s - lane, in [-63, w + 126]: mod 2^32, the columns below the
            // lane's first compare above any width
            const uint32_t ucol = col0 + (uint32_t)s;
            valid = ucol - (uint32_t)xlo < (uint32_t)xlen;
            col = (int)ucol;
        }
        int diag = Ho_up_prev;
        if (EDGE) Ho_up_prev = valid ? Ho_up : Ho_up_prev;
        else Ho_up_prev = Ho_up;
        uint32_t nib = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int Hd;  // H(i-1,j-1) + s(i,j)
            if (PROF) Hd = diag + XO + (int)__builtin_amdgcn_perm(a.hi[k], a.lo[k], b);
            else Hd = diag + ((a.lo[k] == b) ? MO : XO);
            diag = Ho[k];              // H(i,j-1) + O: diagonal of the next row
            const int Ee = E[k] + Ev;
            const int En = max(Ee, Ho[k]);
            const int F = max(Fe_up, Ho_up);
            const int H = max(max(Hd, 0), max(En, F));
            if (MODE == kDirs) {
                // the cell's 4-bit code: bit 0 H != H(i-1,j-1) + s (not diagonal), bit 1 H != F,
                // bit 2 E extends, bit 3 F extends (the walk knows where H is 0 from its values)
                nib |= ((uint32_t)(H != Hd) | ((uint32_t)(H != F) << 1) | ((uint32_t)(Ee > Ho[k]) << 2) |
                        ((uint32_t)(Fe_up > Ho_up) << 3))
                       << (4 * k);
            }
            const int Hon = H + O;
            if (EDGE) {
                if (valid) {
                    E[k] = En;
                    Ho[k] = Hon;
                }
            } else {
                E[k] = En;
                Ho[k] = Hon;
            }
            if (MODE == kStage1) {
                if (!EDGE || valid) rmax[k] = max(rmax[k], H);
            }
            if (MODE == kSearch) {
                if (valid && row0 + k == T.target_row && H == T.target && found_col < 0) found_col = (int)col;
            }
            Ho_up = Hon;
            Fe_up = F + Ev;
        }
        Ho_bot = Ho_up;
        Fe_bot = Fe_up;
        s_out.h[s] = Ho_bot;  // every lane stores; only lane 63's slots are read
        s_out.f[s] = Fe_bot;
        if (MODE == kDirs) dcol[s * 64] = (uint16_t)nib;
        if (DUMP) {  // one select per value and step
            const bool at = s == sdump;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                svH[k] = at ? Ho[k] : svH[k];
                svE[k] = at ? E[k] : svE[k];
            }
        }
    }
    if (DUMP && sdump >= 0) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (row0 + k <= T.r1) ddst[k] = make_int2(svH[k] - O, svE[k]);
    }
}

// Per-lane hand-off between the segments of a skewed stripe: lane l's state at its last
// column of segment g (column C0 + W - 1 - l), which it resumes from in segment g + 1.  R + 1
// int2 per lane: {H, E} of its R rows, then {F + gap_extend of its last row, H of the cell
// above its first row}: the latter is the lane's diagonal carry, the H above-left of its first
// cell in segment g + 1.
template <int R>
constexpr int lane_slot() { return R + 1; }

// Everything a wave needs to sweep one tile (see make_tile).
struct TileSetup {
    TileCtx T;
    const int2* left;     // rectangular tiles: {H,E} of column C0-1 at row r is left[r - left_r0]
    int64_t left_r0;      // (nullptr = zero column)
    int corner;           // H(R0-1, C0-1)
    int2* right;          // rectangular tiles (nullable): the last column {H,E} at row - right_r0
    int64_t right_r0;
    const int2* lane_in;  // skewed tiles after a stripe's first: the previous segment's lane slots
    int2* lane_out;       // skewed tiles before a stripe's last: this segment's lane slots
    bool first_segment;
    int32_t* found;
};

// Sweeps one tile (see TileSetup for its inputs and outputs).
template <int R, int MODE, bool PROF>
__device__ void sweep_tile(const SweepParams& p, const TileSetup& ts, SweepLds& lds) {
    const TileCtx& T = ts.T;
    const BottomSlots s_slots = bottom_slots(lds);
    const int lane = wave_lane();
    const int O = p.sc.O;
    const int64_t w = T.w;
    const int64_t row0 = T.R0 + (int64_t)lane * R;
    constexpr int LS = lane_slot<R>();
    const int2* lin = ts.lane_in ? ts.lane_in + (int64_t)lane * LS : nullptr;

    RowSyms<R> a;
    int Ho[R], E[R], rmax[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int64_t row = row0 + k;
        const bool in = row <= T.r1;
        row_symbol<PROF>(p, in, in ? (uint32_t)T.A[row] : 0u, a.lo[k], a.hi[k]);
        if (lin) {
            const int2 v = lin[k];
            Ho[k] = v.x + O;
            E[k] = v.y;
        } else if (ts.left && in) {
            const int2 v = ts.left[row - ts.left_r0];
            Ho[k] = v.x + O;
            E[k] = v.y;
        } else {
            Ho[k] = O;
            E[k] = 0;
        }
        rmax[k] = 0;
    }
    // H + O of the cell above-left of the lane's first cell: lane 0 takes the corner, the
    // other lanes the left column (rectangular) or their slot (skewed).
    int Ho_up_prev = ts.corner + O;
    if (lin) Ho_up_prev = lin[R].y + O;
    else if (lane > 0) Ho_up_prev = (ts.left && row0 - 1 <= T.r1) ? ts.left[row0 - 1 - ts.left_r0].x + O : O;
    // The bottom row {H+O, F+E} of the lane's previous column, which the lane below takes at
    // its first step; a skewed tile's lanes all start at step 0, so theirs come from the slots.
    int Ho_bot = O, Fe_bot = p.sc.E;
    int bl = PROF ? kPadSel : kPadByte;  // symbol of the lane's previous column, likewise
    if (lin) {
        Ho_bot = Ho[R - 1];
        Fe_bot = lin[R].x;
        bl = col_symbol<PROF>(p, T.B[T.C0 - lane - 1]);
    }
    int found_col = -1;

    // steps: lanes end at step w - 1 + lane in a last segment, at step w - 1 otherwise
    const int64_t nsteps = T.last ? w + 63 : w;
    const int64_t nchunks = (nsteps + 63) / 64;
    // chunks where every lane is inside the tile at every step
    const int64_t first_interior = T.first ? 1 : 0;
    const int64_t last_interior = T.last ? (w >= 64 ? (w - 64) / 64 : -1) : nchunks - 1;

    int2 in_reg;
    int b_reg;
    load_chunk_inputs<PROF>(p, T, 0, in_reg, b_reg);
    int Qh = in_reg.x, Qf = in_reg.y, Qb = b_reg;
    // kStage1: the row maxima of the stripe's earlier segments, read as the last chunk starts
    // so that the read is in flight during its steps rather than after them (rows belong to
    // one stripe, whose one tile per launch is the only access to them)
    int pmax[R];
#pragma unroll
    for (int k = 0; k < R; ++k) pmax[k] = 0;

    for (int64_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks) load_chunk_inputs<PROF>(p, T, c + 1, in_reg, b_reg);
        if (MODE == kStage1 && c + 1 == nchunks && !ts.first_segment) {
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (row0 + k <= T.r1) pmax[k] = T.rowmax[row0 + k];
        }
        const bool interior = c >= first_interior && c <= last_interior;
        const bool starts = T.first && c == 0 && (!T.last || w >= 64);   // lanes start here, none ends
        const bool ends = c >= first_interior && T.last;                  // lanes end here, none starts
        bool dump = false;
        if (MODE == kStage1 && T.scol) {
            // checkpoint columns among the columns the chunk's lanes take, but each lane's last,
            // which goes out after the tile (no per-step dump when Kc is a multiple of W)
            const int64_t lo = T.C0 + max(64 * c - 63, T.first ? (int64_t)0 : (int64_t)-63);
            const int64_t hi = T.C0 + min(64 * c + 63, w - 2);
            dump = (lo | (T.Kc - 1)) <= hi;  // smallest x >= lo with x = -1 mod Kc (a power of two)
        }
        if (!dump) {
            if (interior)
                chunk_steps<R, MODE, PROF, false, false>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                                   Qh, Qf, Qb, bl, s_slots, found_col);
            else if (starts)
                chunk_steps<R, MODE, PROF, false, false, true>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot,
                                                         Fe_bot, Qh, Qf, Qb, bl, s_slots, found_col);
            else if (ends)
                chunk_steps<R, MODE, PROF, false, false, false, true>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev,
                                                                Ho_bot, Fe_bot, Qh, Qf, Qb, bl, s_slots, found_col);
            else
                chunk_steps<R, MODE, PROF, true, false>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                                  Qh, Qf, Qb, bl, s_slots, found_col);
        } else if (interior) {
            chunk_steps<R, MODE, PROF, false, true>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot, Qh, Qf,
                                              Qb, bl, s_slots, found_col);
        } else if (ends) {
            chunk_steps<R, MODE, PROF, false, true, false, true>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot,
                                                           Fe_bot, Qh, Qf, Qb, bl, s_slots, found_col);
        } else {
            chunk_steps<R, MODE, PROF, true, true>(p, T, c, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot, Qh, Qf,
                                             Qb, bl, s_slots, found_col);
        }
        // Bottom row of the tile: lane 63's outputs of the chunk's steps 64c + lane, tile
        // column 64c + lane - 63 where lane 63 was inside the tile.  The LDS hand-off from lane
        // 63 to the other lanes needs no hardware barrier (one wave, LDS operations in order)
        // but must not be reordered by the compiler.
        __builtin_amdgcn_wave_barrier();
        {
            const int2 o = make_int2(lds.h[lane], lds.f[lane]);
            const int64_t col = 64 * c + lane - 63;
            if (col >= (T.first ? 0 : -63) && col < (T.last ? w : w - 63)) {
                const int2 v = make_int2(o.x - O, o.y - p.sc.E);
                if (T.bot) T.bot[col] = v;
                if (MODE == kStage1 && T.srow) T.srow[col] = v;
            }
        }
        Qh = in_reg.x;
        Qf = in_reg.y;
        Qb = b_reg;
        __builtin_amdgcn_wave_barrier();
    }
    if (ts.right) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (row0 + k <= T.r1) ts.right[row0 + k - ts.right_r0] = make_int2(Ho[k] - O, E[k]);
    }
    if (ts.lane_out) {  // a skewed segment's every lane ran to step w - 1
        int2* lo = ts.lane_out + (int64_t)lane * LS;
#pragma unroll
        for (int k = 0; k < R; ++k) lo[k] = make_int2(Ho[k] - O, E[k]);
        lo[R] = make_int2(Fe_bot, Ho_up_prev - O);
    }
    if (MODE == kStage1 && T.scol) {
        // the lane's last column (C0 + w - 1, or C0 + w - 1 - lane in a skewed segment before
        // the last) is a checkpoint column: the lane's state is that column's
        const int64_t xl = T.C0 + w - 1 - (T.last ? 0 : lane);
        if (w > 0 && ((xl + 1) & (T.Kc - 1)) == 0) {
            int2* dst = T.scol + (((xl + 1) >> __builtin_ctzll((unsigned long long)T.Kc)) - 1) * T.m;
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (row0 + k <= T.r1) dst[row0 + k] = make_int2(Ho[k] - O, E[k]);
        }
    }
    if (MODE == kStage1) {
        // segments of a stripe run in increasing order (one launch each), so a strictly
        // larger value keeps the first segment that holds the row maximum
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int64_t row = row0 + k;
            if (row <= T.r1 && (ts.first_segment || rmax[k] > pmax[k])) {
                T.rowmax[row] = rmax[k];
                if (T.rowseg) T.rowseg[row] = T.seg;
            }
        }
    }
    if (MODE == kSearch) {
        if (found_col >= 0) atomicMin(ts.found, (int32_t)(T.C0 + found_col));
    }
}

// Locates tile x of the launch (binary search over the launch's spans) and its inputs and
// outputs; false if the tile does not exist.  Rectangular tiles: tile (s, g) of a span at
// diagonal d has g = d - s; it hands the next segment of the stripe its corner (the top
// boundary at this tile's last column).  Skewed stage-1 tiles (Region.skew): g = d - 2s, since
// segment g of stripe s needs segments g and g + 1 of stripe s - 1 (the bottom row of the
// stripe above at its columns [C0, C0 + w) ends 63 columns into the next segment).
// SHIFT: the kernel takes shifted stripes (the int32 sweep; the packed one holds 29 more
// VGPRs with that path compiled in, and its sweeps run skewed stripes instead).
template <int R, int MODE, bool SHIFT = true>
__device__ bool make_tile(const SweepParams& p, int64_t x, TileSetup& ts, int* span_io = nullptr) {
    constexpr int64_t HR = 64 * R;
    if (x >= p.ntiles) return false;
    // span holding tile x: from the caller's span (x follows a tile of it) when given, whose
    // successor usually starts beyond x, else by binary search over the launch's spans
    int lo = span_io && *span_io >= 0 ? *span_io : 0, hi = p.nspans - 1;
    if (span_io && *span_io >= 0 && (lo == hi || p.spans[lo + 1].first_tile > x)) hi = lo;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.spans[mid].first_tile <= x) lo = mid;
        else hi = mid - 1;
    }
    if (span_io) *span_io = lo;
    const TileSpan sp = p.spans[lo];
    const Region Rg = p.regions[sp.region];
    const bool skew = MODE == kStage1 && Rg.skew;
    const bool shifted = SHIFT && skew && Rg.skew == 2;
    const int64_t s = sp.s_lo + (x - sp.first_tile);
    // diagonal d: s + g (rectangular, shifted), 2s + g (skewed)
    const int64_t g = sp.d - (skew && !shifted ? 2 * s : s);
    const int64_t rows = Rg.r1 - Rg.r0 + 1, cols = Rg.c1 - Rg.c0 + 1;
    const int64_t W = Rg.W;
    const int64_t nstr = (rows + HR - 1) / HR, nseg = (cols + W - 1) / W;
    // lane 0's columns [x0, x0 + W) relative to c0, clipped to [0, cols); a segment exists
    // while they reach the region (a narrower region has fewer segments)
    int64_t x0 = g * W;
    if (shifted) {
        x0 -= 64 * s;
        if (s >= nstr || x0 + W <= 0 || x0 >= cols) return false;
    } else if (s >= nstr || g < 0 || g >= nseg) {
        return false;
    }
    TileCtx& T = ts.T;
    T.A = Rg.A;
    T.B = Rg.B;
    T.m = Rg.m;
    T.scol = Rg.scol;
    T.Kc = Rg.Kc;
    T.rowmax = Rg.rowmax;
    T.rowseg = Rg.rowseg;
    T.seg = (int32_t)g;
    T.target_row = Rg.target_row;
    T.target = Rg.target;
    T.R0 = Rg.r0 + s * HR;
    T.r1 = Rg.r1;
    if (shifted) {  // a first segment starts at column c0, a last one ends at c1
        T.C0 = Rg.c0 + max((int64_t)0, x0);
        T.w = min(x0 + W, cols) - max((int64_t)0, x0);
        T.first = x0 <= 0;
        T.last = x0 + W >= cols;
    } else {  // the same values for g >= 0, in the forms that keep the packed kernel's registers
        T.C0 = Rg.c0 + g * W;
        T.w = min(W, Rg.c1 - T.C0 + 1);
        T.first = !skew || g == 0;
        T.last = !skew || g + 1 == nseg;
    }
    const int64_t coff = T.C0 - Rg.c0;
    T.top = s == 0 ? (Rg.top ? Rg.top + coff : nullptr) : Rg.rb + (s & 1) * cols + coff;
    T.bot = s + 1 < nstr ? Rg.rb + ((s + 1) & 1) * cols + coff : nullptr;
    T.srow = nullptr;
    if (MODE == kStage1 && Rg.srow) {
        const int64_t brow = T.R0 + HR - 1;
        if (brow <= Rg.r1 && (brow + 1) % Rg.K == 0) T.srow = Rg.srow + ((brow + 1) / Rg.K - 1) * Rg.n + T.C0;
    }
    T.dirs = MODE == kDirs ? Rg.dirs + (s * Rg.G + g) * Rg.tstride * 64 : nullptr;
    ts.lane_in = nullptr;
    ts.lane_out = nullptr;
    ts.right = nullptr;
    ts.right_r0 = Rg.r0;
    if (skew) {
        // lane slots: [segment parity][stripe][lane]; the bottom row of the stripe above (or
        // the region's top row) holds the corner H(R0-1, C0-1) one column left of C0
        ts.left = T.first ? Rg.left : nullptr;
        ts.left_r0 = 0;
        if (T.first) ts.corner = (Rg.left && T.R0 >= 1) ? Rg.left[T.R0 - 1].x : 0;
        else ts.corner = T.top ? T.top[-1].x : 0;
        constexpr int64_t LS = lane_slot<R>();
        if (!T.first) ts.lane_in = Rg.cb + ((g & 1) * nstr + s) * 64 * LS;
        if (!T.last) ts.lane_out = Rg.cb + (((g + 1) & 1) * nstr + s) * 64 * LS;
    } else {
        // Left column and corner H(R0-1, C0-1).
        if (g == 0) {
            ts.left = Rg.left;
            ts.left_r0 = 0;
            ts.corner = (Rg.left && T.R0 >= 1) ? Rg.left[T.R0 - 1].x : 0;
        } else {
            ts.left = Rg.cb + (g & 1) * rows;
            ts.left_r0 = Rg.r0;
            if (s == 0) ts.corner = Rg.top ? Rg.top[coff - 1].x : 0;
            else ts.corner = Rg.cor[2 * s + (g & 1)];
        }
        if (s > 0 && g + 1 < nseg && wave_lane() == 0) Rg.cor[2 * s + ((g + 1) & 1)] = T.top[T.w - 1].x;
        if (g + 1 < nseg) ts.right = Rg.cb + ((g + 1) & 1) * rows;
    }
    ts.first_segment = x0 <= 0;
    ts.found = Rg.found;
    return true;
}

// One wave = one tile of the launch; tile index x = wave number in the grid.  Waves of a
// workgroup are independent.
template <int R, int MODE, bool PROF>
__device__ void sweep_main(const SweepParams& p, SweepLds& lds) {
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
    const int64_t x = (int64_t)blockIdx.x * kSweepWaves + wid;
    TileSetup ts;
    if (!make_tile<R, MODE>(p, x, ts)) return;
    sweep_tile<R, MODE, PROF>(p, ts, lds);
}

// ---- packed 16-bit stage 1: two tiles per wave --------------------------------------
// The low 16 bits of every value belong to tile 0 of the wave and the high 16 bits to tile
// 1 (any two tiles of the launch: tiles of one anti-diagonal are independent), so one
// packed operation (v_pk_*) serves two cells.  Values are held relative to a base per
// tile, the H of the cell above and left of the tile's first cell: H changes by at most
// match - gap_open between neighbouring cells (in either direction), so every H, E, F of a
// tile and of its inputs lies within (64R + W + 2) * (match - gap_open) + |gap_open| +
// |gap_extend| of the base, and the host runs this kernel only when that fits int16 (the
// absolute scores may be far larger).  The zero floor of H becomes max(-base, -32768): when
// the base exceeds 32768 no cell of the tile is near zero.  Inputs and outputs in memory
// stay int32 and absolute.
typedef short s16x2 __attribute__((vector_size(4)));
typedef unsigned short u16x2 __attribute__((vector_size(4)));
// Per half: 1 if the (non-negative, below 65536) difference d is not zero, else 0.
// One v_pk_min_u16, written as inline asm: the compiler recognises min(x, 1) as "x != 0" and
// would lower it, per half, to two compares and two selects.
__device__ __forceinline__ uint32_t differs(s16x2 d) {
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(__builtin_bit_cast(uint32_t, d)), "v"(0x00010001u));
    return r;
}

__device__ __forceinline__ int pk_bits(s16x2 v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ s16x2 pk_from(int v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ s16x2 pk2(int lo, int hi) {
    s16x2 v;
    v[0] = (short)lo;
    v[1] = (short)hi;
    return v;
}
__device__ __forceinline__ s16x2 pk_max(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 pk_min(s16x2 a, s16x2 b) { return __builtin_elementwise_min(a, b); }
// Per-half select: halves whose mask bits are set take `a`, the others keep `b`.
__device__ __forceinline__ s16x2 pk_sel(int mask, s16x2 a, s16x2 b) {
    return pk_from((pk_bits(a) & mask) | (pk_bits(b) & ~mask));
}

// 64 steps of two tiles (see chunk_steps).  EDGE: per-half validity masks.
template <int R, int MODE, bool PROF, bool EDGE, bool DUMP, bool START = false, bool END = false>
__device__ __forceinline__ void chunk_steps_pk(const SweepParams& p, const TileCtx (&T)[2], int64_t c,
                                               const int64_t (&row0)[2], const int (&a)[R], const int (&a2)[R],
                                               s16x2 (&Ho)[R],
                                               s16x2 (&E)[R], s16x2 (&rmax)[R], s16x2& Ho_up_prev, s16x2& Ho_bot,
                                               s16x2& Fe_bot, int& Qh, int& Qf, int& Qb, int& bl, BottomSlots s_out,
                                               const s16x2 zero, const int (&base)[2]) {
    const int lane = wave_lane();
    const s16x2 vO = pk2(p.sc.O, p.sc.O), vE = pk2(p.sc.E, p.sc.E), one = pk2(1, 1);
    const s16x2 vMO = pk2(p.sc.M - p.sc.O, p.sc.M - p.sc.O);
    const s16x2 vD = pk2(p.sc.M - p.sc.X, p.sc.M - p.sc.X);  // match - mismatch
    // kDirs: both tiles' code columns.  A wave without a second tile stores the second half's
    // codes into the first tile's slot before the first half's own, which the later store
    // overwrites (one lane's stores to one address complete in program order): no branch.
    uint16_t* dcol[2] = {nullptr, nullptr};
    if (MODE == kDirs) {
        dcol[0] = T[0].dirs + (64 * c) * 64 + lane;
        dcol[1] = T[1].dirs ? T[1].dirs + (64 * c) * 64 + lane : dcol[0];
    }
    const uint32_t col0 = (uint32_t)(64 * c) - (uint32_t)lane;  // the lane's tile column at step 0
    const int s_end = END ? (int)(T[0].w - (64 * c - lane)) : 64;  // first step past column w - 1
    // each tile's columns of the lane: [xlo, xlo + xlen) (see TileCtx)
    int xlo[2], xlen[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        xlo[t] = T[t].first ? 0 : -lane;
        xlen[t] = (int)(T[t].last ? T[t].w : T[t].w - lane) - xlo[t];
    }
    int2* ddst[2] = {nullptr, nullptr};
    int sdump[2] = {-1, -1};
    s16x2 svH[R], svE[R];  // DUMP: each half's state at its checkpoint column, stored after the chunk
    if (DUMP) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
            sdump[t] = dump_step(T[t].C0, xlo[t], xlen[t], T[t].Kc, T[t].scol, T[t].m, c, row0[t], ddst[t]);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            svH[k] = Ho[k];
            svE[k] = E[k];
        }
    }
    SWK_UNROLL(UNROLL_STEPS_PK)
    for (int s = 0; s < 64; ++s) {
        s16x2 Ho_up = pk_from(shr1(Qh, pk_bits(Ho_bot)));
        s16x2 Fe_up = pk_from(shr1(Qf, pk_bits(Fe_bot)));
        bl = shr1(Qb, bl);
        Qh = shl1(Qh);
        Qf = shl1(Qf);
        Qb = shl1(Qb);
        if (START && lane > s) continue;  // before the lane's first column (see chunk_steps)
        if (END && s >= s_end) continue;  // past the lane's last column (both tiles: equal widths)
        int vmask = -1;
        if (EDGE) {
            // tile column 64c + s - lane mod 2^32: the columns below the lane's first compare
            // above any width
            const uint32_t col = col0 + (uint32_t)s;
            vmask = (col - (uint32_t)xlo[0] < (uint32_t)xlen[0] ? 0x0000FFFF : 0) |
                    (col - (uint32_t)xlo[1] < (uint32_t)xlen[1] ? (int)0xFFFF0000 : 0);
        }
        s16x2 diag = Ho_up_prev;
        Ho_up_prev = EDGE ? pk_sel(vmask, Ho_up, Ho_up_prev) : Ho_up;
        uint32_t nib = 0;  // kDirs: both tiles' codes, tile 0 in the low 16 bits
#pragma unroll
        for (int k = 0; k < R; ++k) {
            s16x2 Hd;  // H(i-1,j-1) + s(i,j)
            if (PROF) {
                // a[k]: the two rows' profiles (byte c of each half = s - O for code c); bl:
                // the selector of the two columns' codes
                Hd = diag + pk_from((int)__builtin_amdgcn_perm((uint32_t)a2[k], (uint32_t)a[k], (uint32_t)bl));
            } else {
                // substitution: the halves of a ^ b are zero exactly where the bytes agree
                const s16x2 ne = pk_min(pk_from(a[k] ^ bl), one);
                Hd = diag + vMO - ne * vD;
            }
            diag = Ho[k];
            const s16x2 Ee = E[k] + vE;
            const s16x2 En = pk_max(Ee, Ho[k]);
            const s16x2 F = pk_max(Fe_up, Ho_up);
            const s16x2 H = pk_max(pk_max(Hd, zero), pk_max(En, F));
            if (MODE == kDirs) {
                // the 4-bit code of each half, as k_sweep's: bit 0 H != Hd, bit 1 H != F, bit
                // 2 E extends, bit 3 F extends.  Differences H - x are >= 0, so min(., 1) is
                // the half's "differs" flag; the sign bit of Ho - Ee is "E extends" (all values
                // of a tile lie within 32000 of each other).  Halves never carry into each
                // other (values <= 15).  The sign bits go straight to their nibble positions
                // (one shift and one and-or each).
                const uint32_t nd = differs(H - Hd), nf = differs(H - F);
                const uint32_t eb = 0x00010001u << (4 * k + 2), fb = 0x00010001u << (4 * k + 3);
                nib = (((uint32_t)pk_bits(Ho[k] - Ee) >> (13 - 4 * k)) & eb) | nib;
                nib = (((uint32_t)pk_bits(Ho_up - Fe_up) >> (12 - 4 * k)) & fb) | nib;
                nib = ((nd | (nf << 1)) << (4 * k)) | nib;
            }
            const s16x2 Hon = H + vO;
            if (EDGE) {
                E[k] = pk_sel(vmask, En, E[k]);
                Ho[k] = pk_sel(vmask, Hon, Ho[k]);
                if (MODE == kStage1) rmax[k] = pk_sel(vmask, pk_max(rmax[k], H), rmax[k]);
            } else {
                E[k] = En;
                Ho[k] = Hon;
                if (MODE == kStage1) rmax[k] = pk_max(rmax[k], H);
            }
            Ho_up = Hon;
            Fe_up = F + vE;
        }
        Ho_bot = Ho_up;
        Fe_bot = Fe_up;
        s_out.h[s] = pk_bits(Ho_bot);  // see chunk_steps
        s_out.f[s] = pk_bits(Fe_bot);
        if (MODE == kDirs) {
            dcol[1][s * 64] = (uint16_t)(nib >> 16);  // first: see dcol
            dcol[0][s * 64] = (uint16_t)nib;
        }
        if (DUMP) {
            // one bitwise select per value and step: the halves at their dump step take the state
            const int dm = (s == sdump[0] ? 0x0000FFFF : 0) | (s == sdump[1] ? (int)0xFFFF0000 : 0);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                svH[k] = pk_sel(dm, Ho[k], svH[k]);
                svE[k] = pk_sel(dm, E[k], svE[k]);
            }
        }
    }
    if (DUMP) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (sdump[t] < 0) continue;
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (row0[t] + k <= T[t].r1) ddst[t][k] = make_int2(svH[k][t] - p.sc.O + base[t], svE[k][t] + base[t]);
        }
    }
}

// Stage 1 over two tiles at once (has[1] false: the second half runs on empty input and
// stores nothing).
template <int R, int MODE, bool PROF>
__device__ void sweep_tile_pk(const SweepParams& p, const TileSetup (&ts)[2], SweepLds& lds) {
    const BottomSlots s_slots = bottom_slots(lds);
    const int lane = wave_lane();
    const int O = p.sc.O, Ev = p.sc.E;
    TileCtx T[2] = {ts[0].T, ts[1].T};
    const int64_t row0[2] = {T[0].R0 + (int64_t)lane * R, T[1].R0 + (int64_t)lane * R};
    constexpr int LS = lane_slot<R>();
    const int2* lin[2];  // skewed tiles after a stripe's first: this lane's slot (see sweep_tile)
#pragma unroll
    for (int t = 0; t < 2; ++t) lin[t] = ts[t].lane_in ? ts[t].lane_in + (int64_t)lane * LS : nullptr;
    int a2[R];  // with symbol codes: tile 1's row profiles
    const int base[2] = {ts[0].corner, ts[1].corner};  // every value is held as value - base
    const int floor16[2] = {max(-base[0], -32768), max(-base[1], -32768)};
    const s16x2 zero = pk2(floor16[0], floor16[1]);  // H's zero floor, relative
    int a[R];
    s16x2 Ho[R], E[R], rmax[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int av[2], hv[2], ev[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int64_t row = row0[t] + k;
            const bool in = row <= T[t].r1;
            av[t] = in ? (int)T[t].A[row] : 0x200;
            if (lin[t]) {
                const int2 v = lin[t][k];
                hv[t] = v.x + O - base[t];
                ev[t] = v.y - base[t];
            } else if (ts[t].left && in) {
                const int2 v = ts[t].left[row - ts[t].left_r0];
                hv[t] = v.x + O - base[t];
                ev[t] = v.y - base[t];
            } else {
                hv[t] = O - base[t];
                ev[t] = -base[t];
            }
        }
        if (PROF) {
            // row profiles: tile 0's in a[k] (byte c = s(row, code c) - O), tile 1's in a2[k]
            const int sh = p.code2_shift1 - 1;
            const uint32_t mo = (uint32_t)(p.sc.M - O), xo = (uint32_t)(p.sc.X - O);
            uint32_t pr[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const uint32_t code = ((uint32_t)av[t] >> sh) & 3u;
                pr[t] = (xo * 0x01010101u) & ~(0xFFu << (8 * code));
                pr[t] |= mo << (8 * code);
            }
            a[k] = (int)pr[0];
            a2[k] = (int)pr[1];
        } else {
            a[k] = (av[0] & 0xFFFF) | (av[1] << 16);
            a2[k] = 0;
        }
        Ho[k] = pk2(hv[0], hv[1]);
        E[k] = pk2(ev[0], ev[1]);
        rmax[k] = zero;
    }
    // per tile as in sweep_tile: the cell above-left of the lane's first cell, the bottom row
    // and the column symbol of the lane's previous column
    int up[2], hb[2], fb[2], sym[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        up[t] = ts[t].corner + O - base[t];
        if (lin[t])
            up[t] = lin[t][R].y + O - base[t];
        else if (lane > 0)
            up[t] = ((ts[t].left && row0[t] - 1 <= T[t].r1) ? ts[t].left[row0[t] - 1 - ts[t].left_r0].x : 0) + O -
                    base[t];
        hb[t] = O;
        fb[t] = Ev;
        sym[t] = PROF ? 0x0c : 0x100;
        if (lin[t]) {
            hb[t] = Ho[R - 1][t];
            fb[t] = lin[t][R].x - base[t];
            const uint32_t byte = T[t].B[T[t].C0 - lane - 1];
            sym[t] = PROF ? (int)(((byte >> (p.code2_shift1 - 1)) & 3u) + 4 * t) : (int)byte;
        }
    }
    s16x2 Ho_up_prev = pk2(up[0], up[1]);
    s16x2 Ho_bot = pk2(hb[0], hb[1]), Fe_bot = pk2(fb[0], fb[1]);

    // steps and interior chunks of each tile (see sweep_tile); the wave runs the longer
    int64_t nchunks = 0, first_interior = 0, last_interior = INT64_MAX;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t nst = T[t].last ? T[t].w + 63 : T[t].w;
        nchunks = max(nchunks, (nst + 63) / 64);
        first_interior = max(first_interior, (int64_t)(T[t].first ? 1 : 0));
        last_interior = min(last_interior, T[t].last ? (T[t].w >= 64 ? (T[t].w - 64) / 64 : (int64_t)-1)
                                                     : (nst + 63) / 64 - 1);
    }
    // lanes start in chunk 0 of both tiles and end in neither / end in both tiles alike and
    // start in neither: the EXEC-mask variants; otherwise per-half selects
    const bool both_start = T[0].first && T[1].first && (!T[0].last || T[0].w >= 64) && (!T[1].last || T[1].w >= 64);
    const bool same_end = T[0].last && T[1].last && T[0].w == T[1].w;
    // a skewed tile before its stripe's last ends with the chunk that holds its step w - 1
    // (w a multiple of 64); its bottom row's F + E then goes to the next segment: kept here,
    // since a longer second tile keeps the wave stepping (its H, E and diagonal carry stay
    // as they were: the later chunks' per-half selects leave an ended tile's state alone)
    s16x2 end_fe = Fe_bot;

    auto load = [&](int64_t cc, int& qh, int& qf, int& qb) {
        int h[2], f[2], b[2];
        const int64_t col = 64 * cc + lane;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            int H = 0, F = 0;
            if (col < T[t].w && T[t].top) {
                const int2 v = T[t].top[col];
                H = v.x;
                F = v.y;
            }
            h[t] = H + O - base[t];
            f[t] = F + Ev - base[t];
            if (PROF)  // v_perm_b32 selector bytes: tile 0 reads profile byte code, tile 1 byte 4 + code
                b[t] = (col < T[t].w) ? (int)((((uint32_t)T[t].B[T[t].C0 + col] >> (p.code2_shift1 - 1)) & 3u) + 4 * t)
                                      : 0x0c;
            else
                b[t] = (col < T[t].w) ? (int)T[t].B[T[t].C0 + col] : 0x100;
        }
        qh = pk_bits(pk2(h[0], h[1]));
        qf = pk_bits(pk2(f[0], f[1]));
        qb = PROF ? (b[0] | 0x0c00 | (b[1] << 16) | 0x0c000000) : ((b[0] & 0xFFFF) | (b[1] << 16));
    };
    int Qh, Qf, Qb, nh, nf, nb;
    int bl = PROF ? (sym[0] | 0x0c00 | (sym[1] << 16) | 0x0c000000) : ((sym[0] & 0xFFFF) | (sym[1] << 16));
    load(0, Qh, Qf, Qb);
    // (the row maxima are read after the steps: reading them ahead, as sweep_tile does, holds
    // eight more registers across the loop, which cost more than the read's latency)
    for (int64_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks) load(c + 1, nh, nf, nb);
        const bool interior = c >= first_interior && c <= last_interior;
        const bool starts = both_start && c == 0;
        const bool ends = same_end && c >= first_interior;
        bool dump = false;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (MODE == kStage1 && T[t].scol && T[t].w > 0) {  // as in sweep_tile (each lane's last column: below)
                const int64_t lo = T[t].C0 + max(64 * c - 63, T[t].first ? (int64_t)0 : (int64_t)-63);
                const int64_t hi = T[t].C0 + min(64 * c + 63, T[t].w - 2);
                dump = dump || (lo | (T[t].Kc - 1)) <= hi;
            }
        }
        if (!dump) {
            if (interior)
                chunk_steps_pk<R, MODE, PROF, false, false>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                                      Qh, Qf, Qb, bl, s_slots, zero, base);
            else if (starts)
                chunk_steps_pk<R, MODE, PROF, false, false, true>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev, Ho_bot,
                                                            Fe_bot, Qh, Qf, Qb, bl, s_slots, zero, base);
            else if (ends)
                chunk_steps_pk<R, MODE, PROF, false, false, false, true>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev,
                                                                   Ho_bot, Fe_bot, Qh, Qf, Qb, bl, s_slots, zero, base);
            else
                chunk_steps_pk<R, MODE, PROF, true, false>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                                     Qh, Qf, Qb, bl, s_slots, zero, base);
        } else if (interior) {
            chunk_steps_pk<R, MODE, PROF, false, true>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                                 Qh, Qf, Qb, bl, s_slots, zero, base);
        } else if (ends) {
            chunk_steps_pk<R, MODE, PROF, false, true, false, true>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev,
                                                              Ho_bot, Fe_bot, Qh, Qf, Qb, bl, s_slots, zero, base);
        } else {
            chunk_steps_pk<R, MODE, PROF, true, true>(p, T, c, row0, a, a2, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot, Qh,
                                                Qf, Qb, bl, s_slots, zero, base);
        }
        __builtin_amdgcn_wave_barrier();
        {
            const int2 o = make_int2(lds.h[lane], lds.f[lane]);
            const int64_t col = 64 * c + lane - 63;
            const s16x2 oh = pk_from(o.x), of = pk_from(o.y);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (col >= (T[t].first ? 0 : -63) && col < (T[t].last ? T[t].w : T[t].w - 63)) {
                    const int2 v = make_int2(oh[t] - O + base[t], of[t] - Ev + base[t]);
                    if (T[t].bot) T[t].bot[col] = v;
                    if (MODE == kStage1 && T[t].srow) T[t].srow[col] = v;
                }
            }
            // tiles ending with this chunk (skewed, before their stripe's last)
#pragma unroll
            for (int t = 0; t < 2; ++t)
                if (!T[t].last && c == T[t].w / 64 - 1) end_fe = pk_sel(t ? (int)0xFFFF0000 : 0x0000FFFF, Fe_bot, end_fe);
        }
        Qh = nh;
        Qf = nf;
        Qb = nb;
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if (ts[t].lane_out) {
            int2* lo = ts[t].lane_out + (int64_t)lane * LS;
#pragma unroll
            for (int k = 0; k < R; ++k) lo[k] = make_int2(Ho[k][t] - O + base[t], E[k][t] + base[t]);
            lo[R] = make_int2(end_fe[t] + base[t], Ho_up_prev[t] - O + base[t]);
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int64_t row = row0[t] + k;
            if (row > T[t].r1) continue;
            if (ts[t].right) ts[t].right[row - ts[t].right_r0] = make_int2(Ho[k][t] - O + base[t], E[k][t] + base[t]);
            if (MODE != kStage1) continue;
            // the lane's last column (see sweep_tile) is a checkpoint column
            const int64_t xl = T[t].C0 + T[t].w - 1 - (T[t].last ? 0 : lane);
            if (T[t].scol && T[t].w > 0 && ((xl + 1) & (T[t].Kc - 1)) == 0)
                T[t].scol[(((xl + 1) >> __builtin_ctzll((unsigned long long)T[t].Kc)) - 1) * T[t].m + row] =
                    make_int2(Ho[k][t] - O + base[t], E[k][t] + base[t]);
            const int r = rmax[k][t] + base[t];
            if (ts[t].first_segment || r > T[t].rowmax[row]) {  // first segment holding the maximum
                T[t].rowmax[row] = r;
                if (T[t].rowseg) T[t].rowseg[row] = T[t].seg;
            }
        }
    }
}

// Packed stage 1: wave x takes tiles 2x and 2x+1 of the launch.
template <int R, int MODE, bool PROF>
__device__ void sweep_main_pk(const SweepParams& p, SweepLds& lds) {
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t x = (int64_t)blockIdx.x * kSweepWaves + wid;
    TileSetup ts[2];
    int span = -1;
    // Either tile may be missing (spans built on the device keep a fixed count of tiles per
    // region, some of which a smaller region does not have); a lone second tile runs first.
    const bool has0 = make_tile<R, MODE, false>(p, 2 * x, ts[0], &span);
    const bool has1 = make_tile<R, MODE, false>(p, 2 * x + 1, ts[1], &span);
    if (!has0 && !has1) return;
    if (!has0) ts[0] = ts[1];
    if (!has0 || !has1) {
        // one tile: the second is an empty copy of it (no rows, no inputs, nothing stored),
        // which steps along with the same columns so that the wave keeps the fast variants
        ts[1] = ts[0];
        ts[1].T.r1 = ts[1].T.R0 - 1;
        ts[1].T.top = nullptr;
        ts[1].T.bot = nullptr;
        ts[1].T.srow = nullptr;
        ts[1].T.scol = nullptr;
        ts[1].T.rowseg = nullptr;
        ts[1].T.dirs = nullptr;
        ts[1].left = nullptr;
        ts[1].right = nullptr;
        ts[1].lane_in = nullptr;
        ts[1].lane_out = nullptr;
    }
    sweep_tile_pk<R, MODE, PROF>(p, ts, lds);
}

template <int R, int MODE, bool PROF>
__global__ __launch_bounds__(64 * kSweepWaves) void k_sweep_pk(SweepParams p) {
    __shared__ SweepLds lds[kSweepWaves];
    sweep_main_pk<R, MODE, PROF>(p, lds[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
}

template <int R, int MODE, bool PROF>
__global__ __launch_bounds__(64 * kSweepWaves) void k_sweep(SweepParams p) {
    __shared__ SweepLds lds[kSweepWaves];
    sweep_main<R, MODE, PROF>(p, lds[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
}

// Row argmax of job blockIdx.y over blockIdx.x's share of its rows: the largest score,
// then the smallest row, packed as score << 32 | ~row.
__device__ void argmax_rows_main(const ArgmaxJob& j) {
    unsigned long long best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < j.m; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k =
            ((unsigned long long)(uint32_t)j.rowmax[i] << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)i);
        best = k > best ? k : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best) atomicMax(j.key, best);
}

// Bytes present in a set of byte ranges: bit x of mask[x >> 5] is set if byte x occurs.
// Each workgroup gathers its ranges' bytes in LDS (an atomic only for a byte not yet seen)
// and ors its words into mask once.
__global__ __launch_bounds__(256) void k_alphabet(const ByteRange* ranges, int64_t nranges, uint32_t* mask) {
    __shared__ uint32_t seen[8];
    if (threadIdx.x < 8) seen[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t q = blockIdx.x; q < nranges; q += gridDim.x) {
        const ByteRange r = ranges[q];
        for (int64_t i = threadIdx.x; i < r.n; i += blockDim.x) {
            const uint32_t x = r.p[i];
            const uint32_t bit = 1u << (x & 31u);
            if (!(__hip_atomic_load(&seen[x >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit))
                atomicOr(&seen[x >> 5], bit);
        }
    }
    __syncthreads();
    if (threadIdx.x < 8 && seen[threadIdx.x]) atomicOr(&mask[threadIdx.x], seen[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_argmax_rows(const ArgmaxJob* jobs) {
    const ArgmaxJob j = jobs[blockIdx.y];
    argmax_rows_main(j);
}

// seg = rowseg[end row] of each job whose maximum is positive (one thread per job).
__global__ __launch_bounds__(256) void k_end_segment(const ArgmaxJob* jobs, int njobs) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (x >= njobs) return;
    const ArgmaxJob j = jobs[x];
    if (!j.rowseg || !j.seg) return;
    const unsigned long long key = *j.key;
    if (!key) return;
    const int64_t row = (int64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu));
    if (row >= 0 && row < j.m) *j.seg = j.rowseg[row];
}

// Greedy backward walk over the regions of one traceback round (see WalkArgs).  The codes
// of a tile are stored step by step ([step][lane], one 16-bit word per lane = 4 rows), so
// a window of kWalkSteps consecutive steps of one tile is one contiguous, coalesced read:
// it covers the tile's whole stripe of 64R rows along a band of anti-diagonals, which the
// path (moving up and left, so to lower steps) leaves only by its top, its left edge or
// the window's first step.  Windows are staged in LDS with the sequence bytes they need,
// by all kWalkThreads threads with every load in flight at once (the codes were written
// by the previous launch and are read from HBM); wave 0 walks.  The walk stops at the start of the alignment or at a cell outside every
// region of the round.
// Two window sizes: kWalkSteps for few walks at a time (each walk's latency counts: fewer,
// larger windows), kWalkStepsMany when more walks are in flight than the CUs' LDS can host
// at that size (smaller windows let more walks share a CU; many walks at once, 65,536-pair
// batch with 512-row blocks: 64 steps 147 ms, 128 152 ms, 256 164 ms of traceback).
constexpr int kWalkSteps = 256;
constexpr int kWalkStepsMany = 64;
constexpr int kWalkThreads = 256;
template <int WS>
constexpr int walk_loads() { return WS * 64 * 2 / 16 / kWalkThreads; }  // uint4 loads per thread per window
// Ops of one window: the path stays in the window's 64R rows and WS + 63 columns.
template <int WS>
constexpr int walk_ops() { return 64 * kRowsPerLane + WS + 64; }

template <int R, int WS>
struct WalkLds {
    uint16_t win[WS * 64];                         // steps [t0, t0 + WS) x 64 lanes
    uint8_t a[64 * R], b[WS + 64];                 // rows of the stripe, columns of the band
    uint8_t ops[walk_ops<WS>()];                   // ops of the window, flushed when it ends
    XPoint xp;                                     // crosspoint of the window (at most one)
    int32_t nw, nxw;                               // ops and crosspoints of the window
    int64_t ops_at;                                // where the window's ops go in a.ops
    WalkState sst;
};

// Wave-uniform copies (lane 0's value, in scalar registers).
__device__ __forceinline__ int uni(int v) { return (int)__builtin_amdgcn_readfirstlane((uint32_t)v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Region of the round holding cell (i, j), or nullptr.
__device__ __forceinline__ const WalkRegion* walk_region(const WalkArgs& a, int64_t i, int64_t j) {
    if (i < 0 || j < 0) return nullptr;
    const int64_t da = a.brow - i / a.K, db = a.bcol - j / a.Kc;
    if (da < 0 || db < 0 || da > a.L || db > a.L) return nullptr;
    const WalkRegion* r = a.regs + da * (a.L + 1) + db;
    if (r->rows <= 0 || i < r->r0 || j < r->c0 || i >= r->r0 + r->rows || j >= r->c0 + r->cols) return nullptr;
    return r;
}

template <int R, int WS>
__device__ void walk_main(const WalkArgs& a, WalkLds<R, WS>& lds) {
    constexpr int HR = 64 * R;
    WalkState& sst = lds.sst;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    if (tid == 0) sst = *a.st;
    __syncthreads();
    while (true) {
        WalkState st = sst;
        st.i = uni64(st.i);  // every lane holds the same state; keep it in scalar registers
        st.j = uni64(st.j);
        st.done = uni(st.done);
        if (st.done) break;
        const WalkRegion* rg = walk_region(a, st.i, st.j);
        if (!rg) break;  // outside the round's regions
        const WalkRegion rgn = *rg;
        // tile of the current cell and the window's steps [t0, t1]
        const int64_t il = st.i - rgn.r0, jl = st.j - rgn.c0;
        const int64_t s = il / HR, sg = jl / a.W;
        const int64_t R0 = uni64(rgn.r0 + s * HR), C0 = uni64(rgn.c0 + sg * a.W);  // tile origin (absolute)
        const int ri0 = (int)(st.i - R0), cj0 = (int)(st.j - C0);
        const int t1 = cj0 + ri0 / R;
        const int t0 = t1 >= WS ? t1 - WS + 1 : 0;
        const uint16_t* tile = rgn.dirs + (s * a.G + sg) * a.tstride * 64;
        {
            // steps [t0, t1] are contiguous: (t1 - t0 + 1) * 64 codes, at most
            // walk_loads<WS>() * kWalkThreads uint4; all loads are issued before the first store
            const uint4* src = (const uint4*)(tile + (int64_t)t0 * 64);
            uint4* dst = (uint4*)lds.win;
            const int n16 = (t1 - t0 + 1) * 64 * 2 / 16;
            constexpr int kLoads = walk_loads<WS>();
            uint4 v[kLoads];
#pragma unroll
            for (int q = 0; q < kLoads; ++q) {
                const int k = tid + q * kWalkThreads;
                if (k < n16) v[q] = src[k];
            }
#pragma unroll
            for (int q = 0; q < kLoads; ++q) {
                const int k = tid + q * kWalkThreads;
                if (k < n16) dst[k] = v[q];
            }
        }
        const int64_t rows_here = min((int64_t)HR, rgn.r0 + rgn.rows - R0);
        for (int k = tid; k < HR; k += kWalkThreads) lds.a[k] = k < rows_here ? a.A[R0 + k] : 0;
        // columns of the band: col = t - l for t in [t0, t1], l in [0, 64): [t0 - 63, t1].  The
        // walk only moves left, so columns right of the current one (cj0; C0 + cj0 < n) are
        // never read and not loaded: B may end at the current cell's column.
        const int cj_hi = min(t1, cj0);
        for (int k = tid; k < WS + 64; k += kWalkThreads) {
            const int cj = t0 - 63 + k;
            lds.b[k] = (cj >= 0 && cj <= cj_hi) ? a.B[C0 + cj] : 0;
        }
        __syncthreads();
        if (tid < 64) {
            // The whole wave walks in lockstep on wave-uniform values: each step's code (and
            // whether the pair on the diagonal matches) is read once from LDS and broadcast
            // with readfirstlane, so the loop's control flow is scalar.  Lane 0 alone writes
            // the ops and the crosspoint, to LDS; they are flushed after the window, so the
            // loop issues no global memory operation.
            static_assert(64 * R + WS + 64 <= walk_ops<WS>(), "window ops buffer");
            int ri = uni(ri0), cj = uni(cj0);  // tile-local row and column
            int layer = uni(st.layer);
            int v = uni(st.v);  // value of the current layer's matrix at the current cell
            const int sM = a.sc.M, sX = a.sc.X, sO = a.sc.O, sE = a.sc.E;
            int nw = 0, nxw = 0, done = 0;
            int64_t last_row = uni64(st.last_row);
            int64_t si = uni64(st.start_i), sj = uni64(st.start_j);
            int64_t xrow = ((uni64(st.i) + 1) / a.K) * a.K - 1;  // checkpoint row at or above the row
            const int bofs = t0 - 63;
            while (true) {
                if (ri < 0 || cj < 0) {
                    if (R0 + ri < 0 || C0 + cj < 0) done = 1;  // walked past row or column 0
                    break;
                }
                const int l = ri / R, t = cj + l;
                if (t < t0) break;  // left the window
                if (layer == 0) {
                    // Diagonal run: lane k reads the code of cell (ri - k, cj - k) and its H,
                    // the current value less the scores of the cells before it (matches among
                    // the lanes below k: mbcnt of the ballot); the moves up to the first lane
                    // whose cell is not a diagonal move, holds H = 0 (the alignment starts
                    // after it) or lies outside the tile or the window are taken at once, one
                    // op per lane.
                    const int rik = ri - lane, cjk = cj - lane;
                    const int lk = rik >= 0 ? rik / R : 0, tk = cjk + lk;
                    const bool in = rik >= 0 && cjk >= 0 && tk >= t0;
                    uint32_t ck = 1;
                    bool eqk = false;
                    if (in) {
                        ck = (lds.win[(tk - t0) * 64 + lk] >> (4 * (rik % R))) & 1u;
                        eqk = lds.a[rik] == lds.b[cjk - bofs];
                    }
                    const uint64_t eqm = __ballot(eqk);
                    const int mk = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(eqm >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)eqm, 0u));
                    const int vk = v - mk * sM - (lane - mk) * sX;
                    const uint64_t stop = __ballot(!(in && ck == 0u && vk != 0));
                    const int run = stop ? (int)__builtin_ctzll(stop) : 64;
                    if (run > 0) {
                        const int mr = __popcll(run == 64 ? eqm : (eqm & ((UINT64_C(1) << run) - 1)));
                        v -= mr * sM + (run - mr) * sX;
                        if (lane < run) lds.ops[nw + lane] = eqk ? (uint8_t)'=' : (uint8_t)'X';
                        // checkpoint-row bookkeeping of the rows i, i-1, ..., i-run+1, as the
                        // single steps below would do it (xrow <= i holds after row i)
                        const int64_t i = R0 + ri, ilast = i - (run - 1);
                        if (i != last_row) {
                            if (xrow > i) xrow -= a.K;
                            if (i == xrow && nxw == 0) {
                                if (lane == 0) {
                                    lds.xp.row = i;
                                    lds.xp.col = C0 + cj;
                                    lds.xp.layer = 0;
                                }
                                nxw = 1;
                            }
                        }
                        if (run > 1) {
                            if (xrow > i - 1) xrow -= a.K;
                            if (xrow >= ilast && nxw == 0) {
                                if (lane == 0) {
                                    lds.xp.row = xrow;
                                    lds.xp.col = C0 + cj - (i - xrow);
                                    lds.xp.layer = 0;
                                }
                                nxw = 1;
                            }
                            if (xrow > ilast) xrow -= a.K;
                        }
                        last_row = ilast;
                        si = ilast;
                        sj = C0 + cj - (run - 1);
                        nw += run;
                        ri -= run;
                        cj -= run;
                        continue;
                    }
                }
                const uint32_t word = lds.win[(t - t0) * 64 + l];
                const uint32_t eq = lds.a[ri] == lds.b[cj - bofs] ? 16u : 0u;
                const uint32_t cv = __builtin_amdgcn_readfirstlane(((word >> (4 * (ri % R))) & 15u) | eq);
                const int64_t i = R0 + ri;
                if (i != last_row) {
                    last_row = i;
                    if (xrow > i) xrow -= a.K;
                    if (i == xrow && nxw == 0) {
                        if (lane == 0) {
                            lds.xp.row = i;
                            lds.xp.col = C0 + cj;
                            lds.xp.layer = layer;
                        }
                        nxw = 1;
                    }
                }
                uint32_t op;
                if (layer == 0) {
                    if (v == 0) {  // H = 0: the alignment starts at the last diagonal move
                        done = 1;
                        break;
                    }
                    if (!(cv & 1u)) {  // diagonal: H(i-1, j-1) = H - s(i, j)
                        const bool m = (cv & 16u) != 0;
                        op = m ? (uint32_t)'=' : (uint32_t)'X';
                        v -= m ? sM : sX;
                        si = i;
                        sj = C0 + cj;
                        --ri;
                        --cj;
                    } else if (!(cv & 2u)) {  // H = F(i, j)
                        layer = 2;
                        continue;
                    } else {  // H = E(i, j)
                        layer = 1;
                        continue;
                    }
                } else if (layer == 2) {
                    // F(i, j) = F(i-1, j) + gap_extend, or H(i-1, j) + gap_open
                    op = 'D';
                    if ((cv >> 3) & 1u) {
                        v -= sE;
                    } else {
                        v -= sO;
                        layer = 0;
                    }
                    --ri;
                } else {
                    op = 'I';
                    if ((cv >> 2) & 1u) {
                        v -= sE;
                    } else {
                        v -= sO;
                        layer = 0;
                    }
                    --cj;
                }
                if (lane == 0) lds.ops[nw] = (uint8_t)op;
                ++nw;
            }
            if (lane == 0) {
                lds.nw = nw;
                lds.nxw = nxw;
                lds.ops_at = st.nops;
                st.i = R0 + ri;
                st.j = C0 + cj;
                st.layer = layer;
                st.v = v;
                st.nops += nw;
                st.last_row = last_row;
                st.start_i = si;
                st.start_j = sj;
                st.done = done;
                sst = st;
            }
        }
        __syncthreads();
        {
            const int nw = lds.nw;
            uint8_t* dst = a.ops + lds.ops_at;
            for (int k = tid; k < nw; k += kWalkThreads) dst[k] = lds.ops[k];
            if (tid == 0 && lds.nxw && a.srow && sst.nxp < a.xcap) {
                XPoint xp = lds.xp;
                const int2 v = a.srow[((xp.row + 1) / a.K - 1) * a.n + xp.col];
                xp.score = xp.layer == 2 ? v.y : v.x;
                a.xps[sst.nxp] = xp;
                sst.nxp += 1;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) *a.st = sst;
}

template <int R, int WS>
__global__ __launch_bounds__(kWalkThreads) void k_walk(const WalkArgs* jobs) {
    __shared__ WalkLds<R, WS> lds;
    const WalkArgs a = jobs[blockIdx.x];
    if (!a.st) return;  // an empty slot of a round planned on the device
    walk_main<R, WS>(a, lds);
}

// Run-length CIGAR, one wave per alignment (see CigarJob).  Lane l of a 64-op chunk reads the
// op at reading position r = r0 + l (ops index nops - 1 - r); a run starts where the op
// differs from the previous one (the previous lane's, or the chunk before's last for lane 0).
// Each run start after the first closes the run before it, whose length is the distance to
// the previous start (the nearest start below the lane in the chunk, or the open run carried
// from earlier chunks) and whose op is the previous position's: "<length><op>".  The last run
// closes at nops.  The write pass places each closing lane's text after the texts of the
// closing lanes below it (a wave prefix sum of the text lengths).
constexpr int kCigarWaves = 4;

__device__ __forceinline__ int dec_digits(int64_t v) {
    int d = 1;
    while (v >= 10) {
        v /= 10;
        ++d;
    }
    return d;
}

__device__ __forceinline__ void put_run(char* dst, int64_t len, uint32_t op, int nd) {
    for (int k = nd - 1; k >= 0; --k) {
        dst[k] = (char)('0' + len % 10);
        len /= 10;
    }
    dst[nd] = (char)op;
}

template <bool WRITE>
__global__ __launch_bounds__(64 * kCigarWaves) void k_cigar(CigarJob* jobs, int64_t njobs, char* out) {
    const int lane = wave_lane();
    const int64_t x = (int64_t)blockIdx.x * kCigarWaves + (threadIdx.x >> 6);
    if (x >= njobs) return;
    const uint8_t* ops = jobs[x].ops;
    const int64_t nops = jobs[x].nops;
    char* dst = WRITE ? out + jobs[x].at : nullptr;
    int64_t run0 = 0;           // reading position where the open run started
    uint32_t last = 0;          // op at the last position of the previous chunk
    int64_t written = 0;        // characters of the job's text so far
    for (int64_t r0 = 0; r0 < nops; r0 += 64) {
        const int64_t r = r0 + lane;
        const bool in = r < nops;
        const uint32_t op = in ? (uint32_t)ops[nops - 1 - r] : 0u;
        uint32_t prev = (uint32_t)__shfl_up((int)op, 1);
        if (lane == 0) prev = last;
        const bool start = in && r > 0 && op != prev;  // closes the run before it
        const uint64_t sm = __ballot(start);
        // the nearest start below this lane in the chunk, else the carried one
        const uint64_t below = sm & ((UINT64_C(1) << lane) - 1);
        const int64_t pstart = below ? r0 + 63 - __builtin_clzll(below) : run0;
        int nd = 0;
        int64_t len = 0;
        if (start) {
            len = r - pstart;
            nd = dec_digits(len);
        }
        int chars = start ? nd + 1 : 0;
        // inclusive prefix sum of chars over the wave
        int incl = chars;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (WRITE && start) put_run(dst + written + (incl - chars), len, prev, nd);
        written += __shfl(incl, 63);
        if (sm) run0 = r0 + 63 - __builtin_clzll(sm);
        last = (uint32_t)__shfl((int)op, (int)min<int64_t>(63, nops - 1 - r0));
    }
    if (nops > 0) {  // the last run ends at nops
        const int64_t len = nops - run0;
        const int nd = dec_digits(len);
        if (WRITE && lane == 0) put_run(dst + written, len, last, nd);
        written += nd + 1;
    }
    if (!WRITE && lane == 0) jobs[x].len = written;
}

// One thread per round slot: the checkpoint block of the pair's current cell, cut at the
// cell's row and, in columns, at the wider of this slot's and its partner slot's cells
// (slots 2y and 2y + 1 are the two tiles of one packed wave, which sweeps fastest when both
// have one width; the walk never visits the columns right of the cell).  Sweeping only up to
// the cells, instead of whole tiles, halves the columns of a block entered from below.
__global__ __launch_bounds__(256) void k_tb_plan(TbRoundArgs a) {
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x == 0) *a.nact_out = 0;
    if (x >= a.nslots) return;
    Region& r = a.regions[x];
    WalkArgs& wa = a.wargs[x];
    if (x >= *a.nact_in) {
        Region q{};
        q.r1 = -1;  // no rows or columns: the sweep finds no tile in it
        q.c1 = -1;
        q.W = a.W;
        q.G = a.G;
        q.tstride = a.tstride;
        r = q;
        wa.st = nullptr;
        return;
    }
    const int32_t k = a.act_in[x];
    const TbPair P = a.pairs[k];
    WalkState* stp = a.states + k;
    const int64_t i = stp->i, j = stp->j;
    const int64_t brow = i / P.K, bcol = j / P.Kc;
    const int64_t r0 = brow * P.K, c0 = bcol * P.Kc;
    const int64_t c_end = min(P.n, c0 + P.Kc) - 1;
    const int64_t cols = j - c0 + 1;
    int64_t want = cols;
    const int64_t xp = x ^ 1;
    if (xp < *a.nact_in) {
        const int32_t kp = a.act_in[xp];
        const int64_t jp = a.states[kp].j, Kcp = a.pairs[kp].Kc;
        want = max(want, jp - jp / Kcp * Kcp + 1);
    }
    uint16_t* dirs = a.dirs + x * a.dirs_stride;
    Region q{};
    q.A = P.A;
    q.B = P.B;
    q.m = P.m;
    q.n = P.n;
    q.r0 = r0;
    q.r1 = i;
    q.c0 = c0;
    q.c1 = min(c_end, c0 + want - 1);
    q.W = a.W;
    q.G = a.G;
    q.tstride = a.tstride;
    q.top = r0 > 0 && P.srow ? P.srow + (r0 / P.K - 1) * P.n + c0 : nullptr;
    q.left = c0 > 0 && P.scol ? P.scol + (c0 / P.Kc - 1) * P.m : nullptr;
    q.rb = a.rb + x * a.rb_stride;
    q.cb = a.cb + x * a.cb_stride;
    q.cor = a.cor + x * a.cor_stride;
    q.K = P.K;
    q.Kc = P.Kc;
    q.dirs = dirs;
    r = q;
    WalkRegion w;
    w.r0 = r0;
    w.c0 = c0;
    w.rows = i - r0 + 1;
    w.cols = cols;
    w.dirs = dirs;
    a.wregs[x] = w;
    WalkArgs g{};
    g.A = P.A;
    g.B = P.B;
    g.regs = a.wregs + x;
    g.L = 0;
    g.brow = brow;
    g.bcol = bcol;
    g.tstride = a.tstride;
    g.G = a.G;
    g.W = a.W;
    g.K = P.K;
    g.Kc = P.Kc;
    g.srow = P.srow;
    g.n = P.n;
    g.st = stp;
    g.ops = P.ops;
    g.xps = P.xps;
    g.xcap = P.xcap;
    g.sc = a.sc;
    wa = g;
    TbPre pv;
    pv.i = i;
    pv.j = j;
    pv.layer = stp->layer;
    pv._pad = 0;
    a.pre[x] = pv;
}

// One thread per round slot, after the walk: slots whose pair still walks append it to
// act_out, one atomic per wave.
__global__ __launch_bounds__(256) void k_tb_compact(TbRoundArgs a) {
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool keep = false;
    int32_t k = -1;
    if (x < a.nslots && x < *a.nact_in) {
        k = a.act_in[x];
        WalkState* st = a.states + k;
        if (!st->done) {
            const TbPre pv = a.pre[x];
            if (st->i == pv.i && st->j == pv.j && st->layer == pv.layer) st->done = -1;  // no progress
            else keep = true;
        }
    }
    const uint64_t b = __ballot(keep);
    if (!b) return;
    const int lane = (int)(threadIdx.x & 63u);
    int base = 0;
    if (lane == 0) base = atomicAdd(a.nact_out, (int32_t)__popcll(b));
    base = __shfl(base, 0);
    if (keep) a.act_out[base + __popcll(b & ((UINT64_C(1) << lane) - 1))] = k;
}

}  // namespace

hipError_t launch_sweep(int mode, const SweepParams& p, size_t lds_pad, hipStream_t s) {
    constexpr int R = kRowsPerLane;
    const dim3 grid((unsigned)((p.ntiles + kSweepWaves - 1) / kSweepWaves)), block(64 * kSweepWaves);
    const bool prof = p.code_shift1 != 0;
    switch (mode * 2 + (prof ? 1 : 0)) {
        case kStage1 * 2: hipLaunchKernelGGL((k_sweep<R, kStage1, false>), grid, block, lds_pad, s, p); break;
        case kDirs * 2: hipLaunchKernelGGL((k_sweep<R, kDirs, false>), grid, block, lds_pad, s, p); break;
        case kSearch * 2: hipLaunchKernelGGL((k_sweep<R, kSearch, false>), grid, block, lds_pad, s, p); break;
        case kStage1 * 2 + 1: hipLaunchKernelGGL((k_sweep<R, kStage1, true>), grid, block, lds_pad, s, p); break;
        case kDirs * 2 + 1: hipLaunchKernelGGL((k_sweep<R, kDirs, true>), grid, block, lds_pad, s, p); break;
        case kSearch * 2 + 1: hipLaunchKernelGGL((k_sweep<R, kSearch, true>), grid, block, lds_pad, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sweep_pk(int mode, const SweepParams& p, size_t lds_pad, hipStream_t s) {
    const int64_t pairs = (p.ntiles + 1) / 2;
    const dim3 grid((unsigned)((pairs + kSweepWaves - 1) / kSweepWaves)), block(64 * kSweepWaves);
    constexpr int R = kRowsPerLane;
    const bool prof = p.code2_shift1 != 0;
    if (mode == kStage1 && prof) hipLaunchKernelGGL((k_sweep_pk<R, kStage1, true>), grid, block, lds_pad, s, p);
    else if (mode == kStage1) hipLaunchKernelGGL((k_sweep_pk<R, kStage1, false>), grid, block, lds_pad, s, p);
    else if (mode == kDirs && prof) hipLaunchKernelGGL((k_sweep_pk<R, kDirs, true>), grid, block, lds_pad, s, p);
    else if (mode == kDirs) hipLaunchKernelGGL((k_sweep_pk<R, kDirs, false>), grid, block, lds_pad, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

size_t sweep_static_lds() { return sizeof(SweepLds) * kSweepWaves; }
size_t walk_lds_bytes() { return sizeof(WalkLds<kRowsPerLane, kWalkSteps>); }

hipError_t sweep_allow_lds(size_t dyn_bytes) {
    constexpr int R = kRowsPerLane;
    const int v = (int)dyn_bytes;
    const void* fns[] = {(const void*)k_sweep<R, kStage1, false>, (const void*)k_sweep<R, kDirs, false>,
                         (const void*)k_sweep<R, kSearch, false>, (const void*)k_sweep<R, kStage1, true>,
                         (const void*)k_sweep<R, kDirs, true>,    (const void*)k_sweep<R, kSearch, true>};
    hipError_t e = hipSuccess;
    for (const void* f : fns)
        if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_sweep_pk<R, kStage1, false>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_sweep_pk<R, kStage1, true>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_sweep_pk<R, kDirs, false>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_sweep_pk<R, kDirs, true>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    return e;
}

hipError_t launch_argmax_rows(const ArgmaxJob* jobs, int njobs, int blocks_per_job, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_argmax_rows, dim3((unsigned)(blocks_per_job > 1 ? blocks_per_job : 1), (unsigned)njobs), dim3(256), 0, s,
                       jobs);
    return hipGetLastError();
}

hipError_t launch_end_segment(const ArgmaxJob* jobs, int njobs, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_end_segment, dim3((unsigned)((njobs + 255) / 256)), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_alphabet(const ByteRange* ranges, int64_t nranges, uint32_t* mask, hipStream_t s) {
    if (nranges <= 0) return hipSuccess;
    const unsigned grid = (unsigned)(nranges < 4096 ? nranges : 4096);
    hipLaunchKernelGGL(k_alphabet, dim3(grid), dim3(256), 0, s, ranges, nranges, mask);
    return hipGetLastError();
}

hipError_t launch_cigar(CigarJob* jobs, int64_t njobs, char* out, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((njobs + kCigarWaves - 1) / kCigarWaves);
    if (out) hipLaunchKernelGGL(k_cigar<true>, dim3(blocks), dim3(64 * kCigarWaves), 0, s, jobs, njobs, out);
    else hipLaunchKernelGGL(k_cigar<false>, dim3(blocks), dim3(64 * kCigarWaves), 0, s, jobs, njobs, out);
    return hipGetLastError();
}

hipError_t launch_walk(const WalkArgs* jobs, int njobs, bool many, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    if (many)
        hipLaunchKernelGGL((k_walk<kRowsPerLane, kWalkStepsMany>), dim3((unsigned)njobs), dim3(kWalkThreads), 0, s,
                           jobs);
    else
        hipLaunchKernelGGL((k_walk<kRowsPerLane, kWalkSteps>), dim3((unsigned)njobs), dim3(kWalkThreads), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_tb_plan(const TbRoundArgs& a, hipStream_t s) {
    if (a.nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tb_plan, dim3((unsigned)((a.nslots + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tb_compact(const TbRoundArgs& a, hipStream_t s) {
    if (a.nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tb_compact, dim3((unsigned)((a.nslots + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace swk
