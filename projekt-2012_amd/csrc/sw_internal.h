// Internal types shared by the HIP kernels (sw_kernels.hip) and the host side of the
// C-ABI (sw_api.hip).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swk {

// Rows per lane of the forward sweep; a wave owns a stripe of 64*kRowsPerLane rows.
constexpr int kRowsPerLane = 4;
constexpr int kStripeRows = 64 * kRowsPerLane;

// Sweep modes.
enum SweepMode : int {
    kStage1 = 0,   // whole matrix: row maxima, checkpoint rows and columns
    kDirs = 1,     // block recomputation writing 4-bit direction codes per cell
    kSearch = 2    // block recomputation looking for the first column holding a value
};

struct Score {
    int32_t M, X, O, E;  // match, mismatch, gap open ("first"), gap extension
};

// One independent rectangle of the DP matrix swept by a set of stripes.
// Rows [r0, r1] and columns [c0, c1] are absolute (inclusive).
struct Tile {
    int64_t r0, r1, c0, c1;
    int64_t stripe0;     // first ticket of this tile
    int64_t nstripes;
    int2* bnd;           // (c1-c0+1) {H,F}: bottom row of the last published stripe
    int32_t* prog;       // nstripes: columns of each stripe's bottom row published
    const int2* top;     // {H,F} of row r0-1 at columns c0..c1, nullptr = zero boundary
    const int2* left;    // {H,E} of column c0-1 indexed by absolute row, nullptr = zero
    uint16_t* dirs;      // kDirs: [stripe][step][lane] packed direction codes
    int64_t tstride;     // kDirs: steps stored per stripe
    int32_t* found;      // kSearch: smallest absolute column holding the target value
};

struct SweepParams {
    const uint8_t* A;  // rows
    int64_t m;
    const uint8_t* B;  // columns
    int64_t n;
    Score sc;
    const Tile* tiles;
    int32_t ntiles;
    int32_t _pad;
    int64_t total_stripes;
    unsigned long long* ticket;
    // kStage1 outputs
    int32_t* rowmax;   // m
    int2* srow;        // (m/K) x n {H,F}
    int64_t K;         // multiple of kStripeRows
    int2* scol;        // (n/Kc) x m {H,E}
    int64_t Kc;        // power of two >= 64
    // kSearch
    int64_t target_row;
    int32_t target;
    // Bounded waits: a wave that polls a predecessor's progress more than spin_limit
    // times records {1, tile, stripe, need, seen} in err[0..4] and abandons its work;
    // waves that see err[0] != 0 stop as well.  err must be zeroed before the launch.
    uint32_t spin_limit;
    int32_t* err;
};

constexpr uint32_t kDefaultSpinLimit = 1u << 20;

struct WalkState {
    int64_t i, j;              // current cell (absolute)
    int64_t start_i, start_j;  // first aligned pair seen so far
    int64_t nops;              // ops written (reverse order)
    int64_t nxp;               // crosspoints written
    int64_t last_row;          // last row visited (crosspoint detection)
    int32_t layer;             // 0 = H, 1 = E, 2 = F
    int32_t done;              // 1 when the start of the alignment was reached
};

struct XPoint {
    int64_t row, col;
    int32_t score;
    int32_t layer;
};

struct WalkArgs {
    const uint8_t* A;
    const uint8_t* B;
    const uint16_t* dirs;
    int64_t tstride;
    int64_t r0, c0, rows, cols;  // recomputed region: rows [r0, r0+rows), cols [c0, c0+cols)
    int64_t K;                   // checkpoint row spacing (crosspoint detection)
    const int2* srow;            // checkpoint rows (crosspoint values), may be null
    int64_t n;
    WalkState* st;
    uint8_t* ops;
    XPoint* xps;
    int64_t xcap;
};

// Launchers (sw_kernels.hip).
hipError_t launch_sweep(int mode, const SweepParams& p, int grid, hipStream_t s);
hipError_t launch_argmax_rows(const int32_t* rowmax, int64_t m, unsigned long long* key,
                              hipStream_t s);
hipError_t launch_walk(const WalkArgs& a, hipStream_t s);
int sweep_occupancy(int mode);  // resident 64-thread workgroups per CU

}  // namespace swk
