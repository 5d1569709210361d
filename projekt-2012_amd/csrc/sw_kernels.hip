// HIP kernels (gfx950) for the SWalign hot path.
//
// Forward sweep.  The DP matrix (rows = A, columns = B) is cut into stripes of
// kStripeRows = 64*R rows; one 64-lane wavefront owns a stripe and sweeps it left to
// right.  Lane l holds rows [l*R, l*R+R) of the stripe and, at step t, updates column
// t-l: the stripe advances along anti-diagonals, so lane l needs the bottom-row {H,F}
// of lane l-1 from the previous step, which it receives with one DPP wave_shr:1 per
// value.  Lane 0 receives the stripe's top boundary (the previous stripe's bottom row)
// from LDS; lane 63's bottom row is staged in LDS and written once per 64 steps with a
// coalesced store.  Stripes run as a persistent grid: a wave takes the next stripe by
// ticket, and waits on the predecessor stripe's published column count before it reads
// the boundary row -- so a stripe only ever waits on a stripe that is already running.
//
// Cell update (Sterbic/Projekt-2012 src/FindAlignment.cuh:50-56, zero boundaries as in
// src/SWutils.cpp:131-136), kept in "plus open" form Ho = H + O so that one add serves
// both gap recurrences:
//     E  = max(E + ext, Ho_left)               F = max(F_up + ext, Ho_up)
//     H  = max(0, Hdiag + s, E, F)             Ho = H + O
// with Hdiag + s computed as Ho_diag + (s - O).
#include "sw_internal.h"

namespace swk {

namespace {

#ifndef UNROLL_STEPS
#define UNROLL_STEPS 4
#endif
#define SWK_DO_PRAGMA(x) _Pragma(#x)
#define SWK_UNROLL(n) SWK_DO_PRAGMA(unroll n)

constexpr int kDppWaveShr1 = 0x138;  // lane l <- lane l-1, lane 0 keeps `old`

__device__ __forceinline__ int shr1(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShr1, 0xF, 0xF, false);
}

// Copies a wave-uniform value into a VGPR the compiler cannot see through (a DPP
// row_mirror of a uniform value is the value itself), so that per-step selects keep
// both operands in registers instead of re-materialising them every step.
constexpr int kDppRowMirror = 0x140;
__device__ __forceinline__ int to_vgpr(int v) {
    return __builtin_amdgcn_mov_dpp(v, kDppRowMirror, 0xF, 0xF, false);
}

__device__ __forceinline__ uint64_t pack2(int x, int y) {
    return (uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 32);
}

// Agent-coherent 64-bit load/store for the stripe-to-stripe boundary row (the producer
// and consumer waves may sit on different XCDs).
__device__ __forceinline__ int2 load_coherent(const int2* p) {
    const uint64_t v = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_int2((int)(uint32_t)v, (int)(uint32_t)(v >> 32));
}
__device__ __forceinline__ void store_coherent(int2* p, int x, int y) {
    __hip_atomic_store((uint64_t*)p, pack2(x, y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform reads of the cross-wave flags: every lane takes lane 0's value, so all
// lanes leave a wait loop on the same iteration.
__device__ __forceinline__ int32_t load_flag_uniform(const int32_t* q) {
    return (int32_t)__builtin_amdgcn_readfirstlane(
        (uint32_t)__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ bool sweep_failed(const SweepParams& p) { return load_flag_uniform(p.err) != 0; }

// Waits until stripe `sl-1` of tile T has published at least `need` columns.  The wait
// is bounded: after p.spin_limit polls (or once another wave has failed) it records the
// failure and returns false.  Every lane reads the same values, so the outcome is
// uniform across the wave.
__device__ __forceinline__ bool wait_prog(const SweepParams& p, const Tile& T, int tile_idx, int64_t sl,
                                          int64_t need) {
    if (sl == 0) return true;
    const int32_t* q = T.prog + (sl - 1);
    uint32_t spins = 0;
    int32_t seen;
    while ((seen = load_flag_uniform(q)) < need) {
        if (++spins > p.spin_limit || sweep_failed(p)) {
            if (threadIdx.x == 0) {
                int32_t expected = 0;
                if (__hip_atomic_compare_exchange_strong(p.err, &expected, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                    p.err[1] = tile_idx;
                    p.err[2] = (int32_t)sl;
                    p.err[3] = (int32_t)need;
                    p.err[4] = seen;
                }
            }
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return true;
}

// Publishes `cols` columns of stripe sl after all of this wave's boundary stores have
// completed (vmcnt(0)): data stores are agent-coherent, so completion makes them
// visible to the consumer's agent-coherent loads.
__device__ __forceinline__ void publish(const Tile& T, int64_t sl, int64_t cols) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) __hip_atomic_store(T.prog + sl, (int32_t)cols, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads the inputs of chunk cc (steps [64cc, 64cc+64)): lane 0's top boundary for
// columns 64cc + lane, and the column bytes 64cc-64+lane and 64cc+lane (0x100 pads
// never equal a byte).
__device__ __forceinline__ void load_chunk_inputs(const SweepParams& p, const Tile& T, int64_t sl,
                                                  int64_t cc, int64_t w, int2& in, uint32_t& b0,
                                                  uint32_t& b1) {
    const int lane = threadIdx.x;
    const int64_t col = 64 * cc + lane;
    int H = 0, F = 0;
    if (col < w) {
        if (sl == 0) {
            if (T.top) {
                const int2 v = T.top[col];
                H = v.x;
                F = v.y;
            }
        } else {
            const int2 v = load_coherent(T.bnd + col);
            H = v.x;
            F = v.y;
        }
    }
    in = make_int2(H + p.sc.O, F + p.sc.E);
    const int64_t cb0 = 64 * cc - 64 + lane, cb1 = 64 * cc + lane;
    b0 = (cb0 >= 0 && cb0 < w) ? (uint32_t)p.B[T.c0 + cb0] : 0x100u;
    b1 = (cb1 < w) ? (uint32_t)p.B[T.c0 + cb1] : 0x100u;
}

// 64 steps of one stripe.  EDGE: some (lane, step) pairs fall outside [0, w) and must
// leave the lane state untouched.  DUMP: stage-1 chunk containing a checkpoint column.
template <int R, int MODE, bool EDGE, bool DUMP>
__device__ __forceinline__ void chunk_steps(const SweepParams& p, const Tile& T, int64_t sl, int64_t c,
                                            int64_t w, int64_t row0, const uint32_t (&a)[R],
                                            int (&Ho)[R], int (&E)[R], int (&rmax)[R], int& Ho_up_prev,
                                            int& Ho_bot, int& Fe_bot, const int2* s_in,
                                            const uint32_t* s_b, int2* s_out, int& found_col) {
    const int lane = threadIdx.x;
    const int O = p.sc.O, Ev = p.sc.E;
    const int MO = to_vgpr(p.sc.M - O), XO = to_vgpr(p.sc.X - O);
    uint16_t* dcol = nullptr;
    if (MODE == kDirs) dcol = T.dirs + ((sl * T.tstride) + 64 * c) * 64 + lane;
    SWK_UNROLL(UNROLL_STEPS)
    for (int s = 0; s < 64; ++s) {
        const int2 in = s_in[s];
        const uint32_t b = s_b[64 + s - lane];
        int Ho_up = shr1(in.x, Ho_bot);
        int Fe_up = shr1(in.y, Fe_bot);
        bool valid = true;
        int64_t col = 0;
        if (EDGE || MODE == kSearch) {
            col = 64 * c + s - lane;
            valid = col >= 0 && col < w;
        }
        int diag = Ho_up_prev;
        if (EDGE) Ho_up_prev = valid ? Ho_up : Ho_up_prev;
        else Ho_up_prev = Ho_up;
        uint32_t nib = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int sp = (a[k] == b) ? MO : XO;
            const int Hd = diag + sp;  // H(i-1,j-1) + s(i,j)
            diag = Ho[k];              // H(i,j-1) + O: diagonal of the next row
            const int Ee = E[k] + Ev;
            const int En = max(Ee, Ho[k]);
            const int F = max(Fe_up, Ho_up);
            const int H = max(max(Hd, 0), max(En, F));
            if (MODE == kDirs) {
                const uint32_t h = (H == 0) ? 0u : (H == Hd ? 1u : (H == F ? 2u : 3u));
                nib |= (h | ((uint32_t)(Ee > Ho[k]) << 2) | ((uint32_t)(Fe_up > Ho_up) << 3)) << (4 * k);
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
                if (valid && row0 + k == p.target_row && H == p.target && found_col < 0) found_col = (int)col;
            }
            Ho_up = Hon;
            Fe_up = F + Ev;
        }
        Ho_bot = Ho_up;
        Fe_bot = Fe_up;
        if (lane == 63) s_out[s] = make_int2(Ho_bot, Fe_bot);
        if (MODE == kDirs) dcol[s * 64] = (uint16_t)nib;
        if (DUMP) {
            const int64_t x = 64 * c + s - lane;
            if (x >= 0 && x < w && ((x + 1) & (p.Kc - 1)) == 0) {
                int2* dst = p.scol + ((x + 1) / p.Kc - 1) * p.m;
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (row0 + k < p.m) dst[row0 + k] = make_int2(Ho[k] - O, E[k]);
            }
        }
    }
}

// Per-wave LDS of the sweep: lane 0's top-boundary inputs and the column bytes of the
// current and next chunk, and lane 63's bottom-row outputs of the current chunk.
struct SweepLds {
    int2 in[2][64];
    uint32_t b[2][128];
    int2 out[64];
};

template <int R, int MODE>
__device__ void process_stripe(const SweepParams& p, const Tile& T, int tile_idx, int64_t sl, SweepLds& lds) {
    int2 (*s_in)[64] = lds.in;
    uint32_t (*s_b)[128] = lds.b;
    int2* s_out = lds.out;
    const int lane = threadIdx.x;
    const int O = p.sc.O;
    constexpr int64_t HR = 64 * R;
    const int64_t w = T.c1 - T.c0 + 1;
    const int64_t row0 = T.r0 + sl * HR + (int64_t)lane * R;

    uint32_t a[R];
    int Ho[R], E[R], rmax[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int64_t row = row0 + k;
        const bool in = row <= T.r1;
        a[k] = in ? (uint32_t)p.A[row] : 0x200u;
        if (T.left && in) {
            const int2 v = T.left[row];
            Ho[k] = v.x + O;
            E[k] = v.y;
        } else {
            Ho[k] = O;
            E[k] = 0;
        }
        rmax[k] = 0;
    }
    int Ho_up_prev = O;  // H(row0-1, c0-1) + O
    if (T.left && row0 - 1 >= 0 && row0 - 1 < p.m) Ho_up_prev = T.left[row0 - 1].x + O;
    int Ho_bot = O, Fe_bot = p.sc.E;
    int found_col = -1;

    const int64_t nsteps = w + 63;
    const int64_t nchunks = (nsteps + 63) / 64;
    const int64_t last_interior = (w - 64) / 64;
    int64_t srow_idx = -1;
    if (MODE == kStage1 && p.srow) {
        const int64_t brow = T.r0 + sl * HR + HR - 1;
        if (brow < p.m && (brow + 1) % p.K == 0) srow_idx = (brow + 1) / p.K - 1;
    }

    if (!wait_prog(p, T, tile_idx, sl, min((int64_t)64, w))) return;
    int2 in_reg;
    uint32_t b0, b1;
    load_chunk_inputs(p, T, sl, 0, w, in_reg, b0, b1);
    s_in[0][lane] = in_reg;
    s_b[0][lane] = b0;
    s_b[0][64 + lane] = b1;
    // LDS handoffs between lanes of the wave (here, after lane 63's last output of a
    // chunk and after the next chunk's inputs) need no hardware barrier, the wave runs
    // in lockstep, but must not be reordered by the compiler.
    __builtin_amdgcn_wave_barrier();

    for (int64_t c = 0; c < nchunks; ++c) {
        const int cur = (int)(c & 1);
        if (c + 1 < nchunks) {
            if (!wait_prog(p, T, tile_idx, sl, min(64 * (c + 2), w))) return;
            load_chunk_inputs(p, T, sl, c + 1, w, in_reg, b0, b1);
        }
        const bool interior = c >= 1 && c <= last_interior;
        bool dump = false;
        if (MODE == kStage1 && p.scol) {
            const int64_t lo = max(64 * c - 63, (int64_t)0), hi = min(64 * c + 63, w - 1);
            const int64_t x = ((lo + p.Kc) / p.Kc) * p.Kc - 1;  // smallest x >= lo, x = -1 mod Kc
            dump = x <= hi;
        }
        if (!dump) {
            if (interior)
                chunk_steps<R, MODE, false, false>(p, T, sl, c, w, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot,
                                                   Fe_bot, s_in[cur], s_b[cur], s_out, found_col);
            else
                chunk_steps<R, MODE, true, false>(p, T, sl, c, w, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot,
                                                  Fe_bot, s_in[cur], s_b[cur], s_out, found_col);
        } else {
            chunk_steps<R, MODE, true, true>(p, T, sl, c, w, row0, a, Ho, E, rmax, Ho_up_prev, Ho_bot, Fe_bot,
                                             s_in[cur], s_b[cur], s_out, found_col);
        }
        // Bottom row of the stripe for columns 64c-63+lane.
        __builtin_amdgcn_wave_barrier();
        {
            const int2 o = s_out[lane];
            const int64_t col = 64 * c + lane - 63;
            if (col >= 0 && col < w) {
                const int H = o.x - O, F = o.y - p.sc.E;
                if (sl + 1 < T.nstripes) store_coherent(T.bnd + col, H, F);
                if (MODE == kStage1 && srow_idx >= 0) p.srow[srow_idx * p.n + T.c0 + col] = make_int2(H, F);
            }
            if (sl + 1 < T.nstripes) publish(T, sl, min(64 * c + 1, w));
        }
        if (c + 1 < nchunks) {
            s_in[cur ^ 1][lane] = in_reg;
            s_b[cur ^ 1][lane] = b0;
            s_b[cur ^ 1][64 + lane] = b1;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (sl + 1 < T.nstripes) publish(T, sl, w);
    if (MODE == kStage1) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (row0 + k < p.m) p.rowmax[row0 + k] = rmax[k];
    }
    if (MODE == kSearch) {
        if (found_col >= 0) atomicMin(T.found, (int32_t)(T.c0 + found_col));
    }
}

// Persistent sweep: each wave takes stripes by ticket until none are left.
template <int R, int MODE>
__device__ void sweep_main(const SweepParams& p, SweepLds& lds) {
    while (true) {
        unsigned long long g = 0;
        if (threadIdx.x == 0) g = atomicAdd(p.ticket, 1ull);
        const uint32_t glo = __builtin_amdgcn_readfirstlane((uint32_t)g);
        const uint32_t ghi = __builtin_amdgcn_readfirstlane((uint32_t)(g >> 32));
        const int64_t gi = (int64_t)(((uint64_t)ghi << 32) | glo);
        if (gi >= p.total_stripes || sweep_failed(p)) break;
        int lo = 0, hi = p.ntiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.tiles[mid].stripe0 <= gi) lo = mid;
            else hi = mid - 1;
        }
        const Tile T = p.tiles[lo];
        process_stripe<R, MODE>(p, T, lo, gi - T.stripe0, lds);
    }
}

template <int R, int MODE>
__global__ __launch_bounds__(64) void k_sweep(SweepParams p) {
    __shared__ SweepLds lds;
    sweep_main<R, MODE>(p, lds);
}

__device__ void argmax_rows_main(const int32_t* __restrict__ rowmax, int64_t m, unsigned long long* key) {
    unsigned long long best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        // larger score first, then smaller row
        const unsigned long long k =
            ((unsigned long long)(uint32_t)rowmax[i] << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)i);
        best = k > best ? k : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(key, best);
}

__global__ __launch_bounds__(256) void k_argmax_rows(const int32_t* __restrict__ rowmax, int64_t m,
                                                     unsigned long long* key) {
    argmax_rows_main(rowmax, m, key);
}

// Greedy backward walk over one recomputed region (see WalkArgs).  The region's
// direction codes are staged through LDS in 128x128-cell windows; lane 0 walks.
constexpr int kWalkWinRows = 128, kWalkWinCols = 128;

template <int R>
struct WalkLds {
    uint16_t win[kWalkWinRows / R][kWalkWinCols];
    WalkState sst;
};

template <int R>
__device__ void walk_main(const WalkArgs& a, WalkLds<R>& lds) {
    constexpr int WR = kWalkWinRows, WC = kWalkWinCols, NG = WR / R;
    constexpr int64_t HR = 64 * R;
    auto& win = lds.win;
    WalkState& sst = lds.sst;
    const int lane = threadIdx.x;
    if (lane == 0) sst = *a.st;
    __syncthreads();
    while (true) {
        WalkState st = sst;
        if (st.done) break;
        const int64_t il = st.i - a.r0, jl = st.j - a.c0;
        if (st.i < 0 || st.j < 0 || il < 0 || jl < 0) break;  // left the region
        // rows [wr0, wr0+WR) with wr0 a multiple of R and il inside the window
        int64_t wr0 = il - (WR - R);
        if (wr0 < 0) wr0 = 0;
        wr0 -= wr0 % R;
        int64_t wc0 = jl - WC + 1;
        if (wc0 < 0) wc0 = 0;
        for (int idx = lane; idx < NG * WC; idx += 64) {
            const int g = idx / WC, cc = idx % WC;
            const int64_t row = wr0 + (int64_t)g * R, col = wc0 + cc;
            uint16_t v = 0;
            if (row < a.rows && col < a.cols) {
                const int64_t s = row / HR, l = (row % HR) / R, t = col + l;
                v = a.dirs[(s * a.tstride + t) * 64 + l];
            }
            win[g][cc] = v;
        }
        __syncthreads();
        if (lane == 0) {
            while (true) {
                if (st.i < 0 || st.j < 0) {
                    st.done = 1;
                    break;
                }
                const int64_t ri = st.i - a.r0, rj = st.j - a.c0;
                if (ri < 0 || rj < 0) break;          // left the region
                if (ri < wr0 || rj < wc0) break;      // left the window
                const int64_t gi = ri - wr0;
                const uint32_t d = (win[gi / R][rj - wc0] >> (4 * (gi % R))) & 15u;
                if (st.i != st.last_row) {
                    st.last_row = st.i;
                    if (a.srow && (st.i + 1) % a.K == 0 && st.nxp < a.xcap) {
                        const int2 v = a.srow[((st.i + 1) / a.K - 1) * a.n + st.j];
                        XPoint xp;
                        xp.row = st.i;
                        xp.col = st.j;
                        xp.score = st.layer == 2 ? v.y : v.x;
                        xp.layer = st.layer;
                        a.xps[st.nxp++] = xp;
                    }
                }
                if (st.layer == 0) {
                    const uint32_t h = d & 3u;
                    if (h == 1u) {
                        a.ops[st.nops++] = (a.A[st.i] == a.B[st.j]) ? (uint8_t)'=' : (uint8_t)'X';
                        st.start_i = st.i;
                        st.start_j = st.j;
                        --st.i;
                        --st.j;
                    } else if (h == 2u) {
                        st.layer = 2;
                    } else if (h == 3u) {
                        st.layer = 1;
                    } else {
                        st.done = 1;
                        break;
                    }
                } else if (st.layer == 2) {
                    a.ops[st.nops++] = (uint8_t)'D';
                    st.layer = (d >> 3) & 1u ? 2 : 0;
                    --st.i;
                } else {
                    a.ops[st.nops++] = (uint8_t)'I';
                    st.layer = (d >> 2) & 1u ? 1 : 0;
                    --st.j;
                }
            }
            sst = st;
        }
        __syncthreads();
    }
    __syncthreads();
    if (lane == 0) *a.st = sst;
}

template <int R>
__global__ __launch_bounds__(64) void k_walk(WalkArgs a) {
    __shared__ WalkLds<R> lds;
    walk_main<R>(a, lds);
}

}  // namespace

hipError_t launch_sweep(int mode, const SweepParams& p, int grid, hipStream_t s) {
    constexpr int R = kRowsPerLane;
    switch (mode) {
        case kStage1: hipLaunchKernelGGL((k_sweep<R, kStage1>), dim3(grid), dim3(64), 0, s, p); break;
        case kDirs: hipLaunchKernelGGL((k_sweep<R, kDirs>), dim3(grid), dim3(64), 0, s, p); break;
        case kSearch: hipLaunchKernelGGL((k_sweep<R, kSearch>), dim3(grid), dim3(64), 0, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int sweep_occupancy(int mode) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    constexpr int R = kRowsPerLane;
    switch (mode) {
        case kStage1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<R, kStage1>, 64, 0); break;
        case kDirs: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<R, kDirs>, 64, 0); break;
        case kSearch: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<R, kSearch>, 64, 0); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_argmax_rows(const int32_t* rowmax, int64_t m, unsigned long long* key, hipStream_t s) {
    int64_t blocks = (m + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_argmax_rows, dim3((unsigned)blocks), dim3(256), 0, s, rowmax, m, key);
    return hipGetLastError();
}

hipError_t launch_walk(const WalkArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_walk<kRowsPerLane>, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace swk
