// Request server (swh_server_*): continuous batching of alignment jobs on one device.
//
// A server thread runs ticks.  Each tick admits queued jobs while the HBM pool has room
// for their working memory, then advances every job in flight by one step of its own
// schedule in shared launches: one sweep launch per mode (stage 1, end search, traceback
// blocks) covering the current tile anti-diagonal of every job in that mode -- jobs
// admitted at different times sit at different diagonals of the same launch -- plus one
// argmax launch for jobs whose stage 1 just ended and one walk launch for jobs whose
// traceback blocks are ready.  One read-back per tick brings the few bytes that decide each
// job's next step; before waiting for it the tick also launches the next tick's diagonals of
// every job whose sweep continues, so that the device sweeps them while the host waits and
// plans.  A job's work therefore overlaps with every other job in flight, and jobs join and
// leave between ticks.  No kernel waits on another workgroup.
//
// Each job's device memory (sequences, row maxima, checkpoints, its sweeps' regions, tile
// hand-off buffers, direction codes, ops) is one block of a pool allocated once at creation;
// admission is bounded by the pool (the HBM budget).
#include <algorithm>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sw_host.h"

using namespace swk;
using namespace swh_host;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) of the device pool (host bookkeeping only).
struct Pool {
    int64_t size = 0;
    std::map<int64_t, int64_t> free_;  // offset -> length
    void init(int64_t n) {
        size = n;
        free_.clear();
        free_[0] = n;
    }
    int64_t alloc(int64_t n) {
        n = ceil_div(n, 256) * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const int64_t off = it->first, rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t n) {
        n = ceil_div(n, 256) * 256;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t in_use() const {
        int64_t f = 0;
        for (auto& kv : free_) f += kv.second;
        return size - f;
    }
};

enum Phase { kQueued, kSweep, kArgmax, kWalk, kFetch, kDone };

// The server's small transfers are batched per tick: uploads (sequences of new jobs, reset
// values) are packed into one pinned buffer, copied with one transfer and scattered by one
// kernel; read-backs (end keys, found columns, walk states, finished jobs' ops) are gathered
// by one kernel into one buffer and copied back with one transfer.  One copy dispatch per
// value cost ~4 us of device time each (tens of thousands per second of small jobs).
struct Xfer {
    uint8_t* dst;
    const uint8_t* src;
    int64_t len;
};
constexpr int64_t kXferChunk = INT64_C(64) << 10;  // bytes per copy workgroup

// One workgroup per entry; 16-byte words when the entry is aligned.
__global__ __launch_bounds__(256) void k_copy_list(const Xfer* x) {
    const Xfer e = x[blockIdx.x];
    if ((((uintptr_t)e.dst | (uintptr_t)e.src | (uintptr_t)e.len) & 15) == 0) {
        const uint4* src = (const uint4*)e.src;
        uint4* dst = (uint4*)e.dst;
        for (int64_t k = threadIdx.x; k < e.len / 16; k += 256) dst[k] = src[k];
    } else {
        for (int64_t k = threadIdx.x; k < e.len; k += 256) e.dst[k] = e.src[k];
    }
}

// A grow-only pinned host buffer.
struct PinBuf {
    uint8_t* h = nullptr;
    size_t cap = 0;
};
hipError_t ensure_pin(PinBuf& b, size_t bytes) {
    if (b.h && b.cap >= bytes) return hipSuccess;
    if (b.h) (void)hipHostFree(b.h);
    b = PinBuf{};
    const hipError_t e = hipHostMalloc((void**)&b.h, std::max<size_t>(bytes, 4096), 0);
    if (e == hipSuccess) b.cap = std::max<size_t>(bytes, 4096);
    return e;
}

int64_t al16(int64_t x) { return (x + 15) & ~INT64_C(15); }

// Device and host copies of one set of sweep launches' descriptors.
struct SweepSet {
    DevBuf spans[3];
    std::vector<TileSpan> hspans[3];
};

// One tick's descriptors: the sweeps of jobs whose diagonals were not launched ahead, and
// the sweeps launched ahead for the next tick.  Ticks that need no read-back are not
// synchronised, so the host runs up to kTickRing ticks ahead of the device; a slot is
// reused once the event recorded after its launches has completed.
struct TickSlot {
    SweepSet now, ahead;
    // this tick's uploads: data, then the copy list (pageable: on this platform small
    // pageable copies delayed the tick's sweeps less than copies from pinned memory)
    std::vector<uint8_t> up_h;
    DevBuf up_d;
    hipEvent_t done = nullptr;
    bool pending = false;
};
constexpr int kTickRing = 4;

struct SJob {
    int64_t id = 0;
    std::vector<uint8_t> a, b;  // host copies, oriented (a = rows)
    uint32_t alpha[8] = {};     // bytes present in a and b (bit x of word x >> 5)
    bool swapped = false;
    int64_t m = 0, n = 0;
    Plan pl{};
    Clock::time_point t_submit, t_done;
    // device block
    int64_t off = -1, bytes = 0;
    uint8_t* dA = nullptr;
    uint8_t* dB = nullptr;
    int32_t* rowmax = nullptr;
    int2* srow = nullptr;
    int2* scol = nullptr;
    unsigned long long* key = nullptr;
    int32_t* rowseg = nullptr;  // stage 1: each row's first tile segment holding its maximum
    int32_t* seg = nullptr;     // the end row's, after the argmax
    int32_t* found = nullptr;
    WalkState* walk = nullptr;
    uint8_t* ops = nullptr;
    XPoint* xps = nullptr;
    int64_t xcap = 0;
    uint8_t* work = nullptr;  // per-phase hand-off buffers and direction codes
    int64_t work_bytes = 0;
    int64_t max_regs = 1;     // regions of the job's largest sweep
    int64_t reg_at = 0;       // its regions, in the pool seen as an array of Region
    // schedule
    Phase phase = kQueued;
    int mode = kStage1;
    std::vector<Region> regs;
    std::vector<int64_t> Sr, Gr;
    int64_t D = 0, d = 0;
    int64_t launched_to = 0;  // diagonals of the current sweep launched so far (may run ahead of d)
    int L = 0;                      // traceback: checkpoint blocks recomputed ahead per direction
    std::vector<WalkRegion> wtab;   // the round's (L+1)^2 blocks, rows == 0 where none
    std::vector<int> wtab_reg;      // region of each block (-1 where none)
    int32_t best = 0;
    int64_t bi = -1, bj = -1;
    WalkState st{};
    unsigned long long key_h = 0;
    int32_t seg_h = -1;
    int32_t found_h = -1;
    std::vector<uint8_t> ops_h;      // kFetch: the walk's ops, read back
    std::vector<XPoint> xps_h;
    int status = SWH_OK;
    bool published = false;  // result visible to swh_server_poll (set under the server mutex)
    swh_result res{};
};

}  // namespace

struct swh_server {
    swh_ctx* ctx = nullptr;
    swh_scoring sc{};
    swh_opts opts{};
    int32_t max_in_flight = 0;
    int wps = 1;  // sweep waves per SIMD of the shared launches
    int64_t w16 = 0;      // widest tile whose values fit int16 around its base (0: none)
    bool packed = false;  // stage-1 and codes sweeps in packed 16-bit arithmetic
    uint8_t* pool_base = nullptr;
    Pool pool;
    TickSlot ring[kTickRing];
    int ring_at = 0;
    DevBuf argjobs, wargs, wregs;
    // batched transfers: uploads queued for the next tick (a value copied here, or a host
    // buffer that outlives the tick), read-backs of the current tick
    struct Up {
        uint8_t* dst;
        const uint8_t* host;  // nullptr: up_vals[val]
        int64_t len, val;
    };
    std::vector<Up> up;
    std::vector<uint8_t> up_vals;
    struct Rb {
        const uint8_t* dev;
        int64_t len;
        void* host;
    };
    std::vector<Rb> rb;
    PinBuf rb_h, rbx_h;
    hipEvent_t rb_done = nullptr;  // recorded after a tick's read-back transfer
    DevBuf rb_d, rbx_d;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SJob*> queue;
    std::map<int64_t, SJob*> jobs;  // every job not yet collected by swh_server_poll
    std::vector<SJob*> flight;
    int64_t next_id = 1;
    bool stop = false;
    std::thread loop;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, cells_done = 0;
    int64_t peak_in_flight = 0;     // most jobs in flight at once
    int64_t admitted_mid_flight = 0;  // jobs admitted while another job in flight had started
    int64_t sweep_launches = 0, shared_launches = 0;  // sweep launches; those covering two or more jobs
    int64_t fail_tick = -1;         // test hook (SWALIGN_SERVE_FAIL_TICK): tick that reports a failure
    int64_t ticks = 0;
    // host time of the loop (SWALIGN_TRACE prints it when the server is destroyed): job set-up,
    // tick planning and launches, waiting for read-backs, advancing the jobs
    double ms_setup = 0, ms_launch = 0, ms_wait = 0, ms_advance = 0;
    int64_t nticks = 0, nreadbacks = 0;
    bool loop_exited = false;
    std::vector<double> lat_ms;
    Clock::time_point t_first{}, t_last{};
    bool any = false;
    int loop_status = SWH_OK;
};

namespace {

// Tile width of a server sweep over `cols` columns: the options', else at least the batch
// width (512: a tile's 63 fill steps cost about a tenth) and wider for long rows; never
// wider than the packed sweep's int16 bound, so that every stage-1 and codes launch can run
// packed (two tiles per wave).
int64_t server_tile(const swh_server* s, int64_t cols);

// Working bytes of one job: everything it keeps, plus the largest per-phase set.
int64_t server_tile(const swh_server* s, int64_t cols) {
    if (s->opts.tile_cols > 0) return s->opts.tile_cols;
    // one tile per resident wave across the row, at least 512 columns, rounded down to a power
    // of two (measured: a lone 1 Mbp job took 545 ms at 512 columns against 611 ms at 977 and
    // 685 ms at 768)
    const int64_t want = std::max<int64_t>(512, tile_cols(s->ctx, cols, 0));
    int64_t W = 512;
    while (W * 2 <= want) W *= 2;
    if (s->packed) W = std::min(W, s->w16);
    return W;
}

void layout(swh_server* s, SJob& j) {
    const int64_t m = j.m, n = j.n;
    auto al = [](int64_t x) { return ceil_div(x, 256) * 256; };
    int64_t persist = al(m) + al(n) + al(4 * m) + al(8 * j.pl.nsr * n) + al(8 * j.pl.nsc * m) + al(8) + al(4) +
                      al(sizeof(WalkState)) + al(m + n + 16) + al(4 * m) + al(4);
    j.xcap = j.pl.nsr + 2;
    persist += al((int64_t)sizeof(XPoint) * j.xcap);
    if (m == 0 || n == 0) {  // nothing to sweep: score 0
        j.work_bytes = 0;
        j.bytes = persist;
        return;
    }
    const int64_t S = ceil_div(m, kStripeRows);
    const int64_t stage1 = al(16 * n) + al(16 * m) + al(8 * S);
    const int64_t bw = std::min(n, j.pl.Kc), bh = std::min(m, j.pl.K), nq = ceil_div(n, bw);
    const int64_t search = al(16 * n) + nq * (al(16 * bh) + al(8 * ceil_div(bh, kStripeRows)));
    const int64_t W = server_tile(s, bw);
    const int64_t tstride = ceil_div(W + 63, 64) * 64, G = ceil_div(bw, W);
    const int64_t dirs = ceil_div(bh, kStripeRows) * G * tstride * 64 * 2;
    const int64_t blk = al(dirs) + al(16 * bw) + al(16 * bh) + al(8 * ceil_div(bh, kStripeRows));
    // lookahead: up to 7 blocks in each direction (a band along the diagonal beyond 3) when
    // the job has several, within 1 GiB and an eighth of the pool (so that a small pool can
    // still admit the job)
    j.L = (j.pl.nsr > 0 || j.pl.nsc > 0) ? 7 : 0;
    const int64_t la_cap = std::min<int64_t>(INT64_C(1) << 30, s->pool.size / 8);
    while (j.L > 0 && lookahead_blocks(j.L) * blk > la_cap) --j.L;
    const int64_t tb = lookahead_blocks(j.L) * blk;
    j.work_bytes = std::max({stage1, search, tb});
    // the regions of the job's sweeps stay in its block (one more Region of room to align
    // them to a multiple of sizeof(Region) from the pool's start)
    j.max_regs = std::max<int64_t>({1, nq, lookahead_blocks(j.L)});
    persist += al((int64_t)sizeof(Region) * (j.max_regs + 1));
    j.bytes = persist + j.work_bytes;
}

void queue_value(swh_server* s, void* dst, const void* v, int64_t len) {
    const int64_t at = (int64_t)s->up_vals.size();
    s->up_vals.insert(s->up_vals.end(), (const uint8_t*)v, (const uint8_t*)v + len);
    s->up.push_back(swh_server::Up{(uint8_t*)dst, nullptr, len, at});
}
void finish(swh_server*, SJob& j, int status);

// The job's current sweep regions, copied to its block at the start of the next tick (its
// layout holds max_regs of them; a job whose sweep would need more fails instead).
void upload_regions(swh_server* s, SJob& j) {
    if ((int64_t)j.regs.size() > j.max_regs) {
        finish(s, j, SWH_ERR_INTERNAL);
        return;
    }
    queue_value(s, (Region*)s->pool_base + j.reg_at, j.regs.data(), (int64_t)(sizeof(Region) * j.regs.size()));
}

void queue_ref(swh_server* s, void* dst, const void* host, int64_t len) {
    if (len > 0) s->up.push_back(swh_server::Up{(uint8_t*)dst, (const uint8_t*)host, len, 0});
}

// Appends the copy of [src, src + len) to dst to xs, in pieces of at most kXferChunk bytes.
void add_xfer(std::vector<Xfer>& xs, uint8_t* dst, const uint8_t* src, int64_t len) {
    for (int64_t o = 0; o < len; o += kXferChunk)
        xs.push_back(Xfer{dst + o, src + o, std::min(kXferChunk, len - o)});
}

// Issues the uploads queued since the last tick (one transfer, one scatter kernel).
int flush_uploads(swh_server* s, TickSlot& slot) {
    swh_ctx* c = s->ctx;
    if (s->up.empty()) return SWH_OK;
    int64_t data = 0;
    for (const auto& u : s->up) data += al16(u.len);
    int64_t nx = 0;
    for (const auto& u : s->up) nx += ceil_div(u.len, kXferChunk);
    const int64_t total = data + nx * (int64_t)sizeof(Xfer);
    slot.up_h.resize((size_t)total);
    HCK(c, ensure(slot.up_d, (size_t)total));
    std::vector<Xfer> xs;
    xs.reserve((size_t)nx);
    int64_t off = 0;
    for (const auto& u : s->up) {
        memcpy(slot.up_h.data() + off, u.host ? u.host : s->up_vals.data() + u.val, (size_t)u.len);
        add_xfer(xs, u.dst, (const uint8_t*)slot.up_d.p + off, u.len);
        off += al16(u.len);
    }
    memcpy(slot.up_h.data() + data, xs.data(), sizeof(Xfer) * xs.size());
    HCK(c, hipMemcpyAsync(slot.up_d.p, slot.up_h.data(), (size_t)total, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_copy_list, dim3((unsigned)xs.size()), dim3(256), 0, c->stream,
                       (const Xfer*)((const uint8_t*)slot.up_d.p + data));
    HCK(c, hipGetLastError());
    s->up.clear();
    s->up_vals.clear();
    return SWH_OK;
}

// Gathers this tick's read-backs (one kernel, one transfer) and records an event after
// them; finish_readbacks waits for it (not for later work on the stream) and copies the
// values to their destinations.
int issue_readbacks(swh_server* s) {
    swh_ctx* c = s->ctx;
    int64_t data = 0;
    for (const auto& r : s->rb) data += al16(r.len);
    std::vector<Xfer> xs;
    HCK(c, ensure(s->rb_d, (size_t)std::max<int64_t>(data, 16)));
    int64_t off = 0;
    for (const auto& r : s->rb) {
        add_xfer(xs, (uint8_t*)s->rb_d.p + off, r.dev, r.len);
        off += al16(r.len);
    }
    if (!xs.empty()) {
        HCK(c, ensure_pin(s->rbx_h, sizeof(Xfer) * xs.size()));
        HCK(c, ensure(s->rbx_d, sizeof(Xfer) * xs.size()));
        HCK(c, ensure_pin(s->rb_h, (size_t)std::max<int64_t>(data, 16)));
        memcpy(s->rbx_h.h, xs.data(), sizeof(Xfer) * xs.size());
        HCK(c, hipMemcpyAsync(s->rbx_d.p, s->rbx_h.h, sizeof(Xfer) * xs.size(), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_copy_list, dim3((unsigned)xs.size()), dim3(256), 0, c->stream, (const Xfer*)s->rbx_d.p);
        HCK(c, hipGetLastError());
        HCK(c, hipMemcpyAsync(s->rb_h.h, s->rb_d.p, (size_t)data, hipMemcpyDeviceToHost, c->stream));
    }
    HCK(c, hipEventRecord(s->rb_done, c->stream));
    return SWH_OK;
}

int finish_readbacks(swh_server* s) {
    swh_ctx* c = s->ctx;
    HCK(c, hipEventSynchronize(s->rb_done));
    int64_t off = 0;
    for (const auto& r : s->rb) {
        memcpy(r.host, s->rb_h.h + off, (size_t)r.len);
        off += al16(r.len);
    }
    s->rb.clear();
    return SWH_OK;
}

void place(swh_server* s, SJob& j) {
    uint8_t* p = s->pool_base + j.off;
    auto take = [&](int64_t bytes) {
        uint8_t* q = p;
        p += ceil_div(bytes, 256) * 256;
        return q;
    };
    j.dA = take(j.m);
    j.dB = take(j.n);
    j.rowmax = (int32_t*)take(4 * j.m);
    j.srow = j.pl.nsr ? (int2*)take(8 * j.pl.nsr * j.n) : nullptr;
    if (!j.pl.nsr) take(0);
    j.scol = j.pl.nsc ? (int2*)take(8 * j.pl.nsc * j.m) : nullptr;
    j.key = (unsigned long long*)take(8);
    j.rowseg = (int32_t*)take(4 * j.m);
    j.seg = (int32_t*)take(4);
    j.found = (int32_t*)take(4);
    j.walk = (WalkState*)take(sizeof(WalkState));
    j.ops = take(j.m + j.n + 16);
    j.xps = (XPoint*)take((int64_t)sizeof(XPoint) * j.xcap);
    if (j.m > 0 && j.n > 0) {
        const int64_t at = (int64_t)(take((int64_t)sizeof(Region) * (j.max_regs + 1)) - s->pool_base);
        j.reg_at = ceil_div(at, (int64_t)sizeof(Region));
    }
    j.work = take(j.work_bytes);
}

// Carves each region's tile hand-off buffers (and direction codes) out of the job's
// work area and computes the schedule of the phase.
void start_sweep(SJob& j, int mode, int64_t W, int64_t G_fix, int64_t tstride) {
    j.phase = kSweep;
    j.mode = mode;
    j.d = 0;
    j.D = 0;
    j.launched_to = 0;
    j.Sr.assign(j.regs.size(), 0);
    j.Gr.assign(j.regs.size(), 0);
    uint8_t* p = j.work;
    auto take = [&](int64_t bytes) {
        uint8_t* q = p;
        p += ceil_div(bytes, 256) * 256;
        return q;
    };
    int64_t G = G_fix;
    for (size_t q = 0; q < j.regs.size(); ++q) G = std::max(G, ceil_div(j.regs[q].c1 - j.regs[q].c0 + 1, W));
    if (mode == kDirs)
        for (Region& r : j.regs) r.dirs = (uint16_t*)take(ceil_div(r.r1 - r.r0 + 1, kStripeRows) * G * tstride * 64 * 2);
    for (size_t q = 0; q < j.regs.size(); ++q) {
        Region& r = j.regs[q];
        const int64_t rows = r.r1 - r.r0 + 1, cols = r.c1 - r.c0 + 1;
        j.Sr[q] = ceil_div(rows, kStripeRows);
        j.Gr[q] = ceil_div(cols, W);
        j.D = std::max(j.D, j.Sr[q] + j.Gr[q] - 1);
        r.W = W;
        r.G = G;
        r.tstride = tstride;
        r.rb = (int2*)take(16 * cols);
        r.cb = (int2*)take(16 * rows);
        r.cor = (int32_t*)take(8 * j.Sr[q]);
    }
}

Region base_region(const SJob& j) {
    Region r{};
    r.A = j.dA;
    r.B = j.dB;
    r.m = j.m;
    r.n = j.n;
    r.K = j.pl.K;
    r.Kc = j.pl.Kc;
    return r;
}

void setup_stage1(swh_server* s, SJob& j) {
    Region r = base_region(j);
    r.r0 = 0;
    r.r1 = j.m - 1;
    r.c0 = 0;
    r.c1 = j.n - 1;
    r.rowmax = j.rowmax;
    r.srow = j.srow;
    r.rowseg = j.rowseg;
    r.scol = j.scol;
    j.regs.assign(1, r);
    start_sweep(j, kStage1, server_tile(s, j.n), 0, 0);
    upload_regions(s, j);
}

void setup_search(swh_server* s, SJob& j) {
    const int64_t br0 = (j.bi / j.pl.K) * j.pl.K, bw = std::min(j.pl.Kc, j.n), nq = ceil_div(j.n, bw);
    // only the blocks over the first stage-1 segment holding the end row's maximum (columns
    // [gW, (g+1)W) of rectangular tiles): no column left of it holds the score in that row
    int64_t q0 = 0, q1 = nq - 1;
    const int64_t W1 = server_tile(s, j.n), g = j.seg_h;
    if (g >= 0 && g * W1 < j.n) {
        q0 = (g * W1) / bw;
        q1 = (std::min(j.n, (g + 1) * W1) - 1) / bw;
    }
    j.regs.clear();
    for (int64_t q = q0; q <= q1; ++q) {
        Region r = base_region(j);
        r.r0 = br0;
        r.r1 = j.bi;
        r.c0 = q * bw;
        r.c1 = std::min(j.n, r.c0 + bw) - 1;
        r.top = br0 > 0 ? j.srow + (br0 / j.pl.K - 1) * j.n + r.c0 : nullptr;
        r.left = r.c0 > 0 ? j.scol + (r.c0 / j.pl.Kc - 1) * j.m : nullptr;
        r.target_row = j.bi;
        r.target = j.best;
        r.found = j.found;
        j.regs.push_back(r);
    }
    start_sweep(j, kSearch, server_tile(s, bw), 0, 0);
    upload_regions(s, j);
}

// A traceback round: the block of the walk's current cell (cut at the cell, region 0) and
// the full blocks up to L above and to the left, each from its own checkpoints.
void setup_dirs(swh_server* s, SJob& j) {
    const int64_t bw = std::min(j.pl.Kc, j.n);
    const int64_t W = server_tile(s, bw);
    const int64_t tstride = ceil_div(W + 63, 64) * 64, G = ceil_div(bw, W);
    const int64_t brow = j.st.i / j.pl.K, bcol = j.st.j / j.pl.Kc, L = j.L;
    j.regs.clear();
    j.wtab.assign((size_t)((L + 1) * (L + 1)), WalkRegion{});
    j.wtab_reg.assign(j.wtab.size(), -1);
    for (int64_t da = 0; da <= L; ++da) {
        for (int64_t db = 0; db <= L; ++db) {
            if (brow - da < 0 || bcol - db < 0 || !lookahead_block(L, da, db)) continue;
            Region r = base_region(j);
            r.r0 = (brow - da) * j.pl.K;
            r.r1 = da == 0 ? j.st.i : std::min(j.m, r.r0 + j.pl.K) - 1;
            r.c0 = (bcol - db) * j.pl.Kc;
            r.c1 = db == 0 ? j.st.j : std::min(j.n, r.c0 + j.pl.Kc) - 1;
            r.top = r.r0 > 0 ? j.srow + (r.r0 / j.pl.K - 1) * j.n + r.c0 : nullptr;
            r.left = r.c0 > 0 ? j.scol + (r.c0 / j.pl.Kc - 1) * j.m : nullptr;
            j.wtab_reg[(size_t)(da * (L + 1) + db)] = (int)j.regs.size();
            j.regs.push_back(r);
        }
    }
    start_sweep(j, kDirs, W, G, tstride);
    upload_regions(s, j);
    for (size_t k = 0; k < j.wtab.size(); ++k) {
        if (j.wtab_reg[k] < 0) continue;
        const Region& r = j.regs[(size_t)j.wtab_reg[k]];
        j.wtab[k] = WalkRegion{r.r0, r.c0, r.r1 - r.r0 + 1, r.c1 - r.c0 + 1, r.dirs};
    }
}

// Marks a job finished; its pool block is released when the loop publishes it.
void finish(swh_server*, SJob& j, int status) {
    j.status = status;
    j.phase = kDone;
    j.t_done = Clock::now();
}

const int32_t kNotFound = INT32_MAX;  // initial value of a search's found column

// Diagonals each sweeping job advances per tick: the host's per-tick work (uploads, the
// synchronisation) is shared by that many launches per mode.
constexpr int kDiagonalsPerTick = 8;

// Per mode, kDiagonalsPerTick launches over the sweeping jobs that `from` selects: launch k
// covers diagonal from(j) + k of every selected job j (below its D); from(j) < 0 skips j.
// The descriptors go to `set`; each selected job's launched_to moves past its launches.
template <class From>
int launch_sweeps(swh_server* s, SweepSet& set, SweepParams p, From from) {
    swh_ctx* c = s->ctx;
    std::vector<TileSpan>* spans = set.hspans;
    for (int mo = 0; mo < 3; ++mo) spans[mo].clear();
    std::vector<int64_t> span_at[3], tiles_at[3];
    std::vector<int64_t> d0(s->flight.size(), -1);
    bool any[3] = {false, false, false};
    for (size_t x = 0; x < s->flight.size(); ++x) {
        SJob* j = s->flight[x];
        if (j->phase != kSweep) continue;
        d0[x] = from(j);
        if (d0[x] < 0 || d0[x] >= j->D) {
            d0[x] = -1;
            continue;
        }
        any[j->mode] = true;
        j->launched_to = std::min(j->D, d0[x] + kDiagonalsPerTick);
    }
    for (int mo = 0; mo < 3; ++mo) {
        if (!any[mo]) continue;
        for (int step = 0; step < kDiagonalsPerTick; ++step) {
            span_at[mo].push_back((int64_t)spans[mo].size());
            int64_t t = 0;
            int64_t jobs_here = 0;
            for (size_t x = 0; x < s->flight.size(); ++x) {
                SJob* j = s->flight[x];
                if (d0[x] < 0 || j->mode != mo) continue;
                const int64_t d = d0[x] + step;
                if (d >= j->D) continue;
                ++jobs_here;
                for (size_t q = 0; q < j->regs.size(); ++q) {
                    const int64_t Sq = j->Sr[q], Gq = j->Gr[q];
                    if (d > Sq + Gq - 2) continue;
                    const int64_t s_lo = std::max<int64_t>(0, d - (Gq - 1)), s_hi = std::min(Sq - 1, d);
                    spans[mo].push_back(TileSpan{t, (int32_t)(j->reg_at + (int64_t)q), (int32_t)s_lo, d});
                    t += s_hi - s_lo + 1;
                }
            }
            tiles_at[mo].push_back(t);
            if (t) {
                ++s->sweep_launches;
                if (jobs_here > 1) ++s->shared_launches;
            }
        }
        span_at[mo].push_back((int64_t)spans[mo].size());
    }
    p.regions = (const Region*)s->pool_base;  // every job's regions lie in its pool block
    for (int mo = 0; mo < 3; ++mo) {
        if (spans[mo].empty()) continue;
        HCK(c, ensure(set.spans[mo], sizeof(TileSpan) * spans[mo].size()));
        HCK(c, hipMemcpyAsync(set.spans[mo].p, spans[mo].data(), sizeof(TileSpan) * spans[mo].size(),
                              hipMemcpyHostToDevice, c->stream));
        for (int step = 0; step < kDiagonalsPerTick; ++step) {
            if (!tiles_at[mo][(size_t)step]) continue;
            p.spans = (const TileSpan*)set.spans[mo].p + span_at[mo][(size_t)step];
            p.nspans = (int32_t)(span_at[mo][(size_t)step + 1] - span_at[mo][(size_t)step]);
            p.ntiles = tiles_at[mo][(size_t)step];
            // packed (two tiles per wave) when the launch has tiles enough to give the SIMDs
            // two waves each or more; fewer tiles run faster one per wave in int32
            if (s->packed && mo != kSearch && packed_launch(c, p.ntiles))
                HCK(c, launch_sweep_pk(mo, p, c->lds_pad[s->wps], c->stream));
            else HCK(c, launch_sweep(mo, p, c->lds_pad[s->wps], c->stream));
        }
    }
    return SWH_OK;
}

int tick(swh_server* s) {
    swh_ctx* c = s->ctx;
    const auto t_tick = Clock::now();
    ++s->nticks;
    if (s->fail_tick >= 0 && s->ticks++ == s->fail_tick) return SWH_ERR_INTERNAL;  // injected from the host
    TickSlot& slot = s->ring[s->ring_at];
    s->ring_at = (s->ring_at + 1) % kTickRing;
    if (slot.pending) {
        HCK(c, hipEventSynchronize(slot.done));
        slot.pending = false;
    }
    {
        const int rc = flush_uploads(s, slot);
        if (rc) return rc;
    }
    SweepParams p{};
    p.sc = Score{s->sc.match, s->sc.mismatch, s->sc.gap_open, s->sc.gap_extend};
    if (!c->codes_off) {  // symbol codes when one mapping separates the bytes of every job in flight
        uint32_t mask[8] = {};
        for (SJob* j : s->flight)
            for (int w = 0; w < 8; ++w) mask[w] |= j->alpha[w];
        p.code_shift1 = code_shift1_for(mask, &s->sc);
        p.code2_shift1 = code2_shift1_for(mask, &s->sc);
    }
    // ---- sweeps of the jobs whose next diagonals were not launched ahead last tick (jobs
    // that entered a sweep this tick)
    {
        const int rc = launch_sweeps(s, slot.now, p, [](const SJob* j) { return j->launched_to <= j->d ? j->d : -1; });
        if (rc) return rc;
    }
    // ---- argmax for jobs whose stage 1 ended last tick ----------------------------------
    std::vector<ArgmaxJob> aj;
    int64_t max_m = 1;
    for (SJob* j : s->flight)
        if (j->phase == kArgmax) {
            aj.push_back(ArgmaxJob{j->rowmax, j->m, j->key, j->rowseg, j->seg});
            max_m = std::max(max_m, j->m);
        }
    if (!aj.empty()) {
        HCK(c, ensure(s->argjobs, sizeof(ArgmaxJob) * aj.size()));
        HCK(c, hipMemcpyAsync(s->argjobs.p, aj.data(), sizeof(ArgmaxJob) * aj.size(), hipMemcpyHostToDevice,
                              c->stream));
        const int bpj = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(max_m, 16384)));
        HCK(c, launch_argmax_rows((const ArgmaxJob*)s->argjobs.p, (int)aj.size(), bpj, c->stream));
        HCK(c, launch_end_segment((const ArgmaxJob*)s->argjobs.p, (int)aj.size(), c->stream));
    }
    // ---- walks for jobs whose block codes are ready ---------------------------------------
    std::vector<WalkArgs> wa;
    std::vector<WalkRegion> wr;
    for (SJob* j : s->flight)
        if (j->phase == kWalk) wr.insert(wr.end(), j->wtab.begin(), j->wtab.end());
    if (!wr.empty()) {
        HCK(c, ensure(s->wregs, sizeof(WalkRegion) * wr.size()));
        size_t x = 0;
        for (SJob* j : s->flight) {
            if (j->phase != kWalk) continue;
            const Region& r = j->regs[0];
            WalkArgs a{};
            a.A = j->dA;
            a.B = j->dB;
            a.regs = (const WalkRegion*)s->wregs.p + x;
            x += j->wtab.size();
            a.L = j->L;
            a.brow = j->st.i / j->pl.K;
            a.bcol = j->st.j / j->pl.Kc;
            a.tstride = r.tstride;
            a.G = r.G;
            a.W = r.W;
            a.K = j->pl.K;
            a.Kc = j->pl.Kc;
            a.srow = j->srow;
            a.n = j->n;
            a.st = j->walk;
            a.ops = j->ops;
            a.xps = j->xps;
            a.xcap = j->xcap;
            a.sc = Score{s->sc.match, s->sc.mismatch, s->sc.gap_open, s->sc.gap_extend};
            wa.push_back(a);
        }
        HCK(c, ensure(s->wargs, sizeof(WalkArgs) * wa.size()));
        HCK(c, hipMemcpyAsync(s->wregs.p, wr.data(), sizeof(WalkRegion) * wr.size(), hipMemcpyHostToDevice,
                              c->stream));
        HCK(c, hipMemcpyAsync(s->wargs.p, wa.data(), sizeof(WalkArgs) * wa.size(), hipMemcpyHostToDevice,
                              c->stream));
        HCK(c, launch_walk((const WalkArgs*)s->wargs.p, (int)wa.size(), many_walks(c, (int64_t)wa.size()), c->stream));
    }
    // ---- read back what decides each job's next step (only when some job needs it) -------
    bool readback = false;
    for (SJob* j : s->flight) {
        readback = readback || j->phase == kArgmax || j->phase == kWalk || j->phase == kFetch ||
                   (j->phase == kSweep && j->mode == kSearch && j->d + kDiagonalsPerTick >= j->D);
    }
    for (SJob* j : s->flight) {
        if (!readback) break;
        if (j->phase == kArgmax) {
            s->rb.push_back(swh_server::Rb{(const uint8_t*)j->key, 8, &j->key_h});
            s->rb.push_back(swh_server::Rb{(const uint8_t*)j->seg, 4, &j->seg_h});
        } else if (j->phase == kSweep && j->mode == kSearch && j->d + kDiagonalsPerTick >= j->D) {
            s->rb.push_back(swh_server::Rb{(const uint8_t*)j->found, 4, &j->found_h});
        } else if (j->phase == kWalk) {
            s->rb.push_back(swh_server::Rb{(const uint8_t*)j->walk, (int64_t)sizeof(WalkState), &j->st});
        } else if (j->phase == kFetch) {
            j->ops_h.resize((size_t)j->st.nops + 1);
            if (j->st.nops) s->rb.push_back(swh_server::Rb{j->ops, j->st.nops, j->ops_h.data()});
            const int64_t nx = s->opts.want_xpoints ? std::min(j->st.nxp, j->xcap) : 0;
            j->xps_h.resize((size_t)std::max<int64_t>(nx, 1));
            if (nx) s->rb.push_back(swh_server::Rb{(const uint8_t*)j->xps, nx * (int64_t)sizeof(XPoint), j->xps_h.data()});
        }
    }
    if (readback) {
        const int rc = issue_readbacks(s);
        if (rc) return rc;
    }
    // ---- the next tick's diagonals of every job whose sweep continues past this tick,
    // launched now: the device sweeps them while the host waits for the read-back and
    // plans the next tick
    {
        const int rc = launch_sweeps(s, slot.ahead, p, [](const SJob* j) {
            return j->d + kDiagonalsPerTick < j->D ? j->d + kDiagonalsPerTick : -1;
        });
        if (rc) return rc;
    }
    HCK(c, hipEventRecord(slot.done, c->stream));
    slot.pending = true;
    const auto t_launched = Clock::now();
    s->ms_launch += std::chrono::duration<double, std::milli>(t_launched - t_tick).count();
    if (readback) {
        const int rc = finish_readbacks(s);
        if (rc) return rc;
        ++s->nreadbacks;
    }
    const auto t_read = Clock::now();
    s->ms_wait += std::chrono::duration<double, std::milli>(t_read - t_launched).count();
    // ---- advance ------------------------------------------------------------------------
    for (SJob* j : s->flight) {
        switch (j->phase) {
            case kSweep:
                j->d = std::min(j->D, j->d + kDiagonalsPerTick);
                if (j->d < j->D) break;
                if (j->mode == kStage1) {
                    j->phase = kArgmax;
                    const unsigned long long zero = 0;
                    queue_value(s, j->key, &zero, 8);
                    const int32_t none = -1;
                    queue_value(s, j->seg, &none, 4);
                } else if (j->mode == kSearch) {
                    if (j->found_h == INT32_MAX || j->found_h < 0) {
                        finish(s, *j, SWH_ERR_INTERNAL);
                        break;
                    }
                    j->bj = j->found_h;
                    j->res.end_row = j->bi;
                    j->res.end_col = j->bj;
                    if (s->opts.score_only) {
                        finish(s, *j, SWH_OK);
                        break;
                    }
                    j->st = WalkState{};
                    j->st.i = j->bi;
                    j->st.j = j->bj;
                    j->st.start_i = j->bi;
                    j->st.start_j = j->bj;
                    j->st.last_row = j->bi;
                    j->st.v = j->best;
                    queue_value(s, j->walk, &j->st, sizeof(WalkState));
                    setup_dirs(s, *j);
                } else {
                    j->phase = kWalk;
                }
                break;
            case kArgmax: {
                j->best = (int32_t)(j->key_h >> 32);
                j->bi = j->key_h ? (int64_t)(0xFFFFFFFFu - (uint32_t)(j->key_h & 0xFFFFFFFFu)) : -1;
                j->res.score = j->best > 0 ? j->best : 0;
                if (j->best <= 0) {
                    finish(s, *j, SWH_OK);
                    break;
                }
                queue_value(s, j->found, &kNotFound, 4);
                setup_search(s, *j);
                break;
            }
            case kWalk: {
                const Region& r = j->regs[0];
                if (j->st.done) {
                    j->phase = kFetch;  // its ops are read back next tick
                } else if (j->st.i >= r.r0 && j->st.j >= r.c0) {
                    finish(s, *j, SWH_ERR_INTERNAL);  // no progress
                } else {
                    setup_dirs(s, *j);
                }
                break;
            }
            case kFetch: {
                const int rc = fill_alignment(j->res, j->st, j->ops_h.data(), nullptr, 0, j->xps_h.data(),
                                              std::min(j->st.nxp, j->xcap), j->best, s->opts.want_xpoints != 0);
                std::vector<uint8_t>().swap(j->ops_h);
                std::vector<XPoint>().swap(j->xps_h);
                finish(s, *j, rc);
                break;
            }
            default:
                break;
        }
    }
    s->ms_advance += std::chrono::duration<double, std::milli>(Clock::now() - t_read).count();
    return SWH_OK;
}

void serve_loop(swh_server* s) {
    if (hipSetDevice(s->ctx->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->loop_status = SWH_ERR_HIP;
        s->loop_exited = true;
        for (SJob* j : s->queue) {
            finish(s, *j, SWH_ERR_HIP);
            j->published = true;
            ++s->failed;
        }
        s->queue.clear();
        s->cv_done.notify_all();
        return;
    }
    while (true) {
        std::vector<SJob*> done;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // admission, in arrival order, while the pool and the slot count allow
            while (!s->queue.empty() && (int64_t)s->flight.size() < s->max_in_flight) {
                SJob* j = s->queue.front();
                if (j->bytes > s->pool.size) {
                    s->queue.pop_front();
                    finish(s, *j, SWH_ERR_BUDGET);
                    done.push_back(j);
                    continue;
                }
                const int64_t off = s->pool.alloc(j->bytes);
                if (off < 0) break;
                s->queue.pop_front();
                j->off = off;
                bool mid = false;  // another job in flight has already run part of its work
                for (SJob* o : s->flight) mid = mid || o->phase != kQueued;
                if (mid) ++s->admitted_mid_flight;
                s->flight.push_back(j);
                s->peak_in_flight = std::max<int64_t>(s->peak_in_flight, (int64_t)s->flight.size());
            }
            if (s->flight.empty() && done.empty()) {
                if (s->stop && s->queue.empty()) {
                    s->loop_exited = true;
                    s->cv_done.notify_all();
                    return;
                }
                s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
                continue;
            }
        }
        int rc = SWH_OK;
        const auto t_setup = Clock::now();
        for (SJob* j : s->flight) {
            if (j->phase != kQueued) continue;
            place(s, *j);
            queue_ref(s, j->dA, j->a.data(), j->m);  // uploaded at the start of the next tick
            queue_ref(s, j->dB, j->b.data(), j->n);
            j->res.m = j->m;
            j->res.n = j->n;
            j->res.swapped = j->swapped ? 1 : 0;
            j->res.end_row = j->res.end_col = j->res.start_row = j->res.start_col = -1;
            j->res.precision_used = s->packed ? 16 : 32;
            j->res.checkpoint_rows = j->pl.nsr ? j->pl.K : 0;
            j->res.checkpoint_cols = j->pl.nsc ? j->pl.Kc : 0;
            if (j->m == 0 || j->n == 0) finish(s, *j, SWH_OK);
            else setup_stage1(s, *j);
        }
        s->ms_setup += std::chrono::duration<double, std::milli>(Clock::now() - t_setup).count();
        if (!rc) rc = tick(s);
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) {
            // a runtime failure ends every job in flight and every queued one: each is
            // published with the failure, so no poll waits for a job the loop will not run
            s->loop_status = rc;
            for (SJob* j : s->flight)
                if (j->phase != kDone) finish(s, *j, rc);
            for (SJob* j : s->queue) {
                finish(s, *j, rc);
                done.push_back(j);
            }
            s->queue.clear();
        }
        std::vector<SJob*> keep;
        for (SJob* j : s->flight) (j->phase == kDone ? done : keep).push_back(j);
        s->flight.swap(keep);
        for (SJob* j : done) {
            if (j->off >= 0) s->pool.release(j->off, j->bytes);
            j->off = -1;
            j->published = true;
            const double ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
            j->res.ms_total = ms;
            if (j->status == SWH_OK) {
                ++s->completed;
                s->cells_done += j->m * j->n;
                s->lat_ms.push_back(ms);
            } else {
                ++s->failed;
            }
            s->t_last = j->t_done;
        }
        if (rc) s->loop_exited = true;
        if (!done.empty() || rc) s->cv_done.notify_all();
        if (rc) return;
    }
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = (size_t)std::min<double>((double)v.size() - 1, q * (double)(v.size() - 1) + 0.5);
    return v[k];
}

}  // namespace

extern "C" {

int swh_server_create(int device, const swh_scoring* sc, const swh_opts* opts, int64_t hbm_budget_bytes,
                      int32_t max_in_flight, swh_server** out) {
    if (!out) return SWH_ERR_INVALID_ARG;
    *out = nullptr;
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    swh_ctx* c = nullptr;
    int rc = swh_ctx_create(device, &c);
    if (rc) return rc;
    swh_server* s = new swh_server();
    s->ctx = c;
    s->sc = *sc;
    if (opts) s->opts = *opts;
    s->max_in_flight = max_in_flight > 0 ? max_in_flight : 1024;
    if (const char* pm = getenv("SWALIGN_PACKED_MIN_TILES")) c->packed_min_tiles = *pm ? std::max<int64_t>(0, atoll(pm)) : -1;
    // test hook: SWALIGN_SERVE_FAIL_TICK=t makes tick t report a runtime failure (no device
    // work is affected), to exercise how a failure ends queued and running jobs
    if (const char* ft = getenv("SWALIGN_SERVE_FAIL_TICK")) s->fail_tick = atoll(ft);
    // a tick's launches hold the tiles of every job in flight: four waves per SIMD hide the
    // waits of one (100-job stream, all submitted at once: 815 / 946 / 968 GCUPS at 1 / 2 / 4)
    s->wps = usable_wps(c, s->opts.waves_per_simd > 0 ? s->opts.waves_per_simd : (c->wps_fixed ? c->wps_fixed : 4));
    // packed 16-bit stage 1 and codes sweeps, as for batches (values held relative to each
    // tile's base, so any absolute score), unless int32 is asked for or no tile width fits
    s->w16 = int16_tile_cols(sc);
    s->packed = s->opts.precision != 32 && s->w16 > 0 && (s->opts.tile_cols <= 0 || s->opts.tile_cols <= s->w16);
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    int64_t budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : (int64_t)(freeb / 2);
    budget = std::min<int64_t>(budget, (int64_t)(freeb * 0.9));
    // tile spans address the jobs' regions by an int32 index into the pool
    budget = std::min<int64_t>(budget, (int64_t)INT32_MAX * (int64_t)sizeof(Region));
    if (hipMalloc((void**)&s->pool_base, (size_t)budget) != hipSuccess) {
        (void)hipGetLastError();
        swh_ctx_destroy(c);
        delete s;
        return SWH_ERR_OUT_OF_MEMORY;
    }
    s->pool.init(budget);
    for (TickSlot& t : s->ring) {
        if (hipEventCreateWithFlags(&t.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            swh_server_destroy(s);
            return SWH_ERR_HIP;
        }
    }
    if (hipEventCreateWithFlags(&s->rb_done, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        swh_server_destroy(s);
        return SWH_ERR_HIP;
    }
    s->loop = std::thread(serve_loop, s);
    *out = s;
    return SWH_OK;
}

int swh_server_submit(swh_server* s, const uint8_t* seq1, int64_t len1, const uint8_t* seq2, int64_t len2,
                      int64_t* job_id) {
    if (!s || !job_id || len1 < 0 || len2 < 0 || (len1 && !seq1) || (len2 && !seq2)) return SWH_ERR_INVALID_ARG;
    SJob* j = new SJob();
    j->swapped = len1 < len2;  // src/SWutils.cpp:84-88
    const uint8_t* A = j->swapped ? seq2 : seq1;
    const uint8_t* B = j->swapped ? seq1 : seq2;
    j->m = j->swapped ? len2 : len1;
    j->n = j->swapped ? len1 : len2;
    j->a.assign(A, A + j->m);
    j->b.assign(B, B + j->n);
    {
        bool seen[256] = {};
        for (uint8_t x : j->a) seen[x] = true;
        for (uint8_t x : j->b) seen[x] = true;
        for (int x = 0; x < 256; ++x)
            if (seen[x]) j->alpha[x >> 5] |= 1u << (x & 31);
    }
    // checkpoints every max(m, n)/128 rows and columns (512 to 8192, or the options: short
    // jobs keep their end-search bands and traceback blocks small, long ones need fewer
    // rounds), doubled until they take at most an eighth of the pool.  Saturated streams of
    // the 100 bp - 20 kbp batch mix: 619 GCUPS at a 1024 minimum, 802 at 512 (814 at 256,
    // whose extra traceback rounds lengthen small jobs' latency: p50 45 ms against 34 at 95%
    // load of the large-job mix)
    swh_opts o = s->opts;
    if (!o.score_only && o.checkpoint_rows <= 0) {
        int64_t K = std::min<int64_t>(8192, std::max<int64_t>(512, next_pow2(std::max(j->m, j->n) / 128)));
        while (plan_mem(j->m, j->n, K, K) > s->pool.size / 8 && K < (INT64_C(1) << 40)) K *= 2;
        o.checkpoint_rows = K;
        o.checkpoint_cols = K;
    }
    if (choose_plan(j->m, j->n, &o, INT64_MAX / 4, &j->pl) != SWH_OK) {
        delete j;
        return SWH_ERR_BUDGET;
    }
    layout(s, *j);
    j->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop || s->loop_status) {
            const int st = s->loop_status ? s->loop_status : SWH_ERR_INVALID_ARG;
            delete j;
            return st;
        }
        j->id = s->next_id++;
        *job_id = j->id;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        ++s->submitted;
        if (!s->any) {
            s->any = true;
            s->t_first = j->t_submit;
        }
    }
    s->cv_work.notify_one();
    return SWH_OK;
}

int swh_server_poll(swh_server* s, int64_t job_id, int32_t wait, swh_result* out) {
    if (!s || !out) return SWH_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return SWH_ERR_NOT_FOUND;
    SJob* j = it->second;
    if (!j->published) {
        if (!wait && !s->loop_exited) return SWH_PENDING;
        // a job is published by the loop; once the loop has exited, an unpublished job
        // (submitted as it stopped) never will be, and reports the loop's status
        s->cv_done.wait(lk, [&] { return j->published || s->loop_exited; });
        if (!j->published) {
            const int st = s->loop_status ? s->loop_status : SWH_ERR_INTERNAL;
            s->jobs.erase(it);
            delete j;
            return st;
        }
    }
    *out = j->res;
    const int st = j->status;
    s->jobs.erase(it);
    delete j;
    return st;
}

int swh_server_stats(swh_server* s, swh_serve_stats* st) {
    if (!s || !st) return SWH_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->in_flight = (int64_t)s->flight.size();
    st->queued = (int64_t)s->queue.size();
    st->cells_completed = s->cells_done;
    const double sec = std::chrono::duration<double>(s->t_last - s->t_first).count();
    st->gcups = sec > 0 ? (double)s->cells_done / sec / 1e9 : 0.0;
    st->p50_ms = percentile(s->lat_ms, 0.50);
    st->p99_ms = percentile(s->lat_ms, 0.99);
    st->max_ms = s->lat_ms.empty() ? 0.0 : *std::max_element(s->lat_ms.begin(), s->lat_ms.end());
    st->hbm_pool_bytes = s->pool.size;
    st->hbm_in_use_bytes = s->pool.in_use();
    st->loop_status = s->loop_status;
    st->peak_in_flight = s->peak_in_flight;
    st->admitted_mid_flight = s->admitted_mid_flight;
    st->sweep_launches = s->sweep_launches;
    st->shared_launches = s->shared_launches;
    return SWH_OK;
}

int swh_server_destroy(swh_server* s) {
    if (!s) return SWH_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->loop.joinable()) s->loop.join();
    if (s->ctx->trace)
        fprintf(stderr, "swalign: server: %lld ticks (%lld with read-backs); host ms: set-up %.1f, planning and "
                "launches %.1f, waiting for read-backs %.1f, advancing %.1f\n", (long long)s->nticks,
                (long long)s->nreadbacks, s->ms_setup, s->ms_launch, s->ms_wait, s->ms_advance);
    for (auto& kv : s->jobs) {
        swh_result_free(&kv.second->res);
        delete kv.second;
    }
    for (SJob* j : s->queue)
        if (s->jobs.find(j->id) == s->jobs.end()) delete j;
    (void)hipStreamSynchronize(s->ctx->stream);
    for (TickSlot& t : s->ring) {
        for (SweepSet* w : {&t.now, &t.ahead})
            for (int mo = 0; mo < 3; ++mo)
                if (w->spans[mo].p) (void)hipFree(w->spans[mo].p);
        if (t.up_d.p) (void)hipFree(t.up_d.p);
        if (t.done) (void)hipEventDestroy(t.done);
    }
    if (s->rb_done) (void)hipEventDestroy(s->rb_done);
    for (PinBuf* b : {&s->rb_h, &s->rbx_h})
        if (b->h) (void)hipHostFree(b->h);
    DevBuf* bufs[] = {&s->argjobs, &s->wargs, &s->wregs, &s->rb_d, &s->rbx_d};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (s->pool_base) (void)hipFree(s->pool_base);
    swh_ctx_destroy(s->ctx);
    delete s;
    return SWH_OK;
}

}  // extern "C"
