// Allocation-free variant of swh_align_pair_device (swh_align_pair_device_noalloc).
//
// Every device buffer the alignment needs beyond its inputs -- row maxima, checkpoint rows
// and columns, the sweeps' tile hand-off buffers and launch descriptors, the traceback's
// direction codes, walk state and ops -- is a slice of one static device array, `scratch`,
// of SWH_NOALLOC_SCRATCH_BYTES per device.  The host copy of the ops and the CIGAR go to the
// caller's buffer.  So the call makes no hipMalloc, hipHostMalloc or hipFree, and no other
// device or pinned host allocation; its host-side bookkeeping (launch lists) uses ordinary
// heap vectors.  The stage driver and the kernels are swh_align_pair_device's own: the
// results are identical for the same checkpoint spacing.
//
// Every call on a device shares that device's array, which keeps the last call's contents;
// nothing is read before the call writes it (the end-search keys and found columns are
// cleared by the driver, row maxima and checkpoints are written by stage 1 before they are
// read).  Calls that overlap on one device race on the array and must be serialised by the
// caller.
#include <algorithm>
#include <mutex>

#include "sw_host.h"

using namespace swk;

namespace {

// The scratch array: one copy per device, allocated when the code object is loaded there
// and held until the process ends.  Host code never uses this name as an address (it is a
// host-side stand-in); it uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[SWH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_status[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

int64_t al256(int64_t x) { return (std::max<int64_t>(x, 64) + 255) / 256 * 256; }

// Byte sizes of the fixed slices for one pair of m rows by n columns with plan pl and
// traceback lookahead L (upper bounds of what the stage driver asks of each buffer).
struct Layout {
    int64_t rowmax, rowseg, segs, key, found, srow, scol, cb, rbuf, cor, regions, spans, argjobs, alpha,
        ranges, ops, xps, walk, wregs, wargs, dirs;
    int64_t total() const {
        return rowmax + rowseg + segs + key + found + srow + scol + cb + rbuf + cor + regions + spans + argjobs +
               alpha + ranges + ops + xps + walk + wregs + wargs + dirs;
    }
};

Layout layout(swh_ctx* c, int64_t m, int64_t n, const swh_host::Plan& pl, const swh_host::EngineOpts& eo,
              const swh_scoring* sc, int64_t L, bool traceback) {
    using swh_host::ceil_div;
    const int64_t S = ceil_div(m, kStripeRows);
    const int64_t w16 = swh_host::int16_tile_cols(sc);
    // narrowest tile the driver can choose for stage 1 and for the block sweeps
    int64_t W1 = eo.tile_cols > 0 ? eo.tile_cols : 256;
    if (eo.tile_cols <= 0 && eo.precision == 16 && w16 > 0) W1 = std::min<int64_t>(W1, w16);
    const int64_t bh = std::min(m, pl.K), bw = std::min(n, pl.Kc), sb = ceil_div(bh, kStripeRows);
    const int64_t nq = ceil_div(n, bw);
    const int64_t Ws = swh_host::tile_cols(c, bw, eo.tile_cols);  // end search and traceback tiles
    const int64_t tstride = ceil_div(Ws + 63, 64) * 64, Gblk = ceil_div(bw, Ws);
    const int64_t LB = traceback ? swh_host::lookahead_blocks(L) : 0;
    const int64_t nreg = std::max<int64_t>({1, nq, LB});
    Layout y{};
    y.rowmax = al256(4 * m);
    y.rowseg = al256(4 * m);
    y.segs = al256(4);
    y.key = al256(8);
    y.found = al256(4);
    y.srow = al256(8 * pl.nsr * n);
    y.scol = al256(8 * pl.nsc * m);
    y.cb = al256(16 * std::max({m, nq * bh, LB * bh}));
    y.rbuf = al256(16 * std::max(n, LB * bw));
    y.cor = al256(8 * std::max({S, nq * sb, LB * sb}));
    y.regions = al256((int64_t)sizeof(Region) * nreg);
    y.spans = al256((int64_t)sizeof(TileSpan) *
                    std::max({S + ceil_div(n, W1), nq * (sb + Gblk), LB * (sb + Gblk)}));
    y.argjobs = al256((int64_t)sizeof(ArgmaxJob));
    y.alpha = al256(32);
    y.ranges = al256((int64_t)sizeof(ByteRange) * (ceil_div(m, 65536) + ceil_div(n, 65536)));
    y.ops = al256(m + n + 16);
    y.xps = al256((int64_t)sizeof(XPoint) * (pl.nsr + 2));
    y.walk = al256((int64_t)sizeof(WalkState));
    y.wregs = al256((int64_t)sizeof(WalkRegion) * (L + 1) * (L + 1));
    y.wargs = al256((int64_t)sizeof(WalkArgs));
    y.dirs = al256(LB * sb * Gblk * tstride * 64 * (int64_t)sizeof(uint16_t));
    return y;
}

}  // namespace

namespace swh_host {

int scratch_resolve(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return SWH_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            g_scratch_status[device] = SWH_ERR_HIP;
        } else {
            g_scratch_ptr[device] = p;
            g_scratch_status[device] = SWH_OK;
        }
        g_scratch_done[device] = true;
    }
    *out = g_scratch_ptr[device];
    return g_scratch_status[device];
}

}  // namespace swh_host

using namespace swh_host;

extern "C" int swh_align_pair_device_noalloc(swh_ctx* c, const uint8_t* d1, int64_t len1, const uint8_t* d2,
                                             int64_t len2, const swh_scoring* sc, const swh_opts* o,
                                             swh_result* out, char* cigar, int64_t cigar_cap) {
    if (!c || !out || len1 < 0 || len2 < 0 || (len1 && !d1) || (len2 && !d2)) return SWH_ERR_INVALID_ARG;
    result_init(out);
    if (!scoring_ok(sc)) return SWH_ERR_SCORING;
    // options that would allocate: crosspoints and host-resident checkpoints
    if (o && (o->want_xpoints || o->host_checkpoints)) return SWH_ERR_INVALID_ARG;
    const bool traceback = !(o && o->score_only);
    // the CIGAR (at most 2 (len1 + len2) + 1 bytes) goes to the front of the caller's buffer
    // and the traceback's ops (len1 + len2 + 16 bytes) are staged in its back
    if (traceback && (!cigar || cigar_cap < 3 * (len1 + len2) + 64)) return SWH_ERR_INVALID_ARG;
    if (c->scratch_status != SWH_OK || !c->scratch) return c->scratch_status ? c->scratch_status : SWH_ERR_HIP;
    if (traceback) cigar[0] = 0;  // stays empty when the score is 0
    HCK(c, hipSetDevice(c->device));

    std::vector<Job> jobs{make_job(d1, len1, d2, len2, traceback)};
    Job& j = jobs[0];
    EngineOpts eo = engine_opts(o, false);
    eo.skew = false;  // the layout below is sized for rectangular stage-1 tiles
    const int64_t L_want = traceback ? std::max(0, eo.lookahead) : 0;
    const int64_t cap = SWH_NOALLOC_SCRATCH_BYTES;
    // checkpoint plan and lookahead: the options' spacing, or the default spacing doubled
    // until the pair's working set fits the scratch array; the widest lookahead that fits
    const bool given = o && (o->checkpoint_rows > 0 || o->checkpoint_cols > 0);
    swh_opts po = o ? *o : swh_opts{};
    Layout y{};
    int64_t L = -1;
    for (int attempt = 0; attempt < 48 && L < 0; ++attempt) {
        Plan pl{};
        const int rc = choose_plan(j.m, j.n, &po, INT64_MAX / 4, &pl);
        if (rc) return rc;
        for (int64_t l = L_want; l >= 0 && L < 0; --l) {
            y = layout(c, j.m, j.n, pl, eo, sc, l, traceback);
            if (y.total() <= cap) L = l;
        }
        if (L >= 0) {
            j.pl = pl;
            break;
        }
        if (given || !traceback || (pl.K >= 2 * std::max<int64_t>(j.m, 1) && pl.Kc >= 2 * std::max<int64_t>(j.n, 1)))
            return SWH_ERR_SCRATCH;
        po.checkpoint_rows = 2 * pl.K;
        po.checkpoint_cols = 2 * pl.Kc;
    }
    if (L < 0) return SWH_ERR_SCRATCH;
    eo.lookahead = (int)L;

    // A copy of the context whose buffers are slices of the scratch array.
    swh_ctx lc = *c;
    uint8_t* p = (uint8_t*)c->scratch;
    auto slice = [&](DevBuf& b, int64_t bytes) {
        b.p = p;
        b.cap = (size_t)bytes;
        b.fixed = true;
        p += bytes;
    };
    DevBuf none{};
    none.fixed = true;  // capacity 0: any use fails instead of allocating
    lc.seqA = lc.seqB = lc.seqs = none;
    lc.tb_pairs = lc.tb_states = lc.tb_act = lc.tb_pre = none;  // batches only
    slice(lc.rowmax, y.rowmax);
    slice(lc.rowseg, y.rowseg);
    slice(lc.segs, y.segs);
    slice(lc.key, y.key);
    slice(lc.found, y.found);
    slice(lc.srow, y.srow);
    slice(lc.scol, y.scol);
    slice(lc.cb, y.cb);
    slice(lc.rbuf, y.rbuf);
    slice(lc.cor, y.cor);
    slice(lc.regions, y.regions);
    slice(lc.spans, y.spans);
    slice(lc.argjobs, y.argjobs);
    slice(lc.alpha, y.alpha);
    slice(lc.ranges, y.ranges);
    slice(lc.ops, y.ops);
    slice(lc.xps, y.xps);
    slice(lc.walk, y.walk);
    slice(lc.wregs, y.wregs);
    slice(lc.wargs, y.wargs);
    slice(lc.dirs, y.dirs);
    lc.hsrow = lc.hscol = HostBuf{};
    lc.staging = nullptr;
    lc.staging_cap = 0;
    lc.host_ops = nullptr;
    lc.host_ops_cap = 0;
    lc.pinned_spans = nullptr;
    lc.pinned_spans_cap = 0;
    lc.dirs_cap_fixed = y.dirs;
    lc.fixed_memory = true;
    if (traceback) {
        const int64_t ops_bytes = len1 + len2 + 16;
        lc.ops_dst = (uint8_t*)cigar + (cigar_cap - ops_bytes);
        lc.ops_dst_cap = (size_t)ops_bytes;
        lc.cigar_dst = cigar;
        lc.cigar_dst_cap = cigar_cap - ops_bytes;
    }
    int rc = run_jobs(&lc, jobs, sc, eo, out);
    c->last = lc.last;
    if (rc == SWH_ERR_OUT_OF_MEMORY) rc = SWH_ERR_SCRATCH;  // a slice was too small: never allocated
    return rc ? rc : jobs[0].status;
}
