/*
 * swalign_hip.h -- C-ABI of the MI355X (gfx950) SWalign library, libswalign_hip.so.
 *
 * Optimal local alignment of two byte sequences with affine gaps (Smith-Waterman-Gotoh),
 * as computed by Sterbic/Projekt-2012 ("SWalign" v0.6.3), with the forward pass,
 * checkpoint rows/columns, end-cell search and traceback run as hand-written HIP
 * kernels.  Every function returns an int status (SWH_OK = 0); nothing in the library
 * calls exit().  All sizes are int64.  Inputs are borrowed for the duration of a call;
 * results are owned by the library until swh_result_free().
 *
 * Orientation follows the reference (src/SWutils.cpp:84-88): the longer sequence
 * becomes the rows ("A"), the shorter the columns ("B"); with equal lengths the
 * argument order is kept.  Coordinates in results are 0-based [row in A, column in B].
 *
 * Scoring: s(a,b) = match if a == b (byte equality, no alphabet mapping) else mismatch;
 * a gap of length L costs gap_open + (L-1)*gap_extend (the reference's "first" and
 * "extension", src/SWutils.h:15-20).  Valid iff match >= 1 and the three penalties
 * are <= -1 (src/SWutils.cpp:36).
 *
 * Canonical tie rules (the reference's own order depends on its launch configuration):
 *   end cell : the first cell in row-major order holding the maximal score;
 *   traceback: greedy backward walk, H: zero (start) > diagonal > vertical gap >
 *              horizontal gap; gaps: open > extend.
 * CIGAR ops: '=' match, 'X' mismatch, 'D' base of A against a gap (vertical move),
 *            'I' base of B against a gap (horizontal move).
 */
#ifndef SWALIGN_HIP_H_
#define SWALIGN_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
enum {
    SWH_OK = 0,
    SWH_ERR_INVALID_ARG = 1,     /* null pointer, negative length, bad option            */
    SWH_ERR_SCORING = 2,         /* scoring values not usable (SWutils.cpp:36-37)        */
    SWH_ERR_NO_DEVICE = 3,       /* no HIP device (SWutils.cpp:323-324)                  */
    SWH_ERR_HIP = 4,             /* a HIP runtime call failed                            */
    SWH_ERR_OUT_OF_MEMORY = 5,   /* device allocation failed or budget too small         */
    SWH_ERR_BUDGET = 6,          /* request rejected: exceeds the HBM admission budget   */
    SWH_ERR_INTERNAL = 7,        /* traceback inconsistency (should never happen)        */
    SWH_ERR_NOT_FOUND = 8,       /* unknown job id / nothing to poll                     */
    SWH_PENDING = 9,             /* swh_server_poll: the job has not finished (no error)  */
    SWH_ERR_SCRATCH = 10         /* the pair's working set exceeds the fixed scratch memory of */
                                 /* swh_align_pair_device_noalloc                             */
};

/* Mirrors `scoring` (src/SWutils.h:15-20) field for field. */
typedef struct {
    int32_t match;
    int32_t mismatch;
    int32_t gap_open;     /* "first"     */
    int32_t gap_extend;   /* "extension" */
} swh_scoring;

/* Options; zero-initialise for defaults. */
typedef struct {
    int64_t checkpoint_rows;   /* K: checkpoint row spacing (multiple of 256), 0 = auto     */
    int64_t checkpoint_cols;   /* Kc: checkpoint column spacing (power of two >= 64), 0 = K */
    int64_t hbm_budget_bytes;  /* cap on checkpoint memory, 0 = half of free HBM            */
    int32_t score_only;        /* 1: score + end cell only (no traceback)                   */
    int32_t precision;         /* stage 1: 16 = packed int16 (values relative to each tile's  */
                               /* base) where a tile's range fits int16, else int32; 32 = int32;*/
                               /* 0 = 16 for batches; single pairs 16 where a model expects it */
                               /* faster (1.5 Mbp and more on each side), else 32              */
    int32_t want_xpoints;      /* 1: return the path nodes on checkpoint rows                */
    int32_t tile_cols;         /* columns per sweep tile (any width >= 1); 0 = auto: int32      */
                               /* single pairs max(cols / stripes, cols / (4 x SIMDs)) - 64,   */
                               /* a multiple of 64 >= 192 (shifted stripes); batches 512;      */
                               /* packed single pairs by a model; traceback blocks about four  */
                               /* tiles per CU across the block, >= 256                        */
    int32_t tb_lookahead;      /* traceback: checkpoint blocks recomputed ahead of the walk in  */
                               /* each direction per round (beyond 3, only blocks within 2 of  */
                               /* the diagonal); 0 = auto (7; batches: per round, as many as  */
                               /* the walking pairs' blocks leave room for), -1 = none        */
    int32_t waves_per_simd;    /* sweep waves resident per SIMD (1 to 4), 0 = auto (batches 3,  */
                               /* the request server 4, int32 single pairs 4, packed ones 2   */
                               /* or 3 by a model)                                             */
    int32_t autotune;          /* batch: 1 = time candidate settings on a sample first          */
    int32_t host_checkpoints;  /* 1: checkpoint rows and columns live in pinned host memory    */
                               /* mapped into the device (the reference's temp/row_R side      */
                               /* channel, src/Builders.cpp:27-104, kept in memory): HBM holds */
                               /* none of them and hbm_budget_bytes caps that host memory      */
                               /* instead (0 = half the available host memory); the request    */
                               /* server keeps its checkpoints in its HBM pool                 */
} swh_opts;

/* Path node on a checkpoint row (the analogue of the reference's crosspoints,
 * src/Crosspointer.cpp:104-108, printed by src/main.cu:245-248). */
typedef struct {
    int64_t row, col;
    int32_t score;   /* forward value of the path at that node */
    int32_t layer;   /* 0 = H (match/mismatch state), 2 = inside a vertical gap */
} swh_xpoint;

typedef struct {
    int32_t score;
    int32_t swapped;          /* 1 if seq2 was longer and became the rows            */
    int64_t end_row, end_col; /* 0-based; -1 when score == 0                         */
    int64_t start_row, start_col;
    int64_t m, n;             /* rows (A) and columns (B) after orientation          */
    char* cigar;              /* NUL-terminated run-length CIGAR, NULL if score-only */
    int64_t cigar_len;
    int64_t n_ops;            /* alignment columns                                   */
    swh_xpoint* xpoints;      /* start -> end, NULL unless want_xpoints             */
    int64_t n_xpoints;
    int64_t checkpoint_rows, checkpoint_cols;  /* K, Kc actually used            */
    double ms_forward;        /* stage 1 (forward pass with checkpoints)          */
    double ms_end_search;     /* end-cell column search                           */
    double ms_traceback;      /* block recomputation + walk                       */
    double ms_total;          /* device work of the call                          */
    int32_t precision_used;   /* 32 or 16                                        */
    int32_t forward_launches; /* kernel launches of the stage-1 sweep             */
    int32_t score_codes;      /* 1: the int32 sweeps took substitution scores from
                                 per-row symbol-code profiles; 0: from byte compares
                                 (always 0 from the request server)               */
    int32_t forward_tile_cols; /* columns per stage-1 tile (segment of a skewed stripe) */
} swh_result;

typedef struct swh_ctx swh_ctx;

const char* swh_strerror(int status);

/* Context = one device + one stream + cached device buffers.  Replaces the
 * reference's device discovery and per-run buffers (src/SWutils.cpp:300-338,
 * src/SWutils.cpp:144-173).  One context is not re-entrant. */
int swh_ctx_create(int device, swh_ctx** out);
int swh_ctx_destroy(swh_ctx* ctx);
/* Device name, CU count and clock (kHz) of the context's device. */
int swh_device_info(swh_ctx* ctx, char* name, int64_t name_cap, int32_t* cus, int32_t* clock_khz,
                    int64_t* hbm_bytes);

/* Full alignment of two host sequences: stage 1 forward with checkpoints, end-cell
 * search and traceback.  Replaces the stage-1 driver and the crosspoint search of
 * src/main.cu:158-248 and src/Crosspointer.cpp:104-338. */
int swh_align_pair(swh_ctx* ctx, const uint8_t* seq1, int64_t len1, const uint8_t* seq2,
                   int64_t len2, const swh_scoring* sc, const swh_opts* opts, swh_result* out);

/* Same, for sequences already resident in device memory (d_seq1/d_seq2 are device
 * pointers on the context's device).  No host<->device copy of the inputs. */
int swh_align_pair_device(swh_ctx* ctx, const uint8_t* d_seq1, int64_t len1, const uint8_t* d_seq2,
                          int64_t len2, const swh_scoring* sc, const swh_opts* opts,
                          swh_result* out);

void swh_result_free(swh_result* r);

/* Allocation-free variant of swh_align_pair_device.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Every device buffer beyond the inputs
 * is a slice of a static device array of SWH_NOALLOC_SCRATCH_BYTES held by each device the
 * library runs on (its address is resolved once per device by swh_ctx_create); the call
 * makes no hipMalloc, hipHostMalloc or hipFree.  Limit: the pair's working set (row maxima,
 * checkpoints, traceback blocks) must fit that array -- about 100 kbp x 100 kbp at the
 * default checkpoint spacing, which is doubled from swh_align_pair_device's default until it
 * fits; larger pairs, or a given spacing that does not fit, return SWH_ERR_SCRATCH before
 * any work.  Options want_xpoints and host_checkpoints are rejected (SWH_ERR_INVALID_ARG).
 * The CIGAR is written NUL-terminated to `cigar`, which must hold at least
 * 3 * (len1 + len2) + 64 bytes (its tail stages the traceback); out->cigar stays NULL and
 * out->cigar_len is the CIGAR's length (swh_result_free on `out` is harmless).  For the same
 * checkpoint spacing the result equals swh_align_pair_device's.  All calls on one device
 * share its array: calls that overlap on one device race and must be serialised. */
#define SWH_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int swh_align_pair_device_noalloc(swh_ctx* ctx, const uint8_t* d_seq1, int64_t len1, const uint8_t* d_seq2,
                                  int64_t len2, const swh_scoring* sc, const swh_opts* opts, swh_result* out,
                                  char* cigar, int64_t cigar_cap);

/* ---- stage-level entry points (tests and tools) ------------------------------------ */

/* Stage 1 only: score, end cell and per-row maxima; optionally copies the checkpoint
 * rows {H,F} (rows K-1, 2K-1, ... < m) and checkpoint columns {H,E} (columns Kc-1,
 * 2Kc-1, ... < n) to host buffers.  The analogue of shortPhase/longPhase + RowBuilder
 * (src/FindAlignment.cuh:9-278, src/Builders.cpp:27-104).  A/B are already oriented
 * (m rows, n columns).  rowmax: m int32 (nullable); srow: (m/K)*n*2 int32 (nullable);
 * scol: (n/Kc)*m*2 int32 (nullable). */
int swh_forward(swh_ctx* ctx, const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                const swh_scoring* sc, int64_t K, int64_t Kc, int32_t* score, int64_t* end_row,
                int64_t* end_col, int32_t* rowmax, int32_t* srow, int32_t* scol);

/* ---- batched pairs ------------------------------------------------------------------ */
typedef struct {
    const uint8_t* seq1;
    int64_t len1;
    const uint8_t* seq2;
    int64_t len2;
} swh_pair;

/* Aligns npairs pairs of uneven length together: every sweep (stage 1, end search, each
 * traceback round) covers all pairs side by side, one kernel launch per tile
 * anti-diagonal, so short pairs fill the device while long ones proceed.  out[npairs]
 * receives one result each (free each with swh_result_free); per-pair errors are
 * reported in statuses[npairs] (nullable) without failing the batch.  Checkpoints are
 * spaced alike for the whole batch (from blocks of 256 rows by 512 columns, doubled to fit
 * the budget); stage
 * times in the results are those of the whole batch.  No reference counterpart (the
 * reference aligns one pair per process, src/main.cu:102-250). */
int swh_batch_align(swh_ctx* ctx, const swh_pair* pairs, int64_t npairs, const swh_scoring* sc,
                    const swh_opts* opts, swh_result* out, int32_t* statuses);

/* Same, for pairs whose sequences are already in device memory (seq1/seq2 are device
 * pointers on the context's device). */
int swh_batch_align_device(swh_ctx* ctx, const swh_pair* pairs, int64_t npairs, const swh_scoring* sc,
                           const swh_opts* opts, swh_result* out, int32_t* statuses);

/* Batch autotuner: times every combination of tile width {256, 384, 512, 768, 1024},
 * sweep waves per SIMD {2, 3, 4} and checkpoint blocks (rows x columns) {256 x 512,
 * 512 x 512, 1024 x 1024, 2048 x 2048} on an evenly spaced
 * sample of `sample` pairs (device-resident pairs if `device` != 0) and reports the
 * fastest, to be passed back in swh_opts (tile_cols, waves_per_simd, checkpoint_rows and
 * checkpoint_cols).  Spacings whose checkpoints for the whole batch exceed hbm_budget_bytes
 * (default half the free HBM), and candidates that run out of memory, are skipped
 * (SWH_ERR_BUDGET if none is left).  swh_opts.autotune = 1 makes the batch entry points do
 * this first. */
typedef struct {
    int32_t tile_cols, waves_per_simd;
    int64_t checkpoint_rows;
    double ms_best, ms_default;   /* sample time of the best and of the default settings */
    int32_t candidates, _pad;
    int64_t checkpoint_cols;
} swh_batch_tuning;
int swh_batch_tune(swh_ctx* ctx, const swh_pair* pairs, int64_t npairs, int32_t device, const swh_scoring* sc,
                   const swh_opts* opts, int64_t sample, swh_batch_tuning* out);

/* ---- request server: continuous batching --------------------------------------------- */
/* A server owns a context on `device` and a pool of hbm_budget_bytes of device memory
 * (0 = half the free memory).  Jobs are submitted at any time from any thread; a server
 * thread admits them while the pool has room for their working memory (sequences, row
 * maxima, checkpoints, tile hand-off buffers, direction codes) and at most max_in_flight
 * (0 = 1024) are in flight, then advances all of them together: every kernel launch
 * covers the current tile anti-diagonal of every job in flight, so jobs that arrive while
 * others run are admitted between launches and share them.  Results are bit-identical to
 * swh_align_pair's.  No reference counterpart (the reference aligns one pair per process,
 * src/main.cu:102-250). */
typedef struct swh_server swh_server;

typedef struct {
    int64_t submitted, completed, failed, in_flight, queued;
    int64_t cells_completed;
    double gcups;             /* completed cells / (last completion - first submission)      */
    double p50_ms, p99_ms, max_ms;  /* job latency: submission to result ready            */
    int64_t hbm_pool_bytes, hbm_in_use_bytes;
    int32_t loop_status;      /* SWH_OK, or the runtime failure that stopped the server       */
    int32_t _pad;
    int64_t peak_in_flight;       /* most jobs in flight at once                                */
    int64_t admitted_mid_flight;  /* jobs admitted while another job in flight was mid-way      */
    int64_t sweep_launches;       /* sweep kernel launches                                      */
    int64_t shared_launches;      /* of those, launches holding tiles of two or more jobs       */
} swh_serve_stats;

int swh_server_create(int device, const swh_scoring* sc, const swh_opts* opts, int64_t hbm_budget_bytes,
                      int32_t max_in_flight, swh_server** out);
/* Copies the inputs and queues the job; *job_id identifies it for swh_server_poll. */
int swh_server_submit(swh_server* s, const uint8_t* seq1, int64_t len1, const uint8_t* seq2,
                      int64_t len2, int64_t* job_id);
/* SWH_PENDING if the job is still running and wait == 0; otherwise blocks until it is
 * done, moves its result to *out (free with swh_result_free) and returns its status
 * (SWH_ERR_BUDGET if it can never fit the pool).  out->ms_total is its latency. */
int swh_server_poll(swh_server* s, int64_t job_id, int32_t wait, swh_result* out);
int swh_server_stats(swh_server* s, swh_serve_stats* st);
/* Finishes every queued and running job, then releases the server. */
int swh_server_destroy(swh_server* s);

/* ---- synthetic workloads ------------------------------------------------------------ */
/* Generates the synthetic pair used by the benchmarks (seeded std::mt19937_64):
 * two uniform (or GC-biased) sequences of lengths len1, len2 sharing a homologous
 * region of length hom_len with substitution rate p_sub and indel rate p_indel
 * (geometric lengths, mean 3).  gc < 0 means uniform ACGT; repeat_frac > 0 plants
 * tandem repeats (period 2..50) over that fraction of each sequence.  Buffers are
 * allocated by the library; free with swh_free. */
int swh_synth_pair(uint64_t seed, int64_t len1, int64_t len2, int64_t hom_len, double p_sub,
                   double p_indel, double gc, double repeat_frac, uint8_t** seq1, uint8_t** seq2);
/* Mutates `src` (substitutions p_sub, indels p_indel) and trims/extends both flanks by
 * up to flank_frac of the length: the query model of the batched workload. */
int swh_synth_mutate(uint64_t seed, const uint8_t* src, int64_t len, double p_sub, double p_indel,
                     double flank_frac, uint8_t** out, int64_t* out_len);
void swh_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
