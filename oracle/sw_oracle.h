/*
 * sw_oracle.h -- CPU oracle for the SWalign hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain C++ restatement of the reference's Smith-Waterman-Gotoh method
 * (Sterbic/Projekt-2012, "SWalign" v0.6.3).  It exists so that tests, the smoke
 * check and the CPU baseline leg of the benchmark can compare the HIP library's
 * results with an independent computation.  Nothing in the product library or
 * the SWalign CLI links, loads or calls this code.
 *
 * Conventions (shared with the HIP library):
 *   A = rows (the longer sequence), B = columns (the shorter one), both raw bytes;
 *   s(a,b) = match if a == b (byte equality) else mismatch;
 *   a gap of length L costs gap_open + (L-1)*gap_extend;
 *   boundaries H = E = F = 0 on row -1 and column -1 (reference convention);
 *   end cell  = the first cell in row-major order with the maximal H;
 *   traceback = greedy backward walk with priorities
 *               H: zero(stop) > diagonal > F (vertical) > E (horizontal),
 *               E/F: open > extend (extend only when strictly better).
 *
 * Parity status of each entry point is documented next to its declaration.
 */
#ifndef SW_ORACLE_H_
#define SW_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t match;       /* >= 1                         */
    int32_t mismatch;    /* <= -1                        */
    int32_t gap_open;    /* <= -1, cost of the first gap base ("first") */
    int32_t gap_extend;  /* <= -1, cost of every further gap base        */
} orc_scoring;

typedef struct {
    int32_t score;
    int32_t _pad;
    int64_t end_i, end_j;      /* 0-based, rows in A, columns in B; -1 if score == 0 */
    int64_t start_i, start_j;  /* 0-based first aligned pair; -1 if score == 0        */
    int64_t cigar_len;         /* characters written to the cigar buffer (no NUL)     */
    int64_t n_ops;             /* alignment columns (=,X,I,D)                         */
} orc_alignment;

typedef struct {
    int64_t row, col;   /* path node on a checkpoint row (last cell of the path on that row) */
    int32_t score;      /* forward value of the path at that node                              */
    int32_t layer;      /* 0 = H, 1 = E, 2 = F : state of the path at that node                 */
} orc_xpoint;

/* Forward pass, linear space, single thread (reference recurrence FindAlignment.cuh:50-62).
 * rowmax (nullable, length m) receives max_j H(i,j). Pinned by KATs 1-3, closed forms,
 * brute force and the textbook restatement. */
int orc_forward(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                const orc_scoring* sc, int32_t* score, int64_t* end_i, int64_t* end_j,
                int32_t* rowmax);

/* Same result, tiled anti-diagonal wavefront over `threads` host threads.  end_j may be
 * NULL (here and in orc_forward): the score, end row and row maxima are then computed
 * without the (single-thread) end-column search. */
int orc_forward_mt(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                   const orc_scoring* sc, int threads, int32_t* score, int64_t* end_i,
                   int64_t* end_j, int32_t* rowmax);

/* Textbook Gotoh local alignment with E/F = -infinity on the boundaries.  Used only to
 * pin that the reference's zero E/F boundary leaves every H value unchanged. */
int orc_forward_textbook_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                              const orc_scoring* sc, int32_t* H /* m*n */);

/* Reference-convention full matrices (H, E, F each m*n, row-major). */
int orc_forward_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                     const orc_scoring* sc, int32_t* H, int32_t* E, int32_t* F);

/* Exhaustive enumeration of every local alignment (tiny inputs only, m,n <= 7). */
int orc_bruteforce_score(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                         const orc_scoring* sc, int32_t* score);

/* Full-matrix alignment with the canonical traceback (O(mn) bytes of direction codes). */
int orc_align_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                   const orc_scoring* sc, orc_alignment* out, char* cigar, int64_t cigar_cap);

/* Linear-space alignment by the checkpoint method: checkpoint rows every K rows,
 * checkpoint columns every Kc columns, exact block recomputation along the path.
 * Must equal orc_align_full for every K, Kc.  xp (nullable) receives the path nodes
 * on checkpoint rows, start -> end. */
int orc_align_linear(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                     const orc_scoring* sc, int64_t K, int64_t Kc, orc_alignment* out,
                     char* cigar, int64_t cigar_cap, orc_xpoint* xp, int64_t xp_cap,
                     int64_t* n_xp);

/* orc_align_linear with its stage 1 (checkpoints and row maxima) on `threads` host threads
 * (tiled wavefront as in orc_forward_mt when K is a multiple of 256 and Kc a multiple of
 * min(2048, Kc), else one thread); the end search and traceback are orc_align_linear's.
 * Same outputs as orc_align_linear, checked against it on the CPU. */
int orc_align_linear_mt(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                        const orc_scoring* sc, int64_t K, int64_t Kc, int threads,
                        orc_alignment* out, char* cigar, int64_t cigar_cap, orc_xpoint* xp,
                        int64_t xp_cap, int64_t* n_xp);

/* Checkpoint rows (H,F) of rows K-1, 2K-1, ... (< m) and checkpoint columns (H,E) of
 * columns Kc-1, 2Kc-1, ... (< n).  srow: nsr*n*2 int32, scol: nsc*m*2 int32. */
int orc_checkpoints(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                    const orc_scoring* sc, int64_t K, int64_t Kc, int32_t* srow,
                    int32_t* scol);

/* Reverse global DP from the end cell (end_i,end_j) with value S up to checkpoint row y
 * (Myers-Miller midpoint algebra as in SWutils.cpp:175-231): for every column split c,
 * scr(c) = H_f(y,c) + RR(y+1,c+1) and aff(c) = F_f(y,c) + SS(y+1,c+1) - open + extend.
 * Writes the maximum over c of both into *best and the value for the given split
 * (cross_c, cross_type 0 = diagonal/open crossing, 1 = gap joined across the row)
 * into *at_cross. */
int orc_crosspoint_algebra(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                           const orc_scoring* sc, int64_t y, int64_t end_i, int64_t end_j,
                           int64_t cross_c, int32_t cross_type, int32_t* best,
                           int32_t* at_cross);

/* Re-score an alignment given by start cell + CIGAR (=,X,I,D,M runs). */
int orc_rescore_cigar(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                      const orc_scoring* sc, int64_t start_i, int64_t start_j,
                      const char* cigar, int32_t* score, int64_t* end_i, int64_t* end_j);

/* Reference-compatible single-record FASTA load (FASTA.cpp:39-86): the first line must
 * start with '>', the name is that line minus '>' and the newline, and the sequence is
 * every following byte except '\n' and '\r'.  Returns 0 on success. Caller frees with
 * orc_free. */
int orc_fasta_load(const char* path, char** name, uint8_t** seq, int64_t* len);
void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
