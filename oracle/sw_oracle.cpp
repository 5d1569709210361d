/*
 * sw_oracle.cpp -- CPU oracle for the SWalign hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * library, and only as the checker.  The HIP product library never links it.
 *
 * Reference method restated here (file:line in Sterbic/Projekt-2012):
 *   cell update        src/FindAlignment.cuh:50-56 (E, F, H with the 0 clamp)
 *   strict-> max       src/FindAlignment.cuh:58-62, src/SWutils.cpp:286-298
 *   zero boundaries    src/SWutils.cpp:131-136 (cudaMemset 0), src/SWkernel.cuh:108-128
 *   orientation        src/SWutils.cpp:84-88 (rows = longer sequence; done by callers)
 *   checkpoint rows    src/Builders.cpp:27-104 ({H,F} of every special row)
 *   crosspoint algebra src/SWutils.cpp:175-231 (scr = H_f + H_r, aff = F_f + F_r - O + E)
 *   FASTA parsing      src/FASTA.cpp:39-86
 * The canonical tie rules (end cell, traceback priorities) are this build's own: the
 * reference's argmax order depends on its launch configuration and it never emits an
 * alignment.  They are stated in sw_oracle.h.
 */
#include "sw_oracle.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

inline int32_t max2(int32_t a, int32_t b) { return a > b ? a : b; }

struct Sc {
    int32_t M, X, O, E;
    explicit Sc(const orc_scoring* s) : M(s->match), X(s->mismatch), O(s->gap_open), E(s->gap_extend) {}
};

/* Direction code of one cell (same encoding as the HIP library):
 *   bits 0-1  H source: 0 = H is zero (alignment starts after this cell),
 *                       1 = diagonal, 2 = F (vertical gap), 3 = E (horizontal gap)
 *   bit 2     E(i,j) extends E(i,j-1) (else it opens from H(i,j-1))
 *   bit 3     F(i,j) extends F(i-1,j) (else it opens from H(i-1,j))            */
inline uint8_t dir_code(int32_t H, int32_t Hd, int32_t F, int32_t e_ext, int32_t f_ext) {
    uint8_t h;
    if (H == 0) h = 0;
    else if (H == Hd) h = 1;
    else if (H == F) h = 2;
    else h = 3;
    return (uint8_t)(h | (e_ext << 2) | (f_ext << 3));
}

/* One rectangle [r0,r1] x [c0,c1] of the forward recurrence (FindAlignment.cuh:50-56)
 * with explicit boundaries:
 *   topH/topF[c - c0]   = H,F of row r0-1,    leftH/leftE[r - r0] = H,E of column c0-1,
 *   corner              = H(r0-1, c0-1).
 * On return topH/topF hold row r1 and leftH/leftE hold column c1.
 * Optional outputs: rowmax[r - r0] (max over the rectangle's columns, merged with max),
 * dirs (row-major (r1-r0+1) x (c1-c0+1) direction codes), lastrow (H of row r1). */
void rect_forward(const uint8_t* A, const uint8_t* B, const Sc& s, int64_t r0, int64_t r1,
                  int64_t c0, int64_t c1, int32_t* topH, int32_t* topF, int32_t* leftH,
                  int32_t* leftE, int32_t corner, int32_t* rowmax, uint8_t* dirs) {
    const int64_t w = c1 - c0 + 1;
    int32_t left_above = corner;  // H(r-1, c0-1) before leftH[r-1] is overwritten
    for (int64_t r = r0; r <= r1; ++r) {
        const uint8_t a = A[r];
        int32_t Hdiag = left_above;
        left_above = leftH[r - r0];
        int32_t Hleft = leftH[r - r0];
        int32_t Eleft = leftE[r - r0];
        int32_t rmax = rowmax ? rowmax[r - r0] : 0;
        uint8_t* drow = dirs ? dirs + (r - r0) * w : nullptr;
        for (int64_t k = 0; k < w; ++k) {
            const int32_t Hup = topH[k], Fup = topF[k];
            const int32_t eext = Eleft + s.E, eopen = Hleft + s.O;
            const int32_t fext = Fup + s.E, fopen = Hup + s.O;
            const int32_t E = max2(eext, eopen);
            const int32_t F = max2(fext, fopen);
            const int32_t Hd = Hdiag + (a == B[c0 + k] ? s.M : s.X);
            const int32_t H = max2(max2(0, E), max2(F, Hd));
            if (drow) drow[k] = dir_code(H, Hd, F, eext > eopen, fext > fopen);
            if (H > rmax) rmax = H;
            Hdiag = Hup;
            topH[k] = H;
            topF[k] = F;
            Hleft = H;
            Eleft = E;
        }
        if (rowmax) rowmax[r - r0] = rmax;
        leftH[r - r0] = Hleft;
        leftE[r - r0] = Eleft;
    }
}

struct Alignment {
    int32_t score = 0;
    int64_t end_i = -1, end_j = -1, start_i = -1, start_j = -1;
    std::string ops;  // reversed op sequence while walking, then forward
    std::vector<orc_xpoint> xp;
};

void finish_cigar(Alignment& al, orc_alignment* out, char* cigar, int64_t cap, int64_t* written) {
    std::reverse(al.ops.begin(), al.ops.end());
    std::string rle;
    for (size_t p = 0; p < al.ops.size();) {
        size_t q = p;
        while (q < al.ops.size() && al.ops[q] == al.ops[p]) ++q;
        rle += std::to_string(q - p);
        rle += al.ops[p];
        p = q;
    }
    out->score = al.score;
    out->end_i = al.end_i;
    out->end_j = al.end_j;
    out->start_i = al.start_i;
    out->start_j = al.start_j;
    out->n_ops = (int64_t)al.ops.size();
    out->cigar_len = (int64_t)rle.size();
    if (cigar && cap > 0) {
        int64_t nw = std::min<int64_t>(cap - 1, (int64_t)rle.size());
        memcpy(cigar, rle.data(), (size_t)nw);
        cigar[nw] = 0;
    }
    if (written) *written = (int64_t)rle.size();
}

/* Greedy backward walk over direction codes supplied by `dir(i,j)`; returns when the
 * position leaves the region where `inside(i,j)` holds or when the start is found.
 * State: (i, j, layer) with layer 0 = H, 1 = E, 2 = F. */
template <class DirFn, class InsideFn>
bool walk(const uint8_t* A, const uint8_t* B, int64_t& i, int64_t& j, int& layer, Alignment& al,
          DirFn dir, InsideFn inside) {
    while (true) {
        if (i < 0 || j < 0) return true;  // boundary (H = 0): start already recorded
        if (!inside(i, j)) return false;
        const uint8_t d = dir(i, j);
        if (layer == 0) {
            const int h = d & 3;
            if (h == 1) {
                al.ops.push_back(A[i] == B[j] ? '=' : 'X');
                al.start_i = i;
                al.start_j = j;
                --i;
                --j;
                if (i < 0 || j < 0) return true;
                if (!inside(i, j)) return false;  // next region decides (zero check there)
                if ((dir(i, j) & 3) == 0) return true;
            } else if (h == 2) {
                layer = 2;
            } else if (h == 3) {
                layer = 1;
            } else {
                return true;  // H == 0 where the walk stands: nothing more to align
            }
        } else if (layer == 2) {
            al.ops.push_back('D');
            layer = ((d >> 3) & 1) ? 2 : 0;
            --i;
        } else {
            al.ops.push_back('I');
            layer = ((d >> 2) & 1) ? 1 : 0;
            --j;
        }
    }
}

}  // namespace

extern "C" {

int orc_forward(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                int32_t* score, int64_t* end_i, int64_t* end_j, int32_t* rowmax) {
    if (m < 0 || n < 0) return 1;
    Sc s(sc);
    std::vector<int32_t> H(n, 0), F(n, 0);
    int32_t best = 0;
    int64_t bi = -1, bj = -1;
    for (int64_t i = 0; i < m; ++i) {
        const uint8_t a = A[i];
        int32_t Hdiag = 0, Hleft = 0, Eleft = 0, rmax = 0;
        for (int64_t j = 0; j < n; ++j) {
            const int32_t Hup = H[j];
            const int32_t E = max2(Eleft + s.E, Hleft + s.O);
            const int32_t Fv = max2(F[j] + s.E, Hup + s.O);
            const int32_t Hd = Hdiag + (a == B[j] ? s.M : s.X);
            const int32_t Hv = max2(max2(0, E), max2(Fv, Hd));
            if (Hv > rmax) rmax = Hv;
            if (Hv > best) { best = Hv; bi = i; bj = j; }
            Hdiag = Hup;
            H[j] = Hv;
            F[j] = Fv;
            Hleft = Hv;
            Eleft = E;
        }
        if (rowmax) rowmax[i] = rmax;
    }
    *score = best;
    *end_i = bi;
    if (end_j) *end_j = bj;
    return 0;
}

int orc_forward_mt(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                   int threads, int32_t* score, int64_t* end_i, int64_t* end_j, int32_t* rowmax) {
    if (m < 0 || n < 0) return 1;
    if (threads <= 1 || m < 1024 || n < 1024)
        return orc_forward(A, m, B, n, sc, score, end_i, end_j, rowmax);
    Sc s(sc);
    const int64_t TR = 256, TC = 2048;
    const int64_t nrt = (m + TR - 1) / TR, nct = (n + TC - 1) / TC;
    std::vector<int32_t> bndH(n, 0), bndF(n, 0);  // bottom row of the last finished tile row
    std::vector<int32_t> rmax_all(m, 0);
    std::unique_ptr<std::atomic<int64_t>[]> prog(new std::atomic<int64_t>[nrt]);
    for (int64_t t = 0; t < nrt; ++t) prog[t].store(0);
    auto worker = [&](int tid) {
        std::vector<int32_t> lH(TR), lE(TR);
        for (int64_t ti = tid; ti < nrt; ti += threads) {
            const int64_t r0 = ti * TR, r1 = std::min(m, r0 + TR) - 1;
            std::fill(lH.begin(), lH.end(), 0);
            std::fill(lE.begin(), lE.end(), 0);
            int32_t corner = 0;  // H(r0-1, c0-1): left boundary column -1 is zero
            for (int64_t tj = 0; tj < nct; ++tj) {
                const int64_t c0 = tj * TC, c1 = std::min(n, c0 + TC) - 1;
                if (ti > 0)
                    while (prog[ti - 1].load(std::memory_order_acquire) < tj + 1) std::this_thread::yield();
                const int32_t next_corner = bndH[c1];  // H(r0-1, c1): corner of the next tile
                rect_forward(A, B, s, r0, r1, c0, c1, bndH.data() + c0, bndF.data() + c0, lH.data(),
                             lE.data(), corner, rmax_all.data() + r0, nullptr);
                corner = (ti > 0) ? next_corner : 0;
                prog[ti].store(tj + 1, std::memory_order_release);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    int32_t best = 0;
    int64_t bi = -1;
    for (int64_t i = 0; i < m; ++i)
        if (rmax_all[i] > best) { best = rmax_all[i]; bi = i; }
    int64_t bj = -1;
    if (best > 0 && end_j) {
        // smallest column of row bi holding the maximum: recompute rows 0..bi (linear space)
        std::vector<int32_t> H(n, 0), F(n, 0);
        for (int64_t i = 0; i <= bi && bj < 0; ++i) {
            const uint8_t a = A[i];
            int32_t Hdiag = 0, Hleft = 0, Eleft = 0;
            for (int64_t j = 0; j < n; ++j) {
                const int32_t Hup = H[j];
                const int32_t E = max2(Eleft + s.E, Hleft + s.O);
                const int32_t Fv = max2(F[j] + s.E, Hup + s.O);
                const int32_t Hd = Hdiag + (a == B[j] ? s.M : s.X);
                const int32_t Hv = max2(max2(0, E), max2(Fv, Hd));
                if (i == bi && Hv == best) { bj = j; break; }
                Hdiag = Hup; H[j] = Hv; F[j] = Fv; Hleft = Hv; Eleft = E;
            }
        }
    }
    if (rowmax) memcpy(rowmax, rmax_all.data(), sizeof(int32_t) * m);
    *score = best;
    *end_i = bi;
    if (end_j) *end_j = bj;
    return 0;
}

int orc_forward_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                     int32_t* H, int32_t* E, int32_t* F) {
    Sc s(sc);
    auto at = [&](int32_t* X, int64_t i, int64_t j) -> int32_t {
        return (i < 0 || j < 0) ? 0 : X[i * n + j];
    };
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            const int32_t e = max2(at(E, i, j - 1) + s.E, at(H, i, j - 1) + s.O);
            const int32_t f = max2(at(F, i - 1, j) + s.E, at(H, i - 1, j) + s.O);
            const int32_t d = at(H, i - 1, j - 1) + (A[i] == B[j] ? s.M : s.X);
            E[i * n + j] = e;
            F[i * n + j] = f;
            H[i * n + j] = max2(max2(0, e), max2(f, d));
        }
    return 0;
}

int orc_forward_textbook_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                              const orc_scoring* sc, int32_t* H) {
    Sc s(sc);
    const int32_t NEG = -(1 << 28);
    std::vector<int32_t> E(m * n), F(m * n);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            const int32_t hl = j ? H[i * n + j - 1] : 0, el = j ? E[i * n + j - 1] : NEG;
            const int32_t hu = i ? H[(i - 1) * n + j] : 0, fu = i ? F[(i - 1) * n + j] : NEG;
            const int32_t hd = (i && j) ? H[(i - 1) * n + j - 1] : 0;
            const int32_t e = max2(el + s.E, hl + s.O);
            const int32_t f = max2(fu + s.E, hu + s.O);
            E[i * n + j] = e;
            F[i * n + j] = f;
            H[i * n + j] = max2(max2(0, e), max2(f, hd + (A[i] == B[j] ? s.M : s.X)));
        }
    return 0;
}

/* Exhaustive: every pair of non-empty substrings, every global alignment between them
 * (ops =/X, I, D; a maximal run of I or of D costs open + (L-1)*extend). */
int orc_bruteforce_score(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                         const orc_scoring* sc, int32_t* score) {
    if (m > 7 || n > 7) return 1;
    Sc s(sc);
    int32_t best = 0;
    // rec(i, j, last, val): aligned A[a..i), B[c..j); last op 0 = pair, 1 = I, 2 = D
    struct R {
        const uint8_t *A, *B;
        const Sc* s;
        int64_t ie, je;
        int32_t* best;
        void go(int64_t i, int64_t j, int last, int32_t val) {
            if (i == ie && j == je) {
                if (last == 0 && val > *best) *best = val;  // must end on an aligned pair
                return;
            }
            if (i < ie && j < je) go(i + 1, j + 1, 0, val + (A[i] == B[j] ? s->M : s->X));
            // the recurrence takes max(E + extend, H + open) with H >= E, so a gap base
            // after another one in the same direction costs max(extend, open)
            const int32_t ext = s->E > s->O ? s->E : s->O;
            if (j < je) go(i, j + 1, 1, val + (last == 1 ? ext : s->O));
            if (i < ie) go(i + 1, j, 2, val + (last == 2 ? ext : s->O));
        }
    };
    for (int64_t a = 0; a < m; ++a)
        for (int64_t ie = a + 1; ie <= m; ++ie)
            for (int64_t c = 0; c < n; ++c)
                for (int64_t je = c + 1; je <= n; ++je) {
                    R r{A, B, &s, ie, je, &best};
                    // must start on an aligned pair as well
                    r.go(a + 1, c + 1, 0, (A[a] == B[c] ? s.M : s.X));
                }
    *score = best;
    return 0;
}

int orc_align_full(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                   orc_alignment* out, char* cigar, int64_t cigar_cap) {
    Sc s(sc);
    Alignment al;
    std::vector<uint8_t> dirs((size_t)(m * n));
    std::vector<int32_t> topH(n, 0), topF(n, 0), leftH(m, 0), leftE(m, 0), rmax(m, 0);
    if (m > 0 && n > 0)
        rect_forward(A, B, s, 0, m - 1, 0, n - 1, topH.data(), topF.data(), leftH.data(),
                     leftE.data(), 0, rmax.data(), dirs.data());
    int64_t bi = -1;
    int32_t best = 0;
    for (int64_t i = 0; i < m; ++i)
        if (rmax[i] > best) { best = rmax[i]; bi = i; }
    al.score = best;
    if (best > 0) {
        // smallest column in row bi with H == best: recompute that row's H from dirs is not
        // possible, so recompute the matrix rows 0..bi (small inputs only).
        std::vector<int32_t> H(n, 0), F(n, 0);
        int64_t bj = -1;
        for (int64_t i = 0; i <= bi && bj < 0; ++i) {
            int32_t Hdiag = 0, Hleft = 0, Eleft = 0;
            for (int64_t j = 0; j < n; ++j) {
                const int32_t Hup = H[j];
                const int32_t E = max2(Eleft + s.E, Hleft + s.O);
                const int32_t Fv = max2(F[j] + s.E, Hup + s.O);
                const int32_t Hd = Hdiag + (A[i] == B[j] ? s.M : s.X);
                const int32_t Hv = max2(max2(0, E), max2(Fv, Hd));
                if (i == bi && Hv == best) { bj = j; break; }
                Hdiag = Hup; H[j] = Hv; F[j] = Fv; Hleft = Hv; Eleft = E;
            }
        }
        al.end_i = bi;
        al.end_j = bj;
        int64_t i = bi, j = bj;
        int layer = 0;
        walk(A, B, i, j, layer, al, [&](int64_t r, int64_t c) { return dirs[r * n + c]; },
             [&](int64_t, int64_t) { return true; });
    }
    finish_cigar(al, out, cigar, cigar_cap, nullptr);
    return 0;
}

int orc_checkpoints(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                    int64_t K, int64_t Kc, int32_t* srow, int32_t* scol) {
    if (K < 1 || Kc < 1) return 1;
    Sc s(sc);
    std::vector<int32_t> H(n, 0), F(n, 0);
    for (int64_t i = 0; i < m; ++i) {
        const uint8_t a = A[i];
        int32_t Hdiag = 0, Hleft = 0, Eleft = 0;
        for (int64_t j = 0; j < n; ++j) {
            const int32_t Hup = H[j];
            const int32_t E = max2(Eleft + s.E, Hleft + s.O);
            const int32_t Fv = max2(F[j] + s.E, Hup + s.O);
            const int32_t Hd = Hdiag + (a == B[j] ? s.M : s.X);
            const int32_t Hv = max2(max2(0, E), max2(Fv, Hd));
            if ((j + 1) % Kc == 0 && scol) {
                int32_t* p = scol + ((j + 1) / Kc - 1) * m * 2 + i * 2;
                p[0] = Hv;
                p[1] = E;
            }
            Hdiag = Hup; H[j] = Hv; F[j] = Fv; Hleft = Hv; Eleft = E;
        }
        if ((i + 1) % K == 0 && srow) {
            int32_t* p = srow + ((i + 1) / K - 1) * n * 2;
            for (int64_t j = 0; j < n; ++j) { p[2 * j] = H[j]; p[2 * j + 1] = F[j]; }
        }
    }
    return 0;
}

}  // extern "C"

namespace {

/* Stage 1 of the linear-space method on `threads` host threads: checkpoint rows {H,F}
 * (rows K-1, 2K-1, ...), checkpoint columns {H,E} (columns Kc-1, 2Kc-1, ...) and row
 * maxima, by the tiled anti-diagonal wavefront of orc_forward_mt.  Tiles are 256 rows by
 * min(2048, Kc) columns, so every checkpoint row is a tile row's last row and every
 * checkpoint column a tile's last column (K a multiple of 256, Kc a multiple of the tile
 * width); other spacings run on one thread. */
void linear_stage1(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                   int64_t K, int64_t Kc, int threads, int32_t* srow, int32_t* scol, int32_t* rmax) {
    const int64_t TR = 256, TC = std::min<int64_t>(2048, Kc);
    if (threads <= 1 || m < 1024 || n < 1024 || K % TR != 0 || Kc % TC != 0) {
        orc_checkpoints(A, m, B, n, sc, K, Kc, srow, scol);
        int32_t sc0;
        int64_t ei, ej;
        orc_forward(A, m, B, n, sc, &sc0, &ei, &ej, rmax);
        return;
    }
    Sc s(sc);
    const int64_t nrt = (m + TR - 1) / TR, nct = (n + TC - 1) / TC;
    std::vector<int32_t> bndH(n, 0), bndF(n, 0);
    std::fill(rmax, rmax + m, 0);
    std::unique_ptr<std::atomic<int64_t>[]> prog(new std::atomic<int64_t>[nrt]);
    for (int64_t t = 0; t < nrt; ++t) prog[t].store(0);
    auto worker = [&](int tid) {
        std::vector<int32_t> lH(TR), lE(TR);
        for (int64_t ti = tid; ti < nrt; ti += threads) {
            const int64_t r0 = ti * TR, r1 = std::min(m, r0 + TR) - 1;
            std::fill(lH.begin(), lH.end(), 0);
            std::fill(lE.begin(), lE.end(), 0);
            int32_t corner = 0;
            for (int64_t tj = 0; tj < nct; ++tj) {
                const int64_t c0 = tj * TC, c1 = std::min(n, c0 + TC) - 1;
                if (ti > 0)
                    while (prog[ti - 1].load(std::memory_order_acquire) < tj + 1) std::this_thread::yield();
                const int32_t next_corner = bndH[c1];
                rect_forward(A, B, s, r0, r1, c0, c1, bndH.data() + c0, bndF.data() + c0, lH.data(),
                             lE.data(), corner, rmax + r0, nullptr);
                if ((r1 + 1) % K == 0 && srow) {  // checkpoint row r1, columns c0..c1
                    int32_t* p = srow + ((r1 + 1) / K - 1) * n * 2;
                    for (int64_t c = c0; c <= c1; ++c) { p[2 * c] = bndH[c]; p[2 * c + 1] = bndF[c]; }
                }
                if ((c1 + 1) % Kc == 0 && scol) {  // checkpoint column c1, rows r0..r1
                    int32_t* p = scol + ((c1 + 1) / Kc - 1) * m * 2;
                    for (int64_t r = r0; r <= r1; ++r) { p[2 * r] = lH[r - r0]; p[2 * r + 1] = lE[r - r0]; }
                }
                corner = (ti > 0) ? next_corner : 0;
                prog[ti].store(tj + 1, std::memory_order_release);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
}

int align_linear_impl(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                      int64_t K, int64_t Kc, int threads, orc_alignment* out, char* cigar, int64_t cigar_cap,
                      orc_xpoint* xp, int64_t xp_cap, int64_t* n_xp) {
    if (K < 1 || Kc < 1) return 1;
    Sc s(sc);
    Alignment al;
    const int64_t nsr = m / K, nsc = n / Kc;  // checkpoint rows K-1,2K-1,..; columns Kc-1,..
    std::vector<int32_t> srow((size_t)(nsr * n * 2)), scol((size_t)(nsc * m * 2)), rmax(m, 0);
    // Stage 1: forward pass writing checkpoints and row maxima.
    linear_stage1(A, m, B, n, sc, K, Kc, threads, srow.data(), scol.data(), rmax.data());
    int32_t best = 0;
    int64_t bi = -1;
    for (int64_t i = 0; i < m; ++i)
        if (rmax[i] > best) { best = rmax[i]; bi = i; }
    al.score = best;
    // Boundary accessors for block (bk, bq): rows [bk*K, ...], columns [bq*Kc, ...].
    auto load_bounds = [&](int64_t r0, int64_t r1, int64_t c0, int64_t c1, std::vector<int32_t>& tH,
                           std::vector<int32_t>& tF, std::vector<int32_t>& lH, std::vector<int32_t>& lE,
                           int32_t& corner) {
        tH.assign(c1 - c0 + 1, 0);
        tF.assign(c1 - c0 + 1, 0);
        lH.assign(r1 - r0 + 1, 0);
        lE.assign(r1 - r0 + 1, 0);
        corner = 0;
        if (r0 > 0) {
            const int32_t* p = srow.data() + (r0 / K - 1) * n * 2;
            for (int64_t c = c0; c <= c1; ++c) { tH[c - c0] = p[2 * c]; tF[c - c0] = p[2 * c + 1]; }
            if (c0 > 0) corner = p[2 * (c0 - 1)];
        }
        if (c0 > 0) {
            const int32_t* p = scol.data() + (c0 / Kc - 1) * m * 2;
            for (int64_t r = r0; r <= r1; ++r) { lH[r - r0] = p[2 * r]; lE[r - r0] = p[2 * r + 1]; }
        }
    };
    std::vector<int32_t> tH, tF, lH, lE;
    int32_t corner;
    if (best > 0) {
        // End cell: smallest column of row bi holding `best`, from exact block recomputes of
        // rows [band start, bi] scanned left to right.
        const int64_t r0 = (bi / K) * K;
        int64_t bj = -1;
        for (int64_t c0 = 0; c0 < n && bj < 0; c0 += Kc) {
            const int64_t c1 = std::min(n, c0 + Kc) - 1;
            load_bounds(r0, bi, c0, c1, tH, tF, lH, lE, corner);
            rect_forward(A, B, s, r0, bi, c0, c1, tH.data(), tF.data(), lH.data(), lE.data(), corner,
                         nullptr, nullptr);
            for (int64_t c = c0; c <= c1; ++c)
                if (tH[c - c0] == best) { bj = c; break; }
        }
        al.end_i = bi;
        al.end_j = bj;
        // Traceback block by block.
        int64_t i = bi, j = bj;
        int layer = 0;
        std::vector<uint8_t> dirs;
        int64_t prev_row = i;
        al.xp.push_back(orc_xpoint{bi, bj, best, 0});
        while (true) {
            const int64_t r0 = (i / K) * K, c0 = (j / Kc) * Kc;
            const int64_t r1 = i, c1 = j;  // only the part up-left of the entry is needed
            load_bounds(r0, r1, c0, c1, tH, tF, lH, lE, corner);
            const int64_t w = c1 - c0 + 1;
            dirs.assign((size_t)((r1 - r0 + 1) * w), 0);
            rect_forward(A, B, s, r0, r1, c0, c1, tH.data(), tF.data(), lH.data(), lE.data(), corner,
                         nullptr, dirs.data());
            // walk, recording the first node reached on each checkpoint row
            auto dir = [&](int64_t r, int64_t c) -> uint8_t {
                if (r != prev_row) {
                    prev_row = r;
                    if ((r + 1) % K == 0) {
                        const int32_t* p = srow.data() + ((r + 1) / K - 1) * n * 2;
                        al.xp.push_back(orc_xpoint{r, c, layer == 2 ? p[2 * c + 1] : p[2 * c], layer});
                    }
                }
                return dirs[(r - r0) * w + (c - c0)];
            };
            auto inside = [&](int64_t r, int64_t c) { return r >= r0 && c >= c0; };
            if (walk(A, B, i, j, layer, al, dir, inside)) break;
        }
        std::reverse(al.xp.begin(), al.xp.end());
    }
    finish_cigar(al, out, cigar, cigar_cap, nullptr);
    if (n_xp) *n_xp = (int64_t)al.xp.size();
    if (xp)
        for (int64_t k = 0; k < (int64_t)al.xp.size() && k < xp_cap; ++k) xp[k] = al.xp[k];
    return 0;
}

}  // namespace

extern "C" {

int orc_align_linear(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                     int64_t K, int64_t Kc, orc_alignment* out, char* cigar, int64_t cigar_cap,
                     orc_xpoint* xp, int64_t xp_cap, int64_t* n_xp) {
    return align_linear_impl(A, m, B, n, sc, K, Kc, 1, out, cigar, cigar_cap, xp, xp_cap, n_xp);
}

int orc_align_linear_mt(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                        int64_t K, int64_t Kc, int threads, orc_alignment* out, char* cigar,
                        int64_t cigar_cap, orc_xpoint* xp, int64_t xp_cap, int64_t* n_xp) {
    return align_linear_impl(A, m, B, n, sc, K, Kc, threads, out, cigar, cigar_cap, xp, xp_cap, n_xp);
}

int orc_crosspoint_algebra(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n,
                           const orc_scoring* sc, int64_t y, int64_t end_i, int64_t end_j,
                           int64_t cross_c, int32_t cross_type, int32_t* best, int32_t* at_cross) {
    if (y < 0 || y >= end_i || end_i >= m || end_j >= n) return 1;
    Sc s(sc);
    const int32_t NEG = -(1 << 28);
    // Forward row y (H_f, F_f) for columns 0..end_j.
    std::vector<int32_t> Hf(n, 0), Ff(n, 0);
    for (int64_t i = 0; i <= y; ++i) {
        int32_t Hdiag = 0, Hleft = 0, Eleft = 0;
        for (int64_t j = 0; j < n; ++j) {
            const int32_t Hup = Hf[j];
            const int32_t E = max2(Eleft + s.E, Hleft + s.O);
            const int32_t Fv = max2(Ff[j] + s.E, Hup + s.O);
            const int32_t Hv = max2(max2(0, E), max2(Fv, Hdiag + (A[i] == B[j] ? s.M : s.X)));
            Hdiag = Hup; Hf[j] = Hv; Ff[j] = Fv; Hleft = Hv; Eleft = E;
        }
    }
    // Reverse global DP over rows y+1..end_i and column splits 0..end_j+1 (TraceBack.cuh
    // recurrence without the zero clamp, starting at the end cell, which closes on a pair):
    //   RR[x][c] best score aligning A[x..end_i] with B[c..end_j], last pair (end_i,end_j)
    //   SS[x][c] the same restricted to alignments that open with A[x] against a gap.
    const int64_t W = end_j + 2;  // c in [0, end_j+1]
    std::vector<int32_t> RR(W), SS(W), RRn(W), SSn(W);
    // row x = end_i + 1 (empty A suffix): only the empty alignment of the empty B suffix
    // is valid, and it is only valid as the "after the end pair" state.
    for (int64_t c = 0; c < W; ++c) { RRn[c] = NEG; SSn[c] = NEG; }
    for (int64_t x = end_i; x >= y + 1; --x) {
        int32_t TT = NEG;  // horizontal-gap-first value at (x, c+1) moving leftwards
        for (int64_t c = W - 1; c >= 0; --c) {
            int32_t rr, ss;
            if (c == W - 1) {
                // B suffix empty: only a vertical gap A[x..end_i], which cannot end on the
                // required final pair -> invalid.
                ss = NEG;
                rr = NEG;
                TT = NEG;
            } else {
                int32_t diag;
                if (x == end_i && c == end_j) diag = (A[x] == B[c] ? s.M : s.X);  // the end pair
                else if (x + 1 <= end_i && c + 1 <= end_j) diag = RRn[c + 1] > NEG / 2 ? RRn[c + 1] + (A[x] == B[c] ? s.M : s.X) : NEG;
                else diag = NEG;
                ss = NEG;
                if (x + 1 <= end_i) {
                    const int32_t a = SSn[c] > NEG / 2 ? SSn[c] + s.E : NEG;
                    const int32_t b = RRn[c] > NEG / 2 ? RRn[c] + s.O : NEG;
                    ss = max2(a, b);
                }
                // horizontal-gap-first: B[c] against a gap, then continue at (x, c+1)
                const int32_t tt_prev = TT;  // value at (x, c+1)
                const int32_t rr_right = RR[c + 1];
                int32_t tt = NEG;
                {
                    const int32_t a = tt_prev > NEG / 2 ? tt_prev + s.E : NEG;
                    const int32_t b = rr_right > NEG / 2 ? rr_right + s.O : NEG;
                    tt = max2(a, b);
                }
                TT = tt;
                rr = max2(diag, max2(ss, tt));
            }
            RR[c] = rr;
            SS[c] = ss;
        }
        std::swap(RR, RRn);
        std::swap(SS, SSn);
    }
    // RRn/SSn now hold row y+1.
    int32_t bst = NEG, atc = NEG;
    for (int64_t c = -1; c <= end_j; ++c) {
        const int32_t hf = c < 0 ? 0 : Hf[c];
        const int32_t ff = c < 0 ? 0 : Ff[c];
        const int32_t rr = RRn[c + 1], ss = SSn[c + 1];
        const int32_t scr = rr > NEG / 2 ? hf + rr : NEG;
        const int32_t aff = (ss > NEG / 2 && c >= 0) ? ff + ss - s.O + s.E : NEG;
        bst = max2(bst, max2(scr, aff));
        if (c == cross_c) atc = cross_type == 0 ? scr : aff;
    }
    *best = bst;
    *at_cross = atc;
    return 0;
}

int orc_rescore_cigar(const uint8_t* A, int64_t m, const uint8_t* B, int64_t n, const orc_scoring* sc,
                      int64_t start_i, int64_t start_j, const char* cigar, int32_t* score,
                      int64_t* end_i, int64_t* end_j) {
    Sc s(sc);
    // a run of L gap bases costs open + (L-1)*max(extend, open): when extending costs more
    // than opening, the recurrence re-opens at every base.
    const int32_t ext = s.E > s.O ? s.E : s.O;
    int64_t i = start_i, j = start_j;
    int32_t v = 0;
    char last = 0;
    for (const char* p = cigar; *p;) {
        int64_t cnt = 0;
        while (*p >= '0' && *p <= '9') cnt = cnt * 10 + (*p++ - '0');
        const char op = *p++;
        for (int64_t k = 0; k < cnt; ++k) {
            if (op == '=' || op == 'X' || op == 'M') {
                if (i >= m || j >= n) return 2;
                if (op == '=' && A[i] != B[j]) return 3;
                if (op == 'X' && A[i] == B[j]) return 3;
                v += A[i] == B[j] ? s.M : s.X;
                ++i; ++j;
                last = 'M';
            } else if (op == 'I') {
                if (j >= n) return 2;
                v += last == 'I' ? ext : s.O;
                ++j;
                last = 'I';
            } else if (op == 'D') {
                if (i >= m) return 2;
                v += last == 'D' ? ext : s.O;
                ++i;
                last = 'D';
            } else {
                return 4;
            }
        }
    }
    *score = v;
    *end_i = i - 1;
    *end_j = j - 1;
    return 0;
}

int orc_fasta_load(const char* path, char** name, uint8_t** seq, int64_t* len) {
    FILE* f = fopen(path, "rb");
    if (!f) return 1;
    std::string all;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) all.append(buf, k);
    fclose(f);
    if (all.empty() || all[0] != '>') return 2;
    // fgets(c, 200) semantics for the name: at most 199 bytes, up to and including '\n';
    // the name drops the leading '>' and the line's last byte (FASTA.cpp:45-56).
    auto fgets_len = [&](size_t cap) {
        size_t k = 0;
        while (k < all.size() && k < cap - 1) {
            if (all[k++] == '\n') break;
        }
        return k;
    };
    const size_t l200 = fgets_len(200);
    std::string nm = l200 >= 2 ? all.substr(1, l200 - 2) : std::string();
    // the sequence starts after a second fgets(c, 300) of the first line (FASTA.cpp:66)
    const size_t eol = fgets_len(300);
    uint8_t* out = (uint8_t*)malloc(all.size() + 1);
    int64_t n = 0;
    for (size_t p = eol; p < all.size(); ++p) {
        const char c = all[p];
        if (c == '\n' || c == '\r') continue;
        out[n++] = (uint8_t)c;
    }
    char* nmc = (char*)malloc(nm.size() + 1);
    memcpy(nmc, nm.data(), nm.size());
    nmc[nm.size()] = 0;
    *name = nmc;
    *seq = out;
    *len = n;
    return 0;
}

void orc_free(void* p) { free(p); }

}  // extern "C"
