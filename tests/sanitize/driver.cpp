// CPU sanitizer driver (test infrastructure only): runs the host-only C++ of this repository
// -- the CPU oracle (oracle/sw_oracle.cpp), the SWalign CLI's FASTA loader and alignment
// renderer (projekt-2012_amd/csrc/cli/sw_fasta.cpp, sw_visual.cpp) and the synthetic
// workload generator (projekt-2012_amd/csrc/sw_synth.cpp) -- built with AddressSanitizer
// and UndefinedBehaviorSanitizer (or ThreadSanitizer for the multithreaded oracle paths)
// so that an out-of-bounds access, a leak, signed overflow or a data race aborts the run.
//
// Besides running every entry point, it checks the oracle's functions against each other:
// linear-space alignment at several checkpoint spacings = full-matrix alignment, the
// multithreaded paths = the single-thread ones, brute force = forward score on tiny pairs,
// textbook Gotoh H = reference-convention H, CIGAR re-scores to the score, crosspoints meet
// the forward + reverse algebra, the CLI loader = the oracle's loader.
//
// usage: sanitize_driver <fasta-dir> <tmp-dir> [mt-only]
#include <dirent.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "sw_fasta.h"
#include "sw_oracle.h"
#include "sw_visual.h"
#include "swalign_hip.h"

static int g_fail = 0;
#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            fprintf(stderr, "%s:%d check failed: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

static std::vector<uint8_t> rand_seq(std::mt19937_64& r, int64_t len, int alpha) {
    static const char kSym[] = "ACGTN";
    std::vector<uint8_t> s((size_t)len);
    for (auto& c : s) c = (uint8_t)kSym[r() % (uint64_t)alpha];
    return s;
}

// b = a copy of a window of a with substitutions and short indels, between random flanks
static std::vector<uint8_t> related(std::mt19937_64& r, const std::vector<uint8_t>& a, int64_t len) {
    std::vector<uint8_t> b = rand_seq(r, len / 4, 4);
    if (!a.empty()) {
        const size_t o = r() % a.size(), w = std::min(a.size() - o, (size_t)len / 2 + 1);
        for (size_t k = o; k < o + w; ++k) {
            const uint64_t u = r() % 100;
            if (u < 8) b.push_back("ACGT"[r() % 4]);
            else if (u < 10) continue;
            else if (u < 12) { b.push_back(a[k]); b.push_back("ACGT"[r() % 4]); }
            else b.push_back(a[k]);
        }
    }
    std::vector<uint8_t> t = rand_seq(r, len / 4, 4);
    b.insert(b.end(), t.begin(), t.end());
    return b;
}

struct Aln {
    orc_alignment al{};
    std::string cigar;
    std::vector<orc_xpoint> xp;
};

static Aln align_full(const std::vector<uint8_t>& A, const std::vector<uint8_t>& B, const orc_scoring& sc) {
    Aln o;
    std::vector<char> cig(2 * (A.size() + B.size()) + 64);
    CHECK(orc_align_full(A.data(), (int64_t)A.size(), B.data(), (int64_t)B.size(), &sc, &o.al, cig.data(),
                         (int64_t)cig.size()) == 0);
    o.cigar.assign(cig.data(), (size_t)o.al.cigar_len);
    return o;
}

static Aln align_linear(const std::vector<uint8_t>& A, const std::vector<uint8_t>& B, const orc_scoring& sc,
                        int64_t K, int64_t Kc, int threads) {
    Aln o;
    std::vector<char> cig(2 * (A.size() + B.size()) + 64);
    o.xp.resize(A.size() / (size_t)K + 2);
    int64_t nxp = 0;
    const int rc = threads > 1
        ? orc_align_linear_mt(A.data(), (int64_t)A.size(), B.data(), (int64_t)B.size(), &sc, K, Kc, threads, &o.al,
                              cig.data(), (int64_t)cig.size(), o.xp.data(), (int64_t)o.xp.size(), &nxp)
        : orc_align_linear(A.data(), (int64_t)A.size(), B.data(), (int64_t)B.size(), &sc, K, Kc, &o.al, cig.data(),
                           (int64_t)cig.size(), o.xp.data(), (int64_t)o.xp.size(), &nxp);
    CHECK(rc == 0);
    o.cigar.assign(cig.data(), (size_t)o.al.cigar_len);
    o.xp.resize((size_t)nxp);
    return o;
}

static void same(const Aln& x, const Aln& y) {
    CHECK(x.al.score == y.al.score);
    CHECK(x.al.end_i == y.al.end_i && x.al.end_j == y.al.end_j);
    CHECK(x.al.start_i == y.al.start_i && x.al.start_j == y.al.start_j);
    CHECK(x.al.n_ops == y.al.n_ops && x.cigar == y.cigar);
}

// The single-thread oracle on one pair: every entry point, checked against each other.
static void oracle_pair(const std::vector<uint8_t>& A, const std::vector<uint8_t>& B, const orc_scoring& sc,
                        const char* tmp) {
    const int64_t m = (int64_t)A.size(), n = (int64_t)B.size();
    int32_t s1 = 0;
    int64_t ei = -1, ej = -1;
    std::vector<int32_t> rm((size_t)std::max<int64_t>(m, 1));
    CHECK(orc_forward(A.data(), m, B.data(), n, &sc, &s1, &ei, &ej, rm.data()) == 0);
    const Aln full = align_full(A, B, sc);
    CHECK(full.al.score == s1 && full.al.end_i == ei && full.al.end_j == ej);
    if (m * n <= 40000) {
        std::vector<int32_t> H((size_t)(m * n) + 1), E(H.size()), F(H.size()), T(H.size());
        CHECK(orc_forward_full(A.data(), m, B.data(), n, &sc, H.data(), E.data(), F.data()) == 0);
        CHECK(orc_forward_textbook_full(A.data(), m, B.data(), n, &sc, T.data()) == 0);
        for (int64_t k = 0; k < m * n; ++k) CHECK(H[(size_t)k] == T[(size_t)k]);
    }
    if (m <= 6 && n <= 6) {
        int32_t bf = -1;
        CHECK(orc_bruteforce_score(A.data(), m, B.data(), n, &sc, &bf) == 0);
        CHECK(bf == s1);
    }
    const int64_t sp[][2] = {{1, 1}, {3, 5}, {16, 16}, {64, 32}, {1 << 20, 1 << 20}};
    for (const auto& k : sp) {
        const Aln lin = align_linear(A, B, sc, k[0], k[1], 1);
        same(lin, full);
        // crosspoints lie on checkpoint rows and meet the forward + reverse algebra; the last
        // node is the end cell
        CHECK(lin.al.score == 0 || (!lin.xp.empty() && lin.xp.back().row == lin.al.end_i &&
                                    lin.xp.back().col == lin.al.end_j && lin.xp.back().score == lin.al.score));
        for (size_t q = 0; q + 1 < lin.xp.size(); ++q) {
            const orc_xpoint& x = lin.xp[q];
            CHECK((x.row + 1) % k[0] == 0 && x.layer >= 0 && x.layer <= 2);
            if (m * n > 200000) continue;
            int32_t best = 0, at = 0;
            CHECK(orc_crosspoint_algebra(A.data(), m, B.data(), n, &sc, x.row, lin.al.end_i, lin.al.end_j, x.col,
                                         x.layer == 2 ? 1 : 0, &best, &at) == 0);
            CHECK(best == lin.al.score && at == lin.al.score);
        }
    }
    if (s1 > 0) {
        const int64_t nsr = (m - 1) / 7 + 1, nsc = (n - 1) / 5 + 1;
        std::vector<int32_t> srow((size_t)(nsr * n * 2) + 2), scol((size_t)(nsc * m * 2) + 2);
        CHECK(orc_checkpoints(A.data(), m, B.data(), n, &sc, 7, 5, srow.data(), scol.data()) == 0);
        int32_t rs = 0;
        int64_t ri = 0, rj = 0;
        CHECK(orc_rescore_cigar(A.data(), m, B.data(), n, &sc, full.al.start_i, full.al.start_j,
                                full.cigar.c_str(), &rs, &ri, &rj) == 0);
        CHECK(rs == s1 && ri == ei && rj == ej);
        // the CLI renderer and gap lists on the same alignment
        std::vector<GapRun> ga, gb;
        cigar_gap_lists(full.cigar.c_str(), full.al.start_i, full.al.start_j, &ga, &gb);
        CHECK(write_gap_list((std::string(tmp) + "/g.first.gaps").c_str(), ga));
        CHECK(write_gap_list((std::string(tmp) + "/g.second.gaps").c_str(), gb));
        FILE* f = fopen("/dev/null", "w");
        print_alignment(f, A.data(), B.data(), full.al.start_i, full.al.start_j, full.cigar.c_str());
        fclose(f);
    }
}

// The multithreaded oracle paths against the single-thread ones.
static void oracle_mt(std::mt19937_64& r) {
    const orc_scoring sc = {1, -3, -5, -2};
    for (int it = 0; it < 3; ++it) {
        const std::vector<uint8_t> A = rand_seq(r, 1500 + (int64_t)(r() % 1200), 4);
        const std::vector<uint8_t> B = related(r, A, 1200 + (int64_t)(r() % 900));
        const std::vector<uint8_t>& R = A.size() >= B.size() ? A : B;
        const std::vector<uint8_t>& C = A.size() >= B.size() ? B : A;
        const int64_t m = (int64_t)R.size(), n = (int64_t)C.size();
        int32_t s1 = 0, s2 = 0;
        int64_t i1 = 0, j1 = 0, i2 = 0, j2 = 0;
        std::vector<int32_t> rm1((size_t)m), rm2((size_t)m);
        CHECK(orc_forward(R.data(), m, C.data(), n, &sc, &s1, &i1, &j1, rm1.data()) == 0);
        CHECK(orc_forward_mt(R.data(), m, C.data(), n, &sc, 4, &s2, &i2, &j2, rm2.data()) == 0);
        CHECK(s1 == s2 && i1 == i2 && j1 == j2 && rm1 == rm2);
        same(align_linear(R, C, sc, 256, 256, 4), align_linear(R, C, sc, 256, 256, 1));
        same(align_linear(R, C, sc, 512, 1024, 3), align_full(R, C, sc));
    }
}

static void fasta_dir(const char* dir) {
    DIR* d = opendir(dir);
    CHECK(d != nullptr);
    if (!d) return;
    int files = 0;
    while (dirent* e = readdir(d)) {
        if (e->d_name[0] == '.') continue;
        const std::string p = std::string(dir) + "/" + e->d_name;
        if (p.size() > 3 && p.compare(p.size() - 3, 3, ".gz") == 0) continue;
        FastaRecord rec;
        char* name = nullptr;
        uint8_t* seq = nullptr;
        int64_t len = 0;
        const bool ok = fasta_load(p.c_str(), &rec);
        const int rc = orc_fasta_load(p.c_str(), &name, &seq, &len);
        CHECK(ok == (rc == 0));
        if (ok && rc == 0) {
            CHECK(rec.name == name);
            CHECK((int64_t)rec.seq.size() == len && (len == 0 || !memcmp(rec.seq.data(), seq, (size_t)len)));
        }
        orc_free(name);
        orc_free(seq);
        ++files;
    }
    closedir(d);
    CHECK(files > 0);
}

static void synth(std::mt19937_64& r) {
    for (int it = 0; it < 6; ++it) {
        const int64_t l1 = (int64_t)(r() % 3000), l2 = (int64_t)(r() % 3000), h = (int64_t)(r() % 1500);
        uint8_t *a = nullptr, *b = nullptr;
        CHECK(swh_synth_pair(r(), l1, l2, h, 0.08, 0.015, it % 2 ? 0.4 : -1.0, it % 3 ? 0.01 : 0.0, &a, &b) == 0);
        uint8_t* q = nullptr;
        int64_t ql = -1;
        CHECK(swh_synth_mutate(r(), a, l1, 0.1, 0.03, 0.2, &q, &ql) == 0);
        CHECK(ql >= 0);
        free(a);
        free(b);
        free(q);
    }
    uint8_t *a = nullptr, *b = nullptr;
    CHECK(swh_synth_pair(1, -1, 5, 0, 0, 0, -1, 0, &a, &b) == SWH_ERR_INVALID_ARG);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <fasta-dir> <tmp-dir> [mt-only]\n", argv[0]);
        return 2;
    }
    std::mt19937_64 r(2012);
    if (argc > 3 && !strcmp(argv[3], "mt-only")) {
        oracle_mt(r);
    } else {
        const orc_scoring scs[] = {{1, -3, -5, -2}, {2, -3, -5, -1}, {1, -1, -1, -1}, {5, -4, -10, -1}};
        int pairs = 0;
        for (const orc_scoring& sc : scs) {
            for (int it = 0; it < 12; ++it) {
                const int64_t la = it < 4 ? (int64_t)(r() % 6) : (int64_t)(r() % 400), alpha = 2 + it % 4;
                std::vector<uint8_t> A = rand_seq(r, la, (int)alpha);
                std::vector<uint8_t> B = it % 2 ? related(r, A, la / 2 + 3) : rand_seq(r, (int64_t)(r() % 300), 4);
                if (A.size() < B.size()) std::swap(A, B);
                if (B.empty()) continue;
                oracle_pair(A, B, sc, argv[2]);
                ++pairs;
            }
        }
        oracle_mt(r);
        fasta_dir(argv[1]);
        synth(r);
        printf("pairs %d\n", pairs);
    }
    printf("%s\n", g_fail ? "FAIL" : "OK");
    return g_fail ? 1 : 0;
}
