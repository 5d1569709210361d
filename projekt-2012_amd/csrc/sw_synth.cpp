// Synthetic DNA workloads for the benchmarks (host code, part of libswalign_hip.so).
// Generator: std::mt19937_64; every draw below is written out explicitly (no standard
// distributions) so that a seed gives the same bytes with any C++ library.
//   * background: i.i.d. bases, uniform ACGT or with a GC fraction (A=T=(1-gc)/2,
//     C=G=gc/2);
//   * tandem repeats: a random unit of period U{2..50} copied U{5..40} times at a random
//     position, until the requested fraction of the sequence is covered;
//   * homology: a window of seq1 is mutated (substitution to one of the other three
//     bases with p_sub per base; an indel event with p_indel per base, insertion or
//     deletion with equal odds, length ~ Geometric(mean 3, support >= 1), inserted bases
//     uniform) and written over a random window of seq2.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/swalign_hip.h"

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    uint64_t u64() { return g(); }
    double unif() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
    int64_t below(int64_t k) { return k <= 1 ? 0 : (int64_t)(g() % (uint64_t)k); }
    int64_t range(int64_t lo, int64_t hi) { return lo + below(hi - lo + 1); }  // inclusive
    int64_t geometric_mean3() {  // P(L=k) = (2/3)^(k-1) * (1/3), k >= 1
        int64_t L = 1;
        while (unif() >= 1.0 / 3.0) ++L;
        return L;
    }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

char draw_base(Rng& r, double gc) {
    if (gc < 0) return kBases[r.below(4)];
    const double u = r.unif();
    const double at = (1.0 - gc) / 2.0;
    if (u < at) return 'A';
    if (u < at + gc / 2.0) return 'C';
    if (u < at + gc) return 'G';
    return 'T';
}

void fill_random(Rng& r, std::vector<char>& s, double gc) {
    for (auto& c : s) c = draw_base(r, gc);
}

void plant_repeats(Rng& r, std::vector<char>& s, double frac, double gc) {
    const int64_t L = (int64_t)s.size();
    if (frac <= 0 || L < 100) return;
    int64_t covered = 0;
    const int64_t target = (int64_t)(frac * (double)L);
    std::vector<char> unit;
    while (covered < target) {
        const int64_t period = r.range(2, 50);
        const int64_t copies = r.range(5, 40);
        unit.resize((size_t)period);
        for (auto& c : unit) c = draw_base(r, gc);
        const int64_t len = std::min<int64_t>(period * copies, L);
        const int64_t pos = r.below(L - len + 1);
        for (int64_t k = 0; k < len; ++k) s[(size_t)(pos + k)] = unit[(size_t)(k % period)];
        covered += len;
    }
}

std::vector<char> mutate(Rng& r, const char* src, int64_t len, double p_sub, double p_indel) {
    std::vector<char> out;
    out.reserve((size_t)(len + len / 8 + 16));
    for (int64_t i = 0; i < len;) {
        if (r.unif() < p_indel) {
            const int64_t L = r.geometric_mean3();
            if (r.unif() < 0.5) {
                for (int64_t k = 0; k < L; ++k) out.push_back(kBases[r.below(4)]);
                // the current base is still copied below
            } else {
                i += L;
                continue;
            }
        }
        char c = src[i];
        if (r.unif() < p_sub) {
            int idx = 0;
            for (int k = 0; k < 4; ++k)
                if (kBases[k] == c) idx = k;
            c = kBases[(idx + 1 + r.below(3)) % 4];
        }
        out.push_back(c);
        ++i;
    }
    return out;
}

uint8_t* to_heap(const std::vector<char>& v) {
    uint8_t* p = (uint8_t*)malloc(v.size() + 1);
    if (!p) return nullptr;
    if (!v.empty()) memcpy(p, v.data(), v.size());
    p[v.size()] = 0;
    return p;
}

}  // namespace

extern "C" {

int swh_synth_pair(uint64_t seed, int64_t len1, int64_t len2, int64_t hom_len, double p_sub, double p_indel,
                   double gc, double repeat_frac, uint8_t** seq1, uint8_t** seq2) {
    if (len1 < 0 || len2 < 0 || hom_len < 0 || !seq1 || !seq2) return SWH_ERR_INVALID_ARG;
    Rng r(seed);
    std::vector<char> a((size_t)len1), b((size_t)len2);
    fill_random(r, a, gc);
    fill_random(r, b, gc);
    plant_repeats(r, a, repeat_frac, gc);
    plant_repeats(r, b, repeat_frac, gc);
    hom_len = std::min<int64_t>(hom_len, std::min(len1, len2));
    if (hom_len > 0) {
        const int64_t o1 = r.below(len1 - hom_len + 1);
        std::vector<char> m = mutate(r, a.data() + o1, hom_len, p_sub, p_indel);
        const int64_t w = std::min<int64_t>((int64_t)m.size(), len2);
        const int64_t o2 = r.below(len2 - w + 1);
        memcpy(b.data() + o2, m.data(), (size_t)w);
    }
    *seq1 = to_heap(a);
    *seq2 = to_heap(b);
    return (*seq1 && *seq2) ? SWH_OK : SWH_ERR_OUT_OF_MEMORY;
}

int swh_synth_mutate(uint64_t seed, const uint8_t* src, int64_t len, double p_sub, double p_indel,
                     double flank_frac, uint8_t** out, int64_t* out_len) {
    if (!src || len < 0 || !out || !out_len) return SWH_ERR_INVALID_ARG;
    Rng r(seed);
    std::vector<char> m = mutate(r, (const char*)src, len, p_sub, p_indel);
    // each flank: trimmed or extended by U(0, flank_frac) * len bases
    for (int side = 0; side < 2; ++side) {
        const int64_t k = (int64_t)(r.unif() * flank_frac * (double)len);
        if (r.unif() < 0.5) {
            const int64_t cut = std::min<int64_t>(k, (int64_t)m.size() / 2);
            if (side == 0) m.erase(m.begin(), m.begin() + cut);
            else m.resize(m.size() - (size_t)cut);
        } else {
            std::vector<char> ext((size_t)k);
            for (auto& c : ext) c = kBases[r.below(4)];
            if (side == 0) m.insert(m.begin(), ext.begin(), ext.end());
            else m.insert(m.end(), ext.begin(), ext.end());
        }
    }
    *out = to_heap(m);
    *out_len = (int64_t)m.size();
    return *out ? SWH_OK : SWH_ERR_OUT_OF_MEMORY;
}

}  // extern "C"
