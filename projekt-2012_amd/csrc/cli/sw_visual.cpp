// Alignment rendering (the reference's unused src/Visualisation.cpp:11-129 renders the
// same 3-line, 50-column blocks from two binary gap lists; here the CIGAR is the source
// and the gap lists are produced from it).
#include "sw_visual.h"

#include <algorithm>
#include <cstring>

namespace {
constexpr int kRowLength = 50;

template <class F>
void for_each_op(const char* cigar, F f) {
    for (const char* p = cigar; p && *p;) {
        int64_t cnt = 0;
        while (*p >= '0' && *p <= '9') cnt = cnt * 10 + (*p++ - '0');
        const char op = *p ? *p++ : 0;
        if (!op) break;
        f(op, cnt);
    }
}
}  // namespace

void cigar_gap_lists(const char* cigar, int64_t start_a, int64_t start_b, std::vector<GapRun>* gaps_a,
                     std::vector<GapRun>* gaps_b) {
    int64_t ia = start_a, ib = start_b;
    for_each_op(cigar, [&](char op, int64_t cnt) {
        if (op == '=' || op == 'X' || op == 'M') {
            ia += cnt;
            ib += cnt;
        } else if (op == 'D') {  // A bases against gaps in B
            gaps_b->push_back(GapRun{(int32_t)ib, (int32_t)cnt});
            ia += cnt;
        } else if (op == 'I') {  // B bases against gaps in A
            gaps_a->push_back(GapRun{(int32_t)ia, (int32_t)cnt});
            ib += cnt;
        }
    });
}

bool write_gap_list(const char* path, const std::vector<GapRun>& gaps) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    for (const GapRun& g : gaps) {
        fwrite(&g.begin, sizeof(int32_t), 1, f);
        fwrite(&g.length, sizeof(int32_t), 1, f);
    }
    fclose(f);
    return true;
}

void print_alignment(FILE* out, const uint8_t* A, const uint8_t* B, int64_t start_a, int64_t start_b,
                     const char* cigar) {
    std::string la, lm, lb;
    int64_t ia = start_a, ib = start_b;
    for_each_op(cigar, [&](char op, int64_t cnt) {
        for (int64_t k = 0; k < cnt; ++k) {
            if (op == '=' || op == 'X' || op == 'M') {
                la.push_back((char)A[ia]);
                lb.push_back((char)B[ib]);
                lm.push_back(A[ia] == B[ib] ? '|' : ' ');
                ++ia;
                ++ib;
            } else if (op == 'D') {
                la.push_back((char)A[ia++]);
                lb.push_back('-');
                lm.push_back(' ');
            } else if (op == 'I') {
                la.push_back('-');
                lb.push_back((char)B[ib++]);
                lm.push_back(' ');
            }
        }
    });
    fprintf(out, "\n\tVisual representation of the alignment:\n\n");
    for (size_t p = 0; p < la.size(); p += kRowLength) {
        const size_t w = std::min<size_t>(kRowLength, la.size() - p);
        fprintf(out, "\t\t%.*s\n\t\t%.*s\n\t\t%.*s\n\n", (int)w, la.data() + p, (int)w, lm.data() + p, (int)w,
                lb.data() + p);
    }
}
