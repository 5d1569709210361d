// Reference-compatible FASTA loading (Sterbic/Projekt-2012 src/FASTA.cpp:39-86):
//  * the header is read with fgets(buf, 200): at most 199 bytes up to and including '\n';
//    it must start with '>', and the name is that buffer minus '>' and minus its last
//    byte (normally the newline);
//  * the sequence starts after a second fgets(buf, 300) of the first line and is every
//    remaining byte except '\n' and '\r', kept verbatim (case, IUPAC codes, spaces).
#include "sw_fasta.h"

#include <cstdio>

namespace {
size_t fgets_len(const std::vector<uint8_t>& all, size_t cap) {
    size_t k = 0;
    while (k < all.size() && k < cap - 1) {
        if (all[k++] == '\n') break;
    }
    return k;
}
}  // namespace

bool fasta_load(const char* path, FastaRecord* out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    std::vector<uint8_t> all;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) all.insert(all.end(), buf, buf + k);
    fclose(f);
    if (all.empty() || all[0] != '>') return false;
    const size_t l200 = fgets_len(all, 200);
    out->name.assign(all.begin() + 1, all.begin() + (l200 >= 2 ? l200 - 1 : 1));
    const size_t start = fgets_len(all, 300);
    out->seq.clear();
    out->seq.reserve(all.size() - start);
    for (size_t p = start; p < all.size(); ++p) {
        const uint8_t c = all[p];
        if (c == '\n' || c == '\r') continue;
        out->seq.push_back(c);
    }
    return true;
}
