// Alignment rendering for the SWalign CLI.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

struct GapRun {
    int32_t begin;   // offset in the sequence before which the gap run sits
    int32_t length;  // gap characters
};

// Gap lists of both sequences from a CIGAR ('D' = gap in B, 'I' = gap in A).
void cigar_gap_lists(const char* cigar, int64_t start_a, int64_t start_b, std::vector<GapRun>* gaps_a,
                     std::vector<GapRun>* gaps_b);

// Writes a gap list as consecutive little-endian {int32 begin, int32 length} records.
bool write_gap_list(const char* path, const std::vector<GapRun>& gaps);

// Prints the alignment in 50-column blocks of three lines (A, match bar, B), each line
// prefixed by two tabs; '|' marks identical bytes and '-' gap characters.
void print_alignment(FILE* out, const uint8_t* A, const uint8_t* B, int64_t start_a, int64_t start_b,
                     const char* cigar);
