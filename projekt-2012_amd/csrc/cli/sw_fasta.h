// Reference-compatible single-record FASTA reader for the SWalign CLI.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct FastaRecord {
    std::string name;           // header line without '>' and without its last byte
    std::vector<uint8_t> seq;   // every byte after the header except '\n' and '\r'
};

// Returns false when the file cannot be opened or does not start with '>'
// (Sterbic/Projekt-2012 src/FASTA.cpp:39-86).
bool fasta_load(const char* path, FastaRecord* out);
