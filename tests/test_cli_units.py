"""The SWalign CLI's own FASTA loader (csrc/cli/sw_fasta.cpp) and alignment renderer
(csrc/cli/sw_visual.cpp), built from their sources with g++ into a small driver and run on
the CPU:

* the loader follows the reference's rules (src/FASTA.cpp:39-86): the name is the first
  fgets(buf, 200) minus '>' and its last byte, the sequence starts after a second
  fgets(buf, 300) of the first line and keeps every byte but CR and LF (case, IUPAC codes
  and spaces verbatim); it agrees with the oracle's loader on every case;
* the gap lists ({int32 begin, int32 length} records, src/Visualisation.cpp:27-43) and the
  50-column blocks (src/Visualisation.cpp:62-127) built from a CIGAR equal a restatement of
  the reference's visualOutput fed with gap lists derived here from the oracle's CIGAR.
"""
import os
import shutil
import struct
import subprocess

import pytest

from conftest import ROOT, orient

CLI_SRC = os.path.join(ROOT, "projekt-2012_amd", "csrc", "cli")

DRIVER = r'''
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "sw_fasta.h"
#include "sw_visual.h"
static std::string slurp(const char* p) {
    FILE* f = fopen(p, "rb"); std::string s; char b[4096]; size_t k;
    while ((k = fread(b, 1, sizeof b, f)) > 0) s.append(b, k);
    fclose(f); return s;
}
int main(int argc, char** argv) {
    if (!strcmp(argv[1], "fasta")) {
        FastaRecord r;
        if (!fasta_load(argv[2], &r)) { printf("FAIL\n"); return 0; }
        printf("%zu\n", r.name.size());
        for (unsigned char c : r.name) printf("%02x", c);
        printf("\n%zu\n", r.seq.size());
        for (unsigned char c : r.seq) printf("%02x", c);
        printf("\n");
        return 0;
    }
    // visual A B start_a start_b cigar prefix
    std::string A = slurp(argv[2]), B = slurp(argv[3]), cig = slurp(argv[6]);
    const long long sa = atoll(argv[4]), sb = atoll(argv[5]);
    std::vector<GapRun> ga, gb;
    cigar_gap_lists(cig.c_str(), sa, sb, &ga, &gb);
    write_gap_list((std::string(argv[7]) + ".first.gaps").c_str(), ga);
    write_gap_list((std::string(argv[7]) + ".second.gaps").c_str(), gb);
    print_alignment(stdout, (const uint8_t*)A.data(), (const uint8_t*)B.data(), sa, sb, cig.c_str());
    return 0;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("cli_units")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CLI_SRC, str(src), os.path.join(CLI_SRC, "sw_fasta.cpp"),
                    os.path.join(CLI_SRC, "sw_visual.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def _load(driver, path):
    out = subprocess.run([driver, "fasta", path], capture_output=True, text=True, check=True).stdout.split("\n")
    if out[0] == "FAIL":
        return None
    return bytes.fromhex(out[1]).decode("latin-1"), bytes.fromhex(out[3])


CASES = {
    "crlf_iupac_lowercase_spaces": b">name one\r\nACGTn\r\nRYK M\nacgt\n",
    "no_trailing_newline": b">x\nACGT",
    "header_only": b">only a header\n",
    "header_without_newline": b">abc",
    "empty_name": b">\nAC\nGT\n",
    "header_250": b">" + b"h" * 250 + b"\nACGT\n",
    "header_198": b">" + b"k" * 198 + b"\nAC\n",
    "header_400": b">" + b"q" * 400 + b"\nTTGA\n",
    "bytes_kept": b">bin\n\x01\x7f\xff ?$#\n",
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cli_fasta_loader_equals_reference_rules(driver, orc, tmp_path, name):
    p = tmp_path / "x.fna"
    p.write_bytes(CASES[name])
    got = _load(driver, str(p))
    assert got == orc.fasta_load(str(p))
    nm, seq = got
    if name == "crlf_iupac_lowercase_spaces":
        assert nm == "name one\r" and seq == b"ACGTnRYK Macgt"
    elif name == "header_250":
        assert len(nm) == 197 and seq == b"ACGT"
    elif name == "header_198":
        # a 200-byte line: fgets(200) stops one byte short of the newline and the last byte
        # it kept is dropped
        assert nm == "k" * 197 and seq == b"AC"
    elif name == "header_400":
        # fgets(300) stops inside the header: the header's tail becomes sequence (the
        # reference's behaviour, kept)
        assert len(nm) == 197 and seq == b"q" * (401 - 299) + b"TTGA"
    elif name == "bytes_kept":
        assert seq == b"\x01\x7f\xff ?$#"
    elif name == "header_only":
        assert seq == b""


def test_cli_fasta_loader_rejects_missing_marker(driver, tmp_path):
    p = tmp_path / "x.fna"
    p.write_bytes(b"ACGT\n>late\n")
    assert _load(driver, str(p)) is None
    p.write_bytes(b"")
    assert _load(driver, str(p)) is None


def gap_lists(cigar, start):
    """Gap lists from a CIGAR: a run of I (B against gaps) opens a gap in the first line
    before A[ia]; a run of D opens one in the second line before B[ib]."""
    import re
    ia, ib = start
    ga, gb = [], []
    for cnt, op in re.findall(r"(\d+)([=XIDM])", cigar):
        cnt = int(cnt)
        if op in "=XM":
            ia += cnt
            ib += cnt
        elif op == "I":
            ga.append((ia, cnt))
            ib += cnt
        else:
            gb.append((ib, cnt))
            ia += cnt
    return ga, gb


def reference_visual(first, second, g1, g2, start, end, row=50):
    """src/Visualisation.cpp:62-127 restated: fill 50 columns of the first line from the
    first sequence, opening the gaps of g1 where the offset reaches a gap's begin, stop past
    the end cell; the same for the second line; the middle line marks equal characters."""
    out = ["\n\tVisual representation of the alignment:\n\n"]
    fo, so = start
    fgap = sgap = 0
    i1 = i2 = 0
    while True:
        done = 0
        lines = []
        for seq, gaps, which in ((first, g1, 0), (second, g2, 1)):
            off = fo if which == 0 else so
            ingap = fgap if which == 0 else sgap
            it = i1 if which == 0 else i2
            v = []
            i = 0
            while i < row:
                if ingap:
                    v.append("-")
                    ingap -= 1
                elif it < len(gaps) and gaps[it][0] == off:
                    ingap = gaps[it][1]
                    it += 1
                    continue
                else:
                    v.append(chr(seq[off]))
                    off += 1
                if off > end[which]:
                    done += 1
                    break
                i += 1
            if which == 0:
                fo, fgap, i1 = off, ingap, it
            else:
                so, sgap, i2 = off, ingap, it
            lines.append(v)
        mid = ["|" if x == y else " " for x, y in zip(lines[0], lines[1] + [""] * len(lines[0]))]
        out.append("\t\t%s\n\t\t%s\n\t\t%s\n\n" % ("".join(lines[0]), "".join(mid), "".join(lines[1])))
        if done:
            break
    return "".join(out)


def test_gap_lists_and_blocks_equal_reference_renderer(driver, orc, tmp_path):
    from conftest import fasta_bytes
    A, B = orient(fasta_bytes("NC_005046.fna"), fasta_bytes("NC_012557.fna"))
    ref = orc.align_full(A, B, orc.scoring(5, -5, -5, -5))
    assert ref["score"] == 1580 and "I" in ref["cigar"] and "D" in ref["cigar"]
    (tmp_path / "a").write_bytes(A)
    (tmp_path / "b").write_bytes(B)
    (tmp_path / "c").write_text(ref["cigar"])
    pre = str(tmp_path / "g")
    shown = subprocess.run([driver, "visual", str(tmp_path / "a"), str(tmp_path / "b"), str(ref["start"][0]),
                            str(ref["start"][1]), str(tmp_path / "c"), pre], capture_output=True, text=True,
                           check=True).stdout
    ga, gb = gap_lists(ref["cigar"], ref["start"])
    for path, want in ((pre + ".first.gaps", ga), (pre + ".second.gaps", gb)):
        data = open(path, "rb").read()
        assert list(struct.iter_unpack("<ii", data)) == want
    assert shown == reference_visual(A, B, ga, gb, ref["start"], ref["end"])
