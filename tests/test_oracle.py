"""CPU checks that pin the oracle before it is trusted as the parity reference.

Each restated function is checked against something independent of its own formula:
known-answer values from the reference's test scripts, an exhaustive enumeration of
alignments, closed forms, the textbook Gotoh recurrence, the Myers-Miller crosspoint
algebra (a reverse DP) and invariances of the optimal score.
"""
import itertools
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, fasta_bytes, orient


def rand_seq(rng, n, alpha="ACGT"):
    return "".join(rng.choice(alpha) for _ in range(n))


def rand_scoring(rng, orc):
    return orc.scoring(rng.randint(1, 4), -rng.randint(1, 4), -rng.randint(1, 7), -rng.randint(1, 3))


def test_bruteforce_exhaustive_small(orc):
    """Every pair of {A,C} strings up to length 4 x 4: DP score == enumeration of all
    local alignments (pins the recurrence, the zero clamp and the gap cost model)."""
    for sc in [orc.scoring(1, -3, -5, -2), orc.scoring(2, -1, -2, -1), orc.scoring(3, -2, -1, -3)]:
        for m in range(1, 5):
            for n in range(1, 5):
                for a in itertools.product("AC", repeat=m):
                    for b in itertools.product("AC", repeat=n):
                        a_, b_ = "".join(a), "".join(b)
                        assert orc.forward(a_, b_, sc)[0] == orc.bruteforce(a_, b_, sc), (a_, b_)


def test_bruteforce_random(orc):
    rng = random.Random(11)
    for _ in range(300):
        a, b = rand_seq(rng, rng.randint(1, 7), "ACG"), rand_seq(rng, rng.randint(1, 7), "ACG")
        sc = rand_scoring(rng, orc)
        assert orc.forward(a, b, sc)[0] == orc.bruteforce(a, b, sc)


def test_zero_boundary_equals_textbook(orc):
    """The reference's zero E/F boundaries (memset 0) give the same H as -inf boundaries."""
    rng = random.Random(3)
    for _ in range(100):
        a, b = rand_seq(rng, rng.randint(1, 60)), rand_seq(rng, rng.randint(1, 60))
        sc = rand_scoring(rng, orc)
        H, E, F = orc.full_matrices(a, b, sc)
        assert np.array_equal(H, orc.textbook_H(a, b, sc))


def test_closed_forms(orc):
    sc = orc.scoring(1, -3, -5, -2)
    # identical sequences: L*match at [L, L]
    s = "ACGTTGCAAGT" * 9
    assert orc.forward(s, s, sc) == (len(s), len(s) - 1, len(s) - 1)
    # reference fixtures seq3/seq4: 100 x 'A' each
    a, b = fasta_bytes("seq3.fna"), fasta_bytes("seq4.fna")
    assert len(a) == len(b) == 100
    assert orc.forward(a, b, orc.scoring(5, -5, -5, -5)) == (500, 99, 99)
    # disjoint alphabets: no positive cell
    assert orc.forward("ACACAC", "GTGTGT", sc)[0] == 0
    # one base inserted into B: L*M + gap_open when that beats the longer ungapped half
    L = 40
    base = "".join("ACGT"[(i * 7 + i // 3) % 4] for i in range(L))
    ins = base[:20] + "T" + base[20:]
    s2 = orc.scoring(2, -3, -5, -1)
    assert orc.forward(ins, base, s2)[0] == L * 2 - 5
    al = orc.align_full(ins, base, s2)
    assert al["cigar"] in ("20=1D20=", "19=1D21=", "21=1D19=") or al["cigar"].count("D") == 1


def test_score_symmetries(orc):
    """Swapping the inputs (transposing the matrix) and reversing both sequences leave
    the optimal local score unchanged."""
    rng = random.Random(5)
    for _ in range(200):
        a, b = rand_seq(rng, rng.randint(1, 80)), rand_seq(rng, rng.randint(1, 80))
        sc = rand_scoring(rng, orc)
        s = orc.forward(a, b, sc)[0]
        assert orc.forward(b, a, sc)[0] == s
        assert orc.forward(a[::-1], b[::-1], sc)[0] == s


def test_end_cell_rule(orc):
    """End cell = first maximal cell in row-major order (checked on the full matrix)."""
    rng = random.Random(9)
    for _ in range(100):
        a, b = rand_seq(rng, rng.randint(1, 50), "AC"), rand_seq(rng, rng.randint(1, 50), "AC")
        sc = rand_scoring(rng, orc)
        s, i, j = orc.forward(a, b, sc)
        H, _, _ = orc.full_matrices(a, b, sc)
        if s > 0:
            ii, jj = np.argwhere(H == H.max())[0]
            assert (i, j) == (ii, jj) and s == H.max()


def test_linear_method_equals_full_matrix(orc):
    """Checkpoint rows/columns + block recomputation give exactly the full-matrix
    canonical alignment for any spacing, and every crosspoint satisfies the reference's
    match algebra (forward + reverse DP, SWutils.cpp:175-231)."""
    rng = random.Random(13)
    for t in range(150):
        m, n = rng.randint(1, 90), rng.randint(1, 90)
        a, b = rand_seq(rng, m), rand_seq(rng, n)
        if m > 20 and n > 20:
            b = b[:5] + a[3:m - 3] + b[5:]
        sc = rand_scoring(rng, orc)
        A, B = orient(a, b)
        full = orc.align_full(A, B, sc)
        for K, Kc in [(1, 1), (2, 3), (5, 7), (16, 16), (64, 32), (1000, 1000)]:
            lin = orc.align_linear(A, B, sc, K, Kc)
            for k in ("score", "end", "start", "cigar"):
                assert lin[k] == full[k], (k, K, Kc)
            for (r, c, v, layer) in lin["xpoints"][:-1]:
                best, at = orc.crosspoint_algebra(A, B, sc, r, full["end"], c, 1 if layer == 2 else 0)
                assert best == full["score"] and at == full["score"]
        if full["score"] > 0:
            assert orc.rescore(A, B, sc, full["start"], full["cigar"]) == (full["score"], full["end"])


def test_checkpoint_rows_match_full_matrices(orc):
    rng = random.Random(17)
    a, b = rand_seq(rng, 70), rand_seq(rng, 50)
    sc = orc.scoring(1, -3, -5, -2)
    H, E, F = orc.full_matrices(a, b, sc)
    srow, scol = orc.checkpoints(a, b, sc, 8, 16)
    for k in range(len(srow)):
        r = (k + 1) * 8 - 1
        assert np.array_equal(srow[k][:, 0], H[r]) and np.array_equal(srow[k][:, 1], F[r])
    for k in range(len(scol)):
        c = (k + 1) * 16 - 1
        assert np.array_equal(scol[k][:, 0], H[:, c]) and np.array_equal(scol[k][:, 1], E[:, c])


def test_multithreaded_forward_equals_single(orc):
    rng = random.Random(21)
    a, b = rand_seq(rng, 5000), rand_seq(rng, 4000)
    b = b[:1000] + a[2000:3500] + b[1000:]
    sc = orc.scoring(1, -3, -5, -2)
    (s1, rm1) = orc.forward(a, b, sc, rowmax=True)
    (s8, rm8) = orc.forward(a, b, sc, rowmax=True, threads=8)
    assert s1 == s8 and np.array_equal(rm1, rm8)
    # without the end-column search (the CPU baseline's mode), small and large inputs
    for x, y in [(a[:300], b[:200]), (a, b)]:
        full = orc.forward(x, y, sc, threads=8)
        part = orc.forward(x, y, sc, threads=8, end_col=False)
        assert part[:2] == full[:2] and part[2] == -1


def test_multithreaded_linear_equals_single(orc):
    """The linear-space method with its stage 1 on several host threads (the golden
    generator's mode for the 1 Mbp pair and the served jobs) gives the single-thread
    result, crosspoints included, at spacings whose checkpoints fall on tile edges, at
    spacings that fall back to one thread, and on pairs below the threading cut-off; the
    full-matrix alignment agrees where it fits."""
    rng = random.Random(23)
    a, b = rand_seq(rng, 3300), rand_seq(rng, 2900)
    b = b[:400] + a[700:2600] + b[400:]
    sc = orc.scoring(1, -3, -5, -2)
    for x, y, K, Kc in [(a, b, 512, 128), (a, b, 256, 4096), (a, b, 768, 2048), (a, b, 300, 100),
                        (a[:900], b[:800], 256, 64)]:
        A, B = orient(x, y)
        one = orc.align_linear(A, B, sc, K, Kc)
        mt = orc.align_linear(A, B, sc, K, Kc, threads=6)
        for k in ("score", "end", "start", "cigar", "xpoints"):
            assert mt[k] == one[k], (k, K, Kc)
    A, B = orient(a[:1500], b[:1400])
    full = orc.align_full(A, B, sc)
    mt = orc.align_linear(A, B, sc, 256, 128, threads=4)
    for k in ("score", "end", "start", "cigar"):
        assert mt[k] == full[k], k


@pytest.mark.parametrize("name", ["test1", "test2"])
def test_known_answers(orc, name):
    k = [x for x in json.load(open(os.path.join(GOLDEN, "kats.json"))) if x["name"] == name][0]
    A, B = orient(fasta_bytes(k["a"]), fasta_bytes(k["b"]))
    s, i, j = orc.forward(A, B, orc.scoring(*k["scoring"]))
    assert s == k["score"] and [i + 1, j + 1] == k["end"]


def test_known_answer_3(orc):
    """162 kbp x 172 kbp: 18 at [44353, 41058] (Tests/test3.sh:3); one row holds the
    maximum, so the row-major tie rule is not exercised beyond the row."""
    k = [x for x in json.load(open(os.path.join(GOLDEN, "kats.json"))) if x["name"] == "test3"][0]
    A, B = orient(fasta_bytes(k["a"]), fasta_bytes(k["b"]))
    (s, i, j), rm = orc.forward(A, B, orc.scoring(*k["scoring"]), rowmax=True, threads=8)
    assert s == k["score"] and [i + 1, j + 1] == k["end"]


def test_golden_fixtures_reproduce(orc):
    g = json.load(open(os.path.join(GOLDEN, "config0.json")))
    pairs = json.load(open(os.path.join(GOLDEN, "small_pairs.json")))
    for p in pairs[:100]:
        A, B = orient(p["a"], p["b"])
        r = orc.align_full(A, B, orc.scoring(*p["scoring"]))
        assert (r["score"], list(r["end"]), list(r["start"]), r["cigar"]) == (
            p["score"], list(p["end"]), list(p["start"]), p["cigar"])
    assert g["score"] > 1000 and g["cigar"]


def test_fasta_reference_rules(orc, tmp_path):
    """FASTA.cpp:39-86: the name is the first fgets(200) minus '>' and its last byte; the
    sequence keeps every byte except CR/LF (case and IUPAC codes verbatim)."""
    p = tmp_path / "x.fna"
    p.write_bytes(b">name one\r\nACGTn\r\nRYK M\n")
    name, seq = orc.fasta_load(str(p))
    assert name == "name one\r"   # the reference drops only the last byte of the line
    assert seq == b"ACGTnRYK M"
    p.write_bytes(b"no header\nACGT\n")
    with pytest.raises(IOError):
        orc.fasta_load(str(p))
    long_header = b">" + b"h" * 250 + b"\nACGT\n"
    p.write_bytes(long_header)
    name, seq = orc.fasta_load(str(p))
    assert len(name) == 197 and seq == b"ACGT"   # fgets(200) keeps 199 bytes
