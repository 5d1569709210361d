"""Batched alignment (swh_batch_align / swh_batch_align_device) on an MI355X: every pair of
a mixed-length batch gets exactly the result it gets alone and the oracle's canonical
alignment (integer arithmetic: exact comparisons)."""
import random

import pytest

from conftest import orient

pytestmark = pytest.mark.gpu

SC = (1, -3, -5, -2)


def _mixed_pairs(sw, seed, n, lens):
    rng = random.Random(seed)
    pairs = []
    for k in range(n):
        L = rng.choice(lens)
        L2 = max(0, L + rng.randint(-L // 3 - 1, L // 3 + 1))
        a, b = sw.synth_pair(seed * 1000 + k, L, L2, max(1, min(L, L2) // 2), 0.1, 0.02)
        if rng.random() < 0.3:
            a, b = b, a
        pairs.append((a, b))
    return pairs


def _check(orc, pairs, res, st, K=256, xpoints=True, Kc=512):
    """Every result equals the oracle's alignment at checkpoint blocks of K x Kc (the batch
    default: 256 x 512); crosspoints sit on checkpoint rows, so they depend on K."""
    for (a, b), r, s in zip(pairs, res, st):
        assert s == 0
        A, B = orient(a, b)
        if not A or not B:
            assert r["score"] == 0
            continue
        ref = orc.align_linear(A, B, orc.scoring(*SC), K, Kc)
        if ref["score"] == 0:
            assert r["score"] == 0
            continue
        for k in ("score", "end", "start", "cigar"):
            assert r[k] == ref[k], (k, len(a), len(b))
        if xpoints:
            assert r["xpoints"] == [tuple(x) for x in ref["xpoints"]]


def test_batch_matches_oracle(sw, orc, ctx):
    pairs = _mixed_pairs(sw, 7, 150, [0, 1, 2, 17, 64, 255, 256, 257, 600, 1500, 3000])
    res, st = ctx.batch(pairs, sw.scoring(*SC), xpoints=True)
    _check(orc, pairs, res, st)


def test_batch_device_cigars_match_oracle(sw, orc, ctx):
    """A batch of 300 alignments (CIGAR text built on the device) with runs of every length
    class: single ops, runs of 9/10/99/100/1000+ (identical and near-identical pairs), long
    gaps, one-base pairs."""
    rng = random.Random(23)
    pairs = _mixed_pairs(sw, 19, 260, [1, 2, 17, 64, 300, 1500])
    for k in range(40):
        L = rng.choice([9, 10, 11, 99, 100, 101, 1000, 2500])
        a = bytes(rng.choice(b"ACGT") for _ in range(L))
        if k % 3 == 0:
            b = a  # one run of L
        elif k % 3 == 1:
            i = rng.randrange(L)
            b = a[:i] + bytes([a[i] ^ 6]) + a[i + 1:]  # a mismatch inside
        else:
            i = rng.randrange(L)
            b = a[:i] + b"T" * rng.choice([1, 9, 10, 120]) + a[i:]  # a gap inside
        pairs.append((a, b))
    res, st = ctx.batch(pairs, sw.scoring(*SC), xpoints=True)
    _check(orc, pairs, res, st)


@pytest.mark.parametrize("kw", [dict(tile_cols=100), dict(lookahead=-1), dict(lookahead=2, tile_cols=37),
                                dict(K=512, Kc=256), dict(waves_per_simd=1), dict(waves_per_simd=2),
                                dict(waves_per_simd=3, tile_cols=256)])
def test_batch_options_same_result(sw, orc, ctx, kw):
    pairs = _mixed_pairs(sw, 11, 40, [5, 300, 1100, 2500])
    res, st = ctx.batch(pairs, sw.scoring(*SC), **kw)
    _check(orc, pairs, res, st, K=1024, xpoints=False)


def test_batch_equals_single(sw, ctx):
    pairs = _mixed_pairs(sw, 13, 30, [50, 900, 4000])
    res, st = ctx.batch(pairs, sw.scoring(*SC))
    for (a, b), r, s in zip(pairs, res, st):
        one = ctx.align(a, b, sw.scoring(*SC))
        assert s == 0
        assert (r["score"], r["end"], r["start"], r["cigar"], r["swapped"]) == \
            (one["score"], one["end"], one["start"], one["cigar"], one["swapped"])


def test_batch_device_resident(sw, orc, ctx):
    pairs = _mixed_pairs(sw, 17, 60, [1, 120, 800, 2000])
    dev = sw.DeviceBuffer(b"".join(a + b for a, b in pairs))
    base, off, dp = dev.address, 0, []
    for a, b in pairs:
        dp.append((base + off, len(a), base + off + len(a), len(b)))
        off += len(a) + len(b)
    res, st = ctx.batch_device(dp, sw.scoring(*SC), xpoints=True)
    _check(orc, pairs, res, st)
    dev.free()


def test_batch_score_only(sw, orc, ctx):
    pairs = _mixed_pairs(sw, 19, 50, [3, 700, 2600])
    res, st = ctx.batch(pairs, sw.scoring(*SC), score_only=True)
    for (a, b), r, s in zip(pairs, res, st):
        assert s == 0
        A, B = orient(a, b)
        if not A or not B:
            assert r["score"] == 0
            continue
        sc, i, j = orc.forward(A, B, orc.scoring(*SC))
        assert r["score"] == sc
        if sc > 0:
            assert r["end"] == (i, j) and r["cigar"] is None


def test_batch_autotune(sw, orc, ctx):
    """The tuner reports a candidate from its grid, and a batch run with autotune gives the
    oracle's alignments like any other setting."""
    pairs = _mixed_pairs(sw, 23, 120, [40, 400, 1500, 3000])
    dev = sw.DeviceBuffer(b"".join(a + b for a, b in pairs))
    base, off, dp = dev.address, 0, []
    for a, b in pairs:
        dp.append((base + off, len(a), base + off + len(a), len(b)))
        off += len(a) + len(b)
    t = ctx.batch_tune(dp, sw.scoring(*SC), sample=64)
    assert t["tile_cols"] in (256, 384, 512, 768, 1024)
    assert t["waves_per_simd"] in (2, 3, 4)
    assert (t["checkpoint_rows"], t["checkpoint_cols"]) in ((256, 512), (512, 512), (1024, 1024), (2048, 2048))
    assert t["candidates"] == 60
    assert 0 < t["ms_best"] <= t["ms_default"] * 1.0001
    # crosspoints sit on checkpoint rows, so they depend on the spacing the tuner picks
    res, st = ctx.batch_device(dp, sw.scoring(*SC), autotune=True)
    _check(orc, pairs, res, st, xpoints=False)
    K, Kc = t["checkpoint_rows"], t["checkpoint_cols"]
    res2, st2 = ctx.batch_device(dp, sw.scoring(*SC), xpoints=True, tile_cols=t["tile_cols"],
                                 waves_per_simd=t["waves_per_simd"], K=K, Kc=Kc)
    _check(orc, pairs, res2, st2, K=K, Kc=Kc)
    dev.free()


@pytest.mark.parametrize("alphabet", [b"ACGT", b"ACGTN", b"AC", b"ACGTacgt"])
@pytest.mark.parametrize("score", [(1, -3, -5, -2), (2, -3, -5, -1), (3, -9, -5, -2)])
def test_batch_packed_scores_by_alphabet(sw, orc, ctx, alphabet, score):
    """The packed stage 1 takes substitution scores from 2-bit symbol-code profiles (one
    v_perm_b32 per cell pair) when the batch's bytes fit 4 codes and the profile bytes
    (match - gap_open, mismatch - gap_open) fit [0, 255], and from byte compares otherwise
    (ACGTN, upper and lower case, mismatch below gap_open): the oracle's alignments either way."""
    rng = random.Random(len(alphabet) * 100 + score[0])
    pairs = []
    for k in range(12):
        L = 300 + 220 * k
        a = bytes(rng.choice(alphabet) for _ in range(L))
        b = bytearray(a[L // 10: L - L // 12])
        for i in range(len(b)):
            if rng.random() < 0.1:
                b[i] = rng.choice(alphabet)
        pairs.append((a, bytes(b)))
    res, st = ctx.batch(pairs, sw.scoring(*score), K=256, Kc=256, tile_cols=128)
    assert all(x == 0 for x in st)
    for (a, b), r in zip(pairs, res):
        assert r["precision"] == 16
        A, B = orient(a, b)
        ref = orc.align_linear(A, B, orc.scoring(*score), 256, 256)
        for k in ("score", "end", "start", "cigar"):
            assert r[k] == ref[k], (k, len(a), len(b))


def test_batch_packed_high_scores_equal_int32(sw, ctx):
    """Pairs whose scores pass 32767 (near-identical 30-45 kbp pairs) in a packed batch, beside
    short pairs: stage 1 and the traceback codes run in int16 relative to each tile's base and
    give exactly the alignments of the int32 single-pair path (itself checked against the
    oracle elsewhere)."""
    pairs = []
    for k, L in enumerate((45000, 38000, 30500, 800, 120, 5000)):
        a, b = sw.synth_pair(900 + k, L, L - L // 50, L - L // 20, 0.01, 0.002)
        pairs.append((a, b))
    res, st = ctx.batch(pairs, sw.scoring(1, -3, -5, -2), xpoints=True, K=1024, Kc=1024)
    assert all(x == 0 for x in st)
    assert res[0]["precision"] == 16 and max(r["score"] for r in res) > 32767
    for (a, b), r in zip(pairs, res):
        one = ctx.align(a, b, sw.scoring(1, -3, -5, -2), xpoints=True, K=1024, Kc=1024, precision=32)
        for k in ("score", "end", "start", "cigar", "xpoints"):
            assert r[k] == one[k], (k, len(a), len(b))


def _traced_batch(sw, monkeypatch, capfd, pairs, **kw):
    """Runs a batch on a fresh context that traces its stages; returns the results and
    whether traceback rounds were planned on the device."""
    monkeypatch.setenv("SWALIGN_TRACE", "1")
    with sw.Context(0) as c:
        capfd.readouterr()
        res, st = c.batch(pairs, sw.scoring(*SC), **kw)
        err = capfd.readouterr().err
    return res, st, "device rounds" in err


def test_batch_device_planned_rounds(sw, orc, monkeypatch, capfd):
    """2600 pairs of 150-1200 bp at 256-row checkpoint blocks: enough walking pairs that the
    traceback rounds are planned, swept, walked and compacted on the device until the rest
    go to host-planned rounds with lookahead; every pair gets the oracle's alignment and
    crosspoints."""
    rng = random.Random(29)
    pairs = []
    for k in range(2600):
        L = rng.randint(150, 1200)
        a, b = sw.synth_pair(7000 + k, L, L + rng.randint(-L // 5, L // 5), max(1, L // 2), 0.1, 0.02)
        pairs.append((a, b) if rng.random() < 0.5 else (b, a))
    res, st, dev = _traced_batch(sw, monkeypatch, capfd, pairs, xpoints=True, K=256, Kc=256)
    assert dev
    _check(orc, pairs, res, st, K=256, Kc=256)


def test_batch_device_rounds_equal_host_rounds(sw, orc, ctx, monkeypatch, capfd):
    """Blocks of two stripes by four tiles (512 x 1024, 256-column tiles: two tiles of one
    pair on some launches): a 2400-pair batch whose rounds start on the device gives exactly
    what the same pairs give in batches of 400, whose rounds are all planned on the host,
    crosspoints included; every CIGAR re-scores to its score, and a sample equals the
    oracle's alignment."""
    rng = random.Random(31)
    pairs = []
    for k in range(2400):
        L = rng.randint(1300, 2600)
        a, b = sw.synth_pair(9000 + k, L, L + rng.randint(-L // 6, L // 6), L // 2, 0.1, 0.02)
        pairs.append((a, b))
    kw = dict(xpoints=True, K=512, Kc=1024, tile_cols=256)
    res, st, dev = _traced_batch(sw, monkeypatch, capfd, pairs, **kw)
    assert dev and all(s == 0 for s in st)
    for q in range(0, len(pairs), 400):
        part, pst = ctx.batch(pairs[q:q + 400], sw.scoring(*SC), **kw)
        assert all(s == 0 for s in pst)
        for x, r in enumerate(part):
            g = res[q + x]
            for k in ("score", "end", "start", "cigar", "xpoints", "swapped"):
                assert g[k] == r[k], (k, q + x)
    osc = orc.scoring(*SC)
    for (a, b), r in zip(pairs, res):
        A, B = orient(a, b)
        assert orc.rescore(A, B, osc, r["start"], r["cigar"]) == (r["score"], r["end"])
    for q in range(0, len(pairs), 97):
        A, B = orient(*pairs[q])
        ref = orc.align_linear(A, B, osc, 512, 1024)
        for k in ("score", "end", "start", "cigar"):
            assert res[q][k] == ref[k], (k, q)


@pytest.mark.parametrize("mode", ["", "2"])
@pytest.mark.parametrize("W,K,Kc,precision", [(64, 256, 64, 0), (192, 512, 256, 0), (320, 256, 128, 0),
                                              (512, 1024, 1024, 0), (256, 256, 64, 32), (448, 512, 512, 32)])
def test_batch_skewed_stripes(sw, orc, ctx, monkeypatch, W, K, Kc, precision, mode):
    """Batches run stage 1 in skewed stripes (segments shifted left by one column per lane)
    whenever the tile width is a multiple of 64, packed and int32, by default and with
    SWALIGN_SKEW=2 (shifted stripes for the int32 batches, skewed ones for the packed): pairs of
    one to many segments, checkpoint columns inside segments and at their ends, give the
    oracle's alignments."""
    if mode:
        monkeypatch.setenv("SWALIGN_SKEW", mode)
    pairs = _mixed_pairs(sw, 29 + W, 60, [5, 70, 300, 700, 1500, 2600])
    res, st = ctx.batch(pairs, sw.scoring(*SC), tile_cols=W, K=K, Kc=Kc, precision=precision, xpoints=True)
    _check(orc, pairs, res, st, K=K, Kc=Kc)
    assert all(r["precision"] == (precision or 16) for r, (a, b) in zip(res, pairs) if a and b)


def test_batch_skewed_equals_rectangular(sw, ctx, monkeypatch):
    """The same batch with skewed and with rectangular stage-1 stripes: identical results."""
    pairs = _mixed_pairs(sw, 31, 80, [40, 600, 2000, 3500])
    skewed, st1 = ctx.batch(pairs, sw.scoring(*SC), xpoints=True)
    monkeypatch.setenv("SWALIGN_SKEW", "0")
    rect, st2 = ctx.batch(pairs, sw.scoring(*SC), xpoints=True)
    assert st1 == st2 == [0] * len(pairs)
    for x, y in zip(skewed, rect):
        for k in ("score", "end", "start", "cigar", "xpoints"):
            assert x[k] == y[k], k


@pytest.mark.parametrize("min_tiles", ["", "64", "1000000"])
def test_sparse_launches_int32_in_packed_sweeps(sw, orc, ctx, monkeypatch, min_tiles):
    """Packed sweeps run a launch of few tiles in int32 (one tile per wave) and the others
    packed: with the default threshold (""), with launches of at least 64 tiles packed (so
    one sweep mixes both kernels, packed and int32 tiles handing state to each other through
    the same buffers), and with every launch in int32, a batch and a packed single pair give
    the oracle's alignments."""
    monkeypatch.setenv("SWALIGN_PACKED_MIN_TILES", min_tiles)
    pairs = _mixed_pairs(sw, 41, 120, [30, 400, 1500, 3000])
    res, st = ctx.batch(pairs, sw.scoring(*SC), tile_cols=256, K=256, Kc=256, xpoints=True)
    _check(orc, pairs, res, st, K=256, Kc=256)
    assert all(r["precision"] == 16 for r, (a, b) in zip(res, pairs) if a and b)
    (a, b) = sw.synth_pair(43, 9000, 8500, 4000, 0.08, 0.015)
    got = ctx.align(a, b, sw.scoring(*SC), K=512, Kc=256, xpoints=True, precision=16, tile_cols=128)
    A, B = orient(a, b)
    ref = orc.align_linear(A, B, orc.scoring(*SC), 512, 256)
    assert got["precision"] == 16
    assert (got["score"], got["end"], got["start"], got["cigar"]) == (ref["score"], tuple(ref["end"]),
                                                                     tuple(ref["start"]), ref["cigar"])
