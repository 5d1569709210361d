"""Parity at the benchmark workloads' own sizes (run on an MI355X), integer arithmetic, so
every comparison is exact:

* the 1 Mbp x 1 Mbp pair against the oracle's alignment of it (golden fixture: score, end,
  start, CIGAR by length and SHA-256, crosspoints at checkpoint rows every 8192 rows), in
  int32 and in packed 16-bit stage 1;
* the 10 Mbp x 10 Mbp pair (last, so that its checkpoints are released when the module
  ends), too large for the CPU oracle (1e14 cells), through properties
  that do not depend on size: its CIGAR re-scores to its score from its start and ends at
  its end cell, and int32 vs packed 16-bit stage 1 and two checkpoint spacings give the
  same score, end, start and CIGAR;
* the whole 65,536-pair mixed-length batch: every pair's CIGAR re-scores to its score
  between its start and end, and 500 pairs at evenly spaced ranks of their cell counts equal
  the oracle's alignments (golden fixture);
* the served stream's first 40 jobs (one 0.70 x 0.71 Mbp job among 1-100 kbp jobs) through
  the request server with a 64 GB pool, submitted as a Poisson stream: every result equals
  the oracle's alignment (golden fixture), and jobs overlapped on the device.
"""
import hashlib
import json
import os
import random
import threading
import time

import pytest

from conftest import GOLDEN, orient

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def default_launches():
    """The workloads run with the library's defaults: a packed sweep's launches of few tiles
    run in int32 (conftest asks other tests to run them packed)."""
    old = os.environ.pop("SWALIGN_PACKED_MIN_TILES", None)
    yield
    if old is not None:
        os.environ["SWALIGN_PACKED_MIN_TILES"] = old


def _digest(r):
    c = r["cigar"] or ""
    return dict(score=r["score"], end=list(r["end"]), start=list(r["start"]), n_ops=r["n_ops"],
                cigar_len=len(c), cigar_sha256=hashlib.sha256(c.encode()).hexdigest())


def _same(got, rec):
    d = _digest(got)
    for k in ("score", "end", "start", "n_ops", "cigar_len", "cigar_sha256"):
        assert d[k] == rec[k], (k, d[k], rec[k])


@pytest.mark.parametrize("precision", [32, 16])
def test_config1_pair_equals_oracle(sw, precision):
    g = json.load(open(os.path.join(GOLDEN, "config1.json")))
    (a, b), sc = sw.workload_pair(1)
    with sw.Context(0) as c:
        got = c.align(a, b, sc, K=g["K"], Kc=g["Kc"], xpoints=True, precision=precision)
    assert got["precision"] == precision and (got["m"], got["n"]) == (g["m"], g["n"])
    _same(got, g)
    assert got["xpoints"] == [tuple(x) for x in g["xpoints"]]


def test_config3_full_batch(sw, orc):
    g = json.load(open(os.path.join(GOLDEN, "batch_sample.json")))
    pairs, sc = sw.workload_batch()
    assert len(pairs) == g["npairs"] == 65536
    dev = sw.DeviceBuffer(b"".join(a + b for a, b in pairs))
    base, off, dp = dev.address, 0, []
    for a, b in pairs:
        dp.append((base + off, len(a), base + off + len(a), len(b)))
        off += len(a) + len(b)
    with sw.Context(0) as c:
        print("config 3: aligning 65,536 pairs", flush=True)
        res, st = c.batch_device(dp, sc, K=g["K"], Kc=g["K"], xpoints=True)
    dev.free()
    assert st == [0] * len(pairs)
    osc = orc.scoring(1, -3, -5, -2)
    for (a, b), r in zip(pairs, res):
        A, B = orient(a, b)
        assert r["score"] > 0
        s, end = orc.rescore(A, B, osc, r["start"], r["cigar"])
        assert s == r["score"] and end == r["end"], (len(a), len(b))
    for rec in g["pairs"]:
        r = res[rec["index"]]
        assert (r["m"], r["n"]) == (rec["m"], rec["n"])
        _same(r, rec)
        assert r["xpoints"] == [tuple(x) for x in rec["xpoints"]], rec["index"]


def test_config4_served_stream(sw):
    g = json.load(open(os.path.join(GOLDEN, "serve_jobs.json")))
    jobs, sc = sw.workload_serve(g["njobs"])
    rng = random.Random(2016)
    print("config 4: serving %d jobs" % len(jobs), flush=True)
    with sw.Server(sc, budget=64 << 30) as srv:
        ids = []

        def submitter():  # Poisson arrivals, mean gap 5 ms
            for a, b in jobs:
                ids.append(srv.submit(a, b))
                time.sleep(rng.expovariate(200.0))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        got = []
        for jid in ids:
            st, r = srv.poll(jid)
            assert st == 0
            got.append(r)
        stats = srv.stats()
    for r, rec in zip(got, g["jobs"]):
        assert (r["m"], r["n"]) == (rec["m"], rec["n"])
        _same(r, rec)
    assert stats["completed"] == len(jobs) and stats["loop_status"] == 0
    assert stats["admitted_mid_flight"] > 0 and stats["peak_in_flight"] > 1
    assert stats["shared_launches"] > 0


@pytest.fixture(scope="module")
def chromosome(sw):
    """The 10 Mbp pair and its int32 alignment at checkpoint rows every 8192 rows (a context
    of its own: its ~146 GB of checkpoints are released when the module's tests end)."""
    (a, b), sc = sw.workload_pair(2)
    c = sw.Context(0)
    print("config 2: aligning the 10 Mbp pair (int32, K = 8192)", flush=True)
    r = c.align(a, b, sc, K=8192, precision=32)
    yield a, b, sc, c, r
    c.close()


def test_config2_cigar_rescores(orc, chromosome):
    a, b, sc, c, r = chromosome
    A, B = orient(a, b)
    assert r["score"] > 3_000_000 and r["precision"] == 32
    s, end = orc.rescore(A, B, orc.scoring(2, -3, -5, -1), r["start"], r["cigar"])
    assert s == r["score"] and end == r["end"]
    # the alignment starts and ends on aligned pairs of the two sequences
    assert r["cigar"].rstrip("0123456789")[-1] in "=X" and r["cigar"].lstrip("0123456789")[0] in "=X"


def test_config2_packed_equals_int32(chromosome):
    """The default precision of this pair is the packed 16-bit stage 1 in skewed stripes
    (the library's occupancy model prefers it for long pairs): the int32 alignment."""
    a, b, sc, c, r = chromosome
    print("config 2: aligning the 10 Mbp pair (default: packed 16-bit stage 1)", flush=True)
    r16 = c.align(a, b, sc, K=8192)
    assert r16["precision"] == 16
    for k in ("score", "end", "start", "cigar"):
        assert r16[k] == r[k], k


def test_config2_checkpoint_spacing_independent(chromosome):
    a, b, sc, c, r = chromosome
    print("config 2: aligning the 10 Mbp pair (K = Kc = 16384)", flush=True)
    r2 = c.align(a, b, sc, K=16384, Kc=16384, precision=32)
    assert r2["checkpoint_rows"] == 16384
    for k in ("score", "end", "start", "cigar"):
        assert r2[k] == r[k], k
