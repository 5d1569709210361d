"""Request server (swh_server_*) on an MI355X: jobs submitted while others run are admitted
between launches and share them; every result is exactly the oracle's canonical alignment
(the result the job gets alone), admission respects the HBM pool, and a runtime failure
ends every queued and running job instead of leaving a poll waiting."""
import os
import random
import threading
import time

import pytest

from conftest import orient

pytestmark = pytest.mark.gpu

SC = (1, -3, -5, -2)


def _pairs(sw, seed, n, lens):
    rng = random.Random(seed)
    out = []
    for k in range(n):
        L = rng.choice(lens)
        L2 = max(0, L + rng.randint(-L // 3 - 1, L // 3 + 1))
        out.append(sw.synth_pair(seed * 1000 + k, L, L2, max(1, min(L, L2) // 2), 0.1, 0.02))
    return out


def _key(r):
    return (r["score"], r["end"], r["start"], r["cigar"], r["swapped"])


def _oracle_keys(orc, pairs, score=SC):
    """The oracle's canonical alignment of each pair (K-independent, so any checkpoint spacing)."""
    out = []
    for a, b in pairs:
        A, B = orient(a, b)
        if not A or not B:
            out.append((0, (-1, -1), (-1, -1), None, len(a) < len(b)))
            continue
        r = orc.align_linear(A, B, orc.scoring(*score), 1024, 1024)
        if r["score"] == 0:
            out.append((0, (-1, -1), (-1, -1), None, len(a) < len(b)))
        else:
            out.append((r["score"], r["end"], r["start"], r["cigar"], len(a) < len(b)))
    return out


def _served_key(r):
    if r["score"] == 0:
        return (0, (-1, -1), (-1, -1), None, r["swapped"])
    return _key(r)


def test_served_equals_oracle(sw, orc, ctx):
    """Jobs submitted from another thread while earlier ones run: each served result is the
    oracle's alignment, and also the one the job gets aligned alone; the server's counters
    show jobs admitted while others were mid-way and sweep launches shared between jobs."""
    pairs = _pairs(sw, 3, 40, [0, 1, 50, 300, 1200, 3000, 6000])
    want = _oracle_keys(orc, pairs)
    alone = [_served_key(ctx.align(a, b, sw.scoring(*SC))) for a, b in pairs]
    assert alone == want
    rng = random.Random(5)
    with sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=16) as srv:
        ids = []

        def submitter():
            for a, b in pairs:
                ids.append(srv.submit(a, b))
                time.sleep(rng.random() * 0.004)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        got = []
        for jid in ids:
            st, r = srv.poll(jid)
            assert st == 0
            got.append(_served_key(r))
        assert got == want
        stats = srv.stats()
        assert stats["completed"] == len(pairs) and stats["failed"] == 0
        assert 0 <= stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
        assert stats["hbm_in_use_bytes"] == 0 and stats["loop_status"] == 0
        assert stats["peak_in_flight"] > 1 and stats["admitted_mid_flight"] > 0
        assert 0 < stats["shared_launches"] <= stats["sweep_launches"]


def test_pool_bounds_admission(sw, orc):
    pairs = _pairs(sw, 7, 12, [100, 300, 500])
    alone = _oracle_keys(orc, pairs)
    big = sw.synth_pair(99, 50_000, 50_000, 1000, 0.1, 0.02)
    with sw.Server(sw.scoring(*SC), budget=1 << 20, max_in_flight=2) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        big_id = srv.submit(*big)
        assert [_served_key(srv.poll(j)[1]) for j in ids] == alone
        st, r = srv.poll(big_id)
        assert st == sw.SWH_ERR_BUDGET and r is None
        assert srv.poll(big_id)[0] == sw.SWH_ERR_NOT_FOUND


def test_poll_without_wait(sw):
    with sw.Server(sw.scoring(*SC), budget=1 << 28) as srv:
        a, b = sw.synth_pair(11, 4000, 3500, 2000, 0.1, 0.02)
        jid = srv.submit(a, b)
        st, r = srv.poll(jid, wait=False)
        assert st in (0, sw.SWH_PENDING)
        deadline = time.time() + 60
        while st == sw.SWH_PENDING and time.time() < deadline:
            time.sleep(0.001)
            st, r = srv.poll(jid, wait=False)
        assert st == 0 and r["score"] > 0


def test_served_mixed_alphabets_equal_oracle(sw, orc):
    """Jobs of upper and lower case bytes beside ACGT jobs: ticks whose jobs' bytes no 3-bit
    code separates score by byte compares, the others by symbol codes; every result equals
    the oracle's alignment."""
    rng = random.Random(11)
    pairs = _pairs(sw, 13, 10, [200, 900, 2500])
    for k in range(4):
        L = 600 + 400 * k
        a = bytes(rng.choice(b"ACGTacgt") for _ in range(L))
        b = bytearray(a[L // 10:])
        for i in range(len(b)):
            if rng.random() < 0.1:
                b[i] = rng.choice(b"ACGTacgt")
        pairs.insert(2 * k + 1, (a, bytes(b)))
    want = _oracle_keys(orc, pairs)
    with sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=6) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        got = []
        for jid in ids:
            st, r = srv.poll(jid)
            assert st == 0
            got.append(_served_key(r))
    assert got == want


@pytest.mark.parametrize("wps", [1, 2])
def test_served_waves_per_simd_same_result(sw, orc, wps):
    """The server's shared launches at one or two sweep waves per SIMD (the default is four)
    give every job the oracle's alignment."""
    pairs = _pairs(sw, 17, 12, [150, 800, 2500, 7000])
    want = _oracle_keys(orc, pairs)
    with sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=8, waves_per_simd=wps) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        got = [_served_key(srv.poll(j)[1]) for j in ids]
    assert got == want


def test_runtime_failure_ends_queued_jobs(sw, monkeypatch):
    """A runtime failure of the server loop (forced from the host at its third tick through
    SWALIGN_SERVE_FAIL_TICK; no kernel is affected) finishes every job: running and queued
    jobs report the failure to their polls, which return instead of waiting, later submits
    are refused, and the statistics carry the loop's status."""
    pairs = _pairs(sw, 29, 8, [2000, 3000])
    monkeypatch.setenv("SWALIGN_SERVE_FAIL_TICK", "2")
    srv = sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=1)
    monkeypatch.delenv("SWALIGN_SERVE_FAIL_TICK")
    try:
        ids = [srv.submit(a, b) for a, b in pairs]
        t0 = time.time()
        statuses = [srv.poll(j)[0] for j in ids]
        assert time.time() - t0 < 60
        assert set(statuses) <= {0, sw.SWH_ERR_INTERNAL}
        assert statuses[-1] == sw.SWH_ERR_INTERNAL  # still queued when the loop stopped
        stats = srv.stats()
        assert stats["loop_status"] == sw.SWH_ERR_INTERNAL
        assert stats["completed"] + stats["failed"] == len(pairs)
        with pytest.raises(sw.SWError):
            srv.submit(*pairs[0])
    finally:
        srv.close()
