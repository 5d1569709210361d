"""Request server (swh_server_*) on an MI355X: jobs submitted while others run are admitted
between launches and share them; every result is exactly the one the job gets alone, and
admission respects the HBM pool."""
import random
import threading
import time

import pytest

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


def test_served_equals_alone(sw, ctx):
    pairs = _pairs(sw, 3, 40, [0, 1, 50, 300, 1200, 3000, 6000])
    alone = [_key(ctx.align(a, b, sw.scoring(*SC))) for a, b in pairs]
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
            got.append(_key(r))
        assert got == alone
        stats = srv.stats()
        assert stats["completed"] == len(pairs) and stats["failed"] == 0
        assert 0 <= stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
        assert stats["hbm_in_use_bytes"] == 0 and stats["loop_status"] == 0


def test_pool_bounds_admission(sw, ctx):
    pairs = _pairs(sw, 7, 12, [100, 300, 500])
    alone = [_key(ctx.align(a, b, sw.scoring(*SC))) for a, b in pairs]
    big = sw.synth_pair(99, 50_000, 50_000, 1000, 0.1, 0.02)
    with sw.Server(sw.scoring(*SC), budget=1 << 20, max_in_flight=2) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        big_id = srv.submit(*big)
        assert [_key(srv.poll(j)[1]) for j in ids] == alone
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


def test_served_mixed_alphabets_equal_alone(sw, ctx):
    """Jobs of upper and lower case bytes beside ACGT jobs: ticks whose jobs' bytes no 3-bit
    code separates score by byte compares, the others by symbol codes; every result equals
    the job aligned alone."""
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
    alone = [_key(ctx.align(a, b, sw.scoring(*SC))) for a, b in pairs]
    with sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=6) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        got = []
        for jid in ids:
            st, r = srv.poll(jid)
            assert st == 0
            got.append(_key(r))
    assert got == alone


@pytest.mark.parametrize("wps", [1, 2])
def test_served_waves_per_simd_same_result(sw, ctx, wps):
    """The server's shared launches at one or two sweep waves per SIMD (the default is four)
    give every job the result it gets alone."""
    pairs = _pairs(sw, 17, 12, [150, 800, 2500, 7000])
    alone = [_key(ctx.align(a, b, sw.scoring(*SC))) for a, b in pairs]
    with sw.Server(sw.scoring(*SC), budget=1 << 30, max_in_flight=8, waves_per_simd=wps) as srv:
        ids = [srv.submit(a, b) for a, b in pairs]
        got = [_key(srv.poll(j)[1]) for j in ids]
    assert got == alone
