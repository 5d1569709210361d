"""The allocation-free entry point (swh_align_pair_device_noalloc) against the allocating
one (swh_align_pair_device), through the C-ABI on an MI355X: bit-exact score, end, start
and CIGAR on two calls in a row, the second with a smaller input (the scratch array keeps
the first call's contents), and a pair whose working set exceeds the scratch array is
rejected with SWH_ERR_SCRATCH before any work."""
import pytest

pytestmark = pytest.mark.gpu

SC = (1, -3, -5, -2)


def _key(r):
    return (r["score"], r["end"], r["start"], r["cigar"], r["n_ops"], r["swapped"])


def _dev(sw, a, b):
    # exact-size device buffers: the library may read nothing past either sequence
    return sw.DeviceBuffer(a), sw.DeviceBuffer(b)


def test_noalloc_equals_allocating_entry_point(sw, ctx):
    sc = sw.scoring(*SC)
    cases = [
        (sw.synth_pair(4101, 61_000, 57_000, 30_000, 0.08, 0.015), dict(K=4096, Kc=4096)),
        (sw.synth_pair(4102, 9_000, 11_500, 5_000, 0.10, 0.02), dict(K=1024, Kc=512, lookahead=3)),
    ]
    for (a, b), kw in cases:
        da, db = _dev(sw, a, b)
        try:
            want = ctx.align_device(da.address, len(a), db.address, len(b), sc, **kw)
            got = ctx.align_device_noalloc(da.address, len(a), db.address, len(b), sc, **kw)
            assert want["score"] > 0
            assert _key(got) == _key(want)
            assert (got["checkpoint_rows"], got["checkpoint_cols"]) == (want["checkpoint_rows"],
                                                                       want["checkpoint_cols"])
            # default spacing: the noalloc plan may be sparser; the alignment is the same
            got2 = ctx.align_device_noalloc(da.address, len(a), db.address, len(b), sc)
            assert _key(got2) == _key(want)
            # score only
            so = ctx.align_device_noalloc(da.address, len(a), db.address, len(b), sc, score_only=True)
            assert (so["score"], so["end"]) == (want["score"], want["end"])
        finally:
            da.free()
            db.free()


def test_noalloc_rejects_pairs_beyond_the_scratch(sw, ctx):
    sc = sw.scoring(*SC)
    a, b = sw.synth_pair(4103, 400_000, 380_000, 1000, 0.1, 0.02)
    da, db = _dev(sw, a, b)
    try:
        with pytest.raises(sw.SWError) as e:
            ctx.align_device_noalloc(da.address, len(a), db.address, len(b), sc)
        assert e.value.status == sw.SWH_ERR_SCRATCH
        # a given checkpoint spacing is kept, and rejected when it does not fit
        with pytest.raises(sw.SWError) as e:
            ctx.align_device_noalloc(da.address, 300_000, db.address, 300_000, sc, K=256, Kc=256)
        assert e.value.status == sw.SWH_ERR_SCRATCH
        with pytest.raises(sw.SWError) as e:
            ctx.align_device_noalloc(da.address, 5000, db.address, 5000, sc, xpoints=True)
        assert e.value.status == sw.SWH_ERR_INVALID_ARG
    finally:
        da.free()
        db.free()
