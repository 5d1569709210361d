"""C-ABI surface of libswalign_hip.so, checked without a GPU: every function declared in
include/swalign_hip.h is exported, error statuses are returned (never exit()), and the
host-side synthetic generator is deterministic and follows its mutation model."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "swalign_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("swh_ctx_create", "swh_align_pair", "swh_align_pair_device", "swh_forward", "swh_batch_align",
                 "swh_batch_align_device", "swh_server_create", "swh_server_submit", "swh_server_poll",
                 "swh_server_stats", "swh_server_destroy", "swh_batch_tune",
                 "swh_result_free", "swh_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol(sw):
    L = C.CDLL(sw.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert {n for n, _, _ in sw.EXPORTS} == set(declared_functions())


def test_status_strings(sw):
    L = sw.lib()
    for code in range(0, 11):
        assert L.swh_strerror(code) and L.swh_strerror(code) != b"Unknown status."
    assert b"scoring" in L.swh_strerror(sw.SWH_ERR_SCORING)


def test_invalid_arguments_return_status(sw):
    L = sw.lib()
    r = sw.Result()
    sc = sw.scoring(1, -3, -5, -2)
    o = sw.make_opts()
    assert L.swh_align_pair(None, b"A", 1, b"A", 1, C.byref(sc), C.byref(o), C.byref(r)) == sw.SWH_ERR_INVALID_ARG
    assert L.swh_ctx_create(0, None) == sw.SWH_ERR_INVALID_ARG
    assert L.swh_align_pair_device_noalloc(None, None, 0, None, 0, C.byref(sc), C.byref(o), C.byref(r),
                                           None, 0) == sw.SWH_ERR_INVALID_ARG


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is present")
def test_no_device_is_reported(sw):
    h = C.c_void_p()
    assert sw.lib().swh_ctx_create(0, C.byref(h)) == sw.SWH_ERR_NO_DEVICE


def test_synthetic_pair_is_deterministic(sw):
    a1, b1 = sw.synth_pair(2012, 10_000, 10_000, 3_000, 0.10, 0.02)
    a2, b2 = sw.synth_pair(2012, 10_000, 10_000, 3_000, 0.10, 0.02)
    assert (a1, b1) == (a2, b2) and len(a1) == len(b1) == 10_000
    assert set(a1) <= set(b"ACGT")
    a3, _ = sw.synth_pair(2013, 10_000, 10_000, 3_000, 0.10, 0.02)
    assert a3 != a1


def test_gc_content_and_repeats(sw):
    a, b = sw.synth_pair(7, 200_000, 200_000, 0, 0.0, 0.0, gc=0.40, repeat_frac=0.01)
    gc = (a.count(b"G") + a.count(b"C")) / len(a)
    assert abs(gc - 0.40) < 0.01


def test_mutation_model(sw):
    import random
    rng = random.Random(1)
    src = bytes(rng.choice(b"ACGT") for _ in range(200_000))
    out = sw.synth_mutate(3, src, 0.10, 0.0, 0.0)
    assert len(out) == len(src)
    diff = sum(1 for x, y in zip(src, out) if x != y) / len(src)
    assert abs(diff - 0.10) < 0.01
    out = sw.synth_mutate(4, src, 0.0, 0.02, 0.0)
    # 2% indel events, mean length 3, insertions and deletions balanced: length ~ unchanged
    assert abs(len(out) - len(src)) < 0.02 * len(src)
    out = sw.synth_mutate(5, src[:10_000], 0.0, 0.0, 0.2)
    assert 0.6 * 10_000 < len(out) < 1.4 * 10_000


def test_struct_layouts_match_the_header(sw, tmp_path):
    """The ctypes mirrors in swalign.py have the size and field offsets of the C structs (a C
    program built with gcc against include/swalign_hip.h prints them)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    mirrors = {"swh_scoring": sw.Scoring, "swh_opts": sw.Opts, "swh_xpoint": sw.XPoint, "swh_result": sw.Result,
               "swh_pair": sw.Pair, "swh_batch_tuning": sw.BatchTuning, "swh_serve_stats": sw.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "swalign_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            if f.startswith("_") or f.startswith("reserved"):
                continue
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for cname, py in mirrors.items():
        assert int(out[cname + " size"]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            if f.startswith("_") or f.startswith("reserved"):
                continue
            assert int(out["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
