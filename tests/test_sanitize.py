"""CPU sanitizer builds of the repository's host-only C++ (tests/sanitize/): the oracle, the
CLI's FASTA loader and alignment renderer, and the synthetic workload generator, built with
g++ under AddressSanitizer + UndefinedBehaviorSanitizer and, for the multithreaded oracle
paths, ThreadSanitizer. The driver exercises every entry point and checks them against each
other; any sanitizer report aborts it. The HIP library's host code needs a device and is not
covered here.
"""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

SAN = os.path.join(ROOT, "tests", "sanitize")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("sanitize")
    r = subprocess.run(["make", "-s", "-C", SAN, "OUT=%s" % out, "-j2"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def _run(args, tmp):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", TSAN_OPTIONS="halt_on_error=1:exitcode=87")
    r = subprocess.run(args + [os.path.join(GOLDEN, "fasta"), str(tmp)], capture_output=True, text=True,
                       env=env, timeout=300)
    return r


def test_asan_ubsan_clean(built, tmp_path):
    r = _run([os.path.join(built, "driver_asan")], tmp_path)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr


def test_tsan_clean(built, tmp_path):
    r = subprocess.run([os.path.join(built, "driver_tsan"), os.path.join(GOLDEN, "fasta"), str(tmp_path),
                        "mt-only"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=87"))
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), (r.stdout[-2000:], r.stderr[-4000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr
