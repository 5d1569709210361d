import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

# Packed sweeps run a launch of few tiles in int32 (one tile per wave) by default; the tests'
# inputs are small, so they ask for every launch of a packed sweep to run packed, which keeps
# the packed kernel under test.  Tests of the default choice (test_gpu_configs.py,
# test_sparse_launches_*) remove this.
os.environ.setdefault("SWALIGN_PACKED_MIN_TILES", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def sw():
    return _load("swalign", os.path.join(ROOT, "projekt-2012_amd", "swalign.py"))


@pytest.fixture(scope="session")
def orc():
    return _load("pyoracle", os.path.join(ROOT, "oracle", "pyoracle.py"))


@pytest.fixture(scope="session")
def ctx(sw):
    c = sw.Context(0)
    yield c
    c.close()


def fasta_bytes(name):
    """Reads a golden FASTA (optionally .gz) with the reference's rules: the first line
    (up to 299 bytes) is the header, the sequence is every later byte except CR/LF."""
    import gzip
    p = os.path.join(GOLDEN, "fasta", name)
    if not os.path.exists(p):
        p += ".gz"
        data = gzip.open(p, "rb").read()
    else:
        data = open(p, "rb").read()
    assert data[:1] == b">"
    k = 0
    while k < len(data) and k < 299:
        k += 1
        if data[k - 1:k] == b"\n":
            break
    return data[k:].replace(b"\n", b"").replace(b"\r", b"")


def orient(a, b):
    """Rows = the longer sequence; equal lengths keep the argument order."""
    return (a, b) if len(a) >= len(b) else (b, a)
