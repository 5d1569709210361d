"""Writes the golden fixtures under tests/golden/.

* fasta/: the reference's own test inputs (Sequences/0..5 of Sterbic/Projekt-2012),
  copied verbatim as data; files over 100 kB are stored gzip-compressed.
* kats.json: the expected "Alignment score: S at [r, c]" lines of Tests/test1..5.sh.
* config0.json / small_pairs.json: oracle outputs (score, end, start, CIGAR) for the
  10 kbp x 10 kbp synthetic pair and for seeded random small pairs.
* config1.json: the oracle's alignment of the 1 Mbp x 1 Mbp synthetic pair (score, end,
  start, CIGAR length and SHA-256, crosspoints at checkpoint rows every 8192 rows).
* batch_sample.json: the oracle's alignments of 500 pairs of the 65,536-pair mixed-length
  batch, picked at evenly spaced ranks of their cell counts (checkpoints every 1024 rows).
* serve_jobs.json: the oracle's alignments of the first 40 jobs of the served stream.

Run from the repository root with the oracle built (make -C oracle) and the library
built (for the synthetic generator):
    python tests/golden/make_golden.py [small] [config1] [batch] [serve]
(no argument: all).  The large fixtures take a few minutes on 8 host threads.
"""
import gzip
import hashlib
import json
import os
import random
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "projekt-2012_amd"))
import pyoracle as O  # noqa: E402
import swalign as S  # noqa: E402

KATS = [
    dict(name="test1", a="1/A00142.fna", b="1/A00145.fna", scoring=[5, -5, -5, -5], score=480, end=[623, 561]),
    dict(name="test2", a="2/NC_005046.fna", b="2/NC_012557.fna", scoring=[5, -5, -5, -5], score=1580, end=[1389, 1351]),
    dict(name="test3", a="3/NC_000898.fna", b="3/NC_007605.fna", scoring=[1, -3, -5, -2], score=18, end=[44353, 41058]),
    dict(name="test4", a="4/AE002160.2.fna", b="4/CP000051.1.fna", scoring=[1, -3, -5, -2], score=88353, end=[1072950, 722725]),
    dict(name="test5", a="5/BA000035.2.fna", b="5/BX927147.1.fna", scoring=[1, -3, -5, -2], score=4226, end=[2689488, 2991493]),
]


def copy_fasta():
    out = os.path.join(HERE, "fasta")
    os.makedirs(out, exist_ok=True)
    for d in range(6):
        src = os.path.join(REF, "Sequences", str(d))
        for f in sorted(os.listdir(src)):
            p = os.path.join(src, f)
            if os.path.getsize(p) > 100_000:
                with open(p, "rb") as fi, gzip.open(os.path.join(out, f + ".gz"), "wb", 9) as fo:
                    shutil.copyfileobj(fi, fo)
            else:
                shutil.copy(p, os.path.join(out, f))


THREADS = min(16, os.cpu_count() or 4)


def digest(r):
    """A fixture record of an oracle alignment: the CIGAR by length and SHA-256."""
    c = r["cigar"] or ""
    return dict(score=r["score"], end=list(r["end"]), start=list(r["start"]), n_ops=r["n_ops"],
                cigar_len=len(c), cigar_sha256=hashlib.sha256(c.encode()).hexdigest())


def config1():
    (a, b), _ = S.workload_pair(1)
    A, B = (a, b) if len(a) >= len(b) else (b, a)
    r = O.align_linear(A, B, O.scoring(1, -3, -5, -2), 8192, 8192, threads=THREADS)
    rec = dict(seed=2013, scoring=[1, -3, -5, -2], m=len(A), n=len(B), K=8192, Kc=8192, **digest(r),
               xpoints=[list(x) for x in r["xpoints"]])
    json.dump(rec, open(os.path.join(HERE, "config1.json"), "w"))
    print("config1:", rec["score"], rec["end"], rec["start"], rec["n_ops"])


def batch_sample(count=500):
    """Pairs of the batched workload at evenly spaced ranks of their cell counts."""
    from concurrent.futures import ThreadPoolExecutor
    pairs, _ = S.workload_batch()
    order = sorted(range(len(pairs)), key=lambda k: (len(pairs[k][0]) * len(pairs[k][1]), k))
    picks = [order[(q * (len(order) - 1)) // (count - 1)] for q in range(count)]
    sc = O.scoring(1, -3, -5, -2)

    def one(k):
        a, b = pairs[k]
        A, B = (a, b) if len(a) >= len(b) else (b, a)
        r = O.align_linear(A, B, sc, 1024, 1024)
        return dict(index=k, m=len(A), n=len(B), **digest(r), xpoints=[list(x) for x in r["xpoints"]])

    with ThreadPoolExecutor(THREADS) as ex:  # ctypes releases the GIL during each alignment
        recs = list(ex.map(one, picks))
    json.dump(dict(npairs=len(pairs), seed=2015, scoring=[1, -3, -5, -2], K=1024, pairs=recs),
              open(os.path.join(HERE, "batch_sample.json"), "w"))
    print("batch sample: %d pairs" % len(recs))


def serve_jobs(count=40):
    jobs, _ = S.workload_serve(count)
    sc = O.scoring(1, -3, -5, -2)
    recs = []
    for k, (a, b) in enumerate(jobs):
        A, B = (a, b) if len(a) >= len(b) else (b, a)
        K = 8192 if len(A) >= 262144 else 1024
        r = O.align_linear(A, B, sc, K, K, threads=THREADS)
        recs.append(dict(index=k, m=len(A), n=len(B), **digest(r)))
    json.dump(dict(njobs=count, seed=2016, scoring=[1, -3, -5, -2], jobs=recs),
              open(os.path.join(HERE, "serve_jobs.json"), "w"))
    print("serve jobs: %d" % len(recs))


def main():
    what = set(sys.argv[1:]) or {"small", "config1", "batch", "serve"}
    if "small" in what:
        small()
    if "config1" in what:
        config1()
    if "batch" in what:
        batch_sample()
    if "serve" in what:
        serve_jobs()


def small():
    copy_fasta()
    kats = [dict(k, a=os.path.basename(k["a"]), b=os.path.basename(k["b"])) for k in KATS]
    json.dump(kats, open(os.path.join(HERE, "kats.json"), "w"), indent=1)

    (a, b), _ = S.workload_pair(0)
    sc = O.scoring(1, -3, -5, -2)
    A, B = (a, b) if len(a) >= len(b) else (b, a)
    r = O.align_full(A, B, sc)
    json.dump(dict(seed=2012, scoring=[1, -3, -5, -2], m=len(A), n=len(B), score=r["score"], end=r["end"],
                   start=r["start"], cigar=r["cigar"]), open(os.path.join(HERE, "config0.json"), "w"))

    rng = random.Random(7)
    pairs = []
    for t in range(300):
        m = rng.randint(1, 300)
        n = rng.randint(1, 300)
        alpha = rng.choice(["AC", "ACGT", "ACGTN", "acgtACGT"])
        a = "".join(rng.choice(alpha) for _ in range(m))
        b = "".join(rng.choice(alpha) for _ in range(n))
        if rng.random() < 0.6 and m > 20 and n > 20:
            i = rng.randint(0, m - 10)
            seg = list(a[i:i + rng.randint(10, m - i)])
            for k in range(len(seg)):
                if rng.random() < 0.1:
                    seg[k] = rng.choice(alpha)
            j = rng.randint(0, n)
            b = (b[:j] + "".join(seg) + b[j:])[: max(n, 1)]
        s = [rng.randint(1, 5), -rng.randint(1, 5), -rng.randint(1, 8), -rng.randint(1, 3)]
        A, B = (a, b) if len(a) >= len(b) else (b, a)
        r = O.align_full(A, B, O.scoring(*s))
        pairs.append(dict(a=a, b=b, scoring=s, score=r["score"], end=r["end"], start=r["start"], cigar=r["cigar"]))
    json.dump(pairs, open(os.path.join(HERE, "small_pairs.json"), "w"))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
