"""Benchmark: SWalign on one MI355X.

The metric is GCUPS on one MI355X for a single long pair and a batched mix, plus served
p50/p99 latency against the plain path.  One default run reports all of it on one JSON line:

* value: the 1 Mbp x 1 Mbp pair (200 kbp homologous region at 8% substitutions / 1.5%
  indels, scoring +1/-3/-5/-2, checkpoint rows every 8192 rows).  A step is one full
  alignment (forward sweep with checkpoints, end-cell search, traceback to the CIGAR) with
  the inputs resident in device memory; value = m*n / (seconds per step) / 1e9;
* chromosome: the 10 Mbp x 10 Mbp pair, full traceback, timed after one untimed alignment;
* batched: the 65,536-pair mixed-length batch as one step, against the plain per-pair path;
* served: the served stream (Poisson arrivals at 50/80/95% of the measured capacity, 200 jobs
  per rate, 64 GB pool) with p50/p99 latency, against the plain path (each job alone);
* served_mixed: the batched workload's pair mix (4096 pairs) as a Poisson stream through the
  same server, against the plain per-pair path.

Each record carries a roofline object for its stage-1 sweep kernel (integer-VALU peak, 12
integer ops per cell, kernel time from HIP events on the library's stream) and the top level
a cpu_baseline: the CPU oracle on the 10 kbp x 10 kbp pair with full traceback (one thread
and the host's cores, median of 5) and its forward pass over a row band of the 1 Mbp pair.
`--workload N` runs one workload alone (the profiles' form).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "projekt-2012_amd"))

OPS_PER_CELL = 12           # 2 substitution + 3 E + 3 F + 4 H integer ops per cell
LANES_PER_CU_CLK = 64       # 4 SIMD x 16 lanes; one 32-bit VALU op per lane per clock
PEAK_CUS = 256
PEAK_CLOCK_GHZ = 2.4
PEAK_INT32_TOPS = PEAK_CUS * LANES_PER_CU_CLK * PEAK_CLOCK_GHZ * 1e9 / 1e12   # 39.32
PEAK_INT16X2_TOPS = 2 * PEAK_INT32_TOPS   # packed 16-bit: two operations per lane per clock

WORKLOADS = {
    0: "10 kbp x 10 kbp pair, 3 kbp homology (10% sub, 2% indel), +1/-3/-5/-2, full traceback",
    1: "1 Mbp x 1 Mbp pair, 200 kbp homology (8% sub, 1.5% indel), +1/-3/-5/-2, "
       "checkpoint rows every 8192, full traceback",
    2: "10 Mbp x 10 Mbp pair, GC 40%, tandem repeats over 1%, 2 Mbp homology (5% sub, 1% indel), "
       "+2/-3/-5/-1, full traceback",
    3: "65,536 pairs, target lengths log-uniform in 100 bp - 20 kbp, query = target at 10% sub / 3% indel "
       "with flanks trimmed or extended by up to 20%, +1/-3/-5/-2, full traceback, all pairs in one batch",
    4: "served stream: Poisson arrivals at 50/80/95% of the measured capacity; each side 70% U[1,10] kbp, "
       "25% U[10,100] kbp, 5% U[100 kbp,1 Mbp]; queries mutated as in workload 3; +1/-3/-5/-2, full traceback; "
       "HBM pool 64 GB",
}
METRIC = "GCUPS on 1 MI355X (single long pair, batched mix) + served p50/p99 vs plain"


PROFILES = os.path.join(ROOT, "profiles")


def _counter_avg(fname, counter, kernel_part):
    """Average per dispatch of `counter` for the kernel whose name contains kernel_part, read
    from a committed rocprofv3 summary (profiles/*.txt), or None."""
    path = os.path.join(PROFILES, fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = [x.strip() for x in line.split("|")]
        if len(f) == 5 and f[0] == counter and kernel_part in f[1]:
            return float(f[4])
    return None


def _stage1_dispatches(fname, counter):
    """Stage-1 sweep dispatches (int32 and packed kernels) a committed counter profile covers,
    or None."""
    path = os.path.join(PROFILES, fname)
    if not os.path.exists(path):
        return None
    n = 0
    for line in open(path):
        f = [x.strip() for x in line.split("|")]
        if len(f) == 5 and f[0] == counter and ("k_sweep<4, 0," in f[1] or "k_sweep_pk<4, 0," in f[1]):
            n += int(f[2])
    return n or None


TRAFFIC_PROFILES = {1: "k_sweep_stage1_1Mbp_pair", 2: "k_sweep_stage1_10Mbp_pair", 3: "k_sweep_pk_65536_pair_batch"}


def measured_traffic(workload, codes=1, precision=32, launches=None):
    """HBM bytes per stage-1 launch from the committed FETCH_SIZE / WRITE_SIZE profiles of
    this workload, each scaled by what the counter reads for a 256 MiB copy: (fetched plus
    written bytes or None, fetched bytes or None, source files or None).  A profile of a
    sweep with another number of launches (another tile layout) describes other launches
    and is not used."""
    if workload not in TRAFFIC_PROFILES:
        return None, None, None
    name = TRAFFIC_PROFILES[workload]
    if launches is not None and _stage1_dispatches(name + "_fetch_size.txt", "FETCH_SIZE") != launches:
        return None, None, None
    k = "k_sweep%s<4, 0, %s>" % ("_pk" if precision == 16 else "", "true" if codes else "false")
    fetch = _counter_avg(name + "_fetch_size.txt", "FETCH_SIZE", k)
    write = _counter_avg(name + "_write_size.txt", "WRITE_SIZE", k)
    cf = _counter_avg("copy_256MiB_fetch_size.txt", "FETCH_SIZE", "copyBuffer")
    cw = _counter_avg("copy_256MiB_write_size.txt", "WRITE_SIZE", "copyBuffer")
    kib = 268435456 / 1024.0
    fetched = int(fetch * kib / cf * 1024) if None not in (fetch, cf) else None
    if fetched is None or None in (write, cw):
        return None, fetched, ("profiles/%s_fetch_size.txt" % name if fetched is not None else None)
    return fetched + int(write * kib / cw * 1024), fetched, "profiles/%s_{fetch,write}_size.txt" % name


def algorithmic_bytes(cells, n_cols_per_tile, K, Kc):
    """Stage-1 HBM bytes a tile of 256 x W must move: top row {H,F} + column bytes (9W),
    left column {H,E} + row bytes (2304), row maxima read back (1024); bottom row (8W),
    right column (2048), row maxima (1024); plus the checkpoint rows and columns."""
    W = max(1, n_cols_per_tile)
    per_cell = (17.0 * W + 6400.0) / (256.0 * W) + 8.0 / K + 8.0 / Kc
    return cells * per_cell


def roofline(cells, fwd_ms, launches, workload=None, W=None, K=None, Kc=None, precision=32, codes=1):
    """Stage-1 sweep against the integer-VALU peak of the precision it ran in (int32, or
    packed int16 pairs at twice the operation rate)."""
    achieved = cells * OPS_PER_CELL / (fwd_ms / 1000.0) / 1e12
    peak = PEAK_INT16X2_TOPS if precision == 16 else PEAK_INT32_TOPS
    traffic, fetched, source = measured_traffic(workload, codes, precision, launches)
    extra = {}
    if W and K and Kc:
        extra["algorithmic_bytes_per_launch"] = int(algorithmic_bytes(cells, W, K, Kc) / max(launches, 1))
    if traffic is None and fetched is not None:
        extra["fetched_bytes_per_launch"] = fetched  # no WRITE_SIZE profile of this workload
    if source:
        extra["traffic_source"] = source
    return {
        "bound": "valu",
        "kernel": ("k_sweep_pk<4> (forward sweep, packed 16-bit)" if precision == 16
                   else "k_sweep<4, kStage1, %s> (forward sweep with checkpoints, %s scores)"
                   % (("true", "symbol-code") if codes else ("false", "byte-compare"))),
        "achieved": round(achieved, 3),
        "peak": round(peak, 2),
        "unit": "TOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": traffic,
        **extra,
        "launches": launches,
        "avg_launch_us": round(fwd_ms * 1000.0 / max(launches, 1), 2),
        "ops_per_launch": int(cells * OPS_PER_CELL / max(launches, 1)),
    }


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def load_oracle():
    import importlib.util
    spec = importlib.util.spec_from_file_location("pyoracle", os.path.join(ROOT, "oracle", "pyoracle.py"))
    orc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(orc)
    return orc


def host_threads():
    """Host threads for the CPU baselines: the CPUs this process may run on, within the
    thread count the environment grants (OMP_NUM_THREADS: a one-GPU box's CPU share)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else max(1, n)


def pair_record(torch, sw, workload, steps, warmup, K=None):
    """Workloads 0-2: `steps` timed full alignments of one synthetic pair held in device
    memory, after `warmup` untimed ones."""
    (a, b), sc = sw.workload_pair(workload)
    if K is None:
        K = 8192 if workload in (1, 2) else 0  # quoted at K = 8192; 0 = library default
    ctx = sw.Context(0)
    info = ctx.device_info()
    dev = torch.device("cuda:0")
    da = torch.frombuffer(bytearray(a), dtype=torch.uint8).to(dev)
    db = torch.frombuffer(bytearray(b), dtype=torch.uint8).to(dev)
    torch.cuda.synchronize()

    def step():
        return ctx.align_device(da.data_ptr(), len(a), db.data_ptr(), len(b), sc, K=K)

    for k in range(warmup):
        t1 = time.perf_counter()
        step()
        log("workload %d warmup step %d done in %.2f s" % (workload, k, time.perf_counter() - t1))
    torch.cuda.synchronize()
    results = []
    t0 = time.perf_counter()
    for k in range(steps):
        results.append(step())
        log("workload %d step %d done (%.2f s elapsed)" % (workload, k, time.perf_counter() - t0))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.close()
    del da, db
    torch.cuda.empty_cache()
    r0 = results[0]
    for r in results[1:]:
        assert (r["score"], r["end"], r["start"], r["cigar"]) == (r0["score"], r0["end"], r0["start"], r0["cigar"])
    m, n = r0["m"], r0["n"]
    cells = m * n
    fwd_ms = statistics.mean(r["ms_forward"] for r in results)
    clock_ghz = info["clock_khz"] / 1e6 if info["clock_khz"] else PEAK_CLOCK_GHZ
    rec = {
        "value": round(cells / (elapsed / steps) / 1e9, 2),
        "unit": "GCUPS",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(elapsed * 1000.0 / steps, 3),
        "dtype": "int16x2 stage 1, int32 end search and traceback" if r0["precision"] == 16 else "int32",
        "data": "synthetic (seeded std::mt19937_64, seed %d)" % (2012 + workload),
        "config": {
            "workload": WORKLOADS[workload],
            "m": m, "n": n, "cells": cells,
            "checkpoint_rows": r0["checkpoint_rows"], "checkpoint_cols": r0["checkpoint_cols"],
            "score": r0["score"], "end": list(r0["end"]), "start": list(r0["start"]),
            "cigar_ops": r0["n_ops"],
            "stage_ms": {
                "forward": round(fwd_ms, 3),
                "end_search": round(statistics.mean(r["ms_end_search"] for r in results), 3),
                "traceback": round(statistics.mean(r["ms_traceback"] for r in results), 3),
            },
            "forward_gcups": round(cells / (fwd_ms / 1000.0) / 1e9, 2),
            "device": info["name"], "cus": info["cus"], "clock_mhz": round(clock_ghz * 1000),
        },
        "roofline": roofline(cells, fwd_ms, r0["forward_launches"], workload, precision=r0["precision"],
                             codes=r0["score_codes"], W=r0["forward_tile_cols"],
                             K=r0["checkpoint_rows"] or m, Kc=r0["checkpoint_cols"] or n),
    }
    return rec, (a, b)


def device_pairs(torch, pairs):
    """The pairs in one device tensor: (tensor, [(ptr1, len1, ptr2, len2)])."""
    blob = b"".join(a + b for a, b in pairs)
    dev = torch.frombuffer(bytearray(blob + b"\0"), dtype=torch.uint8).to("cuda:0")
    base, off, dp = dev.data_ptr(), 0, []
    for a, b in pairs:
        dp.append((base + off, len(a), base + off + len(a), len(b)))
        off += len(a) + len(b)
    return dev, dp


def plain_gcups(ctx, sc, dp, pairs, budget_s):
    """The plain path: swh_align_pair_device pair after pair, on a prefix of the pairs
    taking about budget_s."""
    n, cells = 0, 0
    t0 = time.perf_counter()
    for (p1, l1, p2, l2), (a, b) in zip(dp, pairs):
        ctx.align_device(p1, l1, p2, l2, sc)
        n += 1
        cells += len(a) * len(b)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return (cells / dt / 1e9 if cells else None), n


def batch_record(args, torch, sw, pairs, sc, steps, warmup, cpu=True):
    """Workload 3: the whole batch is one step (swh_batch_align_device), inputs resident in
    device memory, settings from the autotuner; the plain path (swh_align_pair_device pair
    after pair) is timed on the first pairs of the same batch."""
    ctx = sw.Context(0)
    dev, dp = device_pairs(torch, pairs)
    cells = sum(len(a) * len(b) for a, b in pairs)
    torch.cuda.synchronize()
    kw, tuning = {}, None
    if args.batch_settings:
        W, wps, K, Kc = (int(x) for x in args.batch_settings.split(","))
        kw = dict(tile_cols=W, waves_per_simd=wps, K=K, Kc=Kc)
    elif not args.no_autotune:
        # tile width, waves per SIMD and checkpoint spacing, chosen on an evenly spaced eighth
        # of the batch (untimed, like the reference's own launch set-up)
        t0 = time.perf_counter()
        tuning = ctx.batch_tune(dp, sc, sample=len(dp) // 8)
        tuning["tune_s"] = round(time.perf_counter() - t0, 2)
        kw = dict(tile_cols=tuning["tile_cols"], waves_per_simd=tuning["waves_per_simd"],
                  K=tuning["checkpoint_rows"], Kc=tuning["checkpoint_cols"])
        log("batch autotune %s" % tuning)
    for k in range(warmup):
        ctx.batch_device(dp, sc, **kw)
        log("batch warmup step %d done in %.2f s" % (k, ctx.last_call_s))
    times, res = [], None
    for k in range(steps):
        res, st = ctx.batch_device(dp, sc, **kw)
        assert all(x == 0 for x in st)
        times.append(ctx.last_call_s)
        log("batch step %d done in %.2f s" % (k, ctx.last_call_s))
    sec = statistics.median(times)
    r0 = res[0]
    plain, nplain = (None, 0) if args.no_plain else plain_gcups(ctx, sc, dp, pairs, 10.0)
    log("batch plain path: %s GCUPS over %d pairs" % (plain and round(plain, 2), nplain))
    ctx.close()
    del dev
    torch.cuda.empty_cache()
    gcups = cells / sec / 1e9
    rec = {
        "value": round(gcups, 2), "unit": "GCUPS", "steps": steps, "warmup": warmup,
        "ms_per_step": round(sec * 1000.0, 3),
        "dtype": "int16x2 stage 1, int32 end search and traceback" if r0["precision"] == 16 else "int32",
        "data": "synthetic (numpy PCG64 lengths and targets, seed 2015; queries by swh_synth_mutate)",
        "config": {
            "workload": WORKLOADS[3], "pairs": len(pairs), "cells": cells,
            "stage_ms": {"forward": round(r0["ms_forward"], 3), "end_search": round(r0["ms_end_search"], 3),
                         "traceback": round(r0["ms_traceback"], 3)},
            "forward_gcups": round(cells / (r0["ms_forward"] / 1000.0) / 1e9, 2),
            "plain_per_pair_gcups": plain and round(plain, 2), "plain_sample_pairs": nplain,
            "batch_vs_plain": plain and round(gcups / plain, 2),
            "autotune": tuning and {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tuning.items()},
        },
        "roofline": roofline(cells, r0["ms_forward"], r0["forward_launches"], 3, precision=r0["precision"],
                             codes=r0["score_codes"]),
    }
    if cpu:
        rec["cpu_baseline"] = cpu_baseline_pairs(pairs, sc)
    return rec


def serve_record(sw, jobs, sc, workload_name, data, budget=64 << 30, warmup=1, plain_budget_s=None):
    """A request server (swh_server_*) fed by Poisson arrivals.  Capacity is the throughput
    of the same jobs submitted all at once; each rate then runs the stream at that fraction
    of capacity.  A job's latency is submission to result (inputs come from host memory, as
    a request's would).  The plain path aligns each job alone, one after another
    (swh_align_pair): the latency a job has with the device to itself."""
    import numpy as np
    cells = [len(a) * len(b) for a, b in jobs]
    precisions = set()  # stage-1 precisions the served results report

    def run(rate):
        rng = np.random.default_rng(7)
        gaps = rng.exponential(1.0 / rate, len(jobs)) if rate else np.zeros(len(jobs))
        with sw.Server(sc, budget=budget) as srv:
            t0 = time.perf_counter()
            ids = []
            due = t0
            for (a, b), g in zip(jobs, gaps):
                due += g
                while time.perf_counter() < due:
                    time.sleep(min(0.001, max(0.0, due - time.perf_counter())))
                ids.append(srv.submit(a, b))
            lat = []
            for j in ids:
                st, r = srv.poll(j)
                assert st == 0
                lat.append(r["ms_total"])
                precisions.add(r["precision"])
            stats = srv.stats()
        return dict(rate_jobs_per_s=round(rate, 2) if rate else None, gcups=round(stats["gcups"], 2),
                    p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(float(np.percentile(lat, 99)), 2),
                    max_ms=round(max(lat), 2), jobs=len(lat), peak_in_flight=stats["peak_in_flight"],
                    admitted_mid_flight=stats["admitted_mid_flight"],
                    shared_launch_frac=round(stats["shared_launches"] / max(1, stats["sweep_launches"]), 3))

    for k in range(warmup):
        run(0)
        log("%s: warmup done" % workload_name)
    sat = run(0)
    capacity = len(jobs) * sat["gcups"] * 1e9 / sum(cells)  # jobs per second at saturation
    log("%s: saturated %.1f GCUPS, capacity %.1f jobs/s" % (workload_name, sat["gcups"], capacity))
    rates = {}
    for frac in (0.5, 0.8, 0.95):
        rates["%d%%" % round(frac * 100)] = run(frac * capacity)
        log("%s: rate %.0f%% done: %s" % (workload_name, frac * 100, rates["%d%%" % round(frac * 100)]))
    ctx = sw.Context(0)
    plain, pcells = [], 0
    t0 = time.perf_counter()
    for (a, b), c in zip(jobs, cells):
        t1 = time.perf_counter()
        ctx.align(a, b, sc)
        plain.append((time.perf_counter() - t1) * 1000.0)
        pcells += c
        if plain_budget_s and time.perf_counter() - t0 > plain_budget_s:
            break
    plain_gc = pcells / (time.perf_counter() - t0) / 1e9
    ctx.close()
    log("%s: plain path %.1f GCUPS over %d jobs" % (workload_name, plain_gc, len(plain)))
    top = rates["95%"]
    return {
        "value": top["gcups"], "unit": "GCUPS", "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "dtype": ("int16x2 stage 1 (int32 for launches of few tiles), int32 end search and traceback"
                  if 16 in precisions else "int32"), "data": data,
        "config": {"workload": workload_name, "jobs": len(jobs), "cells": sum(cells), "hbm_pool_bytes": budget,
                   "saturated": sat, "capacity_jobs_per_s": round(capacity, 2), "rates": rates,
                   "plain": {"gcups": round(plain_gc, 2), "jobs": len(plain),
                             "p50_ms": round(float(np.percentile(plain, 50)), 2),
                             "p99_ms": round(float(np.percentile(plain, 99)), 2)}},
        "served_vs_plain_gcups": round(top["gcups"] / plain_gc, 2),
        "saturated_vs_plain_gcups": round(sat["gcups"] / plain_gc, 2),
    }


def cpu_baseline_pairs(pairs, sc, budget_s=10.0):
    """The oracle's forward pass (one thread per pair, pairs over the host's cores) on a
    prefix of the batch sized to about budget_s."""
    from concurrent.futures import ThreadPoolExecutor
    orc = load_oracle()
    threads = host_threads()
    osc = orc.scoring(sc.match, sc.mismatch, sc.gap_open, sc.gap_extend)

    def one(p):
        a, b = p if len(p[0]) >= len(p[1]) else (p[1], p[0])
        orc.forward(a, b, osc, end_col=True)
        return len(a) * len(b)

    cells, t0, k = 0, time.perf_counter(), 0
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < budget_s and k < len(pairs):
            chunk = pairs[k:k + 4 * threads]
            cells += sum(ex.map(one, chunk))
            k += len(chunk)
    dt = time.perf_counter() - t0
    return dict(value=round(cells / dt / 1e9, 3), unit="GCUPS", cores=threads, kind="port",
                sample="oracle forward pass (score + end cell) over the first %d pairs of the batch, one pair "
                       "per host thread, %d threads, %.1f s" % (k, threads, dt))


def cpu_baseline(a, b, budget_s=15.0):
    """The oracle's multithreaded forward pass over the first rows of the pair."""
    orc = load_oracle()
    A, B = (a, b) if len(a) >= len(b) else (b, a)
    threads = host_threads()
    sc = orc.scoring(1, -3, -5, -2)
    # calibrate on growing bands until one takes a second, then size the sample to budget_s
    rows = 256
    while True:
        t0 = time.perf_counter()
        orc.forward(A[:rows], B, sc, threads=threads, end_col=False)
        dt = time.perf_counter() - t0
        if dt >= 1.0 or rows >= len(A):
            break
        rows = min(len(A), rows * 4)
    rate = rows * len(B) / max(dt, 1e-6)
    rows = int(min(len(A), max(1024, rate * budget_s / len(B))))
    t0 = time.perf_counter()
    orc.forward(A[:rows], B, sc, threads=threads, end_col=False)
    dt = time.perf_counter() - t0
    return dict(value=round(rows * len(B) / dt / 1e9, 3), unit="GCUPS", cores=threads, kind="port",
                sample="oracle forward pass (score and per-row maxima, no traceback) over the first %d rows x %d "
                       "columns of the same pair, %d host threads, %.1f s" % (rows, len(B), threads, dt))


def cpu_baseline_config0(sw):
    """The oracle on the 10 kbp x 10 kbp pair with full traceback (checkpoints every 1024
    rows and columns, as the library's default for this pair): one thread and all the
    host threads, one untimed run then the median of 5."""
    orc = load_oracle()
    (a, b), sc = sw.workload_pair(0)
    A, B = (a, b) if len(a) >= len(b) else (b, a)
    osc = orc.scoring(1, -3, -5, -2)
    out = {}
    for label, th in (("single_thread", 1), ("all_threads", host_threads())):
        orc.align_linear(A, B, osc, 1024, 1024, threads=th)
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            r = orc.align_linear(A, B, osc, 1024, 1024, threads=th)
            ts.append(time.perf_counter() - t0)
        sec = statistics.median(ts)
        out[label] = dict(value=round(len(A) * len(B) / sec / 1e9, 3), unit="GCUPS", cores=th,
                          ms=round(sec * 1000.0, 2), score=r["score"])
    out["sample"] = ("oracle linear-space alignment with full traceback of the 10 kbp x 10 kbp pair "
                     "(1e8 cells), median of 5 after one untimed run")
    out["kind"] = "port"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=["all"] + [str(k) for k in sorted(WORKLOADS)],
                    help="all (default): the 1 Mbp pair as value plus the chromosome, batched and served "
                         "records; N: workload N alone")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="workload 3: default batch settings")
    ap.add_argument("--no-plain", action="store_true", help="workload 3: skip the pair-after-pair comparison")
    ap.add_argument("--batch-settings", default="", metavar="W,WPS,K,KC",
                    help="workload 3: tile columns, sweep waves per SIMD and checkpoint rows / columns "
                         "instead of the autotuner's choice (the profiles' form)")
    ap.add_argument("--skip", default="", help="with --workload all: comma-separated records to leave out "
                                                "(chromosome, batched, served, served_mixed)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device once (calibrates the FETCH_SIZE / WRITE_SIZE "
                         "counters against a transfer of known size) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one MI355X; the pair alignment is not sharded across "
              "GPUs (see DESIGN.md for how it would shard). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        src = torch.ones(args.copy_probe, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return
    import swalign as sw
    common = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None}
    if args.workload in ("3", "4"):
        if args.workload == "3":
            pairs, sc = sw.workload_batch()
            rec = batch_record(args, torch, sw, pairs, sc, args.steps, args.warmup, cpu=not args.no_cpu_baseline)
        else:
            jobs, sc = sw.workload_serve(max(200, args.steps * 40))
            rec = serve_record(sw, jobs, sc, WORKLOADS[4], "synthetic (numpy PCG64 lengths, targets and "
                               "arrivals, seed 2016 / 7; queries by swh_synth_mutate)", warmup=args.warmup)
            rec.update(steps=1, warmup=args.warmup, ms_per_step=None)
        print(json.dumps({**common, **rec}))
        return
    if args.workload != "all":
        w = int(args.workload)
        rec, (a, b) = pair_record(torch, sw, w, args.steps, args.warmup)
        if not args.no_cpu_baseline:
            rec["cpu_baseline"] = cpu_baseline(a, b) if w else cpu_baseline_config0(sw)
        print(json.dumps({**common, **rec}))
        return

    skip = set(x for x in args.skip.split(",") if x)
    rec, (a, b) = pair_record(torch, sw, 1, args.steps, args.warmup)
    line = {**common, **rec}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a, b)
        line["cpu_baseline"]["config0_full_traceback"] = cpu_baseline_config0(sw)
        log("cpu baselines done")
    if "chromosome" not in skip:
        line["chromosome"], _ = pair_record(torch, sw, 2, 1, 1)
    if not ({"batched", "served_mixed"} <= skip):
        pairs, sc = sw.workload_batch()
        if "batched" not in skip:
            line["batched"] = batch_record(args, torch, sw, pairs, sc, 3, 1, cpu=not args.no_cpu_baseline)
        if "served_mixed" not in skip:
            line["served_mixed"] = serve_record(
                sw, pairs[:4096], sc, "the batched workload's pair mix (first 4096 pairs) as a served stream",
                "synthetic (the batched workload's first 4096 pairs; arrivals numpy PCG64 seed 7)",
                plain_budget_s=10.0)
    if "served" not in skip:
        jobs, sc = sw.workload_serve(200)
        line["served"] = serve_record(sw, jobs, sc, WORKLOADS[4], "synthetic (numpy PCG64 lengths, targets and "
                                      "arrivals, seed 2016 / 7; queries by swh_synth_mutate)")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
