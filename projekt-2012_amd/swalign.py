"""ctypes binding of libswalign_hip.so (include/swalign_hip.h).

The library is the HIP (gfx950) implementation of the SWalign hot path; this module
only marshals arguments.  It raises if the built library is missing: there is no
CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libswalign_hip.so")
_LIB = None

SWH_OK = 0
SWH_ERR_INVALID_ARG = 1
SWH_ERR_SCORING = 2
SWH_ERR_NO_DEVICE = 3
SWH_ERR_HIP = 4
SWH_ERR_OUT_OF_MEMORY = 5
SWH_ERR_BUDGET = 6
SWH_ERR_INTERNAL = 7
SWH_ERR_NOT_FOUND = 8
SWH_ERR_SCRATCH = 10
NOALLOC_SCRATCH_BYTES = 64 * 1024 * 1024


class Scoring(C.Structure):
    _fields_ = [("match", C.c_int32), ("mismatch", C.c_int32),
                ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class Opts(C.Structure):
    _fields_ = [("checkpoint_rows", C.c_int64), ("checkpoint_cols", C.c_int64),
                ("hbm_budget_bytes", C.c_int64), ("score_only", C.c_int32),
                ("precision", C.c_int32), ("want_xpoints", C.c_int32),
                ("tile_cols", C.c_int32), ("tb_lookahead", C.c_int32),
                ("waves_per_simd", C.c_int32), ("autotune", C.c_int32), ("host_checkpoints", C.c_int32)]


class XPoint(C.Structure):
    _fields_ = [("row", C.c_int64), ("col", C.c_int64), ("score", C.c_int32), ("layer", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("score", C.c_int32), ("swapped", C.c_int32),
                ("end_row", C.c_int64), ("end_col", C.c_int64),
                ("start_row", C.c_int64), ("start_col", C.c_int64),
                ("m", C.c_int64), ("n", C.c_int64),
                ("cigar", C.c_char_p), ("cigar_len", C.c_int64), ("n_ops", C.c_int64),
                ("xpoints", C.POINTER(XPoint)), ("n_xpoints", C.c_int64),
                ("checkpoint_rows", C.c_int64), ("checkpoint_cols", C.c_int64),
                ("ms_forward", C.c_double), ("ms_end_search", C.c_double),
                ("ms_traceback", C.c_double), ("ms_total", C.c_double),
                ("precision_used", C.c_int32), ("forward_launches", C.c_int32),
                ("score_codes", C.c_int32), ("forward_tile_cols", C.c_int32)]


class Pair(C.Structure):
    _fields_ = [("seq1", C.c_void_p), ("len1", C.c_int64), ("seq2", C.c_void_p), ("len2", C.c_int64)]


# (name, restype, argtypes) of every exported function declared in include/swalign_hip.h
EXPORTS = [
    ("swh_strerror", C.c_char_p, [C.c_int]),
    ("swh_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("swh_ctx_destroy", C.c_int, [C.c_void_p]),
    ("swh_device_info", C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("swh_align_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                 C.POINTER(Scoring), C.POINTER(Opts), C.POINTER(Result)]),
    ("swh_align_pair_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                        C.POINTER(Scoring), C.POINTER(Opts), C.POINTER(Result)]),
    ("swh_result_free", None, [C.POINTER(Result)]),
    ("swh_align_pair_device_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                                C.POINTER(Scoring), C.POINTER(Opts), C.POINTER(Result),
                                                C.c_char_p, C.c_int64]),
    ("swh_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                              C.POINTER(Scoring), C.c_int64, C.c_int64, C.POINTER(C.c_int32),
                              C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p, C.c_void_p,
                              C.c_void_p]),
    ("swh_batch_align", C.c_int, [C.c_void_p, C.POINTER(Pair), C.c_int64, C.POINTER(Scoring),
                                  C.POINTER(Opts), C.POINTER(Result), C.POINTER(C.c_int32)]),
    ("swh_batch_align_device", C.c_int, [C.c_void_p, C.POINTER(Pair), C.c_int64, C.POINTER(Scoring),
                                         C.POINTER(Opts), C.POINTER(Result), C.POINTER(C.c_int32)]),
    ("swh_batch_tune", C.c_int, [C.c_void_p, C.POINTER(Pair), C.c_int64, C.c_int32, C.POINTER(Scoring),
                                 C.POINTER(Opts), C.c_int64, C.c_void_p]),
    ("swh_server_create", C.c_int, [C.c_int, C.POINTER(Scoring), C.POINTER(Opts), C.c_int64, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    ("swh_server_submit", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                    C.POINTER(C.c_int64)]),
    ("swh_server_poll", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(Result)]),
    ("swh_server_stats", C.c_int, [C.c_void_p, C.c_void_p]),
    ("swh_server_destroy", C.c_int, [C.c_void_p]),
    ("swh_synth_pair", C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("swh_synth_mutate", C.c_int, [C.c_uint64, C.c_void_p, C.c_int64, C.c_double, C.c_double,
                                   C.c_double, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    ("swh_free", None, [C.c_void_p]),
]


class SWError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().swh_strerror(status).decode()
        super().__init__("%s (status %d)%s" % (msg, status, (": " + what) if what else ""))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libswalign_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _bytes(seq):
    if isinstance(seq, str):
        seq = seq.encode()
    return bytes(seq)


def scoring(match, mismatch, gap_open, gap_extend):
    return Scoring(match, mismatch, gap_open, gap_extend)


def _result_dict(r):
    d = dict(score=r.score, swapped=bool(r.swapped), end=(r.end_row, r.end_col),
             start=(r.start_row, r.start_col), m=r.m, n=r.n,
             cigar=r.cigar.decode() if r.cigar else None, n_ops=r.n_ops,
             checkpoint_rows=r.checkpoint_rows, checkpoint_cols=r.checkpoint_cols,
             ms_forward=r.ms_forward, ms_end_search=r.ms_end_search, ms_traceback=r.ms_traceback,
             ms_total=r.ms_total, precision=r.precision_used,
             forward_launches=r.forward_launches, score_codes=r.score_codes,
             forward_tile_cols=r.forward_tile_cols)
    if r.xpoints and r.n_xpoints:
        d["xpoints"] = [(r.xpoints[k].row, r.xpoints[k].col, r.xpoints[k].score, r.xpoints[k].layer)
                        for k in range(r.n_xpoints)]
    else:
        d["xpoints"] = []
    return d


def make_opts(K=0, Kc=0, score_only=False, xpoints=False, budget=0, precision=0, tile_cols=0, lookahead=0,
              waves_per_simd=0, autotune=False, host_checkpoints=False):
    o = Opts()
    o.host_checkpoints = 1 if host_checkpoints else 0
    o.tile_cols = tile_cols
    o.waves_per_simd = waves_per_simd
    o.autotune = 1 if autotune else 0
    o.tb_lookahead = lookahead
    o.checkpoint_rows = K
    o.checkpoint_cols = Kc
    o.hbm_budget_bytes = budget
    o.score_only = 1 if score_only else 0
    o.want_xpoints = 1 if xpoints else 0
    o.precision = precision
    return o


class Context:
    """One HIP device + stream + cached device buffers (swh_ctx)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().swh_ctx_create(device, C.byref(self._h))
        if rc != SWH_OK:
            raise SWError(rc, "swh_ctx_create")

    def close(self):
        if self._h:
            lib().swh_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int32()
        clk = C.c_int32()
        hbm = C.c_int64()
        lib().swh_device_info(self._h, name, 256, C.byref(cus), C.byref(clk), C.byref(hbm))
        return dict(name=name.value.decode(), cus=cus.value, clock_khz=clk.value, hbm_bytes=hbm.value)

    def align(self, seq1, seq2, sc, **kw):
        s1, s2 = _bytes(seq1), _bytes(seq2)
        r = Result()
        o = make_opts(**kw)
        rc = lib().swh_align_pair(self._h, s1, len(s1), s2, len(s2), C.byref(sc), C.byref(o), C.byref(r))
        try:
            if rc != SWH_OK:
                raise SWError(rc, "swh_align_pair")
            return _result_dict(r)
        finally:
            lib().swh_result_free(C.byref(r))

    def align_device(self, ptr1, len1, ptr2, len2, sc, **kw):
        """Sequences already in device memory (e.g. torch uint8 tensors' data_ptr())."""
        r = Result()
        o = make_opts(**kw)
        rc = lib().swh_align_pair_device(self._h, C.c_void_p(ptr1), len1, C.c_void_p(ptr2), len2,
                                         C.byref(sc), C.byref(o), C.byref(r))
        try:
            if rc != SWH_OK:
                raise SWError(rc, "swh_align_pair_device")
            return _result_dict(r)
        finally:
            lib().swh_result_free(C.byref(r))

    def align_device_noalloc(self, ptr1, len1, ptr2, len2, sc, **kw):
        """swh_align_pair_device_noalloc: the library allocates no device memory; the CIGAR
        comes back through a host buffer this binding passes in."""
        r = Result()
        o = make_opts(**kw)
        cap = 3 * (len1 + len2) + 64
        buf = C.create_string_buffer(cap)
        rc = lib().swh_align_pair_device_noalloc(self._h, C.c_void_p(ptr1), len1, C.c_void_p(ptr2), len2,
                                                 C.byref(sc), C.byref(o), C.byref(r), buf, cap)
        try:
            if rc != SWH_OK:
                raise SWError(rc, "swh_align_pair_device_noalloc")
            d = _result_dict(r)
            if not kw.get("score_only"):
                d["cigar"] = buf.value[: r.cigar_len].decode() if r.score > 0 else None
            return d
        finally:
            lib().swh_result_free(C.byref(r))

    def forward(self, A, B, sc, K, Kc, checkpoints=True):
        """Stage 1 on oriented inputs (A rows, B columns); returns score, end, rowmax and
        the checkpoint rows/columns as numpy arrays."""
        import numpy as np
        A, B = _bytes(A), _bytes(B)
        m, n = len(A), len(B)
        score = C.c_int32()
        er = C.c_int64()
        ec = C.c_int64()
        rowmax = np.zeros(max(m, 1), np.int32)
        nsr, nsc = m // K, n // Kc
        srow = np.zeros((max(nsr, 1), n, 2), np.int32) if checkpoints else None
        scol = np.zeros((max(nsc, 1), m, 2), np.int32) if checkpoints else None
        p = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        rc = lib().swh_forward(self._h, A, m, B, n, C.byref(sc), K, Kc, C.byref(score), C.byref(er),
                               C.byref(ec), p(rowmax), p(srow), p(scol))
        if rc != SWH_OK:
            raise SWError(rc, "swh_forward")
        out = dict(score=score.value, end=(er.value, ec.value), rowmax=rowmax[:m])
        if checkpoints:
            out["srow"] = srow[:nsr]
            out["scol"] = scol[:nsc]
        return out

    def batch(self, pairs, sc, **kw):
        """pairs: list of (seq1, seq2).  Returns (results, statuses)."""
        keep = []
        arr = (Pair * len(pairs))()
        for k, (a, b) in enumerate(pairs):
            a, b = _bytes(a), _bytes(b)
            ba = C.create_string_buffer(a, len(a))
            bb = C.create_string_buffer(b, len(b))
            keep += [ba, bb]
            arr[k] = Pair(C.cast(ba, C.c_void_p), len(a), C.cast(bb, C.c_void_p), len(b))
        return self._batch_call("swh_batch_align", arr, len(pairs), sc, kw)

    def batch_device(self, dev_pairs, sc, **kw):
        """dev_pairs: list of (ptr1, len1, ptr2, len2) with device pointers (e.g. slices of
        one torch uint8 tensor).  Returns (results, statuses)."""
        arr = (Pair * len(dev_pairs))()
        for k, (p1, l1, p2, l2) in enumerate(dev_pairs):
            arr[k] = Pair(C.c_void_p(p1), l1, C.c_void_p(p2), l2)
        return self._batch_call("swh_batch_align_device", arr, len(dev_pairs), sc, kw)

    def batch_tune(self, dev_pairs, sc, sample=2048, **kw):
        """swh_batch_tune over device-resident pairs (ptr1, len1, ptr2, len2): the fastest
        tile width, waves per SIMD and checkpoint spacing on a sample of the batch."""
        arr = (Pair * len(dev_pairs))()
        for k, (p1, l1, p2, l2) in enumerate(dev_pairs):
            arr[k] = Pair(C.c_void_p(p1), l1, C.c_void_p(p2), l2)
        t = BatchTuning()
        o = make_opts(**kw)
        rc = lib().swh_batch_tune(self._h, arr, len(dev_pairs), 1, C.byref(sc), C.byref(o), sample, C.byref(t))
        if rc != SWH_OK:
            raise SWError(rc, "swh_batch_tune")
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_ if not k.startswith("_")}

    def _batch_call(self, fn, arr, npairs, sc, kw):
        import time
        res = (Result * npairs)()
        st = (C.c_int32 * npairs)()
        o = make_opts(**kw)
        t0 = time.perf_counter()
        rc = getattr(lib(), fn)(self._h, arr, npairs, C.byref(sc), C.byref(o), res, st)
        self.last_call_s = time.perf_counter() - t0  # the library call alone (results on the host)
        out = []
        for k in range(npairs):
            out.append(_result_dict(res[k]) if rc == SWH_OK and st[k] == SWH_OK else None)
            lib().swh_result_free(C.byref(res[k]))
        if rc != SWH_OK:
            raise SWError(rc, fn)
        return out, list(st)


def synth_pair(seed, len1, len2, hom_len, p_sub, p_indel, gc=-1.0, repeat_frac=0.0):
    a = C.c_void_p()
    b = C.c_void_p()
    rc = lib().swh_synth_pair(seed, len1, len2, hom_len, p_sub, p_indel, gc, repeat_frac, C.byref(a),
                              C.byref(b))
    if rc != SWH_OK:
        raise SWError(rc, "swh_synth_pair")
    try:
        return C.string_at(a, len1), C.string_at(b, len2)
    finally:
        lib().swh_free(a)
        lib().swh_free(b)


def synth_mutate(seed, src, p_sub, p_indel, flank_frac):
    src = _bytes(src)
    out = C.c_void_p()
    ln = C.c_int64()
    rc = lib().swh_synth_mutate(seed, src, len(src), p_sub, p_indel, flank_frac, C.byref(out), C.byref(ln))
    if rc != SWH_OK:
        raise SWError(rc, "swh_synth_mutate")
    try:
        return C.string_at(out, ln.value)
    finally:
        lib().swh_free(out)


class BatchTuning(C.Structure):
    _fields_ = [("tile_cols", C.c_int32), ("waves_per_simd", C.c_int32), ("checkpoint_rows", C.c_int64),
                ("ms_best", C.c_double), ("ms_default", C.c_double), ("candidates", C.c_int32), ("_pad", C.c_int32),
                ("checkpoint_cols", C.c_int64)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("in_flight", C.c_int64), ("queued", C.c_int64), ("cells_completed", C.c_int64),
                ("gcups", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("hbm_pool_bytes", C.c_int64), ("hbm_in_use_bytes", C.c_int64), ("loop_status", C.c_int32),
                ("_pad", C.c_int32), ("peak_in_flight", C.c_int64), ("admitted_mid_flight", C.c_int64),
                ("sweep_launches", C.c_int64), ("shared_launches", C.c_int64)]


SWH_PENDING = 9


class Server:
    """Continuous-batching request server (swh_server_*)."""

    def __init__(self, sc, device=0, budget=0, max_in_flight=0, **kw):
        self._h = C.c_void_p()
        self._sc = sc
        o = make_opts(**kw)
        rc = lib().swh_server_create(device, C.byref(sc), C.byref(o), budget, max_in_flight, C.byref(self._h))
        if rc != SWH_OK:
            raise SWError(rc, "swh_server_create")

    def submit(self, seq1, seq2):
        s1, s2 = _bytes(seq1), _bytes(seq2)
        jid = C.c_int64()
        rc = lib().swh_server_submit(self._h, s1, len(s1), s2, len(s2), C.byref(jid))
        if rc != SWH_OK:
            raise SWError(rc, "swh_server_submit")
        return jid.value

    def poll(self, job_id, wait=True):
        """(status, result dict or None); status SWH_PENDING when not finished and wait=False."""
        r = Result()
        rc = lib().swh_server_poll(self._h, job_id, 1 if wait else 0, C.byref(r))
        if rc == SWH_PENDING:
            return rc, None
        try:
            return rc, (_result_dict(r) if rc == SWH_OK else None)
        finally:
            lib().swh_result_free(C.byref(r))

    def stats(self):
        st = ServeStats()
        lib().swh_server_stats(self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServeStats._fields_ if not k.startswith("_")}

    def close(self):
        if self._h:
            lib().swh_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceBuffer:
    """Device memory holding a copy of `data`, allocated with the HIP runtime the library
    itself uses (for callers and tests that have no other device allocator at hand)."""

    def __init__(self, data):
        lib()  # the library's HIP runtime is loaded first, so the name below resolves to it
        self._hip = C.CDLL("libamdhip64.so.7")
        self._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self._hip.hipFree.argtypes = [C.c_void_p]
        data = _bytes(data)
        self.size = len(data)
        self.ptr = C.c_void_p()
        if self._hip.hipMalloc(C.byref(self.ptr), max(1, self.size)) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.size)
        if self.size and self._hip.hipMemcpy(self.ptr, data, self.size, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to the device failed")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# Synthetic workloads (seed = 2012 + workload index).
def workload_serve(njobs, seed=2016):
    """The served stream's jobs: each side's length drawn 70% U[1, 10] kbp, 25% U[10, 100]
    kbp, 5% U[100 kbp, 1 Mbp]; the query is the target mutated as in workload_batch.
    Returns ([(target, query)], scoring)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"ACGT", np.uint8)
    jobs = []
    for k in range(njobs):
        u = rng.random()
        lo, hi = (1_000, 10_000) if u < 0.70 else ((10_000, 100_000) if u < 0.95 else (100_000, 1_000_000))
        L = int(rng.integers(lo, hi + 1))
        t = alpha[rng.integers(0, 4, L)].tobytes()
        jobs.append((t, synth_mutate(seed * 1_000_003 + k, t, 0.10, 0.03, 0.20)))
    return jobs, scoring(1, -3, -5, -2)


def workload_batch(npairs=65536, seed=2015, lmin=100, lmax=20000):
    """The batched workload: target lengths log-uniform in [lmin, lmax] (numpy PCG64 from
    `seed`), uniform ACGT targets; each query is its target mutated at 10% substitutions
    and 3% indels with flanks trimmed or extended by up to 20% (swh_synth_mutate, seeded
    per pair).  Returns ([(target, query)], scoring)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(np.log(lmin), np.log(lmax), npairs)).astype(np.int64)
    alpha = np.frombuffer(b"ACGT", np.uint8)
    pairs = []
    for k, L in enumerate(lens):
        t = alpha[rng.integers(0, 4, int(L))].tobytes()
        pairs.append((t, synth_mutate(seed * 1_000_003 + k, t, 0.10, 0.03, 0.20)))
    return pairs, scoring(1, -3, -5, -2)


def workload_pair(index):
    """The single-pair workloads: 0 = 10 kbp x 10 kbp (3 kbp homology, 10%/2%),
    1 = 1 Mbp x 1 Mbp (200 kbp homology, 8%/1.5%), 2 = 10 Mbp x 10 Mbp, GC 40%, tandem
    repeats over 1%, 2 Mbp homology at 5%/1%."""
    seed = 2012 + index
    if index == 0:
        return synth_pair(seed, 10_000, 10_000, 3_000, 0.10, 0.02), scoring(1, -3, -5, -2)
    if index == 1:
        return synth_pair(seed, 1_000_000, 1_000_000, 200_000, 0.08, 0.015), scoring(1, -3, -5, -2)
    if index == 2:
        return (synth_pair(seed, 10_000_000, 10_000_000, 2_000_000, 0.05, 0.01, gc=0.40, repeat_frac=0.01),
                scoring(2, -3, -5, -1))
    raise ValueError(index)
