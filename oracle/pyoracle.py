"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module, and only to check or time against the HIP library.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Scoring(C.Structure):
    _fields_ = [("match", C.c_int32), ("mismatch", C.c_int32),
                ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class Alignment(C.Structure):
    _fields_ = [("score", C.c_int32), ("_pad", C.c_int32),
                ("end_i", C.c_int64), ("end_j", C.c_int64),
                ("start_i", C.c_int64), ("start_j", C.c_int64),
                ("cigar_len", C.c_int64), ("n_ops", C.c_int64)]


class XPoint(C.Structure):
    _fields_ = [("row", C.c_int64), ("col", C.c_int64),
                ("score", C.c_int32), ("layer", C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "lib", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run make -C oracle")
        _LIB = C.CDLL(path)
    return _LIB


def _buf(seq):
    if isinstance(seq, str):
        seq = seq.encode()
    arr = np.frombuffer(bytes(seq), dtype=np.uint8)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_uint8)), len(arr)


def scoring(m, x, o, e):
    return Scoring(m, x, o, e)


def forward(a, b, sc, rowmax=False, threads=1, end_col=True):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    s = C.c_int32()
    ei = C.c_int64()
    ej = C.c_int64(-1)
    rm = np.zeros(max(m, 1), dtype=np.int32) if rowmax else None
    prm = rm.ctypes.data_as(C.POINTER(C.c_int32)) if rowmax else None
    if threads > 1:
        rc = lib().orc_forward_mt(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int(threads),
                                  C.byref(s), C.byref(ei), C.byref(ej) if end_col else None, prm)
    else:
        rc = lib().orc_forward(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.byref(s),
                               C.byref(ei), C.byref(ej), prm)
    assert rc == 0
    out = (s.value, ei.value, ej.value)
    return (out, rm[:m]) if rowmax else out


def full_matrices(a, b, sc):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    H = np.zeros((m, n), np.int32)
    E = np.zeros((m, n), np.int32)
    F = np.zeros((m, n), np.int32)
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
    lib().orc_forward_full(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), p(H), p(E), p(F))
    return H, E, F


def textbook_H(a, b, sc):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    H = np.zeros((m, n), np.int32)
    lib().orc_forward_textbook_full(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc),
                                    H.ctypes.data_as(C.POINTER(C.c_int32)))
    return H


def bruteforce(a, b, sc):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    s = C.c_int32()
    assert lib().orc_bruteforce_score(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.byref(s)) == 0
    return s.value


def _result(al, cig):
    return dict(score=al.score, end=(al.end_i, al.end_j), start=(al.start_i, al.start_j),
                cigar=cig.value.decode() if al.score > 0 else "", n_ops=al.n_ops)


def align_full(a, b, sc):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    al = Alignment()
    cap = 4 * (m + n) + 64
    cig = C.create_string_buffer(cap)
    lib().orc_align_full(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.byref(al), cig,
                         C.c_int64(cap))
    return _result(al, cig)


def align_linear(a, b, sc, K, Kc, threads=1):
    """Linear-space alignment (checkpoint method); threads > 1 runs its stage 1 on that many
    host threads (orc_align_linear_mt), with the same result."""
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    al = Alignment()
    cap = 4 * (m + n) + 64
    cig = C.create_string_buffer(cap)
    xcap = m // max(K, 1) + 8
    xp = (XPoint * xcap)()
    nxp = C.c_int64()
    if threads > 1:
        rc = lib().orc_align_linear_mt(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int64(K),
                                       C.c_int64(Kc), C.c_int(threads), C.byref(al), cig, C.c_int64(cap), xp,
                                       C.c_int64(xcap), C.byref(nxp))
    else:
        rc = lib().orc_align_linear(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int64(K),
                                    C.c_int64(Kc), C.byref(al), cig, C.c_int64(cap), xp,
                                    C.c_int64(xcap), C.byref(nxp))
    assert rc == 0
    r = _result(al, cig)
    r["xpoints"] = [(xp[k].row, xp[k].col, xp[k].score, xp[k].layer)
                    for k in range(min(nxp.value, xcap))]
    return r


def checkpoints(a, b, sc, K, Kc):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    nsr, nsc = m // K, n // Kc
    srow = np.zeros((max(nsr, 1), n, 2), np.int32)
    scol = np.zeros((max(nsc, 1), m, 2), np.int32)
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
    lib().orc_checkpoints(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int64(K),
                          C.c_int64(Kc), p(srow), p(scol))
    return srow[:nsr], scol[:nsc]


def crosspoint_algebra(a, b, sc, y, end, cross_c, cross_type):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    best = C.c_int32()
    atc = C.c_int32()
    rc = lib().orc_crosspoint_algebra(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int64(y),
                                      C.c_int64(end[0]), C.c_int64(end[1]), C.c_int64(cross_c),
                                      C.c_int32(cross_type), C.byref(best), C.byref(atc))
    assert rc == 0
    return best.value, atc.value


def rescore(a, b, sc, start, cigar):
    A, pa, m = _buf(a)
    B, pb, n = _buf(b)
    s = C.c_int32()
    ei = C.c_int64()
    ej = C.c_int64()
    rc = lib().orc_rescore_cigar(pa, C.c_int64(m), pb, C.c_int64(n), C.byref(sc), C.c_int64(start[0]),
                                 C.c_int64(start[1]), cigar.encode(), C.byref(s), C.byref(ei),
                                 C.byref(ej))
    if rc != 0:
        raise ValueError("cigar does not fit the sequences (code %d)" % rc)
    return s.value, (ei.value, ej.value)


def fasta_load(path):
    nm = C.c_char_p()
    sq = C.POINTER(C.c_uint8)()
    ln = C.c_int64()
    L = lib()
    L.orc_fasta_load.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.POINTER(C.POINTER(C.c_uint8)),
                                 C.POINTER(C.c_int64)]
    rc = L.orc_fasta_load(path.encode(), C.byref(nm), C.byref(sq), C.byref(ln))
    if rc != 0:
        raise IOError("fasta load failed (%d): %s" % (rc, path))
    name = nm.value.decode("latin-1")
    seq = C.string_at(sq, ln.value)
    return name, seq
