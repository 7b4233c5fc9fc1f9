"""ctypes front end of the CPU oracle (test infrastructure only).

Wraps ``liboracle.so`` (the C restatement of the reference's CPU path, see
``swift_oracle.h``) and, where it was built, ``_ref/libswiftref.so`` (the
reference's own CPU functions compiled from its sources).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import this
module; the product under swift_amd/ never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libswiftref.so")

TIE_FIRST = 0
TIE_GLIBC = 1

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference bridge when the reference
    sources are present)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"], stdout=subprocess.DEVNULL)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class _Rng(C.Structure):
    _fields_ = [("r", C.c_uint32 * 34), ("pos", C.c_int)]


class _Index(C.Structure):
    _fields_ = [("k", C.c_int), ("thr", C.c_int), ("nhash", C.c_int64),
                ("start", C.POINTER(C.c_int32)), ("count", C.POINTER(C.c_int32)),
                ("actual", C.POINTER(C.c_int32)), ("total", C.c_int64),
                ("refIdx", C.POINTER(C.c_int8)), ("refPos", C.POINTER(C.c_int32))]


class _SwRes(C.Structure):
    _fields_ = [("score", C.c_int), ("start", C.c_int), ("end", C.c_int),
                ("len", C.c_int), ("maxRow", C.c_int), ("maxCol", C.c_int)]


class _Rec(C.Structure):
    _fields_ = [("read", C.c_int32), ("strand", C.c_int32), ("refIdx", C.c_int32),
                ("shift", C.c_int32), ("refPos", C.c_int32), ("nBiggest", C.c_int32),
                ("refDistance", C.c_int64), ("score", C.c_int32), ("start", C.c_int32),
                ("end", C.c_int32), ("len", C.c_int32), ("keep", C.c_int32),
                ("nTied", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        L.orc_hash.restype = C.c_int
        L.orc_rng_next.restype = C.c_int
        L.orc_index_build.restype = C.c_int
        L.orc_index_build.argtypes = [C.POINTER(_Index), C.c_int, C.POINTER(C.c_char_p),
                                      C.POINTER(C.c_int64), C.c_int, C.c_int]
        L.orc_map_query.restype = C.c_int
        L.orc_best_hits.restype = C.c_int
        L.orc_region.restype = C.c_int
        L.orc_region.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                 C.c_char_p, C.POINTER(C.c_int64)]
        L.orc_pipeline.restype = C.c_int
    return _lib


def ref():
    """The reference's own CPU functions, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
        _ref.ref_rand.restype = C.c_int
        _ref.ref_hash.restype = C.c_int
        _ref.ref_best_hits.restype = C.c_int
        _ref.ref_region.restype = C.c_int
        _ref.ref_format_record.restype = C.c_int
        _ref.ref_format_cigar.restype = C.c_int
        _ref.ref_format_sam.restype = C.c_int
    return _ref


# ---------------------------------------------------------------- rng / codec
class Rng:
    """glibc rand() replay."""

    def __init__(self, seed: int):
        self.s = _Rng()
        lib().orc_rng_seed(C.byref(self.s), C.c_uint(seed))

    def next(self) -> int:
        return lib().orc_rng_next(C.byref(self.s))


def get_hash(s: bytes, k: int) -> int:
    return lib().orc_hash(s, k)


def revcomp(seq: bytes) -> bytes:
    out = C.create_string_buffer(len(seq) + 1)
    lib().orc_revcomp(seq, len(seq), out)
    return out.raw[:len(seq)]


# ---------------------------------------------------------------- SW
def sw(ref_seq: bytes, qry: bytes, match=2, mismatch=-1, gap_open=-10, gap_ext=-1,
       ref_len: int | None = None, want_dirs: bool = False) -> dict:
    """smithWaterman_cpu restated; ref_len defaults to len(ref_seq) (the
    reference buffer may be NUL-padded beyond the sequence)."""
    rl = len(ref_seq) if ref_len is None else ref_len
    rbuf = C.create_string_buffer(ref_seq, max(rl, len(ref_seq)) + 1)
    n = rl + len(qry) + 2
    rc = C.create_string_buffer(n)
    qc = C.create_string_buffer(n)
    res = _SwRes()
    dirs = None
    dptr = None
    if want_dirs:
        dirs = np.zeros((len(qry) + 1, rl + 1), np.uint8)
        dptr = _p(dirs, C.c_uint8)
    lib().orc_sw(rbuf, rl, qry, len(qry), match, mismatch, gap_open, gap_ext,
                 C.byref(res), rc, qc, dptr)
    out = dict(score=res.score, start=res.start, end=res.end, len=res.len,
               max_row=res.maxRow, max_col=res.maxCol,
               ref_cons=rc.raw[:res.len], qry_cons=qc.raw[:res.len])
    if want_dirs:
        out["dirs"] = dirs
    return out


def sw_batch(qrys, refs, match=2, mismatch=-1, gap_open=-10, gap_ext=-1, nthreads=1):
    """Many pairs; returns (score, start, end, len, qcons list, rcons list)."""
    n = len(qrys)
    qb, qoff, qlen = pack(qrys)
    rb, roff, rlen = pack(refs)
    stride = (qlen + rlen + 2).astype(np.int64)
    coff = np.zeros(n, np.int64)
    if n:
        coff[1:] = np.cumsum(stride)[:-1]
    tot = int(stride.sum()) if n else 1
    qc = np.zeros(tot, np.uint8)
    rc = np.zeros(tot, np.uint8)
    sc, st, en, al = (np.zeros(n, np.int32) for _ in range(4))
    lib().orc_sw_batch(n, _p(qb, C.c_char), _p(qoff, C.c_int64), _p(qlen, C.c_int32),
                       _p(rb, C.c_char), _p(roff, C.c_int64), _p(rlen, C.c_int32),
                       match, mismatch, gap_open, gap_ext, _p(sc, C.c_int32),
                       _p(st, C.c_int32), _p(en, C.c_int32), _p(al, C.c_int32),
                       _p(qc, C.c_char), _p(rc, C.c_char), _p(coff, C.c_int64), nthreads)
    qcons = [qc[coff[i]:coff[i] + al[i]].tobytes() for i in range(n)]
    rcons = [rc[coff[i]:coff[i] + al[i]].tobytes() for i in range(n)]
    return sc, st, en, al, qcons, rcons


def pack(seqs):
    """List of bytes -> (uint8 buffer, int64 offsets, int32 lengths)."""
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    if len(seqs):
        off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    buf = np.frombuffer(b"".join(seqs) + b"\0", np.uint8).copy()
    return buf, off, lens


# ---------------------------------------------------------------- phase 1
class Index:
    """lookupTable5Create restated over in-memory sequences."""

    def __init__(self, seqs, k=12, thr=1000):
        self.seqs = [bytes(s) for s in seqs]
        self._keep = [C.c_char_p(s) for s in self.seqs]
        arr = (C.c_char_p * len(self.seqs))(*self.seqs)
        lens = (C.c_int64 * len(self.seqs))(*[len(s) for s in self.seqs])
        self.ix = _Index()
        rc = lib().orc_index_build(C.byref(self.ix), len(self.seqs), arr, lens, k, thr)
        if rc != 0:
            raise MemoryError("oracle index build failed")
        self.k, self.thr = k, thr

    def __del__(self):
        try:
            lib().orc_index_free(C.byref(self.ix))
        except Exception:
            pass

    @property
    def total(self):
        return int(self.ix.total)

    def arrays(self):
        nh, t = int(self.ix.nhash), int(self.ix.total)
        cp = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        return dict(start=cp(self.ix.start, nh), count=cp(self.ix.count, nh),
                    actual=cp(self.ix.actual, nh), refIdx=cp(self.ix.refIdx, t + 1),
                    refPos=cp(self.ix.refPos, t + 1))

    def map_query(self, seq: bytes, max_hits=3, tie_mode=TIE_FIRST, rng: Rng | None = None):
        oi = (C.c_int8 * max(max_hits, 1))()
        osh = (C.c_int32 * max(max_hits, 1))()
        op = (C.c_int32 * max(max_hits, 1))()
        nb = C.c_int()
        nh = C.c_int()
        n = lib().orc_map_query(C.byref(self.ix), seq, len(seq), max_hits, tie_mode,
                                C.byref(rng.s) if rng else None, oi, osh, op,
                                C.byref(nb), C.byref(nh))
        return [(oi[i], osh[i], op[i]) for i in range(n)], nb.value, nh.value


def best_hits(hits, max_hits=3, tie_mode=TIE_FIRST, rng: Rng | None = None):
    """mapHits6GetBestHits restated; hits = list of (refIdx, shift, refPos).
    Returns (best list, n_biggest, sorted input list)."""
    n = len(hits)
    ri = np.array([h[0] for h in hits], np.int8)
    sh = np.array([h[1] for h in hits], np.int32)
    rp = np.array([h[2] for h in hits], np.int32)
    m = max(max_hits, 1)
    oi = np.zeros(m, np.int8)
    osh = np.zeros(m, np.int32)
    op = np.zeros(m, np.int32)
    nb = C.c_int()
    k = lib().orc_best_hits(max_hits, _p(ri, C.c_int8), _p(sh, C.c_int32), _p(rp, C.c_int32),
                            n, tie_mode, C.byref(rng.s) if rng else None, _p(oi, C.c_int8),
                            _p(osh, C.c_int32), _p(op, C.c_int32), C.byref(nb))
    best = [(int(oi[i]), int(osh[i]), int(op[i])) for i in range(k)]
    srt = [(int(ri[i]), int(sh[i]), int(rp[i])) for i in range(n)]
    return best, nb.value, srt


def region(seq: bytes, ref_pos: int, before=100, win=300):
    out = C.create_string_buffer(win + 1)
    rd = C.c_int64()
    n = lib().orc_region(seq, len(seq), ref_pos, before, win, out, C.byref(rd))
    return out.raw[:n], rd.value


def choose_best(groups, scores, tie_mode=TIE_FIRST, rng: Rng | None = None):
    g = np.ascontiguousarray(groups, np.int32)
    s = np.ascontiguousarray(scores, np.int32)
    keep = np.zeros(len(g), np.int32)
    nt = np.zeros(len(g), np.int32)
    lib().orc_choose_best(len(g), _p(g, C.c_int32), _p(s, C.c_int32), tie_mode,
                          C.byref(rng.s) if rng else None, _p(keep, C.c_int32),
                          _p(nt, C.c_int32))
    return keep, nt


def pipeline(index: Index, reads, max_hits=3, match=2, mismatch=-1, gap_open=-10,
             gap_ext=-1, tie_mode=TIE_FIRST, seed=1):
    """Whole CPU path (searchQueries3 + alignQueries_cpu_mt) over `reads`;
    returns a list of record dicts in emission order."""
    rb, roff, rlen = pack(reads)
    seqs = index.seqs
    arr = (C.c_char_p * len(seqs))(*seqs)
    lens = (C.c_int64 * len(seqs))(*[len(s) for s in seqs])
    cap = 2 * max_hits * max(len(reads), 1)
    recs = (_Rec * cap)()
    stride = int(rlen.max() if len(reads) else 0) + 302
    rc = np.zeros(cap * stride, np.uint8)
    qc = np.zeros(cap * stride, np.uint8)
    n = lib().orc_pipeline(C.byref(index.ix), len(seqs), arr, lens, len(reads),
                           _p(rb, C.c_char), _p(roff, C.c_int64), _p(rlen, C.c_int32),
                           max_hits, match, mismatch, gap_open, gap_ext, tie_mode,
                           C.c_uint(seed), cap, recs, _p(rc, C.c_char), _p(qc, C.c_char),
                           stride)
    if n < 0:
        raise RuntimeError("record capacity exceeded")
    out = []
    for i in range(n):
        r = recs[i]
        d = {f: getattr(r, f) for f, _ in _Rec._fields_}
        d["ref_cons"] = rc[i * stride:i * stride + r.len].tobytes()
        d["qry_cons"] = qc[i * stride:i * stride + r.len].tobytes()
        out.append(d)
    return out


def pipeline_mt(index: Index, reads, nthreads=16, max_hits=3, match=2, mismatch=-1,
                gap_open=-10, gap_ext=-1):
    """pipeline() over `nthreads` host threads, each taking a contiguous
    share of the reads (tie rule "first", so records depend only on their
    own read): the same records in the same order as pipeline(), as the
    reference's CPU path runs its phases on 16 and 8 threads
    (src/search.cu:323-345, src/align.cu:2455)."""
    import threading
    n = len(reads)
    nthreads = max(1, min(nthreads, n))
    bounds = [n * t // nthreads for t in range(nthreads + 1)]
    parts = [None] * nthreads

    def work(t):
        parts[t] = pipeline(index, reads[bounds[t]:bounds[t + 1]], max_hits, match, mismatch,
                            gap_open, gap_ext, TIE_FIRST)

    th = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    out = []
    for t in range(nthreads):
        for d in parts[t]:
            d["read"] += bounds[t]
            out.append(d)
    return out


# ---------------------------------------------------------------- output text
def format_record(qry_name: str, ref_name: str, score: float, ref_distance: int,
                  start: int, end: int, qry_cons: bytes, ref_cons: bytes) -> str:
    """printOutput (src/output.cu:154-201) for one record: consensus arrays
    are stored reversed, so they print back to front."""
    r = ref_cons[::-1].decode("latin-1")
    q = qry_cons[::-1].decode("latin-1")
    mid = []
    for a, b in zip(r, q):
        if a.upper() == b.upper() or a == "N" or b == "N":
            mid.append("|")
        elif a == "-" or b == "-":
            mid.append("-")
        else:
            mid.append(".")
    return ("\n****************************************\n"
            f"Qry: {qry_name} | Ref: {ref_name} | Score: {score:.1f} | "
            f"AlignStart: {start + ref_distance} | AlignEnd: {end + ref_distance} | "
            f"AlignLength: {len(ref_cons)}\n"
            f"R: {r}\n   {''.join(mid)}\nQ: {q}\n")


def cigar(qry_cons: bytes, ref_cons: bytes) -> str:
    """printCigar (src/output.cu:33-79): walks the reversed consensus from
    the end; '-' in ref -> D, '-' in query -> I, otherwise M."""
    out = []
    op, n = " ", 0
    for z in range(len(ref_cons) - 1, -1, -1):
        c = "D" if ref_cons[z:z + 1] == b"-" else "I" if qry_cons[z:z + 1] == b"-" else "M"
        if n > 0 and op != c:
            out.append(f"{n}{op}")
            n = 0
        op = c
        n += 1
    if n > 0:
        out.append(f"{n}{op}")
    return "".join(out)
