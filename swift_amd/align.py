"""Host-side mirror of the reference's phase-2 interface, running on the HIP
library.

* ``align_sequences`` -- the reference's ``alignSequences`` launch boundary
  (src/align.cu:2904-2947): fixed strides, every query byte a row,
  reference columns up to the first NUL, float scores.
* ``sw_batch`` -- the batched, length-binned form (no reference
  counterpart) over pairs of uneven lengths.

Both return per-pair score/start/end/length and the consensus strings in
the reference's storage order (walk order = reversed alignment).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .device import DeviceBuffer


@dataclass
class Scoring:
    match: int = 2        # DEFAULT_MATCH (src/common.h:45)
    mismatch: int = -1    # DEFAULT_MISMATCH (src/common.h:46)
    gap_open: int = -10   # DEFAULT_GAP_OPEN_PENALTY (src/common.h:47)
    gap_ext: int = -1     # DEFAULT_GAP_EXT_PENALTY (src/common.h:48)


@dataclass
class Alignments:
    score: np.ndarray
    start: np.ndarray
    end: np.ndarray
    length: np.ndarray
    qry_cons: list
    ref_cons: list

    def record(self, i):
        return dict(score=int(self.score[i]), start=int(self.start[i]), end=int(self.end[i]),
                    len=int(self.length[i]), qry_cons=self.qry_cons[i], ref_cons=self.ref_cons[i])


def pack(seqs):
    """List of bytes -> (uint8 buffer, int64 offsets, int32 lengths)."""
    lens = np.fromiter((len(s) for s in seqs), dtype=np.int32, count=len(seqs))
    off = np.zeros(len(seqs), np.int64)
    if len(seqs):
        off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    buf = np.frombuffer(b"".join(seqs) + b"\0" * 8, np.uint8)
    return buf, off, lens


def align_sequences(queries, refs, query_len=None, ref_len=None, scoring=Scoring(),
                    status=True) -> Alignments:
    """Lays the pairs out as the reference's GPU driver does
    (src/align.cu:684-737: stride qlen+1 / rlen+1, NUL padded) and calls the
    alignSequences drop-in."""
    n = len(queries)
    ql = query_len if query_len is not None else max((len(q) for q in queries), default=0)
    rl = ref_len if ref_len is not None else max((len(r) for r in refs), default=0)
    qbuf = np.zeros((n, ql + 1), np.uint8)
    rbuf = np.zeros((n, rl + 1), np.uint8)
    for i, (q, r) in enumerate(zip(queries, refs)):
        qbuf[i, :len(q)] = np.frombuffer(q[:ql], np.uint8)
        rbuf[i, :min(len(r), rl)] = np.frombuffer(r[:rl], np.uint8)
    cl = ql + rl + 2
    dq, dr = DeviceBuffer.from_array(qbuf), DeviceBuffer.from_array(rbuf)
    dsc, dst, den, dal = (DeviceBuffer(4 * max(n, 1)) for _ in range(4))
    dqc, drc = DeviceBuffer(n * cl), DeviceBuffer(n * cl)
    args = [C.c_void_p(dq.ptr), C.c_void_p(dr.ptr), ql, rl, C.c_float(scoring.match),
            C.c_float(scoring.mismatch), C.c_float(scoring.gap_open), C.c_float(scoring.gap_ext),
            C.c_void_p(dsc.ptr), C.c_void_p(dst.ptr), C.c_void_p(den.ptr), C.c_void_p(dal.ptr),
            C.c_void_p(dqc.ptr), C.c_void_p(drc.ptr), n]
    if status:
        check(lib().swift_align_sequences(*args, None))
    else:
        lib().alignSequences(*args, 0, 1024)
    score = dsc.download(np.float32, n)
    start, end, alen = (b.download(np.int32, n) for b in (dst, den, dal))
    qc = dqc.download(np.uint8, n * cl).reshape(n, cl) if n else np.zeros((0, cl), np.uint8)
    rc = drc.download(np.uint8, n * cl).reshape(n, cl) if n else np.zeros((0, cl), np.uint8)
    return Alignments(score, start, end, alen,
                      [qc[i, :alen[i]].tobytes() for i in range(n)],
                      [rc[i, :alen[i]].tobytes() for i in range(n)])


def smith_waterman_cpu(ref: bytes, qry: bytes, ref_len: int | None = None, scoring=Scoring(),
                       start_pos: int = 0) -> dict:
    """The smithWaterman_cpu drop-in (src/smithWaterman.h:62-66) on one
    pair: zeroed (q+1) x (ref_len+1) matrices in, and out the score, start,
    end, length, consensus strings (walk order) and the H / rowBacktracker /
    colBacktracker matrices the call leaves.  start_pos is the caller's
    initial *startPos (kept when the alignment is empty)."""
    rl = len(ref) if ref_len is None else ref_len
    rbuf = C.create_string_buffer(ref[:rl], max(rl, 1) + 1)
    qbuf = C.create_string_buffer(qry, len(qry) + 1)
    q = len(qry)
    H = np.zeros((q + 1, rl + 1), np.float32)
    rowbt = np.zeros((q + 1, rl + 1), np.int32)
    colbt = np.zeros((q + 1, rl + 1), np.int32)
    n = q + rl + 2
    rc, qc = C.create_string_buffer(n), C.create_string_buffer(n)
    sc, st, en, ln = C.c_float(), C.c_int(start_pos), C.c_int(), C.c_int()
    lib().smithWaterman_cpu(
        H.ctypes.data_as(C.POINTER(C.c_float)), rowbt.ctypes.data_as(C.POINTER(C.c_int)),
        colbt.ctypes.data_as(C.POINTER(C.c_int)), rbuf, rl, qbuf, q, C.c_float(scoring.match),
        C.c_float(scoring.mismatch), C.c_float(scoring.gap_open), C.c_float(scoring.gap_ext),
        C.byref(sc), C.byref(st), C.byref(en), C.byref(ln), rc, qc)
    return dict(score=sc.value, start=st.value, end=en.value, len=ln.value,
                ref_cons=rc.raw[:max(ln.value, 0)], qry_cons=qc.raw[:max(ln.value, 0)],
                H=H, rowbt=rowbt, colbt=colbt, ref_tail=rc.raw[max(ln.value, 0):])


class PairBatch:
    """Pairs resident on the device in the batched layout."""

    def __init__(self, queries, refs):
        qb, qo, qlens = pack(queries)
        rb, ro, rlens = pack(refs)
        self._setup(qb, qo, qlens, rb, ro, rlens)

    @classmethod
    def from_arrays(cls, qbuf, qoff, qlen, rbuf, roff, rlen):
        self = cls.__new__(cls)
        self._setup(qbuf, np.asarray(qoff, np.int64), np.asarray(qlen, np.int32), rbuf,
                    np.asarray(roff, np.int64), np.asarray(rlen, np.int32))
        return self

    def _setup(self, qb, qo, qlens, rb, ro, rlens):
        self.n = len(qlens)
        self.qlen_h, self.rlen_h = qlens, rlens
        self.maxq = int(qlens.max()) if self.n else 0
        self.maxr = int(rlens.max()) if self.n else 0
        self.cells = int((qlens.astype(np.int64) * rlens).sum())
        cstride = (qlens.astype(np.int64) + rlens + 2 + 3) // 4 * 4
        coff = np.zeros(self.n, np.int64)
        if self.n:
            coff[1:] = np.cumsum(cstride)[:-1]
        self.coff_h = coff
        self.cons_bytes = int(cstride.sum()) if self.n else 1
        self.q, self.qo, self.ql = (DeviceBuffer.from_array(a) for a in (qb, qo, qlens))
        self.r, self.ro, self.rl = (DeviceBuffer.from_array(a) for a in (rb, ro, rlens))
        self.co = DeviceBuffer.from_array(coff)
        self.score, self.start, self.end, self.alen = (DeviceBuffer(4 * max(self.n, 1))
                                                       for _ in range(4))
        self.qc, self.rc = DeviceBuffer(self.cons_bytes), DeviceBuffer(self.cons_bytes)

    def run(self, scoring=Scoring(), stream=None) -> None:
        check(lib().swift_sw_batch(
            C.c_void_p(self.q.ptr), C.c_void_p(self.qo.ptr), C.c_void_p(self.ql.ptr),
            C.c_void_p(self.r.ptr), C.c_void_p(self.ro.ptr), C.c_void_p(self.rl.ptr),
            self.n, self.maxq, self.maxr, scoring.match, scoring.mismatch, scoring.gap_open,
            scoring.gap_ext, C.c_void_p(self.score.ptr), C.c_void_p(self.start.ptr),
            C.c_void_p(self.end.ptr), C.c_void_p(self.alen.ptr), C.c_void_p(self.qc.ptr),
            C.c_void_p(self.rc.ptr), C.c_void_p(self.co.ptr), stream))

    def results(self, strings=True) -> Alignments:
        n = self.n
        sc, st, en, al = (b.download(np.int32, n) for b in (self.score, self.start, self.end,
                                                             self.alen))
        qs, rs = [], []
        if strings:
            qc = self.qc.download(np.uint8, self.cons_bytes)
            rc = self.rc.download(np.uint8, self.cons_bytes)
            for i in range(n):
                o = self.coff_h[i]
                qs.append(qc[o:o + al[i]].tobytes())
                rs.append(rc[o:o + al[i]].tobytes())
        return Alignments(sc, st, en, al, qs, rs)


    def autotune(self, scoring=Scoring(), reps=2) -> dict:
        """Times the batched path's settings on these pairs and applies the
        fastest (see swift_autotune_batch)."""
        table = (C.c_float * (5 * 32))()
        bw, bi, bk, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        bms = C.c_float()
        check(lib().swift_autotune_batch(
            C.c_void_p(self.q.ptr), C.c_void_p(self.qo.ptr), C.c_void_p(self.ql.ptr),
            C.c_void_p(self.r.ptr), C.c_void_p(self.ro.ptr), C.c_void_p(self.rl.ptr),
            self.n, self.maxq, self.maxr, scoring.match, scoring.mismatch, scoring.gap_open,
            scoring.gap_ext, C.c_void_p(self.score.ptr), C.c_void_p(self.start.ptr),
            C.c_void_p(self.end.ptr), C.c_void_p(self.alen.ptr), C.c_void_p(self.qc.ptr),
            C.c_void_p(self.rc.ptr), C.c_void_p(self.co.ptr), reps, None, C.byref(bw),
            C.byref(bi), C.byref(bk), C.byref(bms), table, 32, C.byref(nt)))
        rows = [dict(waves_per_cu=int(table[5 * i]), int16=bool(table[5 * i + 1]),
                     walk_deferred=bool(table[5 * i + 2]), coarse_rows=bool(table[5 * i + 3]),
                     ms=round(float(table[5 * i + 4]), 3))
                for i in range(nt.value)]
        return dict(waves_per_cu=bw.value, allow_int16=bool(bi.value),
                    walk_deferred=bool(bk.value), coarse_rows=get_coarse_rows(),
                    ms=round(bms.value, 3), table=rows)


def sw_batch(queries, refs, scoring=Scoring()) -> Alignments:
    b = PairBatch(queries, refs)
    b.run(scoring)
    return b.results()


def batch_stats():
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    w = C.c_int64()
    check(lib().swift_sw_batch_stats(C.byref(a), C.byref(b), C.byref(c), C.byref(w)))
    return dict(ntasks16=a.value, ntasks32=b.value, nrerun=c.value, trace_words=w.value)


def set_tuning(waves_per_cu=0, allow_int16=True, walk_deferred=False):
    check(lib().swift_set_tuning(int(waves_per_cu), 1 if allow_int16 else 0,
                                 1 if walk_deferred else 0))


def get_tuning() -> dict:
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    check(lib().swift_get_tuning(C.byref(a), C.byref(b), C.byref(c)))
    return dict(waves_per_cu=a.value, allow_int16=bool(b.value), walk_deferred=bool(c.value),
                coarse_rows=get_coarse_rows())


def set_coarse_rows(mode: int) -> None:
    """1: coarse row classes (rows per lane rounded up to a multiple of 5;
    with the call's sizes read back, of 10 under 16,384 pairs, and under
    4,096 pairs one class in the wide form, a 64-lane group per wave);
    0: every R its own class; -1: chosen per call by batch size."""
    check(lib().swift_set_coarse_rows(int(mode)))


def set_plain_chunk(max_pairs: int) -> None:
    """Pairs per chunk of the plain path (0: by its 4 GiB scratch budget);
    the chunks of a call run one after another."""
    check(lib().swift_set_plain_chunk(int(max_pairs)))


def get_coarse_rows() -> int:
    m = C.c_int()
    check(lib().swift_get_coarse_rows(C.byref(m)))
    return m.value
