"""Host-side mirror of the reference's phase-1 interface (k-mer index,
read mapping) and of the phase-1 -> phase-2 pipeline, running on the HIP
library.

* ``Index`` -- lookupTable5Create (src/lookupTable5.cu:101-252) over
  sequences or a FASTA file, resident in device memory; ``arrays()`` returns
  the reference's tables (_lookupTable, _numRepeatsPerTuple,
  _numActualRepeatsPerTuple, _refIdx, _refPos).
* ``Index.map_reads`` -- lookupTable5MapQry + mapHits6GetBestHits
  (src/lookupTable5.cu:557-603, src/mapHits6.cu:40-70) for both strands of
  every read at once.
* ``Index.align_reads`` -- searchQueries3 + alignQueries_cpu_mt +
  chooseBestAlignment (src/search.cu:275-382, src/align.cu:2363-2511,
  844-905) on the device.
* Drop-in wrappers (``lookup_table5_*``, ``map_hits6_best_hits``,
  ``get_hash``, ``revcomp``, ``search_queries3``) call the reference-named
  C entry points.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import SwiftError, check, lib
from .align import Scoring, pack
from .device import DeviceBuffer


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Index:
    """Device-resident k-mer index."""

    def __init__(self, handle: C.c_void_p, k: int, thr: int):
        self._h = handle
        self.k, self.thr = k, thr

    @classmethod
    def from_sequences(cls, seqs, k=12, thr=1000) -> "Index":
        buf, off, lens = pack([bytes(s) for s in seqs])
        h = C.c_void_p()
        check(lib().swift_index_build(_p(buf, C.c_char), _p(off, C.c_int64),
                                      _p(lens.astype(np.int64), C.c_int64), len(seqs), k, thr,
                                      C.byref(h)))
        return cls(h, k, thr)

    @classmethod
    def from_packed(cls, buf, off, lens, k=12, thr=1000) -> "Index":
        """Sequences s = buf[off[s] .. off[s] + lens[s]) of a uint8 array."""
        buf = np.ascontiguousarray(buf, np.uint8)
        off = np.ascontiguousarray(off, np.int64)
        lens = np.ascontiguousarray(lens, np.int64)
        h = C.c_void_p()
        check(lib().swift_index_build(_p(buf, C.c_char), _p(off, C.c_int64), _p(lens, C.c_int64),
                                      len(off), k, thr, C.byref(h)))
        return cls(h, k, thr)

    @classmethod
    def from_fasta(cls, path: str, k=12, thr=1000) -> "Index":
        h = C.c_void_p()
        check(lib().swift_index_build_fasta(str(path).encode(), k, thr, C.byref(h)))
        return cls(h, k, thr)

    def close(self):
        if self._h:
            lib().swift_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        ns, mb = C.c_int(), C.c_int()
        tot, nh = C.c_int64(), C.c_int64()
        check(lib().swift_index_info(self._h, C.byref(ns), C.byref(tot), C.byref(nh),
                                     C.byref(mb)))
        return dict(nseq=ns.value, total=tot.value, nhash=nh.value, max_bucket=mb.value)

    def sequence(self, s: int):
        ln = C.c_int64()
        nm = C.c_char_p()
        check(lib().swift_index_sequence(self._h, s, C.byref(ln), C.byref(nm)))
        return ln.value, nm.value.decode(errors="replace")

    def arrays(self) -> dict:
        inf = self.info()
        nh, t = inf["nhash"], inf["total"]
        start, count, actual = (np.empty(nh, np.int32) for _ in range(3))
        ri = np.empty(t + 1, np.int8)
        rp = np.empty(t + 1, np.int32)
        check(lib().swift_index_export(self._h, _p(start, C.c_int32), _p(count, C.c_int32),
                                       _p(actual, C.c_int32), _p(ri, C.c_int8),
                                       _p(rp, C.c_int32)))
        return dict(start=start, count=count, actual=actual, refIdx=ri, refPos=rp)

    # ------------------------------------------------------------ mapping
    def map_reads(self, reads, max_hits=3, max_read_len=None) -> "StrandHits":
        rb = ReadBatch(reads, max_read_len)
        return self.map_batch(rb, max_hits)

    def map_batch(self, rb: "ReadBatch", max_hits=3) -> "StrandHits":
        ns = 2 * rb.n
        st = max(max_hits, 1)
        ref = DeviceBuffer(ns * st)
        sh, pos = DeviceBuffer(4 * ns * st), DeviceBuffer(4 * ns * st)
        nb, nbig, nh = (DeviceBuffer(4 * ns) for _ in range(3))
        check(lib().swift_map_reads(self._h, rb.seq.cptr, rb.off.cptr, rb.len.cptr, rb.n,
                                    rb.max_len, max_hits, ref.cptr, sh.cptr, pos.cptr, nb.cptr,
                                    nbig.cptr, nh.cptr, None))
        return StrandHits(ref.download(np.int8, ns * st).reshape(max(ns, 0), st),
                          sh.download(np.int32, ns * st).reshape(max(ns, 0), st),
                          pos.download(np.int32, ns * st).reshape(max(ns, 0), st),
                          nb.download(np.int32, ns), nbig.download(np.int32, ns),
                          nh.download(np.int32, ns))

    # ------------------------------------------------------------ pipeline
    def align_reads(self, reads, max_hits=3, scoring=Scoring(), max_read_len=None,
                    strings=True) -> "Records":
        rb = ReadBatch(reads, max_read_len)
        out = RecordBuffers(2 * rb.n * max(max_hits, 0), rb.max_len)
        self.align_batch(rb, out, max_hits, scoring)
        return out.results(strings)

    def align_batch(self, rb: "ReadBatch", out: "RecordBuffers", max_hits=3,
                    scoring=Scoring(), stream=None) -> None:
        check(lib().swift_align_reads(self._h, rb.seq.cptr, rb.off.cptr, rb.len.cptr, rb.n,
                                      rb.max_len, max_hits, scoring.match, scoring.mismatch,
                                      scoring.gap_open, scoring.gap_ext, C.byref(out.c), stream))


class ReadBatch:
    """Reads resident on the device."""

    @classmethod
    def from_packed(cls, buf, off, lens, max_read_len=None) -> "ReadBatch":
        self = cls.__new__(cls)
        self._setup(np.ascontiguousarray(buf, np.uint8), np.ascontiguousarray(off, np.int64),
                    np.ascontiguousarray(lens, np.int32), max_read_len)
        return self

    def __init__(self, reads, max_read_len=None):
        reads = [bytes(r) for r in reads]
        buf, off, lens = pack(reads)
        self._setup(buf, off, lens, max_read_len)

    def _setup(self, buf, off, lens, max_read_len):
        self.n = len(lens)
        self.max_len = int(max_read_len if max_read_len is not None
                           else (lens.max() if self.n else 0))
        self.seq = DeviceBuffer.from_array(buf)
        self.off = DeviceBuffer.from_array(off)
        self.len = DeviceBuffer.from_array(lens)


@dataclass
class StrandHits:
    """Per strand s = 2*read + strand: best hits [s, :nbest[s]]."""
    ref: np.ndarray
    shift: np.ndarray
    pos: np.ndarray
    nbest: np.ndarray
    nbiggest: np.ndarray
    nhits: np.ndarray

    def hits(self, s):
        return [(int(self.ref[s, i]), int(self.shift[s, i]), int(self.pos[s, i]))
                for i in range(int(self.nbest[s]))]


class _CRecords(C.Structure):
    _fields_ = [("nrecords", C.c_void_p), ("read", C.c_void_p), ("strand", C.c_void_p),
                ("ref_idx", C.c_void_p), ("shift", C.c_void_p), ("ref_pos", C.c_void_p),
                ("nbiggest", C.c_void_p), ("ref_distance", C.c_void_p), ("score", C.c_void_p),
                ("start", C.c_void_p), ("end", C.c_void_p), ("alen", C.c_void_p),
                ("keep", C.c_void_p), ("ntied", C.c_void_p), ("qry_cons", C.c_void_p),
                ("ref_cons", C.c_void_p), ("cons_stride", C.c_int64), ("capacity", C.c_int32)]


I32_FIELDS = ("read", "strand", "ref_idx", "shift", "ref_pos", "nbiggest", "score", "start",
              "end", "alen", "keep", "ntied")


class RecordBuffers:
    """Device buffers of a swift_records output."""

    def __init__(self, capacity: int, max_read_len: int):
        self.capacity = int(capacity)
        self.stride = (max_read_len + 302 + 3) // 4 * 4
        cap = max(self.capacity, 1)
        self.bufs = {f: DeviceBuffer(4 * cap) for f in I32_FIELDS}
        self.bufs["ref_distance"] = DeviceBuffer(8 * cap)
        self.bufs["nrecords"] = DeviceBuffer(4)
        self.bufs["qry_cons"] = DeviceBuffer(cap * self.stride)
        self.bufs["ref_cons"] = DeviceBuffer(cap * self.stride)
        self.c = _CRecords(**{k: b.ptr for k, b in self.bufs.items()},
                           cons_stride=self.stride, capacity=self.capacity)

    def results(self, strings=True) -> "Records":
        n = int(self.bufs["nrecords"].download(np.int32, 1)[0])
        cols = {f: self.bufs[f].download(np.int32, n) for f in I32_FIELDS}
        cols["ref_distance"] = self.bufs["ref_distance"].download(np.int64, n)
        qc, rc = [], []
        if strings and n:
            q = self.bufs["qry_cons"].download(np.uint8, n * self.stride).reshape(n, self.stride)
            r = self.bufs["ref_cons"].download(np.uint8, n * self.stride).reshape(n, self.stride)
            for i in range(n):
                qc.append(q[i, :cols["alen"][i]].tobytes())
                rc.append(r[i, :cols["alen"][i]].tobytes())
        return Records(n, cols, qc, rc)


@dataclass
class Records:
    n: int
    cols: dict
    qry_cons: list
    ref_cons: list

    def record(self, i) -> dict:
        d = {k: int(v[i]) for k, v in self.cols.items()}
        if self.qry_cons:
            d["qry_cons"] = self.qry_cons[i]
            d["ref_cons"] = self.ref_cons[i]
        return d


class PipelineGraph:
    """Phase 1 -> phase 2 captured as one hipGraph over fixed device buffers:
    `slots` read slots of up to `max_read_len` bases.  run(reads) rewrites
    the read buffers (unused slots become length-0 reads), replays the graph
    and returns the records."""

    def __init__(self, index: Index, slots: int, max_read_len: int, max_hits=3,
                 scoring=Scoring()):
        self.index, self.slots, self.max_len = index, int(slots), int(max_read_len)
        self.seq = DeviceBuffer(self.slots * self.max_len + 64)
        off = np.arange(self.slots, dtype=np.int64) * self.max_len
        self.off = DeviceBuffer.from_array(off)
        self.len = DeviceBuffer(4 * self.slots)
        self.len.upload(np.zeros(self.slots, np.int32))
        self.out = RecordBuffers(2 * self.slots * max_hits, self.max_len)
        self._h = C.c_void_p()
        check(lib().swift_pipeline_graph_create(
            index._h, self.seq.cptr, self.off.cptr, self.len.cptr, self.slots, self.max_len,
            max_hits, scoring.match, scoring.mismatch, scoring.gap_open, scoring.gap_ext,
            C.byref(self.out.c), C.byref(self._h)))

    def load(self, reads) -> None:
        if len(reads) > self.slots:
            raise SwiftError(f"{len(reads)} reads for {self.slots} slots")
        buf = np.zeros(self.slots * self.max_len, np.uint8)
        lens = np.zeros(self.slots, np.int32)
        for i, r in enumerate(reads):
            r = bytes(r)
            if len(r) > self.max_len:
                raise SwiftError(f"read {i} longer than {self.max_len}")
            buf[i * self.max_len:i * self.max_len + len(r)] = np.frombuffer(r, np.uint8)
            lens[i] = len(r)
        self.seq.upload(buf)
        self.len.upload(lens)

    def launch(self, stream=None) -> None:
        check(lib().swift_pipeline_graph_launch(self._h, stream))

    def check(self) -> None:
        """Length and capacity errors of the last replay (after its stream
        has been synchronised)."""
        check(lib().swift_pipeline_graph_check(self._h))

    def run(self, reads, strings=True) -> "Records":
        from .device import synchronize
        self.load(reads)
        self.launch()
        synchronize()
        self.check()
        return self.out.results(strings)

    def close(self):
        if self._h:
            lib().swift_pipeline_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- drop-ins
class Query(C.Structure):
    """struct queryStruct (src/query.h:29-37)."""
    _fields_ = [("name", C.c_char_p), ("length", C.c_int), ("seq", C.c_char_p),
                ("revCompSeq", C.c_char_p), ("nameOffset", C.c_ulong),
                ("seqOffset", C.c_ulong)]


def get_hash(s: bytes) -> int:
    return lib().getHash(C.c_char_p(s), len(s))


def revcomp(s: bytes) -> bytes:
    out = C.create_string_buffer(len(s) + 1)
    lib().qryGetReverseComplement(C.c_char_p(s), len(s), out)
    return out.raw[:len(s)]


def lookup_table5_create(path: str, k: int, max_hits: int, thr: int) -> int:
    tot = C.c_int()
    lib().lookupTable5Create(str(path).encode(), k, max_hits, thr, C.byref(tot))
    return tot.value


def lookup_table5_map_qry(seq: bytes, rev: bool = False, max_hits: int = 64):
    rc = revcomp(seq)
    q = Query(b"q", len(seq), seq, rc, 0, 0)
    ri = (C.c_char * max(max_hits, 1))()
    sh = (C.c_int * max(max_hits, 1))()
    rp = (C.c_int * max(max_hits, 1))()
    n = lib().lookupTable5MapQry(C.byref(q), len(seq), 1 if rev else 0, ri, sh, rp)
    return [(int.from_bytes(ri[i], "little", signed=True), sh[i], rp[i]) for i in range(n)]


def lookup_table5_delete() -> None:
    lib().lookupTable5Delete()


def map_hits6_best_hits(hits, max_hits=3):
    """Returns (best, sorted input)."""
    n = len(hits)
    ri = np.array([h[0] for h in hits], np.int8)
    sh = np.array([h[1] for h in hits], np.int32)
    rp = np.array([h[2] for h in hits], np.int32)
    m = max(max_hits, 1)
    oi, osh, op = np.zeros(m, np.int8), np.zeros(m, np.int32), np.zeros(m, np.int32)
    k = lib().mapHits6GetBestHits(max_hits, _p(ri, C.c_char), _p(sh, C.c_int), _p(rp, C.c_int), n,
                                  _p(oi, C.c_char), _p(osh, C.c_int), _p(op, C.c_int))
    return ([(int(oi[i]), int(osh[i]), int(op[i])) for i in range(k)],
            [(int(ri[i]), int(sh[i]), int(rp[i])) for i in range(n)])


def search_queries3(qry_path, ref_path, match_path, k=12, max_hits=3, thr=1000,
                    num_qrs=1 << 30) -> int:
    nm = C.c_int()
    lib().searchQueries3(str(qry_path).encode(), num_qrs, 0, str(ref_path).encode(), k,
                         str(match_path).encode(), max_hits, C.byref(nm), 16, thr)
    return nm.value


def set_map_mode(mode: int) -> None:
    """Mapping tier a strand starts in: -1 by the index (default), 0 the
    one-wave tier, 1 the hash-cluster tier.  Results are the same."""
    check(lib().swift_set_map_mode(int(mode)))


def get_map_mode() -> int:
    m = C.c_int()
    check(lib().swift_get_map_mode(C.byref(m)))
    return m.value


def search2_queries(qry_path, ref_path, match_path, k=12, thr=1000) -> int:
    """search2Queries (src/search2.h:26-27): every read, at most 3 hits per
    strand, shift column written as 0."""
    nm = C.c_int()
    lib().search2Queries(str(qry_path).encode(), str(ref_path).encode(),
                         str(match_path).encode(), k, 3, C.byref(nm), thr)
    return nm.value


def align_queries(qry_path, ref_path, match_path, out_path, qry_len, out_format=0,
                  scoring=Scoring(), num_matches=None, gpu_driver=True) -> None:
    """Phase 2 from the match file, as the reference's drivers run it:
    gpu_driver=True follows alignQueries3 (src/align.cu:574-808: the first
    num_matches records, text output keeps each read's best record), False
    alignQueries_cpu_mt (src/align.cu:2363-2511: every record, in rounds of 8
    match lines).  out_format 0 = printOutput text, 1 = SAM."""
    if num_matches is None:
        with open(match_path, "rb") as f:
            num_matches = sum(1 for _ in f)
    check(lib().swift_align_queries(
        str(qry_path).encode(), int(qry_len), str(ref_path).encode(),
        str(match_path).encode(), int(out_format), C.c_float(scoring.match),
        C.c_float(scoring.mismatch), C.c_float(scoring.gap_open), C.c_float(scoring.gap_ext),
        str(out_path).encode(), int(num_matches), 1 if gpu_driver else 0))


__all__ = ["Index", "ReadBatch", "StrandHits", "RecordBuffers", "Records", "PipelineGraph",
           "Query", "get_hash",
           "revcomp", "lookup_table5_create", "lookup_table5_map_qry", "lookup_table5_delete",
           "map_hits6_best_hits", "search_queries3", "align_queries", "SwiftError"]
