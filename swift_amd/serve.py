"""Serving loop over a resident index (continuous batching), on the HIP
library's swift_server_* entry points.

    srv = Server(index, nslots=4, mem_budget=32 << 30)
    rid = srv.submit(reads)          # non-blocking, any thread
    recs, latency_ms = srv.result(rid)

Requests are coalesced into device batches by the library's slots (one host
thread and stream each) under the memory budget; a request's records equal
those of ``Index.align_reads`` over the request alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SwiftError, check, lib
from .align import Scoring, pack
from .search import I32_FIELDS, Index, Records, _CRecords

SWIFT_ERR_TIMEOUT = 6


class Server:
    def __init__(self, index: Index, nslots=4, mem_budget=32 << 30, max_batch_reads=16384,
                 max_read_len=320, max_hits=3, scoring=Scoring()):
        self.index = index  # keeps the index alive
        self.max_read_len = max_read_len
        self._h = C.c_void_p()
        check(lib().swift_server_create(index._h, nslots, C.c_int64(mem_budget), max_batch_reads,
                                        max_read_len, max_hits, scoring.match, scoring.mismatch,
                                        scoring.gap_open, scoring.gap_ext, C.byref(self._h)))

    @staticmethod
    def pack(reads):
        """Reads -> (uint8 buffer, int64 offsets, int32 lengths) for submit_packed."""
        return pack([bytes(r) for r in reads])

    def submit(self, reads) -> int:
        return self.submit_packed(*self.pack(reads))

    def submit_packed(self, buf, off, lens) -> int:
        rid = C.c_uint64()
        check(lib().swift_server_submit(self._h, buf.ctypes.data_as(C.c_char_p),
                                        off.ctypes.data_as(C.POINTER(C.c_int64)),
                                        lens.ctypes.data_as(C.POINTER(C.c_int32)), len(lens),
                                        C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms=-1):
        """(nrecords, latency_ms), or None on timeout."""
        n = C.c_int()
        lat = C.c_double()
        rc = lib().swift_server_wait(self._h, C.c_uint64(rid), timeout_ms, C.byref(n),
                                     C.byref(lat))
        if rc == SWIFT_ERR_TIMEOUT:
            return None
        check(rc)
        return n.value, lat.value

    def fetch(self, rid: int, n: int, strings=True) -> Records:
        cap = max(n, 1)
        stride = self.max_read_len + 304
        cols = {f: np.zeros(cap, np.int32) for f in I32_FIELDS}
        rdist = np.zeros(cap, np.int64)
        qc = np.zeros(cap * stride, np.uint8) if strings else None
        rc = np.zeros(cap * stride, np.uint8) if strings else None
        nrec = np.zeros(1, np.int32)
        ptr = lambda a: a.ctypes.data if a is not None else None
        rec = _CRecords(nrecords=ptr(nrec), ref_distance=ptr(rdist), qry_cons=ptr(qc),
                        ref_cons=ptr(rc), cons_stride=stride, capacity=cap,
                        **{f: ptr(cols[f]) for f in I32_FIELDS})
        check(lib().swift_server_fetch(self._h, C.c_uint64(rid), C.byref(rec)))
        n = int(nrec[0])
        out = {f: v[:n] for f, v in cols.items()}
        out["ref_distance"] = rdist[:n]
        qs = [qc[i * stride:i * stride + out["alen"][i]].tobytes() for i in range(n)] if strings else []
        rs = [rc[i * stride:i * stride + out["alen"][i]].tobytes() for i in range(n)] if strings else []
        return Records(n, out, qs, rs)

    def result(self, rid: int, strings=True):
        n, lat = self.wait(rid)
        return self.fetch(rid, n, strings), lat

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(5)]
        check(lib().swift_server_stats(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("completed", "batches", "batched_reads", "max_coalesced",
                         "device_bytes"), (x.value for x in v)))

    def close(self):
        if self._h:
            lib().swift_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "SwiftError"]
