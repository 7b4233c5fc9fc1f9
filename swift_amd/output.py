"""Alignment output in the reference's formats, on the HIP library: the
printOutput / printOutputSam / printCigar drop-ins (src/output.cu:33-201),
CIGAR strings of device records, and a writer of pipeline records."""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

from ._lib import check, lib
from .search import I32_FIELDS, Index, RecordBuffers, _CRecords

_libc = C.CDLL(None)
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]
NAME_STRIDE = 50  # MAX_SEQ_NAME_LENGTH


def _through_file(call) -> str:
    fd, path = tempfile.mkstemp()
    os.close(fd)
    try:
        f = _libc.fopen(path.encode(), b"w")
        call(C.c_void_p(f))
        _libc.fclose(C.c_void_p(f))
        return open(path, "rb").read().decode("latin-1")
    finally:
        os.unlink(path)


def _names(names):
    buf = np.zeros(len(names) * NAME_STRIDE, np.uint8)
    for i, n in enumerate(names):
        b = n.encode()[:NAME_STRIDE - 1]
        buf[i * NAME_STRIDE:i * NAME_STRIDE + len(b)] = np.frombuffer(b, np.uint8)
    return buf


def _cons(strings, stride):
    buf = np.zeros(max(len(strings), 1) * stride, np.uint8)
    for i, s in enumerate(strings):
        buf[i * stride:i * stride + len(s)] = np.frombuffer(bytes(s), np.uint8)
    return buf


def print_output(qry_names, ref_names, scores, ref_distance, starts, ends, qry_cons, ref_cons,
                 sam=False) -> str:
    """printOutput (or printOutputSam) over host records; returns the text."""
    n = len(qry_names)
    alens = np.array([len(s) for s in qry_cons], np.int32)
    stride = int(alens.max() if n else 0) + 2
    qn, rn = _names(qry_names), _names(ref_names)
    sc = np.asarray(scores, np.float32)
    rd = np.asarray(ref_distance, np.int32)
    st, en = np.asarray(starts, np.int32), np.asarray(ends, np.int32)
    qc, rc = _cons(qry_cons, stride), _cons(ref_cons, stride)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    if sam:
        return _through_file(lambda f: lib().printOutputSam(
            n, p(qn), p(rn), p(sc), p(st), p(en), p(alens), stride, p(qc), p(rc), f))
    return _through_file(lambda f: lib().printOutput(
        n, p(qn), p(rn), p(sc), p(rd), p(st), p(en), p(alens), stride, p(qc), p(rc), f))


def cigar(qry_cons: bytes, ref_cons: bytes) -> str:
    return _through_file(lambda f: lib().printCigar(C.c_char_p(qry_cons), C.c_char_p(ref_cons),
                                                    len(qry_cons), f))


def records_cigar(out: RecordBuffers):
    """CIGAR strings of the records in `out` (computed on the device)."""
    from .device import DeviceBuffer, synchronize
    cap = max(out.capacity, 1)
    stride = 2 * out.stride
    buf, lens = DeviceBuffer(cap * stride), DeviceBuffer(4 * cap)
    check(lib().swift_records_cigar(C.byref(out.c), out.capacity, buf.cptr, C.c_int64(stride),
                                    lens.cptr, None))
    synchronize()
    n = int(out.bufs["nrecords"].download(np.int32, 1)[0])
    ln = lens.download(np.int32, n)
    raw = buf.download(np.uint8, n * stride)
    return [raw[i * stride:i * stride + ln[i]].tobytes().decode() for i in range(n)]


def write_records(path, records, qry_names, index: Index, sam=False, only_kept=False) -> None:
    """Pipeline records (search.Records with strings) to `path`."""
    n = records.n
    stride = max((len(s) for s in records.qry_cons), default=0) + 2
    cols = {f: np.ascontiguousarray(records.cols[f], np.int32) for f in I32_FIELDS}
    rd = np.ascontiguousarray(records.cols["ref_distance"], np.int64)
    qc, rc = _cons(records.qry_cons, stride), _cons(records.ref_cons, stride)
    ptr = lambda a: a.ctypes.data
    rec = _CRecords(nrecords=None, ref_distance=ptr(rd), qry_cons=ptr(qc), ref_cons=ptr(rc),
                    cons_stride=stride, capacity=n, **{f: ptr(cols[f]) for f in I32_FIELDS})
    names = (C.c_char_p * max(len(qry_names), 1))(*[s.encode() for s in qry_names])
    check(lib().swift_write_records(str(path).encode(), 1 if sam else 0, 1 if only_kept else 0,
                                    C.byref(rec), n, names, index._h))
