"""swift_align_sequences_noalloc against swift_align_sequences.

The no-alloc entry point keeps the plain path's traceback in a fixed device
array; its results must equal the allocating entry point's bit for bit on the
same pairs.  Two calls in a row, the second with fewer pairs and other
lengths, so stale contents of the shared array from the first call would
show.  The first call has more pairs than the array holds at once (it runs
in chunks)."""
import ctypes as C
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(rnd, n, q, r):
    qs, rs = [], []
    for _ in range(n):
        x = bytes(rnd.choice(b"ACGT") for _ in range(q))
        y = bytearray(bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(0, 40))) + x)
        for k in range(len(y)):
            if rnd.random() < 0.04:
                y[k] = rnd.choice(b"ACGTN")
        y = bytes(y) + bytes(rnd.choice(b"ACGT") for _ in range(r))
        # some references shorter than the stride: NUL padded
        cut = r if rnd.random() < 0.7 else rnd.randint(1, r)
        rs.append(y[:cut])
        qs.append(x)
    return qs, rs


def _run(fn, qs, rs, q, r, extra):
    from swift_amd._lib import check, lib
    from swift_amd.device import DeviceBuffer
    n = len(qs)
    qa = np.zeros((n, q + 1), np.uint8)
    ra = np.zeros((n, r + 1), np.uint8)
    for i in range(n):
        qa[i, :len(qs[i])] = np.frombuffer(qs[i], np.uint8)
        ra[i, :len(rs[i])] = np.frombuffer(rs[i], np.uint8)
    cl = q + r + 2
    b = dict(q=DeviceBuffer.from_array(qa), r=DeviceBuffer.from_array(ra),
             sc=DeviceBuffer(4 * n), st=DeviceBuffer(4 * n), en=DeviceBuffer(4 * n),
             al=DeviceBuffer(4 * n), qc=DeviceBuffer(n * cl), rc=DeviceBuffer(n * cl))
    for k in ("qc", "rc"):
        b[k].zero()
    check(getattr(lib(), fn)(
        C.c_void_p(b["q"].ptr), C.c_void_p(b["r"].ptr), q, r, C.c_float(2), C.c_float(-1),
        C.c_float(-10), C.c_float(-1), C.c_void_p(b["sc"].ptr), C.c_void_p(b["st"].ptr),
        C.c_void_p(b["en"].ptr), C.c_void_p(b["al"].ptr), C.c_void_p(b["qc"].ptr),
        C.c_void_p(b["rc"].ptr), n, *extra))
    check(lib().swift_device_synchronize())
    return (b["sc"].download(np.float32, n), b["st"].download(np.int32, n),
            b["en"].download(np.int32, n), b["al"].download(np.int32, n),
            b["qc"].download(np.uint8, n * cl), b["rc"].download(np.uint8, n * cl))


def test_noalloc_matches_allocating_entry_point():
    rnd = random.Random(77)
    # 100 x 300 pairs: 23,326 words each (trace, best cell, moves), 719 per
    # chunk of the 64 MiB array, so 3,000 pairs take 5 chunks; then fewer,
    # longer pairs
    for n, q, r in ((3000, 100, 300), (300, 150, 300), (40, 64, 120)):
        qs, rs = _pairs(rnd, n, q, r)
        want = _run("swift_align_sequences", qs, rs, q, r, [None])
        got = _run("swift_align_sequences_noalloc", qs, rs, q, r, [None])
        for a, b in zip(got, want):
            assert np.array_equal(a, b), (n, q, r)


def test_noalloc_rejects_oversize():
    from swift_amd._lib import lib
    L = lib()
    # a query beyond the plain kernel's rows: rejected on the host, no launch
    rc = L.swift_align_sequences_noalloc(None, None, 5000, 300, C.c_float(2), C.c_float(-1),
                                          C.c_float(-10), C.c_float(-1), None, None, None,
                                          None, None, None, 1, None)
    assert rc != 0 and b"beyond" in L.swift_last_error()
