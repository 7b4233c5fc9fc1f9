"""Reference drop-ins called the way the reference calls them.

* smithWaterman_cpu (src/smithWaterman.h:62-66): one pair per call, the
  matrices in the caller's zeroed buffers, from 8 threads at once as
  getAlignments runs it (src/align.cu:2455,2519-2548).  Checked against the
  reference test suite's known answers, the oracle, and -- where
  oracle/_ref is built -- the reference's own compiled smithWaterman_cpu,
  every matrix cell included.
* lookupTable5MapQry from 16 threads at once, as searchQueries3's getHits
  threads call it (src/search.cu:323-345,1479-1488), against the oracle.
* mapHits6GetBestHits over inputs that shrink between calls (its scratch is
  kept from call to call).
* alignSequences' status form on two streams at once (its traceback buffer
  is shared between calls).
"""
import ctypes as C
import json
import os
import random
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))


@pytest.fixture(scope="module")
def A():
    from swift_amd import align
    return align


def rand_seq(rnd, n, alphabet=b"ACGT"):
    return bytes(rnd.choice(alphabet) for _ in range(n))


def mutate(rnd, s, rate, alphabet=b"ACGT"):
    out = bytearray()
    for c in s:
        u = rnd.random()
        if u < rate / 3:
            continue
        if u < 2 * rate / 3:
            out += bytes([rnd.choice(alphabet)])
        if u < rate:
            out += bytes([rnd.choice(alphabet)])
        else:
            out.append(c)
    return bytes(out)


def oracle_matrices(orc, ref, qry, ref_len, sc):
    """rowBacktracker / colBacktracker as the oracle's direction codes give
    them (0 start, 1 left, 2 up, 3 diagonal; src/smithWaterman.cu:154-171),
    over the cells the reference writes."""
    o = orc.sw(ref, qry, *sc, ref_len=ref_len, want_dirs=True)
    d = o["dirs"].astype(np.int32)
    q1, r1 = d.shape
    i = np.arange(q1)[:, None].repeat(r1, 1)
    j = np.arange(r1)[None, :].repeat(q1, 0)
    row = np.where((d == 3) | (d == 2), i - 1, i)
    col = np.where((d == 3) | (d == 1), j - 1, j)
    ncols = (ref + b"\0" * ref_len)[:ref_len].find(b"\0")  # the buffer is NUL padded
    ncols = ref_len if ncols < 0 else ncols
    written = np.zeros_like(d, bool)
    written[1:, 1:ncols + 1] = True
    return o, np.where(written, row, 0), np.where(written, col, 0), written


def ref_matrices(refbridge, ref, qry, ref_len, sc, start0=0):
    q = len(qry)
    H = np.zeros((q + 1, ref_len + 1), np.float32)
    rb = np.zeros((q + 1, ref_len + 1), np.int32)
    cb = np.zeros((q + 1, ref_len + 1), np.int32)
    n = q + ref_len + 2
    rc, qc = C.create_string_buffer(n), C.create_string_buffer(n)
    s, st, en, ln = C.c_float(), C.c_int(start0), C.c_int(), C.c_int()
    rbuf = C.create_string_buffer(ref[:ref_len], ref_len + 1)
    refbridge.ref_sw_matrices(rbuf, ref_len, qry, q, C.c_float(sc[0]), C.c_float(sc[1]),
                              C.c_float(sc[2]), C.c_float(sc[3]),
                              H.ctypes.data_as(C.POINTER(C.c_float)),
                              rb.ctypes.data_as(C.POINTER(C.c_int)),
                              cb.ctypes.data_as(C.POINTER(C.c_int)), C.byref(s), C.byref(st),
                              C.byref(en), C.byref(ln), rc, qc)
    return dict(score=s.value, start=st.value, end=en.value, len=ln.value, H=H, rowbt=rb,
                colbt=cb, ref_cons=rc.raw[:ln.value], qry_cons=qc.raw[:ln.value])


# ------------------------------------------------------------ smithWaterman_cpu
@pytest.mark.parametrize("case", GOLD["sw"], ids=lambda c: c["source"][:40])
def test_smith_waterman_cpu_known(A, case):
    ref, qry = case["ref"].encode(), case["qry"].encode()
    rl = case.get("ref_len", len(ref))
    got = A.smith_waterman_cpu(ref, qry, rl, A.Scoring(*case["scoring"]))
    assert got["score"] == case["score"]
    for k in ("start", "end", "len"):
        if k in case:
            assert got[k] == case[k], k
    if "qry_cons" in case:
        assert got["qry_cons"] == case["qry_cons"].encode()
        assert got["ref_cons"] == case["ref_cons"].encode()
    # cellMax: row-major index of the best cell, the first strict maximum
    assert int(np.argmax(got["H"].ravel())) == case["cell_max"]


def sw_cases(seed, n):
    rnd = random.Random(seed)
    cases = []
    for t in range(n):
        q = rnd.randint(1, 160)
        alpha = b"ACGT" if t % 3 else b"ACGTNnacgtX"
        qry = rand_seq(rnd, q, alpha)
        ref = rand_seq(rnd, rnd.randint(0, 40)) + mutate(rnd, qry, 0.08) + \
            rand_seq(rnd, rnd.randint(0, 40), alpha)
        ref_len = len(ref) + (rnd.randint(1, 20) if t % 4 == 0 else 0)  # NUL-padded buffer
        sc = [(2, -1, -10, -1), (1, -3, -5, -2), (3, -2, -4, -4), (2, -1, -1, -1)][t % 4]
        cases.append((ref, qry, ref_len, sc))
    # no positive cell at all, and empty inputs
    cases += [(b"AAAA", b"CCCC", 4, (2, -1, -10, -1)), (b"", b"ACGT", 0, (2, -1, -10, -1)),
              (b"ACGT", b"", 4, (2, -1, -10, -1)), (b"\0\0ACGT", b"ACGT", 6, (2, -1, -10, -1))]
    return cases


def test_smith_waterman_cpu_vs_oracle(A, orc):
    for ref, qry, rl, sc in sw_cases(7, 120):
        got = A.smith_waterman_cpu(ref, qry, rl, A.Scoring(*sc), start_pos=5)
        o, row, col, written = oracle_matrices(orc, ref, qry, rl, sc)
        assert got["score"] == o["score"], (ref, qry)
        assert got["len"] == o["len"] and got["qry_cons"] == o["qry_cons"]
        assert got["ref_cons"] == o["ref_cons"]
        # an empty alignment leaves *startPos as the caller set it
        want_start = o["start"] if o["len"] > 0 else 5
        assert got["start"] == want_start and got["end"] == want_start + (o["end"] - o["start"])
        assert np.array_equal(got["rowbt"], row) and np.array_equal(got["colbt"], col)
        assert not got["H"][~written].any()
        assert got["ref_tail"] == b"\0" * len(got["ref_tail"])  # nothing past alignLen


def test_smith_waterman_cpu_vs_reference(A, refbridge):
    """Every cell of the three matrices and every output against the
    reference's own compiled smithWaterman_cpu, from the same zeroed
    buffers and initial *startPos."""
    for ref, qry, rl, sc in sw_cases(8, 80):
        got = A.smith_waterman_cpu(ref, qry, rl, A.Scoring(*sc), start_pos=3)
        want = ref_matrices(refbridge, ref, qry, rl, sc, start0=3)
        for k in ("score", "start", "end", "len", "qry_cons", "ref_cons"):
            assert got[k] == want[k], (k, ref, qry)
        for k in ("H", "rowbt", "colbt"):
            assert np.array_equal(got[k], want[k]), (k, ref, qry)


def test_smith_waterman_cpu_threads(A, orc):
    """Eight threads at once, each with its own pairs, as getAlignments."""
    cases = sw_cases(9, 96)
    out = [None] * len(cases)

    def work(t):
        for i in range(t, len(cases), 8):
            ref, qry, rl, sc = cases[i]
            out[i] = A.smith_waterman_cpu(ref, qry, rl, A.Scoring(*sc))

    th = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (ref, qry, rl, sc), got in zip(cases, out):
        o = orc.sw(ref, qry, *sc, ref_len=rl)
        assert (got["score"], got["len"], got["qry_cons"], got["ref_cons"]) == \
            (o["score"], o["len"], o["qry_cons"], o["ref_cons"])
        if o["len"] > 0:
            assert (got["start"], got["end"]) == (o["start"], o["end"])


# ------------------------------------------------------------ phase-1 drop-ins
def test_lookup_table5_map_qry_16_threads(orc, tmp_path):
    """16 host threads call lookupTable5MapQry at once on both strands of
    their reads; every result equals the oracle's (tie rule: first)."""
    from swift_amd import search, synth
    rnd = random.Random(21)
    rng = np.random.default_rng(21)
    seqs = [synth.random_bases(rng, n).tobytes() for n in (60000, 41000, 23000)]
    unit = b"ACGTTGCAGG"
    seqs[1] = seqs[1][:5000] + unit * 300 + seqs[1][8000:]  # repeats: ties and big clusters
    fa = tmp_path / "ref.fa"
    with open(fa, "w") as f:
        for s, seq in enumerate(seqs):
            f.write(f">s{s}\n")
            for p in range(0, len(seq), 50):
                f.write(seq[p:p + 50].decode() + "\n")
    reads, _ = synth.sample_reads(rng, [np.frombuffer(s, np.uint8) for s in seqs], 400, 100,
                                  0.02, 0.002)
    reads += [rand_seq(rnd, rnd.randint(12, 199)) for _ in range(40)]
    reads += [unit * 10, b"ACGT" * 3, b"N" * 50]
    search.lookup_table5_create(fa, 12, 3, 1000)
    oix = orc.Index(seqs, 12, 1000)
    got = [None] * len(reads)
    err = []

    def work(t):
        try:
            for i in range(t, len(reads), 16):
                got[i] = (search.lookup_table5_map_qry(reads[i], False, 3),
                          search.lookup_table5_map_qry(reads[i], True, 3))
        except Exception as e:  # noqa: BLE001 - reported below
            err.append(repr(e))

    th = [threading.Thread(target=work, args=(t,)) for t in range(16)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    search.lookup_table5_delete()
    assert not err, err
    for i, r in enumerate(reads):
        want_f = oix.map_query(r, 3)[0]
        want_r = oix.map_query(orc.revcomp(r), 3)[0]
        assert got[i] == (want_f, want_r), i


def test_map_hits6_shrinking_inputs(orc):
    """Calls of decreasing size after large ones: results never carry over
    from an earlier call's scratch."""
    from swift_amd import search
    rnd = random.Random(33)
    for n in (5000, 3000, 64, 7, 1, 2, 900, 3):
        hits = [(rnd.randint(-2, 2), rnd.randint(-6, 6), rnd.randint(-90, 90))
                for _ in range(n)]
        mh = rnd.randint(1, 4)
        best, srt = search.map_hits6_best_hits(hits, mh)
        obest, _, osrt = orc.best_hits(hits, mh, orc.TIE_FIRST)
        assert srt == osrt and best == obest, n


# ------------------------------------------------------------ alignSequences
def test_align_sequences_two_streams(A, orc):
    """Two calls of the plain drop-in's status form in flight on two streams
    at once (each launch returns before its kernels end): both match the
    oracle, so neither call's traceback was overwritten by the other's."""
    from swift_amd.device import DeviceBuffer
    rnd = random.Random(41)
    groups = []
    for q, r, n in ((100, 300, 3000), (150, 300, 2000)):
        qs = [rand_seq(rnd, q) for _ in range(n)]
        rs = [rand_seq(rnd, 30) + mutate(rnd, x, 0.05) + rand_seq(rnd, 40) for x in qs]
        rs = [(x + b"A" * r)[:r] for x in rs]
        qa = np.frombuffer(b"".join(x + b"\0" for x in qs), np.uint8).reshape(n, q + 1)
        ra = np.frombuffer(b"".join(x + b"\0" for x in rs), np.uint8).reshape(n, r + 1)
        cl = q + r + 2
        b = dict(q=DeviceBuffer.from_array(qa), r=DeviceBuffer.from_array(ra),
                 sc=DeviceBuffer(4 * n), st=DeviceBuffer(4 * n), en=DeviceBuffer(4 * n),
                 al=DeviceBuffer(4 * n), qc=DeviceBuffer(n * cl), rc=DeviceBuffer(n * cl))
        groups.append((q, r, n, qs, rs, b))
    from swift_amd._lib import check, lib
    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        check(lib().swift_stream_create(C.byref(s)))
    check(lib().swift_device_synchronize())
    for rep in range(3):
        for (q, r, n, _, _, b), s in zip(groups, streams):
            check(lib().swift_align_sequences(
                C.c_void_p(b["q"].ptr), C.c_void_p(b["r"].ptr), q, r, C.c_float(2),
                C.c_float(-1), C.c_float(-10), C.c_float(-1), C.c_void_p(b["sc"].ptr),
                C.c_void_p(b["st"].ptr), C.c_void_p(b["en"].ptr), C.c_void_p(b["al"].ptr),
                C.c_void_p(b["qc"].ptr), C.c_void_p(b["rc"].ptr), n, s))
        check(lib().swift_device_synchronize())
        for q, r, n, qs, rs, b in groups:
            sc = b["sc"].download(np.float32, n)
            st = b["st"].download(np.int32, n)
            al = b["al"].download(np.int32, n)
            qc = b["qc"].download(np.uint8, n * (q + r + 2)).reshape(n, q + r + 2)
            for i in range(0, n, 7):
                o = orc.sw(rs[i], qs[i])
                assert (int(sc[i]), int(st[i]), int(al[i])) == (o["score"], o["start"],
                                                               o["len"]), (rep, q, i)
                assert qc[i, :o["len"]].tobytes() == o["qry_cons"], (rep, q, i)
    for s in streams:
        check(lib().swift_stream_destroy(s))
