"""Phase-2 parity on the GPU: the HIP kernels against the oracle on the
same seeded inputs, bit for bit (scores, start/end, lengths, consensus
strings)."""
import json
import os
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))
KEYS = ("score", "start", "end", "len", "qry_cons", "ref_cons")


@pytest.fixture(scope="module")
def A():
    from swift_amd import align
    return align


def compare(got, orc_res, ctx):
    for k in KEYS:
        assert got[k] == orc_res[k], (ctx, k, got, {k2: orc_res[k2] for k2 in KEYS})


def rand_pairs(rnd, n, qmax, rmax, alphas):
    qs, rs = [], []
    for _ in range(n):
        al = rnd.choice(alphas)
        qs.append(bytes(rnd.choice(al) for _ in range(rnd.randint(1, qmax))))
        rs.append(bytes(rnd.choice(al) for _ in range(rnd.randint(1, rmax))))
    return qs, rs


def test_plain_known_answers(A, orc):
    for case in GOLD["sw"]:
        m, x, o, e = case["scoring"]
        ref = case["ref"].encode()
        rl = case.get("ref_len", len(ref))
        res = A.align_sequences([case["qry"].encode()], [ref], ref_len=rl,
                                scoring=A.Scoring(m, x, o, e))
        got = res.record(0)
        assert got["score"] == case["score"]
        if "start" in case:
            assert (got["start"], got["end"], got["len"]) == (case["start"], case["end"],
                                                               case["len"])
            assert got["qry_cons"] == case["qry_cons"].encode()
            assert got["ref_cons"] == case["ref_cons"].encode()


SCORINGS = [(2, -1, -10, -1), (1, -1, -1, -1), (1, 0, -1, 0), (3, -2, -2, -1),
            (2, -1, -1, -2), (1, -1, 0, 0), (2, 1, -3, -1)]


@pytest.mark.parametrize("sc", SCORINGS)
def test_plain_random_vs_oracle(A, orc, sc):
    """alignSequences drop-in: uniform strides, NUL-terminated references of
    varying length, every query byte a row (exotic bytes included)."""
    rnd = random.Random(hash(sc) & 0xFFFF)
    ql, rl = 37, 90
    alphas = [b"ACGT", b"AC", b"ACGTN", b"ACGTacgtNn", b"ACGTRY*"]
    qs = [bytes(rnd.choice(rnd.choice(alphas)) for _ in range(ql)) for _ in range(300)]
    rs = [bytes(rnd.choice(rnd.choice(alphas)) for _ in range(rnd.randint(1, rl)))
          for _ in range(300)]
    res = A.align_sequences(qs, rs, ql, rl, A.Scoring(*sc))
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], *sc, ref_len=rl), (i, sc))


@pytest.mark.parametrize("ql,rl", [(100, 300), (64, 300), (65, 128), (1, 50), (250, 450),
                                   (150, 300)])
def test_plain_shapes_vs_oracle(A, orc, ql, rl):
    rnd = random.Random(ql * 1000 + rl)
    qs, rs = [], []
    for _ in range(64):
        reg = bytes(rnd.choice(b"ACGT") for _ in range(rl))
        o = rnd.randint(0, max(0, rl - ql))
        q = bytearray(reg[o:o + ql].ljust(ql, b"A"))
        for k in range(ql):
            if rnd.random() < 0.04:
                q[k] = rnd.choice(b"ACGT")
        qs.append(bytes(q))
        rs.append(reg[:rnd.randint(max(1, rl - 40), rl)])
    res = A.align_sequences(qs, rs, ql, rl)
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], ref_len=rl), (ql, rl, i))


def test_plain_long_paths_vs_oracle(A, orc):
    """Queries of 1,000 bases with indels against 2,000-base references:
    paths past 1,024 moves, so the wave-per-pair walk keeps its moves in more
    than one register per lane (64 words each)."""
    rnd = random.Random(71)
    ql, rl = 1000, 2000
    qs, rs = [], []
    for _ in range(48):
        reg = bytes(rnd.choice(b"ACGT") for _ in range(rl))
        o = rnd.randint(0, 400)
        src = bytearray(reg[o:o + 1400])
        q = bytearray()
        i = 0
        while len(q) < ql and i < len(src):
            r = rnd.random()
            if r < 0.05:
                i += rnd.randint(1, 3)  # deletion from the query
            elif r < 0.10:
                q += bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(1, 3)))
            else:
                q.append(src[i] if rnd.random() > 0.03 else rnd.choice(b"ACGT"))
                i += 1
        qs.append(bytes(q[:ql].ljust(ql, b"A")))
        rs.append(reg)
    res = A.align_sequences(qs, rs, ql, rl)
    assert max(res.length) > 1024
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], ref_len=rl), (ql, rl, i))


def test_plain_chunks_vs_oracle(A, orc):
    """A call of many chunks (7 pairs each) through the one scratch buffer;
    every pair against the oracle, twice in a row (the second call with
    other shapes reuses the buffer)."""
    rnd = random.Random(4242)
    try:
        A.set_plain_chunk(7)
        for ql, rl, n in ((100, 300, 200), (64, 150, 61)):
            qs, rs = [], []
            for _ in range(n):
                reg = bytes(rnd.choice(b"ACGT") for _ in range(rl))
                o = rnd.randint(0, rl - ql)
                q = bytearray(reg[o:o + ql])
                for k in range(ql):
                    if rnd.random() < 0.05:
                        q[k] = rnd.choice(b"ACGTN")
                qs.append(bytes(q))
                rs.append(reg[:rnd.randint(rl // 2, rl)])
            res = A.align_sequences(qs, rs, ql, rl)
            for i in range(n):
                compare(res.record(i), orc.sw(rs[i], qs[i], ref_len=rl), (ql, rl, i))
    finally:
        A.set_plain_chunk(0)


def test_plain_exit_on_bad_score(A):
    with pytest.raises(A.__class__ if False else Exception):
        A.align_sequences([b"ACGT"], [b"ACGT"], scoring=A.Scoring(2.5, -1, -10, -1))


@pytest.mark.parametrize("coarse", [0, -1])
def test_batch_uneven_vs_oracle(A, orc, coarse):
    """Batched path, query 1..320, reference 1..480: every 16-lane row class
    (coarse 0), and the wide form that a batch this small takes by default
    (one 64-lane group per wave, two pairs per task)."""
    rnd = random.Random(11)
    qs, rs = rand_pairs(rnd, 1500, 320, 480, [b"ACGT", b"ACGTN", b"AC", b"ACGTacgtn"])
    before = A.get_coarse_rows()
    try:
        A.set_coarse_rows(coarse)
        res = A.sw_batch(qs, rs)
        st = A.batch_stats()
    finally:
        A.set_coarse_rows(before)
    if coarse == 0:
        assert 0 < st["ntasks16"] <= 1500 // 8 + 20
    else:
        assert st["ntasks16"] == 750  # the wide form: 2 pairs per task
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i]), i)


def test_batch_short_queries_mid_size(A, orc):
    """3,000 pairs whose queries all fit the 10-row class (<= 160 bases):
    one 16-lane class instead of the wide form, same results."""
    rnd = random.Random(31)
    qs, rs = rand_pairs(rnd, 3000, 160, 400, [b"ACGT", b"ACGTN"])
    res = A.sw_batch(qs, rs)
    assert A.batch_stats()["ntasks16"] <= 3000 // 8 + 20  # 8 pairs per task, not 2
    sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=8)
    assert np.array_equal(res.score, sc) and np.array_equal(res.start, st)
    assert np.array_equal(res.end, en) and np.array_equal(res.length, al)
    assert res.qry_cons == qc and res.ref_cons == rc


def test_batch_wide_form_limits(A, orc):
    """The wide form at its limits: 4,095 pairs (the largest batch that takes
    it) with the longest queries and references (320 x 480, 543 steps), and
    packed halves of very different lengths."""
    rnd = random.Random(29)
    qs, rs = [], []
    for i in range(4095):
        if i % 3 == 0:
            ql, rl = 320, 480
        elif i % 3 == 1:
            ql, rl = rnd.randint(1, 40), rnd.randint(1, 60)
        else:
            ql, rl = rnd.randint(200, 320), rnd.randint(300, 480)
        reg = bytes(rnd.choice(b"ACGT") for _ in range(rl))
        s0 = rnd.randint(0, max(0, rl - ql))
        q = bytearray(reg[s0:s0 + ql].ljust(ql, b"A"))
        for _ in range(ql // 30):
            q[rnd.randrange(ql)] = rnd.choice(b"ACGT")
        qs.append(bytes(q))
        rs.append(reg)
    res = A.sw_batch(qs, rs)
    assert A.batch_stats()["ntasks16"] == 2048
    sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=8)
    assert np.array_equal(res.score, sc) and np.array_equal(res.start, st)
    assert np.array_equal(res.end, en) and np.array_equal(res.length, al)
    assert res.qry_cons == qc and res.ref_cons == rc


def test_batch_c3_shapes_vs_oracle(A, orc):
    from swift_amd import synth
    qs, rs = synth.c3_pairs(3000, seed=5)
    res = A.sw_batch(qs, rs)
    sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=8)
    assert np.array_equal(res.score, sc) and np.array_equal(res.start, st)
    assert np.array_equal(res.end, en) and np.array_equal(res.length, al)
    assert res.qry_cons == qc and res.ref_cons == rc


def test_batch_exotic_bytes_route_int32(A, orc):
    """References with bytes outside ACGTNacgtn are flagged by the int16
    kernel and recomputed with byte semantics in the int32 pass."""
    rnd = random.Random(3)
    qs, rs = rand_pairs(rnd, 400, 120, 300, [b"ACGTRY*", b"ACGTacgtNn#"])
    res = A.sw_batch(qs, rs)
    assert A.batch_stats()["nrerun"] > 0
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i]), i)


def test_batch_int16_overflow_rerun(A, orc):
    """match = 200: long exact matches pass the int16 range, the int16 kernel
    flags them and the int32 re-run recomputes them exactly."""
    rnd = random.Random(8)
    qs, rs = [], []
    for _ in range(200):
        L = rnd.randint(20, 250)
        reg = bytes(rnd.choice(b"ACGT") for _ in range(L + rnd.randint(0, 100)))
        o = rnd.randint(0, len(reg) - L)
        qs.append(reg[o:o + L])
        rs.append(reg)
    sc = A.Scoring(200, -1, -10, -1)
    res = A.sw_batch(qs, rs, sc)
    st = A.batch_stats()
    assert st["nrerun"] > 0
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], 200, -1, -10, -1), i)


@pytest.mark.parametrize("match", [7, 8, 9])
def test_batch_int16_key_range(A, orc, match):
    """Best scores up to about 2,900: the int16 lanes keep best cells as keys
    32H + code (unsigned, exact while H <= 2047, past 32767 from H = 1024
    on) and flag a pair whose best reaches 2048 - match for the int32 pass.
    Exact matches of 120-320 bases put the bests on both sides of that
    limit, mixed with ordinary pairs in the same tasks."""
    rnd = random.Random(match)
    qs, rs = [], []
    for i in range(240):
        L = rnd.randint(120, 320) if i % 2 == 0 else rnd.randint(30, 200)
        reg = bytes(rnd.choice(b"ACGT") for _ in range(L + rnd.randint(0, 150)))
        o = rnd.randint(0, len(reg) - L)
        q = bytearray(reg[o:o + L])
        if i % 4 == 1:
            for _ in range(3):
                q[rnd.randrange(L)] = rnd.choice(b"ACGT")
        qs.append(bytes(q))
        rs.append(reg)
    sc = (match, -1, -10, -1)
    res = A.sw_batch(qs, rs, A.Scoring(*sc))
    want = [orc.sw(rs[i], qs[i], *sc) for i in range(len(qs))]
    lim = 2048 - match
    assert any(1024 < w["score"] < lim for w in want)   # keys past 32767, kept in int16
    assert any(w["score"] >= lim for w in want)         # flagged for the int32 pass
    assert A.batch_stats()["nrerun"] > 0
    for i in range(len(qs)):
        compare(res.record(i), want[i], (i, sc))


@pytest.mark.parametrize("sc", SCORINGS)
def test_batch_scorings_vs_oracle(A, orc, sc):
    rnd = random.Random(sum(sc) + 100)
    qs, rs = rand_pairs(rnd, 300, 200, 400, [b"ACGT", b"AC", b"ACGTN"])
    res = A.sw_batch(qs, rs, A.Scoring(*sc))
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], *sc), (i, sc))


@pytest.mark.parametrize("sc", [(2, -1, -10, -1), (1, -1, -1, -1), (1, 0, -1, 0)])
def test_batch_padding_columns(A, orc, sc):
    """Queries of N (which matches every reference base) over references a
    few columns apart in length, so each task sweeps columns past the end of
    most of its pairs: the padding columns must never score as a match, and
    the best cell must stay inside the pair whether or not the steps are
    masked (masking is skipped when mismatch and gaps all lose score)."""
    rnd = random.Random(sum(sc) + 7)
    qs, rs = [], []
    for i in range(400):
        ql = rnd.randint(1, 64)
        rl = rnd.randint(1, 60)
        if i % 3 == 0:
            q = b"N" * ql
        else:
            q = bytes(rnd.choice(b"ACGTNNNN") for _ in range(ql))
        qs.append(q)
        rs.append(bytes(rnd.choice(b"ACGTN") for _ in range(rl)))
    res = A.sw_batch(qs, rs, A.Scoring(*sc))
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i], *sc), (i, sc))


def test_batch_edge_sizes(A, orc):
    qs = [b"A", b"ACGT" * 80, b"T" * 16, b"G" * 17, b"ACGTN", b"CCCCCCCCCC"]
    rs = [b"A", b"ACGT" * 120, b"T", b"G" * 480, b"N", b"GGGGGGGGGG"]
    res = A.sw_batch(qs, rs)
    for i in range(len(qs)):
        compare(res.record(i), orc.sw(rs[i], qs[i]), i)


def test_batch_rejects_oversized(A):
    from swift_amd._lib import SwiftError
    with pytest.raises(SwiftError):
        A.sw_batch([b"ACGT" * 90], [b"ACGT" * 121])   # query 360 > 320, region 484 > 480


def test_batch_repeatable(A):
    from swift_amd import synth
    qs, rs = synth.c3_pairs(2000, seed=9)
    b = A.PairBatch(qs, rs)
    b.run()
    r1 = b.results()
    b.run()
    r2 = b.results()
    assert np.array_equal(r1.score, r2.score) and r1.qry_cons == r2.qry_cons


@pytest.mark.parametrize("deferred,coarse", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_batch_walk_modes(A, orc, deferred, coarse):
    """Fused and deferred walk-back, exact or coarse row classes, all give the
    oracle's results (C3 shapes, an int16-overflow re-run and exotic bytes in
    the same call)."""
    from swift_amd import synth
    before = A.get_tuning()
    try:
        A.set_tuning(before["waves_per_cu"], True, deferred)
        A.set_coarse_rows(coarse)
        qs, rs = synth.c3_pairs(1500, seed=17)
        rnd = random.Random(17)
        q2, r2 = rand_pairs(rnd, 200, 120, 300, [b"ACGTRY*", b"ACGTacgtNn#"])
        qs, rs = qs + q2, rs + r2
        res = A.sw_batch(qs, rs)
        sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=8)
        assert np.array_equal(res.score, sc) and np.array_equal(res.start, st)
        assert np.array_equal(res.end, en) and np.array_equal(res.length, al)
        assert res.qry_cons == qc and res.ref_cons == rc
        big = A.Scoring(200, -1, -10, -1)
        reg = [bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(60, 300))) for _ in range(50)]
        qo = [g[5:5 + rnd.randint(20, 50)] for g in reg]
        res = A.sw_batch(qo, reg, big)
        assert A.batch_stats()["nrerun"] > 0
        for i in range(len(qo)):
            compare(res.record(i), orc.sw(reg[i], qo[i], 200, -1, -10, -1), i)
    finally:
        A.set_tuning(before["waves_per_cu"], before["allow_int16"], before["walk_deferred"])
        A.set_coarse_rows(before["coarse_rows"])


def test_autotune_keeps_results(A, orc):
    """The autotuner only changes speed: after it picks a setting, results
    are still those of the oracle."""
    from swift_amd import synth
    before = A.get_tuning()
    try:
        qs, rs = synth.c3_pairs(2000, seed=23)
        b = A.PairBatch(qs, rs)
        t = b.autotune(reps=1)
        assert len(t["table"]) >= 10 and t["ms"] > 0
        assert A.get_tuning()["walk_deferred"] == t["walk_deferred"]
        assert A.get_coarse_rows() == t["coarse_rows"]
        assert any(r["coarse_rows"] for r in t["table"])
        b.run()
        res = b.results()
        sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=8)
        assert np.array_equal(res.score, sc) and np.array_equal(res.start, st)
        assert res.qry_cons == qc and res.ref_cons == rc
    finally:
        A.set_tuning(before["waves_per_cu"], before["allow_int16"], before["walk_deferred"])
        A.set_coarse_rows(before["coarse_rows"])
