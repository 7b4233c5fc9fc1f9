"""The oracle restatement against the reference's own CPU functions, compiled
from the reference sources into oracle/_ref (skipped where the reference
tree is absent)."""
import ctypes as C
import os
import random

import numpy as np
import pytest


def ref_sw(R, ref, qry, m, x, o, e):
    n = len(ref) + len(qry) + 2
    rc, qc = C.create_string_buffer(n), C.create_string_buffer(n)
    s, st, en, al = C.c_float(), C.c_int(), C.c_int(), C.c_int()
    R.ref_sw_one(ref, len(ref), qry, len(qry), C.c_float(m), C.c_float(x), C.c_float(o),
                 C.c_float(e), C.byref(s), C.byref(st), C.byref(en), C.byref(al), rc, qc)
    return dict(score=int(s.value), start=st.value, end=en.value, len=al.value,
                ref_cons=rc.raw[:al.value], qry_cons=qc.raw[:al.value])


SCORINGS = [(2, -1, -10, -1), (1, -1, -1, -1), (1, 0, -1, 0), (3, -2, -2, -1),
            (2, -1, -1, -2), (1, -1, 0, 0), (200, -1, -10, -1), (2, 1, -3, -1)]


def test_sw_random_vs_reference(orc, refbridge):
    rnd = random.Random(2024)
    for _ in range(2500):
        alpha = rnd.choice([b"ACGT", b"AC", b"ACGTN", b"ACGTacgtNn", b"AG", b"ACGTRY*"])
        q = bytes(rnd.choice(alpha) for _ in range(rnd.randint(1, 40)))
        r = bytes(rnd.choice(alpha) for _ in range(rnd.randint(1, 90)))
        sc = rnd.choice(SCORINGS)
        a, b = orc.sw(r, q, *sc), ref_sw(refbridge, r, q, *sc)
        assert all(a[k] == b[k] for k in b), (q, r, sc, a, b)


def test_sw_realistic_shapes_vs_reference(orc, refbridge):
    """100 x 300 and 250 x 450 pairs with the embedded-read structure of the
    phase-2 windows (zero-cell pointers matter here)."""
    rnd = random.Random(77)
    for trial in range(40):
        ql = rnd.choice([100, 150, 250, 36])
        d = rnd.randint(20, 200)
        refw = bytearray(rnd.choice(b"ACGT") for _ in range(ql + d))
        off = rnd.randint(0, d)
        q = bytearray(refw[off:off + ql])
        for i in range(ql):
            if rnd.random() < 0.03:
                q[i] = rnd.choice(b"ACGT")
        a = orc.sw(bytes(refw), bytes(q))
        b = ref_sw(refbridge, bytes(refw), bytes(q), 2, -1, -10, -1)
        assert all(a[k] == b[k] for k in b)


def test_glibc_rand_replay(orc, refbridge):
    for seed in (0, 1, 42, 20261201, 2 ** 31 + 5, 2 ** 32 - 1):
        refbridge.ref_srand(C.c_uint(seed))
        g = orc.Rng(seed)
        assert [g.next() for _ in range(2000)] == [refbridge.ref_rand() for _ in range(2000)]


def test_hash_and_revcomp_vs_reference(orc, refbridge):
    rnd = random.Random(5)
    for _ in range(2000):
        k = rnd.randint(1, 15)
        s = bytes(rnd.choice(b"ACGTacgtNnX") for _ in range(k))
        assert orc.get_hash(s, k) == refbridge.ref_hash(s, k)
        L = rnd.randint(0, 50)
        t = bytes(rnd.choice(b"ACGTacgtN") for _ in range(L))
        out = C.create_string_buffer(L + 1)
        refbridge.ref_revcomp(t, L, out)
        assert orc.revcomp(t) == out.raw[:L]


def test_map_hits6_glibc_replay_vs_reference(orc, refbridge):
    """Single-threaded replay with the same srand seed: same in-place sort,
    same clusters, same rand()-drawn picks."""
    rnd = random.Random(9)
    seed = 1234
    refbridge.ref_srand(C.c_uint(seed))
    g = orc.Rng(seed)
    for _ in range(500):
        n = rnd.randint(0, 60)
        hits = [(rnd.randint(0, 4), rnd.randint(-3, 6), rnd.randint(0, 50)) for _ in range(n)]
        hits = list(dict.fromkeys(hits))
        mh = rnd.randint(1, 5)
        best, _, srt = orc.best_hits(hits, mh, orc.TIE_GLIBC, g)
        n = len(hits)
        ri = (C.c_char * max(n, 1))(*[bytes([h[0]]) for h in hits])
        sh = (C.c_int * max(n, 1))(*[h[1] for h in hits])
        rp = (C.c_int * max(n, 1))(*[h[2] for h in hits])
        oi, osh, op = (C.c_char * 8)(), (C.c_int * 8)(), (C.c_int * 8)()
        k = refbridge.ref_best_hits(mh, ri, sh, rp, n, oi, osh, op)
        assert best == [(oi[i][0], osh[i], op[i]) for i in range(k)]
        assert srt == [(ri[i][0], sh[i], rp[i]) for i in range(n)]


def test_region_vs_reference(orc, refbridge, tmp_path):
    """Region fetch through the reference's file-offset map and getRefSeq2,
    for several line widths and N runs at window edges."""
    rnd = random.Random(13)
    for width in (50, 60, 1000000):
        seqs = []
        for s in range(3):
            L = rnd.randint(400, 1500)
            b = bytearray(rnd.choice(b"ACGT") for _ in range(L))
            for _ in range(4):
                p = rnd.randint(0, L - 10)
                b[p:p + rnd.randint(1, 8)] = b"N" * 8
            seqs.append(bytes(b[:L]))
        path = str(tmp_path / f"ref{width}.fa")
        with open(path, "wb") as f:
            for i, s in enumerate(seqs):
                f.write(b">chr%d\n" % i)
                for p in range(0, len(s), width):
                    f.write(s[p:p + width] + b"\n")
        for _ in range(120):
            si = rnd.randrange(3)
            pos = rnd.randrange(len(seqs[si]))
            want, wrd = orc.region(seqs[si], pos)
            out = C.create_string_buffer(400)
            rd = C.c_int()
            n = refbridge.ref_region(path.encode(), si, pos, out, C.byref(rd))
            got = out.raw[:n]
            # leading strip of two or more N reads stale memory in the
            # reference (input.cu:399); compare only the defined cases
            raw = seqs[si][wrd:wrd + 300]
            m = len(raw) - len(raw.lstrip(b"Nn"))
            if m >= 2:
                assert got.startswith(want) or want.startswith(got[:len(want)])
                continue
            assert (got, rd.value) == (want, wrd), (si, pos, width)


def test_output_format_vs_reference(orc, refbridge):
    rnd = random.Random(4)
    for _ in range(200):
        q = bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(5, 40)))
        r = bytes(rnd.choice(b"ACGTN") for _ in range(rnd.randint(5, 60)))
        res = orc.sw(r, q)
        want = orc.format_record("read7*", "chr2", res["score"], 321, res["start"],
                                 res["end"], res["qry_cons"], res["ref_cons"])
        out = C.create_string_buffer(4096)
        n = refbridge.ref_format_record(b"read7*", b"chr2", C.c_float(res["score"]), 321,
                                        res["start"], res["end"], res["len"], res["qry_cons"],
                                        res["ref_cons"], len(q) + len(r) + 2, out, 4096)
        assert out.raw[:n].decode() == want
        n = refbridge.ref_format_cigar(res["qry_cons"], res["ref_cons"], res["len"], out, 4096)
        assert out.raw[:n].decode() == orc.cigar(res["qry_cons"], res["ref_cons"])
