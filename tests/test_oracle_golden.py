"""The oracle against the reference test suite's known answers and against
what the mathematics fixes (closed forms, brute force on tiny inputs,
invariants of the printed alignment)."""
import itertools
import json
import os
import random

import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))


def parse_fasta(text):
    seqs, cur = [], None
    for line in text.split("\n"):
        if line.startswith(">"):
            cur = []
            seqs.append(cur)
        elif line and cur is not None:
            cur.append(line)
    return ["".join(c).encode() for c in seqs]


@pytest.mark.parametrize("case", GOLD["sw"], ids=lambda c: c["source"][:40])
def test_sw_known_answers(orc, case):
    m, x, o, e = case["scoring"]
    ref = case["ref"].encode()
    r = orc.sw(ref, case["qry"].encode(), m, x, o, e, ref_len=case.get("ref_len"))
    assert r["score"] == case["score"]
    stride = case.get("cell_max_stride", len(ref) + 1)
    assert r["max_row"] * stride + r["max_col"] == case["cell_max"]
    for k in ("start", "end", "len"):
        if k in case:
            assert r[k] == case[k]
    if "qry_cons" in case:
        assert r["qry_cons"] == case["qry_cons"].encode()
        assert r["ref_cons"] == case["ref_cons"].encode()


def test_is_match_rule(orc):
    for a, b, want in GOLD["is_match"]["cases"]:
        r = orc.sw(b.encode(), a.encode(), 2, -1, -10, -1)
        assert (r["score"] == 2) == want


@pytest.mark.parametrize("case", GOLD["revcomp"], ids=lambda c: c["seq"])
def test_revcomp_known(orc, case):
    assert orc.revcomp(case["seq"].encode()) == case["rc"].encode()


def test_hash_closed_form(orc):
    rnd = random.Random(1)
    code = {"A": 0, "C": 1, "G": 2, "T": 3, "a": 0, "c": 1, "g": 2, "t": 3}
    for _ in range(500):
        k = rnd.randint(1, 15)
        s = "".join(rnd.choice("ACGTacgtNnRX") for _ in range(k))
        want = sum(code.get(ch, 0) * 4 ** i for i, ch in enumerate(s))
        assert orc.get_hash(s.encode(), k) == want
    assert orc.get_hash(b"ACG", 3) == 36  # worked example, src/preprocess.cu:79-86


@pytest.mark.parametrize("case", GOLD["lookup_table5_create"], ids=lambda c: c["source"][:44])
def test_lookup_table5_create_known(orc, case):
    ix = orc.Index(parse_fasta(case["fasta"]), case["k"], case["thr"])
    a = ix.arrays()
    assert len(a["start"]) == 4 ** case["k"]
    for h, v in case["lookup"].items():
        assert a["start"][int(h)] == v
    assert list(a["refIdx"][:len(case["refIdx"])]) == case["refIdx"]
    assert list(a["refPos"][:len(case["refPos"])]) == case["refPos"]
    for h, v in case["count"].items():
        assert a["count"][int(h)] == v


def brute_index(seqs, k, thr, orc):
    occ = {}
    for si, s in enumerate(seqs):
        for p in range(0, len(s) - k + 1, k):
            occ.setdefault(orc.get_hash(s[p:p + k], k), []).append((si, p))
    return {h: v for h, v in occ.items() if len(v) <= thr}, occ


def test_index_brute_force(orc):
    rnd = random.Random(7)
    for trial in range(30):
        k = rnd.choice([2, 3, 4, 5])
        thr = rnd.choice([1, 2, 3, 5, 100])
        seqs = [bytes(rnd.choice(b"ACGTN") for _ in range(rnd.randint(0, 60)))
                for _ in range(rnd.randint(1, 4))]
        kept, occ = brute_index(seqs, k, thr, orc)
        ix = orc.Index(seqs, k, thr)
        a = ix.arrays()
        assert ix.total == sum(len(v) for v in kept.values())
        nxt = 1
        for h in range(4 ** k):
            c = len(occ.get(h, []))
            assert a["count"][h] == c
            if h in kept:
                assert a["start"][h] == nxt and a["actual"][h] == c
                got = [(int(a["refIdx"][nxt + j]), int(a["refPos"][nxt + j])) for j in range(c)]
                assert got == kept[h][::-1]  # reverse file order in the bucket
                nxt += c
            else:
                assert a["start"][h] == 0 and a["actual"][h] == 0


def _check_choice(got, spec):
    if "any_of" in spec:
        assert [list(g) for g in got] in spec["any_of"]
    else:
        assert len(got) == spec["choose"]
        assert all(list(g) in spec["from"] for g in got)
        assert len(set(got)) == len(got)


@pytest.mark.parametrize("case", GOLD["lookup_table5_map"], ids=lambda c: c["source"][:44])
def test_lookup_table5_map_known(orc, case):
    ix = orc.Index(parse_fasta(case["fasta"]), case["k"], case["thr"])
    q = case["qry"].encode()
    for mode in (orc.TIE_FIRST, orc.TIE_GLIBC):
        rng = orc.Rng(99)
        fwd, _, _ = ix.map_query(q, case["max_hits"], mode, rng)
        rev, _, _ = ix.map_query(orc.revcomp(q), case["max_hits"], mode, rng)
        _check_choice(fwd, case["fwd"])
        _check_choice(rev, case["rev"])


@pytest.mark.parametrize("case", GOLD["map_hits6"], ids=lambda c: c["source"][:40])
def test_map_hits6_known(orc, case):
    hits = [tuple(h) for h in case["hits"]]
    best, _, srt = orc.best_hits(hits, case["max_hits"], orc.TIE_GLIBC, orc.Rng(3))
    if "sorted" in case:
        assert [list(s) for s in srt] == case["sorted"]
    if "best" in case:
        assert [list(b) for b in best] == case["best"]
    if "choose" in case:
        _check_choice(best, case)


def brute_best(hits, max_hits):
    clusters = {}
    for ri, sh, rp in hits:
        clusters.setdefault((ri, sh), []).append(rp)
    if not clusters:
        return [], 0
    big = max(len(v) for v in clusters.values())
    keys = sorted(k for k, v in clusters.items() if len(v) == big)
    return [(k[0], k[1], min(clusters[k])) for k in keys], len(keys)


def test_best_hits_brute_force(orc):
    rnd = random.Random(11)
    for _ in range(400):
        n = rnd.randint(0, 40)
        hits = list({(rnd.randint(-1, 3), rnd.randint(-5, 5), rnd.randint(0, 40))
                     for _ in range(n)})
        rnd.shuffle(hits)
        mh = rnd.randint(1, 5)
        cand, nbig = brute_best(hits, mh)
        best, nb, srt = orc.best_hits(hits, mh, orc.TIE_FIRST)
        assert srt == sorted(hits)
        assert nb == nbig
        assert best == cand[:mh]
        gbest, _, _ = orc.best_hits(hits, mh, orc.TIE_GLIBC, orc.Rng(rnd.randint(1, 1000)))
        assert len(gbest) == min(mh, nbig) and all(g in cand for g in gbest)
        assert gbest == sorted(gbest, key=cand.index)


@pytest.mark.parametrize("case", GOLD["search_queries3"], ids=lambda c: c["source"][:40])
def test_search_queries3_lines(orc, case):
    """searchQueries3 match-file lines: name, strand, refIdx, shift, refPos,
    nameOffset, seqOffset (src/search.cu:2236-2238)."""
    seqs = parse_fasta(case["fasta"])
    ix = orc.Index(seqs, case["k"], 1000)
    name, seq = case["reads"][0]
    lines = {0: [], 1: []}
    for strand in (0, 1):
        q = seq.encode() if strand == 0 else orc.revcomp(seq.encode())
        hits, _, _ = ix.map_query(q, case["max_hits"], orc.TIE_FIRST)
        for ri, sh, rp in hits:
            lines[strand].append(f"{name}\t{strand}\t{ri}\t{sh}\t{rp}\t0\t{len(name) + 2}")
    if "lines" in case:
        assert lines[0] + lines[1] == case["lines"]
    else:
        assert set(lines[0]) <= set(case["fwd_set"]) and set(lines[1]) <= set(case["rev_set"])
        assert len(lines[0]) + len(lines[1]) == case["num_matches"]


@pytest.mark.parametrize("case", GOLD["choose_best"], ids=lambda c: c["source"][:40])
def test_choose_best_known(orc, case):
    for mode, seed in ((orc.TIE_FIRST, 1), (orc.TIE_GLIBC, 5), (orc.TIE_GLIBC, 6)):
        keep, _ = orc.choose_best(case["groups"], case["scores"], mode,
                                  orc.Rng(seed) if mode else None)
        assert list(keep) in case["any_of"]


def rescore(ref_cons, qry_cons, m, x, o, e):
    """Score recomputed from the printed alignment (the invariant of
    test/testAlignOutput.cu:299-361): a gap column extends when the previous
    column is the same kind of gap."""
    s, prev = 0, None
    for a, b in zip(ref_cons[::-1], qry_cons[::-1]):
        if a == ord("-"):
            s += e if prev == "D" else o
            prev = "D"
        elif b == ord("-"):
            s += e if prev == "I" else o
            prev = "I"
        else:
            A = a - 32 if 97 <= a <= 122 else a
            s += m if (b == A or b == ord("N") or A == ord("N")) else x
            prev = "M"
    return s


def exhaustive_best(ref, qry, m, x, o, e):
    """Optimum over every local alignment path of tiny sequences (paths start
    with an aligned pair and may stop anywhere), by a memoised recursion over
    (position, previous column kind) -- independent of the reference's
    argmax/tie machinery."""
    from functools import lru_cache
    R = bytes(c - 32 if 97 <= c <= 122 else c for c in ref)

    def sub(i, j):
        return m if (qry[i] == R[j] or qry[i] == ord("N") or R[j] == ord("N")) else x

    @lru_cache(maxsize=None)
    def ext(i, j, prev):
        best = 0
        if i < len(qry) and j < len(R):
            best = max(best, sub(i, j) + ext(i + 1, j + 1, "M"))
        if i < len(qry):
            best = max(best, (e if prev == "D" else o) + ext(i + 1, j, "D"))
        if j < len(R):
            best = max(best, (e if prev == "I" else o) + ext(i, j + 1, "I"))
        return best

    return max([0] + [sub(i, j) + ext(i + 1, j + 1, "M")
                      for i, j in itertools.product(range(len(qry)), range(len(R)))])


def test_sw_invariants_and_brute_force(orc):
    """The oracle's score equals an exhaustive search on tiny inputs (for
    scorings where the greedy direction rule is optimal: gapExt == gapOpen),
    and every alignment it prints rescores to its score and spans the
    reference window it reports."""
    rnd = random.Random(3)
    for _ in range(300):
        q = bytes(rnd.choice(b"ACGN") for _ in range(rnd.randint(1, 6)))
        r = bytes(rnd.choice(b"ACGTn") for _ in range(rnd.randint(1, 7)))
        m, x, g = rnd.choice([(2, -1, -2), (1, -1, -1), (3, -2, -1)])
        res = orc.sw(r, q, m, x, g, g)
        assert res["score"] == exhaustive_best(r, q, m, x, g, g)
    for _ in range(300):
        q = bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(1, 40)))
        r = bytes(rnd.choice(b"ACGTacgt") for _ in range(rnd.randint(1, 80)))
        sc = rnd.choice([(2, -1, -10, -1), (1, -1, -2, -1), (2, -3, -3, -1)])
        res = orc.sw(r, q, *sc)
        if res["len"] == 0:
            assert res["score"] == 0
            continue
        assert rescore(res["ref_cons"], res["qry_cons"], *sc) == res["score"]
        span = bytes(c for c in res["ref_cons"][::-1] if c != ord("-"))
        assert span == r[res["start"] - 1:res["end"]]
        assert bytes(c for c in res["qry_cons"][::-1] if c != ord("-")) in q


def test_region_rules(orc):
    seq = b"NNAC" + b"G" * 400 + b"TTnn"
    out, rd = orc.region(seq, 50)
    assert rd == 0 and out == seq[2:300]          # leading N stripped, no coordinate fix
    out, rd = orc.region(seq, 300)
    assert rd == 200 and out == seq[200:406]      # trailing n stripped, clipped at the end
    out, rd = orc.region(b"N" * 50, 10)
    assert out == b"" and rd == 0
