"""Phase-1 parity on the GPU: the device index, the device mapper, the
reference-named drop-ins and the phase-1 -> phase-2 pipeline against the
oracle (tie rule "first") and the reference test suite's known answers."""
import json
import os
import random

import numpy as np
import pytest

from test_oracle_golden import _check_choice, parse_fasta

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))


@pytest.fixture(scope="module")
def S():
    from swift_amd import search
    return search


def assert_index_equal(got, want):
    for key in ("start", "count", "actual", "refIdx", "refPos"):
        assert np.array_equal(got[key], want[key]), key


def rand_genome(rnd, nseq, lo, hi, alphabet=b"ACGT", repeats=0):
    seqs = []
    for _ in range(nseq):
        s = bytearray(rnd.choice(alphabet) for _ in range(rnd.randint(lo, hi)))
        for _ in range(repeats):  # tandem arrays make large buckets and ties
            unit = bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(2, 30)))
            span = min(len(s), len(unit) * rnd.randint(3, 60))
            p = rnd.randint(0, len(s) - span)
            s[p:p + span] = (unit * (span // len(unit) + 1))[:span]
        seqs.append(bytes(s))
    return seqs


# ------------------------------------------------------------------ index
@pytest.mark.parametrize("k,thr", [(4, 3), (8, 4), (12, 1000), (6, 40)])
def test_index_vs_oracle(S, orc, k, thr):
    rnd = random.Random(k * 100 + thr)
    seqs = rand_genome(rnd, 5, 0, 30000, b"ACGTACGTACGTNnacgt", repeats=6)
    want = orc.Index(seqs, k, thr)
    got = S.Index.from_sequences(seqs, k, thr)
    assert got.info()["total"] == want.total
    assert_index_equal(got.arrays(), want.arrays())


def test_index_large_buckets(S, orc):
    """Buckets past the register and wave sorts (thr 5000, tandem repeats)."""
    rnd = random.Random(5)
    unit = b"ACGTTGCAAC"
    seqs = [unit * 3000 + bytes(rnd.choice(b"ACGT") for _ in range(5000)), unit * 1500]
    want = orc.Index(seqs, 10, 5000)
    got = S.Index.from_sequences(seqs, 10, 5000)
    assert got.info()["max_bucket"] > 1024
    assert_index_equal(got.arrays(), want.arrays())


@pytest.mark.parametrize("case", GOLD["lookup_table5_create"], ids=lambda c: c["source"][:44])
def test_index_fasta_known(S, orc, case, tmp_path):
    """lookupTable5Create known answers through the FASTA reader."""
    p = tmp_path / "ref.fa"
    p.write_text(case["fasta"])
    a = S.Index.from_fasta(p, case["k"], case["thr"]).arrays()
    for h, v in case["lookup"].items():
        assert a["start"][int(h)] == v
    assert list(a["refIdx"][:len(case["refIdx"])]) == case["refIdx"]
    assert list(a["refPos"][:len(case["refPos"])]) == case["refPos"]
    for h, v in case["count"].items():
        assert a["count"][int(h)] == v
    assert_index_equal(a, orc.Index(parse_fasta(case["fasta"]), case["k"], case["thr"]).arrays())


def test_index_fasta_line_handling(S, orc, tmp_path):
    """Any line width, blank lines and a missing final newline give the index
    of the concatenated lines."""
    from swift_amd import synth
    rnd = random.Random(2)
    seqs = rand_genome(rnd, 4, 100, 3000)
    want = orc.Index(seqs, 8, 20).arrays()
    for width in (7, 50, 61, 100000):
        text = synth.to_fasta(seqs, width=width)
        text = text.replace(b"\n", b"\n\n", 3).rstrip(b"\n")
        p = tmp_path / f"w{width}.fa"
        p.write_bytes(text)
        ix = S.Index.from_fasta(p, 8, 20)
        assert_index_equal(ix.arrays(), want)
        assert [ix.sequence(s)[0] for s in range(len(seqs))] == [len(s) for s in seqs]


def test_lookup_table5_create_dropin(S, orc, tmp_path):
    case = GOLD["lookup_table5_create"][1]
    p = tmp_path / "ref.fa"
    p.write_text(case["fasta"])
    tot = S.lookup_table5_create(p, case["k"], 3, case["thr"])
    assert tot == orc.Index(parse_fasta(case["fasta"]), case["k"], case["thr"]).total
    S.lookup_table5_delete()


# ------------------------------------------------------------------ mapping
def check_strands(ix_orc, reads, got, max_hits, orc):
    for r, q in enumerate(reads):
        for strand in (0, 1):
            s = 2 * r + strand
            seq = q if strand == 0 else orc.revcomp(q)
            want, nbig, nh = ix_orc.map_query(seq, max_hits, orc.TIE_FIRST)
            assert got.hits(s) == want, (r, strand)
            assert got.nbiggest[s] == nbig and got.nhits[s] == nh, (r, strand)


@pytest.mark.parametrize("max_hits", [1, 3, 10])
def test_map_reads_vs_oracle(S, orc, max_hits):
    from swift_amd import synth
    rnd = random.Random(max_hits)
    seqs = rand_genome(rnd, 6, 2000, 40000, repeats=10)
    g = np.random.default_rng(max_hits)
    reads, _ = synth.sample_reads(g, [np.frombuffer(s, np.uint8) for s in seqs], 400, 100,
                                  0.01, 0.001)
    reads += [bytes(rnd.choice(b"ACGTN") for _ in range(rnd.randint(0, 150))) for _ in range(100)]
    ix = S.Index.from_sequences(seqs, 12, 1000)
    got = ix.map_reads(reads, max_hits)
    check_strands(orc.Index(seqs, 12, 1000), reads, got, max_hits, orc)


@pytest.mark.parametrize("mode", [-1, 1])
def test_map_reads_overflow_strands(S, orc, mode):
    """Strands past the wave's LDS sort (256 hits) take the hash-cluster
    tier up to 6144 hits, the workgroup LDS sort past that, and the global
    scratch path past 16384 hits; every tier in one call.  Mode 1 starts
    every strand in the hash tier: for reads this short, its filter stage
    first (up to 6144 hits, exact counts for repeated cluster ids only), the
    full table for strands with more repeated ids than that stage holds."""
    rnd = random.Random(4)
    unit = b"ACGTTGCAAC"
    seqs = [unit * 2000, bytes(rnd.choice(b"ACGT") for _ in range(20000)) + unit * 300]
    tail = bytes(rnd.choice(b"ACGT") for _ in range(90))
    reads = [unit * 20, (unit * 30)[3:250], seqs[1][100:300], unit[:5] * 40, unit + tail,
             unit * 2 + tail, unit * 5 + tail, unit * 8 + tail]
    before = S.get_map_mode()
    try:
        S.set_map_mode(mode)
        ix = S.Index.from_sequences(seqs, 10, 100000)
        got = ix.map_reads(reads, 3)
    finally:
        S.set_map_mode(before)
    h = got.nhits
    for lo, hi in ((0, 256), (256, 3072), (3072, 6144), (6144, 16384), (16384, 1 << 30)):
        assert ((h > lo) & (h <= hi)).any(), (lo, hi, h)
    check_strands(orc.Index(seqs, 10, 100000), reads, got, 3, orc)


def test_map_hash_table_sizes(S, orc):
    """Hash-first mapping of short reads passes strands on: up to 6144 hits
    the filter stage (strands with more repeated cluster ids than its table
    holds go on to the full hash table), past 6144 the workgroup sort.
    Strands of 15, 1600, 3200, 4800, 5000 and 6600 hits (counts from the
    oracle) cover each hand-over."""
    rnd = random.Random(11)
    ua, ub = b"ACGTTGCAAC", b"TTGACCGTAG"
    seqs = [ua * 5000, ub * 1600, bytes(rnd.choice(b"ACGT") for _ in range(20000))]
    tail = bytes(rnd.choice(b"ACGT") for _ in range(90))
    reads = [ua + tail, ub * 2 + tail, ub + tail, ua + ub + tail, seqs[2][500:650],
             ub * 3 + tail]
    before = S.get_map_mode()
    try:
        S.set_map_mode(1)
        ix = S.Index.from_sequences(seqs, 10, 100000)
        got = ix.map_reads(reads, 3)
    finally:
        S.set_map_mode(before)
    assert sorted(got.nhits[0::2]) == [15, 1600, 3200, 4800, 5000, 6600]
    check_strands(orc.Index(seqs, 10, 100000), reads, got, 3, orc)


def filter_tier_reads():
    """Workloads of the hash-first filter tier's paths (all counted by the
    oracle on the CPU): (a) k=6 over 300 kbp, about 1,700 hits per strand,
    a true locus of about 25 hits among singleton clusters; (b) random reads
    on a k=8 index of 1 Mbp, a few hundred hits that are all clusters of one;
    (c) a 24-base motif at 100 positions on the k=12 tuple grid: 100
    clusters tied at two hits, more than one wave lists; (d) reads on a k=6
    index of 1 Mbp, about 40 entries per bucket: 100 bases (3,800 hits per
    strand), 250 bases (about 9,800 hits: the 1024-thread filter stage for
    strands past 6,144) and 480 bases (about 19,000: past that stage's
    18,432)."""
    rnd = random.Random(123)
    out = []
    g6 = bytes(rnd.choice(b"ACGT") for _ in range(300_000))
    reads = [g6[p:p + 150] for p in (rnd.randrange(0, 299_000) for _ in range(40))]
    reads += [bytes(rnd.choice(b"ACGT") for _ in range(150)) for _ in range(10)]
    out.append(([g6], 6, reads))
    g8 = bytes(rnd.choice(b"ACGT") for _ in range(1_000_000))
    out.append(([g8], 8, [bytes(rnd.choice(b"ACGT") for _ in range(120)) for _ in range(40)]))
    motif = bytes(rnd.choice(b"ACGT") for _ in range(24))
    gm = b"".join(motif + bytes(rnd.choice(b"ACGT") for _ in range(120)) for _ in range(100))
    reads = [bytes(rnd.choice(b"ACGT") for _ in range(30)) + motif +
             bytes(rnd.choice(b"ACGT") for _ in range(40)) for _ in range(6)]
    out.append(([gm], 12, reads))
    g6 = bytes(rnd.choice(b"ACGT") for _ in range(1_000_000))
    reads = [g6[p:p + 100] for p in (rnd.randrange(0, 999_000) for _ in range(30))]
    reads += [bytes(rnd.choice(b"ACGT") for _ in range(100)) for _ in range(5)]
    # about 9,800 hits: past the first filter's 6,144, into the 1024-thread
    # filter stage; about 19,000: past its 18,432, on to the sorting tiers
    reads += [g6[p:p + 250] for p in (rnd.randrange(0, 999_000) for _ in range(12))]
    reads += [g6[p:p + 480] for p in (rnd.randrange(0, 999_000) for _ in range(4))]
    out.append(([g6], 6, reads))
    return out


def test_map_filter_tier_paths(S, orc):
    """Hash-first mapping (mode 1) of short reads in the filter tier: the
    biggest clusters of two or more hits from the exact candidate table
    (listed by one wave, or by table scans past 64 tied clusters), strands
    whose clusters are all single hits, and strands past the first stage's
    cap through the 1024-thread stage and beyond; every strand against the
    oracle, tie-set sizes and hit counts included."""
    before = S.get_map_mode()
    try:
        S.set_map_mode(1)
        seen = set()
        for seqs, k, reads in filter_tier_reads():
            ix = S.Index.from_sequences(seqs, k, 100000)
            got = ix.map_reads(reads, 3)
            oix = orc.Index(seqs, k, 100000)
            check_strands(oix, reads, got, 3, orc)
            seen.update(int(b) for b in got.nbiggest)
    finally:
        S.set_map_mode(before)
    assert 100 in seen   # the 100 tied clusters of case (c)


@pytest.mark.parametrize("mode,long_read", [(0, False), (1, False), (1, True)])
def test_map_modes_vs_oracle(S, orc, mode, long_read):
    """Both entry tiers give the oracle's hits: mode 1 sends every strand
    (a handful of hits, empty reads, thousands of hits) through the
    hash-cluster tier first, mode 0 through the one-wave tier; the pipeline
    records match the oracle's too.  A read past 511 bases moves the batch
    from the short-read hash kernels to the general one (mapping only)."""
    from swift_amd import synth
    rnd = random.Random(90 + mode)
    seqs = rand_genome(rnd, 4, 3000, 30000, repeats=10)
    unit = b"ACGTTGCAAC"
    seqs.append(unit * 700 + bytes(rnd.choice(b"ACGT") for _ in range(3000)))
    g = np.random.default_rng(90 + mode)
    reads, _ = synth.sample_reads(g, [np.frombuffer(s, np.uint8) for s in seqs[:4]], 200, 100,
                                  0.01, 0.001)
    tail = bytes(rnd.choice(b"ACGT") for _ in range(90))
    reads += [b"", b"ACGTA", unit * 2 + tail, unit * 6 + tail, unit * 12, seqs[4][-120:]]
    if long_read:
        reads.append(seqs[0][1000:1700])
    before = S.get_map_mode()
    try:
        S.set_map_mode(mode)
        assert S.get_map_mode() == mode
        ix = S.Index.from_sequences(seqs, 10, 100000)
        got = ix.map_reads(reads, 3)
        assert (got.nhits > 1024).any() and (got.nhits <= 64).any()
        oix = orc.Index(seqs, 10, 100000)
        check_strands(oix, reads, got, 3, orc)
        if long_read:
            return  # the pipeline's windows take reads up to 320 bases
        recs = ix.align_reads(reads, 3)
        want = orc.pipeline(oix, reads, 3)
        assert recs.n == len(want)
        for i, w in enumerate(want):
            r = recs.record(i)
            for gk, wk in zip(REC_KEYS, ORC_KEYS):
                assert r[gk] == w[wk], (i, gk)
    finally:
        S.set_map_mode(before)


def test_map_reads_rejects_long_reads(S):
    ix = S.Index.from_sequences([b"ACGT" * 100], 8, 10)
    with pytest.raises(S.SwiftError):
        ix.map_reads([b"ACGT" * 20], 3, max_read_len=40)


@pytest.mark.parametrize("case", GOLD["lookup_table5_map"], ids=lambda c: c["source"][:44])
def test_lookup_table5_map_qry_known(S, case, tmp_path):
    p = tmp_path / "ref.fa"
    p.write_text(case["fasta"])
    S.lookup_table5_create(p, case["k"], case["max_hits"], case["thr"])
    q = case["qry"].encode()
    _check_choice(S.lookup_table5_map_qry(q, False), case["fwd"])
    _check_choice(S.lookup_table5_map_qry(q, True), case["rev"])
    S.lookup_table5_delete()


@pytest.mark.parametrize("case", GOLD["map_hits6"], ids=lambda c: c["source"][:40])
def test_map_hits6_known(S, case):
    best, srt = S.map_hits6_best_hits([tuple(h) for h in case["hits"]], case["max_hits"])
    if "sorted" in case:
        assert [list(s) for s in srt] == case["sorted"]
    if "best" in case:
        assert [list(b) for b in best] == case["best"]
    if "choose" in case:
        _check_choice(best, case)


def test_map_hits6_vs_oracle(S, orc):
    rnd = random.Random(12)
    for n in list(range(1, 40)) + [300, 3000]:
        hits = [(rnd.randint(-3, 3), rnd.randint(-8, 8), rnd.randint(-50, 50)) for _ in range(n)]
        mh = rnd.randint(1, 5)
        best, srt = S.map_hits6_best_hits(hits, mh)
        obest, _, osrt = orc.best_hits(hits, mh, orc.TIE_FIRST)
        assert srt == osrt and best == obest, n


@pytest.mark.parametrize("case", GOLD["search_queries3"], ids=lambda c: c["source"][:40])
def test_search_queries3_known(S, case, tmp_path):
    ref, qry, out = tmp_path / "r.fa", tmp_path / "q.fa", tmp_path / "m.txt"
    ref.write_text(case["fasta"])
    qry.write_text("".join(f">{n}\n{s}\n" for n, s in case["reads"]))
    nm = S.search_queries3(qry, ref, out, case["k"], case["max_hits"], 1000)
    lines = out.read_text().splitlines()
    assert nm == len(lines)
    if "lines" in case:
        assert lines == case["lines"]
    else:
        assert set(lines) <= set(case["fwd_set"]) | set(case["rev_set"])
        assert nm == case["num_matches"]


def test_search2_queries_vs_oracle(S, orc, tmp_path):
    """search2Queries: every read, both strands, at most 3 hits each, the
    shift column 0 (src/search2.cu:64,341-373).  Expected lines come from the
    oracle's per-strand mapping, with the name and offsets of each read's
    header line (qryGetNext2, src/query.cu:174-215: nameOffset at the '>',
    seqOffset just past the header line)."""
    from swift_amd import synth
    rnd = random.Random(41)
    seqs = rand_genome(rnd, 3, 3000, 20000, repeats=6)
    ref = tmp_path / "ref.fa"
    with open(ref, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b">chr%d\n" % i + b"".join(s[p:p + 50] + b"\n" for p in range(0, len(s), 50)))
    g = np.random.default_rng(41)
    reads, _ = synth.sample_reads(g, [np.frombuffer(s, np.uint8) for s in seqs], 80, 100,
                                  0.02, 0.002)
    reads += [bytes(rnd.choice(b"ACGT") for _ in range(100)) for _ in range(5)]
    reads += [seqs[2][:70] + b"NNNN" + seqs[2][74:100], b"ACGT" * 3]
    qry = tmp_path / "reads.fa"
    blob = bytearray()
    want = []
    ix = orc.Index(seqs, 12, 1000)
    for i, r in enumerate(reads):
        name = b"read %d x" % i
        name_off = len(blob)
        blob += b">" + name + b"\n"
        seq_off = len(blob)
        blob += r + b"\n"
        for strand, s in enumerate((r, orc.revcomp(r))):
            hits, _, _ = ix.map_query(s, 3)
            want += ["%s\t%d\t%d\t0\t%d\t%d\t%d" % (name.decode(), strand, h[0], h[2],
                                                     name_off, seq_off) for h in hits]
    qry.write_bytes(bytes(blob))
    out = tmp_path / "m.txt"
    nm = S.search2_queries(qry, ref, out, 12, 1000)
    lines = out.read_text().splitlines()
    assert len(want) > 80
    assert nm == len(lines)
    assert lines == want


# ------------------------------------------------------------------ pipeline
REC_KEYS = ("read", "strand", "ref_idx", "shift", "ref_pos", "nbiggest", "ref_distance", "score",
            "start", "end", "alen", "keep", "ntied", "qry_cons", "ref_cons")
ORC_KEYS = ("read", "strand", "refIdx", "shift", "refPos", "nBiggest", "refDistance", "score",
            "start", "end", "len", "keep", "nTied", "qry_cons", "ref_cons")


def test_align_reads_vs_oracle_pipeline(S, orc):
    """Reads -> both strands mapped -> 300-bp windows (N-trimmed) -> SW ->
    best-hit choice, record by record against the oracle's CPU pipeline."""
    from swift_amd import synth
    rnd = random.Random(21)
    seqs = rand_genome(rnd, 4, 5000, 60000, repeats=8)
    seqs[1] = b"N" * 150 + seqs[1][150:-200] + b"n" * 200   # N runs at the ends
    g = np.random.default_rng(21)
    reads, _ = synth.sample_reads(g, [np.frombuffer(s, np.uint8) for s in seqs], 300, 100,
                                  0.02, 0.002)
    reads += [bytes(rnd.choice(b"ACGT") for _ in range(100)) for _ in range(20)]
    reads += [seqs[1][:100], seqs[1][-100:], seqs[0][:60] + b"acgtn" + seqs[0][65:100]]
    ix = S.Index.from_sequences(seqs, 12, 1000)
    got = ix.align_reads(reads, 3)
    want = orc.pipeline(orc.Index(seqs, 12, 1000), reads, 3)
    assert got.n == len(want)
    for i, w in enumerate(want):
        g_rec = got.record(i)
        for gk, wk in zip(REC_KEYS, ORC_KEYS):
            assert g_rec[gk] == w[wk], (i, gk, g_rec, w)


def test_align_reads_uneven_lengths(S, orc):
    from swift_amd import synth
    rnd = random.Random(33)
    seqs = rand_genome(rnd, 3, 20000, 50000)
    g = np.random.default_rng(33)
    arr = [np.frombuffer(s, np.uint8) for s in seqs]
    reads = []
    for L in (50, 77, 150, 220, 300):
        rr, _ = synth.sample_reads(g, arr, 40, L, 0.015, 0.0015)
        reads += rr
    rnd.shuffle(reads)
    ix = S.Index.from_sequences(seqs, 12, 1000)
    got = ix.align_reads(reads, 3)
    want = orc.pipeline(orc.Index(seqs, 12, 1000), reads, 3)
    assert got.n == len(want)
    for i, w in enumerate(want):
        g_rec = got.record(i)
        for gk, wk in zip(REC_KEYS, ORC_KEYS):
            assert g_rec[gk] == w[wk], (i, gk)


def test_c1_config_vs_oracle_pipeline(S, orc):
    """The small config in full (1 Mbp uniform reference, 1,000 reads of
    100 bp at 2 % substitutions and 0.2 % indels, k = 12, threshold 1000,
    maxHits 3, default scoring): every record of the device pipeline equals
    the oracle's CPU pipeline, field by field and string by string."""
    from swift_amd import synth
    seqs, reads, truth = synth.c1_workload()
    ix = S.Index.from_sequences(seqs, 12, 1000)
    got = ix.align_reads(reads, 3)
    want = orc.pipeline(orc.Index(seqs, 12, 1000), reads, 3)
    assert got.n == len(want) > len(reads)
    for i, w in enumerate(want):
        g_rec = got.record(i)
        for gk, wk in zip(REC_KEYS, ORC_KEYS):
            assert g_rec[gk] == w[wk], (i, gk)
