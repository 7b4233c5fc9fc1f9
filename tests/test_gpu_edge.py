"""Empty batches and zero-length sequences through every device entry point.
The oracle's answers for them (smithWaterman_cpu, src/smithWaterman.cu:
44-335, over an empty query or reference: score 0, start 0, end -1, no
strings; lookupTable5MapQry, src/lookupTable5.cu:557-603, over a read
shorter than k: no hits) must come back, with no error."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("score", "start", "end", "len", "qry_cons", "ref_cons")


@pytest.fixture(scope="module")
def A():
    from swift_amd import align
    return align


@pytest.fixture(scope="module")
def S():
    from swift_amd import search
    return search


@pytest.fixture(scope="module")
def genome():
    rnd = random.Random(77)
    return [bytes(rnd.choice(b"ACGT") for _ in range(n)) for n in (30000, 12000)]


def test_sw_batch_no_pairs(A):
    res = A.sw_batch([], [])
    assert len(res.score) == 0 and len(res.start) == 0 and res.qry_cons == []


@pytest.mark.parametrize("sc", [(2, -1, -10, -1), (1, 0, -1, 0)])
def test_sw_batch_zero_length_sequences(A, orc, sc):
    """Empty queries and empty references among ordinary pairs, on the
    packed int16 path and (second scoring) the int32 one."""
    rnd = random.Random(5)
    qs, rs = [], []
    for i in range(300):
        q = bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(1, 200)))
        r = bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(1, 300)))
        if i % 7 == 0:
            q = b""
        if i % 11 == 0:
            r = b""
        qs.append(q)
        rs.append(r)
    res = A.sw_batch(qs, rs, A.Scoring(*sc))
    for i in range(len(qs)):
        got, want = res.record(i), orc.sw(rs[i], qs[i], *sc)
        for k in KEYS:
            assert got[k] == want[k], (i, k, got, want)


def test_align_sequences_empty_references(A, orc):
    """alignSequences: references whose first byte is NUL have no columns."""
    rnd = random.Random(9)
    qs = [bytes(rnd.choice(b"ACGT") for _ in range(40)) for _ in range(64)]
    rs = [b"" if i % 3 == 0 else bytes(rnd.choice(b"ACGT") for _ in range(rnd.randint(1, 90)))
          for i in range(64)]
    res = A.align_sequences(qs, rs, 40, 90)
    for i in range(len(qs)):
        got, want = res.record(i), orc.sw(rs[i], qs[i], ref_len=90)
        for k in KEYS:
            assert got[k] == want[k], (i, k)


def test_map_reads_empty_and_short(S, orc, genome):
    ix = S.Index.from_sequences(genome, 12, 1000)
    got = ix.map_reads([], 3)
    assert len(got.nhits) == 0
    reads = [b"", b"ACG", genome[0][100:111], genome[0][500:600], b""]
    got = ix.map_reads(reads, 3)
    oix = orc.Index(genome, 12, 1000)
    for r, q in enumerate(reads):
        for strand in (0, 1):
            seq = q if strand == 0 else orc.revcomp(q)
            want, nbig, nh = oix.map_query(seq, 3, orc.TIE_FIRST)
            assert got.hits(2 * r + strand) == want, (r, strand)
            assert got.nhits[2 * r + strand] == nh


def test_align_reads_empty_and_short(S, orc, genome):
    ix = S.Index.from_sequences(genome, 12, 1000)
    assert ix.align_reads([], 3).n == 0
    reads = [b"", genome[1][2000:2100], b"AC", b"", genome[0][29900:30000]]
    got = ix.align_reads(reads, 3)
    want = orc.pipeline(orc.Index(genome, 12, 1000), reads, 3)
    assert got.n == len(want) > 0
    for i, w in enumerate(want):
        g = got.record(i)
        assert (g["read"], g["strand"], g["ref_pos"], g["score"], g["start"], g["end"],
                g["qry_cons"], g["ref_cons"]) == (w["read"], w["strand"], w["refPos"], w["score"],
                                                  w["start"], w["end"], w["qry_cons"],
                                                  w["ref_cons"]), i


def test_search_drivers_empty_query_file(S, genome, tmp_path):
    from swift_amd import synth
    ref = tmp_path / "ref.fa"
    ref.write_bytes(synth.to_fasta(genome, width=50))
    qry = tmp_path / "empty.fa"
    qry.write_bytes(b"")
    m3, m2 = tmp_path / "m3.txt", tmp_path / "m2.txt"
    assert S.search_queries3(qry, ref, m3) == 0
    assert S.search2_queries(qry, ref, m2) == 0
    assert m3.read_bytes() == b"" and m2.read_bytes() == b""


@pytest.mark.parametrize("gpu", [True, False])
def test_align_queries_empty_match_file(S, genome, tmp_path, gpu):
    from swift_amd import synth
    ref = tmp_path / "ref.fa"
    ref.write_bytes(synth.to_fasta(genome, width=50))
    qry = tmp_path / "reads.fa"
    qry.write_bytes(b">r0\n" + genome[0][:100] + b"\n")
    match = tmp_path / "match.txt"
    match.write_bytes(b"")
    out = tmp_path / "out.txt"
    S.align_queries(qry, ref, match, out, 100, gpu_driver=gpu)
    assert out.read_bytes() == b""


def test_server_empty_request(S, genome):
    from swift_amd import serve
    ix = S.Index.from_sequences(genome, 12, 1000)
    with serve.Server(ix, nslots=2, mem_budget=1 << 30, max_batch_reads=1024) as srv:
        a = srv.submit([])
        b = srv.submit([genome[0][1000:1100]])
        c = srv.submit([b""])
        (ra, _), (rb, _), (rc, _) = srv.result(a), srv.result(b), srv.result(c)
    assert ra.n == 0 and rc.n == 0
    assert rb.n >= 1 and rb.record(0)["score"] == 200
