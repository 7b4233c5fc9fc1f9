"""Parity on the benchmark configurations' own workloads.

Each test runs the HIP path on the workload generator that bench.py uses for
that configuration and compares every result with the oracle (the CPU
restatement of smithWaterman_cpu and of the searchQueries3 ->
alignQueries_cpu_mt -> chooseBestAlignment pipeline, src/smithWaterman.cu:44-335,
src/search.cu:275-382, src/align.cu:844-905,2363-2548), and, for the
batched pairs, with the reference's own compiled smithWaterman_cpu where
oracle/_ref is built:

* C3: pairs from synth.c3_pairs_fast, the headline pairs' generator;
* C2: the full 50 Mbp / 24-sequence genome of synth.c2_workload, with reads
  including the fully random 5 %, and its phase 2 also through the plain
  per-pair path;
* C4: the full 3.1 Gbp genome of synth.genome_c4 (built once for the
  module), 150-bp reads through the eager pipeline and the hipGraph;
* C5: serving requests from synth.c5_requests (log-uniform 1..4096 reads,
  lengths U{50..300}, 1.5 % error) submitted at once to a server over the
  resident C4 index; each request's records against the oracle pipeline
  over that request alone.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from parity_util import assert_records_match

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ C3
def test_c3_fast_pairs_vs_oracle_and_reference(orc):
    from swift_amd import align, synth
    n = 50_000
    qbuf, qoff, qlen, rbuf, roff, rlen = synth.c3_pairs_fast(n, seed=20261213)
    batch = align.PairBatch.from_arrays(qbuf, qoff, qlen, rbuf, roff, rlen)
    batch.run()
    got = batch.results()
    qs = [qbuf[o:o + k].tobytes() for o, k in zip(qoff, qlen)]
    rs = [rbuf[o:o + k].tobytes() for o, k in zip(roff, rlen)]
    sc, st, en, al, qc, rc = orc.sw_batch(qs, rs, nthreads=16)
    assert np.array_equal(got.score, sc) and np.array_equal(got.length, al)
    assert np.array_equal(got.start, st) and np.array_equal(got.end, en)
    assert got.qry_cons == qc and got.ref_cons == rc
    R = orc.ref()
    if R is None:
        return
    # the reference's smithWaterman_cpu itself, strings on the first 5,000
    m = 5000
    cs = (qlen.astype(np.int64) + rlen + 2)
    co = np.zeros(n, np.int64)
    co[1:] = np.cumsum(cs)[:-1]
    rq, rr = np.zeros(int(cs.sum()), np.uint8), np.zeros(int(cs.sum()), np.uint8)
    out = [np.zeros(n, np.int32) for _ in range(4)]
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    qo, ro = qoff.astype(np.int64), roff.astype(np.int64)
    ql, rl = qlen.astype(np.int32), rlen.astype(np.int32)
    R.ref_sw_batch_mt2(n, p(qbuf, C.c_char), p(qo, C.c_long), p(ql, C.c_int),
                       p(rbuf, C.c_char), p(ro, C.c_long), p(rl, C.c_int), C.c_float(2),
                       C.c_float(-1), C.c_float(-10), C.c_float(-1), *(p(a, C.c_int) for a in out),
                       p(rq, C.c_char), p(rr, C.c_char), p(co, C.c_long), m, 16)
    for a, b in zip((got.score, got.start, got.end, got.length), out):
        assert np.array_equal(a, b)
    for i in range(m):
        o, k = int(co[i]), int(al[i])
        assert rq[o:o + k].tobytes() == got.qry_cons[i] and rr[o:o + k].tobytes() == \
            got.ref_cons[i], i


# ------------------------------------------------------------------ C2
def region(seq, ref_distance, width=300):
    """The record's window as the pipeline cuts it (alignQueries_cpu_mt's
    getRefSeq2, src/input.cu:348-403): up to `width` bases from
    ref_distance, leading N/n stripped, else trailing N/n."""
    w = seq[ref_distance:ref_distance + width]
    lead = len(w) - len(w.lstrip(b"Nn"))
    return w[lead:] if lead else w.rstrip(b"Nn")


def revcomp(read):
    """qryGetReverseComplement (src/query.cu:265-282): A->T, C->G, G->C,
    every other byte -> A."""
    m = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C")}
    return bytes(m.get(c, ord("A")) for c in reversed(read))


def test_c2_workload_vs_oracle(orc):
    """The C2 pipeline against the oracle's; then its phase 2 through the
    plain per-pair path (the configuration's "plain phase-2"): every
    record's (query, window) pair through alignSequences, one call per read
    length, against the oracle's record."""
    from swift_amd import align, search, synth
    seqs, reads, truth = synth.c2_workload(n_reads=2000)   # + 100 fully random reads
    assert sum(1 for t in truth if t[0] < 0) == 100
    ix = search.Index.from_sequences(seqs, 12, 1000)
    got = ix.align_reads(reads, 3)
    want = orc.pipeline_mt(orc.Index(seqs, 12, 1000), reads, 16)
    assert_records_match(got, want, "C2")
    assert got.n > 2 * 2000
    by_len = {}
    for i, w in enumerate(want):
        r = reads[w["read"]]
        q = revcomp(r) if w["strand"] else r
        by_len.setdefault(len(q), []).append((i, q, region(seqs[w["refIdx"]], w["refDistance"])))
    for ql, items in by_len.items():
        res = align.align_sequences([q for _, q, _ in items], [x for _, _, x in items], ql, 300)
        for k, (i, _, _) in enumerate(items):
            g, w = res.record(k), want[i]
            assert (g["score"], g["start"], g["end"], g["len"]) == \
                (w["score"], w["start"], w["end"], w["len"]), ("C2 plain", i)
            assert g["qry_cons"] == w["qry_cons"] and g["ref_cons"] == w["ref_cons"], \
                ("C2 plain", i)


# ------------------------------------------------------------------ C4 / C5
@pytest.fixture(scope="module")
def c4():
    from swift_amd import search, synth
    t0 = time.time()
    g, off, lens = synth.genome_c4()
    log(f"[c4] genome {len(g):.3e} bp in {time.time() - t0:.1f} s")
    ix = search.Index.from_packed(g, off, lens, 12, 1000)
    seqs = [g[o:o + n].tobytes() for o, n in zip(off, lens)]
    t0 = time.time()
    oix = orc_index(seqs)
    log(f"[c4] oracle index in {time.time() - t0:.1f} s")
    return g, off, lens, ix, oix


def orc_index(seqs):
    from oracle import oracle as O
    O.build()
    return O.Index(seqs, 12, 1000)


def test_c4_eager_and_graph_vs_oracle(orc, c4):
    from swift_amd import search, synth
    g, off, lens, ix, oix = c4
    rng = np.random.default_rng(20261214)
    buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, 300, 150, 0.01)
    reads = [buf[o:o + 150].tobytes() for o in ro]
    want = orc.pipeline_mt(oix, reads, 16)
    assert_records_match(ix.align_reads(reads, 3), want, "C4 eager")
    pg = search.PipelineGraph(ix, 320, 150, 3)
    try:
        assert_records_match(pg.run(reads), want, "C4 graph")
        assert_records_match(pg.run(reads[:123]), want[:sum(1 for w in want if w["read"] < 123)],
                             "C4 graph, fewer reads")
    finally:
        pg.close()


def test_c5_requests_on_c4_index_vs_oracle(orc, c4):
    from swift_amd import synth
    from swift_amd.serve import Server
    g, off, lens, ix, oix = c4
    rng = np.random.default_rng(20261215)
    reqs = synth.c5_requests(rng, g, off, lens, 24, max_reads=1024)
    total = sum(len(r[2]) for r in reqs)
    log(f"[c5] {len(reqs)} requests, {total} reads")
    with Server(ix, nslots=4, mem_budget=48 << 30, max_batch_reads=8192) as srv:
        ids = [srv.submit_packed(*r) for r in reqs]      # all at once: batched together
        got = [srv.result(i)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["max_coalesced"] > 1
    for k, ((buf, ro, rl), gk) in enumerate(zip(reqs, got)):
        reads = [buf[o:o + n].tobytes() for o, n in zip(ro, rl)]
        assert_records_match(gk, orc.pipeline_mt(oix, reads, 16), f"C5 request {k}")
