"""Serving loop on the GPU: requests batched together (continuous batching
across concurrent slots) get exactly the records they get alone, and those
are the oracle's CPU pipeline records over the request."""
import threading

import numpy as np
import pytest

from parity_util import assert_records_match

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    from swift_amd import search, synth
    rng = np.random.default_rng(55)
    seqs = [synth.random_bases(rng, n, 0.41) for n in (300_000, 180_000, 90_000)]
    reqs = []
    for _ in range(36):
        n = int(np.exp(rng.uniform(0, np.log(400))))
        L = int(rng.integers(50, 301))
        reads, _ = synth.sample_reads(rng, seqs, n, L, 0.015, 0.0015)
        if rng.random() < 0.3:
            reads.append(synth.random_bases(rng, L).tobytes())
        reqs.append(reads)
    ix = search.Index.from_sequences([s.tobytes() for s in seqs], 12, 1000)
    return ix, reqs, [s.tobytes() for s in seqs]


def same(a, b, ctx):
    assert a.n == b.n, ctx
    for k in b.cols:
        assert np.array_equal(np.asarray(a.cols[k]), np.asarray(b.cols[k])), (ctx, k)
    assert a.qry_cons == b.qry_cons and a.ref_cons == b.ref_cons, ctx


def test_results_same_alone_and_in_mix(world, orc):
    from swift_amd.serve import Server
    ix, reqs, _ = world
    alone = [ix.align_reads(r, 3) for r in reqs]
    with Server(ix, nslots=4, mem_budget=16 << 30, max_batch_reads=4096) as srv:
        ids = [None] * len(reqs)

        def submit(lo):  # several client threads submit at once
            for i in range(lo, len(reqs), 3):
                ids[i] = srv.submit(reqs[i])

        th = [threading.Thread(target=submit, args=(k,)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        got = [srv.result(i)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["batches"] < len(reqs) and st["max_coalesced"] > 1  # requests were coalesced
    for i, (g, a) in enumerate(zip(got, alone)):
        same(g, a, i)
    oix = orc.Index(world[2], 12, 1000)
    for i, (g, r) in enumerate(zip(got, reqs)):
        assert_records_match(g, orc.pipeline_mt(oix, r, 8), i)


def test_large_requests_served_in_parts(world, orc):
    """Requests above 1,024 reads are served as parts over the slots; each
    one's records are still those of the request alone, in read order, and
    the oracle's."""
    from swift_amd import synth
    from swift_amd.serve import Server
    ix, _, seqs = world
    rng = np.random.default_rng(77)
    reqs = []
    for n in (1025, 2500, 4000, 700):
        reads, _ = synth.sample_reads(rng, [np.frombuffer(s, np.uint8) for s in seqs], n,
                                      int(rng.integers(80, 301)), 0.015, 0.0015)
        reqs.append(reads)
    alone = [ix.align_reads(r, 3) for r in reqs]
    with Server(ix, nslots=4, mem_budget=16 << 30, max_batch_reads=4096) as srv:
        ids = [srv.submit(r) for r in reqs]
        got = [srv.result(i)[0] for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    for i, (g, a) in enumerate(zip(got, alone)):
        same(g, a, i)
    oix = orc.Index(seqs, 12, 1000)
    for i, (g, r) in enumerate(zip(got, reqs)):
        assert_records_match(g, orc.pipeline_mt(oix, r, 8), i)


def test_served_int32_reruns_vs_oracle(orc):
    """References holding bytes outside ACGTNacgtn (IUPAC R and Y, about 1 %
    of the bases): their windows are re-run with byte semantics in the
    int32 pass.  A serving slot runs its classes on one stream, where each
    pass's walk follows the pass directly, so the re-run pairs take the
    second pass's walk.  Every request's records against the oracle's."""
    from swift_amd import search, synth
    from swift_amd.serve import Server
    rng = np.random.default_rng(57)
    seqs = []
    for n in (120_000, 80_000):
        s = bytearray(synth.random_bases(rng, n, 0.41).tobytes())
        for i in rng.choice(n, n // 100, replace=False):
            s[i] = ord("R") if rng.random() < 0.5 else ord("Y")
        seqs.append(bytes(s))
    arrs = [np.frombuffer(x, np.uint8) for x in seqs]
    reqs = []
    for _ in range(8):
        n = int(rng.integers(1, 60))
        reads, _ = synth.sample_reads(rng, arrs, n, int(rng.integers(60, 201)), 0.01, 0.001)
        reqs.append(reads)
    ix = search.Index.from_sequences(seqs, 12, 1000)
    oix = orc.Index(seqs, 12, 1000)
    with Server(ix, nslots=2, mem_budget=4 << 30, max_batch_reads=1024) as srv:
        ids = [srv.submit(r) for r in reqs]
        got = [srv.result(i)[0] for i in ids]
    nrec = 0
    for i, (g, r) in enumerate(zip(got, reqs)):
        assert_records_match(g, orc.pipeline_mt(oix, r, 8), i)
        nrec += g.n
    assert nrec > 0


def test_small_budget_many_batches(world):
    from swift_amd.serve import Server
    from swift_amd._lib import SwiftError
    ix, reqs, _ = world
    small = [r for r in reqs if len(r) <= 64][:10]
    with Server(ix, nslots=2, mem_budget=1 << 30, max_batch_reads=128) as srv:
        ids = [srv.submit(r) for r in small]
        got = [srv.result(i)[0] for i in ids]
        with pytest.raises(SwiftError):
            srv.submit([b"ACGT" * 25] * 129)   # larger than one slot admits
    for g, r in zip(got, small):
        same(g, ix.align_reads(r, 3), len(r))


def test_one_slot_in_order(world):
    from swift_amd.serve import Server
    ix, reqs, _ = world
    with Server(ix, nslots=1, mem_budget=8 << 30, max_batch_reads=4096) as srv:
        for r in reqs[:6]:
            got, lat = srv.result(srv.submit(r))
            assert lat > 0
            same(got, ix.align_reads(r, 3), len(r))
