"""Phase 1 -> phase 2 as one hipGraph: every replay, with other device work
and synchronisations between replays and new reads loaded each time, gives
the records of the eager pipeline and of the oracle's CPU pipeline."""
import numpy as np
import pytest

from parity_util import assert_records_match

pytestmark = pytest.mark.gpu


def same(a, b, ctx):
    assert a.n == b.n, ctx
    for k in b.cols:
        assert np.array_equal(np.asarray(a.cols[k]), np.asarray(b.cols[k])), (ctx, k)
    assert a.qry_cons == b.qry_cons and a.ref_cons == b.ref_cons, ctx


def test_graph_replays_match_eager(orc):
    from swift_amd import search, synth
    rng = np.random.default_rng(77)
    seqs = [synth.random_bases(rng, n, 0.41) for n in (250_000, 120_000)]
    ix = search.Index.from_sequences([s.tobytes() for s in seqs], 12, 1000)
    oix = orc.Index([s.tobytes() for s in seqs], 12, 1000)
    sets = []
    for L, n in ((100, 300), (150, 260), (60, 310)):
        reads, _ = synth.sample_reads(rng, seqs, n, L, 0.015, 0.0015)
        reads.append(synth.random_bases(rng, L).tobytes())
        sets.append(reads)
    g = search.PipelineGraph(ix, 320, 150, 3)
    try:
        for rep, reads in enumerate(sets + [sets[0]]):
            got = g.run(reads)
            padded = list(reads) + [b""] * (g.slots - len(reads))
            want = ix.align_reads(padded, 3, max_read_len=150)   # eager work between replays
            same(got, want, rep)
            assert got.n > 0
            assert_records_match(got, orc.pipeline(oix, reads, 3), rep)
    finally:
        g.close()
