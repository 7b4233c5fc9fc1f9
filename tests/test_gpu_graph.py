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


def test_two_graphs_on_two_streams_match_oracle(orc):
    """Two pipeline graphs, each over its own read and record buffers,
    launched in turn on two streams so that their replays overlap on the
    device: every replay's records equal the oracle's for its reads."""
    from swift_amd import search, synth
    from swift_amd.device import Stream
    rng = np.random.default_rng(78)
    seqs = [synth.random_bases(rng, n, 0.41) for n in (200_000, 150_000)]
    ix = search.Index.from_sequences([s.tobytes() for s in seqs], 12, 1000)
    oix = orc.Index([s.tobytes() for s in seqs], 12, 1000)
    graphs = [search.PipelineGraph(ix, 400, 150, 3) for _ in range(2)]
    streams = [Stream() for _ in range(2)]
    try:
        for rnd in range(3):
            sets = []
            for gi, g in enumerate(graphs):
                reads, _ = synth.sample_reads(rng, seqs, 400, 150, 0.01, 0.001)
                g.load(reads)
                g.launch(streams[gi].cptr)
                sets.append(reads)
            for gi, g in enumerate(graphs):
                streams[gi].synchronize()
                g.check()
                got = g.out.results(True)
                assert got.n > 0
                assert_records_match(got, orc.pipeline(oix, sets[gi], 3), (rnd, gi))
    finally:
        for g in graphs:
            g.close()
