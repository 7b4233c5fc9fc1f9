"""The oracle's threaded pipeline (the CPU baseline of the pipeline configs)
gives the records of the single-threaded pipeline, in the same order."""
import numpy as np


def test_pipeline_mt_equals_pipeline(orc):
    from swift_amd import synth
    rng = np.random.default_rng(5)
    seqs = [synth.random_bases(rng, n).tobytes() for n in (30000, 20000)]
    reads, _ = synth.sample_reads(rng, [np.frombuffer(s, np.uint8) for s in seqs], 150, 100,
                                  0.02, 0.002)
    reads += [synth.random_bases(rng, 100).tobytes() for _ in range(10)]
    ix = orc.Index(seqs, 12, 1000)
    want = orc.pipeline(ix, reads, 3)
    for t in (1, 3, 16):
        assert orc.pipeline_mt(ix, reads, t) == want, t
