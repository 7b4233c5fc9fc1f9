"""Output of pipeline records: CIGAR strings computed on the device and the
text / SAM writer, against the oracle's pipeline and formatting."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(orc):
    from swift_amd import search, synth
    rng = np.random.default_rng(91)
    seqs = [synth.random_bases(rng, n, 0.41).tobytes() for n in (80_000, 50_000)]
    reads, _ = synth.sample_reads(rng, [np.frombuffer(s, np.uint8) for s in seqs], 120, 100,
                                  0.03, 0.01)
    ix = search.Index.from_sequences(seqs, 12, 1000)
    want = orc.pipeline(orc.Index(seqs, 12, 1000), reads, 3)
    return ix, reads, want


def test_records_cigar_on_device(world, orc):
    from swift_amd import output, search
    ix, reads, want = world
    rb = search.ReadBatch(reads)
    out = search.RecordBuffers(2 * rb.n * 3, rb.max_len)
    ix.align_batch(rb, out, 3)
    cig = output.records_cigar(out)
    assert len(cig) == len(want)
    for c, w in zip(cig, want):
        assert c == orc.cigar(w["qry_cons"], w["ref_cons"])


@pytest.mark.parametrize("sam", [False, True])
def test_write_records(world, orc, tmp_path, sam):
    from swift_amd import output
    ix, reads, want = world
    recs = ix.align_reads(reads, 3)
    names = [f"r{i:09d}" for i in range(len(reads))]
    path = tmp_path / "out.txt"
    output.write_records(path, recs, names, ix, sam=sam, only_kept=not sam)
    got = path.read_text()
    exp = []
    for w in want:
        qn = names[w["read"]] + ("*" if w["strand"] else "")
        if sam:
            exp.append(f"{qn}\t*\t\t{w['start']}\t255\t{orc.cigar(w['qry_cons'], w['ref_cons'])}"
                       f"\t*\t0\t0\t{w['qry_cons'][::-1].decode()}\t*\n")
        elif w["keep"]:
            exp.append(orc.format_record(qn, "", w["score"], w["refDistance"], w["start"],
                                         w["end"], w["qry_cons"], w["ref_cons"]))
    assert got == "".join(exp)
