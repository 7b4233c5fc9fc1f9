"""Output drop-ins (host code of the HIP library, no GPU needed) against the
reference's own writers (compiled from its sources) and the oracle."""
import random

import pytest


@pytest.fixture(scope="module")
def OUT():
    import os
    from swift_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    from swift_amd import output
    return output


def alignments(orc, n, seed):
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        q = bytes(rnd.choice(b"ACGTn") for _ in range(rnd.randint(5, 60)))
        r = bytes(rnd.choice(b"ACGTNa") for _ in range(rnd.randint(5, 90)))
        out.append(orc.sw(r, q))
    return out


def test_cigar_vs_reference(OUT, orc, refbridge):
    import ctypes as C
    for res in alignments(orc, 300, 1):
        buf = C.create_string_buffer(8192)
        n = refbridge.ref_format_cigar(res["qry_cons"], res["ref_cons"], res["len"], buf, 8192)
        assert OUT.cigar(res["qry_cons"], res["ref_cons"]) == buf.raw[:n].decode()


def test_print_output_vs_reference(OUT, orc, refbridge):
    import ctypes as C
    als = alignments(orc, 150, 2)
    got = OUT.print_output([f"read{i}*" for i in range(len(als))],
                           [f"chr{i % 3}" for i in range(len(als))], [a["score"] for a in als],
                           [17 * i for i in range(len(als))], [a["start"] for a in als],
                           [a["end"] for a in als], [a["qry_cons"] for a in als],
                           [a["ref_cons"] for a in als])
    want = []
    for i, a in enumerate(als):
        buf = C.create_string_buffer(8192)
        n = refbridge.ref_format_record(f"read{i}*".encode(), f"chr{i % 3}".encode(),
                                        C.c_float(a["score"]), 17 * i, a["start"], a["end"],
                                        a["len"], a["qry_cons"], a["ref_cons"],
                                        len(a["qry_cons"]) + 2, buf, 8192)
        want.append(buf.raw[:n].decode("latin-1"))
    assert got == "".join(want)


def test_print_output_skips_dropped_and_matches_oracle(OUT, orc):
    als = alignments(orc, 40, 3)
    scores = [a["score"] if i % 4 else -1 for i, a in enumerate(als)]
    got = OUT.print_output([f"q{i}" for i in range(40)], ["r"] * 40, scores, [5] * 40,
                           [a["start"] for a in als], [a["end"] for a in als],
                           [a["qry_cons"] for a in als], [a["ref_cons"] for a in als])
    want = "".join(orc.format_record(f"q{i}", "r", a["score"], 5, a["start"], a["end"],
                                     a["qry_cons"], a["ref_cons"])
                   for i, a in enumerate(als) if i % 4)
    assert got == want


def test_print_output_sam(OUT, orc):
    als = alignments(orc, 60, 4)
    got = OUT.print_output([f"q{i}" for i in range(60)], ["chrZ"] * 60, [1] * 60, [0] * 60,
                           [a["start"] for a in als], [a["end"] for a in als],
                           [a["qry_cons"] for a in als], [a["ref_cons"] for a in als], sam=True)
    # printOutputSam (src/output.cu:98-130): name, *, ref, start, 255, CIGAR,
    # *, 0, 0, the query as aligned (reversed walk order), *
    want = "".join(f"q{i}\t*\tchrZ\t{a['start']}\t255\t{orc.cigar(a['qry_cons'], a['ref_cons'])}"
                   f"\t*\t0\t0\t{a['qry_cons'][::-1].decode()}\t*\n" for i, a in enumerate(als))
    assert got == want
