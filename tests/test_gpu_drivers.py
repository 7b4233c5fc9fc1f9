"""The phase-2 drivers (alignQueries3, alignQueries_cpu_mt) against the
reference's own functions: each match line's window (refMapCreate +
refMapGetOffsets + getRefSeq2), query and names (getSeq, getSeqName,
qryGetReverseComplement) and output records (printOutput, printOutputSam)
come from the reference sources compiled under oracle/_ref; the scores and
strings from the oracle's smithWaterman_cpu restatement, and the best-record
choice from the oracle (first of the tied records).  The output files must be
byte-identical."""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu

QLEN = 100
BASES = b"ACGT"


def _fasta(path, recs, width):
    with open(path, "wb") as f:
        for name, seq in recs:
            f.write(b">" + name + b"\n")
            for i in range(0, len(seq), width):
                f.write(seq[i:i + width] + b"\n")


def _revcomp(s):
    t = {ord("A"): b"T", ord("C"): b"G", ord("G"): b"C", ord("T"): b"A"}
    return b"".join(t.get(c, b"A") for c in reversed(s))


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    from swift_amd import search
    d = tmp_path_factory.mktemp("drivers")
    rnd = random.Random(31)
    seqs = []
    for n in (20000, 15000, 3000):
        seqs.append(bytearray(rnd.choice(BASES) for _ in range(n)))
    seqs[1][:180] = b"N" * 180            # windows that start in an N run
    seqs[0][9000:9100] = b"N" * 100       # and ones that end in one
    seqs = [bytes(s) for s in seqs]
    ref = d / "ref.fa"
    _fasta(ref, [(b"chr1 first", seqs[0]), (b"chr2", seqs[1]), (b"chrM", seqs[2])], 50)
    reads = []
    starts = [(0, 8890), (0, 9100), (1, 180), (1, 250), (2, 2900), (0, 0)]
    for _ in range(60):
        s = rnd.randrange(3)
        starts.append((s, rnd.randrange(len(seqs[s]) - QLEN)))
    for s, p in starts:
        q = bytearray(seqs[s][p:p + QLEN].replace(b"N", b"A").ljust(QLEN, b"A"))
        for i in range(QLEN):
            if rnd.random() < 0.01:
                q[i] = rnd.choice(BASES)
        q = bytes(q)
        reads.append(_revcomp(q) if rnd.random() < 0.5 else q)
    for _ in range(6):
        reads.append(bytes(rnd.choice(BASES) for _ in range(QLEN)))
    qry = d / "reads.fa"
    _fasta(qry, [(b"r%09d" % i, r) for i, r in enumerate(reads)], 1000)
    match = d / "match.txt"
    search.search_queries3(qry, ref, match, k=12, max_hits=3, thr=1000)
    lines = open(match, "rb").read().splitlines(keepends=True)
    if len(lines) % 8 == 0:      # the CPU driver's last round must be partial
        lines = lines[:-3]
        open(match, "wb").write(b"".join(lines))
    assert len(lines) > 40 and len(lines) % 8
    return d, str(qry), str(ref), str(match), len(lines)


def _expected(O, qry, ref, match, gpu, fmt):
    R = O.ref()
    R.ref_get_seq.argtypes = [C.c_char_p, C.c_ulong, C.c_char_p, C.c_int]
    R.ref_get_seq_name.argtypes = [C.c_char_p, C.c_ulong, C.c_char_p]
    R.ref_map_offsets.argtypes = [C.c_char_p, C.c_int, C.c_uint, C.POINTER(C.c_uint),
                                  C.POINTER(C.c_uint)]
    lines = open(match, "rb").read().splitlines()
    if not gpu:
        # rounds of 8 fgets calls; past the end the last line is read again
        lines = lines + [lines[-1]] * ((-len(lines)) % 8)
    recs = []
    for ln in lines:
        f = ln.split(b"\t")
        rev, ri, rp, qno, qso = int(f[1]), int(f[2]), int(f[4]), int(f[5]), int(f[6])
        win = C.create_string_buffer(400)
        rd = C.c_int()
        R.ref_region(ref.encode(), ri, rp, win, C.byref(rd))
        name_off, seq_off = C.c_uint(), C.c_uint()
        R.ref_map_offsets(ref.encode(), ri, rd.value, C.byref(name_off), C.byref(seq_off))
        rn = C.create_string_buffer(1000)
        R.ref_get_seq_name(ref.encode(), name_off.value, rn)
        qb = C.create_string_buffer(QLEN + 8)
        R.ref_get_seq(qry.encode(), qso, qb, QLEN)
        q = qb.value
        if rev == 1:
            out = C.create_string_buffer(QLEN + 8)
            R.ref_revcomp(qb, QLEN, out)
            q = out.value
        qn = C.create_string_buffer(1000)
        R.ref_get_seq_name(qry.encode(), qno, qn)
        name = qn.value + (b"*" if rev == 1 else b"")
        res = O.sw(win.value, q)
        recs.append((name, rn.value, rd.value, res))
    keep = [True] * len(recs)
    if gpu and fmt == 0:
        base = [r[0].rstrip(b"*") for r in recs]
        g = 0
        while g < len(recs):
            h = g + 1
            while h < len(recs) and base[h] == base[g]:
                h += 1
            best = max([0] + [recs[i][3]["score"] for i in range(g, h)])
            first = next(i for i in range(g, h) if recs[i][3]["score"] == best)
            for i in range(g, h):
                keep[i] = i == first
            g = h
    out = bytearray()
    buf = C.create_string_buffer(8192)
    for k, (qn, rn, rd, r) in enumerate(recs):
        comb = 300 + QLEN + 2
        if fmt == 1:
            n = R.ref_format_sam(qn, rn, C.c_float(r["score"]), r["start"], r["end"], r["len"],
                                 r["qry_cons"], r["ref_cons"], comb, buf, len(buf))
        else:
            if not keep[k]:
                continue
            n = R.ref_format_record(qn, rn, C.c_float(r["score"]), rd, r["start"], r["end"],
                                    r["len"], r["qry_cons"], r["ref_cons"], comb, buf, len(buf))
        out += buf.raw[:n]
    return bytes(out)


@pytest.mark.parametrize("gpu,fmt", [(True, 0), (True, 1), (False, 0), (False, 1)])
def test_driver_output_matches_reference(files, orc, gpu, fmt):
    from swift_amd import search
    if orc.ref() is None:
        pytest.skip("oracle/_ref (the reference's compiled CPU functions) is not built")
    d, qry, ref, match, nlines = files
    out = d / f"out_{int(gpu)}_{fmt}.txt"
    search.align_queries(qry, ref, match, out, QLEN, out_format=fmt, gpu_driver=gpu)
    got = open(out, "rb").read()
    want = _expected(orc, qry, ref, match, gpu, fmt)
    assert got.count(b"\n") > 40
    assert got == want


def test_driver_drop_in_symbols_write_the_same_file(files, tmp_path):
    """alignQueries3 / alignQueries_cpu_mt through their reference
    signatures give what the status-returning form gives."""
    from swift_amd import search
    from swift_amd._lib import lib
    d, qry, ref, match, nlines = files
    for gpu in (True, False):
        a, b = tmp_path / f"a{int(gpu)}.txt", tmp_path / f"b{int(gpu)}.txt"
        search.align_queries(qry, ref, match, a, QLEN, gpu_driver=gpu)
        f = C.c_float
        if gpu:
            lib().alignQueries3(qry.encode(), QLEN, ref.encode(), match.encode(), 0, f(2),
                                f(-1), f(-10), f(-1), str(b).encode(), nlines)
        else:
            lib().alignQueries_cpu_mt(qry.encode(), 0, QLEN, ref.encode(), match.encode(), 0,
                                      f(2), f(-1), f(-10), f(-1), str(b).encode())
        assert open(a, "rb").read() == open(b, "rb").read()
