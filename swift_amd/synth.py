"""Seeded synthetic workloads of the five benchmark configurations.

No network, no datasets: every genome, read set and pair list is generated
here from a named PRNG (numpy PCG64) with seed 20261200 + config number.
Reads are sampled from the generated reference with substitutions and
short indels, named r%09d.
"""
from __future__ import annotations

import numpy as np

BASES = np.frombuffer(b"ACGT", np.uint8)


def seed_for(config: int) -> int:
    return 20261200 + config


def random_bases(rng, n, gc=0.5):
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return BASES[rng.choice(4, size=n, p=p)]


def mutate(rng, seq: np.ndarray, sub: float, indel: float, max_indel=3) -> np.ndarray:
    """Substitutions (uniform among the other 3 bases) and indel events of
    length 1..max_indel (insertion or deletion with equal odds)."""
    s = seq.copy()
    m = rng.random(len(s)) < sub
    if m.any():
        shift = rng.integers(1, 4, size=m.sum())
        code = np.searchsorted(BASES, s[m])
        s[m] = BASES[(code + shift) % 4]
    if indel <= 0:
        return s
    ev = np.nonzero(rng.random(len(s)) < indel)[0]
    if len(ev) == 0:
        return s
    out, prev = [], 0
    for p in ev:
        out.append(s[prev:p])
        k = int(rng.integers(1, max_indel + 1))
        if rng.random() < 0.5:
            out.append(BASES[rng.integers(0, 4, size=k)])
            prev = p
        else:
            prev = min(len(s), p + k)
    out.append(s[prev:])
    return np.concatenate(out)


def revcomp_np(seq: np.ndarray) -> np.ndarray:
    lut = np.full(256, ord("A"), np.uint8)
    lut[ord("A")], lut[ord("C")], lut[ord("G")] = ord("T"), ord("G"), ord("C")
    return lut[seq[::-1]]


# ---------------------------------------------------------------- config 3
def c3_pairs(n_pairs=1_000_000, seed=None, qmin=36, qmax=250, dmin=20, dmax=200,
             sub=0.018, indel=0.002):
    """Batched uneven pairs: query length U{qmin..qmax}; region = query
    embedded at offset U{0..d} in q + d random bases, d ~ U{dmin..dmax}, the
    query carrying 1.8 % substitutions and 0.2 % 1-bp indels relative to the
    region.  Returns (queries, regions) as lists of bytes."""
    rng = np.random.default_rng(seed_for(3) if seed is None else seed)
    ql = rng.integers(qmin, qmax + 1, size=n_pairs)
    d = rng.integers(dmin, dmax + 1, size=n_pairs)
    off = (rng.random(n_pairs) * (d + 1)).astype(np.int64)
    tot = int((ql + d).sum())
    pool = BASES[rng.integers(0, 4, size=tot)]
    qs, rs = [], []
    pos = 0
    subs_all = rng.random(int(ql.sum()) + 16)
    sub_base = rng.integers(1, 4, size=len(subs_all))
    sp = 0
    for i in range(n_pairs):
        L, D, o = int(ql[i]), int(d[i]), int(off[i])
        reg = pool[pos:pos + L + D]
        pos += L + D
        q = reg[o:o + L].copy()
        m = subs_all[sp:sp + L] < sub
        if m.any():
            code = np.searchsorted(BASES, q[m])
            q[m] = BASES[(code + sub_base[sp:sp + L][m]) % 4]
        sp += L
        if indel > 0 and rng.random() < indel * L:
            p = int(rng.integers(0, L))
            if rng.random() < 0.5:
                q = np.concatenate([q[:p], BASES[rng.integers(0, 4, size=1)], q[p:]])[:L]
            else:
                q = np.concatenate([q[:p], q[p + 1:], reg[o + L:o + L + 1]])[:L]
        qs.append(q.tobytes())
        rs.append(reg.tobytes())
    return qs, rs


def c3_pairs_fast(n_pairs=1_000_000, seed=None, qmin=36, qmax=250, dmin=20, dmax=200,
                  sub=0.018, indel=0.002):
    """Vectorised generator of the same distribution for large n (returns
    packed arrays: qbuf, qoff, qlen, rbuf, roff, rlen)."""
    rng = np.random.default_rng(seed_for(3) if seed is None else seed)
    ql = rng.integers(qmin, qmax + 1, size=n_pairs).astype(np.int32)
    d = rng.integers(dmin, dmax + 1, size=n_pairs).astype(np.int32)
    rl = (ql + d).astype(np.int32)
    off = (rng.random(n_pairs) * (d + 1)).astype(np.int64)
    roff = np.zeros(n_pairs, np.int64)
    roff[1:] = np.cumsum(rl.astype(np.int64))[:-1]
    qoff = np.zeros(n_pairs, np.int64)
    qoff[1:] = np.cumsum(ql.astype(np.int64))[:-1]
    rbuf = BASES[rng.integers(0, 4, size=int(rl.sum()), dtype=np.uint8)]
    # query = region[off : off + ql] with substitutions
    idx = np.repeat(roff + off - qoff, ql) + np.arange(int(ql.sum()), dtype=np.int64)
    qbuf = rbuf[idx].copy()
    m = rng.random(len(qbuf)) < sub
    code = np.searchsorted(BASES, qbuf[m])
    qbuf[m] = BASES[(code + rng.integers(1, 4, size=m.sum())) % 4]
    # 1-bp indels: delete one base and shift the rest left within the query
    nind = rng.random(n_pairs) < indel * ql
    for i in np.nonzero(nind)[0]:
        a, L = int(qoff[i]), int(ql[i])
        p = int(rng.integers(0, L))
        seg = qbuf[a:a + L].copy()
        if rng.random() < 0.5:
            seg = np.concatenate([seg[:p], BASES[rng.integers(0, 4, size=1)], seg[p:-1]])
        else:
            nxt = rbuf[roff[i] + off[i] + L] if off[i] + L < rl[i] else BASES[0]
            seg = np.concatenate([seg[:p], seg[p + 1:], [nxt]])
        qbuf[a:a + L] = seg
    return (np.concatenate([qbuf, np.zeros(8, np.uint8)]), qoff, ql,
            np.concatenate([rbuf, np.zeros(8, np.uint8)]), roff, rl)


# ---------------------------------------------------------------- genomes
def genome_uniform(length, seed, gc=0.5):
    rng = np.random.default_rng(seed)
    return random_bases(rng, length, gc)


def genome_c2(total=50_000_000, nseq=24, seed=None, gc=0.41, tandem_frac=0.05, min_len=1_000_000):
    """24 sequences, GC 41 %, 5 % of bases in exact tandem arrays of period
    2..50."""
    rng = np.random.default_rng(seed_for(2) if seed is None else seed)
    cuts = np.sort(rng.choice(np.arange(1, total - nseq * min_len), nseq - 1, replace=False))
    lens = np.diff(np.concatenate([[0], cuts, [total - nseq * min_len]])) + min_len
    seqs = []
    for L in lens:
        s = random_bases(rng, int(L), gc)
        target = int(tandem_frac * L)
        done = 0
        while done < target:
            period = int(rng.integers(2, 51))
            copies = int(rng.integers(3, 40))
            span = min(period * copies, target - done + period)
            p = int(rng.integers(0, max(1, L - span)))
            unit = random_bases(rng, period, gc)
            s[p:p + span] = np.resize(unit, span)
            done += span
        seqs.append(s)
    return seqs


def sample_reads(rng, seqs, n, length, sub, indel, name0=0):
    """Reads of exactly `length` bp at uniform positions and strands."""
    lens = np.array([len(s) for s in seqs], np.int64)
    w = lens / lens.sum()
    which = rng.choice(len(seqs), size=n, p=w)
    reads = []
    truth = []
    for i in range(n):
        s = seqs[which[i]]
        p = int(rng.integers(0, len(s) - length - 10))
        frag = mutate(rng, s[p:p + length + 10], sub, indel)[:length]
        if len(frag) < length:
            frag = np.concatenate([frag, BASES[rng.integers(0, 4, size=length - len(frag))]])
        strand = int(rng.integers(0, 2))
        if strand:
            frag = revcomp_np(frag)
        reads.append(frag.tobytes())
        truth.append((int(which[i]), p, strand))
    return reads, truth


def c1_workload(seed=None, n_reads=1000, ref_len=1_000_000, read_len=100):
    rng = np.random.default_rng(seed_for(1) if seed is None else seed)
    ref = random_bases(rng, ref_len)
    reads, truth = sample_reads(rng, [ref], n_reads, read_len, 0.02, 0.002)
    return [ref.tobytes()], reads, truth


def c2_workload(seed=None, n_reads=100_000, total=50_000_000, read_len=100):
    rng = np.random.default_rng(seed_for(2) if seed is None else seed)
    seqs = genome_c2(total=total, seed=int(rng.integers(1 << 31)))
    reads, truth = sample_reads(rng, seqs, n_reads, read_len, 0.01, 0.001)
    nrand = n_reads // 20
    for _ in range(nrand):
        reads.append(random_bases(rng, read_len, 0.41).tobytes())
        truth.append((-1, -1, -1))
    return [s.tobytes() for s in seqs], reads, truth


def to_fasta(seqs, width=50, names=None) -> bytes:
    out = []
    for i, s in enumerate(seqs):
        out.append(b">" + (names[i] if names else b"chr%d" % (i + 1)) + b"\n")
        for p in range(0, len(s), width):
            out.append(s[p:p + width] + b"\n")
    return b"".join(out)


def reads_fasta(reads) -> bytes:
    return b"".join(b">r%09d\n%s\n" % (i, r) for i, r in enumerate(reads))


# ---------------------------------------------------------------- config 4
def genome_c4(total=3_100_000_000, nseq=25, seed=None, gc=0.41, repeat_frac=0.45, nfam=1000,
              chunk=1 << 27):
    """25 sequences totalling `total` bp (GC 41 %); 45 % of the bases are
    copies of repeat families (consensus length U[300, 6000]) at U[10 %, 20 %]
    substitution divergence, dropped at uniform positions.  Returns (packed
    uint8 buffer, int64 offsets, int64 lengths).  Generated in chunks so the
    peak memory stays near the genome's own size."""
    rng = np.random.default_rng(seed_for(4) if seed is None else seed)
    g = np.empty(total, np.uint8)
    lut = np.empty(1000, np.uint8)
    at = int(round(500 * (1 - gc)))
    lut[:at], lut[at:500] = ord("A"), ord("C")
    lut[500:1000 - at], lut[1000 - at:] = ord("G"), ord("T")
    for c0 in range(0, total, chunk):
        n = min(chunk, total - c0)
        g[c0:c0 + n] = lut[rng.integers(0, 1000, n, dtype=np.int16)]
    fams = []
    for _ in range(nfam):
        L = int(rng.integers(300, 6001))
        fams.append(lut[rng.integers(0, 1000, L, dtype=np.int16)])
    covered, target = 0, int(repeat_frac * total)
    while covered < target:
        f = fams[int(rng.integers(0, nfam))]
        div = rng.uniform(0.10, 0.20)
        cp = f.copy()
        m = rng.random(len(cp)) < div
        if m.any():
            code = np.searchsorted(BASES, cp[m])
            cp[m] = BASES[(code + rng.integers(1, 4, m.sum())) % 4]
        p = int(rng.integers(0, total - len(cp)))
        g[p:p + len(cp)] = cp
        covered += len(cp)
    cuts = np.sort(rng.choice(np.arange(1, nseq * 1000), nseq - 1, replace=False))
    w = np.diff(np.concatenate([[0], cuts, [nseq * 1000]])).astype(np.float64)
    lens = (w / w.sum() * total).astype(np.int64)
    lens[-1] += total - lens.sum()
    off = np.zeros(nseq, np.int64)
    off[1:] = np.cumsum(lens)[:-1]
    return g, off, lens


def sample_reads_fast(rng, genome, off, lens, n, length, sub):
    """n reads of `length` bp at uniform positions inside one sequence each,
    uniform strand, substitutions at rate `sub` (vectorised; no indels).
    Returns (packed uint8 buffer, int64 offsets, int32 lengths)."""
    w = (lens - length).astype(np.float64)
    which = rng.choice(len(lens), size=n, p=w / w.sum())
    pos = off[which] + (rng.random(n) * (lens[which] - length)).astype(np.int64)
    out = genome[pos[:, None] + np.arange(length)[None, :]]
    m = rng.random(out.shape) < sub
    if m.any():
        code = np.searchsorted(BASES, out[m])
        out[m] = BASES[(code + rng.integers(1, 4, m.sum())) % 4]
    rc = rng.random(n) < 0.5
    if rc.any():
        lut = np.full(256, ord("A"), np.uint8)
        lut[ord("A")], lut[ord("C")], lut[ord("G")] = ord("T"), ord("G"), ord("C")
        out[rc] = lut[out[rc][:, ::-1]]
    return (out.reshape(-1), np.arange(n, dtype=np.int64) * length,
            np.full(n, length, np.int32))


def sample_reads_var(rng, genome, off, lens, n, lmin, lmax, sub):
    """n reads of length U{lmin..lmax} at uniform positions inside one
    sequence each, uniform strand, substitutions at rate `sub` (vectorised;
    no indels).  Returns (packed uint8 buffer, int64 offsets, int32
    lengths)."""
    rl = rng.integers(lmin, lmax + 1, size=n).astype(np.int32)
    w = (lens - lmax).astype(np.float64)
    which = rng.choice(len(lens), size=n, p=w / w.sum())
    pos = off[which] + (rng.random(n) * (lens[which] - lmax)).astype(np.int64)
    out = genome[pos[:, None] + np.arange(lmax)[None, :]]
    m = rng.random(out.shape) < sub
    if m.any():
        code = np.searchsorted(BASES, out[m])
        out[m] = BASES[(code + rng.integers(1, 4, m.sum())) % 4]
    # reverse strand: the reverse complement of the read's own bases
    rc = rng.random(n) < 0.5
    if rc.any():
        lut = np.full(256, ord("A"), np.uint8)
        lut[ord("A")], lut[ord("C")], lut[ord("G")] = ord("T"), ord("G"), ord("C")
        idx = np.nonzero(rc)[0]
        cols = rl[idx][:, None] - 1 - np.arange(lmax)[None, :]
        flipped = np.take_along_axis(out[idx], np.maximum(cols, 0), axis=1)
        out[idx] = np.where(cols >= 0, lut[flipped], out[idx])
    keep = np.arange(lmax)[None, :] < rl[:, None]
    roff = np.zeros(n, np.int64)
    roff[1:] = np.cumsum(rl.astype(np.int64))[:-1]
    return out[keep], roff, rl


def c5_requests(rng, genome, off, lens, n_requests, max_reads=4096, lmin=50, lmax=300,
                err=0.015):
    """Serving requests of config 5: log-uniform 1..max_reads reads each, read
    lengths U{lmin..lmax}, substitutions at rate `err`, every read drawn
    afresh.  Returns a list of (buffer, offsets, lengths) per request."""
    sizes = np.exp(rng.uniform(0, np.log(max_reads), n_requests)).astype(int)
    sizes = np.clip(sizes, 1, max_reads)
    return [sample_reads_var(rng, genome, off, lens, int(k), lmin, lmax, err) for k in sizes]
