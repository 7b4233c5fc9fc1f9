#!/usr/bin/env python3
"""Benchmark of the phase-2 hot path on one MI355X.

Workload (config 3 of the benchmark plan): 1,000,000 query/region pairs,
query length U{36..250}, region = query + U{20..200} bases with the query
embedded after 2 % divergence, reference default scoring.  One step = one
swift_sw_batch pass over all pairs resident in HBM (device binning, packed
int16 forward sweep with int32 re-run of overflow suspects, traceback
walk), i.e. scores, coordinates and consensus strings for every pair.

Also measured on the same pairs: the plain per-pair drop-in
(alignSequences, one call per query length) and the reference CPU
smithWaterman_cpu on a bounded sample.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SW GCUPS + reads/s at 1 MI355X, batched vs per-pair; p50/p99 latency"
CELL_OPS = 16                 # algorithmic integer ops per DP cell
PEAK_INT16_TOPS = 78.6        # 256 CU x 64 lanes x 2 (packed) x 2.4 GHz
PEAK_INT32_TOPS = 39.3
# HBM bytes per DP cell of one swift_sw_batch forward pass on this workload,
# from rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) counts, each scaled by
# the same counter's reading for a device copy of known size
# (profiles/sw_batch_c3_pmc.txt)
TRAFFIC_B_PER_CELL = 0.6308


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baselines run on: its CPU model, its logical CPU
    count and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "host_cpus": os.cpu_count() or 1, "allowed_cpus": allowed}


# Host threads of the CPU baselines: the reference's own CPU path runs
# phase 1 on 16 threads and phase 2 on 8 (src/common.h:97, src/align.cu:32);
# 16 is also the CPU share a one-GPU box gives a process
CPU_THREADS = 16


def cpu_baseline(batch, qbuf, qoff, qlen, rbuf, roff, rlen, npairs, nstrings=20000,
                 ncalib=100000):
    """The reference's own smithWaterman_cpu (oracle/_ref, compiled from the
    reference sources) on the first `npairs` pairs with CPU_THREADS threads.
    Its results are compared with the batched GPU pass's on the same pairs
    (scores, coordinates and lengths of every pair, consensus strings of the
    first `nstrings`), and the oracle port is timed beside it on the first
    `ncalib` pairs (the port-to-reference factor that scales the port-timed
    pipeline baselines).  Falls back to the port when oracle/_ref is absent."""
    from oracle import oracle as O
    n = min(npairs, len(qlen))
    threads = max(1, min(CPU_THREADS, host_info()["allowed_cpus"]))
    cells = float((qlen[:n].astype(np.int64) * rlen[:n]).sum())
    R = O.ref()
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    qo = qoff[:n].astype(np.int64).copy()
    ro = roff[:n].astype(np.int64).copy()
    ql = qlen[:n].astype(np.int32).copy()
    rl = rlen[:n].astype(np.int32).copy()
    res = {"unit": "GCUPS", "cores": threads, "threads": threads, **host_info()}
    if R is None:
        qs = [qbuf[qoff[i]:qoff[i] + qlen[i]].tobytes() for i in range(n)]
        rs = [rbuf[roff[i]:roff[i] + rlen[i]].tobytes() for i in range(n)]
        t0 = time.perf_counter()
        O.sw_batch(qs, rs, nthreads=threads)
        dt = time.perf_counter() - t0
        res.update({"value": round(cells / dt / 1e9, 4), "kind": "port",
                    "sample": f"first {n} pairs of the workload ({cells:.3e} cells), "
                              f"{threads} threads, {dt:.1f} s"})
        return res
    ns = min(nstrings, n)
    cstride = ql[:ns].astype(np.int64) + rl[:ns] + 2
    co = np.zeros(ns, np.int64)
    co[1:] = np.cumsum(cstride)[:-1]
    rq = np.zeros(int(cstride.sum()) + 1, np.uint8)
    rr = np.zeros(int(cstride.sum()) + 1, np.uint8)
    sc, st, en, al = (np.zeros(n, np.int32) for _ in range(4))
    t0 = time.perf_counter()
    R.ref_sw_batch_mt2(n, p(qbuf, C.c_char), p(qo, C.c_long), p(ql, C.c_int),
                       p(rbuf, C.c_char), p(ro, C.c_long), p(rl, C.c_int), C.c_float(2),
                       C.c_float(-1), C.c_float(-10), C.c_float(-1), p(sc, C.c_int),
                       p(st, C.c_int), p(en, C.c_int), p(al, C.c_int), p(rq, C.c_char),
                       p(rr, C.c_char), p(co, C.c_long), ns, threads)
    dt = time.perf_counter() - t0
    res.update({"value": round(cells / dt / 1e9, 4), "kind": "reference",
                "sample": f"first {n} pairs of the workload ({cells:.3e} cells), "
                          f"{threads} threads, {dt:.1f} s"})
    # the GPU pass's results for the same pairs
    g = batch.results(strings=False)
    same = (np.array_equal(g.score[:n], sc) and np.array_equal(g.start[:n], st) and
            np.array_equal(g.end[:n], en) and np.array_equal(g.length[:n], al))
    gq = batch.qc.download(np.uint8, batch.cons_bytes)
    gr = batch.rc.download(np.uint8, batch.cons_bytes)
    sdiff = 0
    for i in range(ns):
        a, k, b = int(batch.coff_h[i]), int(al[i]), int(co[i])
        if gq[a:a + k].tobytes() != rq[b:b + k].tobytes() or \
                gr[a:a + k].tobytes() != rr[b:b + k].tobytes():
            sdiff += 1
    res["parity_vs_reference"] = {
        "pairs": n, "fields": "score, start, end, length", "identical": bool(same),
        "string_pairs": ns, "strings_identical": sdiff == 0}
    # port-to-reference factor on one sample, same threads
    m = min(ncalib, n)
    mcells = float((ql[:m].astype(np.int64) * rl[:m]).sum())
    qs = [qbuf[qoff[i]:qoff[i] + qlen[i]].tobytes() for i in range(m)]
    rs = [rbuf[roff[i]:roff[i] + rlen[i]].tobytes() for i in range(m)]
    t0 = time.perf_counter()
    O.sw_batch(qs, rs, nthreads=threads)
    tp = time.perf_counter() - t0
    scm = np.zeros(m, np.int32)
    t0 = time.perf_counter()
    R.ref_sw_batch_mt(m, p(qbuf, C.c_char), p(qo, C.c_long), p(ql, C.c_int),
                      p(rbuf, C.c_char), p(ro, C.c_long), p(rl, C.c_int), C.c_float(2),
                      C.c_float(-1), C.c_float(-10), C.c_float(-1), p(scm, C.c_int), threads)
    tr = time.perf_counter() - t0
    res["calibration"] = {"pairs": m, "threads": threads,
                          "port_gcups": round(mcells / tp / 1e9, 4),
                          "reference_gcups": round(mcells / tr / 1e9, 4),
                          "reference_over_port": round(tp / tr, 4)}
    return res


def plain_pass_setup(qbuf, qoff, qlen, rbuf, roff, rlen):
    """Reference-layout buffers for the plain drop-in: one alignSequences
    call per query length, stride qlen+1 / rmax+1 (NUL padded)."""
    from swift_amd.device import DeviceBuffer
    groups = []
    for q in np.unique(qlen):
        idx = np.nonzero(qlen == q)[0]
        q = int(q)
        rmax = int(rlen[idx].max())
        n = len(idx)
        qa = np.zeros((n, q + 1), np.uint8)
        ra = np.zeros((n, rmax + 1), np.uint8)
        qa[:, :q] = qbuf[(qoff[idx][:, None] + np.arange(q)[None, :])]
        cols = np.arange(rmax)[None, :]
        mask = cols < rlen[idx][:, None]
        src = np.minimum(roff[idx][:, None] + cols, len(rbuf) - 1)
        ra[:, :rmax] = np.where(mask, rbuf[src], 0)
        cl = q + rmax + 2
        bufs = dict(q=DeviceBuffer.from_array(qa), r=DeviceBuffer.from_array(ra),
                    sc=DeviceBuffer(4 * n), st=DeviceBuffer(4 * n), en=DeviceBuffer(4 * n),
                    al=DeviceBuffer(4 * n), qc=DeviceBuffer(n * cl), rc=DeviceBuffer(n * cl))
        groups.append((q, rmax, n, bufs))
    return groups


def plain_pass(groups):
    from swift_amd._lib import check, lib
    for q, rmax, n, b in groups:
        check(lib().swift_align_sequences(
            C.c_void_p(b["q"].ptr), C.c_void_p(b["r"].ptr), q, rmax, C.c_float(2), C.c_float(-1),
            C.c_float(-10), C.c_float(-1), C.c_void_p(b["sc"].ptr), C.c_void_p(b["st"].ptr),
            C.c_void_p(b["en"].ptr), C.c_void_p(b["al"].ptr), C.c_void_p(b["qc"].ptr),
            C.c_void_p(b["rc"].ptr), n, None))


HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
K = 12                       # seed length of every pipeline config


def phase1_bytes(read_lens, nhits, nbest):
    """Algorithmic HBM bytes of phase-1 mapping, per read strand: a bucket
    start and count per k-mer offset (8 B), refIdx + refPos per hit (5 B), and
    the best-hit outputs (9 B each)."""
    lens = np.repeat(np.asarray(read_lens, np.int64), 2)
    return float((8 * np.maximum(lens - K + 1, 0) + 5 * np.asarray(nhits, np.int64) +
                  9 * np.asarray(nbest, np.int64)).sum())


def record_cells(recs, read_lens, seq_lens):
    """Smith-Waterman cells of pipeline records: read length x window
    length, the window being 300 bases from max(0, refPos - 100) clipped at
    the sequence end (these genomes have no N to trim)."""
    c = recs.cols
    if recs.n == 0:
        return 0.0
    wl = np.minimum(300, np.asarray(seq_lens, np.int64)[c["ref_idx"]] - c["ref_distance"])
    return float((np.asarray(read_lens, np.int64)[c["read"]] * np.maximum(wl, 0)).sum())


def pipeline_roofline(p1_bytes, cells, secs):
    """The two-phase pipeline against its roofline: phase 1's algorithmic
    bytes at the HBM peak plus the SW cells' integer ops at the packed-int16
    VALU peak; frac = that ideal time / the measured time."""
    t_hbm = p1_bytes / (HBM_PEAK_GBS * 1e9)
    t_valu = CELL_OPS * cells / (PEAK_INT16_TOPS * 1e12)
    return {"bound": "hbm (phase 1) + valu (phase 2)", "phase1_bytes": round(p1_bytes),
            "sw_cells": round(cells), "ideal_ms": round((t_hbm + t_valu) * 1e3, 4),
            "measured_ms": round(secs * 1e3, 4), "frac": round((t_hbm + t_valu) / secs, 4),
            "phase1_hbm_ms": round(t_hbm * 1e3, 4), "sw_valu_ms": round(t_valu * 1e3, 4)}


def pipeline_parity(ix, reads, want, mh):
    """The device pipeline on `reads` against the oracle pipeline's records
    `want`, record by record, strings included."""
    got = ix.align_reads(reads, mh)
    keys = (("read", "read"), ("strand", "strand"), ("ref_idx", "refIdx"), ("shift", "shift"),
            ("ref_pos", "refPos"), ("nbiggest", "nBiggest"), ("ref_distance", "refDistance"),
            ("score", "score"), ("start", "start"), ("end", "end"), ("alen", "len"),
            ("keep", "keep"), ("ntied", "nTied"), ("qry_cons", "qry_cons"),
            ("ref_cons", "ref_cons"))
    same = got.n == len(want) and all(
        got.record(i)[a] == w[b] for i, w in enumerate(want) for a, b in keys)
    return {"reads": len(reads), "records": len(want), "identical": bool(same)}


def oracle_pipeline_baseline(oix, reads, mh, what):
    """The oracle's CPU pipeline (searchQueries3 -> alignQueries_cpu_mt
    restated) over `reads` on CPU_THREADS threads, as the reference's CPU
    path runs on 16 and 8; returns (baseline dict, records)."""
    from oracle import oracle as O
    threads = max(1, min(CPU_THREADS, host_info()["allowed_cpus"]))
    t0 = time.perf_counter()
    want = O.pipeline_mt(oix, reads, threads, mh)
    dt = time.perf_counter() - t0
    return ({"value": round(len(reads) / dt, 1), "unit": "reads/s", "cores": threads,
             "threads": threads, "kind": "port", **host_info(),
             "sample": f"{what}: {len(reads)} reads, oracle pipeline on {threads} threads "
                       "(index build not included); scale by cpu_baseline.calibration."
                       "reference_over_port for the reference's own speed"}, want)


def c1_leg(steps, cpu):
    """The small config (1 Mbp uniform reference, 1,000 reads of 100 bp at 2 %
    substitutions and 0.2 % indels, k=12, thr 1000, maxHits 3): reads/s of the
    whole device pipeline with its roofline fraction, phase 2 through the
    plain per-pair path (plain_phase2; phase 1 timed by swift_map_reads into
    preallocated outputs), and the oracle's CPU pipeline over the same reads
    timed beside it and compared record by record."""
    from swift_amd import search, synth
    from swift_amd.device import synchronize
    seqs, reads, _ = synth.c1_workload()
    t0 = time.perf_counter()
    ix = search.Index.from_sequences(seqs, 12, 1000)
    build_s = time.perf_counter() - t0
    rb = search.ReadBatch(reads)
    mh = 3
    out = search.RecordBuffers(2 * rb.n * mh, rb.max_len)
    ix.align_batch(rb, out, mh)
    synchronize()
    reps = max(steps, 1) * 4
    t0 = time.perf_counter()
    for _ in range(reps):
        ix.align_batch(rb, out, mh)
    synchronize()
    pipe_s = (time.perf_counter() - t0) / reps
    recs = out.results(strings=True)
    hits = ix.map_batch(rb, mh)
    # phase 1 alone into preallocated outputs (swift_map_reads synchronises)
    from swift_amd._lib import check, lib
    from swift_amd.device import DeviceBuffer
    ns = 2 * rb.n
    mbufs = [DeviceBuffer(ns * mh), DeviceBuffer(4 * ns * mh), DeviceBuffer(4 * ns * mh),
             DeviceBuffer(4 * ns)]
    t0 = time.perf_counter()
    for _ in range(reps):
        check(lib().swift_map_reads(ix._h, rb.seq.cptr, rb.off.cptr, rb.len.cptr, rb.n,
                                    rb.max_len, mh, *[b.cptr for b in mbufs], None, None,
                                    None))
    map_s = (time.perf_counter() - t0) / reps
    lens = [len(r) for r in reads]
    res = {"workload": "C1: 1 Mbp uniform reference, 1000 reads of 100 bp (2 % subs, 0.2 % "
                       "indels), k=12, thr 1000, maxHits 3",
           "index_build_s": round(build_s, 3),
           "pipeline_reads_per_s": round(len(reads) / pipe_s, 1),
           "pipeline_ms": round(pipe_s * 1e3, 3), "records": recs.n,
           "roofline": pipeline_roofline(phase1_bytes(lens, hits.nhits, hits.nbest),
                                         record_cells(recs, lens, [len(s) for s in seqs]),
                                         pipe_s),
           "plain_phase2": plain_phase2(reads, seqs, recs, map_s, reps)}
    if cpu:
        from oracle import oracle as O
        oix = O.Index(seqs, 12, 1000)
        res["cpu_baseline"], want = oracle_pipeline_baseline(oix, reads, mh, "all C1 reads")
        res["parity_vs_oracle"] = pipeline_parity(ix, reads, want, mh)
    return res


def c2_leg(n_reads, steps, cpu_reads):
    """Phase 1 + phase 2 over the C2 workload: 50 Mbp / 24 sequences, k=12,
    tuple threshold 1000, maxHits 3.  Reports the device index build time,
    phase-1 reads/s (both strands mapped) against the HBM roofline, reads/s
    of the whole device pipeline (map -> 300-bp windows -> batched SW ->
    best-hit choice) against its roofline, phase 2 through the plain
    per-pair path (plain_phase2), and the oracle's CPU pipeline timed and
    compared on a sample."""
    from swift_amd import search, synth
    from swift_amd._lib import check, lib
    from swift_amd.device import DeviceBuffer, synchronize
    t0 = time.perf_counter()
    seqs, reads, _ = synth.c2_workload(n_reads=n_reads)
    log(f"[bench] C2 workload: {sum(map(len, seqs)):.3e} bp, {len(reads)} reads "
        f"({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    ix = search.Index.from_sequences(seqs, 12, 1000)
    build_s = time.perf_counter() - t0
    info = ix.info()
    rb = search.ReadBatch(reads)
    ns, mh = 2 * rb.n, 3
    ref, sh, pos = DeviceBuffer(ns * mh), DeviceBuffer(4 * ns * mh), DeviceBuffer(4 * ns * mh)
    nb, nbig, nh = (DeviceBuffer(4 * ns) for _ in range(3))

    def map_once():
        check(lib().swift_map_reads(ix._h, rb.seq.cptr, rb.off.cptr, rb.len.cptr, rb.n,
                                    rb.max_len, mh, ref.cptr, sh.cptr, pos.cptr, nb.cptr,
                                    nbig.cptr, nh.cptr, None))

    map_once()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        map_once()
    synchronize()
    map_s = (time.perf_counter() - t0) / steps
    hits = nh.download(np.int32, ns)
    best = nb.download(np.int32, ns)
    lens = [len(r) for r in reads]
    p1 = phase1_bytes(lens, hits, best)
    out = search.RecordBuffers(2 * rb.n * mh, rb.max_len)
    ix.align_batch(rb, out, mh)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ix.align_batch(rb, out, mh)
    synchronize()
    pipe_s = (time.perf_counter() - t0) / steps
    recs = out.results(strings=True)
    res = {"workload": "C2: 50 Mbp / 24 sequences (GC 41 %, 5 % tandem repeats), "
                       f"{len(reads)} reads of 100 bp (5 % random), k=12, thr 1000, maxHits 3",
           "index_build_s": round(build_s, 3), "kept_tuples": info["total"],
           "phase1_reads_per_s": round(len(reads) / map_s, 1),
           "phase1_ms": round(map_s * 1e3, 3),
           "phase1_roofline": {"bound": "hbm", "achieved": round(p1 / map_s / 1e9, 2),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(p1 / map_s / 1e9 / HBM_PEAK_GBS, 5),
                               "bytes": round(p1)},
           "pipeline_reads_per_s": round(len(reads) / pipe_s, 1),
           "pipeline_ms": round(pipe_s * 1e3, 3), "records": recs.n,
           "pipeline_roofline": pipeline_roofline(p1, record_cells(recs, lens,
                                                                  [len(x) for x in seqs]),
                                                  pipe_s),
           "hits_per_strand": round(float(hits.mean()), 2)}
    res["plain_phase2"] = plain_phase2(reads, seqs, recs, map_s, steps)
    log(f"[bench] C2 plain phase 2: {res['plain_phase2']}")
    if cpu_reads > 0:
        from oracle import oracle as O
        # the first reads and the fully random 5 % at the end
        nrand = min(cpu_reads // 20, len(reads) - n_reads)
        sample = reads[:cpu_reads - nrand] + reads[len(reads) - nrand:]
        oix = O.Index(seqs, 12, 1000)
        res["cpu_baseline"], want = oracle_pipeline_baseline(
            oix, sample, mh, f"first {cpu_reads - nrand} C2 reads + {nrand} random reads")
        res["parity_vs_oracle"] = pipeline_parity(ix, sample, want, mh)
    return res


def plain_phase2(reads, seqs, recs, phase1_s, steps):
    """Phase 2 of C1 and C2 through the plain per-pair path, as the configs
    name it:
    every pipeline record's (query, window) pair -- the read or its reverse
    complement (A->T, C->G, G->C, any other byte -> A), and up to 300 bases
    from max(0, refPos - 100) with leading N stripped, else trailing N --
    through swift_align_sequences, one call per read length (stride qlen + 1
    and 301, NUL padded), inputs resident in HBM.  Reports the plain path's
    GCUPS, the pipeline rate with it (phase-1 time + plain SW time; the
    windows are cut on the host here, untimed), and whether its scores,
    coordinates and strings equal the batched pipeline's records."""
    from swift_amd._lib import check, lib
    from swift_amd.device import DeviceBuffer, synchronize
    n = recs.n
    c = recs.cols
    lut = np.full(256, ord("A"), np.uint8)
    lut[ord("A")], lut[ord("C")], lut[ord("G")] = ord("T"), ord("G"), ord("C")
    rlen_max = 300
    win = np.zeros((n, rlen_max + 1), np.uint8)
    wl = np.zeros(n, np.int64)
    arrs = [np.frombuffer(x, np.uint8) for x in seqs]
    for i in range(n):
        sq = arrs[c["ref_idx"][i]]
        rd = int(c["ref_distance"][i])
        w = sq[rd:rd + rlen_max]
        if len(w) and (w[0] in (78, 110) or w[-1] in (78, 110)):  # N / n at an end
            isn = (w == 78) | (w == 110)
            lead = int(np.argmin(isn)) if not isn.all() else len(w)
            if lead:
                w = w[lead:]
            else:
                keep = len(w) - int(np.argmin(isn[::-1]))
                w = w[:keep]
        win[i, :len(w)] = w
        wl[i] = len(w)
    qlens = np.array([len(reads[r]) for r in c["read"]], np.int64)
    groups, cells = [], 0.0
    for q in np.unique(qlens):
        idx = np.nonzero(qlens == q)[0]
        q = int(q)
        qa = np.zeros((len(idx), q + 1), np.uint8)
        for k, i in enumerate(idx):
            rdx = np.frombuffer(reads[c["read"][i]], np.uint8)
            qa[k, :q] = lut[rdx[::-1]] if c["strand"][i] else rdx
        cl = q + rlen_max + 2
        m = len(idx)
        b = dict(q=DeviceBuffer.from_array(qa), r=DeviceBuffer.from_array(win[idx]),
                 sc=DeviceBuffer(4 * m), st=DeviceBuffer(4 * m), en=DeviceBuffer(4 * m),
                 al=DeviceBuffer(4 * m), qc=DeviceBuffer(m * cl), rc=DeviceBuffer(m * cl))
        groups.append((q, idx, m, cl, b))
        cells += float(q * wl[idx].sum())

    def run():
        for q, idx, m, cl, b in groups:
            check(lib().swift_align_sequences(
                C.c_void_p(b["q"].ptr), C.c_void_p(b["r"].ptr), q, rlen_max, C.c_float(2),
                C.c_float(-1), C.c_float(-10), C.c_float(-1), C.c_void_p(b["sc"].ptr),
                C.c_void_p(b["st"].ptr), C.c_void_p(b["en"].ptr), C.c_void_p(b["al"].ptr),
                C.c_void_p(b["qc"].ptr), C.c_void_p(b["rc"].ptr), m, None))

    run()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    synchronize()
    sw_s = (time.perf_counter() - t0) / steps
    same, sdiff = True, 0
    for q, idx, m, cl, b in groups:
        sc = b["sc"].download(np.float32, m).astype(np.int64)
        st, en, al = (b[k].download(np.int32, m) for k in ("st", "en", "al"))
        same &= bool(np.array_equal(sc, np.asarray(c["score"])[idx]) and
                     np.array_equal(st, np.asarray(c["start"])[idx]) and
                     np.array_equal(en, np.asarray(c["end"])[idx]) and
                     np.array_equal(al, np.asarray(c["alen"])[idx]))
        qc = b["qc"].download(np.uint8, m * cl)
        rc = b["rc"].download(np.uint8, m * cl)
        for k, i in enumerate(idx):
            a = int(al[k])
            if (qc[k * cl:k * cl + a].tobytes() != recs.qry_cons[i] or
                    rc[k * cl:k * cl + a].tobytes() != recs.ref_cons[i]):
                sdiff += 1
    return {"pairs": n, "calls": len(groups), "cells": cells,
            "gcups": round(cells / sw_s / 1e9, 2), "sw_ms": round(sw_s * 1e3, 3),
            "pipeline_reads_per_s": round(len(reads) / (phase1_s + sw_s), 1),
            "identical_to_batched": {"fields": same, "strings": sdiff == 0}}


def c5_leg(index, g, off, lens, n_requests, seed=20261205, p1_per_kmer_read=None):
    """Serving: requests of log-uniform 1..4096 reads (lengths U{50..300},
    1.5 % substitutions, every read drawn afresh; synth.c5_requests)
    submitted to a swift_server with the index resident.  First every
    request at once (capacity, reads/s), then Poisson arrivals at the fixed
    rate C5_RATE (the headline p50/p99) and at 70 % of the measured
    capacity; latency = submit -> records ready, measured by the server."""
    from swift_amd import synth
    rng = np.random.default_rng(seed)
    packed = synth.c5_requests(rng, g, off, lens, n_requests)
    sizes = np.array([len(p[2]) for p in packed])
    total_reads = int(sizes.sum())
    log(f"[bench] C5: {n_requests} requests, {total_reads} reads")
    res = {"workload": f"C5 serving: {n_requests} requests of log-uniform 1..4096 reads, "
                       "lengths U{50..300}, 1.5 % substitutions, index resident",
           "requests": n_requests, "reads": total_reads}
    # serving runs with the library defaults, not with what the C3 autotuner
    # chose for one 1M-pair batch (the settings are process-wide); the server
    # keeps its own 8 persistent waves per CU either way
    from swift_amd import align
    saved = align.get_tuning()
    align.set_tuning(0, True, True)
    align.set_coarse_rows(-1)
    try:
        res = _c5_serve(index, packed, sizes, total_reads, n_requests, rng, res, lens)
    finally:
        align.set_tuning(saved["waves_per_cu"], saved["allow_int16"], saved["walk_deferred"])
        align.set_coarse_rows(saved["coarse_rows"])
    if p1_per_kmer_read is not None:
        # phase-1 bytes per read scale with its k-mer count (L - k + 1)
        kmers = float(sum(np.maximum(p[2].astype(np.int64) - K + 1, 0).sum() for p in packed))
        rf = pipeline_roofline(p1_per_kmer_read * kmers, res.pop("capacity_cells"),
                               res["capacity_s"])
        rf["note"] = ("capacity run (every request at once); SW cells from the served "
                      "records, phase-1 bytes per k-mer from the C4 stream")
        res["capacity_roofline"] = rf
    return res


# the fixed arrival rate of the serving leg's headline latencies (reads/s)
C5_RATE = 1_000_000
C5_SLOTS = 4  # server slots of the serving leg (--c5-slots)


def _c5_serve(index, packed, sizes, total_reads, n_requests, rng, res, seq_lens):
    import threading
    from swift_amd.serve import Server
    with Server(index, nslots=C5_SLOTS, mem_budget=48 << 30, max_batch_reads=8192) as srv:
        for p in packed[:8]:  # warm the slots
            srv.result(srv.submit_packed(*p), strings=False)
        t0 = time.perf_counter()
        ids = [srv.submit_packed(*p) for p in packed]
        recs = [srv.result(i, strings=False)[0] for i in ids]
        cap_s = time.perf_counter() - t0
        res["capacity_cells"] = sum(record_cells(r, p[2], seq_lens)
                                    for r, p in zip(recs, packed))
        capacity = total_reads / cap_s
        mean_size = total_reads / n_requests

        def poisson_run(lam):
            """Poisson arrivals of the requests at lam reads/s; latency =
            submit -> records ready, measured by the server."""
            rate = lam / mean_size                      # requests per second
            gaps = rng.exponential(1.0 / rate, n_requests)
            lat = [0.0] * n_requests
            ids = [None] * n_requests
            ready = threading.Semaphore(0)

            def collect():
                for k in range(n_requests):
                    ready.acquire()
                    _, lat[k] = srv.result(ids[k], strings=False)

            th = threading.Thread(target=collect)
            th.start()
            b0 = srv.stats()["batches"]
            t0 = time.perf_counter()
            due = t0
            for k in range(n_requests):
                due += gaps[k]
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[k] = srv.submit_packed(*packed[k])
                ready.release()
            th.join()
            run_s = time.perf_counter() - t0
            st = srv.stats()
            lat = np.array(lat)
            return {"arrival_rate_reads_per_s": round(lam, 1),
                    "arrival_rate_requests_per_s": round(rate, 2),
                    "served_reads_per_s": round(total_reads / run_s, 1),
                    "p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "p99_ms": round(float(np.percentile(lat, 99)), 3),
                    "mean_ms": round(float(lat.mean()), 3),
                    "batches": st["batches"] - b0}

        fixed = poisson_run(float(C5_RATE))
        loaded = poisson_run(0.7 * capacity)
        st = srv.stats()
    res.update({"capacity_reads_per_s": round(capacity, 1), "capacity_s": round(cap_s, 4),
                **fixed, "at_70pct_capacity": loaded,
                "max_coalesced": st["max_coalesced"], "slots": C5_SLOTS})
    return res


def c4_graph(ix, g, off, lens, rng, n_reads, batch, depth=2):
    """The C4 pipeline as hipGraph replays: `depth` graphs of batch // depth
    read slots of 150 bases, each on its own stream, take the batches in
    turn.  A graph's next batch is uploaded (timed, as for the eager
    batches) once its previous replay has finished, while the other graphs'
    replays run, so one batch's phase 1 overlaps the phase 2 of the batch
    before.  The reads are drawn before the timed region.  A 20,000-read
    batch is compared record by record, strings included, with the eager
    pipeline."""
    from swift_amd import search, synth
    from swift_amd.device import Stream, synchronize
    slots = batch // depth
    pgs = [search.PipelineGraph(ix, slots, 150, 3) for _ in range(depth)]
    sts = [Stream() for _ in range(depth)]
    work, done = [], 0
    while done < n_reads:
        nb = min(slots, n_reads - done)
        buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, nb, 150, 0.01)
        slot_len = np.zeros(slots, np.int32)
        slot_len[:nb] = rl
        work.append((buf, slot_len))
        done += nb
    nrec, busy = 0, [False] * depth

    def retire(s):
        nonlocal nrec
        sts[s].synchronize()
        pgs[s].check()
        nrec += int(pgs[s].out.bufs["nrecords"].download(np.int32, 1)[0])
        busy[s] = False

    synchronize()
    t0 = time.perf_counter()
    for b, (buf, slot_len) in enumerate(work):
        s = b % depth
        if busy[s]:
            retire(s)
        pgs[s].seq.upload(buf)
        pgs[s].len.upload(slot_len)
        pgs[s].launch(sts[s].cptr)
        busy[s] = True
        if (b + 1) % (16 * depth) == 0:
            log(f"[bench] C4 graph {(b + 1) * slots}/{n_reads} reads, "
                f"{time.perf_counter() - t0:.2f} s on the device path")
    for s in range(depth):
        if busy[s]:
            retire(s)
    dt = time.perf_counter() - t0
    log(f"[bench] C4 graph {n_reads} reads, {dt:.2f} s on the device path")
    del work
    buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, 20000, 150, 0.01)
    reads = [buf[o:o + 150].tobytes() for o in ro]
    a = pgs[0].run(reads)
    b = ix.align_reads(reads, 3)
    same = a.n == b.n and all(a.record(i) == b.record(i) for i in range(a.n))
    for pg in pgs:
        pg.close()
    for st in sts:
        st.close()
    return {"reads_per_s": round(n_reads / dt, 1), "records": nrec, "seconds": round(dt, 3),
            "graphs": depth, "batch_reads": slots,
            "vs_eager": {"reads": len(reads), "records": a.n, "identical": bool(same)}}


def c4_leg(n_reads, batch, parity_reads, c5_requests=0, graph=True):
    """Genome scale: 3.1 Gbp / 25 sequences (GC 41 %, 45 % interspersed
    repeats), k=12 index resident in HBM, `n_reads` reads of 150 bp at 1 %
    substitutions streamed through the device pipeline in batches (upload of
    each batch included), eager and as hipGraph replays, each with its
    roofline fraction; `parity_reads` reads compared with the oracle
    pipeline over the same genome; then the C5 serving leg on this resident
    index."""
    from swift_amd import search, synth
    from swift_amd.device import synchronize
    t0 = time.perf_counter()
    g, off, lens = synth.genome_c4()
    log(f"[bench] C4 genome {len(g):.3e} bp in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    ix = search.Index.from_packed(g, off, lens, 12, 1000)
    build_s = time.perf_counter() - t0
    info = ix.info()
    log(f"[bench] C4 index: {info} in {build_s:.2f} s")
    rng = np.random.default_rng(20261204)
    nb0 = min(batch, n_reads)
    buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, nb0, 150, 0.01)
    rb0 = search.ReadBatch.from_packed(buf, ro, rl, 150)
    sh = ix.map_batch(rb0, 3)
    hits = sh.nhits
    hit_stats = {"mean": round(float(hits.mean()), 1), "p50": int(np.percentile(hits, 50)),
                 "p99": int(np.percentile(hits, 99)), "max": int(hits.max()),
                 "frac_over_1024": round(float((hits > 1024).mean()), 4),
                 "frac_over_16384": round(float((hits > 16384).mean()), 5)}
    log(f"[bench] C4 hits per strand: {hit_stats}")
    # per-read algorithmic phase-1 bytes and per-record SW cells, from this
    # first batch (the stream's reads are drawn alike)
    p1_per_read = phase1_bytes(rl, hits, sh.nbest) / nb0
    out = search.RecordBuffers(2 * batch * 3, 150)
    ix.align_batch(rb0, out, 3)
    r0 = out.results(strings=False)
    cells_per_rec = record_cells(r0, rl, lens) / max(r0.n, 1)
    dt, nrec, done = 0.0, 0, 0
    while done < n_reads:
        nb = min(batch, n_reads - done)
        buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, nb, 150, 0.01)
        synchronize()
        t0 = time.perf_counter()
        rb = search.ReadBatch.from_packed(buf, ro, rl, 150)
        ix.align_batch(rb, out, 3)
        dt += time.perf_counter() - t0
        nrec += int(out.bufs["nrecords"].download(np.int32, 1)[0])
        done += nb
        if done % (8 * batch) == 0 or done == n_reads:
            log(f"[bench] C4 {done}/{n_reads} reads, {dt:.2f} s on the device path")
    res = {"workload": "C4: 3.1 Gbp / 25 sequences (GC 41 %, 45 % repeats of 300-6000 bp at "
                       f"10-20 % divergence), {n_reads} reads of 150 bp, 1 % substitutions, "
                       "k=12, thr 1000, maxHits 3",
           "index_build_s": round(build_s, 2), "kept_tuples": info["total"],
           "largest_bucket": info["max_bucket"], "reads_per_s": round(n_reads / dt, 1),
           "seconds": round(dt, 3), "records": nrec, "batch_reads": batch,
           "hits_per_strand": hit_stats,
           "roofline": pipeline_roofline(p1_per_read * n_reads, cells_per_rec * nrec, dt)}
    res["roofline"]["note"] = ("phase-1 bytes per read and SW cells per record from the "
                               "first batch, scaled to the stream")
    if graph:
        res["graph"] = c4_graph(ix, g, off, lens, rng, n_reads, batch)
        res["graph"]["roofline"] = pipeline_roofline(
            p1_per_read * n_reads, cells_per_rec * res["graph"]["records"],
            res["graph"]["seconds"])
    if parity_reads > 0:
        from oracle import oracle as O
        buf, ro, rl = synth.sample_reads_fast(rng, g, off, lens, parity_reads, 150, 0.01)
        reads = [buf[o:o + 150].tobytes() for o in ro]
        t0 = time.perf_counter()
        oix = O.Index([g[o:o + n].tobytes() for o, n in zip(off, lens)], 12, 1000)
        log(f"[bench] C4 oracle index in {time.perf_counter() - t0:.1f} s")
        base, want = oracle_pipeline_baseline(oix, reads, 3, "C4 reads")
        res["cpu_baseline"] = base
        res["parity_vs_oracle"] = pipeline_parity(ix, reads, want, 3)
        del oix
    if c5_requests > 0:
        res["c5"] = c5_leg(ix, g, off, lens, c5_requests,
                           p1_per_kmer_read=p1_per_read / (150 - K + 1))
        res["c5"]["index"] = "C4 index (3.1 Gbp / 25 sequences, k=12) resident"
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pairs", type=int, default=800000)
    ap.add_argument("--c2-reads", type=int, default=100000)
    ap.add_argument("--c2-cpu-reads", type=int, default=16000,
                    help="C2 reads (95 %% mapped, 5 %% random) the oracle pipeline times and "
                         "checks")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the C4 hipGraph stream")
    ap.add_argument("--c5-requests", type=int, default=1500)
    ap.add_argument("--c5-slots", type=int, default=4, help="server slots of the serving leg")
    ap.add_argument("--c4-only", action="store_true", help="only the genome-scale C4 leg "
                    "(and C5 on its index)")
    ap.add_argument("--c4-reads", type=int, default=10_000_000)
    ap.add_argument("--c4-batch", type=int, default=250_000)
    ap.add_argument("--c4-parity", type=int, default=300, metavar="READS")
    ap.add_argument("--c2-only", action="store_true", help="only the C2 pipeline leg (no JSON "
                    "bench line contract; for profiling)")
    ap.add_argument("--autotune", type=int, default=-1, metavar="PAIRS",
                    help="autotune the batched path on this many workload pairs first "
                         "(default -1: the whole workload; 0: off)")
    ap.add_argument("--walk", choices=["default", "fused", "deferred"], default="default",
                    help="traceback walk mode (default: the library's tuned setting)")
    ap.add_argument("--waves", type=int, default=-1, metavar="N",
                    help="persistent waves per CU of the row-class kernels (0: occupancy "
                         "limit; default: the library's tuned setting)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB 5 times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        n = args.calib_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": n, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the alignment "
              "hot path has no cross-GPU exchange: read batches and pairs shard across "
              "GPUs as independent processes, each with its own copy of the index.",
              file=sys.stderr)
        sys.exit(2)

    global C5_SLOTS
    C5_SLOTS = args.c5_slots
    if args.walk != "default" or args.waves >= 0:  # applies to every leg below
        from swift_amd import align as _align
        t = _align.get_tuning()
        _align.set_tuning(args.waves if args.waves >= 0 else t["waves_per_cu"], t["allow_int16"],
                          t["walk_deferred"] if args.walk == "default"
                          else args.walk == "deferred")
    if args.c4_only:
        print(json.dumps(c4_leg(args.c4_reads, args.c4_batch,
                                0 if args.no_cpu else args.c4_parity,
                                0 if args.no_c5 else args.c5_requests, not args.no_graph)),
              flush=True)
        return
    if args.c2_only:
        print(json.dumps(c2_leg(args.c2_reads, args.steps,
                                0 if args.no_cpu else args.c2_cpu_reads)), flush=True)
        return

    from swift_amd import align, synth
    from swift_amd._lib import lib
    from swift_amd.device import synchronize

    t0 = time.perf_counter()
    qbuf, qoff, qlen, rbuf, roff, rlen = synth.c3_pairs_fast(args.pairs)
    cells = int((qlen.astype(np.int64) * rlen).sum())
    log(f"[bench] generated {args.pairs} pairs, {cells:.3e} cells in "
        f"{time.perf_counter() - t0:.1f} s")

    batch = align.PairBatch.from_arrays(qbuf, qoff, qlen, rbuf, roff, rlen)
    tuned = None
    n_tune = args.pairs if args.autotune < 0 else min(args.autotune, args.pairs)
    if n_tune > 0 and args.walk == "default" and args.waves < 0:
        # the settings are timed on the workload itself: a small prefix does
        # not rank them the way the full batch does (the walk placement in
        # particular)
        sample = batch if n_tune == args.pairs else align.PairBatch.from_arrays(
            qbuf, qoff[:n_tune], qlen[:n_tune], rbuf, roff[:n_tune], rlen[:n_tune])
        t0 = time.perf_counter()
        tuned = sample.autotune()
        log(f"[bench] autotune on {n_tune} pairs ({time.perf_counter() - t0:.1f} s): "
            f"{ {x: tuned[x] for x in ('waves_per_cu', 'allow_int16', 'walk_deferred', 'ms')} }")
        del sample
    ms4 = (C.c_float * 4)()
    lib().swift_sw_batch_timing(1, ms4)
    for i in range(args.warmup):
        batch.run()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    synchronize()
    fwd, walk, binning, rerun = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        batch.run()
        lib().swift_sw_batch_timing(1, ms4)
        binning.append(ms4[0])
        fwd.append(ms4[1])
        rerun.append(ms4[2])
        walk.append(ms4[3])
        log(f"[bench] step {i + 1}/{args.steps}: forward {ms4[1]:.2f} ms, walk {ms4[3]:.2f} ms")
    synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt * 1e3 / args.steps
    gcups = cells / (dt / args.steps) / 1e9
    stats = align.batch_stats()
    fwd_ms = float(np.mean(fwd))
    achieved = CELL_OPS * cells / (fwd_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16",
        "data": "synthetic (seeded generator, swift_amd/synth.py)",
        "config": {"workload": "C3 batched uneven pairs: 1M query/region pairs, query "
                               "U{36..250} bp, region = query + U{20..200} bp, 2% divergence; "
                               "scores, coordinates and consensus strings for every pair",
                   "pairs": args.pairs, "cells": cells, "scoring": [2, -1, -10, -1],
                   "lanes": "packed int16 with int32 re-run", "tuning": align.get_tuning()},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS,
                     "unit": "TOP/s", "frac": round(achieved / PEAK_INT16_TOPS, 4),
                     "traffic": round(TRAFFIC_B_PER_CELL * cells) if args.pairs == 1_000_000
                     else None,
                     "traffic_unit": "bytes per forward pass (PMC, calibrated)",
                     "kernel": "sw_batch_kernel<R,int16> (all row classes)",
                     "ops_per_cell": CELL_OPS, "forward_ms": round(fwd_ms, 3)},
        "phases_ms": {"binning": round(float(np.mean(binning)), 3),
                      "forward": round(fwd_ms, 3), "int32_rerun": round(float(np.mean(rerun)), 3),
                      "walk": round(float(np.mean(walk)), 3)},
        "batch_stats": stats,
    }
    if tuned is not None:
        out["autotune"] = {"sample_pairs": n_tune,
                           "chosen": {x: tuned[x] for x in ("waves_per_cu", "allow_int16",
                                                             "walk_deferred", "coarse_rows")},
                           "candidates_ms": [[r["waves_per_cu"], r["int16"], r["walk_deferred"],
                                              r["coarse_rows"], r["ms"]] for r in tuned["table"]]}
    if not args.no_plain:
        groups = plain_pass_setup(qbuf, qoff, qlen, rbuf, roff, rlen)
        plain_pass(groups)
        synchronize()
        t0 = time.perf_counter()
        for i in range(args.plain_steps):
            plain_pass(groups)
            synchronize()
            log(f"[bench] plain pass {i + 1}/{args.plain_steps} done")
        pdt = (time.perf_counter() - t0) / args.plain_steps
        out["plain_per_pair"] = {"gcups": round(cells / pdt / 1e9, 2),
                                 "ms_per_pass": round(pdt * 1e3, 3),
                                 "calls": len(groups)}
        out["speedup_vs_plain"] = round(gcups / (cells / pdt / 1e9), 3)
        del groups
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(batch, qbuf, qoff, qlen, rbuf, roff, rlen,
                                           args.cpu_pairs)
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    del batch
    if not args.no_c1:
        out["c1"] = c1_leg(args.steps, not args.no_cpu)
        log(f"[bench] C1: {out['c1']}")
    if not args.no_c2:
        out["c2"] = c2_leg(args.c2_reads, args.steps, 0 if args.no_cpu else args.c2_cpu_reads)
        log(f"[bench] C2: {out['c2']}")
    if not args.no_c4:
        out["c4"] = c4_leg(args.c4_reads, args.c4_batch, 0 if args.no_cpu else args.c4_parity,
                           0 if args.no_c5 else args.c5_requests, not args.no_graph)
        c5 = out["c4"].pop("c5", None)
        if c5 is not None:
            out["c5"] = c5
        log(f"[bench] C4: {out['c4']}")
        if c5 is not None:
            log(f"[bench] C5: {out['c5']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
