// phase1_host.cpp -- host side of phase 1: FASTA readers with the
// reference's line handling, the device-resident index, read mapping, the
// phase-1 -> phase-2 pipeline, and the reference's phase-1 entry points
// (lookupTable5*, mapHits6GetBestHits, getHash, qryGetReverseComplement,
// searchQueries3) on top of them.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "phase1_host.hpp"
#include "sw_types.hpp"

namespace swift {

// ---------------------------------------------------------------------------
// FASTA readers
// ---------------------------------------------------------------------------
struct RefSeqs {
    std::vector<char> bases;
    std::vector<int64_t> off, len;
    std::vector<std::string> names;
};

static bool slurp(const char *path, std::vector<char> *buf) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    buf->clear();
    char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof(tmp), f)) > 0) buf->insert(buf->end(), tmp, tmp + n);
    fclose(f);
    return true;
}

// lookupTable5Create's reading of the reference (src/lookupTable5.cu:130-185):
// a '>' line starts the next sequence; the bases of the following lines are
// concatenated (tuples run across line breaks through the carried tail, and
// a blank line keeps that tail, because the line length the reference tests
// includes it).  The sequence as indexed is therefore the concatenation of
// its lines without the newline bytes.
static int read_reference_fasta(const char *path, RefSeqs *out) {
    std::vector<char> buf;
    if (!slurp(path, &buf)) {
        set_error("cannot read reference file %s", path);
        return kErrIO;
    }
    out->bases.clear();
    out->off.clear();
    out->len.clear();
    out->names.clear();
    size_t i = 0;
    while (i < buf.size()) {
        size_t e = i;
        while (e < buf.size() && buf[e] != '\n') ++e;
        if (e > i) {
            if (buf[i] == '>') {
                out->names.emplace_back(buf.data() + i + 1, e - i - 1);
                out->off.push_back((int64_t)out->bases.size());
                out->len.push_back(0);
            } else {
                if (out->off.empty()) {
                    set_error("%s: sequence data before the first '>' header", path);
                    return kErrInvalid;
                }
                out->bases.insert(out->bases.end(), buf.begin() + i, buf.begin() + e);
                out->len.back() += (int64_t)(e - i);
            }
        }
        i = e + 1;
    }
    return kOk;
}

// qryGetNext2 (src/query.cu:174-214): a '>' line names the next read and
// sets its offsets (nameOffset = start of the header line, seqOffset = the
// byte after it); each sequence line is one read.  Blank lines are skipped.
struct QueryFile {
    std::vector<char> bases;
    std::vector<int64_t> off;
    std::vector<int32_t> len;
    std::vector<std::string> names;
    std::vector<unsigned long> name_off, seq_off;
};

static int read_query_fasta(const char *path, QueryFile *q) {
    std::vector<char> buf;
    if (!slurp(path, &buf)) {
        set_error("cannot read query file %s", path);
        return kErrIO;
    }
    std::string name;
    unsigned long noff = 0, soff = 0;
    size_t i = 0;
    while (i < buf.size()) {
        size_t e = i;
        while (e < buf.size() && buf[e] != '\n') ++e;
        size_t next = e < buf.size() ? e + 1 : e;
        if (e > i) {
            if (buf[i] == '>') {
                name.assign(buf.data() + i + 1, e - i - 1);
                noff = (unsigned long)(next - (e - i) - 1);
                soff = (unsigned long)next;
            } else {
                q->off.push_back((int64_t)q->bases.size());
                q->len.push_back((int32_t)(e - i));
                q->bases.insert(q->bases.end(), buf.begin() + i, buf.begin() + e);
                q->names.push_back(name);
                q->name_off.push_back(noff);
                q->seq_off.push_back(soff);
            }
        }
        i = next;
    }
    return kOk;
}

// ---------------------------------------------------------------------------
// Index
// ---------------------------------------------------------------------------
static int64_t next_pow2_64(int64_t n) {
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

static int index_build(const char *bases, const int64_t *off, const int64_t *len, int nseq, int k,
                       int thr, IndexObj *ix) {
    if (k < 1 || k > kMaxSeeds) {
        set_error("seed length %d outside 1..%d", k, kMaxSeeds);
        return kErrInvalid;
    }
    if (nseq < 0 || nseq > kMaxIndexSeqs) {
        set_error("%d sequences: the index holds at most %d (refIdx is a signed char)", nseq,
                  kMaxIndexSeqs);
        return kErrCapacity;
    }
    int rc;
    ix->k = k;
    ix->thr = thr;
    ix->nseq = nseq;
    ix->nhash = (int64_t)1 << (2 * k);
    ix->h_off.assign(nseq, 0);
    ix->h_len.assign(nseq, 0);
    std::vector<int64_t> tup0(std::max(nseq, 1), 0);
    int64_t nb = 0, nt = 0;
    for (int s = 0; s < nseq; ++s) {
        if (len[s] < 0 || len[s] > INT32_MAX) {
            set_error("sequence %d length %lld outside 0..2^31-1", s, (long long)len[s]);
            return kErrCapacity;
        }
        ix->h_off[s] = nb;
        ix->h_len[s] = len[s];
        tup0[s] = nt;
        nb += len[s];
        nt += len[s] / k;
    }
    ix->ntup = nt;
    if ((rc = ix->seq.alloc((size_t)nb + 64))) return rc;
    for (int s = 0; s < nseq; ++s)
        if (len[s] > 0)
            SWIFT_HIP_TRY(hipMemcpy(ix->seq.as<char>() + ix->h_off[s], bases + off[s], (size_t)len[s],
                                    hipMemcpyHostToDevice));
    {
        // ACGTN bases only? (then no alignment window can need the int32
        // re-run for its reference bytes)
        DevArr flag;
        if ((rc = flag.alloc(64))) return rc;
        SWIFT_HIP_TRY(hipMemset(flag.ptr, 0, 4));
        SWIFT_HIP_TRY(launch_seq_plain_scan(ix->seq.as<char>(), nb, flag.as<int32_t>(), nullptr));
        int32_t any = 0;
        SWIFT_HIP_TRY(hipMemcpy(&any, flag.ptr, 4, hipMemcpyDeviceToHost));
        ix->plain_bases = any ? 0 : 1;
    }
    {
        // cluster-id bases: cid = cid0[s] + shift + kShiftBias stays below
        // 2^32 - 1 (the empty-slot marker of the hash tier)
        std::vector<uint32_t> c0(std::max(nseq, 1), 0);
        uint64_t acc = 0;
        for (int s = 0; s < nseq; ++s) {
            c0[s] = (uint32_t)std::min<uint64_t>(acc, 0xFFFFFFFFull);
            acc += (uint64_t)len[s] + 2 * kShiftBias;
        }
        ix->cid_ok = acc < 0xFFFFFFFFull ? 1 : 0;
        if ((rc = ix->cid0.alloc(4 * c0.size()))) return rc;
        SWIFT_HIP_TRY(hipMemcpy(ix->cid0.ptr, c0.data(), 4 * c0.size(), hipMemcpyHostToDevice));
    }
    if ((rc = ix->seq_off.alloc(8 * (size_t)std::max(nseq, 1)))) return rc;
    if ((rc = ix->seq_len.alloc(8 * (size_t)std::max(nseq, 1)))) return rc;
    if ((rc = ix->tup0.alloc(8 * (size_t)std::max(nseq, 1)))) return rc;
    if (nseq > 0) {
        SWIFT_HIP_TRY(hipMemcpy(ix->seq_off.ptr, ix->h_off.data(), 8 * nseq, hipMemcpyHostToDevice));
        SWIFT_HIP_TRY(hipMemcpy(ix->seq_len.ptr, ix->h_len.data(), 8 * nseq, hipMemcpyHostToDevice));
        SWIFT_HIP_TRY(hipMemcpy(ix->tup0.ptr, tup0.data(), 8 * nseq, hipMemcpyHostToDevice));
    }
    const size_t nh = (size_t)ix->nhash;
    if ((rc = ix->count.alloc(4 * nh))) return rc;
    if ((rc = ix->start.alloc(4 * nh))) return rc;
    if ((rc = ix->actual.alloc(4 * nh))) return rc;
    if ((rc = ix->cursor.alloc(4 * nh))) return rc;
    if ((rc = ix->bucket.alloc(8 * nh))) return rc;
    SWIFT_HIP_TRY(hipMemset(ix->count.ptr, 0, 4 * nh));

    const int nblocks = (int)((ix->nhash + kScanChunk - 1) / kScanChunk);
    DevArr block_sum, small, mid_list, big_list, big_scratch;
    if ((rc = block_sum.alloc(8 * (size_t)nblocks))) return rc;
    if ((rc = small.alloc(64))) return rc;
    SWIFT_HIP_TRY(hipMemset(small.ptr, 0, 64));
    IndexBuildScratch sc{};
    sc.block_sum = block_sum.as<int64_t>();
    sc.total = small.as<int64_t>();
    sc.max_actual = small.as<int32_t>() + 2;
    sc.mid_count = small.as<int32_t>() + 3;
    sc.big_count = small.as<int32_t>() + 4;
    IndexDev d = ix->dev();
    SWIFT_HIP_TRY(launch_index_build(d, ix->ntup, sc, nullptr));
    int32_t head[6];
    SWIFT_HIP_TRY(hipMemcpy(head, small.ptr, sizeof(head), hipMemcpyDeviceToHost));
    int64_t total;
    memcpy(&total, head, 8);
    ix->total = total;
    ix->max_actual = head[2];
    if (total >= INT32_MAX) {
        set_error("%lld kept tuples: bucket starts are 32-bit", (long long)total);
        return kErrCapacity;
    }
    if ((rc = ix->entry.alloc(8 * (size_t)(total + 1)))) return rc;
    SWIFT_HIP_TRY(hipMemset(ix->entry.ptr, 0xFF, 8));  // sentinel (-1, -1)
    const int64_t nlist = total / (kSmallBucket + 1) + 1;
    if ((rc = mid_list.alloc(4 * (size_t)nlist))) return rc;
    if ((rc = big_list.alloc(4 * (size_t)nlist))) return rc;
    sc.mid_list = mid_list.as<int32_t>();
    sc.big_list = big_list.as<int32_t>();
    if (ix->max_actual > kWaveBucket) {
        sc.big_cap = (int32_t)next_pow2_64(ix->max_actual);
        int64_t slots = ((int64_t)256 << 20) / ((int64_t)sc.big_cap * 8);
        sc.big_slots = (int32_t)std::max<int64_t>(1, std::min<int64_t>(slots, 256));
        if ((rc = big_scratch.alloc((size_t)sc.big_slots * sc.big_cap * 8))) return rc;
        sc.big_scratch = big_scratch.as<uint64_t>();
    }
    d = ix->dev();
    SWIFT_HIP_TRY(launch_index_fill(d, ix->ntup, sc, nullptr));
    if (ix->cid_ok) {
        // the hash tier's 4-byte form of the (now ordered) entries
        if ((rc = ix->gpos.alloc(4 * (size_t)(total + 1)))) return rc;
        SWIFT_HIP_TRY(launch_index_gpos(d, total, ix->gpos.as<uint32_t>(), nullptr));
    }
    SWIFT_HIP_TRY(hipDeviceSynchronize());
    return kOk;
}

static int index_export(const IndexObj &ix, int32_t *start, int32_t *count, int32_t *actual,
                        int8_t *ref_idx, int32_t *ref_pos) {
    const size_t nh = (size_t)ix.nhash;
    if (start) SWIFT_HIP_TRY(hipMemcpy(start, ix.start.ptr, 4 * nh, hipMemcpyDeviceToHost));
    if (count) SWIFT_HIP_TRY(hipMemcpy(count, ix.count.ptr, 4 * nh, hipMemcpyDeviceToHost));
    if (actual) SWIFT_HIP_TRY(hipMemcpy(actual, ix.actual.ptr, 4 * nh, hipMemcpyDeviceToHost));
    if (ref_idx || ref_pos) {
        std::vector<uint64_t> e((size_t)ix.total + 1);
        SWIFT_HIP_TRY(hipMemcpy(e.data(), ix.entry.ptr, 8 * e.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < e.size(); ++i) {
            if (ref_idx) ref_idx[i] = i ? (int8_t)(e[i] >> 32) : -1;
            if (ref_pos) ref_pos[i] = i ? (int32_t)(uint32_t)e[i] : -1;
        }
    }
    return kOk;
}

// ---------------------------------------------------------------------------
// Mapping
// ---------------------------------------------------------------------------
static std::mutex g_p1_mu;
static MapWork &map_work() {
    static MapWork *w = new MapWork;  // lives until the process ends
    return *w;
}

// grows by half again, so batches of varying size settle after a few calls
// (hipFree synchronises the device)
int ensure(DevArr &a, size_t bytes) {
    return bytes <= a.bytes ? (int)kOk : a.alloc(bytes + bytes / 2 + 256);
}

// Fills `m` for a launch over `rd`; out arrays may come from the caller
// (any of them null: the workspace's are used).
// -1: hash tier first when the index predicts it (mean_hits >= kHashFirstHits
// in prepare_map); 0: wave tier first; 1: hash tier first
static std::atomic<int> g_map_mode{-1};

static int prepare_map(const IndexObj &ix, const ReadsDev &rd, int max_read_len, int max_hits,
                       MapArgs *m, MapWork &w, bool want_strands) {
    if (max_read_len < 0 || max_read_len > kMaxMapRead) {
        set_error("reads up to %d bases (got max_read_len %d)", kMaxMapRead, max_read_len);
        return kErrCapacity;
    }
    if (max_hits < 0) {
        set_error("negative max_hits");
        return kErrInvalid;
    }
    const int ns = 2 * rd.nreads;
    const int stride = std::max(max_hits, 1);
    int rc;
    const int nq = std::max(max_read_len - ix.k + 1, 0);
    const int64_t bound = (int64_t)nq * ix.max_actual;
    if (bound >= INT32_MAX) {
        set_error("up to %lld hits per strand: beyond the 32-bit hit counters", (long long)bound);
        return kErrCapacity;
    }
    if ((rc = ensure(w.ref, (size_t)ns * stride + 64))) return rc;
    if ((rc = ensure(w.shift, 4 * ((size_t)ns * stride + 16)))) return rc;
    if ((rc = ensure(w.pos, 4 * ((size_t)ns * stride + 16)))) return rc;
    if ((rc = ensure(w.nbest, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.nbig, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.nhits, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.ovl_list, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.flags, 64))) return rc;
    const int qstride = (max_read_len + 15) & ~15;
    if (want_strands && (rc = ensure(w.strand_seq, (size_t)ns * qstride + 64))) return rc;
    int slots = 0;
    int64_t cap = 0;
    if (bound > kMidCap) {
        // the workgroup sort pads a strand's keys to a power of two in its slot
        cap = next_pow2_64(bound);
        slots = (int)std::max<int64_t>(1, std::min<int64_t>(512, (((int64_t)512 << 20) / (cap * 8))));
        if ((rc = ensure(w.scratch, (size_t)slots * cap * 8))) return rc;
    }
    if ((rc = ensure(w.big_list, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.mid_list, 4 * ((size_t)ns + 16)))) return rc;
    if ((rc = ensure(w.mid2_list, 4 * ((size_t)ns + 16)))) return rc;
    m->ix = ix.dev();
    m->rd = rd;
    m->out = MapOut{w.ref.as<int8_t>(), w.shift.as<int32_t>(), w.pos.as<int32_t>(),
                    w.nbest.as<int32_t>(), w.nbig.as<int32_t>(), w.nhits.as<int32_t>(), stride};
    m->max_hits = max_hits;
    m->max_len = max_read_len;
    m->strand_seq = want_strands ? w.strand_seq.as<char>() : nullptr;
    m->qstride = qstride;
    m->scratch = w.scratch.as<uint64_t>();
    m->scratch_cap = cap;
    m->scratch_slots = slots;
    m->ovl_list = w.ovl_list.as<int32_t>();
    m->ovl_count = w.flags.as<int32_t>();
    m->error = w.flags.as<int32_t>() + 1;
    m->big_list = w.big_list.as<int32_t>();
    m->big_count = w.flags.as<int32_t>() + 2;
    m->mid_list = w.mid_list.as<int32_t>();
    m->mid_count = w.flags.as<int32_t>() + 3;
    m->mid2_list = w.mid2_list.as<int32_t>();
    m->mid2_count = w.flags.as<int32_t>() + 4;
    m->mid_blocks = bound > kMapCap ? 2 * 256 : 0;
    // Hash tier first when the index's mean bucket size predicts more hits
    // per strand than the wave tier holds: the wave tier would only count
    // them and hand them on, and the hash tier counts them again.
    const double mean_hits = (double)ix.total / (double)std::max<int64_t>(ix.nhash, 1) * nq;
    const int mode = g_map_mode.load();
    m->hash_first = mode == 1 || (mode == -1 && mean_hits >= kHashFirstHits);
    return kOk;
}

static int run_map(const MapArgs &m, hipStream_t st) {
    // overflow count, error flags, big-, mid- and mid2-strand counts
    SWIFT_HIP_TRY(launch_fill(m.ovl_count, 5, 0, st));
    SWIFT_HIP_TRY(launch_map(m, st, m.mid_blocks));
    return kOk;
}

int map_error(int flags) {
    if (flags & 2) {
        set_error("a read is longer than max_read_len");
        return kErrCapacity;
    }
    if (flags & 1) {
        set_error("a strand's hits exceeded the overflow scratch");
        return kErrCapacity;
    }
    return kOk;
}

static int map_reads(const IndexObj &ix, const ReadsDev &rd, int max_read_len, int max_hits,
                     int8_t *ref, int32_t *shift, int32_t *pos, int32_t *nbest, int32_t *nbig,
                     int32_t *nhits, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_p1_mu);
    MapWork &w = map_work();
    MapArgs m;
    int rc = prepare_map(ix, rd, max_read_len, max_hits, &m, w, false);
    if (rc) return rc;
    m.out.ref = ref;
    m.out.shift = shift;
    m.out.pos = pos;
    m.out.nbest = nbest;
    if (nbig) m.out.nbig = nbig;
    if (nhits) m.out.nhits = nhits;
    if ((rc = run_map(m, st))) return rc;
    int32_t flags[2];
    SWIFT_HIP_TRY(hipMemcpyAsync(flags, m.ovl_count, 8, hipMemcpyDeviceToHost, st));
    SWIFT_HIP_TRY(hipStreamSynchronize(st));
    return map_error(flags[1]);
}

// Phase 1 -> phase 2.
constexpr int kBefore = 100, kWin = 300;  // src/align.cu:2388,2435

int64_t align_bytes_estimate(int nreads, int max_read_len, int max_hits) {
    const int64_t ns = 2 * (int64_t)nreads, cap = ns * std::max(max_hits, 0);
    const int64_t qstride = (max_read_len + 15) & ~15;
    const int64_t cstride = max_read_len + kWin + 2;
    return ns * (std::max(max_hits, 1) * 9 + 16 + qstride) + cap * (13 * 4 + 8 + 2 * cstride) +
           cap * 36 + batch_bytes_estimate((int)std::max<int64_t>(cap, 1), max_read_len, kWin);
}

static int align_reads(const IndexObj &ix, const ReadsDev &rd, int max_read_len, int max_hits,
                       const Scoring &sc, const swift_records &rec, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_p1_mu);
    return align_reads_in(map_work(), nullptr, ix, rd, max_read_len, max_hits, sc, rec, st);
}

// Validates a pipeline call and sizes its scratch; fills the launch
// arguments of the mapping (m) and of the records (p).
int align_prepare(MapWork &w, const IndexObj &ix, const ReadsDev &rd, int max_read_len,
                  int max_hits, const swift_records &rec, MapArgs *m, PairsDev *p) {
    if (max_read_len > kMaxBatchQ) {
        set_error("the pipeline aligns reads up to %d bases (got %d)", kMaxBatchQ, max_read_len);
        return kErrCapacity;
    }
    const int64_t cap = 2 * (int64_t)rd.nreads * std::max(max_hits, 0);
    if (cap > rec.capacity || cap >= INT32_MAX) {
        set_error("record capacity %d below 2 * nreads * max_hits = %lld", rec.capacity,
                  (long long)cap);
        return kErrCapacity;
    }
    if (rec.cons_stride < max_read_len + kWin + 2) {
        set_error("cons_stride %lld below max_read_len + 302", (long long)rec.cons_stride);
        return kErrInvalid;
    }
    int rc = prepare_map(ix, rd, max_read_len, max_hits, m, w, true);
    if (rc) return rc;
    const size_t c = (size_t)std::max<int64_t>(cap, 1);
    if ((rc = ensure(w.pair0, 4 * (2 * (size_t)rd.nreads + 16)))) return rc;
    if ((rc = ensure(w.scan_sums, 4 * (2 * (size_t)rd.nreads / 2048 + 16)))) return rc;
    if ((rc = ensure(w.qoff, 8 * c))) return rc;
    if ((rc = ensure(w.roff, 8 * c))) return rc;
    if ((rc = ensure(w.coff, 8 * c))) return rc;
    if ((rc = ensure(w.qlen, 4 * c))) return rc;
    if ((rc = ensure(w.rlen, 4 * c))) return rc;
    p->npairs = rec.nrecords;
    p->pair0 = w.pair0.as<int32_t>();
    p->read = rec.read;
    p->strand = rec.strand;
    p->ref = rec.ref_idx;
    p->shift = rec.shift;
    p->pos = rec.ref_pos;
    p->nbig = rec.nbiggest;
    p->rdist = rec.ref_distance;
    p->qoff = w.qoff.as<int64_t>();
    p->roff = w.roff.as<int64_t>();
    p->coff = w.coff.as<int64_t>();
    p->qlen = w.qlen.as<int32_t>();
    p->rlen = w.rlen.as<int32_t>();
    p->capacity = (int)cap;
    p->cstride = rec.cons_stride;
    p->scan_sums = w.scan_sums.as<int32_t>();
    return kOk;
}

// Enqueues map -> records/windows -> batched SW -> best choice on `st`.
// sync: the batched SW sizes its trace on the device (synchronising);
// otherwise it runs from a reservation and the sequence is capturable.
int align_enqueue(const MapArgs &m, const PairsDev &p, Workspace *sws, int nreads, int max_read_len,
                  const Scoring &sc, const swift_records &rec, hipStream_t st, bool sync) {
    int rc;
    if ((rc = run_map(m, st))) return rc;
    SWIFT_HIP_TRY(launch_records(m, p, kBefore, kWin, st));
    BatchIO io;
    io.qseq = m.strand_seq;
    io.qoff = p.qoff;
    io.qlen = p.qlen;
    io.rseq = m.ix.seq;
    io.roff = p.roff;
    io.rlen = p.rlen;
    io.npairs = p.capacity;
    io.maxq = max_read_len;
    io.maxr = kWin;
    io.score = rec.score;
    io.start = rec.start;
    io.end = rec.end;
    io.alen = rec.alen;
    io.qcons = rec.qry_cons;
    io.rcons = rec.ref_cons;
    io.cons_off = p.coff;
    io.refs_plain = m.ix.plain_bases;
    if (sync)
        rc = sws ? sw_batch_run_in(sws, io, sc, st) : sw_batch_run(io, sc, st, true);
    else
        rc = sw_batch_enqueue_in(sws, io, sc, st);
    if (rc) return rc;
    SWIFT_HIP_TRY(launch_choose_best(p, nreads, rec.score, rec.keep, rec.ntied, st));
    return kOk;
}

// mapping error flags of the last run on `st` (synchronises)
int map_check(const MapArgs &m, hipStream_t st) {
    int32_t flags[2];
    SWIFT_HIP_TRY(hipMemcpyAsync(flags, m.ovl_count, 8, hipMemcpyDeviceToHost, st));
    SWIFT_HIP_TRY(hipStreamSynchronize(st));
    return map_error(flags[1]);
}

int align_reads_in(MapWork &w, Workspace *sws, const IndexObj &ix, const ReadsDev &rd,
                   int max_read_len, int max_hits, const Scoring &sc, const swift_records &rec,
                   hipStream_t st) {
    MapArgs m;
    PairsDev p;
    int rc = align_prepare(w, ix, rd, max_read_len, max_hits, rec, &m, &p);
    if (rc) return rc;
    if ((rc = align_enqueue(m, p, sws, rd.nreads, max_read_len, sc, rec, st, true))) return rc;
    return map_check(m, st);
}

// ---------------------------------------------------------------------------
// The reference's module state (src/lookupTable5.cu:39-51): one index
// shared by lookupTable5MapQry calls.
// ---------------------------------------------------------------------------
static IndexObj *g_lt5 = nullptr;
static int g_lt5_max_hits = 0;
// Guards g_lt5 and the drop-ins' scratch.  The reference calls
// lookupTable5MapQry and mapHits6GetBestHits from 16 threads at once
// (src/search.cu:323-345); here such calls run one after another.  Lock
// order: g_lt5_mu, then g_p1_mu (taken inside map_reads).
static std::mutex g_lt5_mu;

// Pinned host block, grow-only: the drop-ins stage their inputs and results
// through it, one copy each way per call.
struct PinnedArr {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes) return kOk;
        if (ptr) SWIFT_HIP_TRY(hipHostFree(ptr));
        ptr = nullptr;
        bytes = 0;
        const size_t nb = b + b / 2 + 256;
        SWIFT_HIP_TRY(hipHostMalloc(&ptr, nb, hipHostMallocDefault));
        bytes = nb;
        return kOk;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Scratch of lookupTable5MapQry / mapHits6GetBestHits: grows to the largest
// call seen and is kept, so a steady stream of calls allocates nothing.
struct DropinWork {
    DevArr in, out, kbuf;
    PinnedArr hin, hout;
};
static DropinWork &dropin_work() {
    static DropinWork *w = new DropinWork;  // lives until the process ends
    return *w;
}

static int build_from_fasta(const char *path, int k, int thr, IndexObj **out) {
    RefSeqs r;
    int rc = read_reference_fasta(path, &r);
    if (rc) return rc;
    std::unique_ptr<IndexObj> ix(new IndexObj);
    rc = index_build(r.bases.data(), r.off.data(), r.len.data(), (int)r.off.size(), k, thr,
                     ix.get());
    if (rc) return rc;
    ix->names = std::move(r.names);
    *out = ix.release();
    return kOk;
}

static void die(const char *fn) {
    fprintf(stderr, "Error: %s: %s\n", fn, last_error());
    exit(1);
}

}  // namespace swift

using namespace swift;

extern "C" {

void preprocessCreate(void) {}
void preprocessDelete(void) {}

int getHash(char *s, int length) {
    int h = 0;
    for (int i = length - 1; i >= 0; --i) {
        int c;
        switch (s[i]) {
        case 'C': case 'c': c = 1; break;
        case 'G': case 'g': c = 2; break;
        case 'T': case 't': c = 3; break;
        default: c = 0;
        }
        h = h * 4 + c;
    }
    return h;
}

void qryGetReverseComplement(char *seq, int seqLength, char *revComp) {
    for (int i = 0; i < seqLength; ++i) {
        char c = seq[seqLength - 1 - i];
        revComp[i] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
    }
    revComp[seqLength < 0 ? 0 : seqLength] = '\0';
}

void lookupTable5Create(const char *refFile, int seedLen, int maxHitsPerQry,
                        int tupleIgnoreThreshold, int *totalTuples) {
    std::lock_guard<std::mutex> lk(g_lt5_mu);
    IndexObj *ix = nullptr;
    {
        std::lock_guard<std::mutex> lk1(g_p1_mu);
        delete g_lt5;
        g_lt5 = nullptr;
        if (build_from_fasta(refFile, seedLen, tupleIgnoreThreshold, &ix)) die(__func__);
    }
    g_lt5 = ix;
    g_lt5_max_hits = maxHitsPerQry;
    if (totalTuples) *totalTuples = (int)ix->total;
}

int lookupTable5MapQry(const Query *query, unsigned int qryLen, unsigned int isRevComp,
                       char *refIdx_bestHits, int *shift_bestHits, int *refPos_bestHits) {
    std::lock_guard<std::mutex> lk(g_lt5_mu);
    if (!g_lt5) {
        set_error("lookupTable5MapQry before lookupTable5Create");
        die(__func__);
    }
    if (qryLen > (unsigned)kMaxMapRead) {
        set_error("reads up to %d bases (got qryLen %u)", kMaxMapRead, qryLen);
        die(__func__);
    }
    const char *seq = isRevComp == 1 ? query->revCompSeq : query->seq;
    const int L = (int)qryLen;
    const int mh = std::max(g_lt5_max_hits, 0);
    const size_t stride = (size_t)std::max(mh, 1);
    DropinWork &w = dropin_work();
    // in: [offset int64 | length int32 | 4 pad | bases]
    // out (both strands of the one read; strand 0 is the sequence asked for):
    // [nbest int32 x 2 | shift int32 x 2 stride | refPos int32 x 2 stride | refIdx int8 x 2 stride]
    const size_t in_bytes = 16 + (size_t)L, out_bytes = 8 + 18 * stride;
    if (ensure(w.in, in_bytes + 64) || ensure(w.out, out_bytes) || w.hin.ensure(in_bytes) ||
        w.hout.ensure(out_bytes))
        die(__func__);
    char *hin = w.hin.as<char>();
    const int64_t off0 = 16;
    memcpy(hin, &off0, 8);
    memcpy(hin + 8, &L, 4);
    memset(hin + 12, 0, 4);
    if (L > 0) memcpy(hin + 16, seq, L);
    if (hipMemcpyAsync(w.in.ptr, hin, in_bytes, hipMemcpyHostToDevice, nullptr) != hipSuccess)
        die(__func__);
    char *din = w.in.as<char>(), *dout = w.out.as<char>();
    ReadsDev rd{din, reinterpret_cast<const int64_t *>(din),
                reinterpret_cast<const int32_t *>(din + 8), 1};
    int32_t *d_n = reinterpret_cast<int32_t *>(dout);
    int32_t *d_sh = d_n + 2, *d_pos = d_sh + 2 * stride;
    int8_t *d_ref = reinterpret_cast<int8_t *>(d_pos + 2 * stride);
    if (map_reads(*g_lt5, rd, L, mh, d_ref, d_sh, d_pos, d_n, nullptr, nullptr, nullptr))
        die(__func__);
    char *hout = w.hout.as<char>();
    if (hipMemcpy(hout, dout, out_bytes, hipMemcpyDeviceToHost) != hipSuccess) die(__func__);
    const int32_t *h_n = reinterpret_cast<const int32_t *>(hout);
    const int32_t *h_sh = h_n + 2, *h_pos = h_sh + 2 * stride;
    const int8_t *h_ref = reinterpret_cast<const int8_t *>(h_pos + 2 * stride);
    const int n = h_n[0];
    for (int i = 0; i < n; ++i) {
        refIdx_bestHits[i] = (char)h_ref[i];
        shift_bestHits[i] = h_sh[i];
        refPos_bestHits[i] = h_pos[i];
    }
    return n;
}

void lookupTable5Delete(void) {
    std::lock_guard<std::mutex> lk(g_lt5_mu);
    std::lock_guard<std::mutex> lk1(g_p1_mu);
    delete g_lt5;
    g_lt5 = nullptr;
}

int mapHits6GetBestHits(int numBestHits, char *refIdx, int *shift, int *refPos, int size,
                        char *refIdx_bestMatch, int *shift_bestMatch, int *refPos_bestMatch) {
    if (size <= 0) return 0;  // src/mapHits6.cu:45-46
    std::lock_guard<std::mutex> lk(g_lt5_mu);
    const int mh = std::max(numBestHits, 0);
    const size_t n = (size_t)size, m = (size_t)std::max(mh, 1);
    DropinWork &w = dropin_work();
    // one device block, staged both ways through one pinned block:
    // [count int32 x 2 | best shift x m | best refPos x m | best refIdx x m | pad to 16 |
    //  shift x n | refPos x n | refIdx x n]   (the hits are sorted in place)
    const size_t head = (8 + 9 * m + 15) & ~(size_t)15;
    const size_t bytes = head + 9 * n;
    if (ensure(w.out, bytes) || w.hout.ensure(bytes) ||
        ensure(w.kbuf, 16 * (size_t)next_pow2_64(size)))
        die(__func__);
    char *h = w.hout.as<char>();
    memcpy(h + head, shift, 4 * n);
    memcpy(h + head + 4 * n, refPos, 4 * n);
    memcpy(h + head + 8 * n, refIdx, n);
    char *d = w.out.as<char>();
    if (hipMemcpyAsync(d + head, h + head, 9 * n, hipMemcpyHostToDevice, nullptr) != hipSuccess)
        die(__func__);
    int32_t *d_n = reinterpret_cast<int32_t *>(d);
    int32_t *d_osh = d_n + 2, *d_opos = d_osh + m;
    int8_t *d_oref = reinterpret_cast<int8_t *>(d_opos + m);
    int32_t *d_sh = reinterpret_cast<int32_t *>(d + head), *d_pos = d_sh + n;
    int8_t *d_ref = reinterpret_cast<int8_t *>(d_pos + n);
    if (launch_best_hits(d_ref, d_sh, d_pos, size, mh, d_oref, d_osh, d_opos, d_n, d_n + 1,
                         w.kbuf.ptr, nullptr) != hipSuccess ||
        hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        die(__func__);
    memcpy(shift, h + head, 4 * n);
    memcpy(refPos, h + head + 4 * n, 4 * n);
    memcpy(refIdx, h + head + 8 * n, n);
    const int32_t *h_n = reinterpret_cast<const int32_t *>(h);
    const int32_t *h_osh = h_n + 2, *h_opos = h_osh + m;
    const int8_t *h_oref = reinterpret_cast<const int8_t *>(h_opos + m);
    const int cnt = h_n[0];
    for (int i = 0; i < cnt; ++i) {
        refIdx_bestMatch[i] = (char)h_oref[i];
        shift_bestMatch[i] = h_osh[i];
        refPos_bestMatch[i] = h_opos[i];
    }
    return cnt;
}

// Phase 1 from files to a match file: index refFileName, map both strands of
// the first numQrs reads of qryFileName, one line per hit.  zero_shift writes
// the shift column as 0, as search2Queries' printHeuristicMatches does
// (src/search2.cu:341-373); searchQueries3's printHeuristicMatches4 writes
// the hit's shift (src/search.cu:2230-2240).
static void search_to_match_file(const char *fn, const char *qryFileName, int numQrs,
                                 const char *refFileName, int seedLen, const char *matchFile,
                                 int maxNumHits, bool zero_shift, int *numMatches,
                                 int tupleIgnoreThres) {
    IndexObj *ix = nullptr;
    if (build_from_fasta(refFileName, seedLen, tupleIgnoreThres, &ix)) die(fn);
    std::unique_ptr<IndexObj> hold(ix);
    QueryFile q;
    if (read_query_fasta(qryFileName, &q)) die(fn);
    const int nreads = std::min<int>(numQrs, (int)q.len.size());
    int maxlen = 0;
    for (int r = 0; r < nreads; ++r) maxlen = std::max(maxlen, q.len[r]);
    if (maxlen > kMaxMapRead) {
        set_error("reads up to %d bases (found %d)", kMaxMapRead, maxlen);
        die(fn);
    }
    FILE *out = fopen(matchFile, "w");
    if (!out) {
        set_error("cannot write %s", matchFile);
        die(fn);
    }
    const int mh = maxNumHits;
    const int stride = std::max(mh, 1);
    const int batch = 1 << 18;
    int total_lines = 0;
    DevArr dseq, doff, dlen, dref, dsh, dpos, dn;
    std::vector<int8_t> href;
    std::vector<int32_t> hsh, hpos, hn;
    for (int r0 = 0; r0 < nreads; r0 += batch) {
        const int nr = std::min(batch, nreads - r0);
        const int64_t b0 = q.off[r0];
        const int64_t b1 = q.off[r0 + nr - 1] + q.len[r0 + nr - 1];
        std::vector<int64_t> off(nr);
        for (int i = 0; i < nr; ++i) off[i] = q.off[r0 + i] - b0;
        const size_t ns = 2 * (size_t)nr;
        if (dseq.alloc(std::max<int64_t>(b1 - b0, 1)) || doff.alloc(8 * nr) ||
            dlen.alloc(4 * nr) || dref.alloc(ns * stride) || dsh.alloc(4 * ns * stride) ||
            dpos.alloc(4 * ns * stride) || dn.alloc(4 * ns))
            die(fn);
        if ((b1 > b0 &&
             hipMemcpy(dseq.ptr, q.bases.data() + b0, b1 - b0, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(doff.ptr, off.data(), 8 * nr, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dlen.ptr, q.len.data() + r0, 4 * nr, hipMemcpyHostToDevice) != hipSuccess)
            die(fn);
        ReadsDev rd{dseq.as<char>(), doff.as<int64_t>(), dlen.as<int32_t>(), nr};
        if (map_reads(*ix, rd, maxlen, mh, dref.as<int8_t>(), dsh.as<int32_t>(), dpos.as<int32_t>(),
                      dn.as<int32_t>(), nullptr, nullptr, nullptr))
            die(fn);
        href.resize(ns * stride);
        hsh.resize(ns * stride);
        hpos.resize(ns * stride);
        hn.resize(ns);
        if (hipMemcpy(href.data(), dref.ptr, href.size(), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hsh.data(), dsh.ptr, 4 * hsh.size(), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hpos.data(), dpos.ptr, 4 * hpos.size(), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hn.data(), dn.ptr, 4 * hn.size(), hipMemcpyDeviceToHost) != hipSuccess)
            die(fn);
        for (int i = 0; i < nr; ++i) {
            const int r = r0 + i;
            for (int s = 0; s < 2; ++s) {
                const size_t sidx = 2 * (size_t)i + s;
                for (int h = 0; h < hn[sidx]; ++h) {
                    const size_t o = sidx * stride + h;
                    fprintf(out, "%s\t%d\t%d\t%d\t%d\t%lu\t%lu\n", q.names[r].c_str(), s,
                            (int)href[o], zero_shift ? 0 : hsh[o], hpos[o], q.name_off[r], q.seq_off[r]);
                    ++total_lines;
                }
            }
        }
    }
    fclose(out);
    if (numMatches) *numMatches = total_lines;
}

void searchQueries3(char *qryFileName, int numQrs, int qryLen, char *refFileName, int seedLen,
                    char *matchFile, unsigned int maxNumHits, int *numMatches, int numThreads,
                    int tupleIgnoreThres) {
    (void)qryLen;
    (void)numThreads;
    // src/search.cu:279-281
    if (qryFileName == NULL || refFileName == NULL || seedLen <= 0 || matchFile == NULL ||
        numQrs <= 0)
        return;
    search_to_match_file(__func__, qryFileName, numQrs, refFileName, seedLen, matchFile,
                         (int)maxNumHits, false, numMatches, tupleIgnoreThres);
}

void search2Queries(char *qryFile, char *refFile, char *matchFile, int seedLen, int maxNumHits,
                    int *numMatches, int tupleIgnoreThres) {
    (void)maxNumHits;  // src/search2.cu:64 sets it to 3 whatever is passed
    if (qryFile == NULL || refFile == NULL || seedLen <= 0 || matchFile == NULL) {
        set_error("search2Queries: null file name or seedLen <= 0");
        die(__func__);
    }
    search_to_match_file(__func__, qryFile, INT32_MAX, refFile, seedLen, matchFile, 3, true,
                         numMatches, tupleIgnoreThres);
}

// ---- status-returning API --------------------------------------------------
int swift_set_map_mode(int mode) {
    if (mode < -1 || mode > 1) {
        set_error("map mode must be -1, 0 or 1");
        return kErrInvalid;
    }
    g_map_mode = mode;
    return kOk;
}

int swift_get_map_mode(int *mode) {
    if (mode) *mode = g_map_mode.load();
    return kOk;
}

int swift_index_build(const char *seq, const int64_t *off, const int64_t *len, int nseq,
                      int seed_len, int tuple_ignore_threshold, swift_index **out) {
    if (!out || (nseq > 0 && (!seq || !off || !len))) {
        set_error("null argument");
        return kErrInvalid;
    }
    std::lock_guard<std::mutex> lk(g_p1_mu);
    std::unique_ptr<swift_index> h(new swift_index);
    int rc = index_build(seq, off, len, nseq, seed_len, tuple_ignore_threshold, &h->ix);
    if (rc) return rc;
    h->ix.names.assign(nseq, std::string());
    *out = h.release();
    return kOk;
}

int swift_index_build_fasta(const char *path, int seed_len, int tuple_ignore_threshold,
                            swift_index **out) {
    if (!out || !path) {
        set_error("null argument");
        return kErrInvalid;
    }
    std::lock_guard<std::mutex> lk(g_p1_mu);
    RefSeqs r;
    int rc = read_reference_fasta(path, &r);
    if (rc) return rc;
    std::unique_ptr<swift_index> h(new swift_index);
    rc = index_build(r.bases.data(), r.off.data(), r.len.data(), (int)r.off.size(), seed_len,
                     tuple_ignore_threshold, &h->ix);
    if (rc) return rc;
    h->ix.names = std::move(r.names);
    *out = h.release();
    return kOk;
}

int swift_index_free(swift_index *ix) {
    std::lock_guard<std::mutex> lk(g_p1_mu);
    delete ix;
    return kOk;
}

int swift_index_info(const swift_index *ix, int *nseq, int64_t *total_tuples, int64_t *nhash,
                     int *max_bucket) {
    if (!ix) {
        set_error("no index");
        return kErrNoIndex;
    }
    if (nseq) *nseq = ix->ix.nseq;
    if (total_tuples) *total_tuples = ix->ix.total;
    if (nhash) *nhash = ix->ix.nhash;
    if (max_bucket) *max_bucket = ix->ix.max_actual;
    return kOk;
}

int swift_index_export(const swift_index *ix, int32_t *start, int32_t *count, int32_t *actual,
                       int8_t *ref_idx, int32_t *ref_pos) {
    if (!ix) {
        set_error("no index");
        return kErrNoIndex;
    }
    return index_export(ix->ix, start, count, actual, ref_idx, ref_pos);
}

int swift_index_sequence(const swift_index *ix, int s, int64_t *len, const char **name) {
    if (!ix) {
        set_error("no index");
        return kErrNoIndex;
    }
    if (s < 0 || s >= ix->ix.nseq) {
        set_error("sequence %d outside 0..%d", s, ix->ix.nseq - 1);
        return kErrInvalid;
    }
    if (len) *len = ix->ix.h_len[s];
    if (name) *name = ix->ix.names[s].c_str();
    return kOk;
}

int swift_map_reads(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                    const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                    int8_t *ref_idx_d, int32_t *shift_d, int32_t *ref_pos_d, int32_t *nbest_d,
                    int32_t *nbiggest_d, int32_t *nhits_d, void *stream) {
    if (!ix) {
        set_error("no index");
        return kErrNoIndex;
    }
    if (nreads < 0 || (nreads > 0 && (!reads_d || !off_d || !len_d || !ref_idx_d || !shift_d ||
                                      !ref_pos_d || !nbest_d))) {
        set_error("null or negative argument");
        return kErrInvalid;
    }
    if (nreads == 0) return kOk;
    ReadsDev rd{reads_d, off_d, len_d, nreads};
    return map_reads(ix->ix, rd, max_read_len, max_hits, ref_idx_d, shift_d, ref_pos_d, nbest_d,
                     nbiggest_d, nhits_d, (hipStream_t)stream);
}

int swift_align_reads(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                      const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                      int match, int mismatch, int gap_open, int gap_ext,
                      const swift_records *rec, void *stream) {
    if (!ix) {
        set_error("no index");
        return kErrNoIndex;
    }
    if (!rec || nreads < 0 || (nreads > 0 && (!reads_d || !off_d || !len_d))) {
        set_error("null or negative argument");
        return kErrInvalid;
    }
    Scoring sc{match, mismatch, gap_open, gap_ext};
    ReadsDev rd{reads_d, off_d, len_d, nreads};
    if (nreads == 0) {
        int32_t zero = 0;
        SWIFT_HIP_TRY(hipMemcpy(rec->nrecords, &zero, 4, hipMemcpyHostToDevice));
        return kOk;
    }
    return align_reads(ix->ix, rd, max_read_len, max_hits, sc, *rec, (hipStream_t)stream);
}

}  // extern "C"
