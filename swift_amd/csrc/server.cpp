// server.cpp -- continuous-batching serving loop over a device-resident
// index (no reference counterpart: the reference is a batch CLI).
//
// Requests (read batches of any size) are queued on submit.  Each of
// `nslots` slots is a host thread with its own HIP stream, mapping scratch
// and batched-SW workspace.  A free slot admits queued requests in arrival
// order and coalesces them into one batch while the batch's device-memory
// estimate stays within the slot's share of the budget (budget / nslots),
// then runs phase 1 -> phase 2 for the batch on its stream and hands every
// request its own records.  Slots run concurrently, so batches admitted
// while others are on the GPU overlap there; a request's results do not
// depend on what it was batched with (every read is mapped and aligned
// independently, records stay in read order).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "phase1_host.hpp"
#include "sw_types.hpp"

namespace swift {

using Clock = std::chrono::steady_clock;

constexpr int kRecI32 = 12;  // read strand ref_idx shift ref_pos nbiggest score start end alen keep ntied
constexpr int kServeWavesPerCu = 8;  // persistent waves per CU of a slot's row-class kernels

struct SrvRequest {
    uint64_t id = 0;
    std::vector<char> bases;
    std::vector<int64_t> off;
    std::vector<int32_t> len;
    int max_len = 0;
    int64_t est = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = kOk;
    std::string err;
    int nrec = 0;
    std::vector<int32_t> col[kRecI32];
    std::vector<int64_t> rdist;
    std::vector<char> qcons, rcons;  // packed: record k at cpos[k], alen bytes
    std::vector<int64_t> cpos;
    // A large request is served as parts: consecutive ranges of its reads,
    // queued as requests of their own so that idle slots take them at once.
    // A part knows its parent and its first read; the parent is done when its
    // last part is, and its records are its parts' in order.
    SrvRequest *parent = nullptr;
    int read0 = 0;
    std::vector<SrvRequest *> parts;
    int parts_left = 0;
    SrvRequest() = default;
    SrvRequest(const SrvRequest &) = delete;
    SrvRequest &operator=(const SrvRequest &) = delete;
    ~SrvRequest() {
        for (SrvRequest *q : parts) delete q;
    }
};

// requests above this many reads are split into parts over the slots
constexpr int kSplitReads = 1024;

// Pinned host staging (hipHostMalloc), grown on demand.
struct PinBuf {
    char *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes) return kOk;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        const size_t want = b + b / 4 + 4096;
        SWIFT_HIP_TRY(hipHostMalloc((void **)&p, want, hipHostMallocDefault));
        bytes = want;
        return kOk;
    }
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() {
        if (p) (void)hipHostFree(p);
    }
};

struct SrvSlot {
    std::thread th;
    Workspace *sws = nullptr;
    MapWork mw;
    hipStream_t st = nullptr;
    DevArr inb, nrec, rdist, qc, rc, pos, sums;
    DevArr col[kRecI32];
    PinBuf pin_in, pin_out;
};

struct Server {
    const IndexObj *ix = nullptr;
    int device = 0, nslots = 1, max_batch_reads = 1, max_read_len = 0, max_hits = 3;
    Scoring sc{2, -1, -10, -1};
    int64_t budget = 0, slot_cap = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SrvRequest *> queue;
    std::unordered_map<uint64_t, SrvRequest *> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<SrvSlot *> slots;
    int64_t completed = 0, batches = 0, batched_reads = 0, max_coalesced = 0;
};

static_assert(kRecI32 == kServeCols, "record columns of the serving readback");

static int64_t cstride_for(int max_len) { return ((int64_t)max_len + 302 + 3) & ~3; }

// device bytes one slot needs for a batch of nreads reads of <= max_len
static int64_t slot_bytes(int nreads, int max_len, int max_hits) {
    const int64_t cap = 2 * (int64_t)nreads * std::max(max_hits, 1);
    return align_bytes_estimate(nreads, max_len, max_hits) + (int64_t)nreads * (max_len + 12) +
           cap * (kRecI32 * 4 + 16 + 2 * cstride_for(max_len));
}

static size_t align8(size_t x) { return (x + 7) & ~(size_t)7; }

static int run_batch(Server &S, SrvSlot &sl, std::vector<SrvRequest *> &batch) {
    int nreads = 0, max_len = 0;
    int64_t nbases = 0;
    for (SrvRequest *r : batch) {
        nreads += (int)r->len.size();
        max_len = std::max(max_len, r->max_len);
        nbases += (int64_t)r->bases.size();
    }
    // reads -> pinned staging: bases | int64 offsets | int32 lengths
    const size_t o_off = align8((size_t)nbases), o_len = o_off + 8 * (size_t)nreads;
    int rc;
    if ((rc = sl.pin_in.ensure(o_len + 4 * (size_t)nreads + 64))) return rc;
    {
        char *bases = sl.pin_in.p;
        int64_t *off = reinterpret_cast<int64_t *>(sl.pin_in.p + o_off);
        int32_t *len = reinterpret_cast<int32_t *>(sl.pin_in.p + o_len);
        int64_t b = 0;
        int k = 0;
        for (SrvRequest *r : batch) {
            memcpy(bases + b, r->bases.data(), r->bases.size());
            for (size_t i = 0; i < r->len.size(); ++i, ++k) {
                off[k] = b + r->off[i];
                len[k] = r->len[i];
            }
            b += (int64_t)r->bases.size();
        }
    }
    const int64_t cap = 2 * (int64_t)nreads * S.max_hits;
    const int64_t cs = cstride_for(max_len);
    const size_t in_bytes = o_len + 4 * (size_t)nreads;
    if ((rc = ensure(sl.inb, in_bytes + 64)) || (rc = ensure(sl.nrec, 64)) ||
        (rc = ensure(sl.rdist, 8 * (size_t)cap + 64)) || (rc = ensure(sl.qc, (size_t)cap * cs + 64)) ||
        (rc = ensure(sl.rc, (size_t)cap * cs + 64)) || (rc = ensure(sl.pos, 8 * (size_t)cap + 64)) ||
        (rc = ensure(sl.sums, 8 * ((size_t)cap / 2048 + 4))))
        return rc;
    for (int c = 0; c < kRecI32; ++c)
        if ((rc = ensure(sl.col[c], 4 * (size_t)cap + 64))) return rc;
    // the staged reads in one copy: bases | offsets | lengths
    SWIFT_HIP_TRY(hipMemcpyAsync(sl.inb.ptr, sl.pin_in.p, in_bytes, hipMemcpyHostToDevice, sl.st));
    swift_records rec;
    rec.nrecords = sl.nrec.as<int32_t>();
    int32_t **cols[kRecI32] = {&rec.read, &rec.strand, &rec.ref_idx, &rec.shift, &rec.ref_pos,
                               &rec.nbiggest, &rec.score, &rec.start, &rec.end, &rec.alen,
                               &rec.keep, &rec.ntied};
    for (int c = 0; c < kRecI32; ++c) *cols[c] = sl.col[c].as<int32_t>();
    rec.ref_distance = sl.rdist.as<int64_t>();
    rec.qry_cons = sl.qc.as<char>();
    rec.ref_cons = sl.rc.as<char>();
    rec.cons_stride = cs;
    rec.capacity = (int32_t)cap;
    char *inb = sl.inb.as<char>();
    ReadsDev rd{inb, reinterpret_cast<const int64_t *>(inb + o_off),
                reinterpret_cast<const int32_t *>(inb + o_len), nreads};
    // the pipeline without its own final readback: the mapping error flags
    // come back with the record count below
    MapArgs m;
    PairsDev p;
    if ((rc = align_prepare(sl.mw, *S.ix, rd, max_len, S.max_hits, rec, &m, &p)) ||
        (rc = align_enqueue(m, p, sl.sws, nreads, max_len, S.sc, rec, sl.st, true)))
        return rc;
    // the records, distances, packed strings, count and mapping flags
    // written by one kernel into pinned memory sized for the worst case; one
    // synchronisation brings the batch back
    const ServeLayout worst = serve_layout((int)cap, cap * cs);
    if ((rc = sl.pin_out.ensure((size_t)worst.end + 64))) return rc;
    {
        ServeOut so;
        so.nrec = rec.nrecords;
        so.flags = m.ovl_count;
        for (int c = 0; c < kRecI32; ++c) so.col[c] = *cols[c];
        so.alen_col = 9;
        so.rdist = rec.ref_distance;
        so.qc = rec.qry_cons;
        so.rc = rec.ref_cons;
        so.cs = cs;
        so.pos = sl.pos.as<int64_t>();
        so.sums = sl.sums.as<int64_t>();
        so.total = sl.nrec.as<int64_t>() + 1;
        so.out = sl.pin_out.p;
        SWIFT_HIP_TRY(launch_serve_out(so, (int)cap, sl.st));
    }
    SWIFT_HIP_TRY(hipStreamSynchronize(sl.st));
    int32_t n, mflags[2];
    int64_t packed;
    memcpy(&n, sl.pin_out.p, 4);
    memcpy(&packed, sl.pin_out.p + 8, 8);
    memcpy(mflags, sl.pin_out.p + 16, 8);
    if ((rc = map_error(mflags[1]))) return rc;
    const ServeLayout lay = serve_layout(n, packed);
    const int32_t *col[kRecI32];
    for (int c = 0; c < kRecI32; ++c)
        col[c] = reinterpret_cast<const int32_t *>(sl.pin_out.p + lay.o_col + (size_t)c * 4 * n);
    const int64_t *hd = reinterpret_cast<const int64_t *>(sl.pin_out.p + lay.o_rd);
    const char *hq = sl.pin_out.p + lay.o_q, *hr = sl.pin_out.p + lay.o_r;
    // split by request: records are in read order, requests are consecutive
    int r0 = 0, i = 0;
    int64_t p0 = 0;
    for (SrvRequest *r : batch) {
        const int nr = (int)r->len.size();
        int j = i;
        while (j < n && col[0][j] < r0 + nr) ++j;
        const int m = j - i;
        r->nrec = m;
        for (int c = 0; c < kRecI32; ++c) r->col[c].assign(col[c] + i, col[c] + j);
        for (int k = 0; k < m; ++k) r->col[0][k] -= r0;  // read index within the request
        r->rdist.assign(hd + i, hd + j);
        r->cpos.resize(m);
        int64_t p = p0;
        for (int k = 0; k < m; ++k) {
            r->cpos[k] = p - p0;
            p += col[9][i + k];
        }
        r->qcons.assign(hq + p0, hq + p);
        r->rcons.assign(hr + p0, hr + p);
        p0 = p;
        i = j;
        r0 += nr;
    }
    return kOk;
}

static void slot_main(Server *S, SrvSlot *sl) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<SrvRequest *> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->stop) return;
            int reads = 0, max_len = 0;
            while (!S->queue.empty()) {
                SrvRequest *r = S->queue.front();
                const int nr = reads + (int)r->len.size();
                const int ml = std::max(max_len, r->max_len);
                if (!batch.empty() &&
                    (nr > S->max_batch_reads || slot_bytes(nr, ml, S->max_hits) > S->slot_cap))
                    break;
                batch.push_back(r);
                S->queue.pop_front();
                reads = nr;
                max_len = ml;
            }
            const auto now = Clock::now();
            for (SrvRequest *r : batch) r->t_start = now;
            S->max_coalesced = std::max<int64_t>(S->max_coalesced, (int64_t)batch.size());
        }
        int rc = run_batch(*S, *sl, batch);
        const std::string err = rc ? last_error() : std::string();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            const auto now = Clock::now();
            int nr = 0;
            for (SrvRequest *r : batch) {
                r->status = rc;
                r->err = err;
                r->done = true;
                r->t_done = now;
                nr += (int)r->len.size();
                SrvRequest *P = r->parent;
                if (!P) {
                    S->completed += 1;
                    continue;
                }
                if (rc && !P->status) {
                    P->status = rc;
                    P->err = err;
                }
                if (--P->parts_left == 0) {
                    int n = 0;
                    for (SrvRequest *q : P->parts) n += q->nrec;
                    P->nrec = n;
                    P->t_done = now;
                    P->done = true;
                    S->completed += 1;
                }
            }
            S->batches += 1;
            S->batched_reads += nr;
        }
        S->cv_done.notify_all();
    }
}

}  // namespace swift

using namespace swift;

struct swift_server {
    Server s;
};

extern "C" {

int swift_server_create(const swift_index *ix, int nslots, int64_t mem_budget_bytes,
                        int max_batch_reads, int max_read_len, int max_hits, int match,
                        int mismatch, int gap_open, int gap_ext, swift_server **out) {
    if (!ix || !out || nslots < 1 || nslots > 16 || max_batch_reads < 1 || max_hits < 1 ||
        max_read_len < 1) {
        set_error("invalid server arguments");
        return kErrInvalid;
    }
    if (max_read_len > kMaxBatchQ) {
        set_error("the pipeline aligns reads up to %d bases", kMaxBatchQ);
        return kErrCapacity;
    }
    Scoring sc{match, mismatch, gap_open, gap_ext};
    bool f32;
    int rc = validate_scoring(sc, max_read_len, 300, &f32);
    if (rc) return rc;
    auto *h = new swift_server;
    Server &S = h->s;
    S.ix = &ix->ix;
    (void)hipGetDevice(&S.device);
    S.nslots = nslots;
    S.max_batch_reads = max_batch_reads;
    S.max_read_len = max_read_len;
    S.max_hits = max_hits;
    S.sc = sc;
    S.budget = mem_budget_bytes;
    S.slot_cap = mem_budget_bytes / nslots;
    if (slot_bytes(1, max_read_len, max_hits) > S.slot_cap) {
        set_error("budget %lld B leaves %lld B per slot, below one read (%lld B)",
                  (long long)mem_budget_bytes, (long long)S.slot_cap,
                  (long long)slot_bytes(1, max_read_len, max_hits));
        delete h;
        return kErrCapacity;
    }
    for (int i = 0; i < nslots; ++i) {
        auto *sl = new SrvSlot;
        // the slot's classes run on its own stream: with several slots, side
        // streams would outnumber the hardware queues, and a class kernel of
        // one slot would then hold up another slot's work queued behind it
        sl->sws = workspace_new(0);
        // Serving batches keep 8 persistent waves per CU whatever a batch
        // autotuner chose for the process: with several slots' kernels
        // sharing the GPU, the occupancy limit gave p99 latencies of 6.7 to
        // 32.6 ms over three runs where 8 waves gave 5.6 to 6.6 ms (C2 index,
        // 400 requests at 70 % of capacity).
        if (sl->sws) workspace_set_waves(sl->sws, kServeWavesPerCu);
        hipError_t e = hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("hipStreamCreate: %s", hipGetErrorString(e));
            workspace_free(sl->sws);
            delete sl;
            for (SrvSlot *o : S.slots) {
                (void)hipStreamDestroy(o->st);
                workspace_free(o->sws);
                delete o;
            }
            delete h;
            return kErrHip;
        }
        S.slots.push_back(sl);
    }
    for (SrvSlot *sl : S.slots) sl->th = std::thread(slot_main, &S, sl);
    *out = h;
    return kOk;
}

int swift_server_submit(swift_server *h, const char *reads, const int64_t *off, const int32_t *len,
                        int nreads, uint64_t *id) {
    if (!h || nreads < 0 || (nreads > 0 && (!reads || !off || !len)) || !id) {
        set_error("invalid submit arguments");
        return kErrInvalid;
    }
    Server &S = h->s;
    auto *r = new SrvRequest;
    for (int i = 0; i < nreads; ++i) {
        if (len[i] < 0 || len[i] > S.max_read_len) {
            set_error("read %d has length %d (server limit %d)", i, len[i], S.max_read_len);
            delete r;
            return kErrCapacity;
        }
        r->max_len = std::max(r->max_len, len[i]);
    }
    if (nreads > S.max_batch_reads || slot_bytes(nreads, r->max_len, S.max_hits) > S.slot_cap) {
        set_error("request of %d reads exceeds one slot (%d reads, %lld B)", nreads,
                  S.max_batch_reads, (long long)S.slot_cap);
        delete r;
        return kErrCapacity;
    }
    // reads [a, b) of the call into request q
    auto fill = [&](SrvRequest *q, int a, int b) {
        int64_t nbq = 0;
        for (int i = a; i < b; ++i) nbq += len[i];
        q->bases.resize((size_t)nbq);
        q->off.resize(b - a);
        q->len.assign(len + a, len + b);
        int64_t o = 0;
        for (int i = a; i < b; ++i) {
            memcpy(q->bases.data() + o, reads + off[i], (size_t)len[i]);
            q->off[i - a] = o;
            q->max_len = std::max(q->max_len, len[i]);
            o += len[i];
        }
    };
    const int nparts =
        S.nslots > 1 && nreads > kSplitReads
            ? std::min(S.nslots, (nreads + kSplitReads - 1) / kSplitReads)
            : 1;
    if (nparts == 1) {
        r->max_len = 0;
        fill(r, 0, nreads);
    } else {
        for (int k = 0; k < nparts; ++k) {
            auto *q = new SrvRequest;
            const int a = (int)((int64_t)nreads * k / nparts);
            const int b = (int)((int64_t)nreads * (k + 1) / nparts);
            fill(q, a, b);
            q->parent = r;
            q->read0 = a;
            r->parts.push_back(q);
        }
        r->parts_left = nparts;
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S.mu);
        r->id = S.next_id++;
        S.reqs[r->id] = r;
        if (nparts == 1) {
            S.queue.push_back(r);
        } else {
            for (SrvRequest *q : r->parts) {
                q->t_submit = r->t_submit;
                S.queue.push_back(q);
            }
        }
    }
    if (nparts == 1) S.cv_work.notify_one();
    else S.cv_work.notify_all();
    *id = r->id;
    return kOk;
}

int swift_server_wait(swift_server *h, uint64_t id, int timeout_ms, int *nrecords,
                      double *latency_ms) {
    Server &S = h->s;
    std::unique_lock<std::mutex> lk(S.mu);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) {
        set_error("unknown request %llu", (unsigned long long)id);
        return kErrInvalid;
    }
    SrvRequest *r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) S.cv_done.wait(lk, ready);
    else if (!S.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        set_error("request %llu not finished after %d ms", (unsigned long long)id, timeout_ms);
        return kErrTimeout;
    }
    if (r->status) {
        set_error("request %llu failed: %s", (unsigned long long)id, r->err.c_str());
        return r->status;
    }
    if (nrecords) *nrecords = r->nrec;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    return kOk;
}

int swift_server_fetch(swift_server *h, uint64_t id, const swift_records *out) {
    Server &S = h->s;
    SrvRequest *r;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        if (it == S.reqs.end() || !it->second->done) {
            set_error("request %llu unknown or not finished", (unsigned long long)id);
            return kErrInvalid;
        }
        r = it->second;
        if (r->status) return r->status;
        if (out && out->capacity < r->nrec) {
            set_error("capacity %d below %d records", out->capacity, r->nrec);
            return kErrCapacity;
        }
        S.reqs.erase(it);
    }
    if (out) {
        int32_t *dst[kRecI32] = {out->read, out->strand, out->ref_idx, out->shift, out->ref_pos,
                                 out->nbiggest, out->score, out->start, out->end, out->alen,
                                 out->keep, out->ntied};
        // the request's records, or its parts' one after another (their read
        // indices shifted by each part's first read)
        int at = 0;
        auto put = [&](const SrvRequest *q, int read0) {
            const int n = q->nrec;
            for (int c = 0; c < kRecI32; ++c)
                if (dst[c] && n) memcpy(dst[c] + at, q->col[c].data(), 4 * (size_t)n);
            if (read0 && dst[0])
                for (int i = 0; i < n; ++i) dst[0][at + i] += read0;
            if (out->ref_distance && n)
                memcpy(out->ref_distance + at, q->rdist.data(), 8 * (size_t)n);
            for (int i = 0; i < n; ++i) {
                const int64_t L = std::min<int64_t>(q->col[9][i], out->cons_stride);
                if (out->qry_cons)
                    memcpy(out->qry_cons + (int64_t)(at + i) * out->cons_stride,
                           q->qcons.data() + q->cpos[i], (size_t)L);
                if (out->ref_cons)
                    memcpy(out->ref_cons + (int64_t)(at + i) * out->cons_stride,
                           q->rcons.data() + q->cpos[i], (size_t)L);
            }
            at += n;
        };
        if (r->parts.empty()) put(r, 0);
        else
            for (const SrvRequest *q : r->parts) put(q, q->read0);
        if (out->nrecords) *out->nrecords = at;
    }
    delete r;
    return kOk;
}

int swift_server_stats(swift_server *h, int64_t *completed, int64_t *batches,
                       int64_t *batched_reads, int64_t *max_coalesced, int64_t *device_bytes) {
    Server &S = h->s;
    std::lock_guard<std::mutex> lk(S.mu);
    if (completed) *completed = S.completed;
    if (batches) *batches = S.batches;
    if (batched_reads) *batched_reads = S.batched_reads;
    if (max_coalesced) *max_coalesced = S.max_coalesced;
    if (device_bytes) {
        int64_t b = 0;
        for (SrvSlot *sl : S.slots) b += workspace_bytes(sl->sws);
        *device_bytes = b;
    }
    return kOk;
}

int swift_server_destroy(swift_server *h) {
    if (!h) return kOk;
    Server &S = h->s;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        S.stop = true;
    }
    S.cv_work.notify_all();
    for (SrvSlot *sl : S.slots)
        if (sl->th.joinable()) sl->th.join();
    for (SrvSlot *sl : S.slots) {
        (void)hipStreamSynchronize(sl->st);
        (void)hipStreamDestroy(sl->st);
        workspace_free(sl->sws);
        delete sl;
    }
    for (auto &kv : S.reqs) delete kv.second;
    delete h;
    return kOk;
}

}  // extern "C"
