// sw_host.cpp -- host orchestration of the Smith-Waterman kernels: the
// alignSequences drop-in (plain per-pair kernel) and the batched
// length-binned path (device-side binning, persistent int16 kernel, int32
// re-run of overflow-suspect pairs, walk-back).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <utility>
#include <vector>

#include "sw_host.hpp"
#include "sw_types.hpp"

namespace swift {

// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";
const char *last_error() { return g_err; }
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int DevBuf::ensure(size_t want) {
    if (want <= bytes) return kOk;
    if (ptr) SWIFT_HIP_TRY(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    size_t b = want + want / 8 + 256;
    SWIFT_HIP_TRY(hipMalloc(&ptr, b));
    bytes = b;
    return kOk;
}
DevBuf::~DevBuf() {
    // buffers live until the process ends; freeing at exit races the
    // runtime's own teardown
}

Tuning &tuning() {
    static Tuning t;
    return t;
}

// ---------------------------------------------------------------------------
// Per-device workspace.
// ---------------------------------------------------------------------------
struct Timing {
    bool on = false;
    hipEvent_t ev[6] = {};
    float ms[5] = {};
};
static Timing g_timing;

static hipError_t mark(int i, hipStream_t st, bool timed) {
    if (!g_timing.on || !timed) return hipSuccess;
    if (!g_timing.ev[i]) {
        hipError_t e = hipEventCreate(&g_timing.ev[i]);
        if (e != hipSuccess) return e;
    }
    return hipEventRecord(g_timing.ev[i], st);
}

void set_batch_timing(bool on) { g_timing.on = on; }
const float *batch_timing_ms() { return g_timing.ms; }

struct Workspace {
    // plain path: scratch of a chunk (trace, best cells, moves)
    DevBuf p_trace;
    // Marks the end of the last plain call's kernels on its stream.  Every
    // call writes and walks p_trace; a call on another stream waits for this
    // event first, so kernels of calls on different streams never share the
    // trace at the same time.
    hipEvent_t p_done = nullptr;
    bool p_pending = false;
    // batched path
    DevBuf trace, res, ptr, bin_of, order, order2, tasks, task_words, scan_sums, moves;
    DevBuf bins;      // bin_count, bin_cursor (x2 passes)
    DevBuf segs;      // seg_task0 / seg_first / seg_len (x2 passes)
    DevBuf counters;  // int32 counters
    DevBuf totals;    // int64 total trace words per pass
    // Mapped pinned host block the sizes are gathered into (one stream
    // synchronisation per pass instead of a copy each)
    struct HostStage {
        int64_t words;
        int64_t steps[2 * kRBins];
        int32_t seg[2 * kRBins + 1];
        int32_t cnt[16];
    } *stage = nullptr;
    int cu_count = 0;
    BatchStats stats{};
    // Side streams the row-class kernels are spread over (fork/join by
    // events).  Three: with the caller's stream that is the four hardware
    // queues a process gets by default, so no two side streams share one
    // (kernels of streams on one hardware queue run one after another).
    static constexpr int kSide = 3;
    int nside = kSide;        // side streams in use; 0: every class on the caller's stream
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr, join[kSide] = {};
    bool streams_ready = false;
    bool timed = false;       // records the HIP-event phase marks (the shared workspace)
    int waves_per_cu = -1;    // persistent waves per CU for this workspace; -1: tuning()'s
    std::mutex mu;            // one batched run at a time per workspace

    DevBuf *all_bufs[18] = {&p_trace, &trace, &res, &ptr, &bin_of, &order,
                            &order2, &tasks, &task_words, &scan_sums, &moves, &bins, &segs,
                            &counters, &totals, nullptr};
};

static void init_workspace(Workspace *w) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) w->cu_count = prop.multiProcessorCount;
    if (w->cu_count <= 0) w->cu_count = 256;
}

static std::mutex g_mu;

// the workspace shared by the swift_sw_batch / alignSequences entry points
static Workspace &ws() {
    static std::vector<Workspace *> all;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if ((int)all.size() <= dev) all.resize(dev + 1, nullptr);
    if (!all[dev]) {
        all[dev] = new Workspace();
        init_workspace(all[dev]);
        all[dev]->timed = true;
    }
    return *all[dev];
}

// private workspaces (one per serving slot)
Workspace *workspace_new(int side_streams) {
    Workspace *w = new Workspace();
    init_workspace(w);
    w->nside = std::max(0, std::min(side_streams, (int)Workspace::kSide));
    return w;
}

void workspace_set_waves(Workspace *w, int waves_per_cu) { w->waves_per_cu = waves_per_cu; }

void workspace_free(Workspace *w) {
    if (!w) return;
    for (int i = 0; w->all_bufs[i]; ++i) {
        if (w->all_bufs[i]->ptr) (void)hipFree(w->all_bufs[i]->ptr);
        w->all_bufs[i]->ptr = nullptr;
        w->all_bufs[i]->bytes = 0;
    }
    if (w->streams_ready) {
        for (int i = 0; i < w->nside; ++i) {
            (void)hipStreamDestroy(w->side[i]);
            (void)hipEventDestroy(w->join[i]);
        }
        (void)hipEventDestroy(w->fork);
    }
    if (w->stage) (void)hipHostFree(w->stage);
    delete w;
}

int64_t workspace_bytes(const Workspace *w) {
    int64_t b = 0;
    for (int i = 0; w->all_bufs[i]; ++i) b += (int64_t)w->all_bufs[i]->bytes;
    return b;
}

BatchStats last_batch_stats() { return ws().stats; }

static int ensure_streams(Workspace &w) {
    if (w.streams_ready) return kOk;
    // only the streams in use: an idle stream would still take a hardware queue
    for (int i = 0; i < w.nside; ++i) {
        SWIFT_HIP_TRY(hipStreamCreateWithFlags(&w.side[i], hipStreamNonBlocking));
        SWIFT_HIP_TRY(hipEventCreateWithFlags(&w.join[i], hipEventDisableTiming));
    }
    SWIFT_HIP_TRY(hipEventCreateWithFlags(&w.fork, hipEventDisableTiming));
    w.streams_ready = true;
    return kOk;
}

// ---------------------------------------------------------------------------
// Scoring checks.  Every key must fit the int32 lanes: scores stay below
// 2^19 (the best-cell key packs H << 12 with the column), so the largest
// reachable score (query length x largest positive score, plus gaps when a
// gap "penalty" is positive) is bounded up front.
// ---------------------------------------------------------------------------
int validate_scoring(const Scoring &s, int maxq, int maxr, bool *force32) {
    int pos = std::max(std::max(s.match, s.mismatch), std::max(s.gap_open, s.gap_ext));
    pos = std::max(pos, 0);
    int neg = std::min(std::min(s.match, s.mismatch), std::min(s.gap_open, s.gap_ext));
    int64_t bound = (int64_t)(maxq + maxr) * pos;
    if (bound >= kMaxScore32 || neg < -(1 << 20)) {
        set_error("scores outside the supported range (max reachable score %lld)",
                  (long long)bound);
        return kErrInvalid;
    }
    // int16 lanes: keys 4*H + 3 within int16, penalties small enough that
    // the lowest candidate stays above -32768; and mismatch and gap moves
    // that all lose score (the packed sweep runs its padding columns
    // unmasked and clamps the left exit at 0)
    const bool losing = s.mismatch < 0 && s.gap_open <= -1 && s.gap_ext <= -1;
    *force32 = pos > 2000 || neg < -2000 || !losing;
    return kOk;
}

static int to_int_score(float f, int *out, const char *what) {
    if (!(std::fabs(f) < 1e7f) || std::floor(f) != f) {
        set_error("%s = %g is not an integer score (the reference parses scores with atoi)",
                  what, (double)f);
        return kErrInvalid;
    }
    *out = (int)f;
    return kOk;
}

int scoring_from_floats(float match, float mismatch, float gap_open, float gap_ext, Scoring *sc) {
    int rc;
    if ((rc = to_int_score(match, &sc->match, "match")) ||
        (rc = to_int_score(mismatch, &sc->mismatch, "mismatch")) ||
        (rc = to_int_score(gap_open, &sc->gap_open, "gapOpenPenalty")) ||
        (rc = to_int_score(gap_ext, &sc->gap_ext, "gapExtPenalty")))
        return rc;
    return kOk;
}

// ---------------------------------------------------------------------------
// Plain path (alignSequences drop-in).
// ---------------------------------------------------------------------------
// scratch of a plain chunk: 4 GiB of trace, moves and best cells (a chunk's
// walk is a latency-bound tail between sweeps, so fewer, larger chunks)
static size_t g_plain_budget_words = (size_t)1 << 30;
static int g_plain_chunk_pairs = 0;                     // > 0: at most this many pairs a chunk

int align_sequences_status(char *q_d, char *r_d, int qlen, int rlen, float match, float mismatch,
                           float gap_open, float gap_ext, float *score_d, int *start_d,
                           int *end_d, int *alen_d, char *qcons_d, char *rcons_d, int npairs,
                           hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_mu);
    Scoring sc;
    int rc;
    if ((rc = to_int_score(match, &sc.match, "match")) ||
        (rc = to_int_score(mismatch, &sc.mismatch, "mismatch")) ||
        (rc = to_int_score(gap_open, &sc.gap_open, "gapOpenPenalty")) ||
        (rc = to_int_score(gap_ext, &sc.gap_ext, "gapExtPenalty")))
        return rc;
    if (npairs < 0 || qlen < 0 || rlen < 0) {
        set_error("negative sizes");
        return kErrInvalid;
    }
    if (npairs == 0) return kOk;
    if (plain_rows_for(qlen) < 0 || rlen > kMaxCols - 64) {
        set_error("query length %d / reference length %d beyond the plain kernel (<= 1024 / %d)",
                  qlen, rlen, kMaxCols - 64);
        return kErrCapacity;
    }
    bool f32;
    if ((rc = validate_scoring(sc, qlen, rlen, &f32))) return rc;
    Workspace &w = ws();
    const int64_t wpp = plain_scratch_words(qlen, rlen);
    int chunk = (int)std::min<int64_t>(npairs, std::max<int64_t>(1, g_plain_budget_words / wpp));
    if (g_plain_chunk_pairs > 0) chunk = std::min(chunk, g_plain_chunk_pairs);
    const size_t need = (size_t)chunk * wpp * 4;
    if (!w.p_done) SWIFT_HIP_TRY(hipEventCreateWithFlags(&w.p_done, hipEventDisableTiming));
    if (w.p_pending) {
        // a larger trace replaces the buffer the last call's kernels may
        // still use; otherwise this stream queues behind them
        if (need > w.p_trace.bytes)
            SWIFT_HIP_TRY(hipEventSynchronize(w.p_done));
        else
            SWIFT_HIP_TRY(hipStreamWaitEvent(st, w.p_done, 0));
    }
    if ((rc = w.p_trace.ensure(need))) return rc;
    // The chunks run one after another on the caller's stream.  (Walking
    // chunk c on a second stream while chunk c + 1 sweeps was slower: the
    // walk's latency-bound waves took CU slots the sweep's single round of
    // waves needed, and both kernels stretched.)
    const PlainOut out{score_d, start_d, end_d, alen_d, qcons_d, rcons_d, (int64_t)qlen + rlen + 2};
    for (int c0 = 0; c0 < npairs; c0 += chunk) {
        const int cn = std::min(chunk, npairs - c0);
        SWIFT_HIP_TRY(launch_plain(q_d, r_d, qlen, rlen, sc, cn, c0, w.p_trace.as<uint32_t>(),
                                   out, st));
    }
    SWIFT_HIP_TRY(hipEventRecord(w.p_done, st));
    w.p_pending = true;
    return kOk;
}

void set_plain_chunk_pairs(int pairs) { g_plain_chunk_pairs = std::max(pairs, 0); }

// ---------------------------------------------------------------------------
// Batched path.
// ---------------------------------------------------------------------------
// 2-bit walk moves per pair id: room for the longest path any binned pair can
// have (the kernel limits, not the caller's maxq/maxr, bound what is binned)
constexpr int kMoveWords = (kMaxBatchQ + kMaxBatchR) / 16 + 1;
// tasks of one pass: the int32 class's 4 pairs per task, or the wide form's
// 2 (batches under kCoarseTinyPairs)
static int max_tasks_for(int npairs) {
    return std::max((npairs + kSlots32 - 1) / kSlots32,
                    (std::min(npairs, kCoarseTinyPairs) + kSlotsWide - 1) / kSlotsWide) +
           2 * kRBins + 1;
}

static int64_t worst_trace_words(int npairs, int maxq, int maxr) {
    int R = std::max(1, (maxq + kBatchG - 1) / kBatchG);
    // coarse row classes round R up (bound for every mode)
    R = std::min(kRBins, (R + kCoarseRowsSmall - 1) / kCoarseRowsSmall * kCoarseRowsSmall);
    int64_t steps = (maxr + kBatchG - 1 + 3) & ~3;
    int64_t w16 = (R + 7) / 8, w32 = (r32_class(R) + 15) / 16;
    int64_t t16 = (npairs + kSlots16 - 1) / kSlots16 + 2 * kRBins;
    int64_t t32 = (npairs + kSlots32 - 1) / kSlots32 + 2 * kRBins;
    int64_t words = (t16 * w16 + 2 * t32 * w32) * steps * 64;
    if (npairs < kCoarseTinyPairs) {
        // one class (R = kRBins): the wide form, one trace word per step
        // and lane, 2 pairs per task, rlen + 63 steps; int32 tasks R = 20
        int64_t wsteps = (maxr + kWideG - 1 + 3) & ~3;
        int64_t tw = (npairs + kSlotsWide - 1) / kSlotsWide;
        int64_t w32t = (r32_class(kRBins) + 15) / 16;
        words = std::max(words, tw * wsteps * 64 + 2 * t32 * w32t * steps * 64);
    }
    return words;
}

// totals layout (int64): [pass] trace words, then from kTotSteps the sweep
// steps of each (pass, segment)
constexpr int kTotSteps = 8;

static int reserve_common(Workspace &w, int npairs) {
    int rc;
    int mt = max_tasks_for(npairs);
    if ((rc = w.res.ensure((size_t)std::max(npairs, 1) * sizeof(PairResult)))) return rc;
    if ((rc = w.ptr.ensure((size_t)std::max(npairs, 1) * sizeof(PairTrace)))) return rc;
    if ((rc = w.bin_of.ensure((size_t)std::max(npairs, 1) * 4))) return rc;
    if ((rc = w.order.ensure((size_t)std::max(npairs, 1) * 4))) return rc;
    if ((rc = w.order2.ensure((size_t)std::max(npairs, 1) * 4))) return rc;
    if ((rc = ensure_streams(w))) return rc;
    if ((rc = w.tasks.ensure((size_t)mt * sizeof(Task)))) return rc;
    if ((rc = w.task_words.ensure((size_t)mt * 8))) return rc;
    if ((rc = w.scan_sums.ensure(((size_t)mt / 2048 + 2) * 8))) return rc;
    if ((rc = w.bins.ensure((size_t)4 * kNumBins * 4))) return rc;
    if ((rc = w.segs.ensure((size_t)2 * 3 * (2 * kRBins + 1) * 4))) return rc;
    if ((rc = w.counters.ensure(256 * 4))) return rc;
    if ((rc = w.totals.ensure((size_t)(kTotSteps + 2 * 2 * kRBins) * 8))) return rc;
    if (!w.stage) {
        void *p = nullptr;
        SWIFT_HIP_TRY(hipHostMalloc(&p, sizeof(Workspace::HostStage), hipHostMallocDefault));
        w.stage = static_cast<Workspace::HostStage *>(p);
    }
    return kOk;
}

// device bytes a batched run of this shape may hold (workspace, worst case)
int64_t batch_bytes_estimate(int npairs, int maxq, int maxr) {
    const int64_t n = std::max(npairs, 1), mt = max_tasks_for(npairs);
    return n * (int64_t)(sizeof(PairResult) + sizeof(PairTrace) + 12) +
           mt * (int64_t)(sizeof(Task) + 8) + 4 * (int64_t)worst_trace_words(npairs, maxq, maxr) +
           (int64_t)4 * kNumBins * 4 + (1 << 16);
}

static int reserve_all(Workspace &w, int npairs, int maxq, int maxr) {
    int rc = reserve_common(w, npairs);
    if (rc) return rc;
    if ((rc = w.moves.ensure((size_t)std::max(npairs, 1) * kMoveWords * 4))) return rc;
    return w.trace.ensure((size_t)worst_trace_words(npairs, maxq, maxr) * 4);
}

int reserve_batch(int npairs, int maxq, int maxr) {
    std::lock_guard<std::mutex> lk(g_mu);
    return reserve_all(ws(), npairs, maxq, maxr);
}

int reserve_batch_in(Workspace *w, int npairs, int maxq, int maxr) {
    std::lock_guard<std::mutex> lk(w->mu);
    return reserve_all(*w, npairs, maxq, maxr);
}

// counters layout (int32): [4*pass] ntasks, [4*pass+1] binned pairs, then
// one work-queue cursor per (pass, segment) from kCntQueues on
constexpr int kCntError = 8;
constexpr int kCntFlagged = 9;
constexpr int kCntQueues = 16;
constexpr int kQueuesPerPass = 2 * kRBins;
// sync_sizes: the sizes of each pass are read back (synchronising); then
// final_sync also waits for the pass's last kernels before returning
static int batch_run(Workspace &w, const BatchIO &io, const Scoring &sc, hipStream_t st,
                     bool sync_sizes, bool final_sync = true);

int sw_batch_run(const BatchIO &io, const Scoring &sc, hipStream_t st, bool sync_sizes) {
    std::lock_guard<std::mutex> lk(g_mu);
    return batch_run(ws(), io, sc, st, sync_sizes);
}

int sw_batch_run_in(Workspace *w, const BatchIO &io, const Scoring &sc, hipStream_t st) {
    std::lock_guard<std::mutex> lk(w->mu);
    // the caller's later work is on the same stream, and it synchronises
    // before it reads results
    return batch_run(*w, io, sc, st, true, false);
}

int sw_batch_enqueue_in(Workspace *w, const BatchIO &io, const Scoring &sc, hipStream_t st) {
    std::lock_guard<std::mutex> lk(w->mu);
    return batch_run(*w, io, sc, st, false);
}

int sw_batch_check_in(Workspace *w, hipStream_t st) {
    int32_t c[16];
    SWIFT_HIP_TRY(hipMemcpyAsync(c, w->counters.ptr, sizeof(c), hipMemcpyDeviceToHost, st));
    SWIFT_HIP_TRY(hipStreamSynchronize(st));
    if (c[kCntError]) {
        set_error("pair lengths beyond the batched limits (%d/%d)", kMaxBatchQ, kMaxBatchR);
        return kErrCapacity;
    }
    return kOk;
}

static int batch_run(Workspace &w, const BatchIO &io, const Scoring &sc, hipStream_t st,
                     bool sync_sizes, bool final_sync) {
    if (io.npairs < 0) {
        set_error("negative pair count");
        return kErrInvalid;
    }
    if (io.maxq > kMaxBatchQ || io.maxr > kMaxBatchR) {
        set_error("batched kernels take queries <= %d and references <= %d (got %d, %d)",
                  kMaxBatchQ, kMaxBatchR, io.maxq, io.maxr);
        return kErrCapacity;
    }
    bool force32 = false;
    int rc = validate_scoring(sc, io.maxq, io.maxr, &force32);
    if (rc) return rc;
    if (!tuning().int16) force32 = true;
    w.stats = BatchStats{};
    if (io.npairs == 0) return kOk;
    if ((rc = reserve_common(w, io.npairs))) return rc;
    const int n = io.npairs;
    const int mt = max_tasks_for(n);
    int32_t *cnt = w.counters.as<int32_t>();
    int64_t *tot = w.totals.as<int64_t>();
    int32_t *bins = w.bins.as<int32_t>();
    int32_t *segs = w.segs.as<int32_t>();
    const int NS = 2 * kRBins + 1;

    // Coarse row classes for small batches, streams without side streams and
    // graph capture: a class kernel whose tasks all fit one round of waves
    // lasts about one task, so the batch's latency is the number of classes
    // run one after another (or over the side streams) times a task; at most
    // 4 classes instead of 20.  Large batches keep every R (no padding rows).
    const int cmode = tuning().coarse_rows;
    // the smallest batches with host-known sizes round to kCoarseRowsSmall,
    // or under kCoarseTinyPairs to a single class: their class kernels run
    // about one task each, one after another, so fewer classes is lower
    // latency.  From kWideShortPairs on, the single (wide) class runs two of
    // its waves per SIMD; a batch whose queries all fit the 10-row class
    // (maxq <= 160) is then faster in that one 16-lane class.
    const bool short_q = io.maxq <= kBatchG * kCoarseRowsSmall;
    const int coarse =
        cmode == 1 || (cmode < 0 && (w.nside == 0 || !sync_sizes || n < kCoarsePairs))
            ? (!sync_sizes || n >= kCoarseSmallPairs                     ? kCoarseRows
               : n >= kCoarseTinyPairs || (n >= kWideShortPairs && short_q) ? kCoarseRowsSmall
                                                                           : kRBins)
            : 0;
    // a single class (the smallest batches) runs the wide form: its few
    // tasks are each one wave, and a wave's sweep is the batch's latency
    const int wide = coarse == kRBins ? 1 : 0;
    auto pass_args = [&](int pass, const PairResult *prev, int f32, BinArgs &b, TaskArgs &t) {
        b = BinArgs{io.rseq, io.roff, io.qlen, io.rlen, prev, n, f32, w.bin_of.as<int32_t>(),
                    bins + (2 * pass) * kNumBins, cnt + kCntError, io.score, io.start, io.end,
                    io.alen, coarse};
        t = TaskArgs{bins + (2 * pass) * kNumBins, bins + (2 * pass + 1) * kNumBins,
                     segs + (3 * pass) * NS, segs + (3 * pass + 1) * NS,
                     segs + (3 * pass + 2) * NS, cnt + 4 * pass, cnt + 4 * pass + 1,
                     reinterpret_cast<unsigned long long *>(tot + kTotSteps + pass * 2 * kRBins),
                     wide};
    };
    KeyConsts kc16 = make_consts(sc, kOneHot), kc32 = make_consts(sc, kBytes);
    int pos = std::max(std::max(std::max(sc.match, sc.mismatch), std::max(sc.gap_open, sc.gap_ext)), 0);
    // int16 lanes: a best score at this limit flags the pair for the int32
    // pass.  The keys 4H + direction stay below 32768 and the best-cell keys
    // 32H + 31 below 65536 for every cell of an unflagged pair: H rises by
    // at most pos per cell, so a pair with any larger cell has one in
    // [2048 - pos, 2047] whose key is exact.
    uint32_t h_limit = (uint32_t)std::min(32767 - (4 * pos + 3) - 4, 4 * (2048 - pos));

    // one launch zeroes the results (they start at "score 0": a pair no
    // kernel reaches never walks), the error flag, the flagged count, the
    // work-queue cursors and both passes' bin histograms
    {
        ZeroSpans z{};
        z.p[0] = w.res.as<int32_t>();
        z.n[0] = n * (int)(sizeof(PairResult) / 4);
        z.p[1] = cnt + kCntError;
        z.n[1] = kCntQueues + 2 * kQueuesPerPass - kCntError;
        z.p[2] = bins;
        z.n[2] = kNumBins;
        z.p[3] = bins + 2 * kNumBins;
        z.n[3] = kNumBins;
        z.nspans = 4;
        SWIFT_HIP_TRY(launch_zero(z, st));
    }
    SWIFT_HIP_TRY(mark(0, st, w.timed));
    // ---- pass 1: bin every pair
    BinArgs b1;
    TaskArgs t1;
    pass_args(0, nullptr, force32 ? 1 : 0, b1, t1);
    // the smallest batches with host-known sizes: one workgroup bins them and
    // writes the sizes into the staging block (no gather)
    const bool tiny_bin = sync_sizes && n <= kBinTinyPairs;
    if (tiny_bin) {
        const TinyStage hs{&w.stage->words, w.stage->steps, w.stage->seg, w.stage->cnt};
        SWIFT_HIP_TRY(launch_binning_tiny(b1, t1, w.order.as<int32_t>(), w.tasks.as<Task>(),
                                          w.task_words.as<int64_t>(), tot + 0, cnt, hs, st));
    } else {
        SWIFT_HIP_TRY(launch_binning(b1, t1, w.order.as<int32_t>(), w.tasks.as<Task>(),
                                     w.task_words.as<int64_t>(), nullptr, tot + 0,
                                     w.scan_sums.as<int64_t>(), mt, st));
    }
    int64_t words1 = 0;
    int32_t seg1[2 * kRBins + 1], seg2[2 * kRBins + 1];
    int64_t steps1[2 * kRBins], steps2[2 * kRBins];
    // Gathers up to four device arrays into the host staging block and waits
    // for them: one kernel and one synchronisation.
    using HS = Workspace::HostStage;
    auto gather = [&](std::initializer_list<std::pair<const void *, int>> spans,
                      std::initializer_list<int> dst) -> int {
        Gather g{};
        for (const auto &sp : spans) {
            g.src[g.nspans] = static_cast<const uint32_t *>(sp.first);
            g.nwords[g.nspans++] = sp.second;
        }
        int k = 0;
        for (int d : dst) g.dst[k++] = d / 4;
        g.out = reinterpret_cast<uint32_t *>(w.stage);
        SWIFT_HIP_TRY(launch_gather(g, st));
        SWIFT_HIP_TRY(hipStreamSynchronize(st));
        return kOk;
    };
    // trace size, per-class task counts (only non-empty classes are
    // launched) and per-class sweep steps (their cost estimate)
    auto read_sizes = [&](const int64_t *words, const int32_t *seg, const int64_t *steps,
                          const int32_t *c, int nc, int64_t *words_h, int32_t *seg_h,
                          int64_t *steps_h) -> int {
        int rc2 = gather({{words, 2}, {seg, 2 * kRBins + 1}, {steps, 4 * kRBins}, {c, nc}},
                         {(int)offsetof(HS, words), (int)offsetof(HS, seg),
                          (int)offsetof(HS, steps), (int)offsetof(HS, cnt)});
        if (rc2) return rc2;
        *words_h = w.stage->words;
        std::memcpy(seg_h, w.stage->seg, sizeof(w.stage->seg));
        std::memcpy(steps_h, w.stage->steps, sizeof(w.stage->steps));
        return kOk;
    };
    if (sync_sizes) {
        // and the error flag, in the same round trip
        if (tiny_bin) {
            SWIFT_HIP_TRY(hipStreamSynchronize(st));
            words1 = w.stage->words;
            std::memcpy(seg1, w.stage->seg, sizeof(w.stage->seg));
            std::memcpy(steps1, w.stage->steps, sizeof(w.stage->steps));
        } else if ((rc = read_sizes(tot + 0, t1.seg_task0, tot + kTotSteps, cnt, 16, &words1, seg1,
                                    steps1))) {
            return rc;
        }
        if (w.stage->cnt[kCntError]) {
            set_error("pair lengths beyond maxq/maxr or the batched limits (%d/%d)", kMaxBatchQ,
                      kMaxBatchR);
            return kErrCapacity;
        }
        if ((rc = w.trace.ensure((size_t)words1 * 4 + 4))) return rc;
        w.stats.ntasks16 = force32 ? 0 : seg1[kRBins];
        w.stats.ntasks32 = seg1[2 * kRBins] - seg1[kRBins];
    }
    if (tuning().walk_deferred && (rc = w.moves.ensure((size_t)n * kMoveWords * 4))) return rc;
    BatchArgs ba{io.qseq, io.qoff, io.qlen, io.rseq, io.roff, io.rlen, w.order.as<int32_t>(),
                 w.tasks.as<Task>(), nullptr, nullptr, nullptr, w.trace.as<uint32_t>(),
                 w.res.as<PairResult>(), w.ptr.as<PairTrace>(), kc16, h_limit,
                 io.score, io.start, io.end, io.alen, io.qcons, io.rcons, io.cons_off,
                 tuning().walk_deferred, cnt + kCntFlagged};
    // One persistent kernel per (class, R) segment, each reading its task
    // range from the device (empty segments exit at once), spread over the
    // side streams so the classes run concurrently and one class's tail
    // overlaps the next class's work.
    // With the host-known sizes (seg_host: task counts, steps_host: sweep
    // steps per segment) empty classes are skipped and the classes are
    // spread longest first onto the least loaded stream (LPT), cost = steps
    // x instructions per step (linear in R), and each class's deferred walk
    // follows its sweep on the same stream, overlapping the sweeps still
    // running on the others.  Without sizes every class kernel is launched
    // round-robin and the walk follows all sweeps.
    // Per-class walks overlap the other classes' sweeps only when the classes
    // spread over side streams.  With every class on the caller's stream
    // (serving slots) one walk after the last class is cheaper: a walk kernel
    // is a latency-bound tail, so one instead of one per class.
    const bool class_walks = ba.walk_deferred && sync_sizes && w.nside > 0;
    // With every class on one stream and host-known sizes, each pass's walk
    // follows its last class directly, ahead of the next readback, so the
    // GPU walks while the host waits for the flagged count.
    const bool pass_walks = ba.walk_deferred && sync_sizes && !class_walks;
    auto launch_classes = [&](const TaskArgs &t, bool with16, int pass, const int32_t *seg_host,
                              const int64_t *steps_host) -> int {
        int live[2 * kRBins], nlive = 0;
        for (int cls = with16 ? 0 : 1; cls < 2; ++cls) {
            for (int R = kRBins; R >= 1; --R) {
                if (cls == 1 && r32_class(R) != R) continue;
                int seg = cls * kRBins + (kRBins - R);
                if (seg_host && seg_host[seg + 1] == seg_host[seg]) continue;
                live[nlive++] = seg;
            }
        }
        if (nlive == 0) return kOk;
        // nside 0: the classes run one after another on the caller's stream
        const int nside = std::min(nlive, w.nside);
        int lane_of[2 * kRBins];
        if (steps_host) {
            double cost[2 * kRBins];
            for (int li = 0; li < nlive; ++li) {
                const int seg = live[li], R = kRBins - seg % kRBins;
                cost[seg] = (double)steps_host[seg] * (26.0 + 20.0 * R);
            }
            std::stable_sort(live, live + nlive, [&](int x, int y) { return cost[x] > cost[y]; });
            double load[Workspace::kSide] = {};
            for (int li = 0; li < nlive; ++li) {
                const int seg = live[li];
                int best = 0;
                for (int i = 1; i < nside; ++i)
                    if (load[i] < load[best]) best = i;
                load[best] += cost[seg];
                lane_of[seg] = best;
            }
        } else {
            for (int li = 0; li < nlive; ++li) lane_of[live[li]] = nside ? li % nside : 0;
        }
        if (nside) SWIFT_HIP_TRY(hipEventRecord(w.fork, st));
        for (int i = 0; i < nside; ++i)
            SWIFT_HIP_TRY(hipStreamWaitEvent(w.side[i], w.fork, 0));
        for (int li = 0; li < nlive; ++li) {
            const int seg = live[li], cls = seg / kRBins, R = kRBins - seg % kRBins;
            BatchArgs b = ba;
            b.task_lo = t.seg_task0 + seg;
            b.task_hi = t.seg_task0 + seg + 1;
            b.next_task = cnt + kCntQueues + pass * kQueuesPerPass + seg;
            b.kc = cls == 0 ? kc16 : kc32;
            hipStream_t s2 = nside ? w.side[lane_of[seg]] : st;
            const int wpc = w.waves_per_cu >= 0 ? w.waves_per_cu : tuning().waves_per_cu;
            const bool wseg = t.wide && seg == 0;
            const int ntk = seg_host ? seg_host[seg + 1] - seg_host[seg] : 0;
            if (wseg)
                SWIFT_HIP_TRY(launch_batch_wide(b, w.cu_count, ntk, s2));
            else
                SWIFT_HIP_TRY(launch_batch(b, cls == 0, R, w.cu_count, wpc, ntk, s2));
            if (class_walks && seg_host) {  // per-class walk
                const int slots = wseg ? kSlotsWide : cls == 0 ? kSlots16 : kSlots32;
                const int maxp = (seg_host[seg + 1] - seg_host[seg]) * slots;
                SWIFT_HIP_TRY(launch_walk(b, b.order, t.seg_first + seg, t.seg_len + seg, maxp,
                                          w.moves.as<uint32_t>(), kMoveWords, s2));
            }
        }
        for (int i = 0; i < nside; ++i) {
            SWIFT_HIP_TRY(hipEventRecord(w.join[i], w.side[i]));
            SWIFT_HIP_TRY(hipStreamWaitEvent(st, w.join[i], 0));
        }
        return kOk;
    };
    SWIFT_HIP_TRY(mark(1, st, w.timed));
    if ((rc = launch_classes(t1, !force32, 0, sync_sizes ? seg1 : nullptr,
                             sync_sizes ? steps1 : nullptr)))
        return rc;
    if (pass_walks)  // pass 1's pairs (the flagged ones wait for pass 2)
        SWIFT_HIP_TRY(launch_walk(ba, w.order.as<int32_t>(), nullptr, cnt + 1, n,
                                  w.moves.as<uint32_t>(), kMoveWords, st));

    SWIFT_HIP_TRY(mark(2, st, w.timed));
    // ---- pass 2: int32 re-run of pairs flagged by the int16 kernel
    int64_t words2 = words1;
    int32_t nflagged = 1;
    // no pair can be flagged: every reference byte is ACGTN and no best score
    // reaches the int16 limit (a score is at most maxq matches); then the
    // flagged-count round trip is skipped
    const bool no_rerun =
        io.refs_plain && (int64_t)io.maxq * std::max(sc.match, 0) * 4 < (int64_t)h_limit;
    if (!force32 && no_rerun) {
        nflagged = 0;
    } else if (!force32 && sync_sizes) {
        // the int16 kernels count what they flag: with none, pass 2 is skipped
        if ((rc = gather({{cnt + kCntFlagged, 1}}, {(int)offsetof(HS, cnt)}))) return rc;
        nflagged = w.stage->cnt[0];
    }
    if (!force32 && nflagged > 0) {
        BinArgs b2;
        TaskArgs t2;
        pass_args(1, w.res.as<PairResult>(), 1, b2, t2);
        SWIFT_HIP_TRY(launch_binning(b2, t2, w.order2.as<int32_t>(), w.tasks.as<Task>(),
                                     w.task_words.as<int64_t>(), tot + 0, tot + 1,
                                     w.scan_sums.as<int64_t>(), mt, st));
        int32_t nre = 0;
        if (sync_sizes) {
            if ((rc = read_sizes(tot + 1, t2.seg_task0, tot + kTotSteps + 2 * kRBins, cnt + 5, 1,
                                 &words2, seg2, steps2)))
                return rc;
            nre = w.stage->cnt[0];
            w.stats.ntasks32 += seg2[2 * kRBins] - seg2[kRBins];
            w.stats.nrerun = nre;
            if ((size_t)words2 * 4 + 4 > w.trace.bytes) {
                // grow, keeping pass-1 traces
                DevBuf bigger;
                if ((rc = bigger.ensure((size_t)words2 * 4 + 4))) return rc;
                SWIFT_HIP_TRY(hipMemcpyAsync(bigger.ptr, w.trace.ptr, (size_t)words1 * 4,
                                             hipMemcpyDeviceToDevice, st));
                SWIFT_HIP_TRY(hipStreamSynchronize(st));
                std::swap(w.trace.ptr, bigger.ptr);
                std::swap(w.trace.bytes, bigger.bytes);
                SWIFT_HIP_TRY(hipFree(bigger.ptr));
                bigger.ptr = nullptr;
                ba.trace = w.trace.as<uint32_t>();
            }
        }
        if (!sync_sizes || nre > 0) {
            ba.order = w.order2.as<int32_t>();
            if ((rc = launch_classes(t2, false, 1, sync_sizes ? seg2 : nullptr,
                                     sync_sizes ? steps2 : nullptr)))
                return rc;
            if (pass_walks)  // the re-run pairs
                SWIFT_HIP_TRY(launch_walk(ba, w.order2.as<int32_t>(), nullptr, cnt + 5, n,
                                          w.moves.as<uint32_t>(), kMoveWords, st));
        }
    }
    w.stats.trace_words = words2;

    // ---- walk-back: fused into the forward kernels; or, deferred, after each
    // class's sweep (host-known class sizes, side streams) or each pass
    // (host-known sizes, one stream), or else here, over every binned pair of
    // pass 1 (re-run pairs carry their pass-2 results by now)
    SWIFT_HIP_TRY(mark(3, st, w.timed));
    if (ba.walk_deferred && !class_walks && !pass_walks) {
        ba.order = w.order.as<int32_t>();
        SWIFT_HIP_TRY(launch_walk(ba, w.order.as<int32_t>(), nullptr, cnt + 1, n,
                                  w.moves.as<uint32_t>(), kMoveWords, st));
    }
    SWIFT_HIP_TRY(mark(4, st, w.timed));
    if (sync_sizes && final_sync) {
        SWIFT_HIP_TRY(hipStreamSynchronize(st));
        if (g_timing.on && w.timed) {
            for (int i = 0; i < 4; ++i)
                SWIFT_HIP_TRY(hipEventElapsedTime(&g_timing.ms[i], g_timing.ev[i], g_timing.ev[i + 1]));
        }
    }
    return kOk;
}

}  // namespace swift
