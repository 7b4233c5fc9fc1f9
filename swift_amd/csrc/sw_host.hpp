// sw_host.hpp -- declarations shared by the host-side translation units.
#pragma once

#include "swift_common.hpp"

namespace swift {

struct KeyConsts;
struct Task;
struct BatchArgs;
struct BinArgs;
struct TaskArgs;

// launch helpers implemented in sw_kernels.hip
struct PlainOut;
// Plain path over pairs [pair0, pair0 + npairs) of the caller's arrays:
// sweep, walk and consensus expansion.  scratch holds
// npairs * plain_scratch_words(qlen, rlen_max) words.
hipError_t launch_plain(const char *q, const char *r, int qlen, int rlen_max, const Scoring &sc,
                        int npairs, int pair0, uint32_t *scratch, const PlainOut &out,
                        hipStream_t st);
int plain_rows_for(int qlen);
int plain_trace_words(int R);
int64_t plain_scratch_words(int qlen, int rlen);  // device words per pair
void set_plain_chunk_pairs(int pairs);             // 0: chunks by the scratch budget alone
hipError_t launch_fill(int32_t *p, int n, int32_t v, hipStream_t st);
// up to four int32 arrays zeroed by one kernel launch
struct ZeroSpans {
    int32_t *p[4];
    int n[4];
    int nspans;
};
hipError_t launch_zero(const ZeroSpans &z, hipStream_t st);

// up to four device arrays of 32-bit words -> one host staging block
struct Gather {
    const uint32_t *src[4];
    int nwords[4];
    int dst[4];      // word offset of each array in `out`
    uint32_t *out;   // mapped pinned host memory
    int nspans;
};
hipError_t launch_gather(const Gather &g, hipStream_t st);

// Device buffer that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t want);
    template <class T> T *as() const { return static_cast<T *>(ptr); }
    ~DevBuf();
};

// Batched SW over pairs already resident on the device.  Lengths live on
// the device; maxq/maxr bound them (host-side validation and buffer sizing).
// Results: int32 score/start/end/alen and reversed consensus strings at
// cons_off[p] (capacity qlen + rlen + 2 each).  When `sync_sizes` is false
// the call allocates nothing and relies on `reserve_batch` having sized the
// workspace for (npairs, maxq, maxr); that form is graph-capturable.
struct BatchIO {
    const char *qseq;
    const int64_t *qoff;
    const int32_t *qlen;
    const char *rseq;
    const int64_t *roff;
    const int32_t *rlen;
    int npairs, maxq, maxr;
    int32_t *score, *start, *end, *alen;
    char *qcons, *rcons;
    const int64_t *cons_off;
    // 1: every reference byte is one of ACGTNacgtn (the caller knows it, as
    // the pipeline does from its index); with scores that cannot reach the
    // int16 limit either, no pair can take the int32 re-run
    int refs_plain = 0;
};

int validate_scoring(const Scoring &s, int maxq, int maxr, bool *force32);
// float scores of the reference's interfaces -> integers (each must be integral)
int scoring_from_floats(float match, float mismatch, float gap_open, float gap_ext, Scoring *sc);
int sw_batch_run(const BatchIO &io, const Scoring &sc, hipStream_t st, bool sync_sizes);
int reserve_batch(int npairs, int maxq, int maxr);

// Private batched-SW workspaces (device buffers, side streams), so that
// several batches can be in flight at once, each driven by its own host
// thread on its own stream.  sw_batch_run_in sizes the trace on the device
// (it synchronises `st`).
struct Workspace;
// side_streams: row-class kernels spread over that many private streams
// (0..3); 0 runs them one after another on the caller's stream
Workspace *workspace_new(int side_streams = 3);
void workspace_free(Workspace *w);
// persistent waves per CU of this workspace's row-class kernels, in place of
// the process-wide tuning (-1: follow it)
void workspace_set_waves(Workspace *w, int waves_per_cu);
int64_t workspace_bytes(const Workspace *w);
int sw_batch_run_in(Workspace *w, const BatchIO &io, const Scoring &sc, hipStream_t st);
// Capturable form: no host synchronisation, no allocation (reserve_batch_in
// first, for at least this npairs / maxq / maxr); every row class is launched
// and empty ones exit at once.  sw_batch_check_in reads the length-error flag
// afterwards (it synchronises `st`).
int reserve_batch_in(Workspace *w, int npairs, int maxq, int maxr);
int sw_batch_enqueue_in(Workspace *w, const BatchIO &io, const Scoring &sc, hipStream_t st);
int sw_batch_check_in(Workspace *w, hipStream_t st);
int64_t batch_bytes_estimate(int npairs, int maxq, int maxr);

// statistics of the last batched run (for tests and the bench)
struct BatchStats {
    int ntasks16, ntasks32, nrerun;
    int64_t trace_words;
};
BatchStats last_batch_stats();

// HIP-event phase timing of swift_sw_batch (binning, int16/int32 forward
// sweep, re-run pass, walk-back), in ms, recorded on the call's stream
void set_batch_timing(bool on);
const float *batch_timing_ms();

// tuning knobs set by the autotuner (see tune.cpp)
struct Tuning {
    int waves_per_cu = 0;       // cap on persistent waves per CU (0: occupancy limit)
    int int16 = 1;              // allow the int16 packed path
    int walk_deferred = 1;      // 1: walk-back kernels after each class's sweep (moves,
                                // then strings); 0: the walk fused into each wave task
    int coarse_rows = -1;       // 1: rows per lane rounded up to a multiple of 5 (at most
                                // 4 row classes); 0: every R its own class; -1: per call
};
// batches below this many pairs take the coarse row classes under coarse_rows -1
constexpr int kCoarsePairs = 131072;
constexpr int kCoarseSmallPairs = 16384;  // ... rounded to kCoarseRowsSmall below this
constexpr int kCoarseTinyPairs = 4096;    // ... and to one class (R = kRBins) below this
// ... but from this many pairs only when some query needs more than
// kCoarseRowsSmall rows per 16-lane group
constexpr int kWideShortPairs = 2048;
Tuning &tuning();

}  // namespace swift
