// tune.cpp -- autotuner of the batched Smith-Waterman path (no reference
// counterpart).  The knobs are the batched kernels' launch shape and lane
// width: persistent waves per CU (occupancy of the row-class kernels), packed
// int16 lanes on/off (two pairs per lane, int32 re-run of overflow suspects;
// off = every pair in int32), where the traceback walk runs (fused into each
// wave task, or deferred to walk kernels), and whether rows per lane are
// rounded up to coarse row classes.  Results are
// identical under every setting, so the tuner only times: each candidate
// runs the caller's own pairs (or a prefix of them) `reps` times after a
// warm-up, timed with HIP events on the caller's stream, and the fastest
// setting is kept and applied.
#include <cfloat>
#include <vector>

#include "../../include/swift_abi.h"
#include "sw_host.hpp"

namespace swift {

struct Candidate {
    int waves_per_cu, int16, walk, coarse;
};

static std::vector<Candidate> candidates(int cu_waves_max) {
    std::vector<Candidate> c;
    const int waves[] = {0, 8, 12, 16, 20, 24};
    for (int walk = 0; walk < 2; ++walk)
        for (int w : waves)
            if (w == 0 || w < cu_waves_max) c.push_back({w, 1, walk, 0});
    for (int w : {0, 8}) {
        c.push_back({w, 1, 1, 1});  // coarse row classes
        c.push_back({w, 1, 0, 1});
    }
    c.push_back({0, 0, 1, 0});  // int32 lanes only
    c.push_back({0, 0, 0, 0});
    return c;
}

}  // namespace swift

using namespace swift;

extern "C" {

int swift_autotune_batch(const char *qseq_d, const int64_t *qoff_d, const int32_t *qlen_d,
                         const char *rseq_d, const int64_t *roff_d, const int32_t *rlen_d,
                         int npairs, int maxq, int maxr, int match, int mismatch, int gap_open,
                         int gap_ext, int32_t *score_d, int32_t *start_d, int32_t *end_d,
                         int32_t *alen_d, char *qcons_d, char *rcons_d,
                         const int64_t *cons_off_d, int reps, void *stream, int *best_waves,
                         int *best_int16, int *best_walk, float *best_ms, float *table_ms,
                         int table_cap, int *ntable) {
    if (npairs <= 0 || reps < 1) {
        set_error("autotune needs pairs and reps >= 1");
        return kErrInvalid;
    }
    hipStream_t st = (hipStream_t)stream;
    BatchIO io{qseq_d, qoff_d, qlen_d, rseq_d, roff_d, rlen_d, npairs, maxq, maxr,
               score_d, start_d, end_d, alen_d, qcons_d, rcons_d, cons_off_d};
    Scoring sc{match, mismatch, gap_open, gap_ext};
    const Tuning saved = tuning();
    hipEvent_t e0, e1;
    SWIFT_HIP_TRY(hipEventCreate(&e0));
    SWIFT_HIP_TRY(hipEventCreate(&e1));
    const auto cand = candidates(32);
    float best = FLT_MAX;
    Candidate pick{saved.waves_per_cu, saved.int16, saved.walk_deferred, saved.coarse_rows};
    int k = 0, rc = kOk;
    for (const Candidate &c : cand) {
        tuning().waves_per_cu = c.waves_per_cu;
        tuning().int16 = c.int16;
        tuning().walk_deferred = c.walk;
        tuning().coarse_rows = c.coarse;
        if ((rc = sw_batch_run(io, sc, st, true))) break;  // warm-up (sizes buffers)
        float ms = 0;
        for (int r = 0; r < reps && rc == kOk; ++r) {
            float t = 0;
            if (hipEventRecord(e0, st) != hipSuccess) rc = kErrHip;
            if (rc == kOk) rc = sw_batch_run(io, sc, st, true);
            if (rc == kOk && hipEventRecord(e1, st) != hipSuccess) rc = kErrHip;
            if (rc == kOk && hipEventSynchronize(e1) != hipSuccess) rc = kErrHip;
            if (rc == kOk && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = kErrHip;
            ms += t;
        }
        if (rc) break;
        ms /= reps;
        if (table_ms && k < table_cap) {
            table_ms[5 * k + 0] = (float)c.waves_per_cu;
            table_ms[5 * k + 1] = (float)c.int16;
            table_ms[5 * k + 2] = (float)c.walk;
            table_ms[5 * k + 3] = (float)c.coarse;
            table_ms[5 * k + 4] = ms;
        }
        ++k;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        tuning() = saved;
        if (rc == kErrHip) set_error("autotune: HIP event failure");
        return rc;
    }
    tuning().waves_per_cu = pick.waves_per_cu;
    tuning().int16 = pick.int16;
    tuning().walk_deferred = pick.walk;
    // the row-class choice is kept as "automatic" (-1) where the automatic
    // rule picks the same for a batch of this size: the tuned batch gets its
    // pick and calls of other sizes (serving slots, small batches) keep theirs
    const int auto_coarse = npairs < kCoarsePairs ? 1 : 0;
    tuning().coarse_rows = pick.coarse == auto_coarse ? -1 : pick.coarse;
    if (best_waves) *best_waves = pick.waves_per_cu;
    if (best_int16) *best_int16 = pick.int16;
    if (best_walk) *best_walk = pick.walk;
    if (best_ms) *best_ms = best;
    if (ntable) *ntable = k;
    return kOk;
}

}  // extern "C"
