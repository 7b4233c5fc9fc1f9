// abi.cpp -- extern "C" entry points of libswift_amd.so (declared in
// include/swift_abi.h).
#include <cstdio>
#include <cstdlib>

#include "../../include/swift_abi.h"
#include "sw_host.hpp"
#include "sw_types.hpp"

namespace swift {
int align_sequences_status(char *q_d, char *r_d, int qlen, int rlen, float match, float mismatch,
                           float gap_open, float gap_ext, float *score_d, int *start_d,
                           int *end_d, int *alen_d, char *qcons_d, char *rcons_d, int npairs,
                           hipStream_t st);
}

using namespace swift;

extern "C" {

const char *swift_last_error(void) { return last_error(); }
const char *swift_version(void) { return "swift_amd 0.1 (gfx950)"; }

int swift_device_count(int *count) {
    SWIFT_HIP_TRY(hipGetDeviceCount(count));
    return kOk;
}
int swift_device_alloc(void **ptr, size_t bytes) {
    SWIFT_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return kOk;
}
int swift_device_free(void *ptr) {
    SWIFT_HIP_TRY(hipFree(ptr));
    return kOk;
}
int swift_memcpy_htod(void *dst, const void *src, size_t bytes) {
    SWIFT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return kOk;
}
int swift_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    SWIFT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return kOk;
}
int swift_memset_device(void *dst, int value, size_t bytes) {
    SWIFT_HIP_TRY(hipMemset(dst, value, bytes));
    return kOk;
}
int swift_device_synchronize(void) {
    SWIFT_HIP_TRY(hipDeviceSynchronize());
    return kOk;
}
int swift_stream_create(void **stream) {
    if (!stream) {
        set_error("null stream pointer");
        return kErrInvalid;
    }
    hipStream_t s = nullptr;
    SWIFT_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return kOk;
}
int swift_stream_synchronize(void *stream) {
    SWIFT_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return kOk;
}
int swift_stream_destroy(void *stream) {
    SWIFT_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return kOk;
}

void alignSequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
                    float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
                    float *score_d, int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                    char *queryConsensus_d, char *refConsensus_d, int queriesPerIteration,
                    int refsPerIteration, int maxThreadsPerBlock) {
    (void)refsPerIteration;
    (void)maxThreadsPerBlock;
    int rc = align_sequences_status(querySeq_d, refSeq_d, querySeqLength, refSeqLength, match,
                                    mismatch, gapOpenPenalty, gapExtPenalty, score_d,
                                    alignStart_d, alignEnd_d, alignLength_d, queryConsensus_d,
                                    refConsensus_d, queriesPerIteration, nullptr);
    if (rc != kOk) {
        fprintf(stderr, "Error: %s:%s: %s\n", __FILE__, __func__, last_error());
        exit(1);
    }
}

int swift_align_sequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
                          float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
                          float *score_d, int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                          char *queryConsensus_d, char *refConsensus_d, int queriesPerIteration,
                          void *stream) {
    return align_sequences_status(querySeq_d, refSeq_d, querySeqLength, refSeqLength, match,
                                  mismatch, gapOpenPenalty, gapExtPenalty, score_d, alignStart_d,
                                  alignEnd_d, alignLength_d, queryConsensus_d, refConsensus_d,
                                  queriesPerIteration, (hipStream_t)stream);
}

int swift_sw_batch(const char *qseq_d, const int64_t *qoff_d, const int32_t *qlen_d,
                   const char *rseq_d, const int64_t *roff_d, const int32_t *rlen_d, int npairs,
                   int maxq, int maxr, int match, int mismatch, int gap_open, int gap_ext,
                   int32_t *score_d, int32_t *start_d, int32_t *end_d, int32_t *alen_d,
                   char *qcons_d, char *rcons_d, const int64_t *cons_off_d, void *stream) {
    BatchIO io{qseq_d, qoff_d, qlen_d, rseq_d, roff_d, rlen_d, npairs, maxq, maxr,
               score_d, start_d, end_d, alen_d, qcons_d, rcons_d, cons_off_d};
    Scoring sc{match, mismatch, gap_open, gap_ext};
    return sw_batch_run(io, sc, (hipStream_t)stream, true);
}

int swift_sw_batch_stats(int *ntasks16, int *ntasks32, int *nrerun, int64_t *trace_words) {
    BatchStats s = last_batch_stats();
    if (ntasks16) *ntasks16 = s.ntasks16;
    if (ntasks32) *ntasks32 = s.ntasks32;
    if (nrerun) *nrerun = s.nrerun;
    if (trace_words) *trace_words = s.trace_words;
    return kOk;
}

int swift_sw_batch_timing(int enable, float *ms4) {
    set_batch_timing(enable != 0);
    if (ms4)
        for (int i = 0; i < 4; ++i) ms4[i] = batch_timing_ms()[i];
    return kOk;
}

int swift_set_tuning(int waves_per_cu, int allow_int16, int walk_deferred) {
    if (waves_per_cu < 0 || waves_per_cu > 64) {
        set_error("waves_per_cu out of range");
        return kErrInvalid;
    }
    tuning().waves_per_cu = waves_per_cu;
    tuning().int16 = allow_int16 ? 1 : 0;
    tuning().walk_deferred = walk_deferred ? 1 : 0;
    return kOk;
}

int swift_get_tuning(int *waves_per_cu, int *allow_int16, int *walk_deferred) {
    if (waves_per_cu) *waves_per_cu = tuning().waves_per_cu;
    if (allow_int16) *allow_int16 = tuning().int16;
    if (walk_deferred) *walk_deferred = tuning().walk_deferred;
    return kOk;
}

int swift_set_coarse_rows(int mode) {
    if (mode < -1 || mode > 1) {
        set_error("coarse row class mode must be -1, 0 or 1");
        return kErrInvalid;
    }
    tuning().coarse_rows = mode;
    return kOk;
}

int swift_get_coarse_rows(int *mode) {
    if (mode) *mode = tuning().coarse_rows;
    return kOk;
}

int swift_set_plain_chunk(int max_pairs) {
    if (max_pairs < 0) {
        set_error("plain chunk size must be >= 0");
        return kErrInvalid;
    }
    set_plain_chunk_pairs(max_pairs);
    return kOk;
}

}  // extern "C"
