/*
 * swift_abi.h -- C ABI of libswift_amd.so, the MI355X (gfx950) implementation
 * of pgngp/swift's two-phase alignment hot path.
 *
 * Two kinds of entry points:
 *  - drop-ins that keep a reference function's name, signature, argument
 *    meaning and error behaviour (each cites the reference declaration it
 *    replaces);
 *  - swift_* entry points (batched, status-returning, stream-aware) that
 *    have no reference counterpart.
 *
 * Plain C types only (pointers, sizes, ints, floats).  Device pointers are
 * HIP device allocations; "stream" arguments are hipStream_t passed as
 * void* (NULL = the default stream).  swift_* functions return 0 on success
 * and a nonzero status otherwise (swift_last_error() describes it); the
 * drop-ins print the error and exit(1), as the reference's
 * PRINT_CUDA_ERROR does (src/common.h:155-167).
 *
 * Scores: the reference passes integer-valued floats (parsed with atoi,
 * src/main.cu:307-326); the drop-ins reject non-integral values.
 */
#ifndef SWIFT_ABI_H_
#define SWIFT_ABI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    SWIFT_OK = 0,
    SWIFT_ERR_INVALID = 1,
    SWIFT_ERR_HIP = 2,
    SWIFT_ERR_NO_INDEX = 3,
    SWIFT_ERR_IO = 4,
    SWIFT_ERR_CAPACITY = 5
};

const char *swift_last_error(void);
const char *swift_version(void);

/* ---------------------------------------------------- device helpers (no
 * reference counterpart: plumbing for hosts that bind the ABI directly) */
int swift_device_count(int *count);
int swift_device_alloc(void **ptr, size_t bytes);
int swift_device_free(void *ptr);
int swift_memcpy_htod(void *dst, const void *src, size_t bytes);
int swift_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int swift_memset_device(void *dst, int value, size_t bytes);
int swift_device_synchronize(void);

/* ---------------------------------------------------- phase 2: plain SW */

/*
 * Drop-in for alignSequences (src/align.h:75-80, src/align.cu:2904-2947).
 * querySeq_d holds queriesPerIteration queries at stride querySeqLength+1,
 * refSeq_d the references at stride refSeqLength+1 (NUL padded).  Every one
 * of the querySeqLength query bytes is a row and reference columns stop at
 * the first NUL, as in smithWaterman_cpu.  Outputs per pair: score, 1-based
 * start/end column of the alignment in the reference, alignment length and
 * the two consensus strings written in walk order (reversed alignment) at
 * stride querySeqLength+refSeqLength+2.  Results follow the reference CPU
 * path (computeMaxScore_cpu/backtrack_cpu), including its direction rules at
 * zero-score cells.  refsPerIteration and maxThreadsPerBlock are accepted
 * and ignored.  Launches on the default stream; results are visible after
 * the next synchronising copy, as with the reference launch.
 */
void alignSequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
                    float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
                    float *score_d, int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                    char *queryConsensus_d, char *refConsensus_d, int queriesPerIteration,
                    int refsPerIteration, int maxThreadsPerBlock);

/* Status-returning form of alignSequences on a given stream (no reference
 * counterpart). */
int swift_align_sequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
                          float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
                          float *score_d, int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                          char *queryConsensus_d, char *refConsensus_d, int queriesPerIteration,
                          void *stream);

/* ---------------------------------------------------- phase 2: batched SW
 * No reference counterpart (the reference aligns fixed 100 x 300 tiles
 * only).  npairs query/reference pairs of uneven lengths, all on the device:
 * pair p's query is qseq_d[qoff_d[p] .. +qlen_d[p]), its reference
 * rseq_d[roff_d[p] .. +rlen_d[p]).  maxq/maxr bound the lengths
 * (maxq <= 320, maxr <= 480).  Scores follow smithWaterman_cpu exactly.
 * Outputs: score/start/end/alen per pair and the consensus strings in walk
 * order at cons_off_d[p] (room for qlen+rlen+2 bytes each, NUL terminated).
 * Pairs are length-binned on the device, run two per lane in packed int16
 * by a persistent kernel, and any pair that could leave the int16 range is
 * re-run in int32.  The call synchronises the stream to size its trace
 * buffer; results are complete when it returns. */
int swift_sw_batch(const char *qseq_d, const int64_t *qoff_d, const int32_t *qlen_d,
                   const char *rseq_d, const int64_t *roff_d, const int32_t *rlen_d, int npairs,
                   int maxq, int maxr, int match, int mismatch, int gap_open, int gap_ext,
                   int32_t *score_d, int32_t *start_d, int32_t *end_d, int32_t *alen_d,
                   char *qcons_d, char *rcons_d, const int64_t *cons_off_d, void *stream);

/* Counters of the last swift_sw_batch call: wave tasks run with int16 and
 * int32 lanes, pairs re-run in int32, trace words used. */
int swift_sw_batch_stats(int *ntasks16, int *ntasks32, int *nrerun, int64_t *trace_words);

/* Phase timing of swift_sw_batch measured with HIP events on its stream:
 * enable != 0 turns recording on for later calls; ms4 receives the last
 * call's binning, forward sweep (all row classes, int16 and int32), int32
 * re-run pass and walk-back times in ms. */
int swift_sw_batch_timing(int enable, float *ms4);

/* Tuning knobs (set by the autotuner): persistent waves per CU and whether
 * the int16 path may be used. */
int swift_set_tuning(int waves_per_cu, int allow_int16);

#ifdef __cplusplus
}
#endif
#endif /* SWIFT_ABI_H_ */
