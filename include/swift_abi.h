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
#include <stdio.h>

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
    SWIFT_ERR_CAPACITY = 5,
    SWIFT_ERR_TIMEOUT = 6
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
/* A non-blocking HIP stream (hipStream_t as void*) at the default priority. */
int swift_stream_create(void **stream);
int swift_stream_synchronize(void *stream);
int swift_stream_destroy(void *stream);

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

/*
 * Drop-in for smithWaterman_cpu (src/smithWaterman.h:62-66,
 * src/smithWaterman.cu:44-335): one pair, host buffers.  The caller owns the
 * (qryLen+1) x (refLen+1) matrices HMatrix, rowBacktracker and
 * colBacktracker (row-major, index row * (refLen+1) + col) and zeroes them,
 * as the reference's callers do (src/align.cu:2519-2548).  On return they
 * hold what the reference leaves there: for rows 1..qryLen and the columns
 * before the reference's first NUL, the cell's score and the row/column of
 * the cell its best candidate came from (itself for a zero start); other
 * cells are not written.  Score, 1-based start/end column, length and the
 * two consensus strings in walk order (alignLen bytes each, no NUL
 * written); startPos is left unchanged when the alignment is empty, and
 * endPos is then startPos - 1.  The pair runs on the GPU; calls from
 * several host threads run concurrently.  qryLen <= 4096.  Prints and exits
 * on error.
 */
void smithWaterman_cpu(float *HMatrix, int *rowBacktracker, int *colBacktracker, char *refSeq,
                       int refLen, char *qrySeq, int qryLen, float match, float mismatch,
                       float gapOpenPenalty, float gapExtPenalty, float *alignScore, int *startPos,
                       int *endPos, int *alignLen, char *refConsensus, char *qryConsensus);

/* Status-returning form of alignSequences on a given stream (no reference
 * counterpart). */
int swift_align_sequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
                          float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
                          float *score_d, int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                          char *queryConsensus_d, char *refConsensus_d, int queriesPerIteration,
                          void *stream);

/* swift_align_sequences that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, limits and results as swift_align_sequences;
 * the traceback scratch (trace, best cells, walk moves) is a fixed 64 MiB
 * device array per GPU (held until the process ends) instead of
 * library-allocated memory, and the pairs run in chunks that fit it.  SWIFT_ERR_CAPACITY when one pair's traceback
 * would not fit (not reachable within the plain kernel's length limits).
 * Every call on one GPU shares that GPU's array: calls whose kernels
 * overlap on one GPU race, so issue them on one stream or wait for one
 * before the next.  Launches only; results are complete once the stream is
 * synchronised. */
int swift_align_sequences_noalloc(char *querySeq_d, char *refSeq_d, int querySeqLength,
                                  int refSeqLength, float match, float mismatch,
                                  float gapOpenPenalty, float gapExtPenalty, float *score_d,
                                  int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                                  char *queryConsensus_d, char *refConsensus_d,
                                  int queriesPerIteration, void *stream);

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

/* Tuning knobs (set by the autotuner): persistent waves per CU (0 = the
 * occupancy limit), whether the packed int16 path may be used, and whether
 * the traceback walk runs fused into each wave task (0) or as one kernel
 * after the sweeps (1).  Results are identical in every setting. */
int swift_set_tuning(int waves_per_cu, int allow_int16, int walk_deferred);
int swift_get_tuning(int *waves_per_cu, int *allow_int16, int *walk_deferred);

/* Row classes of a batched call: each pair runs in the kernel of its rows
 * per lane R = ceil(query / 16).  1 = coarse classes: R rounded up to a
 * multiple of 5 (at most 4 class kernels: lower latency for small batches,
 * some padding rows); when the call reads its sizes back (not under graph
 * capture), to a multiple of 10 under 16384 pairs, and under 4096 pairs one
 * class in the wide form (one 64-lane group per wave, 5 rows per lane, two
 * pairs per task).  0 = every R its own class; -1 (default) = 1 for batches
 * under 131072 pairs, inside graph capture and on serving slots, else 0.
 * No counterpart in the reference; results are identical in every mode. */
int swift_set_coarse_rows(int mode);

/* Pairs per chunk of a swift_align_sequences call (0, the default: as many
 * as 4 GiB of traceback scratch holds; the chunks run one after another).
 * No counterpart in the reference; results are identical for every size. */
int swift_set_plain_chunk(int max_pairs);
int swift_get_coarse_rows(int *mode);

/* Autotuner of the batched path: times every candidate setting (waves per
 * CU in {max, 8, 12, 16, 20, 24} x walk fused/deferred with int16 lanes, the
 * coarse row classes, and int32-only lanes) on the given pairs, reps timed
 * runs each after a warm-up, with HIP events on `stream`, then applies the
 * fastest (swift_get_coarse_rows reports its row-class mode: -1 when the
 * automatic rule picks the same for a batch of this size).  Outputs are
 * written by the timed runs (identical under every setting).  table_ms
 * (optional, 5 floats per candidate: waves, int16, walk, coarse rows, ms)
 * receives up to table_cap rows; *ntable the number of candidates timed.
 * Changes the process-wide tuning: call it while no other batched call is
 * running. */
int swift_autotune_batch(const char *qseq_d, const int64_t *qoff_d, const int32_t *qlen_d,
                         const char *rseq_d, const int64_t *roff_d, const int32_t *rlen_d,
                         int npairs, int maxq, int maxr, int match, int mismatch, int gap_open,
                         int gap_ext, int32_t *score_d, int32_t *start_d, int32_t *end_d,
                         int32_t *alen_d, char *qcons_d, char *rcons_d,
                         const int64_t *cons_off_d, int reps, void *stream, int *best_waves,
                         int *best_int16, int *best_walk, float *best_ms, float *table_ms,
                         int table_cap, int *ntable);

/* ======================================================== phase 1: seeding
 * Reference drop-ins.  Same names, arguments and meaning as the reference's
 * CPU phase 1; the index they share lives in device memory. */

/* Query record (src/query.h:29-37). */
typedef struct queryStruct {
    char *name;
    int length;
    char *seq;
    char *revCompSeq;
    unsigned long nameOffset;
    unsigned long seqOffset;
} Query;

/* preprocessCreate / preprocessDelete (src/preprocess.h:21,29).  The codes
 * are fixed tables here; the calls exist for callers that make them. */
void preprocessCreate(void);
void preprocessDelete(void);
/* getHash (src/preprocess.h:23, src/preprocess.cu:70-91): sum_i 4^i code(s[i]),
 * A/a 0, C/c 1, G/g 2, T/t 3, any other byte 0. */
int getHash(char *s, int length);
/* qryGetReverseComplement (src/query.cu:265-282): reversed, A->T C->G G->C,
 * any other byte -> A; revComp receives seqLength+1 bytes (NUL). */
void qryGetReverseComplement(char *seq, int seqLength, char *revComp);

/* lookupTable5Create (src/lookupTable5.h:28-29): reads the reference FASTA
 * (any line width) and builds the non-overlapping k-mer index on the device;
 * tuples seen more than tupleIgnoreThreshold times are dropped.
 * *totalTuples receives the number of kept tuples.  Exits on error, like the
 * reference. */
void lookupTable5Create(const char *refFile, int seedLen, int maxHitsPerQry,
                        int tupleIgnoreThreshold, int *totalTuples);
/* lookupTable5MapQry (src/lookupTable5.h:42-43): best hits of one strand
 * (query->seq, or query->revCompSeq when isRevComp == 1) over qryLen bases;
 * returns the number of hits written (<= maxHitsPerQry).  Thread safe: calls
 * from several host threads are serialised. */
int lookupTable5MapQry(const Query *query, unsigned int qryLen, unsigned int isRevComp,
                       char *refIdx_bestHits, int *shift_bestHits, int *refPos_bestHits);
/* lookupTable5Delete (src/lookupTable5.h:40). */
void lookupTable5Delete(void);
/* mapHits6GetBestHits (src/mapHits6.h:34): sorts the size hits in place by
 * (refIdx, shift, refPos), clusters equal (refIdx, shift) runs and writes the
 * first hit of each biggest cluster, at most numBestHits of them in sorted
 * order.  Where the reference draws among more tied clusters with rand(),
 * this keeps the first numBestHits.  Returns the count. */
int mapHits6GetBestHits(int numBestHits, char *refIdx, int *shift, int *refPos, int size,
                        char *refIdx_bestMatch, int *shift_bestMatch, int *refPos_bestMatch);
/* searchQueries3 (src/search.h:33-35): builds the index from refFileName,
 * maps both strands of the first numQrs reads of qryFileName on the device
 * and writes one match line per hit to matchFile,
 * "name\tstrand\trefIdx\tshift\trefPos\tnameOffset\tseqOffset" (reads in file
 * order, forward strand first).  qryLen and numThreads are accepted and
 * ignored.  *numMatches receives the line count. */
void searchQueries3(char *qryFileName, int numQrs, int qryLen, char *refFileName, int seedLen,
                    char *matchFile, unsigned int maxNumHits, int *numMatches, int numThreads,
                    int tupleIgnoreThres);
/* search2Queries (src/search2.h:26-27): the reference's GPU phase-1 driver.
 * Maps both strands of every read of qryFile and writes the same line format
 * with the shift column 0 (src/search2.cu:341-373); maxNumHits is taken as 3
 * whatever is passed (src/search2.cu:64).  The hits are those of
 * searchQueries3 (lookupTable5 + mapHits6 selection): the reference's
 * lookupTable13 kernel behind this driver draws its picks from an
 * uninitialised generator and has no deterministic output to reproduce.
 * Prints and exits on error, as the reference does. */
void search2Queries(char *qryFile, char *refFile, char *matchFile, int seedLen, int maxNumHits,
                    int *numMatches, int tupleIgnoreThres);

/* ---- status-returning phase-1 API (no reference counterpart) ----------- */

/* Which mapping tier a strand starts in.  -1 (default): the hash-cluster
 * tier when the index's mean bucket size times the k-mers of a
 * max_read_len read reaches 1024 hits (the wave tier would only count such
 * strands and pass them on), else the one-wave tier; 0: always the one-wave
 * tier first; 1: always the hash-cluster tier first (when the index has
 * 32-bit cluster ids).  Results are identical in every mode. */
int swift_set_map_mode(int mode);
int swift_get_map_mode(int *mode);

typedef struct swift_index swift_index;

/* Index over nseq sequences in host memory (sequence s = seq[off[s] ..
 * +len[s])), k = seed_len <= 15, nseq <= 127. */
int swift_index_build(const char *seq, const int64_t *off, const int64_t *len, int nseq,
                      int seed_len, int tuple_ignore_threshold, swift_index **out);
/* Index over a FASTA file read as lookupTable5Create reads it. */
int swift_index_build_fasta(const char *path, int seed_len, int tuple_ignore_threshold,
                            swift_index **out);
int swift_index_free(swift_index *ix);
/* nseq, kept tuples, 4^k and the largest kept bucket. */
int swift_index_info(const swift_index *ix, int *nseq, int64_t *total_tuples, int64_t *nhash,
                     int *max_bucket);
/* Copies the tables to host memory (any pointer may be NULL): start, count
 * and actual ([4^k]: _lookupTable, _numRepeatsPerTuple,
 * _numActualRepeatsPerTuple) and ref_idx/ref_pos ([total + 1], element 0
 * the -1 sentinel: _refIdx, _refPos). */
int swift_index_export(const swift_index *ix, int32_t *start, int32_t *count, int32_t *actual,
                       int8_t *ref_idx, int32_t *ref_pos);
/* Length of sequence s as indexed, and its name (FASTA builds; "" otherwise). */
int swift_index_sequence(const swift_index *ix, int s, int64_t *len, const char **name);

/* Maps both strands of nreads device-resident reads (read r =
 * reads_d[off_d[r] .. +len_d[r]), len <= max_read_len <= 2047).  Strand
 * s = 2r (forward) or 2r+1 (reverse complement): up to max_hits best hits
 * at [s * max_hits ..] of ref_idx_d/shift_d/ref_pos_d, their count in
 * nbest_d[s], the number of clusters tied at the biggest size in
 * nbiggest_d[s] and the k-mer hits gathered in nhits_d[s] (nbiggest_d and
 * nhits_d may be NULL).  Synchronises the stream before returning. */
int swift_map_reads(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                    const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                    int8_t *ref_idx_d, int32_t *shift_d, int32_t *ref_pos_d, int32_t *nbest_d,
                    int32_t *nbiggest_d, int32_t *nhits_d, void *stream);

/* Phase 1 -> phase 2 on the device: records (one per best hit, read order,
 * forward strand first, hits in cluster order) with the region window, the
 * Smith-Waterman result against it and the per-read best-hit choice.  All
 * pointers are device memory with room for `capacity` records
 * (>= 2 * nreads * max_hits); cons_stride >= max_read_len + 302. */
typedef struct {
    int32_t *nrecords;        /* [1] records written */
    int32_t *read, *strand, *ref_idx, *shift, *ref_pos, *nbiggest;
    int64_t *ref_distance;    /* window start: max(0, refPos - 100) */
    int32_t *score, *start, *end, *alen;
    int32_t *keep;            /* 1 for the record chooseBestAlignment keeps */
    int32_t *ntied;           /* records sharing the read's best score */
    char *qry_cons, *ref_cons; /* record i at i * cons_stride, walk order */
    int64_t cons_stride;
    int32_t capacity;
} swift_records;

/* Reads up to 320 bases; windows of 300 bases from max(0, refPos - 100)
 * (alignQueries_cpu_mt, src/align.cu:2363-2511; chooseBestAlignment,
 * src/align.cu:844-905).  Synchronises the stream before returning. */
int swift_align_reads(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                      const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                      int match, int mismatch, int gap_open, int gap_ext,
                      const swift_records *rec, void *stream);

/* Phase 1 -> phase 2 captured as one hipGraph for a fixed batch shape: the
 * read buffers (nreads slots; an unused slot holds a length-0 read) and the
 * record buffers are fixed device addresses; the caller rewrites their
 * contents between launches.  Scratch is reserved for the worst case when
 * the graph is made.  After a launch and a synchronisation of its stream,
 * swift_pipeline_graph_check reports length and capacity errors of that
 * run.  Records equal swift_align_reads on the same reads. */
typedef struct swift_graph swift_graph;
int swift_pipeline_graph_create(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                                const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                                int match, int mismatch, int gap_open, int gap_ext,
                                const swift_records *rec, swift_graph **out);
int swift_pipeline_graph_launch(swift_graph *g, void *stream);
int swift_pipeline_graph_check(swift_graph *g);
int swift_pipeline_graph_free(swift_graph *g);

/* ======================================================== output formats
 * Drop-ins for the reference's writers (src/output.h:29-55), same output
 * byte for byte: names at stride MAX_SEQ_NAME_LENGTH (50), consensus strings
 * in walk order at stride combinedLength; printOutput skips score < 0. */
void printOutput(int numQueries, char *queryName, char *refName, float *score, int *refDistance,
                 int *alignStart, int *alignEnd, int *alignLength, int combinedLength,
                 char *queryConsensus, char *refConsensus, FILE *outputFilePtr);
void printOutputSam(int numQueries, char *queryName, char *refName, float *score,
                    int *alignStart, int *alignEnd, int *alignLength, int combinedLength,
                    char *queryConsensus, char *refConsensus, FILE *outputFilePtr);
void printCigar(char *queryConsensus, char *refConsensus, int alignLength, FILE *outputFilePtr);

/* CIGAR strings of device records i < *rec->nrecords (printCigar's rule) at
 * cigar_d + i * cigar_stride (NUL terminated, cigar_stride >= 2 *
 * cons_stride), lengths in cigar_len_d; max_records bounds the count. */
int swift_records_cigar(const swift_records *rec, int max_records, char *cigar_d,
                        int64_t cigar_stride, int32_t *cigar_len_d, void *stream);
/* Writes pipeline records (host copies: a swift_records of host pointers) to
 * `path` as printOutput text (format 0) or SAM (format 1).  Record i is named
 * qry_names[read[i]] (+ '*' on the reverse strand) against the FASTA name of
 * its index sequence; only_kept writes the records the best-hit choice
 * keeps. */
int swift_write_records(const char *path, int format, int only_kept, const swift_records *rec,
                        int nrec, const char *const *qry_names, const swift_index *ix);

/* ======================================================== phase-2 drivers
 * Whole-phase drop-ins of the reference's phase-2 drivers: read the phase-1
 * match file (name\tstrand\trefIdx\tshift\trefPos\tnameOffset\tseqOffset),
 * fetch each query and its reference window (refPos - 100, up to 300 bases)
 * from the FASTA files by file offset, align every pair on the device and
 * write printOutput text (outFormat 0) or SAM (outFormat 1).  Scores must be
 * integral.  On error they print and exit(1), as the reference does on its
 * I/O errors. */
/* alignQueries3 (src/align.h:33-36): numMatches records; for text output
 * the best record of each read is kept (chooseBestAlignment; ties: the
 * first). */
void alignQueries3(char *queryFileName, int maxQryLen, char *refFileName, char *matchFile,
                   int outFormat, float match, float mismatch, float gapOpenPenalty,
                   float gapExtPenalty, const char *outputFile, int numMatches);
/* alignQueries_cpu_mt (src/align.h:54-56): every record, in the reference's
 * rounds of 8 match lines (the last round repeats the last match when the
 * line count is not a multiple of 8, as the reference's fgets loop does). */
void alignQueries_cpu_mt(char *qryFile, int numQrs, int qryLen, char *refFile, char *matchFile,
                         int outFormat, float match, float mismatch, float gapOpenPenalty,
                         float gapExtPenalty, const char *outputFile);
/* Status-returning form of both: gpuDriver 1 = alignQueries3's rules
 * (numMatches records), 0 = alignQueries_cpu_mt's (numMatches ignored). */
int swift_align_queries(const char *qryFile, int qryLen, const char *refFile,
                        const char *matchFile, int outFormat, float match, float mismatch,
                        float gapOpenPenalty, float gapExtPenalty, const char *outputFile,
                        int numMatches, int gpuDriver);

/* ======================================================== serving loop
 * No reference counterpart (the reference is a batch command-line tool).
 * A server keeps an index resident and aligns read batches submitted from
 * any host thread.  Each of nslots slots (a host thread with its own stream
 * and scratch) takes queued requests in arrival order and coalesces them
 * into one device batch while the batch's device-memory estimate stays
 * within mem_budget_bytes / nslots and max_batch_reads; slots run
 * concurrently, so batches admitted while others are on the GPU overlap
 * there.  A request's records are the same whatever it is batched with:
 * swift_align_reads over the request alone gives identical records. */
typedef struct swift_server swift_server;

int swift_server_create(const swift_index *ix, int nslots, int64_t mem_budget_bytes,
                        int max_batch_reads, int max_read_len, int max_hits, int match,
                        int mismatch, int gap_open, int gap_ext, swift_server **out);
/* Copies the reads (host memory) and queues them; *id names the request.
 * Requests larger than one slot are rejected (SWIFT_ERR_CAPACITY).  With
 * more than one slot, a request of more than 1024 reads is queued as up to
 * one part per slot (consecutive ranges of its reads); its records still
 * come back together, in read order. */
int swift_server_submit(swift_server *s, const char *reads, const int64_t *off, const int32_t *len,
                        int nreads, uint64_t *id);
/* Waits up to timeout_ms (< 0: no limit); *nrecords and the submit-to-done
 * latency in ms on success, SWIFT_ERR_TIMEOUT when still running. */
int swift_server_wait(swift_server *s, uint64_t id, int timeout_ms, int *nrecords,
                      double *latency_ms);
/* Copies a finished request's records into host arrays (swift_records with
 * host pointers; NULL fields are skipped; read indices are relative to the
 * request) and releases the request. */
int swift_server_fetch(swift_server *s, uint64_t id, const swift_records *out);
int swift_server_stats(swift_server *s, int64_t *completed, int64_t *batches,
                       int64_t *batched_reads, int64_t *max_coalesced, int64_t *device_bytes);
/* Stops the slots (queued requests are dropped) and frees the server. */
int swift_server_destroy(swift_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SWIFT_ABI_H_ */
