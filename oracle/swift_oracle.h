/*
 * swift_oracle.h -- CPU restatement of the reference's CPU alignment path
 * (pgngp/swift v0.11.2, `main -cpu`: searchQueries3 + alignQueries_cpu_mt).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (swift_amd/) never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Scores are integers: the reference stores them in float, but its CLI parses
 * them with atoi (src/main.cu:307-326), so every value it can compute is an
 * exactly representable integer and int arithmetic reproduces it bit for bit.
 */
#ifndef SWIFT_ORACLE_H_
#define SWIFT_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- glibc rand() replay (TYPE_3 additive feedback generator) ----------
 * The reference draws its tie-breaking random numbers with rand()
 * (src/array.cu:70, src/align.cu:891).  This is a restatement of glibc's
 * published random_r algorithm so a single-threaded run can be replayed. */
typedef struct {
    uint32_t r[34];
    int pos; /* index of the oldest of the 34 most recent words */
} orc_rng;

void orc_rng_seed(orc_rng *g, unsigned int seed);
int orc_rng_next(orc_rng *g);

/* ---- k-mer codec and reverse complement -------------------------------- */
/* getHash (src/preprocess.cu:70-91): sum_i 4^i * code(s[i]) */
int orc_hash(const char *s, int k);
/* qryGetReverseComplement (src/query.cu:265-282) */
void orc_revcomp(const char *seq, int len, char *out);

/* ---- phase-1 index (lookupTable5Create, src/lookupTable5.cu:101-252) --- */
typedef struct {
    int k, thr;
    int64_t nhash;      /* 4^k */
    int32_t *start;     /* _lookupTable: 1-based bucket start, 0 when empty */
    int32_t *count;     /* _numRepeatsPerTuple: all occurrences */
    int32_t *actual;    /* _numActualRepeatsPerTuple: count if <= thr else 0 */
    int64_t total;      /* _numTotalTuples */
    int8_t *refIdx;     /* [total + 1], element 0 is the -1 sentinel */
    int32_t *refPos;    /* [total + 1], element 0 is the -1 sentinel */
} orc_index;

int orc_index_build(orc_index *ix, int nseq, const char *const *seqs,
                    const int64_t *lens, int k, int thr);
void orc_index_free(orc_index *ix);

/* ---- hit clustering (mapHits6GetBestHits, src/mapHits6.cu:40-70) -------
 * tie_mode ORC_TIE_FIRST: when more than maxHits clusters share the biggest
 * size, keep the first maxHits in sorted order (deterministic).
 * tie_mode ORC_TIE_GLIBC: draw them with the replayed rand() exactly as
 * arrGetRandomNums (src/array.cu:61-77) does.  The inputs are sorted in
 * place, as the reference does.  *nBiggest receives the number of clusters
 * of the biggest size (the size of the tie set). */
enum { ORC_TIE_FIRST = 0, ORC_TIE_GLIBC = 1 };
int orc_best_hits(int maxHits, int8_t *refIdx, int32_t *shift,
                  int32_t *refPos, int n, int tie_mode, orc_rng *rng,
                  int8_t *oIdx, int32_t *oShift, int32_t *oPos,
                  int *nBiggest);

/* lookupTable5MapQry (src/lookupTable5.cu:557-603) for one strand; the
 * caller passes the forward read or its reverse complement. */
int orc_map_query(const orc_index *ix, const char *seq, int len, int maxHits,
                  int tie_mode, orc_rng *rng, int8_t *oIdx, int32_t *oShift,
                  int32_t *oPos, int *nBiggest, int *nHits);

/* ---- region fetch (align.cu:2388,2435 + getRefSeq2, input.cu:348-403) ---
 * Writes the <= winLen bases starting at max(0, refPos - before) of the
 * sequence, leading/trailing N trimmed as getRefSeq2 does, NUL terminated.
 * Returns the region length; *refDistance receives the window start. */
int orc_region(const char *seq, int64_t seqLen, int64_t refPos, int before,
               int winLen, char *out, int64_t *refDistance);

/* ---- Smith-Waterman (smithWaterman_cpu, src/smithWaterman.cu:44-335) ---
 * Rows = query (all qryLen bytes), columns = reference up to its first NUL
 * or refLen.  Direction codes (dirOut, optional, (qryLen+1)*(refLen+1)
 * bytes, row-major): 0 start, 1 left, 2 up, 3 diagonal. */
typedef struct {
    int score, start, end, len, maxRow, maxCol;
} orc_sw_result;

void orc_sw(const char *ref, int refLen, const char *qry, int qryLen,
            int match, int mismatch, int gapOpen, int gapExt,
            orc_sw_result *res, char *refCons, char *qryCons,
            uint8_t *dirOut);

/* Many independent pairs on nthreads threads (CPU baseline and tests).
 * Sequences are packed back to back; cons buffers have stride consStride
 * (>= qlen + rlen + 2). */
void orc_sw_batch(int npairs, const char *qry, const int64_t *qoff,
                  const int32_t *qlen, const char *ref, const int64_t *roff,
                  const int32_t *rlen, int match, int mismatch, int gapOpen,
                  int gapExt, int32_t *score, int32_t *start, int32_t *end,
                  int32_t *alen, char *qcons, char *rcons,
                  const int64_t *consOff, int nthreads);

/* ---- best-hit choice (chooseBestAlignment, src/align.cu:844-905) --------
 * Records of one read form one group (the '*' strand suffix is ignored by
 * qryNameCmp, src/align.cu:1501-1519).  keep[i] = 1 for the chosen record
 * of each group, 0 otherwise.  nTied[i] = size of the tie set of i's group. */
void orc_choose_best(int n, const int32_t *group, const int32_t *score,
                     int tie_mode, orc_rng *rng, int32_t *keep,
                     int32_t *nTied);

/* ---- whole CPU pipeline (searchQueries3 -> alignQueries_cpu_mt) --------
 * Reads are processed in order, forward strand then reverse complement
 * (src/search.cu:325-357, 1479-1488).  Each phase-1 hit becomes one record
 * aligned against its 300-bp window (src/align.cu:2424-2495).  Records are
 * written in that order.  Returns the number of records (-1 on overflow of
 * maxRecords).  consStride >= maxReadLen + winLen + 2. */
typedef struct {
    int32_t read, strand, refIdx, shift, refPos, nBiggest;
    int64_t refDistance;
    int32_t score, start, end, len, keep, nTied;
} orc_record;

int orc_pipeline(const orc_index *ix, int nseq, const char *const *seqs,
                 const int64_t *seqLens, int nreads, const char *reads,
                 const int64_t *readOff, const int32_t *readLen, int maxHits,
                 int match, int mismatch, int gapOpen, int gapExt,
                 int tie_mode, unsigned int seed, int maxRecords,
                 orc_record *rec, char *rcons, char *qcons, int consStride);

#ifdef __cplusplus
}
#endif
#endif /* SWIFT_ORACLE_H_ */
