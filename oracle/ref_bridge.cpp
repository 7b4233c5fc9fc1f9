/*
 * ref_bridge.cpp -- C entry points into the reference's own CPU functions,
 * compiled from the reference sources by oracle/Makefile into
 * oracle/_ref/libswiftref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  It lets the tests check the oracle restatement
 * against the reference itself, and lets bench.py time the reference's
 * smithWaterman_cpu as the CPU baseline.  The declarations below restate the
 * reference signatures (src/smithWaterman.h:62-66, src/mapHits6.h:34,
 * src/preprocess.h:21-23, src/query.h:46, src/refMap.h, src/input.h,
 * src/output.h) so the bridge needs no reference header.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned int uint;
typedef unsigned long ulong;

/* reference functions (C++ linkage, as nvcc compiled them) */
void smithWaterman_cpu(float *HMatrix, int *rowBacktracker,
                       int *colBacktracker, char *refSeq, int refLen,
                       char *qrySeq, int qryLen, float match, float mismatch,
                       float gapOpenPenalty, float gapExtPenalty,
                       float *alignScore, int *startPos, int *endPos,
                       int *alignLen, char *refConsensus, char *qryConsensus);
int mapHits6GetBestHits(int numBestHits, char *refIdx, int *shift,
                        int *refPos, int size, char *refIdx_bestMatch,
                        int *shift_bestMatch, int *refPos_bestMatch);
void preprocessCreate();
void preprocessDelete();
int getHash(char *s, int length);
void qryGetReverseComplement(char *seq, int seqLength, char *revComp);
void refMapCreate(const char *file);
void refMapFree();
void refMapGetOffsets(int index, uint pos, uint *nameOffset, uint *seqOffset);
void getRefSeq2(FILE *filePtr, ulong fileOffset, char *seq, int maxLength);
void getSeq(FILE *filePtr, ulong fileOffset, char *seq, int maxLength);
void getSeqName(FILE *filePtr, ulong fileOffset, char *name);
void printOutput(int numQueries, char *queryName, char *refName, float *score,
                 int *refDistance, int *alignStart, int *alignEnd,
                 int *alignLength, int combinedLength, char *queryConsensus,
                 char *refConsensus, FILE *outputFilePtr);
void printOutputSam(int numQueries, char *queryName, char *refName, float *score,
        int *alignStart, int *alignEnd, int *alignLength, int combinedLength,
        char *queryConsensus, char *refConsensus, FILE *outputFilePtr);
void printCigar(char *queryConsensus, char *refConsensus, int alignLength,
                FILE *outputFilePtr);

extern "C" {

/* One pair, caller matrices zeroed as getAlignments does
 * (src/align.cu:2519-2548). */
void ref_sw_one(const char *ref, int refLen, const char *qry, int qryLen,
                float match, float mismatch, float gapOpen, float gapExt,
                float *score, int *start, int *end, int *alen, char *refCons,
                char *qryCons)
{
    size_t cells = (size_t) (refLen + 1) * (qryLen + 1);
    float *H = (float *) calloc(cells, sizeof(float));
    int *rb = (int *) calloc(cells, sizeof(int));
    int *cb = (int *) calloc(cells, sizeof(int));
    char *r = (char *) calloc((size_t) refLen + 1, 1);
    char *q = (char *) calloc((size_t) qryLen + 1, 1);
    memcpy(r, ref, (size_t) refLen);
    memcpy(q, qry, (size_t) qryLen);
    memset(refCons, 0, (size_t) (refLen + qryLen + 2));
    memset(qryCons, 0, (size_t) (refLen + qryLen + 2));
    *score = 0.0f;
    *start = 0;
    *end = 0;
    *alen = 0;
    smithWaterman_cpu(H, rb, cb, r, refLen, q, qryLen, match, mismatch,
                      gapOpen, gapExt, score, start, end, alen, refCons,
                      qryCons);
    free(H);
    free(rb);
    free(cb);
    free(r);
    free(q);
}

/* One pair through smithWaterman_cpu with the caller's matrices
 * ((qryLen+1) x (refLen+1), zeroed by the caller) and initial *start, so a
 * test can compare every matrix cell and output the reference leaves. */
void ref_sw_matrices(const char *ref, int refLen, const char *qry, int qryLen,
                     float match, float mismatch, float gapOpen, float gapExt,
                     float *H, int *rowBt, int *colBt, float *score, int *start,
                     int *end, int *alen, char *refCons, char *qryCons)
{
    char *r = (char *) calloc((size_t) refLen + 1, 1);
    char *q = (char *) calloc((size_t) qryLen + 1, 1);
    memcpy(r, ref, (size_t) refLen);
    memcpy(q, qry, (size_t) qryLen);
    smithWaterman_cpu(H, rowBt, colBt, r, refLen, q, qryLen, match, mismatch,
                      gapOpen, gapExt, score, start, end, alen, refCons,
                      qryCons);
    free(r);
    free(q);
}

typedef struct {
    int tid, nthreads, npairs, maxCells;
    const char *qry, *ref;
    const long *qoff, *roff;
    const int *qlen, *rlen;
    float match, mismatch, gapOpen, gapExt;
    int *score, *start, *end, *alen;
    char *qcons, *rcons;
    const long *consOff;
    int nstrings;
} mt_arg;

static void *mt_worker(void *p)
{
    mt_arg *a = (mt_arg *) p;
    float *H = (float *) malloc((size_t) a->maxCells * sizeof(float));
    int *rb = (int *) malloc((size_t) a->maxCells * sizeof(int));
    int *cb = (int *) malloc((size_t) a->maxCells * sizeof(int));
    char rc[1200], qc[1200];
    for (int i = a->tid; i < a->npairs; i += a->nthreads) {
        int q = a->qlen[i], r = a->rlen[i];
        size_t cells = (size_t) (q + 1) * (r + 1);
        /* per-pair zeroing, as getAlignments does */
        memset(H, 0, cells * sizeof(float));
        memset(rb, 0, cells * sizeof(int));
        memset(cb, 0, cells * sizeof(int));
        memset(rc, 0, (size_t) (q + r + 2));
        memset(qc, 0, (size_t) (q + r + 2));
        float s = 0.0f;
        int st = 0, en = 0, al = 0;
        smithWaterman_cpu(H, rb, cb, (char *) a->ref + a->roff[i], r,
                          (char *) a->qry + a->qoff[i], q, a->match,
                          a->mismatch, a->gapOpen, a->gapExt, &s, &st, &en,
                          &al, rc, qc);
        a->score[i] = (int) s;
        if (a->start) {
            a->start[i] = st;
            a->end[i] = en;
            a->alen[i] = al;
        }
        if (a->qcons && i < a->nstrings) {
            memcpy(a->qcons + a->consOff[i], qc, (size_t) (q + r + 2));
            memcpy(a->rcons + a->consOff[i], rc, (size_t) (q + r + 2));
        }
    }
    free(H);
    free(rb);
    free(cb);
    return NULL;
}

/* Many pairs on nthreads threads; q + r must stay below 1198.  start, end
 * and alen (optional, NULL to skip) receive the coordinates; qcons/rcons
 * (optional) the consensus strings of the first nstrings pairs at
 * consOff[i] (room for q + r + 2 bytes each). */
void ref_sw_batch_mt2(int npairs, const char *qry, const long *qoff,
                      const int *qlen, const char *ref, const long *roff,
                      const int *rlen, float match, float mismatch,
                      float gapOpen, float gapExt, int *score, int *start,
                      int *end, int *alen, char *qcons, char *rcons,
                      const long *consOff, int nstrings, int nthreads)
{
    int maxCells = 1;
    for (int i = 0; i < npairs; ++i) {
        int c = (qlen[i] + 1) * (rlen[i] + 1);
        if (c > maxCells)
            maxCells = c;
    }
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    pthread_t th[256];
    mt_arg args[256];
    for (int t = 0; t < nthreads; ++t) {
        mt_arg a = {t, nthreads, npairs, maxCells, qry, ref, qoff, roff,
                    qlen, rlen, match, mismatch, gapOpen, gapExt, score,
                    start, end, alen, qcons, rcons, consOff, nstrings};
        args[t] = a;
        pthread_create(&th[t], NULL, mt_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
}

void ref_sw_batch_mt(int npairs, const char *qry, const long *qoff,
                     const int *qlen, const char *ref, const long *roff,
                     const int *rlen, float match, float mismatch,
                     float gapOpen, float gapExt, int *score, int nthreads)
{
    ref_sw_batch_mt2(npairs, qry, qoff, qlen, ref, roff, rlen, match, mismatch,
                     gapOpen, gapExt, score, NULL, NULL, NULL, NULL, NULL,
                     NULL, 0, nthreads);
}

int ref_best_hits(int maxHits, char *refIdx, int *shift, int *refPos, int n,
                  char *oIdx, int *oShift, int *oPos)
{
    return mapHits6GetBestHits(maxHits, refIdx, shift, refPos, n, oIdx,
                               oShift, oPos);
}

void ref_srand(unsigned int seed) { srand(seed); }
int ref_rand(void) { return rand(); }

int ref_hash(const char *s, int k)
{
    static int ready = 0;
    if (!ready) {
        preprocessCreate();
        ready = 1;
    }
    return getHash((char *) s, k);
}

void ref_revcomp(const char *seq, int len, char *out)
{
    qryGetReverseComplement((char *) seq, len, out);
}

/* Region fetch through the reference's file-offset map: window start
 * max(0, refPos - 100), 300 bases (src/align.cu:2435-2439). */
int ref_region(const char *fastaPath, int refIdx, int refPos, char *out,
               int *refDistance)
{
    uint nameOff = 0, seqOff = 0;
    int rd = refPos - 100;
    if (rd < 0)
        rd = 0;
    *refDistance = rd;
    refMapCreate(fastaPath);
    refMapGetOffsets(refIdx, (uint) rd, &nameOff, &seqOff);
    FILE *f = fopen(fastaPath, "r");
    if (!f) {
        refMapFree();
        return -1;
    }
    char buf[400];
    memset(buf, 0, sizeof(buf));
    getRefSeq2(f, seqOff, buf, 300);
    fclose(f);
    refMapFree();
    strcpy(out, buf);
    return (int) strlen(out);
}

/* getSeq / getSeqName on a file by path (the phase-2 drivers' query and
 * name fetch, src/input.cu:37-52,163-204). */
int ref_get_seq(const char *path, unsigned long offset, char *out, int maxLength)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return -1;
    out[0] = '\0';
    getSeq(f, offset, out, maxLength);
    fclose(f);
    return (int) strlen(out);
}

int ref_get_seq_name(const char *path, unsigned long offset, char *out)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return -1;
    getSeqName(f, offset, out);
    fclose(f);
    return (int) strlen(out);
}

/* refMapCreate + refMapGetOffsets (src/refMap.cu:59-155). */
int ref_map_offsets(const char *fastaPath, int refIdx, unsigned int pos,
                    unsigned int *nameOffset, unsigned int *seqOffset)
{
    uint n = 0, sOff = 0;
    refMapCreate(fastaPath);
    refMapGetOffsets(refIdx, pos, &n, &sOff);
    refMapFree();
    *nameOffset = n;
    *seqOffset = sOff;
    return 0;
}

/* printOutput of one record into a memory buffer; returns its length. */
int ref_format_record(const char *qryName, const char *refName, float score,
                      int refDistance, int start, int end, int alen,
                      const char *qryCons, const char *refCons, int combined,
                      char *out, int outCap)
{
    char qn[64], rn[64];
    memset(qn, 0, sizeof(qn));
    memset(rn, 0, sizeof(rn));
    strncpy(qn, qryName, 49);
    strncpy(rn, refName, 49);
    char *qc = (char *) calloc((size_t) combined, 1);
    char *rc = (char *) calloc((size_t) combined, 1);
    memcpy(qc, qryCons, (size_t) alen);
    memcpy(rc, refCons, (size_t) alen);
    FILE *f = tmpfile();
    printOutput(1, qn, rn, &score, &refDistance, &start, &end, &alen,
                combined, qc, rc, f);
    long n = ftell(f);
    rewind(f);
    if (n >= outCap)
        n = outCap - 1;
    size_t got = fread(out, 1, (size_t) n, f);
    out[got] = '\0';
    fclose(f);
    free(qc);
    free(rc);
    return (int) got;
}

/* printOutputSam of one record into a memory buffer; returns its length. */
int ref_format_sam(const char *qryName, const char *refName, float score, int start,
                   int end, int alen, const char *qryCons, const char *refCons,
                   int combined, char *out, int outCap)
{
    char qn[64], rn[64];
    memset(qn, 0, sizeof(qn));
    memset(rn, 0, sizeof(rn));
    strncpy(qn, qryName, 49);
    strncpy(rn, refName, 49);
    char *qc = (char *) calloc((size_t) combined, 1);
    char *rc = (char *) calloc((size_t) combined, 1);
    memcpy(qc, qryCons, (size_t) alen);
    memcpy(rc, refCons, (size_t) alen);
    FILE *f = tmpfile();
    printOutputSam(1, qn, rn, &score, &start, &end, &alen, combined, qc, rc, f);
    long n = ftell(f);
    rewind(f);
    if (n >= outCap)
        n = outCap - 1;
    size_t got = fread(out, 1, (size_t) n, f);
    out[got] = '\0';
    fclose(f);
    free(qc);
    free(rc);
    return (int) got;
}

int ref_format_cigar(const char *qryCons, const char *refCons, int alen,
                     char *out, int outCap)
{
    FILE *f = tmpfile();
    printCigar((char *) qryCons, (char *) refCons, alen, f);
    long n = ftell(f);
    rewind(f);
    if (n >= outCap)
        n = outCap - 1;
    size_t got = fread(out, 1, (size_t) n, f);
    out[got] = '\0';
    fclose(f);
    return (int) got;
}

} /* extern "C" */
