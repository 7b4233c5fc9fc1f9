// dropin_cxx.cpp -- the reference's drop-ins under C++ linkage.
//
// The reference compiles every file as C++ (nvcc on .cu sources) and its
// headers declare these functions without extern "C", so a reference caller
// (main.cu calling alignQueries3, searchQueries3 calling lookupTable5MapQry,
// getAlignments calling smithWaterman_cpu, ...) binds the C++ mangled name,
// e.g. _Z14alignSequencesPcS_iiffffPfPiS1_S1_S_S_iii.  Each definition below
// has the reference declaration's exact parameter types, so it carries that
// mangled name, and forwards to the extern "C" entry point of the same name.
// The C entry points are bound here under other names through asm labels: a
// C++ function and an extern "C" function of one name and signature cannot be
// declared in one translation unit.  Both kinds of symbol are exported.
//
// This file does not include swift_abi.h for that reason; the forwarded
// signatures are the ones declared there.
#include <cstdio>

// Query is the reference's typedef of struct queryStruct (src/query.h:29-37);
// pointers to it mangle by the struct's name.  Its layout is in swift_abi.h.
struct queryStruct;
typedef struct queryStruct Query;
typedef unsigned int uint;

#define SWIFT_C(name) __asm__(#name)

// ---- the extern "C" entry points (swift_abi.h) under local names ----------
extern "C" {
void c_alignSequences(char *, char *, int, int, float, float, float, float, float *, int *, int *,
                      int *, char *, char *, int, int, int) SWIFT_C(alignSequences);
void c_smithWaterman_cpu(float *, int *, int *, char *, int, char *, int, float, float, float,
                         float, float *, int *, int *, int *, char *,
                         char *) SWIFT_C(smithWaterman_cpu);
void c_alignQueries3(char *, int, char *, char *, int, float, float, float, float, const char *,
                     int) SWIFT_C(alignQueries3);
void c_alignQueries_cpu_mt(char *, int, int, char *, char *, int, float, float, float, float,
                           const char *) SWIFT_C(alignQueries_cpu_mt);
void c_preprocessCreate(void) SWIFT_C(preprocessCreate);
void c_preprocessDelete(void) SWIFT_C(preprocessDelete);
int c_getHash(char *, int) SWIFT_C(getHash);
void c_qryGetReverseComplement(char *, int, char *) SWIFT_C(qryGetReverseComplement);
void c_lookupTable5Create(const char *, int, int, int, int *) SWIFT_C(lookupTable5Create);
int c_lookupTable5MapQry(const Query *, unsigned int, unsigned int, char *, int *,
                         int *) SWIFT_C(lookupTable5MapQry);
void c_lookupTable5Delete(void) SWIFT_C(lookupTable5Delete);
int c_mapHits6GetBestHits(int, char *, int *, int *, int, char *, int *,
                          int *) SWIFT_C(mapHits6GetBestHits);
void c_searchQueries3(char *, int, int, char *, int, char *, unsigned int, int *, int,
                      int) SWIFT_C(searchQueries3);
void c_search2Queries(char *, char *, char *, int, int, int *, int) SWIFT_C(search2Queries);
void c_printOutput(int, char *, char *, float *, int *, int *, int *, int *, int, char *, char *,
                   FILE *) SWIFT_C(printOutput);
void c_printOutputSam(int, char *, char *, float *, int *, int *, int *, int, char *, char *,
                      FILE *) SWIFT_C(printOutputSam);
void c_printCigar(char *, char *, int, FILE *) SWIFT_C(printCigar);
}

#define SWIFT_EXPORT __attribute__((visibility("default")))

// ---- phase 2 ---------------------------------------------------------------
// src/align.h:75-80
SWIFT_EXPORT void alignSequences(char *querySeq_d, char *refSeq_d, int querySeqLength,
                                 int refSeqLength, float match, float mismatch,
                                 float gapOpenPenalty, float gapExtPenalty, float *score_d,
                                 int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                                 char *queryConsensus_d, char *refConsensus_d,
                                 int queriesPerIteration, int refsPerIteration,
                                 int maxThreadsPerBlock) {
    c_alignSequences(querySeq_d, refSeq_d, querySeqLength, refSeqLength, match, mismatch,
                     gapOpenPenalty, gapExtPenalty, score_d, alignStart_d, alignEnd_d,
                     alignLength_d, queryConsensus_d, refConsensus_d, queriesPerIteration,
                     refsPerIteration, maxThreadsPerBlock);
}

// src/smithWaterman.h:62-66
SWIFT_EXPORT void smithWaterman_cpu(float *HMatrix, int *rowBacktracker, int *colBacktracker,
                                    char *refSeq, int refLen, char *qrySeq, int qryLen,
                                    float match, float mismatch, float gapOpenPenalty,
                                    float gapExtPenalty, float *alignScore, int *startPos,
                                    int *endPos, int *alignLen, char *refConsensus,
                                    char *qryConsensus) {
    c_smithWaterman_cpu(HMatrix, rowBacktracker, colBacktracker, refSeq, refLen, qrySeq, qryLen,
                        match, mismatch, gapOpenPenalty, gapExtPenalty, alignScore, startPos,
                        endPos, alignLen, refConsensus, qryConsensus);
}

// src/align.h:33-36
SWIFT_EXPORT void alignQueries3(char *queryFileName, int maxQryLen, char *refFileName,
                                char *matchFile, int outFormat, float match, float mismatch,
                                float gapOpenPenalty, float gapExtPenalty,
                                const char *outputFile, int numMatches) {
    c_alignQueries3(queryFileName, maxQryLen, refFileName, matchFile, outFormat, match, mismatch,
                    gapOpenPenalty, gapExtPenalty, outputFile, numMatches);
}

// src/align.h:54-56
SWIFT_EXPORT void alignQueries_cpu_mt(char *qryFile, int numQrs, int qryLen, char *refFile,
                                      char *matchFile, int outFormat, float match,
                                      float mismatch, float gapOpenPenalty, float gapExtPenalty,
                                      const char *outputFile) {
    c_alignQueries_cpu_mt(qryFile, numQrs, qryLen, refFile, matchFile, outFormat, match,
                          mismatch, gapOpenPenalty, gapExtPenalty, outputFile);
}

// ---- phase 1 ---------------------------------------------------------------
// src/preprocess.h:21,23,29
SWIFT_EXPORT void preprocessCreate() { c_preprocessCreate(); }
SWIFT_EXPORT void preprocessDelete() { c_preprocessDelete(); }
SWIFT_EXPORT int getHash(char *s, int length) { return c_getHash(s, length); }

// src/query.h (qryGetReverseComplement)
SWIFT_EXPORT void qryGetReverseComplement(char *seq, int seqLength, char *revComp) {
    c_qryGetReverseComplement(seq, seqLength, revComp);
}

// src/lookupTable5.h:28-29 (the overload of lookupTable5.h:34-35, reading
// keys/values files written by an earlier run, has no definition in the
// reference and none here)
SWIFT_EXPORT void lookupTable5Create(const char *refFile, int seedLen, int maxHitsPerQry,
                                     int tupleIgnoreThreshold, int *totalTuples) {
    c_lookupTable5Create(refFile, seedLen, maxHitsPerQry, tupleIgnoreThreshold, totalTuples);
}

// src/lookupTable5.h:42-43
SWIFT_EXPORT int lookupTable5MapQry(const Query *query, uint qryLen, uint isRevComp,
                                    char *refIdx, int *shift, int *refPos) {
    return c_lookupTable5MapQry(query, qryLen, isRevComp, refIdx, shift, refPos);
}

// src/lookupTable5.h:40
SWIFT_EXPORT void lookupTable5Delete() { c_lookupTable5Delete(); }

// src/mapHits6.h:34
SWIFT_EXPORT int mapHits6GetBestHits(int numBestHits, char *refIdx, int *shift, int *refPos,
                                     int size, char *refIdx_bestMatch, int *shift_bestMatch,
                                     int *refPos_bestMatch) {
    return c_mapHits6GetBestHits(numBestHits, refIdx, shift, refPos, size, refIdx_bestMatch,
                                 shift_bestMatch, refPos_bestMatch);
}

// src/search.h:33-35
SWIFT_EXPORT void searchQueries3(char *qryFileName, int numQrs, int qryLen, char *refFileName,
                                 int seedLen, char *matchFile, uint maxNumHits, int *numMatches,
                                 int numThreads, int tupleIgnoreThres) {
    c_searchQueries3(qryFileName, numQrs, qryLen, refFileName, seedLen, matchFile, maxNumHits,
                     numMatches, numThreads, tupleIgnoreThres);
}

// src/search2.h:26-27
SWIFT_EXPORT void search2Queries(char *qryFile, char *refFile, char *matchFile, int seedLen,
                                 int maxNumHits, int *numMatches, int tupleIgnoreThres) {
    c_search2Queries(qryFile, refFile, matchFile, seedLen, maxNumHits, numMatches,
                     tupleIgnoreThres);
}

// ---- output (src/output.h:29-32,49-56) ------------------------------------
SWIFT_EXPORT void printOutput(int numQueries, char *queryName, char *refName, float *score,
                              int *refDistance, int *alignStart, int *alignEnd, int *alignLength,
                              int combinedLength, char *queryConsensus, char *refConsensus,
                              FILE *outputFilePtr) {
    c_printOutput(numQueries, queryName, refName, score, refDistance, alignStart, alignEnd,
                  alignLength, combinedLength, queryConsensus, refConsensus, outputFilePtr);
}

SWIFT_EXPORT void printOutputSam(int numQueries, char *queryName, char *refName, float *score,
                                 int *alignStart, int *alignEnd, int *alignLength,
                                 int combinedLength, char *queryConsensus, char *refConsensus,
                                 FILE *outputFilePtr) {
    c_printOutputSam(numQueries, queryName, refName, score, alignStart, alignEnd, alignLength,
                     combinedLength, queryConsensus, refConsensus, outputFilePtr);
}

SWIFT_EXPORT void printCigar(char *queryConsensus, char *refConsensus, int alignLength,
                             FILE *outputFilePtr) {
    c_printCigar(queryConsensus, refConsensus, alignLength, outputFilePtr);
}
