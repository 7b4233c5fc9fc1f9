"""The drop-ins link from C++ callers, as the reference's callers are C++.

The reference compiles every source as C++ and declares its functions
without extern "C" (e.g. src/align.h:75-80, src/lookupTable5.h:28-43,
src/search.h:33-35, src/mapHits6.h:34, src/output.h:29-56), so a reference
caller binds mangled names.  These CPU tests compile C++ callers against
libswift_amd.so:

- one with the reference's prototypes written out below (each cites its
  header), which also runs the drop-ins that need no GPU;
- one that includes the reference's own headers, when the reference
  sources are present (compiled as oracle/Makefile compiles them);
- and a check that every drop-in symbol the reference's own object files
  define (oracle/_ref, built from the reference sources) is exported.
No GPU call is made.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "swift_amd")
REF_SRC = "/root/reference/src"
HIPCC = "/opt/rocm/bin/hipcc"

# Prototypes as the reference headers declare them (C++ linkage).
PROTOTYPES = r"""
#include <cstdio>
#include <cstring>
#include <cstdlib>
typedef unsigned int uint;
typedef unsigned long ulong;
typedef struct queryStruct {            /* src/query.h:29-37 */
    char *name; int length; char *seq; char *revCompSeq; ulong nameOffset; ulong seqOffset;
} Query;
/* src/align.h:33-36 */
void alignQueries3(char *queryFileName, int maxQryLen, char *refFileName, char *matchFile,
    int outFormat, float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
    const char *outputFile, int numMatches);
/* src/align.h:54-56 */
void alignQueries_cpu_mt(char *qryFile, int numQrs, int qryLen, char *refFile, char *matchFile,
    int outFormat, float match, float mismatch, float gapOpenPenalty, float gapExtPenalty,
    const char *outputFile);
/* src/align.h:75-80 */
void alignSequences(char *querySeq_d, char *refSeq_d, int querySeqLength, int refSeqLength,
    float match, float mismatch, float gapOpenPenalty, float gapExtPenalty, float *score_d,
    int *alignStart_d, int *alignEnd_d, int *alignLength_d, char *queryConsensus_d,
    char *refConsensus_d, int queriesPerIteration, int refsPerIteration, int maxThreadsPerBlock);
/* src/smithWaterman.h:62-66 */
void smithWaterman_cpu(float *HMatrix, int *rowBacktracker, int *colBacktracker, char *refSeq,
    int refLen, char *qrySeq, int qryLen, float match, float mismatch, float gapOpenPenalty,
    float gapExtPenalty, float *alignScore, int *startPos, int *endPos, int *alignLen,
    char *refConsensus, char *qryConsensus);
/* src/preprocess.h:21,23,29 */
void preprocessCreate();
int getHash(char *s, int length);
void preprocessDelete();
/* src/query.h (qryGetReverseComplement) */
void qryGetReverseComplement(char *seq, int seqLength, char *revComp);
/* src/lookupTable5.h:28-29,40,42-43 */
void lookupTable5Create(const char *refFile, int seedLen, int maxHitsPerQry,
    int tupleIgnoreThreshold, int *totalTuples);
void lookupTable5Delete();
int lookupTable5MapQry(const Query *query, uint qryLen, uint isRevComp, char *refIdx,
    int *shift, int *refPos);
/* src/mapHits6.h:34 */
int mapHits6GetBestHits(int numBestHits, char *refIdx, int *shift, int *refPos, int size,
    char *refIdx_bestMatch, int *shift_bestMatch, int *refPos_bestMatch);
/* src/search.h:33-35 */
void searchQueries3(char *qryFileName, int numQrs, int qryLen, char *refFileName, int seedLen,
    char *matchFile, uint maxNumHits, int *numMatches, int numThreads, int tupleIgnoreThres);
/* src/search2.h:26-27 */
void search2Queries(char *qryFile, char *refFile, char *matchFile, int seedLen, int maxNumHits,
    int *numMatches, int tupleIgnoreThres);
/* src/output.h:29-32,49-56 */
void printOutput(int numQueries, char *queryName, char *refName, float *score, int *refDistance,
    int *alignStart, int *alignEnd, int *alignLength, int combinedLength, char *queryConsensus,
    char *refConsensus, FILE *outputFilePtr);
void printOutputSam(int numQueries, char *queryName, char *refName, float *score,
    int *alignStart, int *alignEnd, int *alignLength, int combinedLength, char *queryConsensus,
    char *refConsensus, FILE *outputFilePtr);
void printCigar(char *queryConsensus, char *refConsensus, int alignLength, FILE *outputFilePtr);
"""

# Every drop-in's address is taken, so the link must resolve each one.
ADDRESSES = r"""
static void *const kAll[] = {
    (void *)&alignQueries3, (void *)&alignQueries_cpu_mt, (void *)&alignSequences,
    (void *)&smithWaterman_cpu, (void *)&preprocessCreate, (void *)&getHash,
    (void *)&preprocessDelete, (void *)&qryGetReverseComplement,
    (void *)(void (*)(const char *, int, int, int, int *))&lookupTable5Create,
    (void *)&lookupTable5Delete, (void *)&lookupTable5MapQry, (void *)&mapHits6GetBestHits,
    (void *)&searchQueries3, (void *)&search2Queries, (void *)&printOutput,
    (void *)&printOutputSam, (void *)&printCigar};
"""

# Host-only drop-ins run through their C++ names.  Expected values: the hash
# closed form sum_i 4^i code(s[i]) (A0 C1 G2 T3, other bytes 0,
# src/preprocess.cu:70-91), the reverse complement rule of
# src/query.cu:265-282 (anything but ACG -> A after reversal), and a printed
# record in printOutput's format (src/output.cu:154-201).
MAIN = r"""
int main(int argc, char **argv) {
    for (unsigned i = 0; i < sizeof(kAll) / sizeof(kAll[0]); ++i)
        if (!kAll[i]) return 10 + i;
    preprocessCreate();
    char s[] = "ACGTN";
    if (getHash(s, 5) != 0 + 1 * 4 + 2 * 16 + 3 * 64 + 0) return 2;
    char r[] = "ACGTNa", rc[8];
    qryGetReverseComplement(r, 6, rc);
    if (strcmp(rc, "AAACGT") != 0) return 3;
    preprocessDelete();
    FILE *f = fopen(argv[1], "w");
    char qn[50] = "q1", rn[50] = "chr1", qc[8] = "AC-G", rcn[8] = "ACTG";
    float score = 5.0f;
    int dist = 10, st = 3, en = 6, len = 4;
    printOutput(1, qn, rn, &score, &dist, &st, &en, &len, 8, qc, rcn, f);
    printCigar(qc, rcn, 4, f);
    fclose(f);
    return 0;
}
"""


def _link(tmp_path, name, src, compiler, extra=()):
    cpp = tmp_path / f"{name}.cpp"
    cpp.write_text(src)
    exe = tmp_path / name
    cmd = [compiler, *extra, str(cpp), "-o", str(exe), f"-L{LIBDIR}", "-lswift_amd",
           f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def built():
    from swift_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from swift_amd import build
        build.build()
    return _lib.LIB_PATH


def test_cxx_caller_with_reference_prototypes_links_and_runs(built, tmp_path):
    exe = _link(tmp_path, "caller", PROTOTYPES + ADDRESSES + MAIN, "g++")
    out = tmp_path / "out.txt"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    text = out.read_text()
    # the same record through the C entry point (pinned byte for byte against
    # the reference's writer in test_output_host.py)
    import ctypes as C
    from swift_amd import _lib
    L = _lib.lib()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    want_path = tmp_path / "want.txt"
    f = libc.fopen(str(want_path).encode(), b"w")
    qn, rn = C.create_string_buffer(b"q1", 50), C.create_string_buffer(b"chr1", 50)
    qc, rc = C.create_string_buffer(b"AC-G", 8), C.create_string_buffer(b"ACTG", 8)
    score, dist = C.c_float(5.0), C.c_int(10)
    st, en, ln = C.c_int(3), C.c_int(6), C.c_int(4)
    L.printOutput(1, qn, rn, C.byref(score), C.byref(dist), C.byref(st), C.byref(en),
                  C.byref(ln), 8, qc, rc, C.c_void_p(f))
    L.printCigar(qc, rc, 4, C.c_void_p(f))
    libc.fclose(f)
    assert text == want_path.read_text() and "Score: 5.0" in text


def test_exports_both_linkages(built):
    """Each drop-in is exported under its C name and its C++ mangled name."""
    r = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True)
    syms = {ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()}
    for name in ("alignSequences", "smithWaterman_cpu", "alignQueries3", "alignQueries_cpu_mt",
                 "getHash", "qryGetReverseComplement", "lookupTable5Create",
                 "lookupTable5MapQry", "lookupTable5Delete", "mapHits6GetBestHits",
                 "searchQueries3", "search2Queries", "printOutput", "printOutputSam",
                 "printCigar", "preprocessCreate", "preprocessDelete"):
        assert name in syms, name
        assert any(s.startswith(f"_Z{len(name)}{name}") for s in syms), name


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources absent")
def test_caller_with_reference_headers_links(built, tmp_path):
    """A caller compiled with the reference's own headers (the recipe of
    oracle/Makefile: hipcc for the __global__ declarations, the standard C
    headers nvcc would have included, common.h and query.h) links.  align.h
    needs CUDA's cudaDeviceProp, which this image lacks, so its three
    functions are declared as above."""
    headers = ["query.h", "preprocess.h", "mapHits6.h", "output.h", "smithWaterman.h",
               "lookupTable5.h", "search.h", "search2.h"]
    align_decls = PROTOTYPES[PROTOTYPES.index("/* src/align.h:33-36 */"):
                             PROTOTYPES.index("/* src/smithWaterman.h:62-66 */")]
    src = "".join(f'#include "{h}"\n' for h in headers) + align_decls + ADDRESSES + \
        "int main() { return kAll[0] ? 0 : 1; }\n"
    inc = [f"-I{REF_SRC}", "-w", "-x", "hip", "--offload-arch=gfx950", "-include",
           "hip/hip_runtime.h"]
    for h in ("stdio.h", "stdlib.h", "string.h", "math.h", "ctype.h",
              f"{REF_SRC}/common.h"):
        inc += ["-include", h]
    exe = _link(tmp_path, "refcaller", src, HIPCC, inc)
    assert exe.exists()


def test_reference_objects_symbols_exported(built):
    """The mangled names the reference's own compiled sources give these
    functions (oracle/_ref, built by `make -C oracle ref`) are exported."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    objs = [os.path.join(ref, f"{n}.o") for n in ("smithWaterman", "mapHits6", "preprocess",
                                                  "query", "output")]
    if not all(os.path.exists(o) for o in objs) or not shutil.which("nm"):
        pytest.skip("reference objects not built")
    wanted = {"smithWaterman_cpu", "mapHits6GetBestHits", "getHash", "preprocessCreate",
              "preprocessDelete", "qryGetReverseComplement", "printOutput", "printOutputSam",
              "printCigar"}
    ref_syms = set()
    for o in objs:
        r = subprocess.run(["nm", "--defined-only", o], capture_output=True, text=True)
        for ln in r.stdout.splitlines():
            parts = ln.split()
            if len(parts) == 3 and parts[1] == "T":
                ref_syms.add(parts[2])
    r = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True)
    ours = {ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()}
    found = {s for s in ref_syms if any(s.startswith(f"_Z{len(n)}{n}") for n in wanted)}
    assert len(found) == len(wanted), sorted(found)
    missing = sorted(found - ours)
    assert not missing, missing
