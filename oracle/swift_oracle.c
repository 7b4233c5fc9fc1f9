/*
 * swift_oracle.c -- CPU restatement of the reference's CPU alignment path.
 *
 * TEST INFRASTRUCTURE ONLY (see swift_oracle.h).  Nothing under swift_amd/
 * links or calls this file; it is the checker the HIP product is compared
 * against, and the "port" CPU baseline.
 *
 * Each function names the reference file:line it restates.  The code is
 * written from the reference's behaviour, not copied: the SW restatement
 * keeps one direction code per cell instead of the reference's two
 * row/column backtracker matrices, the clustering sorts an index array with
 * qsort instead of a hand-written quicksort, and the index is built from
 * in-memory sequences instead of a temporary text file.
 */
#include "swift_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* glibc rand() replay.  glibc's default generator (TYPE_3) is the additive
 * feedback recurrence r[i] = r[i-3] + r[i-31] (mod 2^32) seeded from a
 * 16807 Lehmer sequence; srand discards the first 310 outputs and rand()
 * returns r[i] >> 1.  Used only to replay the reference's tie draws
 * (src/array.cu:70, src/align.cu:891). */
void orc_rng_seed(orc_rng *g, unsigned int seed)
{
    uint32_t s[344];
    int32_t word;
    int i;
    if (seed == 0)
        seed = 1;
    s[0] = seed;
    word = (int32_t) seed;
    for (i = 1; i < 31; ++i) {
        /* Schrage's method for 16807 * word mod (2^31 - 1) */
        long hi = word / 127773;
        long lo = word % 127773;
        long w = 16807 * lo - 2836 * hi;
        if (w < 0)
            w += 2147483647;
        word = (int32_t) w;
        s[i] = (uint32_t) word;
    }
    for (i = 31; i < 34; ++i)
        s[i] = s[i - 31];
    for (i = 34; i < 344; ++i)
        s[i] = s[i - 31] + s[i - 3];
    for (i = 0; i < 34; ++i)
        g->r[i] = s[310 + i];
    g->pos = 0;
}

int orc_rng_next(orc_rng *g)
{
    uint32_t v = g->r[(g->pos + 3) % 34] + g->r[(g->pos + 31) % 34];
    g->r[g->pos] = v;
    g->pos = (g->pos + 1) % 34;
    return (int) (v >> 1);
}

/* ------------------------------------------------------------------------ */
static int base_code(unsigned char c)
{
    /* preprocessCreate (src/preprocess.cu:30-48): A/a 0, C/c 1, G/g 2,
     * T/t 3, N/n and every other byte 0. */
    switch (c) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}

int orc_hash(const char *s, int k)
{
    /* getHash (src/preprocess.cu:70-91): the first base is the least
     * significant base-4 digit. */
    int h = 0, i;
    for (i = k - 1; i >= 0; --i)
        h = h * 4 + base_code((unsigned char) s[i]);
    return h;
}

void orc_revcomp(const char *seq, int len, char *out)
{
    /* qryGetReverseComplement (src/query.cu:265-282): A->T, C->G, G->C and
     * every other byte (T, N, lower case) -> A. */
    int i;
    for (i = 0; i < len; ++i) {
        char c = seq[len - 1 - i];
        out[i] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
    }
    out[len] = '\0';
}

/* ------------------------------------------------------------------------ */
int orc_index_build(orc_index *ix, int nseq, const char *const *seqs,
                    const int64_t *lens, int k, int thr)
{
    /* lookupTable5Create (src/lookupTable5.cu:101-252).  Tuples are the
     * non-overlapping k-mers at positions 0, k, 2k, ... of each sequence
     * (the line-carry logic at :151-181 makes line breaks transparent);
     * a hash seen more than thr times is dropped entirely (:159-173);
     * buckets are laid out in hash order starting at 1 (:199-208) and
     * filled in reverse file order (:219-229). */
    int64_t nh = (int64_t) 1 << (2 * k), h, pos, slot, s;
    int q;
    int32_t *rem;
    memset(ix, 0, sizeof(*ix));
    ix->k = k;
    ix->thr = thr;
    ix->nhash = nh;
    ix->start = (int32_t *) calloc((size_t) nh, sizeof(int32_t));
    ix->count = (int32_t *) calloc((size_t) nh, sizeof(int32_t));
    ix->actual = (int32_t *) calloc((size_t) nh, sizeof(int32_t));
    rem = (int32_t *) calloc((size_t) nh, sizeof(int32_t));
    if (!ix->start || !ix->count || !ix->actual || !rem)
        return -1;
    for (q = 0; q < nseq; ++q)
        for (pos = 0; pos + k <= lens[q]; pos += k)
            ++ix->count[orc_hash(seqs[q] + pos, k)];
    s = 1;
    for (h = 0; h < nh; ++h) {
        if (ix->count[h] > 0 && ix->count[h] <= thr) {
            ix->actual[h] = ix->count[h];
            ix->start[h] = (int32_t) s;
            s += ix->count[h];
            rem[h] = ix->count[h];
        }
    }
    ix->total = s - 1;
    ix->refIdx = (int8_t *) malloc((size_t) (ix->total + 1));
    ix->refPos = (int32_t *) malloc((size_t) (ix->total + 1) * sizeof(int32_t));
    if (!ix->refIdx || !ix->refPos)
        return -1;
    ix->refIdx[0] = -1;
    ix->refPos[0] = -1;
    for (q = 0; q < nseq; ++q) {
        for (pos = 0; pos + k <= lens[q]; pos += k) {
            h = orc_hash(seqs[q] + pos, k);
            if (ix->actual[h] == 0)
                continue;
            slot = ix->start[h] + rem[h] - 1;
            --rem[h];
            ix->refIdx[slot] = (int8_t) q;
            ix->refPos[slot] = (int32_t) pos;
        }
    }
    free(rem);
    return 0;
}

void orc_index_free(orc_index *ix)
{
    free(ix->start);
    free(ix->count);
    free(ix->actual);
    free(ix->refIdx);
    free(ix->refPos);
    memset(ix, 0, sizeof(*ix));
}

/* ------------------------------------------------------------------------ */
typedef struct {
    int8_t idx;
    int32_t shift, pos;
} hit_t;

static int hit_cmp(const void *a, const void *b)
{
    /* ordering of quickSort (src/mapHits6.cu:299-318): refIdx, then shift,
     * then refPos, all ascending; refIdx compares as a signed char. */
    const hit_t *x = (const hit_t *) a, *y = (const hit_t *) b;
    if (x->idx != y->idx)
        return x->idx < y->idx ? -1 : 1;
    if (x->shift != y->shift)
        return x->shift < y->shift ? -1 : 1;
    if (x->pos != y->pos)
        return x->pos < y->pos ? -1 : 1;
    return 0;
}

int orc_best_hits(int maxHits, int8_t *refIdx, int32_t *shift,
                  int32_t *refPos, int n, int tie_mode, orc_rng *rng,
                  int8_t *oIdx, int32_t *oShift, int32_t *oPos,
                  int *nBiggest)
{
    /* mapHits6GetBestHits (src/mapHits6.cu:40-70) with createClusters
     * (:255-272), findBiggestClusterSize (:202-223) and findBiggestClusters
     * (:132-167). */
    hit_t *h;
    int *big, *pick;
    int i, nbig = 0, biggest = 0, run, nbest;
    if (nBiggest)
        *nBiggest = 0;
    if (n <= 0)
        return 0;
    h = (hit_t *) malloc((size_t) n * sizeof(hit_t));
    big = (int *) malloc((size_t) n * sizeof(int));
    for (i = 0; i < n; ++i) {
        h[i].idx = refIdx[i];
        h[i].shift = shift[i];
        h[i].pos = refPos[i];
    }
    qsort(h, (size_t) n, sizeof(hit_t), hit_cmp);
    for (i = 0; i < n; ++i) { /* the reference sorts the caller's arrays */
        refIdx[i] = h[i].idx;
        shift[i] = h[i].shift;
        refPos[i] = h[i].pos;
    }
    /* cluster = maximal run of equal (refIdx, shift); first pass: size */
    for (i = 0; i < n; i += run) {
        run = 1;
        while (i + run < n && h[i + run].idx == h[i].idx
               && h[i + run].shift == h[i].shift)
            ++run;
        if (run > biggest)
            biggest = run;
    }
    for (i = 0; i < n; i += run) {
        run = 1;
        while (i + run < n && h[i + run].idx == h[i].idx
               && h[i + run].shift == h[i].shift)
            ++run;
        if (run == biggest)
            big[nbig++] = i;
    }
    if (nBiggest)
        *nBiggest = nbig;
    nbest = maxHits < nbig ? maxHits : nbig;
    if (nbest < 0)
        nbest = 0;
    pick = (int *) malloc((size_t) (nbest > 0 ? nbest : 1) * sizeof(int));
    if (tie_mode == ORC_TIE_GLIBC) {
        /* arrGetRandomNums (src/array.cu:61-77): distinct draws of
         * rand() % range, then qsort ascending (mapHits6.cu:155). */
        int got = 0, j, r, dup, t;
        while (got < nbest) {
            r = orc_rng_next(rng) % nbig;
            dup = 0;
            for (j = 0; j < got; ++j)
                dup |= pick[j] == r;
            if (!dup)
                pick[got++] = r;
        }
        for (i = 1; i < nbest; ++i)
            for (j = i; j > 0 && pick[j - 1] > pick[j]; --j) {
                t = pick[j];
                pick[j] = pick[j - 1];
                pick[j - 1] = t;
            }
    } else {
        for (i = 0; i < nbest; ++i)
            pick[i] = i;
    }
    for (i = 0; i < nbest; ++i) {
        const hit_t *f = &h[big[pick[i]]]; /* first (smallest refPos) hit */
        oIdx[i] = f->idx;
        oShift[i] = f->shift;
        oPos[i] = f->pos;
    }
    free(pick);
    free(big);
    free(h);
    return nbest;
}

int orc_map_query(const orc_index *ix, const char *seq, int len, int maxHits,
                  int tie_mode, orc_rng *rng, int8_t *oIdx, int32_t *oShift,
                  int32_t *oPos, int *nBiggest, int *nHits)
{
    /* lookupTable5MapQry (src/lookupTable5.cu:557-603): every overlapping
     * query offset i in [0, len-k]; each occurrence of its k-mer in bucket
     * order becomes the hit (refIdx, refPos - i, refPos). */
    int k = ix->k, nq = len - k + 1, i, n = 0, cap = 64, r;
    int64_t j;
    int8_t *ri = (int8_t *) malloc((size_t) cap);
    int32_t *sh = (int32_t *) malloc((size_t) cap * sizeof(int32_t));
    int32_t *rp = (int32_t *) malloc((size_t) cap * sizeof(int32_t));
    for (i = 0; i < nq; ++i) {
        int hsh = orc_hash(seq + i, k);
        int32_t a = ix->actual[hsh];
        if (a == 0)
            continue;
        if (n + a > cap) {
            while (n + a > cap)
                cap *= 2;
            ri = (int8_t *) realloc(ri, (size_t) cap);
            sh = (int32_t *) realloc(sh, (size_t) cap * sizeof(int32_t));
            rp = (int32_t *) realloc(rp, (size_t) cap * sizeof(int32_t));
        }
        for (j = ix->start[hsh]; j < ix->start[hsh] + a; ++j) {
            ri[n] = ix->refIdx[j];
            sh[n] = ix->refPos[j] - i;
            rp[n] = ix->refPos[j];
            ++n;
        }
    }
    if (nHits)
        *nHits = n;
    r = orc_best_hits(maxHits, ri, sh, rp, n, tie_mode, rng, oIdx, oShift,
                      oPos, nBiggest);
    free(ri);
    free(sh);
    free(rp);
    return r;
}

/* ------------------------------------------------------------------------ */
int orc_region(const char *seq, int64_t seqLen, int64_t refPos, int before,
               int winLen, char *out, int64_t *refDistance)
{
    /* alignQueries_cpu_mt: refDistance = max(0, refPos - refLen/3)
     * (src/align.cu:2388,2435); getRefSeq2 (src/input.cu:348-403) reads up
     * to winLen bases, stopping at the end of the sequence, then strips
     * leading N/n.  Trailing N/n are stripped only when no leading N was
     * stripped: with a leading strip the reference's trailing test indexes
     * past the string (input.cu:399 uses strlen of the unstripped copy), which
     * for one stripped N reads the terminator; longer strips read stale
     * bytes and are defined here as "no trailing strip".  refDistance is not
     * adjusted for the strip, as in the reference. */
    int64_t rd = refPos - before, n, m = 0;
    if (rd < 0)
        rd = 0;
    *refDistance = rd;
    n = seqLen - rd;
    if (n > winLen)
        n = winLen;
    if (n < 0)
        n = 0;
    while (m < n && (seq[rd + m] == 'N' || seq[rd + m] == 'n'))
        ++m;
    if (m == 0)
        while (n > 0 && (seq[rd + n - 1] == 'N' || seq[rd + n - 1] == 'n'))
            --n;
    memcpy(out, seq + rd + m, (size_t) (n - m));
    out[n - m] = '\0';
    return (int) (n - m);
}

/* ------------------------------------------------------------------------ */
enum { D_START = 0, D_LEFT = 1, D_UP = 2, D_DIAG = 3 };

static int upper_byte(int ch)
{
    /* toupper_cpu (src/smithWaterman.cu:1194-1199) */
    if ((unsigned int) (ch - 'a') < 26u)
        ch += 'A' - 'a';
    return ch;
}

static void step_back(const uint8_t *dir, int W, int i, int j, int *ni,
                      int *nj)
{
    switch (dir[(int64_t) i * W + j]) {
    case D_DIAG: *ni = i - 1; *nj = j - 1; break;
    case D_UP: *ni = i - 1; *nj = j; break;
    case D_LEFT: *ni = i; *nj = j - 1; break;
    default: *ni = i; *nj = j; break;
    }
}

void orc_sw(const char *ref, int refLen, const char *qry, int qryLen,
            int match, int mismatch, int gapOpen, int gapExt,
            orc_sw_result *res, char *refCons, char *qryCons,
            uint8_t *dirOut)
{
    /* computeMaxScore_cpu (src/smithWaterman.cu:114-183) + findArrMax
     * (:194-210): H = first maximum of [diag, up, left, 0] (ties prefer
     * diag, then up, then left, then 0 -- so a 0-valued diag/up/left still
     * wins over the 0 candidate); a vertical (horizontal) step costs
     * gapExt when the cell above (left) itself came from above (left), else
     * gapOpen; match when bytes are equal or either is 'N', the reference
     * byte upper-cased, the query byte not; columns stop at the first NUL
     * of ref.  The best cell is the first strict maximum in row-major
     * order.  backtrack_cpu (:263-335) then walks the pointers. */
    int W = refLen + 1, ncols = refLen, i, j;
    int64_t cells = (int64_t) (qryLen + 1) * W;
    int *H = (int *) calloc((size_t) cells, sizeof(int));
    uint8_t *D = dirOut ? dirOut : (uint8_t *) malloc((size_t) cells);
    int best = 0, br = 0, bc = 0;
    int ci, cj, ni, nj, tick = 0, nrb = -1, start = 0;
    memset(D, D_START, (size_t) cells);
    for (j = 0; j < refLen; ++j)
        if (ref[j] == '\0') {
            ncols = j;
            break;
        }
    for (i = 1; i <= qryLen; ++i) {
        int *Hr = H + (int64_t) i * W, *Hu = Hr - W;
        uint8_t *Dr = D + (int64_t) i * W, *Du = Dr - W;
        char q = qry[i - 1];
        for (j = 1; j <= ncols; ++j) {
            char R = (char) upper_byte(ref[j - 1]);
            int s = (q == R || q == 'N' || R == 'N') ? match : mismatch;
            int cand[4], c, m = -1, arg = 0;
            cand[0] = Hu[j - 1] + s;
            cand[1] = Hu[j] + (Du[j] == D_UP ? gapExt : gapOpen);
            cand[2] = Hr[j - 1] + (Dr[j - 1] == D_LEFT ? gapExt : gapOpen);
            cand[3] = 0;
            for (c = 0; c < 4; ++c)
                if (m < cand[c]) {
                    m = cand[c];
                    arg = c;
                }
            Hr[j] = m;
            Dr[j] = (uint8_t) (arg == 0 ? D_DIAG : arg == 1 ? D_UP
                               : arg == 2 ? D_LEFT : D_START);
            if (best < m) {
                best = m;
                br = i;
                bc = j;
            }
        }
    }
    /* backtrack_cpu: emit the current cell while its pointer leads to a
     * different cell; a pointer into row 0 or column 0 emits the current
     * cell and stops; a start pointer stops without emitting. */
    ci = br;
    cj = bc;
    step_back(D, W, ci, cj, &ni, &nj);
    while ((ci != ni || cj != nj) && nj != 0 && ni != 0) {
        if (nj == cj) refCons[tick] = '-';
        else { refCons[tick] = ref[cj - 1]; ++nrb; }
        qryCons[tick] = (ni == ci) ? '-' : qry[ci - 1];
        start = cj;
        ++tick;
        ci = ni;
        cj = nj;
        step_back(D, W, ci, cj, &ni, &nj);
    }
    if ((ci != ni || cj != nj) && (nj == 0 || ni == 0)) {
        if (nj == cj) refCons[tick] = '-';
        else { refCons[tick] = ref[cj - 1]; ++nrb; }
        qryCons[tick] = (ni == ci) ? '-' : qry[ci - 1];
        start = cj;
        ++tick;
    }
    refCons[tick] = '\0';
    qryCons[tick] = '\0';
    res->score = best;
    res->maxRow = br;
    res->maxCol = bc;
    res->start = start;
    res->end = start + nrb;
    res->len = tick;
    free(H);
    if (!dirOut)
        free(D);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    int tid, nthreads, npairs;
    const char *qry, *ref;
    const int64_t *qoff, *roff, *consOff;
    const int32_t *qlen, *rlen;
    int match, mismatch, gapOpen, gapExt;
    int32_t *score, *start, *end, *alen;
    char *qcons, *rcons;
} batch_arg;

static void *batch_worker(void *p)
{
    batch_arg *a = (batch_arg *) p;
    int i;
    for (i = a->tid; i < a->npairs; i += a->nthreads) {
        orc_sw_result r;
        orc_sw(a->ref + a->roff[i], a->rlen[i], a->qry + a->qoff[i],
               a->qlen[i], a->match, a->mismatch, a->gapOpen, a->gapExt, &r,
               a->rcons + a->consOff[i], a->qcons + a->consOff[i], NULL);
        a->score[i] = r.score;
        a->start[i] = r.start;
        a->end[i] = r.end;
        a->alen[i] = r.len;
    }
    return NULL;
}

void orc_sw_batch(int npairs, const char *qry, const int64_t *qoff,
                  const int32_t *qlen, const char *ref, const int64_t *roff,
                  const int32_t *rlen, int match, int mismatch, int gapOpen,
                  int gapExt, int32_t *score, int32_t *start, int32_t *end,
                  int32_t *alen, char *qcons, char *rcons,
                  const int64_t *consOff, int nthreads)
{
    pthread_t th[256];
    batch_arg args[256];
    int t;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    for (t = 0; t < nthreads; ++t) {
        batch_arg *a = &args[t];
        a->tid = t;
        a->nthreads = nthreads;
        a->npairs = npairs;
        a->qry = qry;
        a->ref = ref;
        a->qoff = qoff;
        a->roff = roff;
        a->consOff = consOff;
        a->qlen = qlen;
        a->rlen = rlen;
        a->match = match;
        a->mismatch = mismatch;
        a->gapOpen = gapOpen;
        a->gapExt = gapExt;
        a->score = score;
        a->start = start;
        a->end = end;
        a->alen = alen;
        a->qcons = qcons;
        a->rcons = rcons;
        if (nthreads > 1)
            pthread_create(&th[t], NULL, batch_worker, a);
    }
    if (nthreads == 1)
        batch_worker(&args[0]);
    else
        for (t = 0; t < nthreads; ++t)
            pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
void orc_choose_best(int n, const int32_t *group, const int32_t *score,
                     int tie_mode, orc_rng *rng, int32_t *keep,
                     int32_t *nTied)
{
    /* chooseBestAlignment (src/align.cu:844-905): max starts at 0 and rises
     * on strict <; the records equal to it form the tie set; one of them is
     * kept (rand() % j in the reference; the first in FIRST mode). */
    int s = 0, e, i, j, pick, mx;
    while (s < n) {
        e = s + 1;
        while (e < n && group[e] == group[s])
            ++e;
        mx = 0;
        for (i = s; i < e; ++i)
            if (mx < score[i])
                mx = score[i];
        j = 0;
        for (i = s; i < e; ++i)
            j += score[i] == mx;
        pick = 0;
        if (tie_mode == ORC_TIE_GLIBC && j > 0)
            pick = orc_rng_next(rng) % j;
        for (i = s; i < e; ++i) {
            keep[i] = 0;
            if (score[i] == mx) {
                if (pick == 0)
                    keep[i] = 1;
                --pick;
            }
            if (nTied)
                nTied[i] = j;
        }
        s = e;
    }
}

/* ------------------------------------------------------------------------ */
int orc_pipeline(const orc_index *ix, int nseq, const char *const *seqs,
                 const int64_t *seqLens, int nreads, const char *reads,
                 const int64_t *readOff, const int32_t *readLen, int maxHits,
                 int match, int mismatch, int gapOpen, int gapExt,
                 int tie_mode, unsigned int seed, int maxRecords,
                 orc_record *rec, char *rcons, char *qcons, int consStride)
{
    /* searchQueries3 (src/search.cu:275-382) + getHits (:1479-1488) for
     * phase 1, alignQueries_cpu_mt (src/align.cu:2363-2511) for phase 2:
     * window = 300 bases from max(0, refPos - 100) of sequence refIdx; the
     * strand-1 query is the reverse complement.  keep marks the record
     * chooseBestAlignment (src/align.cu:844-905) retains per read. */
    orc_rng rng;
    int r, s, h, n = 0, maxLen = 0, nh;
    int8_t oi[64];
    int32_t osh[64], op[64];
    int32_t *grp, *sc, *kp, *nt;
    char *buf, *win;
    (void) nseq;
    orc_rng_seed(&rng, seed);
    if (maxHits > 64)
        maxHits = 64;
    for (r = 0; r < nreads; ++r)
        if (readLen[r] > maxLen)
            maxLen = readLen[r];
    buf = (char *) malloc((size_t) maxLen + 1);
    win = (char *) malloc(301);
    for (r = 0; r < nreads; ++r) {
        const char *rd = reads + readOff[r];
        int L = readLen[r];
        for (s = 0; s < 2; ++s) {
            int nb = 0;
            const char *q = rd;
            if (s == 1) {
                orc_revcomp(rd, L, buf);
                q = buf;
            } else {
                memcpy(buf, rd, (size_t) L);
                buf[L] = '\0';
                q = buf;
            }
            nh = orc_map_query(ix, q, L, maxHits, tie_mode, &rng, oi, osh,
                               op, &nb, NULL);
            for (h = 0; h < nh; ++h) {
                orc_sw_result res;
                int64_t rdist;
                orc_record *o;
                if (n >= maxRecords) {
                    free(buf);
                    free(win);
                    return -1;
                }
                o = &rec[n];
                orc_region(seqs[oi[h]], seqLens[oi[h]], op[h], 100, 300, win,
                           &rdist);
                orc_sw(win, 300, q, L, match, mismatch, gapOpen, gapExt, &res,
                       rcons + (int64_t) n * consStride,
                       qcons + (int64_t) n * consStride, NULL);
                o->read = r;
                o->strand = s;
                o->refIdx = oi[h];
                o->shift = osh[h];
                o->refPos = op[h];
                o->nBiggest = nb;
                o->refDistance = rdist;
                o->score = res.score;
                o->start = res.start;
                o->end = res.end;
                o->len = res.len;
                ++n;
            }
        }
    }
    free(buf);
    free(win);
    grp = (int32_t *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int32_t));
    sc = (int32_t *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int32_t));
    kp = (int32_t *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int32_t));
    nt = (int32_t *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int32_t));
    for (h = 0; h < n; ++h) {
        grp[h] = rec[h].read;
        sc[h] = rec[h].score;
    }
    orc_choose_best(n, grp, sc, tie_mode, &rng, kp, nt);
    for (h = 0; h < n; ++h) {
        rec[h].keep = kp[h];
        rec[h].nTied = nt[h];
    }
    free(grp);
    free(sc);
    free(kp);
    free(nt);
    return n;
}
