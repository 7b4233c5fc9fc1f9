// sw_matrix.hip -- the smithWaterman_cpu drop-in: one pair per call with the
// full score and backtracker matrices written out, as the reference's CPU
// Smith-Waterman leaves them in its caller's buffers
// (src/smithWaterman.cu:44-335, src/smithWaterman.h:62-66).
//
// The reference's CPU phase 2 (getAlignments, src/align.cu:2519-2548) calls
// it once per pair from 8 threads, with matrices it owns and zeroes.  Here
// each call runs one workgroup over the anti-diagonals of its pair, keeping
// the last three diagonals of scores and directions in LDS, and writes every
// cell the reference writes: H (float), rowBacktracker and colBacktracker at
// the reference's index row * (refLen + 1) + col.  Thread 0 then walks the
// backtracker matrices as backtrack_cpu does.  Calls from several host
// threads run concurrently, each on a stream and buffers of its own.
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/swift_abi.h"
#include "sw_host.hpp"

namespace swift {

namespace {

constexpr int kMatThreads = 256;
constexpr int kMatMaxQ = 4096;  // rows: 3 diagonals of q + 1 cells (5 B each) stay within 64 KB of LDS

// direction of a cell, in the order findArrMax breaks ties
// (src/smithWaterman.cu:194-210): candidates [diag, up, left, 0]
enum : uint8_t { kDiag = 0, kUp = 1, kLeft = 2, kStart = 3 };

struct MatOut {
    int32_t score, start, nref, alen, best_row, best_col;
};

__device__ __forceinline__ bool base_match(unsigned char q, unsigned char r) {
    // isMatch_cpu(qry, toupper_cpu(ref)) (src/smithWaterman.cu:849-855,1194-1199)
    const unsigned char ru = (r >= 'a' && r <= 'z') ? (unsigned char)(r - 32) : r;
    return q == ru || q == 'N' || ru == 'N';
}

// Row-major order of cells: the reference keeps the first strict maximum of
// its row-by-row scan (src/smithWaterman.cu:173-179).
__device__ __forceinline__ bool better(int h, int i, int j, int bh, int bi, int bj) {
    return h > bh || (h == bh && h > 0 && (i < bi || (i == bi && j < bj)));
}

__global__ void __launch_bounds__(kMatThreads)
sw_matrix_kernel(const char *__restrict__ qry, int q, const char *__restrict__ ref, int re,
                 int pitch, Scoring sc, float *__restrict__ H, int *__restrict__ rowbt,
                 int *__restrict__ colbt, MatOut *out, char *__restrict__ rcons,
                 char *__restrict__ qcons) {
    extern __shared__ int32_t lds_mat[];
    int32_t *hd = lds_mat;                                          // [3][q + 1] scores
    uint8_t *dd = reinterpret_cast<uint8_t *>(lds_mat + 3 * (q + 1)); // [3][q + 1] directions
    __shared__ int red_h[kMatThreads], red_i[kMatThreads], red_j[kMatThreads];
    const int t = threadIdx.x;
    int bh = 0, bi = 0, bj = 0;

    for (int d = 2; d <= q + re; ++d) {
        const int cur = d % 3, prev = (d + 2) % 3, prev2 = (d + 1) % 3;
        const int ilo = max(1, d - re), ihi = min(q, d - 1);
        for (int i = ilo + t; i <= ihi; i += kMatThreads) {
            const int j = d - i;
            // neighbours: diag (i-1, j-1) on d-2, up (i-1, j) and left (i, j-1)
            // on d-1; row 0 and column 0 hold the caller's zeroes
            const int hdiag = (i > 1 && j > 1) ? hd[prev2 * (q + 1) + i - 1] : 0;
            const int hup = i > 1 ? hd[prev * (q + 1) + i - 1] : 0;
            const uint8_t dup = i > 1 ? dd[prev * (q + 1) + i - 1] : kStart;
            const int hleft = j > 1 ? hd[prev * (q + 1) + i] : 0;
            const uint8_t dleft = j > 1 ? dd[prev * (q + 1) + i] : kStart;
            const bool m = base_match((unsigned char)qry[i - 1], (unsigned char)ref[j - 1]);
            const int c0 = hdiag + (m ? sc.match : sc.mismatch);
            // a gap extends only when the neighbour points the same way
            const int c1 = hup + (dup == kUp ? sc.gap_ext : sc.gap_open);
            const int c2 = hleft + (dleft == kLeft ? sc.gap_ext : sc.gap_open);
            int h = c0;
            uint8_t dir = kDiag;
            if (c1 > h) { h = c1; dir = kUp; }
            if (c2 > h) { h = c2; dir = kLeft; }
            if (0 > h) { h = 0; dir = kStart; }
            hd[cur * (q + 1) + i] = h;
            dd[cur * (q + 1) + i] = dir;
            const int64_t cell = (int64_t)i * pitch + j;
            H[cell] = (float)h;
            rowbt[cell] = dir == kLeft || dir == kStart ? i : i - 1;
            colbt[cell] = dir == kUp || dir == kStart ? j : j - 1;
            if (better(h, i, j, bh, bi, bj)) {
                bh = h;
                bi = i;
                bj = j;
            }
        }
        __syncthreads();
    }
    red_h[t] = bh;
    red_i[t] = bi;
    red_j[t] = bj;
    __syncthreads();
    for (int s = kMatThreads / 2; s >= 1; s >>= 1) {
        if (t < s && better(red_h[t + s], red_i[t + s], red_j[t + s], red_h[t], red_i[t], red_j[t])) {
            red_h[t] = red_h[t + s];
            red_i[t] = red_i[t + s];
            red_j[t] = red_j[t + s];
        }
        __syncthreads();
    }
    // the backtracker matrices written above by the whole workgroup are read
    // back by thread 0: make them visible first
    __threadfence();
    __syncthreads();
    if (t != 0) return;
    // backtrack_cpu (src/smithWaterman.cu:263-335): follow the pointers from
    // the best cell while they move and the next cell is off row 0 and
    // column 0; that last step is still emitted, the cell it reaches is not
    int cr = red_i[0], cc = red_j[0];
    auto bt = [&](int r, int c, int *nr, int *nc) {
        if (r == 0 || c == 0) {  // the caller's zeroed border
            *nr = 0;
            *nc = 0;
        } else {
            *nr = rowbt[(int64_t)r * pitch + c];
            *nc = colbt[(int64_t)r * pitch + c];
        }
    };
    int nr, nc;
    bt(cr, cc, &nr, &nc);
    int tick = -1, nref = -1, start = 0;
    while (cr != nr || cc != nc) {
        ++tick;
        if (nc == cc) {
            rcons[tick] = '-';
        } else {
            rcons[tick] = ref[cc - 1];
            ++nref;
        }
        qcons[tick] = nr == cr ? '-' : qry[cr - 1];
        start = cc;
        if (nr == 0 || nc == 0) break;
        cr = nr;
        cc = nc;
        bt(cr, cc, &nr, &nc);
    }
    out->score = red_h[0];
    out->start = start;
    out->nref = nref;
    out->alen = tick + 1;
    out->best_row = red_i[0];
    out->best_col = red_j[0];
}

// One call's device buffers and stream; calls take a free one from the pool.
struct MatWork {
    hipStream_t st = nullptr;
    DevBuf seq, mat, res;
};

std::mutex g_mat_mu;
std::vector<MatWork *> g_mat_free;

MatWork *mat_acquire() {
    {
        std::lock_guard<std::mutex> lk(g_mat_mu);
        if (!g_mat_free.empty()) {
            MatWork *w = g_mat_free.back();
            g_mat_free.pop_back();
            return w;
        }
    }
    MatWork *w = new MatWork;
    if (hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        return nullptr;
    }
    return w;
}

void mat_release(MatWork *w) {
    std::lock_guard<std::mutex> lk(g_mat_mu);
    g_mat_free.push_back(w);
}

int matrix_sw(float *Hm, int *rowbt, int *colbt, const char *refSeq, int refLen,
              const char *qrySeq, int qryLen, const Scoring &sc, float *score, int *startPos,
              int *endPos, int *alignLen, char *refCons, char *qryCons) {
    if (qryLen < 0 || refLen < 0) {
        set_error("negative lengths");
        return kErrInvalid;
    }
    if (qryLen > kMatMaxQ) {
        set_error("smithWaterman_cpu: query of %d bases (at most %d)", qryLen, kMatMaxQ);
        return kErrCapacity;
    }
    // columns stop at the reference's first NUL (src/smithWaterman.cu:135-136)
    int re = 0;
    while (re < refLen && refSeq[re] != '\0') ++re;
    bool f32;
    int rc = validate_scoring(sc, qryLen, re, &f32);
    if (rc) return rc;
    const int q = qryLen, pitch = refLen + 1;
    if (q == 0 || re == 0) {
        // no cell is computed: score 0, best cell (0, 0), nothing emitted
        *score = 0.0f;
        *endPos = *startPos - 1;
        *alignLen = 0;
        return kOk;
    }
    MatWork *w = mat_acquire();
    if (!w) {
        set_error("smithWaterman_cpu: no HIP stream");
        return kErrHip;
    }
    struct Release {
        MatWork *w;
        ~Release() { mat_release(w); }
    } release{w};
    const size_t cells = (size_t)(q + 1) * pitch;
    const size_t cons = (size_t)q + re + 2;
    if ((rc = w->seq.ensure((size_t)q + re + 16))) return rc;
    if ((rc = w->mat.ensure(12 * cells))) return rc;
    if ((rc = w->res.ensure(sizeof(MatOut) + 2 * cons + 16))) return rc;
    char *dq = w->seq.as<char>(), *dr = dq + q;
    float *dH = w->mat.as<float>();
    int *drow = reinterpret_cast<int *>(dH + cells), *dcol = drow + cells;
    MatOut *dout = w->res.as<MatOut>();
    char *drc = reinterpret_cast<char *>(dout + 1), *dqc = drc + cons;
    SWIFT_HIP_TRY(hipMemcpyAsync(dq, qrySeq, q, hipMemcpyHostToDevice, w->st));
    SWIFT_HIP_TRY(hipMemcpyAsync(dr, refSeq, re, hipMemcpyHostToDevice, w->st));
    const size_t lds = 3 * (size_t)(q + 1) * 5;
    hipLaunchKernelGGL(sw_matrix_kernel, dim3(1), dim3(kMatThreads), lds, w->st, dq, q, dr, re,
                       pitch, sc, dH, drow, dcol, dout, drc, dqc);
    SWIFT_HIP_TRY(hipGetLastError());
    MatOut o;
    SWIFT_HIP_TRY(hipMemcpyAsync(&o, dout, sizeof(MatOut), hipMemcpyDeviceToHost, w->st));
    SWIFT_HIP_TRY(hipStreamSynchronize(w->st));
    // the cells the reference writes: rows 1..q, columns 1..re
    const size_t pb = (size_t)pitch * 4, off = (size_t)pitch + 1;
    SWIFT_HIP_TRY(hipMemcpy2DAsync(Hm + off, pb, dH + off, pb, 4 * (size_t)re, q,
                                   hipMemcpyDeviceToHost, w->st));
    SWIFT_HIP_TRY(hipMemcpy2DAsync(rowbt + off, pb, drow + off, pb, 4 * (size_t)re, q,
                                   hipMemcpyDeviceToHost, w->st));
    SWIFT_HIP_TRY(hipMemcpy2DAsync(colbt + off, pb, dcol + off, pb, 4 * (size_t)re, q,
                                   hipMemcpyDeviceToHost, w->st));
    if (o.alen > 0) {
        SWIFT_HIP_TRY(hipMemcpyAsync(refCons, drc, o.alen, hipMemcpyDeviceToHost, w->st));
        SWIFT_HIP_TRY(hipMemcpyAsync(qryCons, dqc, o.alen, hipMemcpyDeviceToHost, w->st));
    }
    SWIFT_HIP_TRY(hipStreamSynchronize(w->st));
    *score = (float)o.score;
    // startPos is only written when a cell is emitted (src/smithWaterman.cu:297,329)
    if (o.alen > 0) *startPos = o.start;
    *endPos = *startPos + o.nref;
    *alignLen = o.alen;
    return kOk;
}

}  // namespace

}  // namespace swift

using namespace swift;

extern "C" {

// Drop-in for smithWaterman_cpu (src/smithWaterman.h:62-66).
void smithWaterman_cpu(float *HMatrix, int *rowBacktracker, int *colBacktracker, char *refSeq,
                       int refLen, char *qrySeq, int qryLen, float match, float mismatch,
                       float gapOpenPenalty, float gapExtPenalty, float *alignScore, int *startPos,
                       int *endPos, int *alignLen, char *refConsensus, char *qryConsensus) {
    Scoring sc;
    int rc = scoring_from_floats(match, mismatch, gapOpenPenalty, gapExtPenalty, &sc);
    if (!rc)
        rc = matrix_sw(HMatrix, rowBacktracker, colBacktracker, refSeq, refLen, qrySeq, qryLen,
                       sc, alignScore, startPos, endPos, alignLen, refConsensus, qryConsensus);
    if (rc) {
        fprintf(stderr, "Error: %s: %s\n", __func__, last_error());
        exit(1);
    }
}

}  // extern "C"
