// sw_kernels.hip -- Smith-Waterman forward sweep, traceback walk and
// length binning for gfx950 (MI355X, wave64).
//
// Recurrence and tie rules are those of the reference CPU path,
// computeMaxScore_cpu / findArrMax / backtrack_cpu
// (src/smithWaterman.cu:114-335):
//   H(i,j) = first maximum of [diag, up, left, 0]
//   diag = H(i-1,j-1) + (match(q_i, R_j) ? match : mismatch)
//   up   = H(i-1,j) + (dir(i-1,j) == up   ? gapExt : gapOpen)
//   left = H(i,j-1) + (dir(i,j-1) == left ? gapExt : gapOpen)
// with the best cell the first strict maximum in row-major order.
//
// Kernel design (written for CDNA4, not translated from the reference's
// one-thread-per-row block kernel):
//  * A lane group of G lanes (G = 16 via DPP row_shr, or 64 via wave_shr)
//    owns one pair per 32-bit lane slot; lane g holds query rows
//    [g*R, g*R+R) in registers and lane g sweeps column j = t - g at step t
//    (a staggered wavefront).  The last row's exits move one lane down per
//    step with a single DPP mov, no LDS round trip and no barrier.
//  * Every cell works on keys K = 4*H + dir (dir 3 diag, 2 up, 1 left,
//    0 start): max(diagKey, upKey, leftKey, 0) yields both the score and the
//    reference's tie-broken direction.  Exits carried to the neighbours are
//    already keys: EU = 4*(H + gap_up) + 2, EL = 4*(H + gap_left) + 1,
//    DR = 4*H + 4*mismatch + 3 (diagonal, before the match bonus).
//  * Batched form: two pairs per lane in the two 16-bit halves of each
//    register (v_pk_* int16 arithmetic), 4 groups of 16 lanes per wave,
//    8 pairs per wave task; a persistent grid pulls tasks from an atomic
//    counter in length-binned, longest-first order.  Any pair whose best
//    score nears the int16 range is flagged and re-run with int32 lanes.
//  * Direction codes (2 bits/cell) are packed per lane into 32-bit words and
//    stored coalesced ([step][word][lane]); a separate kernel walks them
//    back with one lane per pair.
#include "sw_types.hpp"
#include "sw_host.hpp"

#include <type_traits>

namespace swift {

// ---------------------------------------------------------------------------
// lane arithmetic: PK = packed 2 x int16 in a uint32, else one int32
// ---------------------------------------------------------------------------
typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned short v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ v2s as_s(uint32_t x) { return __builtin_bit_cast(v2s, x); }
__device__ __forceinline__ v2u as_us(uint32_t x) { return __builtin_bit_cast(v2u, x); }
__device__ __forceinline__ uint32_t as_u(v2s x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t as_u(v2u x) { return __builtin_bit_cast(uint32_t, x); }

template <bool PK> __host__ __device__ constexpr uint32_t splat(int x) {
    return PK ? ((uint32_t)(x & 0xFFFF) | ((uint32_t)(x & 0xFFFF) << 16)) : (uint32_t)x;
}
template <bool PK> __device__ __forceinline__ uint32_t vadd(uint32_t a, uint32_t b) {
    if constexpr (PK) return as_u(as_s(a) + as_s(b));
    else return a + b;
}
template <bool PK> __device__ __forceinline__ uint32_t vsub(uint32_t a, uint32_t b) {
    if constexpr (PK) return as_u(as_s(a) - as_s(b));
    else return a - b;
}
template <bool PK> __device__ __forceinline__ uint32_t vmax(uint32_t a, uint32_t b) {
    if constexpr (PK) return as_u(__builtin_elementwise_max(as_s(a), as_s(b)));
    else return (uint32_t)max((int)a, (int)b);
}
// a * b + c, a in {0, 1}
template <bool PK> __device__ __forceinline__ uint32_t vmad(uint32_t a, uint32_t b, uint32_t c) {
    if constexpr (PK) return as_u(as_s(a) * as_s(b) + as_s(c));
    else return (uint32_t)((int)a * (int)b + (int)c);
}
template <bool PK> __device__ __forceinline__ uint32_t vmin1(uint32_t a) {
    if constexpr (PK) return as_u(__builtin_elementwise_min(as_us(a), (v2u){1, 1}));
    else return min(a, 1u);
}
// per element (k >> s) & 1 for a constant bit pattern k
template <bool PK> __device__ __forceinline__ uint32_t vbit(uint32_t k, uint32_t s) {
    if constexpr (PK) return as_u(as_us(splat<true>(k)) >> as_us(s)) & 0x00010001u;
    else return (k >> s) & 1u;
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_shift(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, 0xF, 0xF, false);
}
constexpr int kRowShr1 = 0x111;   // lane l <- lane l-1 within each row of 16
constexpr int kWaveShr1 = 0x138;  // lane l <- lane l-1 across the wave

// Byte policy codes: low byte = byte value (reference byte upper-cased),
// bits 8..15 = 0xFF unless the byte is 'N' (the wildcard of isMatch_cpu).
__device__ __forceinline__ uint32_t byte_qcode(unsigned char c) {
    return (uint32_t)c | (c == 'N' ? 0u : 0xFF00u);
}
__device__ __forceinline__ uint32_t byte_rcode(unsigned char c) {
    int u = upper_byte(c) & 0xFF;
    return (uint32_t)u | (u == 'N' ? 0u : 0xFF00u);
}

// substitution flag (0/1 per element)
template <bool PK, int POL>
__device__ __forceinline__ uint32_t sub_flag(uint32_t q, uint32_t r) {
    if constexpr (POL == kOneHot) {
        return vmin1<PK>(q & r);  // 1 = match
    } else {
        static_assert(!PK, "byte policy is int32 only");
        uint32_t x = (q ^ r) & ((q & r) >> 8) & 0xFFu;
        return min(x, 1u);        // 1 = mismatch
    }
}

// ---------------------------------------------------------------------------
// The staggered sweep of one lane group.  G lanes, R rows per lane.
// Each lane slot carries PK ? 2 : 1 pairs.
// ---------------------------------------------------------------------------
template <int G, int R, bool PK, int POL>
struct Sweep {
    static constexpr int KPW = PK ? 8 : 16;          // rows per trace word
    static constexpr int W = (R + KPW - 1) / KPW;     // trace words per step
    static constexpr uint32_t H4MASK = PK ? 0xFFFCFFFCu : ~3u;
    static constexpr uint32_t CMASK = PK ? 0x00030003u : 3u;
    static constexpr int CTRL = G == 16 ? kRowShr1 : kWaveShr1;

    uint32_t q[R];        // query codes of this lane's rows
    uint32_t dr[R];       // DR of (row, j-1)
    uint32_t el[R];       // EL of (row, j-1)
    uint32_t best[R];     // int32: H<<12 | (4095-j) ; PK: row max of 4H
    uint32_t bcol[R];     // PK only: first column of the row max (both halves)
    uint32_t eu_last, dr_last, dr_top;

    uint32_t c_dr, c_dsub, c_eo2, c_eo1, c_de;

    __device__ __forceinline__ void init(const KeyConsts &k) {
        c_dr = splat<PK>(k.dr);
        c_dsub = splat<PK>(k.dsub);
        c_eo2 = splat<PK>(k.eo2);
        c_eo1 = splat<PK>(k.eo1);
        c_de = splat<PK>(k.de);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dr[r] = c_dr;   // column 0: H = 0, start
            el[r] = c_eo1;  // column 0 is not a left move -> open
            best[r] = 0;
            bcol[r] = 0;
        }
        eu_last = c_eo2;
        dr_last = c_dr;
        dr_top = c_dr;
    }

    // One step.  rr: reference code(s) of column j; j = t - g; vmask: per
    // element all-ones where column j is inside the pair (MASKED only).
    template <bool MASKED>
    __device__ __forceinline__ void step(uint32_t rr, int j, uint32_t vmask, uint32_t *acc) {
        // row 0 of this lane takes the lane above's last row: its up exit at
        // column j (computed last step) and its diagonal value at column j-1
        // (received last step).  Lane 0 of the group sees matrix row 0.
        const uint32_t eu_in = dpp_shift<CTRL>(c_eo2, eu_last);
        const uint32_t dr_next = dpp_shift<CTRL>(c_dr, dr_last);
        const uint32_t diag0 = dr_top;
        dr_top = dr_next;
        if (MASKED && j < 0) return;  // left of column 1: keep column-0 state

        uint32_t eu = eu_in, diag = diag0;
        uint32_t jj = 0, colterm = 0;
        if constexpr (PK) jj = splat<true>(j);
        else colterm = (uint32_t)(kMaxCols - j);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint32_t f = sub_flag<PK, POL>(q[r], rr);
            uint32_t dk = vmad<PK>(f, c_dsub, diag);
            uint32_t K = vmax<PK>(vmax<PK>(dk, eu), vmax<PK>(el[r], 0u));
            uint32_t h4 = K & H4MASK;
            uint32_t c = K & CMASK;
            acc[r / KPW] |= c << (2 * (r % KPW));
            diag = dr[r];
            dr[r] = vadd<PK>(h4, c_dr);
            eu = vmad<PK>(vbit<PK>(4u, c), c_de, vadd<PK>(h4, c_eo2));
            el[r] = vmad<PK>(vbit<PK>(2u, c), c_de, vadd<PK>(h4, c_eo1));
            if constexpr (PK) {
                uint32_t hv = MASKED ? (h4 & vmask) : h4;
                uint32_t d = vsub<true>(best[r], hv);          // < 0 <=> new strict max
                uint32_t m = as_u(as_s(d) >> (v2s){15, 15});  // 0xFFFF per new max
                bcol[r] = (m & jj) | (~m & bcol[r]);
                best[r] = vmax<true>(best[r], hv);
            } else {
                uint32_t key = (h4 << (kColBits - 2)) | colterm;
                if (MASKED) key &= vmask;
                best[r] = max(best[r], key);
            }
        }
        eu_last = eu;
        dr_last = dr[R - 1];
    }
};

// ---------------------------------------------------------------------------
// Result of a lane group: per lane, best (H, row, col) over its valid rows,
// then a group-wide max of (H desc, row asc).  Returns through `out`.
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ void group_best(uint32_t lane_h, uint32_t lane_row, uint32_t lane_col,
                                           int g, PairResult *out, bool flag_overflow,
                                           uint32_t h_limit) {
    // key: H high, then smaller row first
    uint64_t key = lane_h ? (((uint64_t)lane_h << 32) | (0xFFFFFFFFu - lane_row)) : 0;
    uint64_t m = key;
#pragma unroll
    for (int s = G / 2; s >= 1; s >>= 1) {
        uint64_t o = __shfl_xor(m, s, G);
        m = o > m ? o : m;
    }
    if (m == 0) {
        if (g == 0 && out) *out = PairResult{0, 0, 0, 0};
    } else if (key == m && out) {
        int fl = (flag_overflow && lane_h * 4 >= h_limit) ? 1 : 0;
        *out = PairResult{(int32_t)lane_h, (int32_t)lane_row, (int32_t)lane_col, fl};
    }
}

// ---------------------------------------------------------------------------
// Plain per-pair kernel: the alignSequences drop-in.  One wave per pair,
// G = 64, byte semantics, int32.  Inputs use the reference's layout:
// pair p's query at querySeq + p*(qlen+1) (all qlen bytes are rows, as in
// smithWaterman_cpu), reference at refSeq + p*(rlen+1), columns stopping at
// its first NUL.  LDS holds the reference codes of the pair.
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(64)
sw_plain_kernel(const char *__restrict__ qseq, const char *__restrict__ rseq, int qlen, int rlen_max,
                KeyConsts kc, int npairs, int pair0, uint32_t *__restrict__ trace,
                int64_t trace_words_per_pair, PairResult *__restrict__ res,
                PairTrace *__restrict__ ptr) {
    using S = Sweep<64, R, false, kBytes>;
    extern __shared__ uint32_t lds_rr[];
    const int p = blockIdx.x;
    if (p >= npairs) return;
    const int g = threadIdx.x;
    const char *qs = qseq + (int64_t)(pair0 + p) * (qlen + 1);
    const char *rs = rseq + (int64_t)(pair0 + p) * (rlen_max + 1);

    // reference codes into LDS; the column count stops at the first NUL
    int first_nul = rlen_max;
    for (int j = g; j < rlen_max; j += 64) {
        unsigned char c = (unsigned char)rs[j];
        lds_rr[j] = byte_rcode(c);
        if (c == 0) first_nul = min(first_nul, j);
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) first_nul = min(first_nul, __shfl_xor(first_nul, s, 64));
    const int rlen = first_nul;
    __syncthreads();

    S sw;
    sw.init(kc);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int row = g * R + r;
        sw.q[r] = row < qlen ? byte_qcode((unsigned char)qs[row]) : 0u;
    }
    uint32_t *tr = trace + (int64_t)p * trace_words_per_pair;
    const int T = rlen + 63;
    const int main_lo = 63, main_hi = rlen;  // every lane inside [0, rlen)
    uint32_t acc[S::W];
    auto run = [&](auto masked, int t) {
        int j = t - g;
        uint32_t rr = (j >= 0 && j < rlen) ? lds_rr[j] : 0u;
        uint32_t vmask = (j >= 0 && j < rlen) ? ~0u : 0u;
#pragma unroll
        for (int w = 0; w < S::W; ++w) acc[w] = 0;
        sw.template step<decltype(masked)::value>(rr, j, vmask, acc);
#pragma unroll
        for (int w = 0; w < S::W; ++w) tr[((int64_t)t * S::W + w) * 64 + g] = acc[w];
    };
    int t = 0;
    for (; t < min(main_lo, T); ++t) run(std::true_type{}, t);
    for (; t < main_hi; ++t) run(std::false_type{}, t);
    for (; t < T; ++t) run(std::true_type{}, t);

    // lane best over its valid rows: row-major first maximum
    uint32_t bh = 0, brow = 0, bcol = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int row = g * R + r;
        uint32_t h = sw.best[r] >> kColBits;
        if (row < qlen && h > bh) {
            bh = h;
            brow = row + 1;
            bcol = kMaxCols - (sw.best[r] & kMaxCols) + 1;
        }
    }
    group_best<64>(bh, brow, bcol, g, res + p, false, 0);
    if (g == 0)
        ptr[p] = PairTrace{(int64_t)p * trace_words_per_pair, 0, (int16_t)R, (uint8_t)S::W, 0};
}

// ---------------------------------------------------------------------------
// Batched persistent kernels.
// ---------------------------------------------------------------------------
// One wave task: 4 groups of 16 lanes; PK -> 2 pairs per group (halves).
template <int R, bool PK, int POL>
__device__ __forceinline__ void run_task(const BatchArgs &a, const Task &tk, uint32_t *lds) {
    using S = Sweep<kBatchG, R, PK, POL>;
    constexpr int PPG = PK ? 2 : 1;  // pairs per group
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4, g = lane & 15;
    uint32_t *rr_lds = lds + grp * kBatchCols;

    // the group's pairs
    int pid[PPG], ql[PPG], rl[PPG];
    int64_t qo[PPG], ro[PPG];
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        int slot = grp * PPG + h;
        pid[h] = slot < tk.npairs ? a.order[tk.first + slot] : -1;
        ql[h] = pid[h] >= 0 ? a.qlen[pid[h]] : 0;
        rl[h] = pid[h] >= 0 ? a.rlen[pid[h]] : 0;
        qo[h] = pid[h] >= 0 ? a.qoff[pid[h]] : 0;
        ro[h] = pid[h] >= 0 ? a.roff[pid[h]] : 0;
    }
    // reference codes of the group's pairs into LDS (both halves packed)
    const int ncols = tk.steps - (kBatchG - 1);
    for (int j = g; j < ncols; j += kBatchG) {
        uint32_t v = 0;
#pragma unroll
        for (int h = 0; h < PPG; ++h) {
            if (j < rl[h]) {
                unsigned char c = (unsigned char)a.rseq[ro[h] + j];
                uint32_t code = POL == kOneHot ? ref_code(c) : byte_rcode(c);
                v |= code << (16 * h);
            }
        }
        rr_lds[j] = v;
    }
    S sw;
    sw.init(a.kc);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int row = g * R + r;
        uint32_t v = 0;
#pragma unroll
        for (int h = 0; h < PPG; ++h) {
            if (row < ql[h]) {
                unsigned char c = (unsigned char)a.qseq[qo[h] + row];
                uint32_t code = POL == kOneHot ? query_code(c) : byte_qcode(c);
                v |= code << (16 * h);
            }
        }
        sw.q[r] = v;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);  // LDS writes of this wave visible to its lanes
    __builtin_amdgcn_wave_barrier();

    uint32_t *tr = a.trace + tk.trace_base;
    uint32_t vm_full = 0, acc[S::W];
    auto run = [&](auto masked, int t) {
        int j = t - g;
        bool in = j >= 0 && j < ncols;
        uint32_t rr = in ? rr_lds[j] : 0u;
        uint32_t vmask = 0;
        if constexpr (decltype(masked)::value) {
#pragma unroll
            for (int h = 0; h < PPG; ++h)
                if (j >= 0 && j < rl[h]) vmask |= (PK ? 0xFFFFu : 0xFFFFFFFFu) << (16 * h);
        }
#pragma unroll
        for (int w = 0; w < S::W; ++w) acc[w] = 0;
        sw.template step<decltype(masked)::value>(rr, j, vmask, acc);
#pragma unroll
        for (int w = 0; w < S::W; ++w) tr[((int64_t)t * S::W + w) * 64 + lane] = acc[w];
    };
    (void)vm_full;
    const int T = tk.steps;
    const int lo = min(kBatchG - 1, T), hi = max(lo, tk.min_rlen);
    int t = 0;
    for (; t < lo; ++t) run(std::true_type{}, t);
    for (; t < hi; ++t) run(std::false_type{}, t);
    for (; t < T; ++t) run(std::true_type{}, t);

    // per half: lane best over valid rows, then group max
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        uint32_t bh = 0, brow = 0, bcol = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int row = g * R + r;
            uint32_t hv, cv;
            if constexpr (PK) {
                hv = ((sw.best[r] >> (16 * h)) & 0xFFFFu) >> 2;
                cv = ((sw.bcol[r] >> (16 * h)) & 0xFFFFu) + 1;
            } else {
                hv = sw.best[r] >> kColBits;
                cv = kMaxCols - (sw.best[r] & kMaxCols) + 1;
            }
            if (row < ql[h] && hv > bh) {
                bh = hv;
                brow = row + 1;
                bcol = cv;
            }
        }
        PairResult *o = pid[h] >= 0 ? a.res + pid[h] : nullptr;
        group_best<kBatchG>(bh, brow, bcol, g, o, PK, a.h_limit);
        if (g == 0 && pid[h] >= 0)
            a.ptr[pid[h]] = PairTrace{tk.trace_base, grp * kBatchG, (int16_t)R, (uint8_t)S::W,
                                      (int8_t)(h | (PK ? 2 : 0))};
    }
    __builtin_amdgcn_wave_barrier();
}

// Row classes: R per lane with 16-lane groups covers query length 16*R.
// One kernel per R: each gets the register budget its R needs (a single
// kernel switching over all R is allocated for the worst case).
#define SWIFT_R16_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) \
    X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)
#define SWIFT_R32_CASES(X) X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20)

template <int R, bool PK, int POL>
__global__ void __launch_bounds__(256)
sw_batch_kernel(BatchArgs a) {
    __shared__ uint32_t lds[4][4 * kBatchCols];
    const int wave = threadIdx.x >> 6;
    const int lo = *a.task_lo, hi = *a.task_hi;
    if (lo >= hi) return;
    for (;;) {
        int tid = 0;
        if ((threadIdx.x & 63) == 0) tid = atomicAdd(a.next_task, 1);
        tid = __builtin_amdgcn_readfirstlane(__shfl(tid, 0, 64)) + lo;
        if (tid >= hi) break;
        const Task tk = a.tasks[tid];
        run_task<R, PK, POL>(a, tk, lds[wave]);
    }
}

// ---------------------------------------------------------------------------
// Walk-back: one lane per pair (backtrack_cpu, src/smithWaterman.cu:263-335).
// Emits the consensus strings in walk order (the reference stores them
// reversed), start/end (1-based window columns) and the alignment length.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int trace_code(const uint32_t *tr, const PairTrace &pt, int i, int j) {
    const int R = pt.rows;
    const int kpw = (pt.half & 2) ? 8 : 16;
    const int g = (i - 1) / R, k = (i - 1) - g * R;
    const int lane = pt.lane0 + g, t = (j - 1) + g;
    const int w = k / kpw;
    const int bit = (pt.half & 1) * 16 + 2 * (k - w * kpw);
    uint32_t word = tr[pt.base + ((int64_t)t * pt.words + w) * 64 + lane];
    return (word >> bit) & 3;
}

__global__ void __launch_bounds__(256)
sw_walk_kernel(WalkArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    const int gp = a.pair0 + p;
    const int ri = a.res_global ? gp : p;
    const PairResult r = a.res[ri];
    const char *qs = a.qseq + (a.qoff ? a.qoff[gp] : (int64_t)gp * a.qstride);
    const char *rs = a.rseq + (a.roff ? a.roff[gp] : (int64_t)gp * a.rstride);
    const int64_t co = a.coff ? a.coff[gp] : (int64_t)gp * a.cstride;
    char *qc = a.qcons + co;
    char *rc = a.rcons + co;
    int start = 0, nrb = -1, tick = 0;
    const int qlen = a.qlen ? a.qlen[gp] : (int)a.qstride - 1;
    const int rlen = a.rlen ? a.rlen[gp] : (int)a.rstride - 1;
    if (r.score > 0 && r.row >= 1 && r.row <= qlen && r.col >= 1 && r.col <= rlen) {
        const PairTrace pt = a.ptr[ri];
        int ci = r.row, cj = r.col, ni, nj;
        auto back = [&](int i, int j, int &oi, int &oj) {
            int c = trace_code(a.trace, pt, i, j);
            oi = (c & 2) ? i - 1 : i;           // diag (3) and up (2) move up
            oj = (c == 3 || c == 1) ? j - 1 : j; // diag (3) and left (1) move left
        };
        back(ci, cj, ni, nj);
        for (;;) {
            bool moves = ci != ni || cj != nj;
            if (!moves) break;
            if (nj == cj) rc[tick] = '-';
            else { rc[tick] = rs[cj - 1]; ++nrb; }
            qc[tick] = (ni == ci) ? '-' : qs[ci - 1];
            start = cj;
            ++tick;
            if (ni == 0 || nj == 0) break;  // the target is in row/column 0
            ci = ni;
            cj = nj;
            back(ci, cj, ni, nj);
        }
    }
    qc[tick] = '\0';
    rc[tick] = '\0';
    if (a.score_f) a.score_f[gp] = (float)r.score;
    if (a.score_i) a.score_i[gp] = r.score;
    a.start[gp] = start;
    a.end[gp] = start + nrb;
    a.alen[gp] = tick;
}

// ---------------------------------------------------------------------------
// Device-side length binning for the batched kernels.
//
// Pass 1 (classify): each pair gets a class -- int16 one-hot (fast path),
// int32 byte semantics (reference bytes outside ACGTN, scores beyond the
// int16 range, or an int16 overflow re-run) -- and a rows-per-lane class R.
// Bins are (class, R, rlen/8) and are visited longest-first, so tasks are
// served in decreasing cost (LPT) from the persistent work queue.
// ---------------------------------------------------------------------------
__global__ void sw_bin_classify(BinArgs a) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    int ql = a.qlen[p], rl = a.rlen[p];
    if (ql < 0 || rl < 0 || ql > kMaxBatchQ || rl > kMaxBatchR) {
        // lengths beyond the batched kernels: the pair is left out (score 0)
        // and the call reports the error
        a.bin_of[p] = -1;
        atomicOr(a.error, 1);
        return;
    }
    int cls;
    if (a.prev) {
        if (!(a.prev[p].flags & 1)) { a.bin_of[p] = -1; return; }
        cls = 1;
    } else {
        cls = a.force32 ? 1 : 0;
        if (!cls) {
            const char *r = a.rseq + a.roff[p];
            for (int j = 0; j < rl; ++j)
                if (ref_code((unsigned char)r[j]) == 0) { cls = 1; break; }
        }
    }
    int R = max(1, (ql + kBatchG - 1) / kBatchG);
    if (cls == 1) R = r32_class(R);
    int lb = min(kLenBins - 1, rl >> 3);
    // bin order: class, then R descending, then rlen descending
    int bin = (cls * kRBins + (kRBins - R)) * kLenBins + (kLenBins - 1 - lb);
    a.bin_of[p] = bin;
    atomicAdd(&a.bin_count[bin], 1);
}

// single workgroup: bin prefix sums and per-(class, R) task counts
__global__ void __launch_bounds__(1024) sw_bin_scan(TaskArgs a) {
    __shared__ int32_t part[1024];
    const int tid = threadIdx.x;
    constexpr int per = (kNumBins + 1023) / 1024;
    int32_t loc[per], s = 0;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        int b = tid * per + i;
        loc[i] = b < kNumBins ? a.bin_count[b] : 0;
        s += loc[i];
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int32_t v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int32_t run = part[tid] - s;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        int b = tid * per + i;
        if (b < kNumBins) a.bin_cursor[b] = run;
        run += loc[i];
    }
    __syncthreads();
    if (tid == 0) {
        int32_t t0 = 0;
        for (int seg = 0; seg < 2 * kRBins; ++seg) {
            int cls = seg / kRBins;
            int32_t first = a.bin_cursor[seg * kLenBins];
            int32_t len = 0;
            for (int lb = 0; lb < kLenBins; ++lb) len += a.bin_count[seg * kLenBins + lb];
            int slots = cls == 0 ? 8 : 4;
            a.seg_task0[seg] = t0;
            a.seg_first[seg] = first;
            a.seg_len[seg] = len;
            t0 += (len + slots - 1) / slots;
        }
        a.seg_task0[2 * kRBins] = t0;
        *a.ntasks = t0;
        *a.npairs_binned = part[1023];
    }
}

__global__ void sw_bin_scatter(const int32_t *bin_of, int npairs, int32_t *bin_cursor,
                               int32_t *order) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    int b = bin_of[p];
    if (b < 0) return;
    int pos = atomicAdd(&bin_cursor[b], 1);
    order[pos] = p;
}

// task descriptors (trace sizes filled, bases scanned afterwards)
__global__ void sw_build_tasks(const int32_t *seg_task0, const int32_t *seg_first,
                               const int32_t *seg_len, const int32_t *ntasks_p,
                               const int32_t *order, const int32_t *rlen, Task *tasks,
                               int64_t *task_words) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int ntasks = *ntasks_p;
    if (t >= ntasks) return;
    int seg = 0;
    while (seg + 1 < 2 * kRBins && seg_task0[seg + 1] <= t) ++seg;
    int cls = seg / kRBins;
    int R = kRBins - (seg % kRBins);
    int slots = cls == 0 ? 8 : 4;
    int k = t - seg_task0[seg];
    int first = seg_first[seg] + k * slots;
    int n = min(slots, seg_len[seg] - k * slots);
    int mx = 0, mn = 1 << 30;
    for (int i = 0; i < n; ++i) {
        int rl = rlen[order[first + i]];
        mx = max(mx, rl);
        mn = min(mn, rl);
    }
    Task tk;
    tk.trace_base = 0;
    tk.first = first;
    tk.npairs = (int16_t)n;
    tk.rows = (int16_t)R;
    tk.steps = mx + kBatchG - 1;
    tk.min_rlen = mn;
    tasks[t] = tk;
    int kpw = cls == 0 ? 8 : 16;
    int W = (R + kpw - 1) / kpw;
    task_words[t] = (int64_t)tk.steps * W * 64;
}

// single workgroup exclusive scan of task trace sizes into the descriptors
__global__ void __launch_bounds__(1024)
sw_task_bases(Task *tasks, const int64_t *task_words, const int32_t *ntasks_p,
              const int64_t *base0, int64_t *total_words) {
    __shared__ int64_t part[1024];
    const int tid = threadIdx.x;
    const int n = *ntasks_p;
    const int per = (n + 1023) / 1024;
    int64_t s = 0;
    for (int i = 0; i < per; ++i) {
        int t = tid * per + i;
        if (t < n) s += task_words[t];
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int64_t run = part[tid] - s + (base0 ? *base0 : 0);
    for (int i = 0; i < per; ++i) {
        int t = tid * per + i;
        if (t < n) {
            tasks[t].trace_base = run;
            run += task_words[t];
        }
    }
    if (tid == 1023) *total_words = part[1023] + (base0 ? *base0 : 0);
}

__global__ void fill_i32(int32_t *p, int n, int32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// host-visible launch helpers (used by sw_host.cpp)
// ---------------------------------------------------------------------------
template <int R>
static hipError_t launch_plain_r(const char *q, const char *r, int qlen, int rlen_max,
                                 const KeyConsts &kc, int npairs, int pair0, uint32_t *trace,
                                 int64_t wpp, PairResult *res, PairTrace *ptr, hipStream_t st) {
    size_t lds = sizeof(uint32_t) * (size_t)max(rlen_max, 1);
    hipLaunchKernelGGL(sw_plain_kernel<R>, dim3(npairs), dim3(64), lds, st, q, r, qlen, rlen_max,
                       kc, npairs, pair0, trace, wpp, res, ptr);
    return hipGetLastError();
}

int plain_rows_for(int qlen) {
    int R = (qlen + 63) / 64;
    if (R < 1) R = 1;
    if (R <= 8) return R;
    if (R <= 10) return 10;
    if (R <= 12) return 12;
    if (R <= 16) return 16;
    return -1;
}

int plain_trace_words(int R) { return (R + 15) / 16; }

hipError_t launch_plain(const char *q, const char *r, int qlen, int rlen_max, const Scoring &sc,
                        int npairs, int pair0, uint32_t *trace, int64_t wpp, PairResult *res,
                        PairTrace *ptr, hipStream_t st) {
    KeyConsts kc = make_consts(sc, kBytes);
    switch (plain_rows_for(qlen)) {
#define SWIFT_PCASE(RR) \
    case RR: return launch_plain_r<RR>(q, r, qlen, rlen_max, kc, npairs, pair0, trace, wpp, res, ptr, st);
        SWIFT_PCASE(1) SWIFT_PCASE(2) SWIFT_PCASE(3) SWIFT_PCASE(4) SWIFT_PCASE(5)
        SWIFT_PCASE(6) SWIFT_PCASE(7) SWIFT_PCASE(8) SWIFT_PCASE(10) SWIFT_PCASE(12)
        SWIFT_PCASE(16)
#undef SWIFT_PCASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_walk(const WalkArgs &wa, hipStream_t st) {
    if (wa.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(sw_walk_kernel, dim3((wa.npairs + 255) / 256), dim3(256), 0, st, wa);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &ba, bool int16, int rows, int grid, hipStream_t st) {
    if (int16) {
        switch (rows) {
#define SWIFT_L16(RR) \
    case RR: hipLaunchKernelGGL((sw_batch_kernel<RR, true, kOneHot>), dim3(grid), dim3(256), 0, st, ba); break;
            SWIFT_R16_CASES(SWIFT_L16)
#undef SWIFT_L16
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (rows) {
#define SWIFT_L32(RR) \
    case RR: hipLaunchKernelGGL((sw_batch_kernel<RR, false, kBytes>), dim3(grid), dim3(256), 0, st, ba); break;
            SWIFT_R32_CASES(SWIFT_L32)
#undef SWIFT_L32
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_binning(const BinArgs &bin, const TaskArgs &ta, int32_t *order, Task *tasks,
                          int64_t *task_words, const int64_t *base0, int64_t *total_words,
                          int max_tasks,
                          hipStream_t st) {
    int n = bin.npairs;
    hipLaunchKernelGGL(fill_i32, dim3((kNumBins + 255) / 256), dim3(256), 0, st, bin.bin_count,
                       kNumBins, 0);
    if (n > 0)
        hipLaunchKernelGGL(sw_bin_classify, dim3((n + 255) / 256), dim3(256), 0, st, bin);
    hipLaunchKernelGGL(sw_bin_scan, dim3(1), dim3(1024), 0, st, ta);
    if (n > 0)
        hipLaunchKernelGGL(sw_bin_scatter, dim3((n + 255) / 256), dim3(256), 0, st, bin.bin_of,
                           n, ta.bin_cursor, order);
    if (max_tasks > 0) {
        hipLaunchKernelGGL(sw_build_tasks, dim3((max_tasks + 255) / 256), dim3(256), 0, st,
                           ta.seg_task0, ta.seg_first, ta.seg_len, ta.ntasks, order, bin.rlen,
                           tasks, task_words);
    }
    hipLaunchKernelGGL(sw_task_bases, dim3(1), dim3(1024), 0, st, tasks, task_words, ta.ntasks,
                       base0, total_words);
    return hipGetLastError();
}

hipError_t launch_fill(int32_t *p, int n, int32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace swift
