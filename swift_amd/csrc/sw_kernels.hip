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
#include "scan.hpp"

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

// Hides a value's range from the optimiser (no instruction is emitted), so
// that products by 0/1 flags and masked merges stay packed v_pk_* /
// v_bitop3 arithmetic instead of being rewritten into per-half compares and
// selects (there is no packed select).
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
    asm("" : "+v"(x));
    return x;
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

// Byte policy (int32 lanes): 1 = mismatch.
__device__ __forceinline__ uint32_t byte_mismatch(uint32_t q, uint32_t r) {
    uint32_t x = (q ^ r) & ((q & r) >> 8) & 0xFFu;
    return min(x, 1u);
}

// Packed one-hot policy: the match flag of both halves from one v_perm_b32.
// A row holds the query's match table over the reference classes A C G T
// (0-3): 4 bytes per half, 0xFF where the query byte matches the class
// (query N: all four; any other query byte: none).  A column's word holds,
// per half, a perm selector in both bytes: class k -> k (half 0, table in
// bytes 0-3) or 4 + k (half 1, bytes 4-7); reference N -> 13 (0xFF, it
// matches every query byte); a padding column or another byte -> 12 (0x00).
// perm(table half 1, table half 0, selector) is 0xFFFF per matching half.
__host__ __device__ constexpr uint32_t query_table(uint32_t mask) {
    uint32_t t = 0;
    for (int k = 0; k < 4; ++k)
        if ((mask >> k) & 1u) t |= 0xFFu << (8 * k);
    return t;
}
__host__ __device__ constexpr uint32_t column_selector(uint32_t cls, int half) {
    const uint32_t s = cls < 4 ? cls + 4u * (uint32_t)half : cls == 4 ? 13u : 12u;
    return s * 0x0101u;
}

// ---------------------------------------------------------------------------
// The staggered sweep of one lane group.  G lanes, R rows per lane.
// Each lane slot carries PK ? 2 : 1 pairs.
// ---------------------------------------------------------------------------
//
// The packed kernels (PK) run every step unmasked: they take only scorings
// whose mismatch and gap moves all lose score (mismatch < 0, gaps <= -1), so
// cells left of column 1 stay in the column-0 state and cells right of a
// pair's last column (padding selectors: never a match) score strictly below
// the pair's best real cell.  The left exit is kept >= 0 by a saturating
// subtract, which makes the zero candidate of the max implicit.
template <int G, int R, bool PK, int POL>
struct Sweep {
    static_assert(PK == (POL == kOneHot), "packed lanes use the one-hot policy, int32 the bytes");
    static constexpr int KPW = PK ? 8 : 16;          // rows per trace word
    static constexpr int W = (R + KPW - 1) / KPW;     // trace words per step
    static constexpr uint32_t H4MASK = PK ? 0xFFFCFFFCu : ~3u;
    static constexpr uint32_t CMASK = PK ? 0x00030003u : 3u;
    static constexpr int CTRL = G == 16 ? kRowShr1 : kWaveShr1;

    uint32_t q[R];             // int32: query byte codes; PK: match tables of half 0
    uint32_t q2[PK ? R : 1];   // PK: match tables of half 1
    uint32_t dr[R];       // DR of (row, j-1)
    uint32_t el[R];       // EL of (row, j-1)
    // int32: best[r] = max over the row of H<<12 | (4095-j).
    // PK: best[r] = running max of the key 32H + (15 - t%16) (unsigned),
    // sealed to 32H + 31 at each block end; bcol[r] = block << 5 | key code
    // of the row's first strict maximum (see block_end).
    uint32_t best[R];
    uint32_t bcol[R];
    uint32_t eu_last, dr_last, dr_top;

    uint32_t c_dr, c_dsub, c_eo2, c_eo1, c_de, c_sign, c_eight;
    // PK: 16-bit tables indexed by direction code, laid out for v_perm_b32
    // (entry of code c in bytes 2c, 2c+1): the up exit, and the left exit's
    // penalty (subtracted with saturation at 0)
    uint32_t lut_u_lo, lut_u_hi, pen_l_lo, pen_l_hi;

    static __device__ __forceinline__ void lut(int v0, int v1, int v2, int v3, uint32_t &lo,
                                               uint32_t &hi) {
        lo = (uint32_t)(v0 & 0xFFFF) | ((uint32_t)(v1 & 0xFFFF) << 16);
        hi = (uint32_t)(v2 & 0xFFFF) | ((uint32_t)(v3 & 0xFFFF) << 16);
    }

    __device__ __forceinline__ void init(const KeyConsts &k) {
        c_dr = splat<PK>(k.dr);
        c_dsub = splat<PK>(PK ? -k.dsub : k.dsub);  // PK: times a 0xFFFF flag
        c_eo2 = splat<PK>(k.eo2);
        c_eo1 = splat<PK>(k.eo1);
        c_de = splat<PK>(k.de);
        c_sign = splat<PK>(k.sign);
        c_eight = opaque(splat<PK>(8));  // kept a multiplier: one v_pk_mad per key
        // up exit: open unless the cell came from above (code 2); left exit:
        // open unless it came from the left (code 1)
        lut(k.eo2, k.eo2, k.eo2 + k.de, k.eo2, lut_u_lo, lut_u_hi);
        lut(-k.eo1, -(k.eo1 + k.de), -k.eo1, -k.eo1, pen_l_lo, pen_l_hi);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dr[r] = c_dr;   // column 0: H = 0, start
            // column 0 is not a left move -> open; PK: clamped at 0
            el[r] = PK ? 0u : c_eo1;
            best[r] = PK ? 0x001F001Fu : 0u;  // PK: H = 0, sealed
            bcol[r] = 0;
        }
        eu_last = c_eo2;
        dr_last = c_dr;
        dr_top = c_dr;
    }

    // One step.  rr: reference code(s) of column j; j = t - g; vmask: per
    // element all-ones where column j is inside the pair (MASKED only);
    // code (PK): 15 - t%16 in both halves.
    template <bool MASKED>
    __device__ __forceinline__ void step(uint32_t rr, int j, uint32_t vmask, uint32_t code,
                                         uint32_t *acc) {
        // row 0 of this lane takes the lane above's last row: its up exit at
        // column j (computed last step) and its diagonal value at column j-1
        // (received last step).  Lane 0 of the group sees matrix row 0.
        const uint32_t eu_in = dpp_shift<CTRL>(c_eo2, eu_last);
        const uint32_t dr_next = dpp_shift<CTRL>(c_dr, dr_last);
        const uint32_t diag0 = dr_top;
        dr_top = dr_next;
        if (MASKED && j < 0) return;  // left of column 1: keep column-0 state

        uint32_t colterm = 0;
        if constexpr (!PK) colterm = (uint32_t)(kMaxCols - j);
        // The diagonal, left and zero candidates depend only on column j-1:
        // fold them for every row first, so the row-to-row chain through the
        // up exit eu is just max -> code -> exit per row, with the rows'
        // independent work free to fill its gaps (max is order-free: the
        // direction code in the key's low bits breaks ties).
        uint32_t pre[R];
        if constexpr (PK) {
            // all diagonal candidates first, then the maxes: a packed result
            // read by the very next instruction costs a wait state
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t diag = r == 0 ? diag0 : dr[r - 1];
                const uint32_t f = __builtin_amdgcn_perm(q2[r], q[r], rr);  // 0xFFFF: match
                pre[r] = vmad<true>(f, c_dsub, diag);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) pre[r] = vmax<true>(pre[r], el[r]);  // el >= 0
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t diag = r == 0 ? diag0 : dr[r - 1];
                const uint32_t dk = vmad<false>(byte_mismatch(q[r], rr), c_dsub, diag);
                pre[r] = vmax<false>(vmax<false>(dk, el[r]), 0u);
            }
        }
        uint32_t eu = eu_in;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint32_t K = vmax<PK>(pre[r], eu);
            uint32_t h4 = K & H4MASK;
            uint32_t c = K & CMASK;
            acc[r / KPW] |= c << (2 * (r % KPW));
            dr[r] = vadd<PK>(h4, c_dr);
            if constexpr (PK) {
                // per-half gap exits from the 16-bit tables
                const uint32_t sel = c * 0x0202u + 0x01000100u;
                eu = vadd<true>(h4, __builtin_amdgcn_perm(lut_u_hi, lut_u_lo, sel));
                el[r] = as_u(__builtin_elementwise_sub_sat(
                    as_us(h4), as_us(__builtin_amdgcn_perm(pen_l_hi, pen_l_lo, sel))));
            } else {
                // int32: a select of the extension delta (a 32-bit multiply
                // by the 0/1 flag would be a quarter-rate v_mul_lo_u32)
                eu = h4 + c_eo2 + (c == 2u ? c_de : 0u);
                el[r] = h4 + c_eo1 + (c == 1u ? c_de : 0u);
            }
            if constexpr (PK) {
                static_assert(!(PK && MASKED), "packed steps are never masked");
                // key = 32H + code: within a block of 16 steps the largest
                // key is the row's largest H at its earliest step
                best[r] = as_u(__builtin_elementwise_max(
                    as_us(best[r]), as_us(vmad<true>(h4, c_eight, code))));
            } else {
                uint32_t key = (h4 << (kColBits - 2)) | colterm;
                if (MASKED) key &= vmask;
                best[r] = max(best[r], key);
            }
        }
        eu_last = eu;
        dr_last = dr[R - 1];
    }

    // PK, after the last step of a block of 16 (and after the sweep's last
    // step).  The previous block end sealed the running maximum to 32H + 31,
    // which every later key of the same H (code <= 15) leaves in place; a key
    // of larger H replaces it.  So a row's maximum rose in this block exactly
    // where the low 5 bits are no longer 31, and the row's first strict
    // maximum is then in this block, at the key's code.  blk32 = block index
    // << 5 in both halves.
    __device__ __forceinline__ void block_end(uint32_t blk32) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t lo = best[r] & 0x001F001Fu;
            const uint32_t m = as_u(as_s(vsub<true>(lo, 0x001F001Fu)) >> as_s(c_sign));
            bcol[r] = (m & (lo | blk32)) | (~m & bcol[r]);
            best[r] |= 0x001F001Fu;
        }
    }
};

// ---------------------------------------------------------------------------
// Result of a lane group: per lane, best (H, row, col) over its valid rows,
// then a group-wide max of (H desc, row asc) -- the reference's first strict
// maximum in row-major order.  Every lane of the group receives the result.
// ---------------------------------------------------------------------------
struct GroupResult {
    int32_t score, row, col, flag;
};

template <int G>
__device__ __forceinline__ GroupResult group_best(uint32_t lane_h, uint32_t lane_row,
                                                  uint32_t lane_col, int R, bool flag_overflow,
                                                  uint32_t h_limit, uint32_t force_flag) {
    uint64_t key = lane_h ? (((uint64_t)lane_h << 32) | (0xFFFFFFFFu - lane_row)) : 0;
    uint64_t m = key;
#pragma unroll
    for (int s = G / 2; s >= 1; s >>= 1) {
        uint64_t o = __shfl_xor(m, s, G);
        m = o > m ? o : m;
    }
    GroupResult gr;
    gr.score = (int32_t)(m >> 32);
    gr.row = m ? (int32_t)(0xFFFFFFFFu - (uint32_t)m) : 0;
    // the column lives in the lane that owns the best row
    const int base = (int)(threadIdx.x & 63) & ~(G - 1);
    const int src = base + (gr.row > 0 ? (gr.row - 1) / R : 0);
    gr.col = m ? __shfl((int)lane_col, src, 64) : 0;
    gr.flag = ((flag_overflow && (uint32_t)gr.score * 4 >= h_limit) || force_flag) ? 1 : 0;
    return gr;
}

// Output of one consensus string, four bytes per store when its start is
// 4-byte aligned.
struct ConsWriter {
    char *base;
    uint32_t word = 0;
    int n = 0;
    bool aligned;
    __device__ explicit ConsWriter(char *b) : base(b), aligned(((uintptr_t)b & 3) == 0) {}
    __device__ __forceinline__ void put(char c) {
        if (aligned) {
            word |= (uint32_t)(unsigned char)c << (8 * (n & 3));
            if ((n & 3) == 3) {
                *(uint32_t *)(base + (n & ~3)) = word;
                word = 0;
            }
        } else {
            base[n] = c;
        }
        ++n;
    }
    __device__ __forceinline__ void finish() {  // NUL terminator
        put('\0');
        if (aligned && (n & 3)) *(uint32_t *)(base + (n & ~3)) = word;
    }
};

// Trace layout: granules of 4 steps, [t/4][w][lane][t%4] (16 bytes per lane
// and word index), so a walk that stays on one lane for a few columns reads
// one granule instead of one cache line per step.
struct TraceReader {
    const uint4 *tr;
    PairTrace pt;
    uint32_t inv;  // ceil(2^16 / R): x / R == (x * inv) >> 16 for R <= 20, x < 2048
    int64_t cached = -1;
    uint4 gran;
    __device__ TraceReader(const uint4 *t, const PairTrace &p)
        : tr(t), pt(p), inv((65536u + (uint32_t)p.rows - 1) / (uint32_t)p.rows) {}
    __device__ __forceinline__ int code(int i, int j) {
        const int R = pt.rows;
        const int kpw = (pt.half & 2) ? 8 : 16;
        const int g = (int)(((uint32_t)(i - 1) * inv) >> 16), k = (i - 1) - g * R;
        const int lane = pt.lane0 + g, t = (j - 1) + g;
        const int w = k / kpw;
        const int bit = (pt.half & 1) * 16 + 2 * (k - w * kpw);
        const int64_t gi = pt.base / 4 + ((int64_t)(t >> 2) * pt.words + w) * 64 + lane;
        if (gi != cached) {
            gran = tr[gi];
            cached = gi;
        }
        const int u = t & 3;
        uint32_t word = u == 0 ? gran.x : u == 1 ? gran.y : u == 2 ? gran.z : gran.w;
        return (word >> bit) & 3;
    }
};

// TraceReader for the deferred walk kernel, which issues in the gaps the
// sweeps leave: per step only 32-bit arithmetic on (i, j), the pair's first
// granule resolved once, rows per word a shift, and the step's word picked
// by selects, not branches (80 instructions per step against 125).  The
// fused walk inside the sweep kernels keeps TraceReader: this one's extra
// live values spill there.
struct TraceReaderLean {
    const uint4 *tr;  // the pair's lane group at step 0, word 0
    int R, W, kl, bitb;
    uint32_t inv;
    int cached = -1;
    uint4 gran;
    __device__ TraceReaderLean(const uint4 *t, const PairTrace &p)
        : tr(t + p.base / 4 + p.lane0), R(p.rows), W(p.words), kl((p.half & 2) ? 3 : 4),
          bitb((p.half & 1) * 16), inv((65536u + (uint32_t)p.rows - 1) / (uint32_t)p.rows) {}
    __device__ __forceinline__ int code(int i, int j) {
        const int g = (int)(((uint32_t)(i - 1) * inv) >> 16), k = (i - 1) - g * R;
        const int t = (j - 1) + g;
        const int w = k >> kl;  // rows per word: 8 (packed halves) or 16
        const int bit = bitb + 2 * (k & ((1 << kl) - 1));
        // 32-bit granule index (a 24-bit multiply: t < 2^12, W <= 3); the
        // byte offset is formed once, in 64 bits, at the load
        const uint32_t gi = ((__umul24((uint32_t)(t >> 2), (uint32_t)W) + w) << 6) + g;
        if ((int)gi != cached) {
            gran = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(tr) +
                                                     ((size_t)gi << 4));
            cached = (int)gi;
        }
        const uint32_t lo = (t & 1) ? gran.y : gran.x, hi = (t & 1) ? gran.w : gran.z;
        const uint32_t word = (t & 2) ? hi : lo;
        return (word >> bit) & 3;
    }
};

// Word offset of step t, trace word w, lane `lane` in the granule layout.
template <int W>
__device__ __forceinline__ int64_t trace_word(int t, int w, int lane) {
    return (((int64_t)(t >> 2) * W + w) * 64 + lane) * 4 + (t & 3);
}

// Runs steps [0, T): masked before lo and from hi on, unmasked between.
template <class Run>
__device__ __forceinline__ void sweep_steps(int T, int lo, int hi, Run &run) {
    int t = 0;
    for (; t < lo; ++t) run(std::true_type{}, t);
    for (; t < hi; ++t) run(std::false_type{}, t);
    for (; t < T; ++t) run(std::true_type{}, t);
}

struct WalkOut {
    int start, end, len;
};

// backtrack_cpu (src/smithWaterman.cu:263-335) over the packed direction
// codes: emit the current cell while its pointer leads elsewhere; a pointer
// into row 0 or column 0 emits the current cell and stops; a start pointer
// stops without emitting.  Each emitted cell is handed to `emit(k, op, i, j)`
// in walk order (the reference stores the strings reversed, i.e. in this
// order) with op 0 = diagonal (both bases), 1 = up (reference gap), 2 = left
// (query gap).
template <class Reader = TraceReader, class Emit>
__device__ __forceinline__ WalkOut walk_cells(const uint32_t *trace, int score, int row, int col,
                                              const PairTrace &pt, int qlen, int rlen, Emit &emit) {
    int start = 0, nrb = -1, tick = 0;
    if (score > 0 && row >= 1 && row <= qlen && col >= 1 && col <= rlen) {
        int ci = row, cj = col, ni, nj;
        Reader rd(reinterpret_cast<const uint4 *>(trace), pt);
        auto back = [&](int i, int j, int &oi, int &oj) {
            int c = rd.code(i, j);
            oi = (c & 2) ? i - 1 : i;            // diag (3) and up (2) move up
            oj = (c == 3 || c == 1) ? j - 1 : j;  // diag (3) and left (1) move left
        };
        back(ci, cj, ni, nj);
        for (;;) {
            if (ci == ni && cj == nj) break;
            const int op = nj == cj ? 1 : ni == ci ? 2 : 0;
            if (op != 1) ++nrb;
            emit(tick, op, ci, cj);
            start = cj;
            ++tick;
            if (ni == 0 || nj == 0) break;  // the target is in row/column 0
            ci = ni;
            cj = nj;
            back(ci, cj, ni, nj);
        }
    }
    return WalkOut{start, start + nrb, tick};
}

// Walk emitting the consensus strings directly (fused mode), NUL terminated.
__device__ WalkOut walk_pair(const uint32_t *trace, int score, int row, int col,
                             const PairTrace &pt, const char *qs, const char *rs, int qlen,
                             int rlen, char *qcons, char *rcons) {
    ConsWriter qc(qcons), rc(rcons);
    auto emit = [&](int, int op, int i, int j) {
        rc.put(op == 1 ? '-' : rs[j - 1]);
        qc.put(op == 2 ? '-' : qs[i - 1]);
    };
    WalkOut wo = walk_cells(trace, score, row, col, pt, qlen, rlen, emit);
    qc.finish();
    rc.finish();
    return wo;
}

// Makes this wave's earlier global stores (trace words) visible to its own
// later loads from other lanes.
__device__ __forceinline__ void wave_store_fence() {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

// ---------------------------------------------------------------------------
// Plain per-pair kernel: the alignSequences drop-in.  One wave per pair
// sweeps, G = 64, byte semantics, int32; the traceback follows in
// sw_plain_walk_kernel and sw_plain_expand_kernel.  Inputs use the reference's layout:
// pair p's query at querySeq + p*(qlen+1) (all qlen bytes are rows, as in
// smithWaterman_cpu), reference at refSeq + p*(rlen+1), columns stopping at
// its first NUL.  LDS holds the reference codes of the pair.
// ---------------------------------------------------------------------------
constexpr int kPlainPad = 64;  // LDS words before column 0 of the plain kernel's reference

template <int R>
__global__ void __launch_bounds__(64)
sw_plain_kernel(const char *__restrict__ qseq, const char *__restrict__ rseq, int qlen, int rlen_max,
                KeyConsts kc, int npairs, int pair0, uint32_t *__restrict__ trace,
                int64_t trace_words_per_pair, int4 *__restrict__ res) {
    using S = Sweep<64, R, false, kBytes>;
    extern __shared__ uint32_t lds_rr[];
    const int p = blockIdx.x;
    if (p >= npairs) return;
    const int g = threadIdx.x;
    const char *qs = qseq + (int64_t)(pair0 + p) * (qlen + 1);
    const char *rs = rseq + (int64_t)(pair0 + p) * (rlen_max + 1);

    // the column count stops at the first NUL
    int first_nul = rlen_max;
    for (int j = g; j < rlen_max; j += 64)
        if (rs[j] == 0) first_nul = min(first_nul, j);
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) first_nul = min(first_nul, __shfl_xor(first_nul, s, 64));
    const int rlen = first_nul;
    // reference codes into LDS at kPlainPad + j, zero for every column a
    // lane reaches outside [0, rlen) (j from -63 to rlen + 62), so each step
    // reads its column without a branch
    for (int x = g; x < rlen_max + 2 * kPlainPad; x += 64) {
        const int j = x - kPlainPad;
        lds_rr[x] = j >= 0 && j < rlen ? byte_rcode((unsigned char)rs[j]) : 0u;
    }
    __syncthreads();

    S sw;
    sw.init(kc);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int row = g * R + r;
        sw.q[r] = row < qlen ? byte_qcode((unsigned char)qs[row]) : 0u;
    }
    // the pair's trace through a buffer resource: 32-bit byte offsets
    // (a pair's trace is under 2^21 words)
    const __amdgpu_buffer_rsrc_t tr_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        trace + (int64_t)p * trace_words_per_pair, 0, (int)0x7FFFFFFF, 0x00020000);
    const int T = rlen + 63;
    const int main_lo = 63, main_hi = rlen;  // every lane inside [0, rlen)
    auto run = [&](auto masked, int t) {
        int j = t - g;
        uint32_t rr = lds_rr[kPlainPad + j];
        uint32_t vmask = 0;
        if constexpr (decltype(masked)::value) vmask = (j >= 0 && j < rlen) ? ~0u : 0u;
        uint32_t a1[S::W];
#pragma unroll
        for (int w = 0; w < S::W; ++w) a1[w] = 0;
        sw.template step<decltype(masked)::value>(rr, j, vmask, 0u, a1);
        // byte offset of word (t >> 2, w, g, t & 3) of the granule layout
        // (trace_word: 64 lanes x 4 steps = 1 KB per step group and word)
        const uint32_t o = (uint32_t)(t >> 2) * (S::W * 1024u) + (uint32_t)g * 16u +
                           (uint32_t)(t & 3) * 4u;
#pragma unroll
        for (int w = 0; w < S::W; ++w)
            __builtin_amdgcn_raw_buffer_store_b32(a1[w], tr_rsrc, o + (uint32_t)w * 1024u, 0, 0);
    };
    sweep_steps(T, min(main_lo, T), main_hi, run);

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
    const GroupResult gr = group_best<64>(bh, brow, bcol, R, false, 0, 0);
    // the walk runs in sw_plain_walk_kernel, one lane per pair
    if (g == 0) res[p] = make_int4(gr.score, gr.row, gr.col, rlen);
}

// ---------------------------------------------------------------------------
// Batched persistent kernels.
// ---------------------------------------------------------------------------
// LDS words per lane group of G lanes: G zero columns before column 0, the
// reference columns, and G columns of tail
template <int G>
constexpr int group_lds_words() { return G + kBatchCols + G; }

// One wave task: 64 / G groups of G lanes (G = 16: 4 groups; G = 64, the
// wide form: one); PK -> 2 pairs per group (halves).
template <int G, int R, bool PK, int POL>
__device__ __forceinline__ void run_task(const BatchArgs &a, const Task &tk, uint32_t *lds) {
    using S = Sweep<G, R, PK, POL>;
    constexpr int PPG = PK ? 2 : 1;  // pairs per group
    constexpr int NG = 64 / G;       // groups per wave
    constexpr int PAD = G;           // zero columns before column 0
    constexpr int GL = group_lds_words<G>();
    const int lane = threadIdx.x & 63;
    const int grp = lane / G, g = lane % G;
    // reference codes at [PAD + j]; the PAD leading entries and the tail
    // are read by lanes outside the pair's columns (their results unused)
    uint32_t *rr_lds = lds + grp * GL;

    // the group's pairs.  The loads are unconditional (a slot past the
    // task's pairs reads its first pair and is then cleared; byte loads are
    // clamped into the sequence, whose first byte exists for every binned
    // pair) so that each level of the order -> lengths -> bytes chain issues
    // as one batch: one memory round trip per level instead of one per load.
    int pid[PPG], ql[PPG], rl[PPG];
    int64_t qo[PPG], ro[PPG];
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        const int slot = grp * PPG + h;
        pid[h] = a.order[tk.first + (slot < tk.npairs ? slot : 0)];
    }
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        const int p = pid[h];
        ql[h] = a.qlen[p];
        rl[h] = a.rlen[p];
        qo[h] = a.qoff[p];
        ro[h] = a.roff[p];
    }
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        if (grp * PPG + h >= tk.npairs) {
            pid[h] = -1;
            ql[h] = rl[h] = 0;
            qo[h] = ro[h] = 0;
        }
    }
    // the query rows: raw bytes now (in flight during the reference loads
    // below), converted to match tables afterwards
    S sw;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = g * R + r;
        uint32_t v[PPG];
#pragma unroll
        for (int h = 0; h < PPG; ++h) {
            // unconditional load of an in-range byte, then the select: a
            // load under a branch is waited for before the branch joins
            const uint32_t c = (unsigned char)a.qseq[qo[h] + max(0, min(row, ql[h] - 1))];
            v[h] = row < ql[h] ? c : 256u;
        }
        sw.q[r] = v[0];
        if constexpr (PK) sw.q2[r] = v[PPG - 1];
    }
    // reference codes of the group's pairs into LDS (both halves packed),
    // padding up to the last column any lane reaches (ncols + G - 2)
    // and the step-ahead prefetch of it.  Loaded in batches of NB columns
    // per lane, each batch's loads issued together.
    const int ncols = tk.steps - (G - 1);
    constexpr int NB = 8;
    uint32_t exotic = 0;  // bit h: the pair's reference holds a byte outside ACGTNacgtn
    for (int j0 = g; j0 < ncols + G; j0 += NB * G) {
        unsigned char cb[NB][PPG];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int j = j0 + i * G;
#pragma unroll
            for (int h = 0; h < PPG; ++h)
                cb[i][h] = (unsigned char)a.rseq[ro[h] + max(0, min(j, rl[h] - 1))];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int j = j0 + i * G;
            if (j < ncols + G) {
                uint32_t v = 0;
#pragma unroll
                for (int h = 0; h < PPG; ++h) {
                    uint32_t code = PK ? kRefPad : 0u;
                    if (j < rl[h]) {
                        code = PK ? ref_index_bf(cb[i][h]) : byte_rcode(cb[i][h]);
                        if (PK && code == kRefOther) exotic |= 1u << h;
                    }
                    if constexpr (PK) code = column_selector(code, h);
                    v |= code << (16 * h);
                }
                rr_lds[PAD + j] = v;
            }
        }
    }
    if constexpr (PK) {
#pragma unroll
        for (int s2 = G / 2; s2 >= 1; s2 >>= 1) exotic |= __shfl_xor(exotic, s2, G);
    }
    if (g < PAD)
        rr_lds[g] = PK ? column_selector(kRefPad, 0) | column_selector(kRefPad, 1) << 16 : 0u;
    sw.init(a.kc);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        // 256: a row past the query (no match anywhere)
        auto conv = [](uint32_t c) -> uint32_t {
            return c > 255u ? 0u : PK ? query_table(query_mask_bf((unsigned char)c))
                                      : byte_qcode((unsigned char)c);
        };
        sw.q[r] = conv(sw.q[r]);
        if constexpr (PK) sw.q2[r] = conv(sw.q2[r]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);  // LDS writes of this wave visible to its lanes
    __builtin_amdgcn_wave_barrier();

    uint32_t *tr = a.trace + tk.trace_base;
    uint32_t rr_next = rr_lds[PAD - g];  // column of step 0
    auto run = [&](auto masked, int t) {
        int j = t - g;
        uint32_t rr = rr_next;
        // prefetch the next step's column (clamped: steps past the last
        // granule boundary read a stale word whose results are unused)
        rr_next = rr_lds[min(PAD + 1 + j, GL - 1)];
        uint32_t vmask = 0;
        if constexpr (decltype(masked)::value) {
#pragma unroll
            for (int h = 0; h < PPG; ++h)
                if (j >= 0 && j < rl[h]) vmask |= (PK ? 0xFFFFu : 0xFFFFFFFFu) << (16 * h);
        }
        uint32_t a1[S::W];
#pragma unroll
        for (int w = 0; w < S::W; ++w) a1[w] = 0;
        sw.template step<decltype(masked)::value>(rr, j, vmask, splat<PK>(15 - (t & 15)), a1);
#pragma unroll
        for (int w = 0; w < S::W; ++w) tr[trace_word<S::W>(t, w, lane)] = a1[w];
    };
    // int32: masked steps keep the cells outside a pair's columns out of its
    // row maxima.  Packed: no step is masked (see Sweep).
    const int T = tk.steps;
    if constexpr (PK) {
        for (int t = 0; t < T; ++t) {
            run(std::false_type{}, t);
            if ((t & 15) == 15 || t == T - 1) sw.block_end(splat<true>((t >> 4) << 5));
        }
    } else {
        const int lo = min(G - 1, T), hi = max(lo, tk.min_rlen);
        sweep_steps(T, lo, hi, run);
    }

    // per half: lane best over valid rows, then group max
    GroupResult gres[PPG];
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        uint32_t bh = 0, brow = 0, bcol = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int row = g * R + r;
            uint32_t hv, cv;
            if constexpr (PK) {
                // block << 5 | (15 - t%16) -> step t -> column t - g (+1)
                const uint32_t bc = (sw.bcol[r] >> (16 * h)) & 0xFFFFu;
                hv = ((sw.best[r] >> (16 * h)) & 0xFFFFu) >> 5;
                cv = ((bc >> 5) << 4) + 15u - (bc & 15u) - (uint32_t)g + 1;
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
        gres[h] = group_best<G>(bh, brow, bcol, R, PK, a.h_limit, (exotic >> h) & 1);
        if (g == 0 && pid[h] >= 0) {
            a.res[pid[h]] = PairResult{gres[h].score, gres[h].row, gres[h].col, gres[h].flag};
            if (gres[h].flag) atomicAdd(a.nflagged, 1);
        }
    }

    const int s = lane;
    const int sg = min(s / PPG, NG - 1), sh = s % PPG;
    if (a.walk_deferred) {
        // record where the pair's trace lives; sw_walk_kernel walks it
        if (s < tk.npairs)
            a.ptr[a.order[tk.first + s]] = PairTrace{tk.trace_base, sg * G, (int16_t)R,
                                                     (int8_t)S::W, (int8_t)(sh | (PK ? 2 : 0))};
        __builtin_amdgcn_wave_barrier();
        return;
    }
    // Traceback of the task's pairs right after the sweep, while the trace
    // is cache-hot: lane s walks slot s.  A flagged pair is left to the int32
    // re-run.  The walk is latency-bound; other waves' sweeps cover it.
    wave_store_fence();
    int w_score = 0, w_row = 0, w_col = 0, w_flag = 1;
#pragma unroll
    for (int h = 0; h < PPG; ++h) {
        int v0 = __shfl(gres[h].score, sg * G, 64);
        int v1 = __shfl(gres[h].row, sg * G, 64);
        int v2 = __shfl(gres[h].col, sg * G, 64);
        int v3 = __shfl(gres[h].flag, sg * G, 64);
        if (h == sh) {
            w_score = v0;
            w_row = v1;
            w_col = v2;
            w_flag = v3;
        }
    }
    if (s < tk.npairs && !w_flag) {
        const int p = a.order[tk.first + s];
        const PairTrace pt{tk.trace_base, sg * G, (int16_t)R, (uint8_t)S::W,
                           (int8_t)(sh | (PK ? 2 : 0))};
        const int64_t co = a.coff[p];
        WalkOut wo = walk_pair(a.trace, w_score, w_row, w_col, pt, a.qseq + a.qoff[p],
                               a.rseq + a.roff[p], a.qlen[p], a.rlen[p], a.qcons + co,
                               a.rcons + co);
        a.score_out[p] = w_score;
        a.start_out[p] = wo.start;
        a.end_out[p] = wo.end;
        a.alen_out[p] = wo.len;
    }
    __builtin_amdgcn_wave_barrier();
}

// Orders one wave's LDS stores before its later LDS loads from other lanes
// (each wave of sw_expand_kernel works alone on its own LDS region, so no
// workgroup barrier is involved).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Deferred walk-back, in two kernels over the binned pairs after every sweep
// of the call (both passes) has finished.  sw_walk_kernel: one lane per pair
// follows the trace and stores the path as 2-bit moves (16 per word, walk
// order) plus start/end/length; it touches no sequence bytes, so its
// dependent loads are the trace granules alone.  sw_expand_kernel turns the
// moves into the consensus strings, four pairs per wave (below).  A pair
// still flagged for the int32 re-run (none after the re-run pass) is
// skipped.
__global__ void __launch_bounds__(256) sw_walk_kernel(BatchArgs a, const int32_t *order,
                                                      const int32_t *first_p,
                                                      const int32_t *count, uint32_t *moves,
                                                      int move_words) {
    const int n = *count;
    const int32_t *ord = order + (first_p ? *first_p : 0);
    for (int i = blockIdx.x * 256 + (int)threadIdx.x; i < n; i += gridDim.x * 256) {
        const int p = ord[i];
        const PairResult r = a.res[p];
        if (r.flags & 1) continue;
        const PairTrace pt = a.ptr[p];
        uint32_t *mv = moves + (int64_t)p * move_words;
        uint32_t word = 0;
        auto emit = [&](int k, int op, int, int) {
            word |= (uint32_t)op << (2 * (k & 15));
            if ((k & 15) == 15) {
                mv[k >> 4] = word;
                word = 0;
            }
        };
        WalkOut wo = walk_cells<TraceReaderLean>(a.trace, r.score, r.row, r.col, pt, a.qlen[p],
                                                 a.rlen[p], emit);
        if (wo.len & 15) mv[wo.len >> 4] = word;
        a.score_out[p] = r.score;
        a.start_out[p] = wo.start;
        a.end_out[p] = wo.end;
        a.alen_out[p] = wo.len;
    }
}

// Deferred walk of one pair by a whole wave, for calls too small to fill
// the chip (a lane per pair leaves each wave waiting on one dependent
// granule load after another, and every lane's wait is the wave's).  The
// walk's state is wave-uniform; the trace around it is loaded as a window
// of granules, one per lane -- lane groups (g0 - NG, g0], every trace word,
// step blocks (tb0 - NT, tb0].  Lane l reads the code of the cell l steps
// down the diagonal from the window lane holding its granule (ds_bpermute);
// a ballot gives the length of the leading diagonal run, which the walk
// takes at once, then the move off it.  The walk only moves up and left, so
// it stays in a window until its lane group or step block passes the
// window's lower edge; a new window is then loaded at the current cell, one
// load per lane.  Same path, moves and results as walk_cells.  tr: the
// pair's lane group at step 0, word 0.
template <int NMV>
__device__ __forceinline__ WalkOut walk_wave(const uint4 *tr, int R, int W, int kl, int bitb,
                                             int score, int row, int col, int qlen, int rlen,
                                             uint32_t *mv) {
    const int lane = threadIdx.x & 63;
    int start = 0, nrb = -1, tick = 0;
    if (!(score > 0 && row >= 1 && row <= qlen && col >= 1 && col <= rlen))
        return WalkOut{start, start + nrb, tick};
    const uint32_t inv = (65536u + (uint32_t)R - 1) / (uint32_t)R;  // x / R for x < 2048
    // window shape: about 24 rows of lane groups, the rest of the 64 lanes
    // in step blocks; lanes past NG x W x NT load a copy of a window granule
    const int NG = min(8, (24 + R - 1) / R + 1);
    const int WT = 64 / NG, NT = WT / W;
    const int dg = min(lane / WT, NG - 1), r0 = lane - (lane / WT) * WT;
    const int lw = min(r0 / NT, W - 1), dtb = r0 - (r0 / NT) * NT;
    const int kmask = (1 << kl) - 1;
    int g0 = -1, tb0 = 0;
    uint4 v = make_uint4(0, 0, 0, 0);
    // the moves, 16 per word, word m * 64 + lane in mvr[m] (stored at the end:
    // a store inside the walk would be waited for with the next window load)
    uint32_t mvr[NMV];
#pragma unroll
    for (int m = 0; m < NMV; ++m) mvr[m] = 0;
    uint32_t mword = 0;
    auto flush = [&](int word_index, uint32_t value) {
#pragma unroll
        for (int m = 0; m < NMV; ++m)
            if (word_index == m * 64 + lane) mvr[m] = value;
    };
    int ci = row, cj = col;
    for (;;) {
        const int g = (int)(((uint32_t)(ci - 1) * inv) >> 16);
        const int tb = ((cj - 1) + g) >> 2;
        if (g0 < 0 || g0 - g >= NG || tb0 - tb >= NT) {
            g0 = g;
            tb0 = tb;
            v = tr[(max(tb0 - dtb, 0) * W + lw) * 64 + max(g0 - dg, 0)];
        }
        // Lane l reads the code of cell (ci - l, cj - l), l steps down the
        // diagonal, from the window lane holding its granule; the walk then
        // takes every leading diagonal move at once.
        const int il = ci - lane, jl = cj - lane;
        const int gl = (int)(__umul24((uint32_t)max(il - 1, 0), inv) >> 16);
        const int kl_ = (il - 1) - gl * R, tl = (jl - 1) + gl;
        const bool in_win = il >= 1 && jl >= 1 && g0 - gl < NG && tb0 - (tl >> 2) < NT;
        const int L = in_win ? (g0 - gl) * WT + (kl_ >> kl) * NT + (tb0 - (tl >> 2)) : 0;
        const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute(L << 2, (int)v.x);
        const uint32_t y = (uint32_t)__builtin_amdgcn_ds_bpermute(L << 2, (int)v.y);
        const uint32_t z = (uint32_t)__builtin_amdgcn_ds_bpermute(L << 2, (int)v.z);
        const uint32_t u = (uint32_t)__builtin_amdgcn_ds_bpermute(L << 2, (int)v.w);
        const uint32_t word = (tl & 2) ? ((tl & 1) ? u : z) : ((tl & 1) ? y : x);
        const int cl = in_win ? (int)((word >> (bitb + 2 * (kl_ & kmask))) & 3u) : -1;
        const uint64_t stop = __ballot(cl != 3);
        // the first cell off the diagonal run (or outside the window)
        const int first = stop ? (int)__builtin_ctzll(stop) : 64;
        if (first > 0) {
            // `first` diagonal moves (op 0: no bits to set); the move words
            // they complete are flushed as the count passes them
            nrb += first;
            start = cj - first + 1;
            ci -= first;
            cj -= first;
            int rest = first;
            while ((tick & 15) + rest >= 16) {
                const int adv = 16 - (tick & 15);
                flush(tick >> 4, mword);
                mword = 0;
                tick += adv;
                rest -= adv;
            }
            tick += rest;
            if (ci == 0 || cj == 0) break;  // the target is in row/column 0
        }
        if (first >= 64) continue;
        // that cell is now the current one; outside the window, a window is
        // loaded around it first
        if (!__builtin_amdgcn_readlane((int)in_win, first)) continue;
        const int c = __builtin_amdgcn_readlane(cl, first);
        if (c == 0) break;  // a start pointer: stop without emitting
        // up (2) or left (1)
        const int di = c >> 1, dj = c & 1;
        const int op = dj == 0 ? 1 : 2;
        nrb += dj;
        mword |= (uint32_t)op << (2 * (tick & 15));
        if ((tick & 15) == 15) {
            flush(tick >> 4, mword);
            mword = 0;
        }
        start = cj;
        ++tick;
        ci -= di;
        cj -= dj;
        if (ci == 0 || cj == 0) break;  // the target is in row/column 0
    }
    if (tick & 15) {
#pragma unroll
        for (int m = 0; m < NMV; ++m)
            if ((tick >> 4) == m * 64 + lane) mvr[m] = mword;
    }
    const int nwords = (tick + 15) >> 4;
#pragma unroll
    for (int m = 0; m < NMV; ++m)
        if (m * 64 + lane < nwords) mv[m * 64 + lane] = mvr[m];
    return WalkOut{start, start + nrb, tick};
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// sw_walk_kernel's work, one wave per pair (small batches); a path's moves
// fit one word per lane
static_assert((kMaxBatchQ + kMaxBatchR) / 16 + 1 <= 64, "moves of a batched pair");
__global__ void __launch_bounds__(256) sw_walk_wave_kernel(BatchArgs a, const int32_t *order,
                                                           const int32_t *first_p,
                                                           const int32_t *count, uint32_t *moves,
                                                           int move_words) {
    const int n = uni(*count);
    const int32_t *ord = order + (first_p ? uni(*first_p) : 0);
    const int nw = gridDim.x * 4;
    for (int i = uni((int)(blockIdx.x * 4 + (threadIdx.x >> 6))); i < n; i += nw) {
        const int p = uni(ord[i]);
        const PairResult r = a.res[p];
        if (uni(r.flags) & 1) continue;
        const PairTrace pt = a.ptr[p];
        const int half = uni(pt.half), R = uni(pt.rows), W = uni(pt.words);
        const int64_t base = ((int64_t)uni((int)(pt.base >> 32)) << 32) |
                             (uint32_t)uni((int)(uint32_t)pt.base);
        const uint4 *tr = reinterpret_cast<const uint4 *>(a.trace) + base / 4 + uni(pt.lane0);
        const int score = uni(r.score);
        WalkOut wo = walk_wave<1>(tr, R, W, (half & 2) ? 3 : 4, (half & 1) * 16, score, uni(r.row),
                               uni(r.col), uni(a.qlen[p]), uni(a.rlen[p]),
                               moves + (int64_t)p * move_words);
        if ((threadIdx.x & 63) == 0) {
            a.score_out[p] = score;
            a.start_out[p] = wo.start;
            a.end_out[p] = wo.end;
            a.alen_out[p] = wo.len;
        }
    }
}

// One pair's consensus expansion, by one 16-lane row of a wave (lane sl of
// the row).  A pair's moves are taken in chunks of 256 (16 per lane); a
// prefix sum of the row and column steps over the row's 16 lanes gives each
// lane the cell its 16 moves start from, and the lane builds its 16 bytes of
// each string in registers from the pair's bases staged in LDS, then stores
// them as dwords where the strings are dword aligned (bytes at the string's
// end, then the NUL).  live = false: the row only takes part in the wave's
// LDS ordering.
struct ExpandPair {
    bool live = false;
    int len = 0, row = 0, col = 0;   // moves, best cell (1-based)
    int nq = 0, nr = 0;              // bases the path can reach
    const unsigned char *qsrc = nullptr, *rsrc = nullptr;
    char *qc = nullptr, *rc = nullptr;
    const uint32_t *mv = nullptr;    // 2-bit moves, 16 per word, walk order
};

__device__ __forceinline__ void expand_row(const ExpandPair &e, uint32_t *lq4, uint32_t *lr4,
                                           int sl) {
    // the bases into LDS, from the aligned dwords holding them (an aligned
    // dword holding a sequence byte lies within that byte's page)
    const int qm = (int)((uintptr_t)e.qsrc & 3), rm = (int)((uintptr_t)e.rsrc & 3);
    const int nqw = e.nq > 0 ? (qm + e.nq + 3) >> 2 : 0;
    const int nrw = e.nr > 0 ? (rm + e.nr + 3) >> 2 : 0;
    for (int u = sl; u < nqw; u += 16)
        lq4[u] = reinterpret_cast<const uint32_t *>(e.qsrc - qm)[u];
    for (int u = sl; u < nrw; u += 16)
        lr4[u] = reinterpret_cast<const uint32_t *>(e.rsrc - rm)[u];
    wave_lds_sync();
    const unsigned char *lq = (const unsigned char *)lq4 + qm;
    const unsigned char *lr = (const unsigned char *)lr4 + rm;
    char *qc = e.qc, *rc = e.rc;
    const bool aligned = (((uintptr_t)qc | (uintptr_t)rc) & 3) == 0;
    const int len = e.len;
    int ci = e.row, cj = e.col;
    const int nchunk = (len + 255) >> 8;
    for (int c = 0; c < nchunk; ++c) {
        const int k0 = c * 256 + sl * 16;
        const int cnt = max(0, min(16, len - k0));
        const uint32_t w = cnt > 0 ? e.mv[k0 >> 4] : 0u;
        // row steps: every move but left (2); column steps: every move but
        // up (1); ops are 0, 1 or 2
        const uint32_t lo2 = cnt == 16 ? 0x55555555u : ((1u << (2 * cnt)) - 1u) & 0x55555555u;
        const uint32_t hib = (w >> 1) & lo2, lob = w & lo2;
        const int di = cnt - __builtin_popcount(hib & ~lob);
        const int dj = cnt - __builtin_popcount(lob & ~hib);
        int ei = di, ej = dj;  // inclusive scans over the row's 16 lanes
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int oi = __shfl_up(ei, d, 16), oj = __shfl_up(ej, d, 16);
            if (sl >= d) {
                ei += oi;
                ej += oj;
            }
        }
        int xi = ci - (ei - di), xj = cj - (ej - dj);
        uint32_t oq[4] = {0u, 0u, 0u, 0u}, orr[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k < cnt) {
                const int op = (w >> (2 * k)) & 3;
                const uint32_t qb = op == 2 ? (uint32_t)'-' : (uint32_t)lq[xi - 1];
                const uint32_t rb = op == 1 ? (uint32_t)'-' : (uint32_t)lr[xj - 1];
                oq[k >> 2] |= qb << (8 * (k & 3));
                orr[k >> 2] |= rb << (8 * (k & 3));
                xi -= op != 2;
                xj -= op != 1;
            }
        }
        if (cnt == 16 && aligned) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                reinterpret_cast<uint32_t *>(qc + k0)[d] = oq[d];
                reinterpret_cast<uint32_t *>(rc + k0)[d] = orr[d];
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                qc[k0 + k] = (char)(oq[k >> 2] >> (8 * (k & 3)));
                rc[k0 + k] = (char)(orr[k >> 2] >> (8 * (k & 3)));
            }
        }
        ci -= __shfl(ei, 15, 16);
        cj -= __shfl(ej, 15, 16);
    }
    if (e.live && sl == 0) {
        qc[len] = 0;
        rc[len] = 0;
    }
    wave_lds_sync();  // before the next pairs overwrite this row's buffers
}

// sw_expand_kernel: the batched pass's consensus strings, four pairs per
// wave (one per 16-lane row), 16 pairs per workgroup.
__global__ void __launch_bounds__(256) sw_expand_kernel(BatchArgs a, const int32_t *order,
                                                        const int32_t *first_p,
                                                        const int32_t *count,
                                                        const uint32_t *moves, int move_words) {
    constexpr int kQW = (kMaxBatchQ + 6) / 4, kRW = (kMaxBatchR + 6) / 4;  // dwords staged
    constexpr int kPairs = 16;                                           // per workgroup
    __shared__ uint32_t lq4[kPairs][kQW], lr4[kPairs][kRW];
    const int n = *count;
    const int slot = threadIdx.x >> 4, sl = threadIdx.x & 15;
    const int32_t *ord = order + (first_p ? *first_p : 0);
    for (int i0 = blockIdx.x * kPairs; i0 < n; i0 += gridDim.x * kPairs) {
        const int i = i0 + slot;
        ExpandPair e;
        if (i < n) {
            const int p = ord[i];
            const PairResult r = a.res[p];
            if (!(r.flags & 1)) {  // a flagged pair is left to the int32 pass
                e.live = true;
                e.len = a.alen_out[p];
                e.row = r.row;
                e.col = r.col;
                // the path only visits rows <= row and columns <= col
                e.nq = min(a.qlen[p], r.row);
                e.nr = min(a.rlen[p], r.col);
                const int64_t co = a.coff[p];
                e.qsrc = (const unsigned char *)a.qseq + a.qoff[p];
                e.rsrc = (const unsigned char *)a.rseq + a.roff[p];
                e.qc = a.qcons + co;
                e.rc = a.rcons + co;
                e.mv = moves + (int64_t)p * move_words;
            }
        }
        expand_row(e, lq4[slot], lr4[slot], sl);
    }
}

// Plain path, after sw_plain_kernel: one lane per pair walks the pair's
// trace from its best cell (res: score, row, col, columns) and stores the
// path as 2-bit moves plus the drop-in's score, start, end and length.
__global__ void __launch_bounds__(256) sw_plain_walk_kernel(const uint32_t *trace, int64_t wpp,
                                                            int R, int W, int qlen, int npairs,
                                                            int pair0, const int4 *res,
                                                            uint32_t *moves, int move_words,
                                                            PlainOut out) {
    const int p = blockIdx.x * 256 + (int)threadIdx.x;
    if (p >= npairs) return;
    const int4 r = res[p];
    uint32_t *mv = moves + (int64_t)p * move_words;
    uint32_t word = 0;
    auto emit = [&](int k, int op, int, int) {
        word |= (uint32_t)op << (2 * (k & 15));
        if ((k & 15) == 15) {
            mv[k >> 4] = word;
            word = 0;
        }
    };
    const PairTrace pt{(int64_t)p * wpp, 0, (int16_t)R, (int8_t)W, 0};
    WalkOut wo = walk_cells<TraceReaderLean>(trace, r.x, r.y, r.z, pt, qlen, r.w, emit);
    if (wo.len & 15) mv[wo.len >> 4] = word;
    const int gp = pair0 + p;
    out.score[gp] = (float)r.x;
    out.start[gp] = wo.start;
    out.end[gp] = wo.end;
    out.alen[gp] = wo.len;
}

// sw_plain_walk_kernel's work, one wave per pair (small calls); NMV: move
// words per lane (a path's moves fit NMV * 64 words)
template <int NMV>
__global__ void __launch_bounds__(256)
sw_plain_walk_wave_kernel(const uint32_t *trace, int64_t wpp, int R, int W, int qlen, int npairs,
                          int pair0, const int4 *res, uint32_t *moves, int move_words,
                          PlainOut out) {
    const int p = uni((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (p >= npairs) return;
    const int4 r = res[p];
    const int score = uni(r.x);
    const uint4 *tr = reinterpret_cast<const uint4 *>(trace) + (int64_t)p * wpp / 4;
    WalkOut wo = walk_wave<NMV>(tr, R, W, 4, 0, score, uni(r.y), uni(r.z), qlen, uni(r.w),
                                moves + (int64_t)p * move_words);
    if ((threadIdx.x & 63) == 0) {
        const int gp = pair0 + p;
        out.score[gp] = (float)score;
        out.start[gp] = wo.start;
        out.end[gp] = wo.end;
        out.alen[gp] = wo.len;
    }
}

// Plain path: the consensus strings from the walks' moves, four pairs per
// one-wave workgroup (dynamic LDS: 4 x (qw + rw) dwords).
__global__ void __launch_bounds__(64) sw_plain_expand_kernel(const char *qseq, const char *rseq,
                                                             int qlen, int rlen_max, int npairs,
                                                             int pair0, const int4 *res,
                                                             const uint32_t *moves,
                                                             int move_words, int qw, int rw,
                                                             PlainOut out) {
    extern __shared__ uint32_t lds_x[];
    const int slot = threadIdx.x >> 4, sl = threadIdx.x & 15;
    const int p = blockIdx.x * 4 + slot;
    ExpandPair e;
    if (p < npairs) {
        const int gp = pair0 + p;
        const int4 r = res[p];
        e.live = true;
        e.len = out.alen[gp];
        e.row = r.y;
        e.col = r.z;
        e.nq = min(qlen, r.y);
        e.nr = min(r.w, r.z);
        e.qsrc = (const unsigned char *)qseq + (int64_t)gp * (qlen + 1);
        e.rsrc = (const unsigned char *)rseq + (int64_t)gp * (rlen_max + 1);
        e.qc = out.qcons + (int64_t)gp * out.cstride;
        e.rc = out.rcons + (int64_t)gp * out.cstride;
        e.mv = moves + (int64_t)p * move_words;
    }
    expand_row(e, lds_x + slot * (qw + rw), lds_x + slot * (qw + rw) + qw, sl);
}

// Row classes: R per lane with 16-lane groups covers query length 16*R.
// One kernel per R: each gets the register budget its R needs (a single
// kernel switching over all R is allocated for the worst case).
#define SWIFT_R16_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) \
    X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)
#define SWIFT_R32_CASES(X) X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20)

// Occupancy floor per row class (waves per SIMD), set where the natural
// register count sits just above an occupancy step.
__host__ __device__ constexpr int batch_waves_per_eu(int R, bool pk) {
    return pk ? (R <= 7 ? 5 : R <= 12 ? 4 : R <= 16 ? 3 : 2) : (R <= 10 ? 5 : R <= 16 ? 4 : 3);
}

template <int R, bool PK, int POL>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(batch_waves_per_eu(R, PK))))
sw_batch_kernel(BatchArgs a) {
    __shared__ uint32_t lds[4][4 * group_lds_words<kBatchG>()];
    const int wave = threadIdx.x >> 6;
    const int lo = *a.task_lo, hi = *a.task_hi;
    if (lo >= hi) return;
    for (;;) {
        int tid = 0;
        if ((threadIdx.x & 63) == 0) tid = atomicAdd(a.next_task, 1);
        tid = __builtin_amdgcn_readfirstlane(__shfl(tid, 0, 64)) + lo;
        if (tid >= hi) break;
        const Task tk = a.tasks[tid];
        run_task<kBatchG, R, PK, POL>(a, tk, lds[wave]);
    }
}

// The wide form for batches too small to fill the chip: one 64-lane group
// per wave (the lane shift a wave_shr), kWideRows rows per lane, two packed
// pairs per task.  A 16-lane group sweeps R = 20 rows per lane over rlen + 15
// steps; a 64-lane group 5 rows over rlen + 63, so one task's latency, which
// is all a small batch waits for, drops about 3.5x for the same total work.
template <int R>
__global__ void __launch_bounds__(256) sw_batch_wide_kernel(BatchArgs a) {
    __shared__ uint32_t lds[4][group_lds_words<kWideG>()];
    const int wave = threadIdx.x >> 6;
    const int lo = *a.task_lo, hi = *a.task_hi;
    if (lo >= hi) return;
    for (;;) {
        int tid = 0;
        if ((threadIdx.x & 63) == 0) tid = atomicAdd(a.next_task, 1);
        tid = __builtin_amdgcn_readfirstlane(__shfl(tid, 0, 64)) + lo;
        if (tid >= hi) break;
        const Task tk = a.tasks[tid];
        run_task<kWideG, R, true, kOneHot>(a, tk, lds[wave]);
    }
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
// pair -> bin (or -1 when the pair takes no part in this pass)
__device__ __forceinline__ int classify_pair(const BinArgs &a, int p) {
    int ql = a.qlen[p], rl = a.rlen[p];
    if (ql < 0 || rl < 0 || ql > kMaxBatchQ || rl > kMaxBatchR) {
        // lengths beyond the batched kernels: the pair is left out (score 0)
        // and the call reports the error
        atomicOr(a.error, 1);
        return -1;
    }
    if (ql == 0 || rl == 0) {
        // all-zero matrix: start 0, end -1, length 0 (src/smithWaterman.cu
        // :50,333-334)
        if (!a.prev) {
            a.score_out[p] = 0;
            a.start_out[p] = 0;
            a.end_out[p] = -1;
            a.alen_out[p] = 0;
        }
        return -1;
    }
    int cls;
    if (a.prev) {
        if (!(a.prev[p].flags & 1)) return -1;
        cls = 1;
    } else {
        // references with bytes outside ACGTNacgtn are caught by the int16
        // kernel itself and re-run with byte semantics in the int32 pass
        cls = a.force32 ? 1 : 0;
    }
    int R = max(1, (ql + kBatchG - 1) / kBatchG);
    // coarse classes (small batches): fewer class kernels, each a pair's rows
    // spread over fewer lanes of its group; the rows past the query are padding
    if (a.coarse > 0) R = min(kRBins, (R + a.coarse - 1) / a.coarse * a.coarse);
    if (cls == 1) R = r32_class(R);
    int lb = min(kLenBins - 1, rl >> 3);
    // bin order: class, then R descending, then rlen descending
    return (cls * kRBins + (kRBins - R)) * kLenBins + (kLenBins - 1 - lb);
}

// Pairs [blockIdx.x * kBinChunk, +kBinChunk) are counted in an LDS
// histogram first, so a bin costs one global atomic per workgroup however
// many pairs share it (uniform reads put every pair in one bin).
constexpr int kBinPer = 8;
constexpr int kBinChunk = 256 * kBinPer;

__global__ void __launch_bounds__(256) sw_bin_classify(BinArgs a) {
    __shared__ int32_t hist[kNumBins];
    for (int b = threadIdx.x; b < kNumBins; b += 256) hist[b] = 0;
    __syncthreads();
    const int base = blockIdx.x * kBinChunk;
#pragma unroll
    for (int i = 0; i < kBinPer; ++i) {
        int p = base + i * 256 + (int)threadIdx.x;
        if (p < a.npairs) {
            int b = classify_pair(a, p);
            a.bin_of[p] = b;
            if (b >= 0) atomicAdd(&hist[b], 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kNumBins; b += 256)
        if (hist[b]) atomicAdd(&a.bin_count[b], hist[b]);
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
    // segment (class, R) lengths in parallel, then their task offsets
    __shared__ int32_t seg_tasks[2 * kRBins];
    if (tid < 2 * kRBins) a.seg_steps[tid] = 0;  // summed by sw_build_tasks
    if (tid < 2 * kRBins) {
        const int seg = tid, cls = seg / kRBins;
        int32_t len = 0;
        for (int lb = 0; lb < kLenBins; ++lb) len += a.bin_count[seg * kLenBins + lb];
        const int slots = cls == 0 ? (a.wide && seg == 0 ? kSlotsWide : kSlots16) : kSlots32;
        a.seg_first[seg] = a.bin_cursor[seg * kLenBins];
        a.seg_len[seg] = len;
        seg_tasks[seg] = (len + slots - 1) / slots;
    }
    __syncthreads();
    if (tid == 0) {
        int32_t t0 = 0;
        for (int seg = 0; seg < 2 * kRBins; ++seg) {
            a.seg_task0[seg] = t0;
            t0 += seg_tasks[seg];
        }
        a.seg_task0[2 * kRBins] = t0;
        *a.ntasks = t0;
        *a.npairs_binned = part[1023];
    }
}

// Same chunks as sw_bin_classify: ranks within the chunk from LDS atomics,
// one global cursor reservation per (chunk, bin).
__global__ void __launch_bounds__(256) sw_bin_scatter(const int32_t *bin_of, int npairs,
                                                      int32_t *bin_cursor, int32_t *order) {
    __shared__ int32_t hist[kNumBins];
    for (int b = threadIdx.x; b < kNumBins; b += 256) hist[b] = 0;
    __syncthreads();
    const int base = blockIdx.x * kBinChunk;
    int bin[kBinPer], rank[kBinPer];
#pragma unroll
    for (int i = 0; i < kBinPer; ++i) {
        int p = base + i * 256 + (int)threadIdx.x;
        bin[i] = p < npairs ? bin_of[p] : -1;
        rank[i] = bin[i] >= 0 ? atomicAdd(&hist[bin[i]], 1) : 0;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kNumBins; b += 256)
        if (hist[b]) hist[b] = atomicAdd(&bin_cursor[b], hist[b]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kBinPer; ++i)
        if (bin[i] >= 0) order[hist[bin[i]] + rank[i]] = base + i * 256 + (int)threadIdx.x;
}

// task descriptors (trace sizes filled, bases scanned afterwards)
__device__ __forceinline__ void build_task(int t, const int32_t *seg_task0,
                                           const int32_t *seg_first, const int32_t *seg_len,
                                           const int32_t *order, const int32_t *rlen,
                                           Task *tasks, int64_t *task_words,
                                           unsigned long long *steps_acc, int wide) {
    int seg = 0;
    while (seg + 1 < 2 * kRBins && seg_task0[seg + 1] <= t) ++seg;
    int cls = seg / kRBins;
    int R = kRBins - (seg % kRBins);
    int slots = cls == 0 ? kSlots16 : kSlots32;
    int G = kBatchG;
    if (wide && seg == 0) {
        slots = kSlotsWide;
        G = kWideG;
        R = kWideRows;
    }
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
    tk.steps = mx + G - 1;
    tk.min_rlen = mn;
    tasks[t] = tk;
    int kpw = cls == 0 ? 8 : 16;
    int W = (R + kpw - 1) / kpw;
    task_words[t] = (int64_t)((tk.steps + 3) & ~3) * W * 64;  // whole 4-step granules
    atomicAdd(&steps_acc[seg], (unsigned long long)tk.steps);
}

__global__ void __launch_bounds__(256)
sw_build_tasks(const int32_t *seg_task0, const int32_t *seg_first, const int32_t *seg_len,
               const int32_t *ntasks_p, const int32_t *order, const int32_t *rlen, Task *tasks,
               int64_t *task_words, unsigned long long *seg_steps, int wide) {
    // per-segment sums of the tasks' sweep steps (the host's cost estimate
    // of each class kernel), reduced in LDS before one atomic per segment
    __shared__ unsigned long long steps_acc[2 * kRBins];
    if (threadIdx.x < 2 * kRBins) steps_acc[threadIdx.x] = 0;
    __syncthreads();
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int ntasks = *ntasks_p;
    if (t < ntasks) build_task(t, seg_task0, seg_first, seg_len, order, rlen, tasks, task_words,
                               steps_acc, wide);
    __syncthreads();
    if (threadIdx.x < 2 * kRBins && steps_acc[threadIdx.x])
        atomicAdd(&seg_steps[threadIdx.x], steps_acc[threadIdx.x]);
}

// Exclusive scan over one 1,024-thread workgroup (each thread's value v):
// wave scans by shuffles, then the 16 wave totals through LDS.
template <class T>
__device__ __forceinline__ T wg_excl_scan(T v, T *wsum, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const T t = wsum[i];
        before += i < w ? t : T(0);
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// The whole of pass-1 binning for a batch of up to kBinTinyPairs pairs in
// one workgroup (classify, bin counts, their scan, the segments, the
// scatter, the tasks and their trace bases), with the sizes the host reads
// written straight into its mapped staging block: one launch and no gather
// instead of seven launches.  Same outputs as launch_binning + the gather
// (the order within a bin may differ; nothing depends on it).
__global__ void __launch_bounds__(1024)
sw_bin_tiny(BinArgs a, TaskArgs ta, int32_t *order, Task *tasks, int64_t *task_words,
            int64_t *total_words, const int32_t *counters, TinyStage hs) {
    constexpr int PP = kBinTinyPairs / 1024;           // pairs per thread
    constexpr int BP = (kNumBins + 1023) / 1024;       // bins per thread
    constexpr int TP = (kBinTinyPairs / kSlotsWide + 2 * kRBins + 1023) / 1024;  // tasks per thread
    __shared__ int32_t hist[kNumBins];  // counts, then cursors
    __shared__ int32_t s_task0[2 * kRBins + 1], s_first[2 * kRBins], s_len[2 * kRBins];
    __shared__ unsigned long long steps_acc[2 * kRBins];
    __shared__ int32_t wsum32[16];
    __shared__ int64_t wsum64[16];
    const int tid = threadIdx.x, n = a.npairs;
    for (int i = tid; i < kNumBins; i += 1024) hist[i] = 0;
    if (tid < 2 * kRBins) steps_acc[tid] = 0;
    __syncthreads();
    int bin[PP], rank[PP];
#pragma unroll
    for (int k = 0; k < PP; ++k) {
        const int p = k * 1024 + tid;
        bin[k] = p < n ? classify_pair(a, p) : -1;
        rank[k] = bin[k] >= 0 ? atomicAdd(&hist[bin[k]], 1) : 0;
    }
    __syncthreads();
    if (tid < 2 * kRBins) {
        int32_t len = 0;
        for (int lb = 0; lb < kLenBins; ++lb) len += hist[tid * kLenBins + lb];
        s_len[tid] = len;
    }
    // bin cursors: exclusive scan of the counts
    int32_t loc[BP], sum = 0;
#pragma unroll
    for (int i = 0; i < BP; ++i) {
        const int b = tid * BP + i;
        loc[i] = b < kNumBins ? hist[b] : 0;
        sum += loc[i];
    }
    int32_t binned;
    int32_t run = wg_excl_scan<int32_t>(sum, wsum32, &binned);
#pragma unroll
    for (int i = 0; i < BP; ++i) {
        const int b = tid * BP + i;
        if (b < kNumBins) hist[b] = run;
        run += loc[i];
    }
    __syncthreads();
    if (tid == 0) {
        int32_t t0 = 0;
        for (int seg = 0; seg < 2 * kRBins; ++seg) {
            const int cls = seg / kRBins;
            const int slots = cls == 0 ? (ta.wide && seg == 0 ? kSlotsWide : kSlots16) : kSlots32;
            s_first[seg] = hist[seg * kLenBins];
            s_task0[seg] = t0;
            t0 += (s_len[seg] + slots - 1) / slots;
        }
        s_task0[2 * kRBins] = t0;
    }
#pragma unroll
    for (int k = 0; k < PP; ++k)
        if (bin[k] >= 0) order[hist[bin[k]] + rank[k]] = k * 1024 + tid;
    __syncthreads();  // the order and the segments, for every thread of the group
    const int nt = s_task0[2 * kRBins];
    if (tid <= 2 * kRBins) ta.seg_task0[tid] = s_task0[tid];
    if (tid < 2 * kRBins) {
        ta.seg_first[tid] = s_first[tid];
        ta.seg_len[tid] = s_len[tid];
    }
    if (tid == 0) {
        *ta.ntasks = nt;
        *ta.npairs_binned = binned;
    }
    for (int t = tid; t < nt; t += 1024)
        build_task(t, s_task0, s_first, s_len, order, a.rlen, tasks, task_words, steps_acc,
                   ta.wide);
    __syncthreads();  // the task words
    int64_t tw[TP], tsum = 0;
#pragma unroll
    for (int i = 0; i < TP; ++i) {
        const int t = tid * TP + i;
        tw[i] = t < nt ? task_words[t] : 0;
        tsum += tw[i];
    }
    int64_t words;
    int64_t base = wg_excl_scan<int64_t>(tsum, wsum64, &words);
#pragma unroll
    for (int i = 0; i < TP; ++i) {
        const int t = tid * TP + i;
        if (t < nt) tasks[t].trace_base = base;
        base += tw[i];
    }
    if (tid < 2 * kRBins) {
        ta.seg_steps[tid] = steps_acc[tid];
        hs.steps[tid] = (int64_t)steps_acc[tid];
    }
    if (tid <= 2 * kRBins) hs.seg[tid] = s_task0[tid];
    if (tid == 0) {
        *total_words = words;
        *hs.words = words;
    }
    __syncthreads();  // the counters (the error flag of classify_pair included)
    if (tid < 16)
        hs.cnt[tid] = tid == 0   ? nt
                      : tid == 1 ? binned
                                 : __hip_atomic_load(const_cast<int32_t *>(counters) + tid,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct StoreTaskBase {
    Task *t;
    __device__ void operator()(int i, int64_t v) const { t[i].trace_base = v; }
};

__global__ void fill_i32(int32_t *p, int n, int32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// zeroes up to four arrays in one launch: blockIdx.y picks the array
__global__ void __launch_bounds__(256) zero_spans(ZeroSpans z) {
    int32_t *p = z.p[blockIdx.y];
    const int n = z.n[blockIdx.y];
    for (int i = blockIdx.x * 256 + (int)threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// Copies a few small device arrays into one host staging block (mapped
// pinned memory), so the host reads them all after a single stream
// synchronisation instead of one copy each.
__global__ void __launch_bounds__(256) gather_words(Gather g) {
    for (int s = 0; s < g.nspans; ++s)
        for (int i = threadIdx.x; i < g.nwords[s]; i += 256) g.out[g.dst[s] + i] = g.src[s][i];
}

// ---------------------------------------------------------------------------
// host-visible launch helpers (used by sw_host.cpp)
// ---------------------------------------------------------------------------
template <int R>
static hipError_t launch_plain_r(const char *q, const char *r, int qlen, int rlen_max,
                                 const KeyConsts &kc, int npairs, int pair0, uint32_t *trace,
                                 int64_t wpp, int4 *res, hipStream_t st) {
    size_t lds = sizeof(uint32_t) * (size_t)(rlen_max + 2 * kPlainPad);
    hipLaunchKernelGGL(sw_plain_kernel<R>, dim3(npairs), dim3(64), lds, st, q, r, qlen, rlen_max,
                       kc, npairs, pair0, trace, wpp, res);
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

// trace words per pair: whole 4-step granules of (rlen + 63) steps
static int64_t plain_trace_stride(int qlen, int rlen) {
    return (int64_t)((rlen + 63 + 3) & ~3) * plain_trace_words(plain_rows_for(qlen)) * 64;
}

// 2-bit moves per pair: a path has at most qlen + rlen moves
static int plain_move_words(int qlen, int rlen) { return (qlen + rlen) / 16 + 1; }

int64_t plain_scratch_words(int qlen, int rlen) {
    return plain_trace_stride(qlen, rlen) + 4 + plain_move_words(qlen, rlen);
}

// scratch of a plain chunk: [npairs x wpp trace][npairs int4 best cells]
// [npairs x mw moves]; wpp is a multiple of 4 words, so the int4s are
// 16-byte aligned
struct PlainScratch {
    int64_t wpp;
    int mw;
    uint32_t *trace;
    int4 *res;
    uint32_t *moves;
    PlainScratch(int qlen, int rlen_max, int npairs, uint32_t *scratch)
        : wpp(plain_trace_stride(qlen, rlen_max)), mw(plain_move_words(qlen, rlen_max)),
          trace(scratch), res(reinterpret_cast<int4 *>(scratch + (int64_t)npairs * wpp)),
          moves(scratch + (int64_t)npairs * (wpp + 4)) {}
};

// the sweep: trace and best cells into scratch
static hipError_t launch_plain_sweep(const char *q, const char *r, int qlen, int rlen_max,
                                     const Scoring &sc, int npairs, int pair0, uint32_t *scratch,
                                     hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    KeyConsts kc = make_consts(sc, kBytes);
    const PlainScratch ps(qlen, rlen_max, npairs, scratch);
    switch (plain_rows_for(qlen)) {
#define SWIFT_PCASE(RR) \
    case RR: return launch_plain_r<RR>(q, r, qlen, rlen_max, kc, npairs, pair0, ps.trace, ps.wpp, \
                                       ps.res, st);
        SWIFT_PCASE(1) SWIFT_PCASE(2) SWIFT_PCASE(3) SWIFT_PCASE(4) SWIFT_PCASE(5)
        SWIFT_PCASE(6) SWIFT_PCASE(7) SWIFT_PCASE(8) SWIFT_PCASE(10) SWIFT_PCASE(12)
        SWIFT_PCASE(16)
#undef SWIFT_PCASE
    default: return hipErrorInvalidValue;
    }
}

// the walk and consensus expansion, which read them
static hipError_t launch_plain_walk(const char *q, const char *r, int qlen, int rlen_max,
                                    int npairs, int pair0, uint32_t *scratch, const PlainOut &out,
                                    hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    const int R = plain_rows_for(qlen);
    const PlainScratch ps(qlen, rlen_max, npairs, scratch);
    const int nmv = (ps.mw + 63) / 64;
    if (npairs < kWaveWalkPairs && nmv <= 5) {
        auto k = nmv <= 1 ? sw_plain_walk_wave_kernel<1>
                 : nmv == 2 ? sw_plain_walk_wave_kernel<2>
                 : nmv == 3 ? sw_plain_walk_wave_kernel<3>
                 : nmv == 4 ? sw_plain_walk_wave_kernel<4>
                            : sw_plain_walk_wave_kernel<5>;
        hipLaunchKernelGGL(k, dim3((npairs + 3) / 4), dim3(256), 0, st, ps.trace, ps.wpp, R,
                           plain_trace_words(R), qlen, npairs, pair0, ps.res, ps.moves, ps.mw, out);
    } else {
        hipLaunchKernelGGL(sw_plain_walk_kernel, dim3((npairs + 255) / 256), dim3(256), 0, st,
                           ps.trace, ps.wpp, R, plain_trace_words(R), qlen, npairs, pair0, ps.res,
                           ps.moves, ps.mw, out);
    }
    const int qw = (qlen + 6) / 4, rw = (rlen_max + 6) / 4;
    hipLaunchKernelGGL(sw_plain_expand_kernel, dim3((npairs + 3) / 4), dim3(64),
                       sizeof(uint32_t) * 4 * (size_t)(qw + rw), st, q, r, qlen, rlen_max, npairs,
                       pair0, ps.res, ps.moves, ps.mw, qw, rw, out);
    return hipGetLastError();
}

hipError_t launch_plain(const char *q, const char *r, int qlen, int rlen_max, const Scoring &sc,
                        int npairs, int pair0, uint32_t *scratch, const PlainOut &out,
                        hipStream_t st) {
    hipError_t e = launch_plain_sweep(q, r, qlen, rlen_max, sc, npairs, pair0, scratch, st);
    if (e != hipSuccess) return e;
    return launch_plain_walk(q, r, qlen, rlen_max, npairs, pair0, scratch, out, st);
}


// Persistent grid of one row-class kernel: as many 256-thread blocks as fit
// on every CU at once (cached per kernel); `waves_per_cu` > 0 caps it.
template <int RR, bool PK, class K>
static int full_grid(K kernel, int cu_count, int waves_per_cu, int max_tasks) {
    static int cached = -1;  // one per row class and lane width
    if (cached < 0) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, 256, 0) != hipSuccess ||
            blocks < 1)
            blocks = 2;
        cached = blocks;
    }
    int b = cached;
    if (waves_per_cu > 0) b = min(b, max(1, waves_per_cu / 4));
    // with the task count known, no more waves than tasks: every wave of the
    // grid takes at least one turn at the queue's atomic counter, so idle
    // waves lengthen a small class's kernel
    if (max_tasks > 0) return min(b * cu_count, (max_tasks + 3) / 4);
    return b * cu_count;
}

hipError_t launch_batch_wide(const BatchArgs &ba, int cu_count, int max_tasks, hipStream_t st) {
    auto k = sw_batch_wide_kernel<kWideRows>;
    static int blocks = -1;  // per CU, at full occupancy
    if (blocks < 0) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, 256, 0) != hipSuccess || b < 1)
            b = 2;
        blocks = b;
    }
    int grid = blocks * cu_count;
    if (max_tasks > 0) grid = min(grid, (max_tasks + 3) / 4);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, ba);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &ba, bool int16, int rows, int cu_count, int waves_per_cu,
                        int max_tasks, hipStream_t st) {
    if (int16) {
        switch (rows) {
#define SWIFT_L16(RR)                                                                      \
    case RR: {                                                                             \
        auto k = sw_batch_kernel<RR, true, kOneHot>;                                       \
        hipLaunchKernelGGL(k, dim3(full_grid<RR, true>(k, cu_count, waves_per_cu, max_tasks)), dim3(256), 0, st, ba); \
        break;                                                                             \
    }
            SWIFT_R16_CASES(SWIFT_L16)
#undef SWIFT_L16
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (rows) {
#define SWIFT_L32(RR)                                                                      \
    case RR: {                                                                             \
        auto k = sw_batch_kernel<RR, false, kBytes>;                                       \
        hipLaunchKernelGGL(k, dim3(full_grid<RR, false>(k, cu_count, waves_per_cu, max_tasks)), dim3(256), 0, st, ba); \
        break;                                                                             \
    }
            SWIFT_R32_CASES(SWIFT_L32)
#undef SWIFT_L32
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_walk(const BatchArgs &ba, const int32_t *order, const int32_t *first,
                       const int32_t *count, int max_pairs, uint32_t *moves, int move_words,
                       hipStream_t st) {
    if (max_pairs <= 0) return hipSuccess;
    if (max_pairs < kWaveWalkPairs)
        hipLaunchKernelGGL(sw_walk_wave_kernel, dim3((max_pairs + 3) / 4), dim3(256), 0, st, ba,
                           order, first, count, moves, move_words);
    else
        hipLaunchKernelGGL(sw_walk_kernel, dim3((max_pairs + 255) / 256), dim3(256), 0, st, ba,
                           order, first, count, moves, move_words);
    hipLaunchKernelGGL(sw_expand_kernel, dim3((max_pairs + 15) / 16), dim3(256), 0, st, ba,
                       order, first, count, moves, move_words);
    return hipGetLastError();
}

hipError_t launch_binning_tiny(const BinArgs &bin, const TaskArgs &ta, int32_t *order,
                               Task *tasks, int64_t *task_words, int64_t *total_words,
                               const int32_t *counters, const TinyStage &hs, hipStream_t st) {
    if (bin.npairs > kBinTinyPairs) return hipErrorInvalidValue;
    hipLaunchKernelGGL(sw_bin_tiny, dim3(1), dim3(1024), 0, st, bin, ta, order, tasks, task_words,
                       total_words, counters, hs);
    return hipGetLastError();
}

hipError_t launch_binning(const BinArgs &bin, const TaskArgs &ta, int32_t *order, Task *tasks,
                          int64_t *task_words, const int64_t *base0, int64_t *total_words,
                          int64_t *scan_sums, int max_tasks, hipStream_t st) {
    int n = bin.npairs;
    const int chunks = (n + kBinChunk - 1) / kBinChunk;
    if (n > 0) hipLaunchKernelGGL(sw_bin_classify, dim3(chunks), dim3(256), 0, st, bin);
    hipLaunchKernelGGL(sw_bin_scan, dim3(1), dim3(1024), 0, st, ta);
    if (n > 0)
        hipLaunchKernelGGL(sw_bin_scatter, dim3(chunks), dim3(256), 0, st, bin.bin_of, n,
                           ta.bin_cursor, order);
    if (max_tasks > 0) {
        hipLaunchKernelGGL(sw_build_tasks, dim3((max_tasks + 255) / 256), dim3(256), 0, st,
                           ta.seg_task0, ta.seg_first, ta.seg_len, ta.ntasks, order, bin.rlen,
                           tasks, task_words, ta.seg_steps, ta.wide);
    }
    return launch_scan<int64_t>(LoadArr<int64_t>{task_words}, StoreTaskBase{tasks}, ta.ntasks,
                                max_tasks, scan_sums, base0, total_words, nullptr, st);
}

hipError_t launch_zero(const ZeroSpans &z, hipStream_t st) {
    int most = 0;
    for (int k = 0; k < z.nspans; ++k) most = max(most, z.n[k]);
    if (most <= 0) return hipSuccess;
    const int blocks = min((most + 255) / 256, 2048);
    hipLaunchKernelGGL(zero_spans, dim3(blocks, z.nspans), dim3(256), 0, st, z);
    return hipGetLastError();
}

hipError_t launch_fill(int32_t *p, int n, int32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_gather(const Gather &g, hipStream_t st) {
    hipLaunchKernelGGL(gather_words, dim3(1), dim3(256), 0, st, g);
    return hipGetLastError();
}

}  // namespace swift
