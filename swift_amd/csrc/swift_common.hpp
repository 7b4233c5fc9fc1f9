// swift_common.hpp -- shared host/device definitions of the swift_amd HIP
// library (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace swift {

// ----------------------------------------------------------------------------
// Error handling.  The reference prints the CUDA error and exits
// (PRINT_CUDA_ERROR, src/common.h:155-167); the drop-in entry points keep
// that behaviour, the swift_* entry points return a status instead.
// ----------------------------------------------------------------------------
enum Status : int {
    kOk = 0,
    kErrInvalid = 1,   // bad argument (sizes, scoring outside the supported range)
    kErrHip = 2,       // a HIP runtime call failed
    kErrNoIndex = 3,   // phase-1 call before an index was built
    kErrIO = 4,        // file could not be read or written
    kErrCapacity = 5,  // input beyond a documented size limit
    kErrTimeout = 6,   // a wait ended before the work finished
};

const char *last_error();
void set_error(const char *fmt, ...);

#define SWIFT_HIP_TRY(expr)                                                  \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ::swift::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr,   \
                               hipGetErrorString(e_));                       \
            return ::swift::kErrHip;                                         \
        }                                                                    \
    } while (0)

#define SWIFT_HIP_EXIT(expr)                                                 \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            fprintf(stderr, "Error: %s:%s:%d: %s\n", __FILE__, __func__,     \
                    __LINE__, hipGetErrorString(e_));                        \
            exit(1);                                                         \
        }                                                                    \
    } while (0)

// ----------------------------------------------------------------------------
// Scoring.  Reference scores are integer-valued floats (the CLI parses them
// with atoi, src/main.cu:307-326); kernels work in a key space where a cell
// key is 4*H + direction, direction 3 = diagonal, 2 = up, 1 = left,
// 0 = start, so one integer max reproduces findArrMax's first-maximum rule
// over [diag, up, left, 0] (src/smithWaterman.cu:194-210).
// ----------------------------------------------------------------------------
struct Scoring {
    int match, mismatch, gap_open, gap_ext;
};

// Byte classes used by the one-hot match test of the batched kernels:
// match(q, r) <=> (qcode & rcode) != 0.  Query 'N' matches everything and a
// reference 'N'/'n' matches everything (isMatch_cpu, src/smithWaterman.cu
// :849-855 with the reference byte upper-cased); any other query byte
// (lower case included: the query is not upper-cased) only matches a
// reference N.  Reference bytes outside ACGTNacgtn need byte equality and are
// routed to the byte-semantics kernel.
constexpr uint32_t kCodeA = 1, kCodeC = 2, kCodeG = 4, kCodeT = 8;
constexpr uint32_t kCodeOther = 16, kCodeN = 31;

__host__ __device__ inline uint32_t query_code(unsigned char c) {
    switch (c) {
    case 'A': return kCodeA;
    case 'C': return kCodeC;
    case 'G': return kCodeG;
    case 'T': return kCodeT;
    case 'N': return kCodeN;
    default: return kCodeOther;
    }
}

// Shift form of the same rule used by the packed kernels: a reference byte
// becomes a class index (A 0, C 1, G 2, T 3, N 4, anything else 5) and a
// query byte a bit mask over the classes it matches, so
// match(q, r) <=> (query_mask(q) >> ref_index(r)) & 1.
constexpr uint32_t kRefOther = 5;
// Class of the padding columns around a pair's reference in the packed
// kernels: no query mask has bit 6, so a padding column matches nothing.
constexpr uint32_t kRefPad = 6;
__host__ __device__ constexpr uint32_t ref_index(unsigned char c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    case 'N': case 'n': return 4;
    default: return kRefOther;
    }
}
__host__ __device__ constexpr uint32_t query_mask(unsigned char c) {
    switch (c) {
    case 'A': return 0x11;
    case 'C': return 0x12;
    case 'G': return 0x14;
    case 'T': return 0x18;
    case 'N': return 0x3F;
    default: return 0x10;  // only a reference N matches
    }
}

// Branch-free forms of the two classifiers for the kernels (a switch over
// bytes compiles to chains of compares under exec masks).  (b >> 1) & 7
// separates A, C, G, T, N (and a, c, g, t, n): slots 0, 1, 3, 2, 7; a byte
// is the slot's letter or falls in the "other" class.
__host__ __device__ constexpr uint32_t ref_index_bf(unsigned char c) {
    const uint32_t u = c | 0x20u;  // the reference byte, case folded
    const uint32_t h = (u >> 1) & 7u;
    const uint32_t letter = (uint32_t)(0x6E00000067746361ull >> (8 * h)) & 0xFFu;
    return letter == u ? (0x45552310u >> (4 * h)) & 0xFu : kRefOther;
}
__host__ __device__ constexpr uint32_t query_mask_bf(unsigned char c) {
    const uint32_t h = ((uint32_t)c >> 1) & 7u;  // the query byte, case sensitive
    const uint32_t letter = (uint32_t)(0x4E00000047544341ull >> (8 * h)) & 0xFFu;
    return letter == c ? (uint32_t)(0x3F00000014181211ull >> (8 * h)) & 0xFFu : 0x10u;
}
constexpr bool classifiers_agree() {
    for (int c = 0; c < 256; ++c)
        if (ref_index_bf((unsigned char)c) != ref_index((unsigned char)c) ||
            query_mask_bf((unsigned char)c) != query_mask((unsigned char)c))
            return false;
    return true;
}
static_assert(classifiers_agree(), "branch-free byte classifiers differ from the tables");

// 0 marks a reference byte the one-hot kernels cannot represent.
__host__ __device__ inline uint32_t ref_code(unsigned char c) {
    switch (c) {
    case 'A': case 'a': return kCodeA;
    case 'C': case 'c': return kCodeC;
    case 'G': case 'g': return kCodeG;
    case 'T': case 't': return kCodeT;
    case 'N': case 'n': return kCodeN;
    default: return 0;
    }
}

// toupper_cpu (src/smithWaterman.cu:1194-1199)
__host__ __device__ inline int upper_byte(int ch) {
    if ((unsigned int)(ch - 'a') < 26u) ch += 'A' - 'a';
    return ch;
}

// ----------------------------------------------------------------------------
// Trace / pair-result layout shared by the forward kernels and the walk-back.
// A forward kernel writes, per pair, the best cell and where its traceback
// bits live: words of 2-bit direction codes indexed
//   base + ((t * W + w) * 64 + lane)
// where t = step (= column index j0 + lane offset g), lane = lane0 + g holds
// query rows [g*R, g*R + R), word w holds rows [w*KPW, w*KPW + KPW) of the
// lane and the code of row k sits at bit (half * 16 + 2 * (k % KPW)).
// ----------------------------------------------------------------------------
struct PairTrace {
    int64_t base;     // word offset of the pair's task in the trace buffer
    int32_t lane0;    // first lane of the pair's lane group
    int16_t rows;     // R: query rows per lane
    int8_t words;     // W: trace words per lane per step
    int8_t half;      // 0/1: which 16-bit half (packed int16), 0 for int32
};

struct PairResult {
    int32_t score;    // best H (0 when no positive cell)
    int32_t row;      // 1-based row of the best cell (0 when score == 0)
    int32_t col;      // 1-based column of the best cell
    int32_t flags;    // bit 0: int16 overflow suspected -> rerun in int32
};

}  // namespace swift
