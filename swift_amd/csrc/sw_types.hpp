// sw_types.hpp -- kernel argument structures and layout constants shared by
// sw_kernels.hip and the host code.
#pragma once

#include "swift_common.hpp"

namespace swift {

// per-launch constants in key space (key = 4*H + direction)
struct KeyConsts {
    int dr;    // added to 4H to form the diagonal-ready value
    int dsub;  // added when the substitution flag is set
    int eo2;   // 4*gapOpen + 2 (up exit, open)
    int eo1;   // 4*gapOpen + 1 (left exit, open)
    int de;    // 4*(gapExt - gapOpen)
    // 15, passed at run time: an arithmetic shift by it leaves the optimiser
    // no sign-splat to turn a masked merge of packed halves into per-half
    // compares and selects
    int sign;
};

// ONEHOT: flag = match,    dr = 4*mismatch + 3, dsub = 4*(match - mismatch)
// BYTES : flag = mismatch, dr = 4*match + 3,    dsub = 4*(mismatch - match)
enum Policy { kOneHot = 0, kBytes = 1 };

inline KeyConsts make_consts(const Scoring &s, int policy) {
    KeyConsts k;
    if (policy == kOneHot) {
        k.dr = 4 * s.mismatch + 3;
        k.dsub = 4 * (s.match - s.mismatch);
    } else {
        k.dr = 4 * s.match + 3;
        k.dsub = 4 * (s.mismatch - s.match);
    }
    k.eo2 = 4 * s.gap_open + 2;
    k.eo1 = 4 * s.gap_open + 1;
    k.de = 4 * (s.gap_ext - s.gap_open);
    k.sign = 15;
    return k;
}

constexpr int kColBits = 12;                   // int32 best key: H << 12 | (4095 - j)
constexpr int kMaxCols = (1 << kColBits) - 1;  // longest reference window (plain kernel)
constexpr int kMaxScore32 = 1 << 19;           // scores must stay below this
constexpr int kBatchG = 16;                    // lanes per pair group (batched kernels)
constexpr int kBatchCols = 480;                // reference columns per group (LDS)
constexpr int kRBins = 20;                     // R = 1..20 rows per lane (query <= 320)
constexpr int kLenBins = kBatchCols / 8;       // bins of 8 reference columns
constexpr int kNumBins = 2 * kRBins * kLenBins;
constexpr int kMaxBatchQ = kBatchG * kRBins;   // 320
constexpr int kMaxBatchR = kBatchCols;         // 480
constexpr int kSlots16 = 8;                    // pairs per wave task, int16 packed
constexpr int kSlots32 = 4;                    // pairs per wave task, int32
// wide form (small batches): one 64-lane group per wave, packed int16
constexpr int kWideG = 64;
constexpr int kWideRows = kMaxBatchQ / kWideG;  // 5
constexpr int kSlotsWide = 2;

__host__ __device__ inline int r32_class(int r) { return r <= 2 ? 2 : (r + 1) & ~1; }

struct Task {
    int64_t trace_base;  // word offset
    int32_t first;       // index of the task's first pair in the sorted order
    int16_t npairs;      // pairs in the task (<= slots per wave)
    int16_t rows;        // R
    int32_t steps;       // max rlen + G - 1
    int32_t min_rlen;    // smallest rlen among the task's pairs
};

struct BatchArgs {
    const char *qseq;
    const int64_t *qoff;
    const int32_t *qlen;
    const char *rseq;
    const int64_t *roff;
    const int32_t *rlen;
    const int32_t *order;     // sorted pair order
    const Task *tasks;
    const int32_t *task_lo;   // device: first task of this kernel's class
    const int32_t *task_hi;   // device: one past its last task
    int32_t *next_task;       // work-queue cursor (zeroed before launch)
    uint32_t *trace;
    PairResult *res;
    PairTrace *ptr;
    KeyConsts kc;
    uint32_t h_limit;         // 4*H at or above this flags an int16 overflow risk
    // traceback outputs, written by the fused walk
    int32_t *score_out, *start_out, *end_out, *alen_out;
    char *qcons, *rcons;
    const int64_t *coff;
    int walk_deferred;        // 1: record PairTrace descriptors, walk in sw_walk_kernel
    int32_t *nflagged;        // counts the pairs flagged for the int32 re-run
};

// outputs of the plain (alignSequences) kernel: reference layout
struct PlainOut {
    float *score;
    int32_t *start, *end, *alen;
    char *qcons, *rcons;
    int64_t cstride;
};


struct BinArgs {
    const char *rseq;
    const int64_t *roff;
    const int32_t *qlen;
    const int32_t *rlen;
    const PairResult *prev;   // rerun pass: only flagged pairs take part
    int npairs;
    int force32;              // scoring outside the int16 range: all pairs int32
    int32_t *bin_of;          // per pair bin (or -1 when not in this pass)
    int32_t *bin_count;       // kNumBins
    int32_t *error;           // set to nonzero when a pair's lengths are out of range
    // empty pairs (query or reference of length 0) get their all-zero result
    // here and take no part in the kernels
    int32_t *score_out, *start_out, *end_out, *alen_out;
    // > 0: rows per lane rounded up to a multiple of `coarse` (kCoarseRows,
    // or kCoarseRowsSmall for the smallest batches: at most kRBins / coarse
    // classes; exact, the extra rows are padding); 0: every R its own class
    int coarse;
};

constexpr int kCoarseRows = 5;
constexpr int kCoarseRowsSmall = 10;

struct TaskArgs {
    const int32_t *bin_count;
    int32_t *bin_cursor;      // exclusive prefix of bin_count (scatter cursor)
    int32_t *seg_task0;       // first task of each (class, R) segment (+1 sentinel)
    int32_t *seg_first;       // first sorted pair of each segment
    int32_t *seg_len;         // pairs in each segment
    int32_t *ntasks;
    int32_t *npairs_binned;
    unsigned long long *seg_steps;  // sweep steps summed over each segment's tasks
    // 1: segment 0 (int16, R = kRBins) runs the wide form: kSlotsWide pairs
    // per task, kWideRows rows per lane, rlen + kWideG - 1 steps
    int wide;
};

// launch helpers (sw_kernels.hip)
// max_tasks > 0: the class's task count, known on the host (the grid then
// has no more waves than tasks); 0: a full persistent grid
hipError_t launch_batch(const BatchArgs &ba, bool int16, int rows, int cu_count, int waves_per_cu,
                        int max_tasks, hipStream_t st);
// the wide form's kernel (int16, kWideRows rows per lane)
hipError_t launch_batch_wide(const BatchArgs &ba, int cu_count, int max_tasks, hipStream_t st);
// deferred walk-back over pairs order[*first .. *first + *count) (first may
// be null: 0; max_pairs bounds *count); moves holds move_words words per pair
// id, move_words >= (kMaxBatchQ + kMaxBatchR) / 16 + 1.  Under
// kWaveWalkPairs pairs a wave walks each pair, else a lane.
constexpr int kWaveWalkPairs = 16384;
hipError_t launch_walk(const BatchArgs &ba, const int32_t *order, const int32_t *first,
                       const int32_t *count, int max_pairs, uint32_t *moves, int move_words,
                       hipStream_t st);
// pass-1 binning in one workgroup for batches up to kBinTinyPairs pairs,
// the sizes written into the host's mapped staging block (TinyStage)
constexpr int kBinTinyPairs = 4096;
struct TinyStage {
    int64_t *words;   // total trace words
    int64_t *steps;   // 2 * kRBins: sweep steps per segment
    int32_t *seg;     // 2 * kRBins + 1: first task of each segment
    int32_t *cnt;     // 16: the counters (ntasks, binned pairs, error flag, ...)
};
hipError_t launch_binning_tiny(const BinArgs &bin, const TaskArgs &ta, int32_t *order,
                               Task *tasks, int64_t *task_words, int64_t *total_words,
                               const int32_t *counters, const TinyStage &hs, hipStream_t st);
// scan_sums: room for max_tasks / 2048 + 1 int64 values; bin.bin_count
// must be zero on entry
hipError_t launch_binning(const BinArgs &bin, const TaskArgs &ta, int32_t *order, Task *tasks,
                          int64_t *task_words, const int64_t *base0, int64_t *total_words,
                          int64_t *scan_sums, int max_tasks, hipStream_t st);

}  // namespace swift
