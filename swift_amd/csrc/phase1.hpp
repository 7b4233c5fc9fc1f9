// phase1.hpp -- device-side data layout of phase 1 (k-mer index, read
// mapping, region extraction, best-hit choice) shared by phase1.hip and the
// host code.
#pragma once

#include "swift_common.hpp"

namespace swift {

// Limits of the device mapper.  Hit keys pack (refIdx, shift, offset) into
// 64 bits: refIdx < 2^7 (the reference stores it in a signed char),
// shift + kShiftBias < 2^32, query offset < 2^11.
constexpr int kMaxSeeds = 15;          // k <= 15 keeps 4^k buckets < 2^30
constexpr int kMaxIndexSeqs = 127;     // refIdx is a signed char in the reference
constexpr int kMaxMapRead = 2047;      // query offsets fit 11 bits
constexpr int kShortRead = 511;        // batches of reads up to this long take the short-read
                                       // hash tier kernels
constexpr int64_t kShiftBias = 2048;   // shift >= -(kMaxMapRead - 1)
constexpr int kMapCap = 256;           // hits per strand sorted in LDS by one wave (2 KB of
                                       // keys: 8 waves per SIMD fit in LDS)
constexpr int kHashFirstHits = 1024;   // predicted hits per strand from which strands start
                                       // in the hash tier
constexpr int kMidCap1 = 4096;         // ... by one 256-thread workgroup (32 KB LDS)
constexpr int kHashSlots = 8192;       // cluster table of the hash tier (64 KB LDS)
constexpr int kHashCap = 6144;         // hits per strand the hash tier counts: table load
                                       // <= 0.75, so every insertion finds its slot
constexpr int kMidCap = 16384;         // ... by one 1024-thread workgroup (128 KB LDS)
constexpr int kLongFilterCap = 18432;  // short-read strands past kHashCap: the filter stage
                                       // at 1024 threads (18 hits per thread)
constexpr int kSmallBucket = 32;       // buckets sorted by one thread
constexpr int kWaveBucket = 1024;      // buckets sorted by one wave in LDS
constexpr int kScanChunk = 4096;       // hashes per scan workgroup

// Index arrays in HBM.  Sequences are stored as the index sees them (the
// reference's line handling applied, see the FASTA reader), back to back.
struct IndexDev {
    const char *seq;          // bases of every sequence
    const int64_t *seq_off;   // [nseq] offset of sequence s in seq
    const int64_t *seq_len;   // [nseq]
    const int64_t *tup0;      // [nseq] first tuple number of sequence s
    int nseq, k, thr;
    int64_t nhash;            // 4^k
    int32_t *count;           // [nhash] occurrences (_numRepeatsPerTuple)
    int32_t *start;           // [nhash] 1-based bucket start, 0 if empty/dropped
    int32_t *actual;          // [nhash] kept occurrences (0 if count > thr)
    int32_t *cursor;          // [nhash] fill cursor
    int2 *bucket;             // [nhash] (start, actual) for the mapper
    uint64_t *entry;          // [total + 1] refIdx << 32 | refPos, [0] sentinel
    // 32-bit cluster ids: cid = cid0[refIdx] + shift + kShiftBias, with
    // cid0[s] = sum over r < s of (len[r] + 2 * kShiftBias); increasing in
    // (refIdx, shift).  cid_ok = 0 when the sum does not fit (> 4 Gbp).
    const uint32_t *cid0;
    int cid_ok;
    // [total + 1] cid0[refIdx] + refPos of each entry (cid_ok only): a hit's
    // cluster id is this minus its query offset plus kShiftBias, so the hash
    // tier reads 4 bytes per hit instead of the 8-byte entry
    const uint32_t *gpos;
    // 1: every base of seq is one of ACGTNacgtn (so no window can send a pair
    // to the int32 re-run for its reference bytes)
    int plain_bases;
};

struct IndexBuildScratch {
    int64_t *block_sum;       // [nhash / kScanChunk + 1]
    int64_t *total;           // [1] kept tuples
    int32_t *max_actual;      // [1]
    int32_t *mid_list, *mid_count;   // buckets for the wave sort
    int32_t *big_list, *big_count;   // buckets for the block sort
    uint64_t *big_scratch;    // [big_slots * big_cap]
    int32_t big_slots, big_cap;
};

struct ReadsDev {
    const char *seq;
    const int64_t *off;
    const int32_t *len;
    int nreads;
};

// Per strand s = 2 * read + strand (strand 1 = reverse complement): up to
// `stride` best hits and the counters of the mapping.
struct MapOut {
    int8_t *ref;
    int32_t *shift, *pos;
    int32_t *nbest;           // hits reported (<= max_hits)
    int32_t *nbig;            // clusters of the biggest size (tie-set size)
    int32_t *nhits;           // k-mer hits gathered
    int stride;
};

struct MapArgs {
    IndexDev ix;
    ReadsDev rd;
    MapOut out;
    int max_hits;
    int max_len;              // longer reads (or negative lengths) map as empty
    char *strand_seq;         // optional: strand bases at s * qstride (>= max_len)
    int qstride;
    uint64_t *scratch;        // overflow strands: slots of scratch_cap keys
    int64_t scratch_cap;
    int scratch_slots;
    int32_t *ovl_list, *ovl_count;   // strands past kMapCap hits
    int32_t *mid_list, *mid_count;   // strands past the hash tier (kHashCap hits, or
                                     // kMidCap1 without 32-bit cluster ids)
    int32_t *big_list, *big_count;   // strands past kMidCap hits
    int32_t *mid2_list, *mid2_count; // short reads: strands the long filter stage leaves to
                                     // the workgroup sort
    int mid_blocks;                  // workgroups of map_strands_mid (0: none needed)
    int hash_first;                  // 1: every strand starts in the hash tier (no wave tier)
    int32_t *error;           // bit 0: a strand exceeded scratch_cap, bit 1: a read
                              // longer than max_len
};

// Records of the phase-1 -> phase-2 pipeline, one per reported hit, in read
// order, forward strand first, hits in cluster order.
struct PairsDev {
    int32_t *npairs;          // [1]
    int32_t *pair0;           // [2 * nreads + 1] first record of each strand
    int32_t *read, *strand, *ref, *shift, *pos, *nbig;
    int64_t *rdist;           // window start in the sequence
    int64_t *qoff, *roff, *coff;
    int32_t *qlen, *rlen;
    int capacity;
    int64_t cstride;
    int32_t *scan_sums;       // [2 * nreads / 2048 + 2]
};

hipError_t launch_index_build(const IndexDev &ix, int64_t ntup, const IndexBuildScratch &s,
                              hipStream_t st);
hipError_t launch_index_gpos(const IndexDev &ix, int64_t total, uint32_t *gpos, hipStream_t st);
hipError_t launch_index_fill(const IndexDev &ix, int64_t ntup, const IndexBuildScratch &s,
                             hipStream_t st);
// mid_blocks: persistent workgroups of map_strands_mid (0 when no strand can
// pass kMapCap hits)
hipError_t launch_map(const MapArgs &a, hipStream_t st, int mid_blocks);
// *flag |= 1 when seq[0, n) holds a byte outside ACGTNacgtn
hipError_t launch_seq_plain_scan(const char *seq, int64_t n, int32_t *flag, hipStream_t st);
hipError_t launch_records(const MapArgs &m, const PairsDev &p, int before, int win,
                          hipStream_t st);
hipError_t launch_choose_best(const PairsDev &p, int nreads, const int32_t *score,
                              int32_t *keep, int32_t *ntied, hipStream_t st);
// A served batch's results in host-visible (pinned) memory, written by one
// kernel: a 64-byte header (int32 records at 0, int64 packed string bytes at
// 8, the two int32 mapping flags at 16), then kServeCols int32 record
// columns of n entries each, the n int64 distances, and the query and
// reference consensus strings packed back to back (record i's alen bytes at
// the exclusive prefix of alen).
constexpr int kServeCols = 12;
struct ServeLayout {
    int64_t o_col, o_rd, o_q, o_r, end;
};
__host__ __device__ inline ServeLayout serve_layout(int n, int64_t packed) {
    ServeLayout l;
    l.o_col = 64;
    l.o_rd = l.o_col + (((int64_t)kServeCols * 4 * n + 7) & ~(int64_t)7);
    l.o_q = l.o_rd + 8 * (int64_t)n;
    l.o_r = l.o_q + ((packed + 7) & ~(int64_t)7);
    l.end = l.o_r + packed;
    return l;
}
struct ServeOut {
    const int32_t *nrec;                // records (device)
    const int32_t *flags;               // mapping flags, 2 ints (device)
    const int32_t *col[kServeCols];     // record columns; col[alen_col] the lengths
    int alen_col;
    const int64_t *rdist;
    const char *qc, *rc;                // consensus strings at stride cs
    int64_t cs;
    int64_t *pos, *sums, *total;        // scan scratch: cap + 1, cap / 2048 + 4, 1
    char *out;                          // pinned host memory, serve_layout(cap, cap * cs).end
};
// the scan of the string lengths, then the kernel that writes everything
hipError_t launch_serve_out(const ServeOut &o, int cap, hipStream_t st);
// mapHits6GetBestHits over caller hits already on the device: sorts the
// triples in place (refIdx, shift, refPos ascending) and writes the best
// clusters' first hits.  kbuf holds next_pow2(n) 16-byte keys.
hipError_t launch_best_hits(int8_t *ref, int32_t *shift, int32_t *pos, int n, int max_hits,
                            int8_t *o_ref, int32_t *o_shift, int32_t *o_pos, int32_t *o_n,
                            int32_t *o_nbig, void *kbuf, hipStream_t st);

}  // namespace swift
