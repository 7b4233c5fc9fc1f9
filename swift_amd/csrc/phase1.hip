// phase1.hip -- k-mer index construction, read mapping, region extraction
// and best-hit choice on gfx950, with the semantics of the reference CPU
// path:
//   lookupTable5Create (src/lookupTable5.cu:101-252): non-overlapping
//     k-mers at positions 0, k, 2k, ... of each sequence, hash
//     sum_i 4^i code(s[i]) (getHash, src/preprocess.cu:70-91), a hash seen
//     more than thr times dropped entirely, buckets in hash order starting
//     at 1, entries in reverse file order within a bucket;
//   lookupTable5MapQry + mapHits6GetBestHits (src/lookupTable5.cu:557-603,
//     src/mapHits6.cu:40-167): every overlapping query k-mer gathers its
//     bucket as hits (refIdx, refPos - i, refPos); clusters are runs of equal
//     (refIdx, shift) after sorting by (refIdx, shift, refPos); the first hit
//     of each biggest cluster is reported, at most maxHits of them, in sorted
//     order.  The reference draws among more than maxHits tied clusters with
//     rand() (src/mapHits6.cu:146-160); the mapper keeps the first maxHits,
//     one member of the reference's tie set, and reports the set's size.
//
// Layout: the index lives in HBM as an int2 bucket table (start, kept count)
// plus 64-bit entries (refIdx << 32 | refPos).  One wave maps one read
// strand: lanes hash query offsets, a wave prefix sum lays the hits out, the
// wave gathers bucket entries cooperatively (one hit per lane), sorts 64-bit
// hit keys in LDS and compacts the biggest clusters with ballots.  Strands
// whose hits exceed LDS take the same steps in a global scratch slot with a
// 256-thread workgroup.
#include <algorithm>

#include "phase1.hpp"
#include "scan.hpp"

namespace swift {

__device__ __forceinline__ uint32_t kcode(unsigned char c) {
    // preprocessCreate (src/preprocess.cu:30-48): A/a 0, C/c 1, G/g 2,
    // T/t 3, every other byte 0
    switch (c) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}

__device__ __forceinline__ unsigned char rc_byte(unsigned char c) {
    // qryGetReverseComplement (src/query.cu:265-282): anything but A/C/G -> A
    return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
}

// first base = least significant base-4 digit
__device__ __forceinline__ uint32_t kmer_hash(const unsigned char *p, int k) {
    uint32_t h = 0;
    for (int i = k - 1; i >= 0; --i) h = h * 4 + kcode(p[i]);
    return h;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Bitonic sort of n2 (a power of two) keys by threads tid, tid + nthr, ...;
// `sync` orders the stages.  DESC = false sorts ascending.
template <bool DESC, class K, class Sync>
__device__ void bitonic(K *key, int n2, int tid, int nthr, Sync sync) {
    for (int kk = 2; kk <= n2; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n2; i += nthr) {
                int p = i ^ j;
                if (p > i) {
                    bool asc = ((i & kk) == 0) != DESC;
                    K x = key[i], y = key[p];
                    if (asc ? (y < x) : (x < y)) {
                        key[i] = y;
                        key[p] = x;
                    }
                }
            }
            sync();
        }
    }
}

__device__ __forceinline__ int next_pow2(int n) {
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    return n2;
}

// ---------------------------------------------------------------------------
// Index build
// ---------------------------------------------------------------------------
__device__ __forceinline__ int seq_of_tuple(const int64_t *tup0, int nseq, int64_t t) {
    int lo = 0, hi = nseq - 1;  // last s with tup0[s] <= t
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tup0[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) idx_count(IndexDev ix, int64_t ntup) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntup;
         t += (int64_t)gridDim.x * blockDim.x) {
        int s = seq_of_tuple(ix.tup0, ix.nseq, t);
        int64_t pos = (t - ix.tup0[s]) * ix.k;
        uint32_t h = kmer_hash((const unsigned char *)ix.seq + ix.seq_off[s] + pos, ix.k);
        atomicAdd(&ix.count[h], 1);
    }
}

__device__ __forceinline__ int kept(int c, int thr) { return (c > 0 && c <= thr) ? c : 0; }

// per-chunk sums of the kept counts and the largest kept count
__global__ void __launch_bounds__(1024) idx_scan_blocks(IndexDev ix, IndexBuildScratch s) {
    __shared__ int64_t red[1024];
    __shared__ int32_t redm[1024];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int64_t sum = 0;
    int32_t mx = 0;
    for (int i = threadIdx.x; i < kScanChunk; i += 1024) {
        int64_t h = base + i;
        if (h < ix.nhash) {
            int c = kept(ix.count[h], ix.thr);
            sum += c;
            mx = max(mx, c);
        }
    }
    red[threadIdx.x] = sum;
    redm[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[threadIdx.x] += red[threadIdx.x + off];
            redm[threadIdx.x] = max(redm[threadIdx.x], redm[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s.block_sum[blockIdx.x] = red[0];
        atomicMax(s.max_actual, redm[0]);
    }
}

// exclusive scan of the chunk sums (one workgroup); the total into *total
__global__ void __launch_bounds__(1024) idx_scan_sums(IndexBuildScratch s, int nblocks) {
    __shared__ int64_t part[1024];
    const int per = (nblocks + 1023) / 1024;
    int64_t sum = 0;
    for (int i = 0; i < per; ++i) {
        int b = threadIdx.x * per + i;
        if (b < nblocks) sum += s.block_sum[b];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - sum;
    for (int i = 0; i < per; ++i) {
        int b = threadIdx.x * per + i;
        if (b < nblocks) {
            int64_t v = s.block_sum[b];
            s.block_sum[b] = run;
            run += v;
        }
    }
    if (threadIdx.x == 1023) *s.total = part[1023];
}

// start = 1 + exclusive prefix of the kept counts (0 for empty or dropped
// hashes), actual = kept count (src/lookupTable5.cu:199-208)
__global__ void __launch_bounds__(1024) idx_scan_apply(IndexDev ix, IndexBuildScratch s) {
    __shared__ int64_t part[1024];
    constexpr int per = kScanChunk / 1024;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * per;
    int32_t loc[per];
    int64_t sum = 0;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        int64_t h = base + i;
        loc[i] = h < ix.nhash ? kept(ix.count[h], ix.thr) : 0;
        sum += loc[i];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = s.block_sum[blockIdx.x] + part[threadIdx.x] - sum;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        int64_t h = base + i;
        if (h < ix.nhash) {
            int32_t st = loc[i] ? (int32_t)(run + 1) : 0;
            ix.start[h] = st;
            ix.actual[h] = loc[i];
            ix.bucket[h] = make_int2(st, loc[i]);
            ix.cursor[h] = 0;
        }
        run += loc[i];
    }
}

// scatter every kept tuple into its bucket (order fixed by the sort below)
__global__ void __launch_bounds__(256) idx_fill(IndexDev ix, int64_t ntup) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntup;
         t += (int64_t)gridDim.x * blockDim.x) {
        int s = seq_of_tuple(ix.tup0, ix.nseq, t);
        int64_t pos = (t - ix.tup0[s]) * ix.k;
        uint32_t h = kmer_hash((const unsigned char *)ix.seq + ix.seq_off[s] + pos, ix.k);
        if (ix.actual[h] == 0) continue;
        int slot = ix.start[h] + atomicAdd(&ix.cursor[h], 1);
        ix.entry[slot] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)pos;
    }
}

// Reverse file order within a bucket (src/lookupTable5.cu:219-229) is
// descending (refIdx, refPos).  Buckets of <= kSmallBucket entries are sorted
// by one thread in registers; larger ones are listed for the wave sort or
// the workgroup sort.
__global__ void __launch_bounds__(256) idx_sort_small(IndexDev ix, IndexBuildScratch s) {
    int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (h >= ix.nhash) return;
    const int a = ix.actual[h];
    if (a < 2) return;
    if (a > kSmallBucket) {
        if (a <= kWaveBucket) s.mid_list[atomicAdd(s.mid_count, 1)] = (int32_t)h;
        else s.big_list[atomicAdd(s.big_count, 1)] = (int32_t)h;
        return;
    }
    uint64_t *e = ix.entry + ix.start[h];
    uint64_t v[kSmallBucket];
#pragma unroll
    for (int i = 0; i < kSmallBucket; ++i) v[i] = i < a ? e[i] : 0;
    // padding zeros sort last; a real entry equal to 0 is indistinguishable
#pragma unroll
    for (int i = 1; i < kSmallBucket; ++i) {
#pragma unroll
        for (int j = i; j > 0; --j) {
            uint64_t x = v[j - 1], y = v[j];
            bool sw = y > x;
            v[j - 1] = sw ? y : x;
            v[j] = sw ? x : y;
        }
    }
#pragma unroll
    for (int i = 0; i < kSmallBucket; ++i)
        if (i < a) e[i] = v[i];
}

// one wave per listed bucket, keys in LDS
__global__ void __launch_bounds__(64) idx_sort_mid(IndexDev ix, IndexBuildScratch s) {
    __shared__ uint64_t key[kWaveBucket];
    const int n = *s.mid_count;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int h = s.mid_list[it];
        const int a = ix.actual[h];
        uint64_t *e = ix.entry + ix.start[h];
        const int n2 = next_pow2(a);
        for (int i = threadIdx.x; i < n2; i += 64) key[i] = i < a ? e[i] : 0;
        __syncthreads();
        bitonic<true>(key, n2, threadIdx.x, 64, [] { __syncthreads(); });
        for (int i = threadIdx.x; i < a; i += 64) e[i] = key[i];
        __syncthreads();
    }
}

// one workgroup per listed bucket, keys in a global scratch slot
__global__ void __launch_bounds__(256) idx_sort_big(IndexDev ix, IndexBuildScratch s) {
    const int n = *s.big_count;
    uint64_t *key = s.big_scratch + (int64_t)blockIdx.x * s.big_cap;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int h = s.big_list[it];
        const int a = ix.actual[h];
        uint64_t *e = ix.entry + ix.start[h];
        const int n2 = next_pow2(a);
        for (int i = threadIdx.x; i < n2; i += 256) key[i] = i < a ? e[i] : 0;
        __syncthreads();
        bitonic<true>(key, n2, threadIdx.x, 256, [] { __syncthreads(); });
        for (int i = threadIdx.x; i < a; i += 256) e[i] = key[i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Mapper
// ---------------------------------------------------------------------------
// hit key: refIdx (bits 43..49) | shift + kShiftBias (bits 11..42) | query
// offset (bits 0..10); ascending keys = the reference's (refIdx, shift,
// refPos) order because refPos = shift + offset
__device__ __forceinline__ uint64_t hit_key(uint64_t entry, int i) {
    uint64_t ref = entry >> 32;
    uint64_t shift_b = (uint64_t)((int64_t)(uint32_t)entry - i + kShiftBias);
    return (ref << 43) | (shift_b << 11) | (uint64_t)i;
}

// Cluster compaction over sorted keys[0..n) by one wave: the biggest run
// size, the number of runs of that size, and the first `max_hits` of them
// in sorted order (findBiggestClusterSize / findBiggestClusters,
// src/mapHits6.cu:202-223,132-167).
template <class K, class ClusterOf, class Emit>
__device__ void wave_best_clusters(const K *key, int n, int max_hits, ClusterOf cl, Emit emit,
                                   int32_t *o_nbest, int32_t *o_nbig) {
    const int lane = threadIdx.x & 63;
    auto run_at = [&](int e) {
        if (e >= n || (e > 0 && cl(key[e]) == cl(key[e - 1]))) return 0;
        int len = 1;
        const auto c = cl(key[e]);
        while (e + len < n && cl(key[e + len]) == c) ++len;
        return len;
    };
    int best = 0;
    for (int b = 0; b < n; b += 64) best = max(best, run_at(b + lane));
#pragma unroll
    for (int sft = 32; sft >= 1; sft >>= 1) best = max(best, __shfl_xor(best, sft, 64));
    int found = 0;
    for (int b = 0; b < n; b += 64) {
        const int e = b + lane;
        const bool big = best > 0 && run_at(e) == best;
        const uint64_t mask = __ballot(big);
        const int slot = found + __popcll(mask & ((1ull << lane) - 1));
        if (big && slot < max_hits) emit(slot, key[e]);
        found += __popcll(mask);
    }
    if (lane == 0) {
        *o_nbest = min(found, max_hits);
        *o_nbig = found;
    }
}

struct HitCluster {
    __device__ uint64_t operator()(uint64_t k) const { return k >> 11; }
};

// Gathers the hits of one strand into kbuf (capacity cap) with one wave.
// Returns the hit count; *ok is false when it exceeds cap (kbuf then holds
// a prefix).  ex/stt: 64-int LDS arrays.
__device__ int gather_hits(const IndexDev &ix, const unsigned char *qb, int nq, uint64_t *kbuf,
                           int64_t cap, int32_t *ex, int32_t *stt, bool *ok) {
    const int lane = threadIdx.x & 63;
    int total = 0;
    bool fits = true;
    for (int b = 0; b < nq; b += 64) {
        const int i = b + lane;
        int cnt = 0, st = 0;
        if (i < nq) {
            int2 bk = ix.bucket[kmer_hash(qb + i, ix.k)];
            st = bk.x;
            cnt = bk.y;
        }
        const int incl = wave_incl_scan(cnt);
        const int tot = __shfl(incl, 63, 64);
        if (fits && (int64_t)total + tot <= cap) {
            ex[lane] = incl - cnt;
            stt[lane] = st;
            wave_sync();
            for (int e = lane; e < tot; e += 64) {
                int o = 0;  // last lane whose exclusive offset is <= e
#pragma unroll
                for (int sft = 32; sft >= 1; sft >>= 1)
                    if (ex[o + sft] <= e) o += sft;
                kbuf[total + e] = hit_key(ix.entry[stt[o] + (e - ex[o])], b + o);
            }
            wave_sync();
        } else {
            fits = false;
        }
        total += tot;
    }
    *ok = fits;
    return total;
}

// The hits gather_hits lays out, for strands of at most kFlatRounds * 64
// k-mers: every bucket lookup is issued first and every entry load of the
// strand in one pass, several in flight per lane, so a strand costs two
// dependent memory round trips instead of two per 64 k-mers.  ex / stt hold
// kFlatRounds * 64 entries.
constexpr int kFlatRounds = 4;
__device__ int gather_hits_flat(const IndexDev &ix, const unsigned char *qb, int nq,
                                uint64_t *kbuf, int64_t cap, int32_t *ex, int32_t *stt, bool *ok) {
    const int lane = threadIdx.x & 63;
    int cnt[kFlatRounds], st[kFlatRounds];
#pragma unroll
    for (int r = 0; r < kFlatRounds; ++r) {
        const int i = r * 64 + lane;
        cnt[r] = 0;
        st[r] = 0;
        if (i < nq) {
            const int2 bk = ix.bucket[kmer_hash(qb + i, ix.k)];
            st[r] = bk.x;
            cnt[r] = bk.y;
        }
    }
    int total = 0;
#pragma unroll
    for (int r = 0; r < kFlatRounds; ++r) {
        const int incl = wave_incl_scan(cnt[r]);
        ex[r * 64 + lane] = total + incl - cnt[r];
        stt[r * 64 + lane] = st[r];
        total += __shfl(incl, 63, 64);
    }
    *ok = (int64_t)total <= cap;
    if (!*ok) return total;
    wave_sync();
    constexpr int U = 4;  // entry loads in flight per lane
    for (int e0 = lane; e0 < total; e0 += U * 64) {
        uint64_t v[U];
        int oo[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 64;
            oo[u] = -1;
            v[u] = 0;
            if (e < total) {
                int o = 0;  // the last k-mer whose exclusive offset is <= e
#pragma unroll
                for (int sft = kFlatRounds * 32; sft >= 1; sft >>= 1)
                    if (ex[o + sft] <= e) o += sft;
                v[u] = ix.entry[stt[o] + (e - ex[o])];
                oo[u] = o;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (oo[u] >= 0) kbuf[e0 + u * 64] = hit_key(v[u], oo[u]);
    }
    wave_sync();
    return total;
}

// Strand sidx of read (sidx >> 1), length L at offset off of the read
// buffer, into qb (reverse-complemented for the odd strand), and into the
// strand buffer when there is one.
__device__ void load_strand_at(const MapArgs &a, int sidx, int L, int64_t off, unsigned char *qb,
                               int *len_out) {
    const int strand = sidx & 1;
    if (L < 0 || L > a.max_len) {
        if (threadIdx.x == 0) atomicOr(a.error, 2);
        L = 0;
    }
    const unsigned char *src = (const unsigned char *)a.rd.seq + off;
    char *dst = a.strand_seq ? a.strand_seq + (int64_t)sidx * a.qstride : nullptr;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        unsigned char c = src[i];
        int at = strand ? L - 1 - i : i;
        unsigned char v = strand ? rc_byte(c) : c;
        qb[at] = v;
        if (dst) dst[at] = (char)v;
    }
    *len_out = L;
}

__device__ void load_strand(const MapArgs &a, int sidx, unsigned char *qb, int *len_out) {
    const int r = sidx >> 1;
    load_strand_at(a, sidx, a.rd.len[r], a.rd.off[r], qb, len_out);
}

__device__ void emit_best(const MapArgs &a, int sidx, const uint64_t *keys, int n) {
    const int64_t o = (int64_t)sidx * a.out.stride;
    wave_best_clusters(
        keys, n, a.max_hits, HitCluster{},
        [&](int slot, uint64_t k) {
            int sh = (int)((int64_t)((k >> 11) & 0xFFFFFFFFull) - kShiftBias);
            a.out.ref[o + slot] = (int8_t)(k >> 43);
            a.out.shift[o + slot] = sh;
            a.out.pos[o + slot] = sh + (int)(k & 2047);
        },
        a.out.nbest + sidx, a.out.nbig + sidx);
}

// one wave per strand; strands whose hits overflow LDS are listed
__global__ void __launch_bounds__(64) map_strands(MapArgs a) {
    __shared__ uint64_t keys[kMapCap];
    __shared__ unsigned char qb[kMaxMapRead + 1];
    __shared__ int32_t ex[kFlatRounds * 64], stt[kFlatRounds * 64];
    const int sidx = blockIdx.x;
    const int lane = threadIdx.x;
    int L;
    load_strand(a, sidx, qb, &L);
    __syncthreads();
    bool ok;
    const int nq = L - a.ix.k + 1;
    const int total = nq <= kFlatRounds * 64
                          ? gather_hits_flat(a.ix, qb, nq, keys, kMapCap, ex, stt, &ok)
                          : gather_hits(a.ix, qb, nq, keys, kMapCap, ex, stt, &ok);
    if (lane == 0) a.out.nhits[sidx] = total;
    if (total == 0) {
        if (lane == 0) {
            a.out.nbest[sidx] = 0;
            a.out.nbig[sidx] = 0;
        }
        return;
    }
    if (!ok) {
        if (lane == 0) a.ovl_list[atomicAdd(a.ovl_count, 1)] = sidx;
        return;
    }
    const int n2 = next_pow2(total);
    for (int i = total + lane; i < n2; i += 64) keys[i] = ~0ull;
    __syncthreads();
    bitonic<false>(keys, n2, lane, 64, [] { __syncthreads(); });
    emit_best(a, sidx, keys, total);
}

// Strands past one tier's capacity are sorted by a workgroup of THREADS
// threads with CAP keys in LDS: thread t gathers query offsets t, t + THREADS,
// ...; a workgroup scan of the per-thread hit counts lays the hits out.
// Strands past CAP hits are listed for the next tier (`next_list`).  The
// tiers: 256 threads / 4096 keys (32 KB LDS, several workgroups per CU),
// 1024 threads / 16384 keys (128 KB).
template <int THREADS, int CAP>
__global__ void __launch_bounds__(THREADS)
map_strands_wg(MapArgs a, const int32_t *list, const int32_t *list_count, int32_t *next_list,
               int32_t *next_count) {
    constexpr int NW = THREADS / 64;
    constexpr int PER = (kMaxMapRead + THREADS) / THREADS;  // offsets per thread
    __shared__ uint64_t keys[CAP];
    __shared__ unsigned char qb[kMaxMapRead + 1];
    __shared__ int32_t wsum[NW];
    const int n = *list_count;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int sidx = list[it];
        int L;
        load_strand(a, sidx, qb, &L);
        __syncthreads();
        const int nq = L - a.ix.k + 1;
        int c[PER], st[PER], v = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int o = t + u * THREADS;
            c[u] = 0;
            st[u] = 0;
            if (o < nq) {
                const int2 b = a.ix.bucket[kmer_hash(qb + o, a.ix.k)];
                st[u] = b.x;
                c[u] = b.y;
            }
            v += c[u];
        }
        const int incl = wave_incl_scan(v);
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int x = wsum[w];
            before += w < wv ? x : 0;
            total += x;
        }
        if (total > CAP) {
            if (t == 0) next_list[atomicAdd(next_count, 1)] = sidx;
            __syncthreads();
            continue;
        }
        int at = before + incl - v;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int o = t + u * THREADS;
            const uint64_t *e = a.ix.entry + st[u];
            int j = 0;
            for (; j + 4 <= c[u]; j += 4) {  // four independent loads in flight
                const uint64_t e0 = e[j], e1 = e[j + 1], e2 = e[j + 2], e3 = e[j + 3];
                keys[at + j] = hit_key(e0, o);
                keys[at + j + 1] = hit_key(e1, o);
                keys[at + j + 2] = hit_key(e2, o);
                keys[at + j + 3] = hit_key(e3, o);
            }
            for (; j < c[u]; ++j) keys[at + j] = hit_key(e[j], o);
            at += c[u];
        }
        const int n2 = next_pow2(total);
        for (int i = total + t; i < n2; i += THREADS) keys[i] = ~0ull;
        __syncthreads();
        bitonic<false>(keys, n2, t, THREADS, [] { __syncthreads(); });
        if (t < 64) emit_best(a, sidx, keys, total);
        __syncthreads();
    }
}

// Hash tier (strands of kMapCap < hits <= CAP when 32-bit cluster ids
// fit): no sort.  Each hit is counted into an LDS open-addressing table keyed
// by its cluster id (count in the high half of a word, by atomic add; the
// smallest query offset in the low half, filled in afterwards for the
// clusters of the biggest count only); the inserting threads keep each hit's
// slot and the largest count they produced.  After the biggest count is known,
// every cluster of that count is listed once, through its hit at the
// cluster's smallest query offset, and one wave picks the first max_hits of
// them in cluster-id order, i.e. the (refIdx, shift) order of
// findBiggestClusters with the first element of each cluster.  The threads
// then clear only the slots their hits used.
// The table has SLOTS entries (load <= 3/4 at CAP hits); reads of at most
// MAXQ bases (the batch's max_len).  A smaller table and k-mer slot array for
// short reads leave room for more workgroups, i.e. more strands' loads in
// flight per CU; strands past CAP are handed on to the next tier.
template <int SLOTS>
__device__ __forceinline__ uint32_t cid_slot(uint32_t cid) {
    return (cid * 2654435761u) >> (32 - __builtin_ctz(SLOTS));  // Fibonacci hashing
}

template <int THREADS>
__device__ __forceinline__ uint64_t block_min_u64(uint64_t v, uint64_t *red) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        uint64_t o = __shfl_xor(v, s, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t m = ~0ull;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) m = red[w] < m ? red[w] : m;
    __syncthreads();
    return m;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, s, 64));
    return v;
}

template <int THREADS, int SLOTS, int CAP, int MAXQ, int WPE>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WPE)))
map_strands_hash(MapArgs a, const int32_t *list, const int32_t *list_count, int32_t *next_list,
                 int32_t *next_count) {
    constexpr int NW = THREADS / 64;
    constexpr int PER = (MAXQ + THREADS) / THREADS;
    constexpr uint32_t kEmpty = 0xFFFFFFFFu;
    constexpr uint32_t kFresh = 0x0000FFFFu;  // count 0, smallest offset 0xFFFF
    constexpr int NS = THREADS * PER;            // k-mer slots, s = thread * PER + u
    constexpr int HPT = CAP / THREADS;           // hits per thread (a strand here has <= CAP)
    constexpr int kCand = 64;                    // biggest clusters picked by one wave
    static_assert(CAP % THREADS == 0 && CAP < 32768, "hit offsets are int16");
    static_assert(4 * CAP <= 3 * SLOTS && (SLOTS & (SLOTS - 1)) == 0,
                  "distinct clusters fill <= 3/4 of a power-of-two table");
    static_assert(MAXQ <= kMaxMapRead, "query offsets fit 11 bits");
    __shared__ uint32_t tkey[SLOTS], tval[SLOTS];
    __shared__ unsigned char qb[MAXQ + 1];
    __shared__ int32_t wsum[NW];
    __shared__ uint64_t red[NW];
    __shared__ int16_t sex[NS];                  // block-wide exclusive hit offset of slot s
    __shared__ int32_t sbase[NS];                // its first entry index minus that offset
    __shared__ uint64_t cand[kCand];             // keys of the biggest clusters
    __shared__ uint32_t sh_best;
    __shared__ int32_t sh_ncand;
    __shared__ uint32_t scid0[kMaxIndexSeqs];   // cluster-id base of each sequence
    // no list: every strand of the batch, in order (hash_first)
    const int n = list ? *list_count : 2 * a.rd.nreads;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = t; i < SLOTS; i += THREADS) {
        tkey[i] = kEmpty;
        tval[i] = kFresh;
    }
    for (int i = t; i < a.ix.nseq; i += THREADS) scid0[i] = a.ix.cid0[i];
    // Software pipeline over the block's strands: while strand i is counted,
    // the bases of the block's next strand are loaded (into registers, then
    // into qb once this strand no longer reads it) and its buckets' heads
    // gathered, so a strand's entries are its one exposed global round trip.
    // Read lengths and offsets run one strand further ahead.
    auto strand_at = [&](int it) { return list ? list[it] : it; };
    const int G = gridDim.x;
    // a strand's checked length and this thread's bytes of it
    auto fetch = [&](int len, int64_t off, unsigned char (&nb)[PER]) -> int {
        if (len < 0 || len > a.max_len) {
            if (t == 0) atomicOr(a.error, 2);
            len = 0;
        }
        const unsigned char *src = (const unsigned char *)a.rd.seq + off;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = t + u * THREADS;
            nb[u] = i < len ? src[i] : 0;
        }
        return len;
    };
    // the strand's bases into qb (reverse complement for strand 1), and into
    // the pipeline's strand copy
    auto store = [&](int sidx, int len, const unsigned char (&nb)[PER]) {
        const int strand = sidx & 1;
        char *dst = a.strand_seq ? a.strand_seq + (int64_t)sidx * a.qstride : nullptr;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = t + u * THREADS;
            if (i < len) {
                const int at = strand ? len - 1 - i : i;
                const unsigned char b = strand ? rc_byte(nb[u]) : nb[u];
                qb[at] = b;
                if (dst) dst[at] = (char)b;
            }
        }
    };
    // bucket (first entry, count) of each k-mer of the strand in qb
    auto heads = [&](int len, int (&c)[PER], int (&st)[PER]) {
        const int nq = len - a.ix.k + 1;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int o = t + u * THREADS;
            c[u] = 0;
            st[u] = 0;
            if (o < nq) {
                const int2 b = a.ix.bucket[kmer_hash(qb + o, a.ix.k)];
                st[u] = b.x;
                c[u] = b.y;
            }
        }
    };
    int c[PER], st[PER], L0 = 0;
    int nx_len = 0;
    int64_t nx_off = 0;
    if ((int)blockIdx.x < n) {
        const int s0 = strand_at(blockIdx.x);
        unsigned char nb[PER];
        L0 = fetch(a.rd.len[s0 >> 1], a.rd.off[s0 >> 1], nb);
        store(s0, L0, nb);
        if ((int)blockIdx.x + G < n) {
            const int r = strand_at(blockIdx.x + G) >> 1;
            nx_len = a.rd.len[r];
            nx_off = a.rd.off[r];
        }
    }
    __syncthreads();
    heads(L0, c, st);
    for (int it = blockIdx.x; it < n; it += G) {
        const int sidx = strand_at(it);
        const bool more = it + G < n;
        const int nsidx = more ? strand_at(it + G) : 0;
        unsigned char nb[PER];
        int nL = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) nb[u] = 0;
        if (more) {
            nL = fetch(nx_len, nx_off, nb);
            if (it + 2 * G < n) {
                const int r = strand_at(it + 2 * G) >> 1;
                nx_len = a.rd.len[r];
                nx_off = a.rd.off[r];
            }
        }
        if (t == 0) {
            sh_best = 0;
            sh_ncand = 0;
        }
        int v = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) v += c[u];
        const int incl = wave_incl_scan(v);
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        int total = 0, before = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            before += w < wv ? wsum[w] : 0;
            total += wsum[w];
        }
        if (!list && t == 0) a.out.nhits[sidx] = total;  // the wave tier's count otherwise
        if (total > CAP) {
            if (t == 0) next_list[atomicAdd(next_count, 1)] = sidx;
            // qb's last reads (this strand's heads) are behind the barrier above
            store(nsidx, nL, nb);
            __syncthreads();
            heads(nL, c, st);
            continue;
        }
        // The strand's hits, flattened over the block: hit e belongs to the
        // last slot whose exclusive offset is <= e.  Each thread then loads
        // its (at most HPT) entries back to back instead of walking one
        // k-mer's bucket serially, which left the block waiting on its
        // largest bucket, one dependent load per hit.
        {
            int run = before + incl - v;
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                sex[t * PER + u] = (int16_t)run;
                sbase[t * PER + u] = st[u] - run;
                run += c[u];
            }
        }
        __syncthreads();
        uint32_t hen[HPT];
        int hq[HPT], hh[HPT];
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            const int e = t + k * THREADS;
            hen[k] = 0;
            hq[k] = -1;
            hh[k] = 0;
            if (e < total) {
                int sl = 0;
#pragma unroll
                for (int step = NS / 2; step >= 1; step >>= 1)
                    if (sex[sl + step] <= e) sl += step;
                hen[k] = a.ix.gpos[sbase[sl] + e];
                hq[k] = sl / PER + (sl % PER) * THREADS;  // the hit's query offset
            }
        }
        // the next strand's bases (qb is free, see above)
        store(nsidx, nL, nb);
        // count every hit into the table; remember its slot and the largest
        // count this thread produced
        uint32_t mybest = 0;
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hq[k] < 0) continue;
            const uint32_t cid = hen[k] - (uint32_t)hq[k] + (uint32_t)kShiftBias;
            uint32_t h = cid_slot<SLOTS>(cid);
            for (;;) {
                const uint32_t old = atomicCAS(&tkey[h], kEmpty, cid);
                if (old == kEmpty || old == cid) break;
                h = (h + 1) & (SLOTS - 1);
            }
            // count in the high half (the low half stays 0xFFFF until the
            // offset pass below)
            const uint32_t cnt = (atomicAdd(&tval[h], 1u << 16) >> 16) + 1;
            mybest = max(mybest, cnt);
            hh[k] = (int)h;
        }
        mybest = wave_max_u32(mybest);
        if (lane == 0) atomicMax(&sh_best, mybest);
        __syncthreads();
        // the next strand's heads, in flight through the rest of this one
        heads(nL, c, st);
        const uint32_t best = sh_best;
        // smallest query offset of each cluster of the biggest count: every
        // update of such a slot carries the same count bits
#pragma unroll
        for (int k = 0; k < HPT; ++k)
            if (hq[k] >= 0 && (tval[hh[k]] >> 16) == best)
                atomicMin(&tval[hh[k]], (best << 16) | (uint32_t)hq[k]);
        __syncthreads();
        // The clusters of the biggest count, each through its hit at the
        // cluster's smallest query offset (one hit per cluster: two hits of
        // one cluster at the same offset would be the same reference position).
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hq[k] < 0) continue;
            const uint32_t tv = tval[hh[k]];
            if ((tv >> 16) == best && (int)(tv & 0xFFFFu) == hq[k]) {
                const int at = atomicAdd(&sh_ncand, 1);
                if (at < kCand) cand[at] = ((uint64_t)tkey[hh[k]] << 16) | (uint32_t)hq[k];
            }
        }
        __syncthreads();
        const int nbig_all = sh_ncand;
        // the first max_hits of them by cluster id
        const int take = min(nbig_all, a.max_hits);
        const int64_t ob = (int64_t)sidx * a.out.stride;
        auto emit = [&](int r, uint64_t k) {
            const uint32_t cid = (uint32_t)(k >> 16);
            int ref = 0;  // last sequence whose cid0 <= cid
            for (int lo = 0, hi = a.ix.nseq - 1; lo <= hi;) {
                const int mid = (lo + hi) >> 1;
                if (scid0[mid] <= cid) {
                    ref = mid;
                    lo = mid + 1;
                } else {
                    hi = mid - 1;
                }
            }
            const int sh = (int)((int64_t)(cid - scid0[ref]) - kShiftBias);
            a.out.ref[ob + r] = (int8_t)ref;
            a.out.shift[ob + r] = sh;
            a.out.pos[ob + r] = sh + (int)(k & 0xFFFFu);
        };
        if (nbig_all <= kCand) {
            if (wv == 0) {  // one wave: repeated wave minimum over the list
                uint64_t kk = lane < nbig_all ? cand[lane] : ~0ull;
                for (int r = 0; r < take; ++r) {
                    uint64_t m = kk;
#pragma unroll
                    for (int s2 = 32; s2 >= 1; s2 >>= 1) {
                        const uint64_t o2 = __shfl_xor(m, s2, 64);
                        m = o2 < m ? o2 : m;
                    }
                    if (lane == 0) emit(r, m);
                    if (kk == m) kk = ~0ull;  // cluster ids are distinct
                }
            }
        } else {
            // more tied clusters than the list holds: scan the table
            uint64_t last = 0;
            for (int r = 0; r < take; ++r) {
                uint64_t mine = ~0ull;
                for (int i = t; i < SLOTS; i += THREADS) {
                    if ((tval[i] >> 16) != best) continue;
                    const uint64_t k = ((uint64_t)tkey[i] << 16) | (tval[i] & 0xFFFFu);
                    if ((r == 0 || k > last) && k < mine) mine = k;
                }
                const uint64_t k = block_min_u64<THREADS>(mine, red);
                last = k;
                if (t == 0) emit(r, k);
            }
        }
        if (t == 0) {
            a.out.nbest[sidx] = take;
            a.out.nbig[sidx] = nbig_all;
        }
        // every table read of this strand is behind a barrier: clear the
        // slots this thread's hits used
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hq[k] < 0) continue;
            tkey[hh[k]] = kEmpty;
            tval[hh[k]] = kFresh;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Filter tier (reads up to MAXQ bases, 32-bit cluster ids): one THREADS-wide
// workgroup per strand of up to CAP hits, in one pass.  Most hits of a
// strand are clusters of one (random matches of its k-mers across the
// genome); only a cluster id seen at least twice can be in a cluster of two
// or more.  So every hit first sets the bit hash(cid) of a bitmap A in LDS
// (atomic or); a hit that finds its bit already set sets the same bit of a
// bitmap B.  After a barrier a hit is a candidate iff its bit in B is set:
// that holds for every hit of every cluster of size >= 2 (the bit depends on
// the cluster id alone) and for a few clusters of one whose bit collides.
// Only candidates are counted exactly, in a small LDS table (count by atomic
// add, smallest query offset by atomic min).  If the largest count M is at
// least 2, the biggest clusters are the table's clusters of count M, each
// listed once through the hit whose insertion returned M; otherwise every
// cluster is a single hit and the first max_hits clusters are the hits with
// the smallest cluster ids.  Strands with more candidate hits than the
// table holds go on to the hash tier (`ovl`), strands past CAP hits to the
// sort tier (`mid`).
//
// Hits are flattened over the workgroup (hit e by thread e % THREADS); the
// k-mer slot of hit e comes from an owner map (slot + 1 written at the
// slot's first hit, then a workgroup max-scan) kept in bitmap A's words
// before the bitmaps are needed.
template <int THREADS, int CAP, int SMALL, int SMALLCAP, int MAXQ, int WPE, bool LIST>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WPE)))
map_strands_filter(MapArgs a, const int32_t *src, const int32_t *src_count, int32_t *ovl,
                   int32_t *ovl_count, int32_t *mid, int32_t *mid_count) {
    constexpr int NW = THREADS / 64;
    constexpr int PER = (MAXQ + THREADS) / THREADS;
    constexpr int NS = THREADS * PER;            // k-mer slots, s = thread * PER + u
    constexpr int HPT = CAP / THREADS;           // hits per thread
    constexpr int kBits = 1 << 16;               // bits of each bitmap
    constexpr int kWords = kBits / 32;
    constexpr uint32_t kEmpty = 0xFFFFFFFFu;
    constexpr int kCand = 64;
    // the owner map (16 bits per hit) lives in bitmap A and B's words while
    // it fits them, else in an array of its own
    constexpr bool kOwnSep = CAP / 2 > 2 * kWords;
    static_assert(CAP % THREADS == 0 && HPT % 2 == 0, "owner entries two to a word");
    static_assert(NS < 65535 && MAXQ < 1024 && HPT % 3 == 0,
                  "owner entries; three 10-bit query offsets per register");
    static_assert(4 * SMALLCAP <= 3 * SMALL && (SMALL & (SMALL - 1)) == 0,
                  "candidate clusters fill <= 3/4 of a power-of-two table");
    __shared__ __align__(16) uint32_t bm[2 * kWords];   // bitmaps A | B; first the owner map
    __shared__ __align__(16) uint32_t own_sep[kOwnSep ? CAP / 2 : 1];
    __shared__ uint32_t skey[SMALL], scnt[SMALL], smin[SMALL];
    __shared__ uint32_t ccid[SMALLCAP];          // candidate hits: cluster id,
    __shared__ uint32_t cqc[SMALLCAP];           // offset | count returned by its insertion << 16
    __shared__ unsigned char qb[MAXQ + 1];
    __shared__ int32_t sbase[NS];                // first entry index of slot s minus its offset
    __shared__ int32_t wsum[NW];
    __shared__ uint32_t wmax[NW];
    __shared__ uint64_t red[NW];
    __shared__ uint64_t cand[kCand];
    __shared__ uint32_t sh_best;
    __shared__ int32_t sh_ncand, sh_nhit;
    __shared__ uint32_t scid0[kMaxIndexSeqs];
    // LIST: the strands of src (src_count of them), else every strand
    const int n = LIST ? *src_count : 2 * a.rd.nreads;
    auto strand_at = [&](int i) { return LIST ? src[i] : i; };
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = t; i < 2 * kWords; i += THREADS) bm[i] = 0;
    if constexpr (kOwnSep)
        for (int i = t; i < CAP / 2; i += THREADS) own_sep[i] = 0;
    for (int i = t; i < SMALL; i += THREADS) {
        skey[i] = kEmpty;
        scnt[i] = 0;
        smin[i] = 0xFFFFFFFFu;
    }
    for (int i = t; i < a.ix.nseq; i += THREADS) scid0[i] = a.ix.cid0[i];
    uint32_t *ow_words = kOwnSep ? own_sep : bm;  // the owner map, two entries a word
    uint16_t *owner = reinterpret_cast<uint16_t *>(ow_words);
    // the 4-byte entries through a buffer resource: one 32-bit offset register
    // per load in flight (the index holds < 2^32 bytes of them when cluster
    // ids fit 32 bits; loads past the range return 0 instead of faulting)
    const __amdgpu_buffer_rsrc_t gpos_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.ix.gpos, 0, (int)0xFFFFFFFFu, 0x00020000);
    const int G = gridDim.x;
    auto fetch = [&](int len, int64_t off, unsigned char (&nb)[PER]) -> int {
        if (len < 0 || len > a.max_len) {
            if (t == 0) atomicOr(a.error, 2);
            len = 0;
        }
        const unsigned char *src = (const unsigned char *)a.rd.seq + off;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = t + u * THREADS;
            nb[u] = i < len ? src[i] : 0;
        }
        return len;
    };
    auto store = [&](int sidx, int len, const unsigned char (&nb)[PER]) {
        const int strand = sidx & 1;
        char *dst = a.strand_seq ? a.strand_seq + (int64_t)sidx * a.qstride : nullptr;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = t + u * THREADS;
            if (i < len) {
                const int at = strand ? len - 1 - i : i;
                const unsigned char b = strand ? rc_byte(nb[u]) : nb[u];
                qb[at] = b;
                if (dst) dst[at] = (char)b;
            }
        }
    };
    auto heads = [&](int len, int (&c)[PER], int (&st)[PER]) {
        const int nq = len - a.ix.k + 1;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int o = t + u * THREADS;
            c[u] = 0;
            st[u] = 0;
            if (o < nq) {
                const int2 b = a.ix.bucket[kmer_hash(qb + o, a.ix.k)];
                st[u] = b.x;
                c[u] = b.y;
            }
        }
    };
    auto bit_of = [](uint32_t cid) { return (cid * 2654435761u) >> 16; };  // 16-bit index
    auto small_slot = [](uint32_t cid) {
        return (cid * 2246822519u) >> (32 - __builtin_ctz(SMALL));
    };
    int c[PER], st[PER], L0 = 0;
    int nx_len = 0;
    int64_t nx_off = 0;
    if ((int)blockIdx.x < n) {
        const int s0 = strand_at(blockIdx.x);
        unsigned char nb[PER];
        L0 = fetch(a.rd.len[s0 >> 1], a.rd.off[s0 >> 1], nb);
        store(s0, L0, nb);
        if ((int)blockIdx.x + G < n) {
            const int r = strand_at(blockIdx.x + G) >> 1;
            nx_len = a.rd.len[r];
            nx_off = a.rd.off[r];
        }
    }
    __syncthreads();
    heads(L0, c, st);
    for (int it = blockIdx.x; it < n; it += G) {
        const int sidx = strand_at(it);
        const bool more = it + G < n;
        const int nsidx = more ? strand_at(it + G) : 0;
        unsigned char nb[PER];
        int nL = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) nb[u] = 0;
        if (more) {
            nL = fetch(nx_len, nx_off, nb);
            if (it + 2 * G < n) {
                const int r = strand_at(it + 2 * G) >> 1;
                nx_len = a.rd.len[r];
                nx_off = a.rd.off[r];
            }
        }
        if (t == 0) {
            sh_best = 0;
            sh_ncand = 0;
            sh_nhit = 0;
        }
        int v = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) v += c[u];
        const int incl = wave_incl_scan(v);
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        int total = 0, before = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            before += w < wv ? wsum[w] : 0;
            total += wsum[w];
        }
        if (t == 0) a.out.nhits[sidx] = total;
        if (total > CAP) {
            if (t == 0) mid[atomicAdd(mid_count, 1)] = sidx;
            store(nsidx, nL, nb);
            __syncthreads();
            heads(nL, c, st);
            continue;
        }
        // owner map: slot + 1 at each non-empty slot's first hit, then a
        // max-scan (thread t: hits [t * HPT, t * HPT + HPT))
        {
            int run = before + incl - v;
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                if (c[u] > 0) owner[run] = (uint16_t)(t * PER + u + 1);
                sbase[t * PER + u] = st[u] - run;
                run += c[u];
            }
        }
        __syncthreads();
        uint32_t ow[HPT / 2];
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < HPT / 2; ++j) {
            ow[j] = ow_words[t * (HPT / 2) + j];
            m = max(m, max(ow[j] & 0xFFFFu, ow[j] >> 16));
        }
        uint32_t wm = m;  // wave inclusive max-scan of the thread maxima
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)wm, d, 64);
            if (lane >= d) wm = max(wm, o);
        }
        if (lane == 63) wmax[wv] = wm;
        __syncthreads();
        uint32_t carry = (uint32_t)__shfl_up((int)wm, 1, 64);
        if (lane == 0) carry = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w)
            if (w < wv) carry = max(carry, wmax[w]);
        if (t * HPT < total) {
#pragma unroll
            for (int j = 0; j < HPT / 2; ++j) {
                const uint32_t lo = max(carry, ow[j] & 0xFFFFu);
                const uint32_t hi = max(lo, ow[j] >> 16);
                carry = hi;
                ow_words[t * (HPT / 2) + j] = lo | (hi << 16);
            }
        }
        __syncthreads();
        // the hits: hit k of this thread (e = t + k * THREADS) exists iff
        // e < total; its query offset (10 bits, three to a register) and
        // 4-byte global position, which less the offset is the cluster id
        uint32_t hcid[HPT], hqp[HPT / 3];
        auto has = [&](int k) { return t + k * THREADS < total; };
        auto qof = [&](int k) { return (int)((hqp[k / 3] >> (10 * (k % 3))) & 0x3FFu); };
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            const int e = t + k * THREADS;
            hcid[k] = 0;
            if (k % 3 == 0) hqp[k / 3] = 0;
            if (e < total) {
                const int sl = (int)owner[e] - 1;
                hcid[k] = __builtin_amdgcn_raw_buffer_load_b32(
                    gpos_rsrc, (uint32_t)(sbase[sl] + e) * 4u, 0, 0);
                hqp[k / 3] |= (uint32_t)(sl / PER + (sl % PER) * THREADS) << (10 * (k % 3));
            }
        }
        // the next strand's bases (qb is free: this strand's heads are in
        // registers) and, once every owner entry is read, the bitmaps' zeros
        store(nsidx, nL, nb);
        __syncthreads();
        heads(nL, c, st);  // in flight through the rest of this strand
        if (t * HPT < total) {
#pragma unroll
            for (int j = 0; j < HPT / 2; ++j) ow_words[t * (HPT / 2) + j] = 0;
        }
        __syncthreads();
        // bitmap A: seen; bitmap B: seen again
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (!has(k)) continue;
            hcid[k] = hcid[k] - (uint32_t)qof(k) + (uint32_t)kShiftBias;
            const uint32_t b = bit_of(hcid[k]);
            const uint32_t msk = 1u << (b & 31);
            if (atomicOr(&bm[b >> 5], msk) & msk) atomicOr(&bm[kWords + (b >> 5)], msk);
        }
        __syncthreads();
        // the candidate hits into a list (order does not matter)
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (!has(k)) continue;
            const uint32_t b = bit_of(hcid[k]);
            if (bm[kWords + (b >> 5)] & (1u << (b & 31))) {
                const int at = atomicAdd(&sh_nhit, 1);
                if (at < SMALLCAP) {
                    ccid[at] = hcid[k];
                    cqc[at] = (uint32_t)qof(k);
                }
            }
        }
        __syncthreads();
        const int ncand_all = sh_nhit;
        if (ncand_all > SMALLCAP) {
            // too many candidates for the table: the hash tier counts them all
            if (t == 0) ovl[atomicAdd(ovl_count, 1)] = sidx;
        } else {
            uint32_t mybest = 0;
            for (int i = t; i < ncand_all; i += THREADS) {
                const uint32_t cid = ccid[i];
                const uint32_t q = cqc[i];
                uint32_t h = small_slot(cid);
                for (;;) {
                    const uint32_t old = atomicCAS(&skey[h], kEmpty, cid);
                    if (old == kEmpty || old == cid) break;
                    h = (h + 1) & (SMALL - 1);
                }
                const uint32_t cnt = atomicAdd(&scnt[h], 1u) + 1;
                atomicMin(&smin[h], q);
                cqc[i] = q | (cnt << 16);
                mybest = max(mybest, cnt);
            }
            mybest = wave_max_u32(mybest);
            if (lane == 0 && mybest) atomicMax(&sh_best, mybest);
            __syncthreads();
            const uint32_t M = sh_best;
            const int64_t ob = (int64_t)sidx * a.out.stride;
            auto emit = [&](int r, uint32_t cid, int q) {
                int ref = 0;  // last sequence whose cid0 <= cid
                for (int lo = 0, hi = a.ix.nseq - 1; lo <= hi;) {
                    const int mid2 = (lo + hi) >> 1;
                    if (scid0[mid2] <= cid) {
                        ref = mid2;
                        lo = mid2 + 1;
                    } else {
                        hi = mid2 - 1;
                    }
                }
                const int sh = (int)((int64_t)(cid - scid0[ref]) - kShiftBias);
                a.out.ref[ob + r] = (int8_t)ref;
                a.out.shift[ob + r] = sh;
                a.out.pos[ob + r] = sh + q;
            };
            // the slot of a listed cluster id (present in the table)
            auto find = [&](uint32_t cid) {
                uint32_t h = small_slot(cid);
                while (skey[h] != cid) h = (h + 1) & (SMALL - 1);
                return h;
            };
            if (M >= 2) {
                // each cluster of count M once: through the hit whose
                // insertion returned M
                for (int i = t; i < ncand_all; i += THREADS) {
                    if ((cqc[i] >> 16) != M) continue;
                    const int at = atomicAdd(&sh_ncand, 1);
                    if (at < kCand) cand[at] = ccid[i];
                }
                __syncthreads();
                const int nbig = sh_ncand;
                const int take = min(nbig, a.max_hits);
                if (nbig <= kCand) {
                    if (wv == 0) {  // first `take` cluster ids: repeated wave minimum
                        uint32_t kk = lane < nbig ? (uint32_t)cand[lane] : 0xFFFFFFFFu;
                        for (int r = 0; r < take; ++r) {
                            uint32_t mn = kk;
#pragma unroll
                            for (int s2 = 32; s2 >= 1; s2 >>= 1)
                                mn = min(mn, (uint32_t)__shfl_xor((int)mn, s2, 64));
                            if (lane == 0) emit(r, mn, (int)smin[find(mn)]);
                            if (kk == mn) kk = 0xFFFFFFFFu;  // cluster ids are distinct
                        }
                    }
                } else {
                    // more tied clusters than the list holds: scan the table
                    uint32_t last = 0;
                    for (int r = 0; r < take; ++r) {
                        uint64_t mine = ~0ull;
                        for (int i = t; i < SMALL; i += THREADS)
                            if (scnt[i] == M && (r == 0 || skey[i] > last) && skey[i] < mine)
                                mine = skey[i];
                        const uint64_t kmin = block_min_u64<THREADS>(mine, red);
                        last = (uint32_t)kmin;
                        if (t == 0) emit(r, last, (int)smin[find(last)]);
                    }
                }
                if (t == 0) {
                    a.out.nbest[sidx] = take;
                    a.out.nbig[sidx] = nbig;
                }
            } else {
                // every cluster is one hit: the hits with the smallest cluster
                // ids (distinct), each emitted by its thread with its offset
                const int take = min(total, a.max_hits);
                uint32_t last = 0;
                for (int r = 0; r < take; ++r) {
                    uint32_t mine = 0xFFFFFFFFu;
#pragma unroll
                    for (int k = 0; k < HPT; ++k)
                        if (has(k) && (r == 0 || hcid[k] > last)) mine = min(mine, hcid[k]);
                    last = (uint32_t)block_min_u64<THREADS>(mine, red);
                    int myq = -1;
#pragma unroll
                    for (int k = 0; k < HPT; ++k)
                        if (has(k) && hcid[k] == last) myq = qof(k);
                    if (myq >= 0) emit(r, last, myq);
                }
                if (t == 0) {
                    a.out.nbest[sidx] = take;
                    a.out.nbig[sidx] = total;
                }
            }
        }
        __syncthreads();
        // clear the bitmaps and, when it was used, the table
#pragma unroll
        for (int i = t; i < 2 * kWords; i += THREADS) bm[i] = 0;
        if constexpr (kOwnSep)
            for (int i = t; i < CAP / 2; i += THREADS) own_sep[i] = 0;
        if (ncand_all > 0 && ncand_all <= SMALLCAP) {
            for (int i = t; i < SMALL; i += THREADS) {
                skey[i] = kEmpty;
                scnt[i] = 0;
                smin[i] = 0xFFFFFFFFu;
            }
        }
        __syncthreads();
    }
}

// strands past kMidCap hits: one 256-thread workgroup each, keys in a
// power-of-two global scratch slot
__global__ void __launch_bounds__(256) map_strands_big(MapArgs a) {
    __shared__ unsigned char qb[kMaxMapRead + 1];
    __shared__ int32_t ex[64], stt[64];
    __shared__ int32_t sh_total, sh_ok;
    const int n = *a.big_count;
    uint64_t *kbuf = a.scratch + (int64_t)blockIdx.x * a.scratch_cap;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int sidx = a.big_list[it];
        int L;
        load_strand(a, sidx, qb, &L);
        __syncthreads();
        if (threadIdx.x < 64) {
            bool ok;
            int total = gather_hits(a.ix, qb, L - a.ix.k + 1, kbuf, a.scratch_cap, ex, stt, &ok);
            if (threadIdx.x == 0) {
                sh_total = total;
                sh_ok = ok;
            }
        }
        __syncthreads();
        const int total = sh_total;
        if (!sh_ok) {  // beyond the scratch slot: reported as a capacity error
            if (threadIdx.x == 0) {
                atomicOr(a.error, 1);
                a.out.nbest[sidx] = 0;
                a.out.nbig[sidx] = 0;
            }
            __syncthreads();
            continue;
        }
        const int n2 = next_pow2(total);
        for (int i = total + threadIdx.x; i < n2; i += 256) kbuf[i] = ~0ull;
        __syncthreads();
        bitonic<false>(kbuf, n2, threadIdx.x, 256, [] { __syncthreads(); });
        if (threadIdx.x < 64) emit_best(a, sidx, kbuf, total);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Records: per-strand prefix, hit -> record, region window per record
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) hits_to_records(MapArgs m, PairsDev p) {
    const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= 2 * m.rd.nreads) return;
    const int nb = m.out.nbest[sidx];
    const int at = p.pair0[sidx];
    const int64_t o = (int64_t)sidx * m.out.stride;
    for (int h = 0; h < nb; ++h) {
        p.read[at + h] = sidx >> 1;
        p.strand[at + h] = sidx & 1;
        p.ref[at + h] = m.out.ref[o + h];
        p.shift[at + h] = m.out.shift[o + h];
        p.pos[at + h] = m.out.pos[o + h];
        p.nbig[at + h] = m.out.nbig[sidx];
    }
}

__device__ __forceinline__ bool is_n(char c) { return c == 'N' || c == 'n'; }

// Window of record i: refDistance = max(0, refPos - before), up to win
// bases clipped at the sequence end (alignQueries_cpu_mt, src/align.cu
// :2388,2435); leading N/n stripped, trailing N/n stripped only when no
// leading strip happened, refDistance left unchanged (getRefSeq2,
// src/input.cu:348-403).  Slots past the record count get empty pairs.
__global__ void __launch_bounds__(256) build_regions(MapArgs m, PairsDev p, int before, int win) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.capacity) return;
    p.coff[i] = (int64_t)i * p.cstride;
    if (i >= *p.npairs) {
        p.qlen[i] = 0;
        p.rlen[i] = 0;
        p.qoff[i] = 0;
        p.roff[i] = 0;
        return;
    }
    const int s = p.ref[i];
    const int64_t slen = m.ix.seq_len[s];
    int64_t rd = (int64_t)p.pos[i] - before;
    if (rd < 0) rd = 0;
    int64_t n = slen - rd;
    if (n > win) n = win;
    if (n < 0) n = 0;
    const char *w = m.ix.seq + m.ix.seq_off[s] + rd;
    int64_t lead = 0;
    while (lead < n && is_n(w[lead])) ++lead;
    if (lead == 0)
        while (n > 0 && is_n(w[n - 1])) --n;
    const int sidx = 2 * p.read[i] + p.strand[i];
    p.rdist[i] = rd;
    p.roff[i] = m.ix.seq_off[s] + rd + lead;
    p.rlen[i] = (int32_t)(n - lead);
    p.qoff[i] = (int64_t)sidx * m.qstride;
    p.qlen[i] = m.rd.len[p.read[i]];
}

// chooseBestAlignment (src/align.cu:844-905) per read: the maximum starts at
// 0 and rises on a strict <, the records equal to it form the tie set, the
// first of them is kept (the reference keeps the rand() % j-th).
__global__ void __launch_bounds__(256) choose_best(PairsDev p, int nreads, const int32_t *score,
                                                   int32_t *keep, int32_t *ntied) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreads) return;
    const int a = p.pair0[2 * r], b = p.pair0[2 * r + 2];
    int mx = 0, j = 0;
    for (int i = a; i < b; ++i) mx = max(mx, score[i]);
    for (int i = a; i < b; ++i) j += score[i] == mx;
    bool done = false;
    for (int i = a; i < b; ++i) {
        bool k = !done && score[i] == mx;
        keep[i] = k;
        done |= k;
        ntied[i] = j;
    }
}

// ---------------------------------------------------------------------------
// Serving readback
// ---------------------------------------------------------------------------
struct LoadLen64 {
    const int32_t *p;
    __device__ int64_t operator()(int i) const { return p[i]; }
};

// One served batch straight into pinned host memory (serve_layout): the
// columns and distances with consecutive threads on consecutive entries (so
// the writes across the bus are whole lines), then the strings, one wave per
// record.  One kernel and one synchronisation instead of a copy per array.
__global__ void __launch_bounds__(256) serve_out(ServeOut o) {
    const int n = *o.nrec;
    const int64_t packed = n > 0 ? *o.total : 0;
    const ServeLayout l = serve_layout(n, packed);
    const int gt = blockIdx.x * 256 + (int)threadIdx.x, G = gridDim.x * 256;
    if (gt == 0) {
        *reinterpret_cast<int32_t *>(o.out) = n;
        *reinterpret_cast<int64_t *>(o.out + 8) = packed;
        reinterpret_cast<int32_t *>(o.out + 16)[0] = o.flags[0];
        reinterpret_cast<int32_t *>(o.out + 16)[1] = o.flags[1];
    }
    int32_t *cols = reinterpret_cast<int32_t *>(o.out + l.o_col);
#pragma unroll
    for (int c = 0; c < kServeCols; ++c)
        for (int k = gt; k < n; k += G) cols[(int64_t)c * n + k] = o.col[c][k];
    int64_t *rd = reinterpret_cast<int64_t *>(o.out + l.o_rd);
    for (int k = gt; k < n; k += G) rd[k] = o.rdist[k];
    const int lane = threadIdx.x & 63;
    const int32_t *alen = o.col[o.alen_col];
    char *oq = o.out + l.o_q, *orr = o.out + l.o_r;
    for (int i = gt >> 6; i < n; i += G >> 6) {
        const int L = alen[i];
        const char *q = o.qc + (int64_t)i * o.cs, *r = o.rc + (int64_t)i * o.cs;
        char *dq = oq + o.pos[i], *dr = orr + o.pos[i];
        for (int k = lane; k < L; k += 64) {
            dq[k] = q[k];
            dr[k] = r[k];
        }
    }
}

hipError_t launch_serve_out(const ServeOut &o, int cap, hipStream_t st) {
    hipError_t e = launch_scan<int64_t>(LoadLen64{o.col[o.alen_col]}, StoreArr<int64_t>{o.pos},
                                        o.nrec, cap, o.sums, nullptr, o.total, nullptr, st);
    if (e != hipSuccess) return e;
    const int blocks = std::max(1, std::min((cap + 3) / 4, 1024));
    hipLaunchKernelGGL(serve_out, dim3(blocks), dim3(256), 0, st, o);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// mapHits6GetBestHits over caller hits (drop-in; one workgroup)
// ---------------------------------------------------------------------------
struct HitTriple {
    uint64_t hi;  // (refIdx + 128) << 32 | (shift ^ 0x80000000)
    uint32_t lo;  // refPos ^ 0x80000000
    uint32_t pad;
    __device__ bool operator<(const HitTriple &o) const {
        return hi < o.hi || (hi == o.hi && lo < o.lo);
    }
};
struct TripleCluster {
    __device__ uint64_t operator()(const HitTriple &t) const { return t.hi; }
};

__global__ void __launch_bounds__(256)
best_hits_kernel(int8_t *ref, int32_t *shift, int32_t *pos, int n, int max_hits, int8_t *o_ref,
                 int32_t *o_shift, int32_t *o_pos, int32_t *o_n, int32_t *o_nbig,
                 HitTriple *key) {
    const int n2 = next_pow2(n);
    for (int i = threadIdx.x; i < n2; i += 256) {
        HitTriple t;
        if (i < n) {
            t.hi = ((uint64_t)(uint32_t)(ref[i] + 128) << 32) | ((uint32_t)shift[i] ^ 0x80000000u);
            t.lo = (uint32_t)pos[i] ^ 0x80000000u;
        } else {
            t.hi = ~0ull;
            t.lo = ~0u;
        }
        t.pad = 0;
        key[i] = t;
    }
    __syncthreads();
    bitonic<false>(key, n2, threadIdx.x, 256, [] { __syncthreads(); });
    for (int i = threadIdx.x; i < n; i += 256) {
        HitTriple t = key[i];
        ref[i] = (int8_t)((int)(t.hi >> 32) - 128);
        shift[i] = (int32_t)((uint32_t)t.hi ^ 0x80000000u);
        pos[i] = (int32_t)(t.lo ^ 0x80000000u);
    }
    if (threadIdx.x < 64)
        wave_best_clusters(
            key, n, max_hits, TripleCluster{},
            [&](int slot, const HitTriple &t) {
                o_ref[slot] = (int8_t)((int)(t.hi >> 32) - 128);
                o_shift[slot] = (int32_t)((uint32_t)t.hi ^ 0x80000000u);
                o_pos[slot] = (int32_t)(t.lo ^ 0x80000000u);
            },
            o_n, o_nbig);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
static int tuple_grid(int64_t ntup) {
    int64_t g = (ntup + 255) / 256;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 1 << 16);
}

hipError_t launch_index_build(const IndexDev &ix, int64_t ntup, const IndexBuildScratch &s,
                              hipStream_t st) {
    if (ntup > 0) hipLaunchKernelGGL(idx_count, dim3(tuple_grid(ntup)), dim3(256), 0, st, ix, ntup);
    const int nblocks = (int)((ix.nhash + kScanChunk - 1) / kScanChunk);
    hipLaunchKernelGGL(idx_scan_blocks, dim3(nblocks), dim3(1024), 0, st, ix, s);
    hipLaunchKernelGGL(idx_scan_sums, dim3(1), dim3(1024), 0, st, s, nblocks);
    hipLaunchKernelGGL(idx_scan_apply, dim3(nblocks), dim3(1024), 0, st, ix, s);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) idx_gpos(IndexDev ix, int64_t total, uint32_t *gpos) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i <= total; i += (int64_t)gridDim.x * 256) {
        const uint64_t e = ix.entry[i];
        gpos[i] = i == 0 ? 0xFFFFFFFFu : ix.cid0[(int)(e >> 32)] + (uint32_t)e;
    }
}

hipError_t launch_index_gpos(const IndexDev &ix, int64_t total, uint32_t *gpos, hipStream_t st) {
    hipLaunchKernelGGL(idx_gpos, dim3(tuple_grid(total + 1)), dim3(256), 0, st, ix, total, gpos);
    return hipGetLastError();
}

hipError_t launch_index_fill(const IndexDev &ix, int64_t ntup, const IndexBuildScratch &s,
                             hipStream_t st) {
    if (ntup > 0) hipLaunchKernelGGL(idx_fill, dim3(tuple_grid(ntup)), dim3(256), 0, st, ix, ntup);
    hipLaunchKernelGGL(idx_sort_small, dim3((unsigned)((ix.nhash + 255) / 256)), dim3(256), 0, st,
                       ix, s);
    hipLaunchKernelGGL(idx_sort_mid, dim3(8192), dim3(64), 0, st, ix, s);
    if (s.big_slots > 0)
        hipLaunchKernelGGL(idx_sort_big, dim3(s.big_slots), dim3(256), 0, st, ix, s);
    return hipGetLastError();
}

// Sets *flag when any byte of seq[0, n) is outside ACGTNacgtn (b | 0x20 is
// one of acgtn exactly for those ten bytes); 16 bytes a thread per step.
__global__ void __launch_bounds__(256) seq_plain_scan(const char *seq, int64_t n, int32_t *flag) {
    const int64_t n16 = n >> 4;
    const uint4 *v = reinterpret_cast<const uint4 *>(seq);
    auto bad = [](uint32_t w) {
        uint32_t any = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t c = ((w >> (8 * b)) & 0xFFu) | 0x20u;
            any |= (c != 'a' && c != 'c' && c != 'g' && c != 't' && c != 'n') ? 1u : 0u;
        }
        return any;
    };
    uint32_t found = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) {
        const uint4 x = v[i];
        found |= bad(x.x) | bad(x.y) | bad(x.z) | bad(x.w);
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 15)) {
        const uint32_t c = (uint32_t)(unsigned char)seq[(n16 << 4) + threadIdx.x] | 0x20u;
        found |= (c != 'a' && c != 'c' && c != 'g' && c != 't' && c != 'n') ? 1u : 0u;
    }
    if (found) atomicOr(flag, 1);
}

hipError_t launch_seq_plain_scan(const char *seq, int64_t n, int32_t *flag, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(((n >> 4) + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL(seq_plain_scan, dim3((unsigned)blocks), dim3(256), 0, st, seq, n, flag);
    return hipGetLastError();
}

hipError_t launch_map(const MapArgs &a, hipStream_t st, int mid_blocks) {
    if (a.rd.nreads <= 0) return hipSuccess;
    const bool hash_first = a.hash_first && mid_blocks > 0 && a.ix.cid_ok;
    // the tier kernels stride over strands (every one, or a list of some):
    // no more workgroups than strands (small batches; each surplus
    // workgroup would still clear its LDS tables)
    const int nstr = 2 * a.rd.nreads;
    auto grid = [nstr](int g) { return dim3((unsigned)std::max(1, std::min(g, nstr))); };
    if (!hash_first) hipLaunchKernelGGL(map_strands, dim3(2 * a.rd.nreads), dim3(64), 0, st, a);
    if (mid_blocks > 0) {
        if (a.ix.cid_ok && hash_first && a.max_len <= kShortRead) {
            // short reads, every strand: first the half-size table (four
            // workgroups per CU), strands past its cap on to the full one
            // through the wave tier's (here unused) overflow list
            constexpr int kWpe = 8;  // waves per SIMD: kWpe / 2 workgroups of 8 waves per CU
            hipLaunchKernelGGL(
                (map_strands_filter<512, kHashCap, 1024, 768, kShortRead, kWpe, false>),
                grid(kWpe / 2 * mid_blocks), dim3(512), 0, st, a, nullptr, nullptr, a.ovl_list,
                a.ovl_count, a.mid_list, a.mid_count);
            hipLaunchKernelGGL((map_strands_hash<512, kHashSlots, kHashCap, kShortRead, 4>),
                               grid(2 * mid_blocks), dim3(512), 0, st, a, a.ovl_list, a.ovl_count,
                               a.mid_list, a.mid_count);
            // strands past kHashCap hits (reads near the short-read limit):
            // the filter stage at 1024 threads, one workgroup per CU; what it
            // cannot take goes on to the workgroup sort
            hipLaunchKernelGGL(
                (map_strands_filter<1024, kLongFilterCap, 2048, 1536, kShortRead, 4, true>),
                grid(mid_blocks / 2), dim3(1024), 0, st, a, a.mid_list, a.mid_count,
                a.mid2_list, a.mid2_count, a.mid2_list, a.mid2_count);
            hipLaunchKernelGGL((map_strands_wg<1024, kMidCap>), grid(mid_blocks), dim3(1024), 0,
                               st, a, a.mid2_list, a.mid2_count, a.big_list, a.big_count);
            if (a.scratch_slots > 0)
                hipLaunchKernelGGL(map_strands_big, grid(a.scratch_slots), dim3(256), 0, st, a);
            return hipGetLastError();
        } else if (a.ix.cid_ok) {
            hipLaunchKernelGGL((map_strands_hash<512, kHashSlots, kHashCap, kMaxMapRead, 4>),
                               grid(2 * mid_blocks), dim3(512), 0, st, a,
                               hash_first ? nullptr : a.ovl_list,
                               hash_first ? nullptr : a.ovl_count, a.mid_list, a.mid_count);
        } else
            hipLaunchKernelGGL((map_strands_wg<256, kMidCap1>), grid(5 * mid_blocks), dim3(256),
                               0, st, a, a.ovl_list, a.ovl_count, a.mid_list, a.mid_count);
        hipLaunchKernelGGL((map_strands_wg<1024, kMidCap>), grid(mid_blocks), dim3(1024), 0, st,
                           a, a.mid_list, a.mid_count, a.big_list, a.big_count);
    }
    if (a.scratch_slots > 0)
        hipLaunchKernelGGL(map_strands_big, grid(a.scratch_slots), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_records(const MapArgs &m, const PairsDev &p, int before, int win,
                          hipStream_t st) {
    const int ns = 2 * m.rd.nreads;
    hipError_t e = launch_scan<int32_t>(LoadArr<int32_t>{m.out.nbest}, StoreArr<int32_t>{p.pair0},
                                        nullptr, ns, p.scan_sums, nullptr, p.pair0 + ns, p.npairs,
                                        st);
    if (e != hipSuccess) return e;
    if (ns > 0)
        hipLaunchKernelGGL(hits_to_records, dim3((ns + 255) / 256), dim3(256), 0, st, m, p);
    if (p.capacity > 0)
        hipLaunchKernelGGL(build_regions, dim3((p.capacity + 255) / 256), dim3(256), 0, st, m, p,
                           before, win);
    return hipGetLastError();
}

hipError_t launch_choose_best(const PairsDev &p, int nreads, const int32_t *score, int32_t *keep,
                              int32_t *ntied, hipStream_t st) {
    if (nreads > 0)
        hipLaunchKernelGGL(choose_best, dim3((nreads + 255) / 256), dim3(256), 0, st, p, nreads,
                           score, keep, ntied);
    return hipGetLastError();
}

hipError_t launch_best_hits(int8_t *ref, int32_t *shift, int32_t *pos, int n, int max_hits,
                            int8_t *o_ref, int32_t *o_shift, int32_t *o_pos, int32_t *o_n,
                            int32_t *o_nbig, void *kbuf, hipStream_t st) {
    hipLaunchKernelGGL(best_hits_kernel, dim3(1), dim3(256), 0, st, ref, shift, pos, n, max_hits,
                       o_ref, o_shift, o_pos, o_n, o_nbig, static_cast<HitTriple *>(kbuf));
    return hipGetLastError();
}

}  // namespace swift
