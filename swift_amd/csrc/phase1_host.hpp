// phase1_host.hpp -- host-side objects of phase 1 shared by the entry points
// (phase1_host.cpp) and the serving loop (server.cpp).
#pragma once

#include <string>
#include <vector>

#include "../../include/swift_abi.h"
#include "phase1.hpp"
#include "sw_host.hpp"

namespace swift {

// Exact-size device array, freed with its owner.
struct DevArr {
    void *ptr = nullptr;
    size_t bytes = 0;
    int alloc(size_t b) {
        release();
        SWIFT_HIP_TRY(hipMalloc(&ptr, b ? b : 1));
        bytes = b;
        return kOk;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
    DevArr() = default;
    DevArr(const DevArr &) = delete;
    DevArr &operator=(const DevArr &) = delete;
    ~DevArr() { release(); }
};

// Device-resident k-mer index.
struct IndexObj {
    int k = 0, thr = 0, nseq = 0, max_actual = 0;
    int64_t nhash = 0, ntup = 0, total = 0;
    DevArr seq, seq_off, seq_len, tup0, count, start, actual, cursor, bucket, entry, cid0, gpos;
    int cid_ok = 0;
    int plain_bases = 0;  // every base one of ACGTNacgtn
    std::vector<int64_t> h_off, h_len;
    std::vector<std::string> names;

    IndexDev dev() const {
        IndexDev d;
        d.seq = seq.as<const char>();
        d.seq_off = seq_off.as<const int64_t>();
        d.seq_len = seq_len.as<const int64_t>();
        d.tup0 = tup0.as<const int64_t>();
        d.nseq = nseq;
        d.k = k;
        d.thr = thr;
        d.nhash = nhash;
        d.count = count.as<int32_t>();
        d.start = start.as<int32_t>();
        d.actual = actual.as<int32_t>();
        d.cursor = cursor.as<int32_t>();
        d.bucket = bucket.as<int2>();
        d.entry = entry.as<uint64_t>();
        d.cid0 = cid0.as<const uint32_t>();
        d.cid_ok = cid_ok;
        d.plain_bases = plain_bases;
        d.gpos = gpos.as<const uint32_t>();
        return d;
    }
};

// Mapping / pipeline scratch (grows, never shrinks).
struct MapWork {
    DevArr ref, shift, pos, nbest, nbig, nhits, strand_seq, ovl_list, mid_list, mid2_list,
        big_list, flags, scratch;
    DevArr pair0, qoff, roff, coff, qlen, rlen, scan_sums;
};

int ensure(DevArr &a, size_t bytes);

// Phase 1 -> phase 2 for one read batch with explicit scratch: `sws` is a
// private batched-SW workspace (nullptr: the shared one).  Not locked: one
// caller per MapWork.  Synchronises `st`.
int align_reads_in(MapWork &w, Workspace *sws, const IndexObj &ix, const ReadsDev &rd,
                   int max_read_len, int max_hits, const Scoring &sc, const swift_records &rec,
                   hipStream_t st);
// The same in steps: align_prepare sizes scratch and fills the launch
// arguments; align_enqueue launches (sync = false: no host synchronisation
// or allocation, sws required and reserved, capturable); map_check reads the
// mapping error flags afterwards.
int align_prepare(MapWork &w, const IndexObj &ix, const ReadsDev &rd, int max_read_len,
                  int max_hits, const swift_records &rec, MapArgs *m, PairsDev *p);
int align_enqueue(const MapArgs &m, const PairsDev &p, Workspace *sws, int nreads, int max_read_len,
                  const Scoring &sc, const swift_records &rec, hipStream_t st, bool sync);
int map_check(const MapArgs &m, hipStream_t st);
// the status of the mapping error flags word (map_check without the readback)
int map_error(int flags);
// device bytes align_reads_in may hold for nreads reads of <= max_read_len
int64_t align_bytes_estimate(int nreads, int max_read_len, int max_hits);

}  // namespace swift

struct swift_index {
    swift::IndexObj ix;
};
