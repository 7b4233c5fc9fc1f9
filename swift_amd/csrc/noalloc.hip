// noalloc.hip -- swift_align_sequences_noalloc: the plain per-pair path with
// its traceback scratch in a fixed device array instead of library-allocated
// memory, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The plain path keeps each pair's 2-bit traceback, best cell and walk
// moves in scratch between its kernels.  Here that scratch is `scratch`, a
// file-scope __device__ array of kNoAllocWords 32-bit words (64 MiB) that
// each GPU holds until the process ends.  Its device address is looked up
// once per device with
// hipGetSymbolAddress and kept; kernels get that address, never the array's
// host-side name.  Pairs run in chunks that fit the array, one after another
// on the caller's stream.  Every call on a GPU shares that GPU's array, so
// calls whose kernels overlap on one GPU race: issue them on one stream, or
// wait for one before issuing the next.
#include <mutex>

#include "../../include/swift_abi.h"
#include "sw_host.hpp"
#include "sw_types.hpp"

namespace swift {
namespace {

constexpr int64_t kNoAllocWords = (int64_t)16 << 20;  // 64 MiB of trace words
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint32_t scratch[kNoAllocWords];

std::mutex g_sym_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's `scratch`, looked up on the first call
// made with the device current.
int scratch_words(uint32_t **out) {
    int dev = 0;
    SWIFT_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d devices this entry point supports", dev, kMaxDevices);
        return kErrInvalid;
    }
    std::lock_guard<std::mutex> lk(g_sym_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        SWIFT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the scratch array");
            return kErrHip;
        }
        g_scratch_ptr[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch_ptr[dev];
    return kOk;
}

}  // namespace
}  // namespace swift

using namespace swift;

extern "C" {

int swift_align_sequences_noalloc(char *querySeq_d, char *refSeq_d, int querySeqLength,
                                  int refSeqLength, float match, float mismatch,
                                  float gapOpenPenalty, float gapExtPenalty, float *score_d,
                                  int *alignStart_d, int *alignEnd_d, int *alignLength_d,
                                  char *queryConsensus_d, char *refConsensus_d,
                                  int queriesPerIteration, void *stream) {
    Scoring sc;
    int rc = scoring_from_floats(match, mismatch, gapOpenPenalty, gapExtPenalty, &sc);
    if (rc) return rc;
    const int qlen = querySeqLength, rlen = refSeqLength, npairs = queriesPerIteration;
    if (npairs < 0 || qlen < 0 || rlen < 0) {
        set_error("negative sizes");
        return kErrInvalid;
    }
    if (plain_rows_for(qlen) < 0 || rlen > kMaxCols - 64) {
        set_error("query length %d / reference length %d beyond the plain kernel (<= 1024 / %d)",
                  qlen, rlen, kMaxCols - 64);
        return kErrCapacity;
    }
    bool f32;
    if ((rc = validate_scoring(sc, qlen, rlen, &f32))) return rc;
    if (npairs == 0) return kOk;
    // words per pair: the trace, the best cell and the walk's moves
    const int64_t wpp = plain_scratch_words(qlen, rlen);
    if (wpp > kNoAllocWords) {
        set_error("one pair's traceback needs %lld words, beyond the %lld-word scratch",
                  (long long)wpp, (long long)kNoAllocWords);
        return kErrCapacity;
    }
    uint32_t *trace = nullptr;
    if ((rc = scratch_words(&trace))) return rc;
    const int chunk = (int)std::min<int64_t>(npairs, kNoAllocWords / wpp);
    hipStream_t st = (hipStream_t)stream;
    for (int c0 = 0; c0 < npairs; c0 += chunk) {
        const int cn = std::min(chunk, npairs - c0);
        PlainOut out{score_d, alignStart_d, alignEnd_d, alignLength_d, queryConsensus_d,
                     refConsensus_d, (int64_t)qlen + rlen + 2};
        SWIFT_HIP_TRY(launch_plain(querySeq_d, refSeq_d, qlen, rlen, sc, cn, c0, trace, out, st));
    }
    return kOk;
}

}  // extern "C"
