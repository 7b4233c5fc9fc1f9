// output.hip -- CIGAR strings of pipeline records on the device, with the
// rule of printCigar (src/output.cu:33-79): walking the walk-order strings
// from the end (the alignment's start), runs of D ('-' in the reference), I
// ('-' in the query) and M, each written as <length><op>.
#include "../../include/swift_abi.h"
#include "swift_common.hpp"

namespace swift {

__device__ __forceinline__ int put_run(char *o, int n, int run, char op) {
    char d[8];
    int k = 0;
    do {
        d[k++] = (char)('0' + run % 10);
        run /= 10;
    } while (run);
    while (k) o[n++] = d[--k];
    o[n++] = op;
    return n;
}

// one thread per record i < *n_dev
__global__ void __launch_bounds__(256)
cigar_kernel(const int32_t *n_dev, const int32_t *alen, const char *qc, const char *rc,
             int64_t cs, char *out, int64_t os, int32_t *olen) {
    const int i = blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= *n_dev) return;
    const char *q = qc + (int64_t)i * cs, *r = rc + (int64_t)i * cs;
    char *o = out + (int64_t)i * os;
    int n = 0, run = 0;
    char op = ' ';
    for (int z = alen[i] - 1; z >= 0; --z) {
        const char c = r[z] == '-' ? 'D' : q[z] == '-' ? 'I' : 'M';
        if (run > 0 && op != c) {
            n = put_run(o, n, run, op);
            run = 0;
        }
        op = c;
        ++run;
    }
    if (run > 0) n = put_run(o, n, run, op);
    o[n] = '\0';
    olen[i] = n;
}

}  // namespace swift

using namespace swift;

extern "C" int swift_records_cigar(const swift_records *rec, int max_records, char *cigar_d,
                                   int64_t cigar_stride, int32_t *cigar_len_d, void *stream) {
    if (!rec || max_records < 0 || (max_records > 0 && (!cigar_d || !cigar_len_d))) {
        set_error("invalid arguments");
        return kErrInvalid;
    }
    // a record of alen <= cons_stride - 2 columns gives at most 2 * alen
    // characters plus the NUL
    if (cigar_stride < 2 * rec->cons_stride) {
        set_error("cigar_stride %lld below 2 * cons_stride", (long long)cigar_stride);
        return kErrInvalid;
    }
    if (max_records == 0) return kOk;
    hipLaunchKernelGGL(cigar_kernel, dim3((max_records + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, rec->nrecords, rec->alen, rec->qry_cons, rec->ref_cons,
                       rec->cons_stride, cigar_d, cigar_stride, cigar_len_d);
    SWIFT_HIP_TRY(hipGetLastError());
    return kOk;
}
