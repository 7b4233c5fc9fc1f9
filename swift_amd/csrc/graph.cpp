// graph.cpp -- phase 1 -> phase 2 captured as one hipGraph (no reference
// counterpart).  The pipeline for a fixed batch shape (nreads read slots of
// <= max_read_len bases, fixed device buffers) is recorded once with stream
// capture and replayed with one hipGraphLaunch: mapping, record/window
// construction, binning, every row-class sweep, the walk-back and the
// best-hit choice, with no host round trip.  Unused read slots hold
// length-0 reads (no hits, no records).  Scratch is reserved for the worst
// case at capture time; counters and flags are cleared by fill kernels
// inside the graph (no memset nodes).
#include "phase1_host.hpp"
#include "sw_types.hpp"

namespace swift {

struct Graph {
    const IndexObj *ix = nullptr;
    MapWork mw;
    Workspace *sws = nullptr;
    hipStream_t cap_st = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    MapArgs m;
    hipStream_t last = nullptr;
    ~Graph() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        if (cap_st) (void)hipStreamDestroy(cap_st);
        workspace_free(sws);
    }
};

}  // namespace swift

using namespace swift;

struct swift_graph {
    Graph gr;
};

extern "C" {

int swift_pipeline_graph_create(const swift_index *ix, const char *reads_d, const int64_t *off_d,
                                const int32_t *len_d, int nreads, int max_read_len, int max_hits,
                                int match, int mismatch, int gap_open, int gap_ext,
                                const swift_records *rec, swift_graph **out) {
    if (!ix || !rec || !out || nreads < 1 || max_hits < 1 || !reads_d || !off_d || !len_d) {
        set_error("invalid graph arguments");
        return kErrInvalid;
    }
    Scoring sc{match, mismatch, gap_open, gap_ext};
    bool f32;
    int rc = validate_scoring(sc, max_read_len, 300, &f32);
    if (rc) return rc;
    auto *h = new swift_graph;
    Graph &G = h->gr;
    G.ix = &ix->ix;
    G.sws = workspace_new();
    ReadsDev rd{reads_d, off_d, len_d, nreads};
    PairsDev p;
    if ((rc = align_prepare(G.mw, *G.ix, rd, max_read_len, max_hits, *rec, &G.m, &p)) ||
        (rc = reserve_batch_in(G.sws, p.capacity, max_read_len, 300))) {
        delete h;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&G.cap_st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(G.cap_st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("stream capture: %s", hipGetErrorString(e));
        delete h;
        return kErrHip;
    }
    rc = align_enqueue(G.m, p, G.sws, nreads, max_read_len, sc, *rec, G.cap_st, false);
    e = hipStreamEndCapture(G.cap_st, &G.g);
    if (rc == kOk && e != hipSuccess) {
        set_error("stream capture: %s", hipGetErrorString(e));
        rc = kErrHip;
    }
    if (rc == kOk) {
        e = hipGraphInstantiate(&G.ge, G.g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
            rc = kErrHip;
        }
    }
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return kOk;
}

int swift_pipeline_graph_launch(swift_graph *h, void *stream) {
    if (!h) {
        set_error("no graph");
        return kErrInvalid;
    }
    SWIFT_HIP_TRY(hipGraphLaunch(h->gr.ge, (hipStream_t)stream));
    h->gr.last = (hipStream_t)stream;
    return kOk;
}

int swift_pipeline_graph_check(swift_graph *h) {
    if (!h) {
        set_error("no graph");
        return kErrInvalid;
    }
    int rc = map_check(h->gr.m, h->gr.last);
    if (rc) return rc;
    return sw_batch_check_in(h->gr.sws, h->gr.last);
}

int swift_pipeline_graph_free(swift_graph *h) {
    if (h) {
        (void)hipDeviceSynchronize();
        delete h;
    }
    return kOk;
}

}  // extern "C"
