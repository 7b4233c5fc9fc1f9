// output.cpp -- alignment output in the reference's formats: the
// printOutput / printOutputSam / printCigar drop-ins (src/output.cu:33-201)
// and a writer of pipeline records.  Lines are assembled in memory and
// written with one fwrite each.
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "output.hpp"
#include "phase1_host.hpp"

namespace swift {

constexpr int kNameStride = 50;  // MAX_SEQ_NAME_LENGTH (src/common.h:37-39)

// printCigar (src/output.cu:33-79): from the end of the walk-order strings
// (the start of the alignment), runs of D ('-' in the reference), I ('-' in
// the query) and M.
void cigar_string(const char *qc, const char *rc, int alen, std::string *out) {
    char op = ' ';
    int run = 0;
    char num[16];
    for (int z = alen - 1; z >= 0; --z) {
        const char c = rc[z] == '-' ? 'D' : qc[z] == '-' ? 'I' : 'M';
        if (run > 0 && op != c) {
            snprintf(num, sizeof(num), "%d%c", run, op);
            *out += num;
            run = 0;
        }
        op = c;
        ++run;
    }
    if (run > 0) {
        snprintf(num, sizeof(num), "%d%c", run, op);
        *out += num;
    }
}

static std::string bounded(const char *s, int cap) {
    size_t n = 0;
    while (n < (size_t)cap && s[n]) ++n;
    return std::string(s, n);
}

// One printOutput record (src/output.cu:154-201).
void text_record(const std::string &qn, const std::string &rn, float score, int start, int end,
                 int alen, const char *qc, const char *rc, std::string *out) {
    char head[512];
    snprintf(head, sizeof(head),
             "\n****************************************\n"
             "Qry: %s | Ref: %s | Score: %.1f | AlignStart: %d | AlignEnd: %d | "
             "AlignLength: %d\n",
             qn.c_str(), rn.c_str(), (double)score, start, end, alen);
    *out += head;
    *out += "R: ";
    for (int z = alen - 1; z >= 0; --z) *out += rc[z];
    *out += "\n   ";
    for (int z = alen - 1; z >= 0; --z) {
        const int r = (unsigned char)rc[z], q = (unsigned char)qc[z];
        if (toupper(r) == toupper(q) || r == 'N' || q == 'N') *out += '|';
        else if (r == '-' || q == '-') *out += '-';
        else *out += '.';
    }
    *out += "\nQ: ";
    for (int z = alen - 1; z >= 0; --z) *out += qc[z];
    *out += "\n";
}

// One printOutputSam record (src/output.cu:98-130).
void sam_record(const std::string &qn, const std::string &rn, int start, int alen, const char *qc,
                const char *rc, std::string *out) {
    char head[256];
    snprintf(head, sizeof(head), "%s\t*\t%s\t%d\t255\t", qn.c_str(), rn.c_str(), start);
    *out += head;
    cigar_string(qc, rc, alen, out);
    *out += "\t*\t0\t0\t";
    for (int z = alen - 1; z >= 0; --z) *out += qc[z];
    *out += "\t*\n";
}

}  // namespace swift

using namespace swift;

extern "C" {

void printCigar(char *queryConsensus, char *refConsensus, int alignLength, FILE *outputFilePtr) {
    std::string s;
    cigar_string(queryConsensus, refConsensus, alignLength, &s);
    fwrite(s.data(), 1, s.size(), outputFilePtr);
}

void printOutputSam(int numQueries, char *queryName, char *refName, float *score,
                    int *alignStart, int *alignEnd, int *alignLength, int combinedLength,
                    char *queryConsensus, char *refConsensus, FILE *outputFilePtr) {
    (void)score;
    (void)alignEnd;
    std::string s;
    for (int y = 0; y < numQueries; ++y) {
        s.clear();
        sam_record(bounded(queryName + y * kNameStride, kNameStride),
                   bounded(refName + y * kNameStride, kNameStride), alignStart[y],
                   alignLength[y], queryConsensus + (int64_t)y * combinedLength,
                   refConsensus + (int64_t)y * combinedLength, &s);
        fwrite(s.data(), 1, s.size(), outputFilePtr);
    }
}

void printOutput(int numQueries, char *queryName, char *refName, float *score, int *refDistance,
                 int *alignStart, int *alignEnd, int *alignLength, int combinedLength,
                 char *queryConsensus, char *refConsensus, FILE *outputFilePtr) {
    std::string s;
    for (int y = 0; y < numQueries; ++y) {
        if (score[y] < 0.0f) continue;  // records dropped by the best-hit choice
        s.clear();
        text_record(bounded(queryName + y * kNameStride, kNameStride),
                    bounded(refName + y * kNameStride, kNameStride), score[y],
                    alignStart[y] + refDistance[y], alignEnd[y] + refDistance[y], alignLength[y],
                    queryConsensus + (int64_t)y * combinedLength,
                    refConsensus + (int64_t)y * combinedLength, &s);
        fwrite(s.data(), 1, s.size(), outputFilePtr);
    }
}

// Pipeline records (host copies, swift_records layout with host pointers) to
// a file: format 0 = printOutput text, 1 = SAM.  The query of a record is
// qry_names[read] with '*' appended on the reverse strand and its reference
// the index sequence's FASTA name, as the reference drivers name them
// (src/align.cu:2437-2452).  only_kept: records chooseBestAlignment keeps
// (the GPU driver), else all (the CPU driver prints every hit).
int swift_write_records(const char *path, int format, int only_kept, const swift_records *rec,
                        int nrec, const char *const *qry_names, const swift_index *ix) {
    if (!path || !rec || !ix || nrec < 0 || (nrec > 0 && !qry_names)) {
        set_error("invalid arguments");
        return kErrInvalid;
    }
    FILE *f = fopen(path, "w");
    if (!f) {
        set_error("cannot write %s", path);
        return kErrIO;
    }
    std::string s;
    const IndexObj &X = ix->ix;
    for (int i = 0; i < nrec; ++i) {
        if (only_kept && !rec->keep[i]) continue;
        const int ri = rec->ref_idx[i];
        std::string qn = qry_names[rec->read[i]];
        if (rec->strand[i]) qn += "*";
        const std::string rn = ri >= 0 && ri < X.nseq ? X.names[ri] : std::string();
        const char *qc = rec->qry_cons + (int64_t)i * rec->cons_stride;
        const char *rc = rec->ref_cons + (int64_t)i * rec->cons_stride;
        s.clear();
        if (format == 1)
            sam_record(qn, rn, rec->start[i], rec->alen[i], qc, rc, &s);
        else
            text_record(qn, rn, (float)rec->score[i], rec->start[i] + (int)rec->ref_distance[i],
                        rec->end[i] + (int)rec->ref_distance[i], rec->alen[i], qc, rc, &s);
        if (fwrite(s.data(), 1, s.size(), f) != s.size()) {
            fclose(f);
            set_error("write to %s failed", path);
            return kErrIO;
        }
    }
    fclose(f);
    return kOk;
}

}  // extern "C"
