// drivers.cpp -- the reference's whole-phase phase-2 drivers as drop-ins:
// alignQueries3 (src/align.cu:574-808, the GPU driver) and
// alignQueries_cpu_mt (src/align.cu:2363-2511, the CPU driver).  Both read
// the match file written by phase 1, fetch each match's query and reference
// window from the FASTA files by file offset, align every pair and write the
// reference's text or SAM output.  Here the files are read once into memory,
// the offsets and windows follow the reference's own rules (refMapCreate,
// getSeq, getRefSeq2), and all pairs of the file go through the batched
// device Smith-Waterman in chunks.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swift_abi.h"
#include "output.hpp"
#include "sw_host.hpp"
#include "sw_types.hpp"

namespace swift {
namespace {

constexpr int kRefLen = 300;        // REF_LENGTH (src/common.h:57)
constexpr int kNameStride = 50;     // MAX_SEQ_NAME_LENGTH (src/common.h:37-39)
constexpr int kLineMax = 100000;    // MAX_LINE_LENGTH (src/common.h:41)
constexpr int kCpuThreads = 8;      // _NUM_THREADS of alignQueries_cpu_mt
constexpr int kChunk = 1 << 20;     // pairs per device batch

bool load_file(const char *path, std::string *out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    out->clear();
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, n);
    fclose(f);
    return true;
}

// One fgets call at `off`: the line (at most kLineMax - 1 bytes, '\n'
// included) and the offset after it; false at end of file.
bool file_line(const std::string &f, size_t off, size_t *b, size_t *e, size_t *next) {
    if (off >= f.size()) return false;
    size_t end = off;
    const size_t cap = std::min(f.size(), off + (size_t)kLineMax - 1);
    while (end < cap && f[end] != '\n') ++end;
    if (end < cap) ++end;  // the '\n' belongs to this fgets
    *b = off;
    *e = end;
    *next = end;
    return true;
}

// refMapCreate (src/refMap.cu:59-117): per sequence, the offsets of its
// header and first base, and the bases per full line (0 for one line).
struct RefMap {
    std::vector<uint32_t> name_off, seq_off;
    std::vector<int> per_line;
};

RefMap ref_map(const std::string &f) {
    RefMap m;
    int len = 0, old_len = 0, lines = 0;
    uint32_t pos = 0;
    size_t off = 0, b, e, next;
    while (file_line(f, off, &b, &e, &next)) {
        off = next;
        if (len > 0) old_len = len;
        len = (int)(e - b);
        if (len > 0 && f[e - 1] == '\n') --len;
        if (len == 0) continue;
        if (f[b] == '>') {
            if (!m.name_off.empty())
                m.per_line.push_back(lines > 1 ? (int)((pos - old_len) / (lines - 1)) : 0);
            m.name_off.push_back((uint32_t)(next - len - 1));
            m.seq_off.push_back((uint32_t)next);
            pos = 0;
            lines = 0;
        } else {
            ++lines;
            pos += (uint32_t)len;
        }
    }
    if (len > 0) old_len = len;
    if (!m.name_off.empty())
        m.per_line.push_back(lines > 1 ? (int)((pos - old_len) / (lines - 1)) : 0);
    return m;
}

// getSeqName (src/input.cu:37-52): the header line at `off`, without '>'
// and the line end.
std::string seq_name(const std::string &f, size_t off) {
    size_t b, e, next;
    if (!file_line(f, off, &b, &e, &next)) return std::string();
    if (e > b && f[e - 1] == '\n') --e;
    return e > b ? f.substr(b + 1, e - b - 1) : std::string();
}

// getSeq / getSeq2 (src/input.cu:163-204,216-258): sequence lines from
// `off`, concatenated up to max_len bases or the next header.
std::string get_seq(const std::string &f, size_t off, int max_len) {
    std::string s;
    int count = 0;
    size_t b, e, next;
    while (file_line(f, off, &b, &e, &next)) {
        off = next;
        if (f[b] == '>') break;
        int len = (int)(e - b);
        if (len > 0 && f[e - 1] == '\n') --len;
        count += len;
        if (count > max_len) len -= count - max_len;
        s.append(f, b, (size_t)std::max(len, 0));
        if (count >= max_len) break;
    }
    return s;
}

// getRefSeq2 (src/input.cu:348-403): like get_seq, then leading N/n
// stripped; trailing N/n stripped only when no leading N was (the
// reference's trailing test indexes past the stripped copy otherwise).
std::string get_ref_seq2(const std::string &f, size_t off, int max_len) {
    std::string s = get_seq(f, off, max_len);
    size_t lead = 0;
    while (lead < s.size() && (s[lead] == 'N' || s[lead] == 'n')) ++lead;
    size_t n = s.size();
    if (lead == 0)
        while (n > 0 && (s[n - 1] == 'N' || s[n - 1] == 'n')) --n;
    return lead < n ? s.substr(lead, n - lead) : std::string();
}

// qryNameCmp (src/align.cu:1501-1519): equal, ignoring one trailing '*'.
bool same_query(const std::string &s, const std::string &t) {
    size_t i = 0;
    for (; i < s.size() && i < t.size(); ++i)
        if (s[i] != t[i]) return false;
    if (s.size() == t.size()) return true;
    const std::string &lng = s.size() > t.size() ? s : t;
    return lng.size() == i + 1 && lng[i] == '*';
}

void revcomp(const std::string &in, std::string *out) {
    // qryGetReverseComplement (src/query.cu:265-282): A<->T, C<->G, anything
    // else -> A
    out->resize(in.size());
    for (size_t i = 0; i < in.size(); ++i) {
        const char c = in[in.size() - 1 - i];
        (*out)[i] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
    }
}

struct Match {
    int rev, ref_idx, ref_pos;
    unsigned qname_off, qseq_off;
};

// getOffsets (src/input.cu:75-82): the same sscanf format
bool parse_match(const std::string &line, Match *m) {
    uint32_t rev = 0, qn = 0, qs = 0;
    int ri = 0, rp = 0;
    if (sscanf(line.c_str(), "%*s\t%u\t%d\t%*d\t%d\t%u\t%u", &rev, &ri, &rp, &qn, &qs) != 5)
        return false;
    *m = Match{(int)rev, ri, rp, qn, qs};
    return true;
}

struct Pair {
    std::string qname, rname, q, r;
    int ref_distance;
};

struct DeviceRun {
    DevBuf q, qo, ql, r, ro, rl, co, sc, st, en, al, qc, rc;
};

// Aligns pairs [0, n) on the device; fills score/start/end/alen and the
// consensus strings (walk order) per pair.
int align_pairs(const std::vector<Pair> &p, size_t p0, size_t n, const Scoring &sc,
                DeviceRun &d, std::vector<int32_t> *score, std::vector<int32_t> *start,
                std::vector<int32_t> *end, std::vector<int32_t> *alen,
                std::vector<char> *qcons, std::vector<char> *rcons, std::vector<int64_t> *coff) {
    std::vector<char> qb, rb;
    std::vector<int64_t> qo(n), ro(n);
    std::vector<int32_t> ql(n), rl(n);
    coff->assign(n, 0);
    int maxq = 1, maxr = 1;
    int64_t cons = 0;
    for (size_t i = 0; i < n; ++i) {
        const Pair &x = p[p0 + i];
        qo[i] = (int64_t)qb.size();
        ro[i] = (int64_t)rb.size();
        qb.insert(qb.end(), x.q.begin(), x.q.end());
        rb.insert(rb.end(), x.r.begin(), x.r.end());
        ql[i] = (int32_t)x.q.size();
        rl[i] = (int32_t)x.r.size();
        maxq = std::max(maxq, ql[i]);
        maxr = std::max(maxr, rl[i]);
        (*coff)[i] = cons;
        cons += (ql[i] + rl[i] + 2 + 3) / 4 * 4;
    }
    qb.resize(qb.size() + 8, 0);
    rb.resize(rb.size() + 8, 0);
    int rc;
    if ((rc = d.q.ensure(qb.size())) || (rc = d.r.ensure(rb.size())) ||
        (rc = d.qo.ensure(8 * n + 8)) || (rc = d.ro.ensure(8 * n + 8)) ||
        (rc = d.ql.ensure(4 * n + 4)) || (rc = d.rl.ensure(4 * n + 4)) ||
        (rc = d.co.ensure(8 * n + 8)) || (rc = d.sc.ensure(4 * n + 4)) ||
        (rc = d.st.ensure(4 * n + 4)) || (rc = d.en.ensure(4 * n + 4)) ||
        (rc = d.al.ensure(4 * n + 4)) || (rc = d.qc.ensure((size_t)cons + 8)) ||
        (rc = d.rc.ensure((size_t)cons + 8)))
        return rc;
    SWIFT_HIP_TRY(hipMemcpy(d.q.ptr, qb.data(), qb.size(), hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.r.ptr, rb.data(), rb.size(), hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.qo.ptr, qo.data(), 8 * n, hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.ro.ptr, ro.data(), 8 * n, hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.ql.ptr, ql.data(), 4 * n, hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.rl.ptr, rl.data(), 4 * n, hipMemcpyHostToDevice));
    SWIFT_HIP_TRY(hipMemcpy(d.co.ptr, coff->data(), 8 * n, hipMemcpyHostToDevice));
    BatchIO io{d.q.as<char>(), d.qo.as<int64_t>(), d.ql.as<int32_t>(), d.r.as<char>(),
               d.ro.as<int64_t>(), d.rl.as<int32_t>(), (int)n, maxq, maxr,
               d.sc.as<int32_t>(), d.st.as<int32_t>(), d.en.as<int32_t>(), d.al.as<int32_t>(),
               d.qc.as<char>(), d.rc.as<char>(), d.co.as<int64_t>()};
    if ((rc = sw_batch_run(io, sc, nullptr, true))) return rc;
    SWIFT_HIP_TRY(hipDeviceSynchronize());
    score->resize(n);
    start->resize(n);
    end->resize(n);
    alen->resize(n);
    qcons->resize((size_t)cons + 8);
    rcons->resize((size_t)cons + 8);
    SWIFT_HIP_TRY(hipMemcpy(score->data(), d.sc.ptr, 4 * n, hipMemcpyDeviceToHost));
    SWIFT_HIP_TRY(hipMemcpy(start->data(), d.st.ptr, 4 * n, hipMemcpyDeviceToHost));
    SWIFT_HIP_TRY(hipMemcpy(end->data(), d.en.ptr, 4 * n, hipMemcpyDeviceToHost));
    SWIFT_HIP_TRY(hipMemcpy(alen->data(), d.al.ptr, 4 * n, hipMemcpyDeviceToHost));
    SWIFT_HIP_TRY(hipMemcpy(qcons->data(), d.qc.ptr, (size_t)cons, hipMemcpyDeviceToHost));
    SWIFT_HIP_TRY(hipMemcpy(rcons->data(), d.rc.ptr, (size_t)cons, hipMemcpyDeviceToHost));
    return kOk;
}

std::string bounded_name(const std::string &s) {
    // the drivers keep names in MAX_SEQ_NAME_LENGTH slots
    return s.size() < (size_t)kNameStride ? s : s.substr(0, kNameStride - 1);
}

// The shared body.  gpu_driver: alignQueries3's rules (best-hit choice over
// consecutive records of one read for the text format); else
// alignQueries_cpu_mt's (every record printed, in rounds of 8 match lines).
int run_driver(const char *qry_file, int qry_len, const char *ref_file, const char *match_file,
               int out_format, float match, float mismatch, float gap_open, float gap_ext,
               const char *output_file, int num_matches, bool gpu_driver) {
    if (!qry_file || !ref_file || !match_file || !output_file || qry_len <= 0) {
        set_error("invalid arguments");
        return kErrInvalid;
    }
    Scoring sc;
    int rc;
    if ((rc = scoring_from_floats(match, mismatch, gap_open, gap_ext, &sc))) return rc;
    std::string qf, rf, mf;
    if (!load_file(qry_file, &qf) || !load_file(ref_file, &rf) || !load_file(match_file, &mf)) {
        set_error("cannot read %s / %s / %s", qry_file, ref_file, match_file);
        return kErrIO;
    }
    const RefMap map = ref_map(rf);
    // match lines, read as getOffsets does: one fgets per record
    std::vector<std::string> lines;
    {
        size_t off = 0, b, e, next;
        while (file_line(mf, off, &b, &e, &next)) {
            lines.emplace_back(mf, b, e - b);
            off = next;
        }
    }
    size_t nrec;
    if (gpu_driver) {
        if (num_matches < 0 || (size_t)num_matches > lines.size()) {
            set_error("numMatches %d but the match file has %zu lines", num_matches, lines.size());
            return kErrInvalid;
        }
        nrec = (size_t)num_matches;
    } else {
        // rounds of _NUM_THREADS records; a round past the last line re-reads
        // the last line (fgets at end of file leaves it in the buffer), so
        // the last match is repeated up to the next multiple of 8
        nrec = lines.empty() ? 0 : (lines.size() + kCpuThreads - 1) / kCpuThreads * kCpuThreads;
    }
    std::vector<Pair> pairs(nrec);
    for (size_t i = 0; i < nrec; ++i) {
        Match m;
        const std::string &ln = lines[std::min(i, lines.size() - 1)];
        if (!parse_match(ln, &m)) {
            set_error("match line %zu is not a match record", i + 1);
            return kErrInvalid;
        }
        if (m.ref_idx < 0 || m.ref_idx >= (int)map.name_off.size()) {
            set_error("match line %zu names reference %d of %zu", i + 1, m.ref_idx,
                      map.name_off.size());
            return kErrInvalid;
        }
        Pair &p = pairs[i];
        p.ref_distance = std::max(0, m.ref_pos - kRefLen / 3);
        const int per_line = map.per_line[m.ref_idx];
        const size_t roff = (size_t)map.seq_off[m.ref_idx] + (uint32_t)p.ref_distance +
                            (per_line > 0 ? (uint32_t)p.ref_distance / (uint32_t)per_line : 0);
        p.r = get_ref_seq2(rf, roff, kRefLen);
        p.rname = bounded_name(seq_name(rf, map.name_off[m.ref_idx]));
        const std::string q = get_seq(qf, m.qseq_off, qry_len);
        p.qname = seq_name(qf, m.qname_off);
        if (m.rev == 1) {
            revcomp(q, &p.q);
            p.qname += "*";
        } else {
            p.q = q;
        }
        p.qname = bounded_name(p.qname);
        if (p.q.size() > (size_t)kMaxBatchQ) {
            set_error("query of match line %zu is longer than %d", i + 1, kMaxBatchQ);
            return kErrCapacity;
        }
    }
    FILE *out = fopen(output_file, "w");
    if (!out) {
        set_error("cannot write %s", output_file);
        return kErrIO;
    }
    DeviceRun dev;
    std::vector<int32_t> score, start, end, alen;
    std::vector<char> qcons, rcons;
    std::vector<int64_t> coff;
    std::string text;
    for (size_t p0 = 0; p0 < nrec;) {
        size_t n = std::min((size_t)kChunk, nrec - p0);
        // a chunk never splits the records of one read (the best-hit choice
        // spans them)
        while (gpu_driver && p0 + n < nrec && same_query(pairs[p0 + n - 1].qname, pairs[p0 + n].qname))
            ++n;
        if ((rc = align_pairs(pairs, p0, n, sc, dev, &score, &start, &end, &alen, &qcons, &rcons,
                              &coff))) {
            fclose(out);
            return rc;
        }
        std::vector<float> fs(n);
        for (size_t i = 0; i < n; ++i) fs[i] = (float)score[i];
        if (gpu_driver && out_format != 1) {
            // chooseBestAlignment (src/align.cu:844-905): per group of
            // consecutive records of one read, the first best score is kept
            // (the reference draws one of the tied records with rand())
            for (size_t g = 0; g < n;) {
                size_t h = g + 1;
                while (h < n && same_query(pairs[p0 + h].qname, pairs[p0 + h - 1].qname)) ++h;
                float best = 0.0f;
                for (size_t i = g; i < h; ++i) best = std::max(best, fs[i]);
                size_t keep = h;
                for (size_t i = g; i < h && keep == h; ++i)
                    if (fs[i] == best) keep = i;
                for (size_t i = g; i < h; ++i)
                    if (i != keep) fs[i] = -1.0f;
                g = h;
            }
        }
        for (size_t i = 0; i < n; ++i) {
            const Pair &p = pairs[p0 + i];
            const char *qc = qcons.data() + coff[i];
            const char *rcs = rcons.data() + coff[i];
            text.clear();
            if (out_format == 1) {
                sam_record(p.qname, p.rname, start[i], alen[i], qc, rcs, &text);
            } else {
                if (fs[i] < 0.0f) continue;  // printOutput skips dropped records
                text_record(p.qname, p.rname, fs[i], start[i] + p.ref_distance,
                            end[i] + p.ref_distance, alen[i], qc, rcs, &text);
            }
            if (fwrite(text.data(), 1, text.size(), out) != text.size()) {
                fclose(out);
                set_error("write to %s failed", output_file);
                return kErrIO;
            }
        }
        p0 += n;
    }
    fclose(out);
    return kOk;
}

void exit_on_error(int rc, const char *who) {
    if (rc == kOk) return;
    fprintf(stderr, "Error: %s: %s\n", who, swift_last_error());
    exit(1);
}

}  // namespace
}  // namespace swift

using namespace swift;

extern "C" {

int swift_align_queries(const char *qryFile, int qryLen, const char *refFile,
                        const char *matchFile, int outFormat, float match, float mismatch,
                        float gapOpenPenalty, float gapExtPenalty, const char *outputFile,
                        int numMatches, int gpuDriver) {
    return run_driver(qryFile, qryLen, refFile, matchFile, outFormat, match, mismatch,
                      gapOpenPenalty, gapExtPenalty, outputFile, numMatches, gpuDriver != 0);
}

void alignQueries3(char *queryFileName, int maxQryLen, char *refFileName, char *matchFile,
                   int outFormat, float match, float mismatch, float gapOpenPenalty,
                   float gapExtPenalty, const char *outputFile, int numMatches) {
    exit_on_error(swift_align_queries(queryFileName, maxQryLen, refFileName, matchFile,
                                      outFormat, match, mismatch, gapOpenPenalty,
                                      gapExtPenalty, outputFile, numMatches, 1),
                  "alignQueries3");
}

void alignQueries_cpu_mt(char *qryFile, int numQrs, int qryLen, char *refFile, char *matchFile,
                         int outFormat, float match, float mismatch, float gapOpenPenalty,
                         float gapExtPenalty, const char *outputFile) {
    (void)numQrs;
    exit_on_error(swift_align_queries(qryFile, qryLen, refFile, matchFile, outFormat, match,
                                      mismatch, gapOpenPenalty, gapExtPenalty, outputFile, -1, 0),
                  "alignQueries_cpu_mt");
}

}  // extern "C"
