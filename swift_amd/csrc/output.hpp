// output.hpp -- record formatters shared by the output drop-ins and the
// phase-2 drivers (output.cpp).
#pragma once
#include <string>

namespace swift {

// printCigar's runs (src/output.cu:33-79) of walk-order consensus strings
void cigar_string(const char *qc, const char *rc, int alen, std::string *out);
// one printOutput record (src/output.cu:154-201)
void text_record(const std::string &qn, const std::string &rn, float score, int start, int end,
                 int alen, const char *qc, const char *rc, std::string *out);
// one printOutputSam record (src/output.cu:98-130)
void sam_record(const std::string &qn, const std::string &rn, int start, int alen, const char *qc,
                const char *rc, std::string *out);

}  // namespace swift
