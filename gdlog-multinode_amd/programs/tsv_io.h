// tsv_io.h -- fact-file input and result output of the command-line drivers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "gdlog/exception.h"
#include "gdlog/tuple.h"

namespace gdlog_cli {

// get_relation_from_file (test/tc.cu:17-57): rows of `%ld<sep>%ld`, the
// separator a tab, a space or a comma.  A missing file ends the process.
inline std::vector<column_type> read_pairs(const std::string &path) {
    std::vector<column_type> rows;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) {
        std::cerr << "cannot open " << path << std::endl;
        std::exit(2);
    }
    long a, b;
    while (std::fscanf(f, "%ld%*[ \t,]%ld", &a, &b) == 2) {
        rows.push_back((column_type)a);
        rows.push_back((column_type)b);
    }
    std::fclose(f);
    return rows;
}

// One id per row (a unary relation's facts).  A missing file ends the process.
inline std::vector<column_type> read_ids(const std::string &path) {
    std::vector<column_type> rows;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) {
        std::cerr << "cannot open " << path << std::endl;
        std::exit(2);
    }
    long a;
    while (std::fscanf(f, "%ld", &a) == 1) rows.push_back((column_type)a);
    std::fclose(f);
    return rows;
}

// Sorted keys ((c0 << 32) | c1) as tab-separated rows; arity 1 writes c0 only.
inline void write_tsv(const std::string &path, const std::vector<uint64_t> &keys,
                      int arity = 2) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + path);
    for (uint64_t k : keys) {
        if (arity == 1) std::fprintf(f, "%u\n", (unsigned)(k >> 32));
        else std::fprintf(f, "%u\t%u\n", (unsigned)(k >> 32), (unsigned)(k & 0xffffffffu));
    }
    if (std::fclose(f) != 0) throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + path);
}

// Keys as little-endian u64.
inline void write_bin(const std::string &path, const std::vector<uint64_t> &keys) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + path);
    const size_t n = keys.size();
    const bool ok = n == 0 || std::fwrite(keys.data(), sizeof(uint64_t), n, f) == n;
    if (std::fclose(f) != 0 || !ok)
        throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + path);
}

}  // namespace gdlog_cli
