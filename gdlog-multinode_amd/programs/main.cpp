// main.cpp -- command-line drivers TC, SG, CSPA and PLEN (the reference's
// executables of the same names, CMakeLists.txt:58-91): read a tab-separated
// edge file (CSPA: a directory holding assign.facts and dereference.facts),
// run the program and print the result size and times.
//
// Options (extensions; the reference prints sizes only):
//   --out PATH        the result as sorted tab-separated rows, one per tuple
//                     (CSPA: a directory receiving ValueFlow.csv,
//                     MemoryAlias.csv and ValueAlias.csv; PLEN: rows
//                     from, to, hop count), for parity diffs
//   --out-bin PATH    the main result as sorted little-endian u64 keys
//                     (c0 << 32) | c1
//   --checkpoint DIR  snapshot each fixpoint to DIR/lie<k>.ckpt (at the end,
//                     and every N iterations with --every N); with --resume,
//                     start each fixpoint from its snapshot there
//   --max-iteration K (TC) stop after iteration K, as the reference's LIE
//                     max_iteration
// and always print the result checksum (sum of fmix64 over the keys).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "gdlog/exception.h"
#include "programs.h"
#include "tsv_io.h"

namespace {

using gdlog_cli::read_pairs;
using gdlog_cli::write_bin;
using gdlog_cli::write_tsv;

EdgeInput host_input(const std::vector<column_type> &rows) {
    EdgeInput in;
    in.host = rows.data();
    in.rows = rows.size() / 2;
    return in;
}

}  // namespace

int main(int argc, char *argv[]) {
    std::string dataset, out_path, out_bin, ckpt_dir;
    int every = 0, max_iteration = -1;
    bool resume = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << a << " needs a value" << std::endl;
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--out") out_path = value();
        else if (a == "--out-bin") out_bin = value();
        else if (a == "--checkpoint") ckpt_dir = value();
        else if (a == "--every") every = std::atoi(value().c_str());
        else if (a == "--resume") resume = true;
        else if (a == "--max-iteration") max_iteration = std::atoi(value().c_str());
        else if (!a.empty() && a[0] == '-') {
            std::cerr << "unknown option " << a << std::endl;
            return 2;
        } else dataset = a;
    }
    if (dataset.empty()) {
        std::cerr << "usage: " << argv[0]
                  << " <dataset> [--out PATH] [--out-bin PATH] [--checkpoint DIR [--every N]"
                     " [--resume]] [--max-iteration K]"
                  << std::endl;
        return 2;
    }
#if !defined(GDL_PROGRAM_TC)
    if (max_iteration >= 0) {
        std::cerr << "--max-iteration applies to TC only" << std::endl;
        return 2;
    }
#endif
    const int block_size = 512, grid_size = 32 * 256;
    ProgramOutput out;
    out.want_keys = !out_path.empty() || !out_bin.empty();
    gdlog::set_checkpoint_defaults(ckpt_dir.empty() ? nullptr : ckpt_dir.c_str(), every, resume,
                                   -1);
    gdlog::reset_checkpoint_sequence();
    try {
#if defined(GDL_PROGRAM_TC)
        std::vector<column_type> g = read_pairs(dataset);
        std::cout << "Input graph rows: " << g.size() / 2 << std::endl;
        tc_program(host_input(g), max_iteration, grid_size, block_size, &out);
        std::cout << "Path counts " << out.n_main << std::endl;
        std::cout << "TC time: " << out.fixpoint_ms / 1000.0 << std::endl;
        if (!out_path.empty()) write_tsv(out_path, out.main_keys);
#elif defined(GDL_PROGRAM_PLEN)
        std::vector<column_type> g = read_pairs(dataset);
        std::cout << "Input graph rows: " << g.size() / 2 << std::endl;
        std::vector<uint32_t> lens;
        out.want_keys = true;
        path_length_program(host_input(g), grid_size, block_size, &out,
                            out_path.empty() ? nullptr : &lens);
        std::cout << "Path counts " << out.n_main << std::endl;
        std::cout << "longest path " << out.sizes[0] << std::endl;
        std::cout << "PLEN time: " << out.fixpoint_ms / 1000.0 << std::endl;
        if (!out_path.empty()) {
            FILE *f = std::fopen(out_path.c_str(), "w");
            if (!f) throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + out_path);
            for (size_t i = 0; i < lens.size(); ++i)
                std::fprintf(f, "%u\t%u\t%u\n", (unsigned)(out.main_keys[i] >> 32),
                             (unsigned)(out.main_keys[i] & 0xffffffffu), lens[i]);
            if (std::fclose(f) != 0) throw gdlog::Error(GDL_ERR_INVALID_ARG, "cannot write " + out_path);
        }
#elif defined(GDL_PROGRAM_SG)
        std::vector<column_type> g = read_pairs(dataset);
        std::cout << "Input graph rows: " << g.size() / 2 << std::endl;
        sg_program(host_input(g), grid_size, block_size, &out);
        std::cout << "sg counts " << out.n_main << std::endl;
        std::cout << "sg time: " << out.fixpoint_ms / 1000.0 << std::endl;
        if (!out_path.empty()) write_tsv(out_path, out.main_keys);
#elif defined(GDL_PROGRAM_CSPA)
        std::vector<column_type> a = read_pairs(dataset + "/assign.facts");
        std::vector<column_type> d = read_pairs(dataset + "/dereference.facts");
        std::cout << "Input assign rows: " << a.size() / 2 << std::endl;
        std::cout << "Input dereference rows: " << d.size() / 2 << std::endl;
        cspa_program(host_input(a), host_input(d), grid_size, block_size, &out);
        std::cout << "ValueFlow " << out.sizes[0] << std::endl;
        std::cout << "MemoryAlias " << out.sizes[1] << std::endl;
        std::cout << "ValueAlias " << out.sizes[2] << std::endl;
        std::cout << "analysis scc time: " << out.fixpoint_ms / 1000.0 << std::endl;
        if (!out_path.empty()) {
            write_tsv(out_path + "/ValueFlow.csv", out.extra_keys[0]);
            write_tsv(out_path + "/MemoryAlias.csv", out.extra_keys[1]);
            write_tsv(out_path + "/ValueAlias.csv", out.main_keys);
        }
#endif
        if (!out_bin.empty()) write_bin(out_bin, out.main_keys);
        std::cout << "iterations " << out.iterations << " derived tuples " << out.derived
                  << std::endl;
        std::printf("checksum 0x%016llx\n", (unsigned long long)out.checksum);
    } catch (const gdlog::Error &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
