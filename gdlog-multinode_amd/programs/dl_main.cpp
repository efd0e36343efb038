// dl_main.cpp -- DL, the command-line driver of the Datalog front end
// (gdlog/datalog.h): runs a .dl program on fact files, in the manner of
// `souffle -F <facts> -D <out> <program.dl>` (the reference runs its Souffle
// programs that way, test/souffle/*.dl).
//
//   DL <program.dl | plan.slogc> [-F DIR] [-D DIR] [--plan]
//                   [--checkpoint DIR [--every N] [--resume]]
//   A .slogc file is a Slog compiler's RA plan (DatalogProgram::compile_slog);
//   its relation names become identifiers (value-flow -> value_flow).
//   -F DIR            each .input relation R is read from DIR/R.facts, or from
//                     its filename attribute (relative to DIR unless absolute);
//                     rows of two ids separated by a tab, space or comma, or
//                     one id for a unary relation; default .
//   -D DIR            each kept relation R is written to DIR/R.csv, sorted
//                     tab-separated rows
//   --plan            print the lowered plan before running (for a .slogc
//                     plan, the rules read back from it first)
//   --checkpoint ...  snapshot each LIE's fixpoint to DIR/lie<k>.ckpt, as the
//                     TC / SG / CSPA drivers do
// Prints each kept relation's size and checksum (sum of fmix64 over its keys),
// the iteration count of each LIE and the run time.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "gdlog/datalog.h"
#include "gdlog/lie.h"
#include "tsv_io.h"

namespace {

// MurmurHash3's 64-bit finalizer; the checksum the engine's gdl_checksum_u64
// computes on the device.
uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

}  // namespace

int main(int argc, char *argv[]) {
    std::string program, facts = ".", out_dir, ckpt_dir;
    bool show_plan = false, resume = false;
    int every = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << a << " needs a value" << std::endl;
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-F") facts = value();
        else if (a == "-D") out_dir = value();
        else if (a == "--plan") show_plan = true;
        else if (a == "--checkpoint") ckpt_dir = value();
        else if (a == "--every") every = std::atoi(value().c_str());
        else if (a == "--resume") resume = true;
        else if (!a.empty() && a[0] == '-') {
            std::cerr << "unknown option " << a << std::endl;
            return 2;
        } else program = a;
    }
    if (program.empty()) {
        std::cerr << "usage: " << argv[0]
                  << " <program.dl> [-F DIR] [-D DIR] [--plan] [--checkpoint DIR [--every N]"
                     " [--resume]]"
                  << std::endl;
        return 2;
    }
    std::ifstream in(program);
    if (!in) {
        std::cerr << "cannot open " << program << std::endl;
        return 2;
    }
    std::stringstream text;
    text << in.rdbuf();

    gdlog::DatalogProgram prog;
    gdl_status st = GDL_OK;
    std::string err;
    const bool slog = program.size() > 6 && program.compare(program.size() - 6, 6, ".slogc") == 0;
    if (!(slog ? prog.compile_slog(text.str(), &st, &err) : prog.compile(text.str(), &st, &err))) {
        std::cerr << program << ": " << err << std::endl;
        return 1;
    }
    if (show_plan) std::cout << (slog ? prog.source() : "") << prog.plan();

    // each .input relation comes from its .facts file
    std::map<std::string, std::vector<uint64_t>> inputs;
    for (const std::string &rel : prog.inputs()) {
        // the .input's filename attribute (relative to -F unless absolute), else
        // <relation>.facts, as Souffle resolves them
        const std::string &attr = prog.input_file(rel);
        const std::string fn = attr.empty() ? facts + "/" + rel + ".facts"
                               : attr[0] == '/' ? attr
                                                : facts + "/" + attr;
        const int ar = prog.arity(rel);
        const std::vector<column_type> rows =
            ar == 1 ? gdlog_cli::read_ids(fn) : gdlog_cli::read_pairs(fn);
        std::vector<uint64_t> &keys = inputs[rel];
        keys.reserve(rows.size() / ar);
        column_type mx = 0;
        for (size_t r = 0; r + ar <= rows.size(); r += ar) {
            const column_type b = ar == 2 ? rows[r + 1] : 0;
            mx |= rows[r] | b;
            keys.push_back(((uint64_t)rows[r] << 32) | (uint64_t)(b & 0xffffffffu));
        }
        if (mx >> 32) {
            std::cerr << rel << ".facts: an id does not fit in 32 bits" << std::endl;
            return 1;
        }
        std::cout << "Input " << rel << " rows: " << keys.size() << std::endl;
    }
    gdlog::set_checkpoint_defaults(ckpt_dir.empty() ? nullptr : ckpt_dir.c_str(), every, resume,
                                   -1);
    try {
        const auto t0 = std::chrono::steady_clock::now();
        prog.run(inputs);
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
        for (const std::string &rel : prog.kept()) {
            const std::vector<uint64_t> &keys = *prog.result(rel);
            uint64_t sum = 0;
            for (uint64_t k : keys) sum += fmix64(k);
            std::printf("%s %zu checksum 0x%016llx\n", rel.c_str(), keys.size(),
                        (unsigned long long)sum);
            if (!out_dir.empty())
                gdlog_cli::write_tsv(out_dir + "/" + rel + ".csv", keys, prog.arity(rel));
        }
        std::cout << "iterations";
        for (int it : prog.iterations()) std::cout << " " << it;
        std::cout << std::endl;
        std::cout << "derived tuples " << prog.derived() << std::endl;
        std::cout << "run time: " << ms / 1000.0 << std::endl;
    } catch (const gdlog::Error &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
