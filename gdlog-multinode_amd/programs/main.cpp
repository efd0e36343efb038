// main.cpp -- command-line drivers TC, SG and CSPA (the reference's
// executables of the same names, CMakeLists.txt:58-91): read a tab-separated
// edge file (CSPA: a directory holding assign.facts and dereference.facts),
// run the program and print the result size and times.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "gdlog/exception.h"
#include "programs.h"

namespace {

// get_relation_from_file (test/tc.cu:17-57): rows of `%ld<sep>%ld`.
std::vector<column_type> read_pairs(const std::string &path) {
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

EdgeInput host_input(const std::vector<column_type> &rows) {
    EdgeInput in;
    in.host = rows.data();
    in.rows = rows.size() / 2;
    return in;
}

}  // namespace

int main(int argc, char *argv[]) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <dataset>" << std::endl;
        return 2;
    }
    const int block_size = 512, grid_size = 32 * 256;
    ProgramOutput out;
    try {
#if defined(GDL_PROGRAM_TC)
        std::vector<column_type> g = read_pairs(argv[1]);
        std::cout << "Input graph rows: " << g.size() / 2 << std::endl;
        tc_program(host_input(g), -1, grid_size, block_size, &out);
        std::cout << "Path counts " << out.n_main << std::endl;
        std::cout << "TC time: " << out.fixpoint_ms / 1000.0 << std::endl;
#elif defined(GDL_PROGRAM_SG)
        std::vector<column_type> g = read_pairs(argv[1]);
        std::cout << "Input graph rows: " << g.size() / 2 << std::endl;
        sg_program(host_input(g), grid_size, block_size, &out);
        std::cout << "sg counts " << out.n_main << std::endl;
        std::cout << "sg time: " << out.fixpoint_ms / 1000.0 << std::endl;
#elif defined(GDL_PROGRAM_CSPA)
        std::string dir = argv[1];
        std::vector<column_type> a = read_pairs(dir + "/assign.facts");
        std::vector<column_type> d = read_pairs(dir + "/dereference.facts");
        std::cout << "Input assign rows: " << a.size() / 2 << std::endl;
        std::cout << "Input dereference rows: " << d.size() / 2 << std::endl;
        cspa_program(host_input(a), host_input(d), grid_size, block_size, &out);
        std::cout << "ValueFlow " << out.sizes[0] << std::endl;
        std::cout << "MemoryAlias " << out.sizes[1] << std::endl;
        std::cout << "ValueAlias " << out.sizes[2] << std::endl;
        std::cout << "analysis scc time: " << out.fixpoint_ms / 1000.0 << std::endl;
#endif
        std::cout << "iterations " << out.iterations << " derived tuples " << out.derived
                  << std::endl;
    } catch (const gdlog::Error &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
