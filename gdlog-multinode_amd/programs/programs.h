// programs.h -- shared input/output plumbing of the TC, SG and CSPA programs.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "gdlog/lie.h"

// The host copy of a __device__ hook pointer (a tuple generator, predicate or
// copy hook), read with hipMemcpyFromSymbol once per symbol and device: the
// reference's programs read it on every run, and each read waits for the
// whole device.
template <typename T>
T host_hook(const T &device_symbol) {
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, T> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const std::pair<const void *, int> key((const void *)&device_symbol, dev);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    T h{};
    if (hipMemcpyFromSymbol(&h, (const void *)&device_symbol, sizeof(T), 0,
                            hipMemcpyDeviceToHost) != hipSuccess)
        return h;  // not cached: the operator's hook check reports the null hook
    cache.emplace(key, h);
    return h;
}

// Edge input of a program: host rows of two 8-byte columns (the reference's
// get_relation_from_file layout) or device-resident packed keys.
struct EdgeInput {
    const column_type *host = nullptr;
    const uint64_t *dev = nullptr;
    tuple_size_t rows = 0;
    // load_relation of the edges (reversed: (c1, c0) rows, the reference's
    // host-side reversal, test/tc.cu:94-101).
    void load(Relation *r, const std::string &name, bool reversed, int grid_size,
              int block_size, bool tmp_flag = false) const;
};

struct ProgramOutput {
    int iterations = 0;
    int init_iterations = 0;
    uint64_t derived = 0;  // sum of |delta| over update relations and iterations
    double load_ms = 0.0;
    double fixpoint_ms = 0.0;
    std::vector<IterationRecord> trace, trace_init;
    bool want_keys = false;
    uint64_t n_main = 0;
    uint64_t checksum = 0;  // relation_checksum of the main output
    std::vector<uint64_t> main_keys;
    std::vector<uint64_t> extra_keys[2];
    uint64_t sizes[4] = {0, 0, 0, 0};
};

double now_ms();  // device-synchronised wall clock
uint64_t sum_delta(const std::vector<IterationRecord> &t);

void tc_program(const EdgeInput &graph, int max_iteration, int grid_size, int block_size,
                ProgramOutput *out);
// All-pairs hop counts (test/path_length.cu): out as TC's, plus *lens[i] the
// hop count of out->main_keys[i] (filled when lens is given); out->sizes[0]
// is the longest hop count.
void path_length_program(const EdgeInput &graph, int grid_size, int block_size,
                         ProgramOutput *out, std::vector<uint32_t> *lens);
// The same result from the reference's own encoding (test/path_length.cu:
// arity-3 path with a dependent column, the hook computing newt[2] =
// outer[2] + 1 run on the device); out->main_keys are the (from, to) pairs.
// TC with hooks that run on the device (variant 0: predicate t[1] != 42;
// 1: generator (inner[1], (outer[1] + 1) mod 1000); 2: a copy hook seeding
// path with (src[0], (src[1] * 7) mod 1000)).
void tc_hooks_program(const EdgeInput &graph, int variant, int grid_size, int block_size,
                      ProgramOutput *out);
void path_length_hooks_program(const EdgeInput &graph, int grid_size, int block_size,
                               ProgramOutput *out, std::vector<uint32_t> *lens);

// One program run (kind 0 TC, 1 SG, 2 CSPA) from host packed keys on ctx,
// which is this thread's engine context for the run; output keys are copied
// to the host when out->want_keys is set, the checksum always.  Errors come
// back as the status and *err.
gdl_status run_program_host(gdl_ctx *ctx, int kind, const uint64_t *a, uint64_t na,
                            const uint64_t *b, uint64_t nb, ProgramOutput *out,
                            std::string *err);
void sg_program(const EdgeInput &graph, int grid_size, int block_size, ProgramOutput *out);
void cspa_program(const EdgeInput &assign, const EdgeInput &dereference, int grid_size,
                  int block_size, ProgramOutput *out);
