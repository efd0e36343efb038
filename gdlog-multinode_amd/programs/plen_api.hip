// plen_api.hip -- the reference's path_length program written as the
// reference writes it (test/path_length.cu:49-124): path is an arity-3
// relation with a dependent hop-count column, and the join's generator hook
// computes newt[2] = outer[2] + 1, so the hook runs on the device for every
// produced tuple (gdlog/hook_exec.h) instead of becoming a column map.  The
// same RA-op list, relation names and hook as the reference; the bespoke
// PLEN program (plen.hip) derives the same result from the TC encoding.
#include <cstring>

#include "gdlog/exception.h"
#include "programs.h"

using namespace gdlog;

namespace {

// test/path_length.cu:49-54
__device__ void reorder_path(tuple_type inner, tuple_type outer, tuple_type newt) {
    newt[0] = inner[1];
    newt[1] = outer[1];
    newt[2] = outer[2] + 1;
}
__device__ tuple_generator_hook reorder_path_device = reorder_path;

}  // namespace

void path_length_hooks_program(const EdgeInput &graph, int grid_size, int block_size,
                               ProgramOutput *out, std::vector<uint32_t> *lens) {
    double t0 = now_ms();
    // the graph's rows on the host (test/path_length.cu:70-95 builds the
    // path rows (from, to, 1) and the reversed edge rows there)
    const tuple_size_t n = graph.rows;
    std::vector<column_type> raw_graph(2 * n);
    if (graph.host) {
        std::memcpy(raw_graph.data(), graph.host, 2 * n * sizeof(column_type));
    } else if (n) {
        std::vector<uint64_t> keys(n);
        check(gdl_memcpy_d2h(context(), keys.data(), graph.dev, n * sizeof(uint64_t)), "edges D2H");
        for (tuple_size_t i = 0; i < n; ++i) {
            raw_graph[2 * i] = keys[i] >> 32;
            raw_graph[2 * i + 1] = keys[i] & 0xffffffffull;
        }
    }
    std::vector<column_type> raw_path(3 * n), raw_reverse(2 * n);
    for (tuple_size_t i = 0; i < n; ++i) {
        raw_path[i * 3] = raw_graph[i * 2];
        raw_path[i * 3 + 1] = raw_graph[i * 2 + 1];
        raw_path[i * 3 + 2] = 1;
    }
    for (tuple_size_t i = 0; i < n; ++i) {
        raw_reverse[i * 2 + 1] = raw_graph[i * 2];
        raw_reverse[i * 2] = raw_graph[i * 2 + 1];
    }

    Relation *edge_2__2_1 = new Relation();
    Relation *path_3__1_2_3 = new Relation();
    path_3__1_2_3->index_flag = false;
    load_relation(path_3__1_2_3, "path_3__1_2_3", 3, raw_path.data(), n, 1, 1, grid_size,
                  block_size);
    load_relation(edge_2__2_1, "edge_2__2_1", 2, raw_reverse.data(), n, 1, 0, grid_size,
                  block_size);

    LIE tc_scc(grid_size, block_size);
    tc_scc.add_relations(edge_2__2_1, true);
    tc_scc.add_relations(path_3__1_2_3, false);
    float join_detail[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    tuple_generator_hook reorder_path_host;
    if (hipMemcpyFromSymbol(&reorder_path_host, (const void *)&reorder_path_device,
                            sizeof(tuple_generator_hook)) != hipSuccess)
        throw Error(GDL_ERR_HIP, "path_length: hipMemcpyFromSymbol of the hook");
    tc_scc.add_ra(RelationalJoin(edge_2__2_1, FULL, path_3__1_2_3, DELTA, path_3__1_2_3,
                                 reorder_path_host, nullptr, LEFT, grid_size, block_size,
                                 join_detail));
    double t1 = now_ms();
    tc_scc.fixpoint_loop();
    double t2 = now_ms();

    // the result: (from, to) pairs in key order and their hop counts
    GHashRelContainer *full = path_3__1_2_3->full;
    const std::vector<uint64_t> keys = relation_keys(full);
    std::vector<uint64_t> pairs(keys.size());
    std::vector<uint32_t> l(keys.size());
    uint64_t longest = 0, cs = 0;
    for (size_t i = 0; i < keys.size(); ++i) {
        column_type c[3];
        unpack_key(path_3__1_2_3->layout, keys[i], c);
        pairs[i] = (c[0] << 32) | c[1];
        l[i] = (uint32_t)c[2];
        longest = std::max<uint64_t>(longest, c[2]);
        uint64_t k = pairs[i];  // checksum over the pairs (fmix64, as gdl_checksum_u64)
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdull;
        k ^= k >> 33;
        k *= 0xc4ceb9fe1a85ec53ull;
        k ^= k >> 33;
        cs += k;
    }
    out->load_ms = t1 - t0;
    out->fixpoint_ms = t2 - t1;
    out->iterations = tc_scc.iteration_count;
    out->trace = tc_scc.trace;
    out->derived = sum_delta(tc_scc.trace);
    out->n_main = keys.size();
    out->checksum = cs;
    out->sizes[0] = longest;
    if (out->want_keys) out->main_keys = std::move(pairs);
    if (lens) *lens = std::move(l);
    free_relation(edge_2__2_1);
    free_relation(path_3__1_2_3);
    delete edge_2__2_1;
    delete path_3__1_2_3;
}
