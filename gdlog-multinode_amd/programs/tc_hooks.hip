// tc_hooks.hip -- TC on the gdlog C++ API with hooks that are not column
// selections, so they run on the device for every produced tuple
// (gdlog/hook_exec.h).  Same RA encoding as the reference's TC program
// (test/tc.cu:80-149); only the hooks differ:
//   variant 0: the generator of test/tc.cu:61-65 and the predicate
//              t[1] != 42 (a comparison with a constant, not between columns):
//              paths into node 42 are never derived;
//   variant 1: the generator newt = (inner[1], (outer[1] + 1) mod 1000): an
//              arithmetic column (the closure of the shifted relation), no
//              predicate;
//   variant 2: variant 1's generator as a copy hook, then TC: a RelationalCopy
//              from the edges FULL with newt = (src[0], (src[1] * 7) mod 1000)
//              seeds path, and the join uses test/tc.cu's generator.
#include "programs.h"

namespace {

__device__ void tc_gen(tuple_type inner, tuple_type outer, tuple_type newt) {
    newt[0] = inner[1];
    newt[1] = outer[1];
}
__device__ bool not_into_42(tuple_type t) { return t[1] != 42; }
__device__ void shift_gen(tuple_type inner, tuple_type outer, tuple_type newt) {
    newt[0] = inner[1];
    newt[1] = (outer[1] + 1) % 1000;
}
__device__ void scale_copy(tuple_type src, tuple_type dst) {
    dst[0] = src[0];
    dst[1] = (src[1] * 7) % 1000;
}
__device__ tuple_generator_hook tc_gen_device = tc_gen;
__device__ tuple_predicate not_into_42_device = not_into_42;
__device__ tuple_generator_hook shift_gen_device = shift_gen;
__device__ tuple_copy_hook scale_copy_device = scale_copy;

}  // namespace

void tc_hooks_program(const EdgeInput &graph, int variant, int grid_size, int block_size,
                      ProgramOutput *out) {
    double t0 = now_ms();
    Relation *edge_2__2_1 = new Relation();
    Relation *path_2__1_2 = new Relation();
    path_2__1_2->index_flag = false;
    graph.load(edge_2__2_1, "edge_2__2_1", true, grid_size, block_size);
    Relation *edge_2__1_2 = nullptr;
    LIE init(grid_size, block_size);
    if (variant == 2) {
        // path := scale_copy(edges): one LIE running the copy once
        edge_2__1_2 = new Relation();
        graph.load(edge_2__1_2, "edge_2__1_2", false, grid_size, block_size);
        load_relation(path_2__1_2, "path_2__1_2", 2, nullptr, 0, 1, 0, grid_size, block_size);
        init.add_relations(edge_2__1_2, true);
        init.add_relations(path_2__1_2, false);
        init.add_ra(RelationalCopy(edge_2__1_2, FULL, path_2__1_2, host_hook(scale_copy_device),
                                   nullptr, grid_size, block_size));
        init.fixpoint_loop();
        out->init_iterations = init.iteration_count;
        out->trace_init = init.trace;
    } else {
        graph.load(path_2__1_2, "path_2__1_2", false, grid_size, block_size);
    }
    LIE tc_scc(grid_size, block_size);
    tc_scc.reload_full_flag = false;
    tc_scc.add_relations(edge_2__2_1, true);
    tc_scc.add_relations(path_2__1_2, false);
    float join_detail[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const tuple_generator_hook gen =
        host_hook(variant == 1 ? shift_gen_device : tc_gen_device);
    const tuple_predicate pred = variant == 0 ? host_hook(not_into_42_device) : nullptr;
    tc_scc.add_ra(RelationalJoin(edge_2__2_1, FULL, path_2__1_2, DELTA, path_2__1_2, gen, pred,
                                 LEFT, grid_size, block_size, join_detail));
    double t1 = now_ms();
    tc_scc.fixpoint_loop();
    double t2 = now_ms();
    out->load_ms = t1 - t0;
    out->fixpoint_ms = t2 - t1;
    out->iterations = tc_scc.iteration_count;
    out->trace = tc_scc.trace;
    out->derived = sum_delta(tc_scc.trace) + sum_delta(init.trace);
    out->n_main = path_2__1_2->full->tuple_counts;
    out->checksum = relation_checksum(path_2__1_2->full);
    if (out->want_keys) out->main_keys = relation_keys(path_2__1_2->full);
    free_relation(edge_2__2_1);
    free_relation(path_2__1_2);
    delete edge_2__2_1;
    delete path_2__1_2;
    if (edge_2__1_2) {
        free_relation(edge_2__1_2);
        delete edge_2__1_2;
    }
}
