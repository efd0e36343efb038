// tc.hip -- transitive closure on the gdlog C++ API.
//
// Same RA encoding as the reference program (test/tc.cu:80-149):
//   path(a, c) :- edge(a, b), path(b, c)
// with edge_2__2_1 (reversed edges, indexed on b) as the FULL inner side and
// path DELTA as the outer side.
#include "programs.h"

__device__ void reorder_path(tuple_type inner, tuple_type outer, tuple_type newt) {
    newt[0] = inner[1];
    newt[1] = outer[1];
}
__device__ tuple_generator_hook reorder_path_device = reorder_path;

void tc_program(const EdgeInput &graph, int max_iteration, int grid_size, int block_size,
                ProgramOutput *out) {
    double t0 = now_ms();
    Relation *edge_2__2_1 = new Relation();
    Relation *path_2__1_2 = new Relation();
    path_2__1_2->index_flag = false;
    graph.load(path_2__1_2, "path_2__1_2", false, grid_size, block_size);
    graph.load(edge_2__2_1, "edge_2__2_1", true, grid_size, block_size);

    LIE tc_scc(grid_size, block_size);
    tc_scc.reload_full_flag = false;
    if (max_iteration >= 0) tc_scc.max_iteration = max_iteration;
    tc_scc.add_relations(edge_2__2_1, true);
    tc_scc.add_relations(path_2__1_2, false);
    float join_detail[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    tuple_generator_hook reorder_path_host = host_hook(reorder_path_device);
    tc_scc.add_ra(RelationalJoin(edge_2__2_1, FULL, path_2__1_2, DELTA, path_2__1_2,
                                 reorder_path_host, nullptr, LEFT, grid_size, block_size,
                                 join_detail));
    double t1 = now_ms();
    tc_scc.fixpoint_loop();
    double t2 = now_ms();

    out->load_ms = t1 - t0;
    out->fixpoint_ms = t2 - t1;
    out->iterations = tc_scc.iteration_count;
    out->trace = tc_scc.trace;
    out->derived = sum_delta(tc_scc.trace);
    out->n_main = path_2__1_2->full->tuple_counts;
    out->checksum = relation_checksum(path_2__1_2->full);
    if (out->want_keys) out->main_keys = relation_keys(path_2__1_2->full);
    free_relation(edge_2__2_1);
    free_relation(path_2__1_2);
    delete edge_2__2_1;
    delete path_2__1_2;
}
