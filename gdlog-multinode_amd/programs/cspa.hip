// cspa.hip -- context-sensitive points-to analysis on the gdlog C++ API.
//
// Same RA encoding as the reference program (test/cspa.cu:100-379), which
// realises the ten rules of test/souffle/cspa.dl:25-39 with two LIEs: an
// init SCC of projections from assign, and the analysis SCC of twelve joins
// (three through tmp relations) plus two ACopy mirror syncs.
#include "programs.h"

__device__ void cp_2_1__1(tuple_type input, tuple_type outpt) {
    outpt[0] = input[0];
    outpt[1] = input[0];
}
__device__ tuple_copy_hook cp_2_1__1_device = cp_2_1__1;
__device__ void cp_2_1__2(tuple_type input, tuple_type outpt) {
    outpt[0] = input[1];
    outpt[1] = input[1];
}
__device__ tuple_copy_hook cp_2_1__2_device = cp_2_1__2;
__device__ void cp_2_1__1_2(tuple_type input, tuple_type outpt) {
    outpt[0] = input[1];
    outpt[1] = input[0];
}
__device__ tuple_copy_hook cp_2_1__1_2_device = cp_2_1__1_2;

__device__ void join_10_11(tuple_type inner, tuple_type outer, tuple_type output) {
    output[1] = inner[1];
    output[0] = outer[1];
}
__device__ tuple_generator_hook join_10_11_device = join_10_11;
__device__ void join_01_11(tuple_type inner, tuple_type outer, tuple_type output) {
    output[0] = inner[1];
    output[1] = outer[1];
}
__device__ tuple_generator_hook join_01_11_device = join_01_11;

void cspa_program(const EdgeInput &assign, const EdgeInput &dereference, int grid_size,
                  int block_size, ProgramOutput *out) {
    double t0 = now_ms();
    Relation *assign_2__2_1 = new Relation();
    assign.load(assign_2__2_1, "assign_2__2_1", true, grid_size, block_size);
    Relation *dereference_2__1_2 = new Relation();
    dereference.load(dereference_2__1_2, "dereference_2__1_2", false, grid_size, block_size);
    Relation *dereference_2__2_1 = new Relation();
    dereference.load(dereference_2__2_1, "dereference_2__2_1", true, grid_size, block_size);

    Relation *value_flow_2__1_2 = new Relation();
    load_relation(value_flow_2__1_2, "value_flow_2__1_2", 2, nullptr, 0, 1, 0, grid_size,
                  block_size);
    Relation *value_flow_2__2_1 = new Relation();
    load_relation(value_flow_2__2_1, "value_flow_2__2_1", 2, nullptr, 0, 1, 0, grid_size,
                  block_size);
    Relation *memory_alias_2__1_2 = new Relation();
    load_relation(memory_alias_2__1_2, "memory_alias_2__1_2", 2, nullptr, 0, 1, 0, grid_size,
                  block_size);
    Relation *memory_alias_2__2_1 = new Relation();
    load_relation(memory_alias_2__2_1, "memory_alias_2__2_1", 2, nullptr, 0, 1, 0, grid_size,
                  block_size);

    tuple_copy_hook cp_2_1__1_host = host_hook(cp_2_1__1_device);
    tuple_copy_hook cp_2_1__2_host = host_hook(cp_2_1__2_device);
    tuple_copy_hook cp_2_1__1_2_host = host_hook(cp_2_1__1_2_device);
    tuple_generator_hook join_10_11_host = host_hook(join_10_11_device);
    tuple_generator_hook join_01_11_host = host_hook(join_01_11_device);

    double t1 = now_ms();
    LIE init_scc(grid_size, block_size);
    init_scc.add_relations(value_flow_2__1_2, false);
    init_scc.add_relations(value_flow_2__2_1, false);
    init_scc.add_relations(memory_alias_2__1_2, false);
    init_scc.add_relations(memory_alias_2__2_1, false);
    init_scc.add_relations(assign_2__2_1, true);
    // ValueFlow(x, x) :- assign(y, x).  ValueFlow(x, x) :- assign(x, y).
    // ValueFlow(y, x) :- assign(y, x).
    init_scc.add_ra(RelationalCopy(assign_2__2_1, FULL, value_flow_2__1_2, cp_2_1__1_host,
                                   nullptr, grid_size, block_size));
    init_scc.add_ra(RelationalCopy(assign_2__2_1, FULL, value_flow_2__1_2, cp_2_1__2_host,
                                   nullptr, grid_size, block_size));
    init_scc.add_ra(RelationalCopy(assign_2__2_1, FULL, value_flow_2__1_2, cp_2_1__1_2_host,
                                   nullptr, grid_size, block_size));
    // MemoryAlias(x, x) :- assign(y, x).  MemoryAlias(x, x) :- assign(x, y).
    init_scc.add_ra(RelationalCopy(assign_2__2_1, FULL, memory_alias_2__1_2, cp_2_1__1_host,
                                   nullptr, grid_size, block_size));
    init_scc.add_ra(RelationalCopy(assign_2__2_1, FULL, memory_alias_2__1_2, cp_2_1__2_host,
                                   nullptr, grid_size, block_size));
    // column-reversed mirrors
    init_scc.add_ra(RelationalCopy(value_flow_2__1_2, DELTA, value_flow_2__2_1, cp_2_1__1_2_host,
                                   nullptr, grid_size, block_size));
    init_scc.add_ra(RelationalCopy(memory_alias_2__1_2, DELTA, memory_alias_2__2_1,
                                   cp_2_1__1_2_host, nullptr, grid_size, block_size));
    init_scc.fixpoint_loop();

    Relation *value_flow_forward_2__1_2 = new Relation();
    load_relation(value_flow_forward_2__1_2, "value_flow_forward_2__1_2", 2, nullptr, 0, 1, 0,
                  grid_size, block_size);
    Relation *value_flow_forward_2__2_1 = new Relation();
    load_relation(value_flow_forward_2__2_1, "value_flow_forward_2__2_1", 2, nullptr, 0, 1, 0,
                  grid_size, block_size);
    Relation *value_alias_2__1_2 = new Relation();
    value_alias_2__1_2->index_flag = false;
    load_relation(value_alias_2__1_2, "value_alias_2__1_2", 2, nullptr, 0, 1, 0, grid_size,
                  block_size);
    Relation *tmp_rel_def = new Relation();
    tmp_rel_def->index_flag = false;
    load_relation(tmp_rel_def, "tmp_rel_def", 2, nullptr, 0, 1, 0, grid_size, block_size);
    Relation *tmp_rel_ma1 = new Relation();
    tmp_rel_ma1->index_flag = false;
    load_relation(tmp_rel_ma1, "tmp_rel_ma1", 2, nullptr, 0, 1, 0, grid_size, block_size, true);
    Relation *tmp_rel_ma2 = new Relation();
    tmp_rel_ma2->index_flag = false;
    load_relation(tmp_rel_ma2, "tmp_rel_ma2", 2, nullptr, 0, 1, 0, grid_size, block_size, true);

    LIE analysis_scc(grid_size, block_size);
    analysis_scc.add_relations(assign_2__2_1, true);
    analysis_scc.add_relations(dereference_2__1_2, true);
    analysis_scc.add_relations(dereference_2__2_1, true);
    analysis_scc.add_relations(value_flow_2__1_2, false);
    analysis_scc.add_relations(value_flow_2__2_1, false);
    analysis_scc.add_relations(memory_alias_2__1_2, false);
    analysis_scc.add_relations(memory_alias_2__2_1, false);
    analysis_scc.add_relations(value_alias_2__1_2, false);
    analysis_scc.add_tmp_relation(tmp_rel_def);
    analysis_scc.add_tmp_relation(tmp_rel_ma1);
    analysis_scc.add_tmp_relation(tmp_rel_ma2);
    float join_detail[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // ValueFlow(x, y) :- ValueFlow(x, z), ValueFlow(z, y).
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, value_flow_2__2_1, DELTA,
                                       value_flow_2__1_2, join_10_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(value_flow_2__2_1, FULL, value_flow_2__1_2, DELTA,
                                       value_flow_2__1_2, join_01_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    // ValueAlias(x, y) :- ValueFlow(z, x), ValueFlow(z, y).
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, value_flow_2__1_2, DELTA,
                                       value_alias_2__1_2, join_01_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, value_flow_2__1_2, DELTA,
                                       value_alias_2__1_2, join_10_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    // ValueFlow(x, y) :- Assign(x, z), MemoryAlias(z, y).
    analysis_scc.add_ra(RelationalJoin(assign_2__2_1, FULL, memory_alias_2__1_2, DELTA,
                                       value_flow_2__1_2, join_01_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    // tmp_rel_def(z, x) :- Dereference(y, x), ValueAlias(y, z)
    analysis_scc.add_ra(RelationalJoin(dereference_2__1_2, FULL, value_alias_2__1_2, DELTA,
                                       tmp_rel_def, join_10_11_host, nullptr, LEFT, grid_size,
                                       block_size, join_detail));
    // MemoryAlias(x, w) :- Dereference(z, w), tmp_rel_def(z, x)
    analysis_scc.add_ra(RelationalJoin(dereference_2__1_2, FULL, tmp_rel_def, NEWT,
                                       memory_alias_2__1_2, join_10_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    // ValueAlias(x, y) :- ValueFlow(z, x), MemoryAlias(z, w), ValueFlow(w, y).
    analysis_scc.add_ra(RelationalJoin(memory_alias_2__1_2, FULL, value_flow_2__1_2, DELTA,
                                       tmp_rel_ma1, join_01_11_host, nullptr, LEFT, grid_size,
                                       block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, memory_alias_2__1_2, DELTA,
                                       tmp_rel_ma1, join_10_11_host, nullptr, LEFT, grid_size,
                                       block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, tmp_rel_ma1, NEWT,
                                       value_alias_2__1_2, join_10_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(memory_alias_2__2_1, FULL, value_flow_2__1_2, DELTA,
                                       tmp_rel_ma2, join_01_11_host, nullptr, LEFT, grid_size,
                                       block_size, join_detail));
    analysis_scc.add_ra(RelationalJoin(value_flow_2__1_2, FULL, tmp_rel_ma2, NEWT,
                                       value_alias_2__1_2, join_01_11_host, nullptr, LEFT,
                                       grid_size, block_size, join_detail));
    analysis_scc.add_ra(RelationalACopy(value_flow_2__1_2, value_flow_2__2_1, cp_2_1__1_2_host,
                                        nullptr, grid_size, block_size));
    analysis_scc.add_ra(RelationalACopy(memory_alias_2__1_2, memory_alias_2__2_1,
                                        cp_2_1__1_2_host, nullptr, grid_size, block_size));
    analysis_scc.fixpoint_loop();
    double t2 = now_ms();

    out->load_ms = t1 - t0;
    out->fixpoint_ms = t2 - t1;
    out->init_iterations = init_scc.iteration_count;
    out->iterations = analysis_scc.iteration_count;
    out->trace_init = init_scc.trace;
    out->trace = analysis_scc.trace;
    out->derived = sum_delta(init_scc.trace) + sum_delta(analysis_scc.trace);
    out->n_main = value_alias_2__1_2->full->tuple_counts;
    out->sizes[0] = value_flow_2__1_2->full->tuple_counts;
    out->sizes[1] = memory_alias_2__1_2->full->tuple_counts;
    out->sizes[2] = value_alias_2__1_2->full->tuple_counts;
    out->sizes[3] = value_flow_2__2_1->full->tuple_counts;
    out->checksum = relation_checksum(value_alias_2__1_2->full);
    if (out->want_keys) {
        out->main_keys = relation_keys(value_alias_2__1_2->full);
        out->extra_keys[0] = relation_keys(value_flow_2__1_2->full);
        out->extra_keys[1] = relation_keys(memory_alias_2__1_2->full);
    }
    for (Relation *r : {assign_2__2_1, dereference_2__1_2, dereference_2__2_1, value_flow_2__1_2,
                        value_flow_2__2_1, memory_alias_2__1_2, memory_alias_2__2_1,
                        value_flow_forward_2__1_2, value_flow_forward_2__2_1, value_alias_2__1_2,
                        tmp_rel_def, tmp_rel_ma1, tmp_rel_ma2}) {
        free_relation(r);
        delete r;
    }
}
