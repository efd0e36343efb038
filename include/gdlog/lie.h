// gdlog/lie.h -- the logical inference engine (semi-naive fixpoint of one
// SCC), same interface as the reference (include/lie.cuh:11-65).
#pragma once

#include <climits>
#include <vector>

#include "gdlog/relational_algebra.h"

// One record per (iteration, update relation), in the order the fixpoint
// visits them: newt size before the set difference, new tuples (delta) and
// full size after the merge.
struct IterationRecord {
    int iteration;
    int relation;
    tuple_size_t newt;
    tuple_size_t delta;
    tuple_size_t full;
};

struct LIE {
    std::vector<ra_op> ra_ops;
    std::vector<Relation *> update_relations;
    std::vector<Relation *> static_relations;
    std::vector<Relation *> tmp_relations;

    int grid_size;
    int block_size;

    bool reload_full_flag = true;
    int max_iteration = INT_MAX;

    // filled by fixpoint_loop
    int iteration_count = 0;
    std::vector<IterationRecord> trace;

    LIE(int grid_size, int block_size) : grid_size(grid_size), block_size(block_size) {}

    void fixpoint_loop();
    void add_relations(Relation *rel, bool static_flag);
    void add_tmp_relation(Relation *rel);
    void add_ra(ra_op op);
};
