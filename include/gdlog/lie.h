// gdlog/lie.h -- the logical inference engine (semi-naive fixpoint of one
// SCC), same interface as the reference (include/lie.cuh:11-65).
#pragma once

#include <climits>
#include <string>
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
    // An extension: an update relation kept as an append log (FULL not read as
    // a join inner) ends the fixpoint with FULL in log order -- its loaded
    // tuples, then each iteration's delta -- instead of sorted
    // (full->sorted == false), so a tuple's position tells the iteration that
    // first derived it.
    bool keep_full_log = false;

    // Checkpoint / resume (an extension: the reference has none).
    // checkpoint_file: the loop writes a snapshot there (every update
    // relation's FULL and delta, the iteration counter, the trace, and a
    // fingerprint of the static relations) every checkpoint_every iterations
    // (0: none in between), when max_iteration or stop_after stops it, and
    // at the fixpoint.  resume_file: when that file exists, the loop starts
    // from its snapshot instead of iteration 0 (a finished snapshot restores
    // the result without running a body); the trace and result are those of
    // an uninterrupted run.  stop_after >= 0: right after the snapshot at that
    // iteration the loop throws gdlog::Error(GDL_ERR_BUSY) (a preempted run).
    // With none of them set, the thread's defaults apply (set_checkpoint_defaults).
    std::string checkpoint_file;
    int checkpoint_every = 0;
    std::string resume_file;
    int stop_after = -1;

    // filled by fixpoint_loop
    int iteration_count = 0;
    std::vector<IterationRecord> trace;

    LIE(int grid_size, int block_size) : grid_size(grid_size), block_size(block_size) {}

    void fixpoint_loop();
    void add_relations(Relation *rel, bool static_flag);
    void add_tmp_relation(Relation *rel);
    void add_ra(ra_op op);
};

namespace gdlog {
// Thread defaults for LIEs that set no checkpoint fields themselves: the
// k-th LIE whose fixpoint_loop runs after reset_checkpoint_sequence() uses
// <dir>/lie<k>.ckpt as its checkpoint file (and, with resume, as its resume
// file).  dir = nullptr or "" turns the defaults off.
void set_checkpoint_defaults(const char *dir, int every, bool resume, int stop_after);
void reset_checkpoint_sequence();
}  // namespace gdlog
