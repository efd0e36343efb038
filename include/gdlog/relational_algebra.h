// gdlog/relational_algebra.h -- RA operators of the gdlog C++ API.
//
// Same constructors, fields and operator() as the reference
// (include/relational_algebra.cuh:18-117): RelationalJoin (binary hash join
// into output_rel->newt), RelationalCopy (projection into dest newt) and
// RelationalACopy (re-indexing copy of one relation's newt into another's).
// The hooks are turned into column maps when the operator is constructed
// (gdlog/hook_probe.h); hooks that are not column selections run through
// kernels of the same translation unit (gdlog/hook_exec.h).  Constructors
// must therefore be compiled with hipcc in the translation unit that defines
// the hooks.
#pragma once

#include <variant>

#include "gdlog/hook_exec.h"
#include "gdlog/hook_probe.h"
#include "gdlog/relation.h"

// Chunk size of the reference's host-side reduction (relational_algebra.cuh:8-10);
// kept for API compatibility (the engine's scan has no 32-bit limit).
#ifndef MAX_REDUCE_SIZE
#define MAX_REDUCE_SIZE 80000000
#endif

struct RelationalJoin {
    Relation *inner_rel;
    RelationVersion inner_ver;
    Relation *outer_rel;
    RelationVersion outer_ver;
    Relation *output_rel;
    tuple_generator_hook tuple_generator;
    tuple_predicate tuple_pred;
    JoinDirection direction;
    int grid_size;
    int block_size;
    // benchmark flag: drop the join result instead of loading it into newt
    bool disable_load = false;
    // accumulated phase times (seconds): [0] count [1] scan [2] materialize
    // [3] sort [4] unique [5] index [6] merge [7] post-merge unique
    float *detail_time;

    gdlog::HookMap gen_map;
    gdlog::PredMap pred_map;
    // Runs the hooks themselves when they are not a column selection and a
    // column comparison (gdlog/hook_exec.h); set by the hook constructor.
    gdlog::HookJoinLaunch hook_join = nullptr;
    bool uses_hooks() const { return !gen_map.ok || !pred_map.ok; }
    // The reference's handling of an empty inner: it zeroes the outer's tuple
    // count (join.cu:42-45), so later operators reading the same outer version
    // in that iteration see nothing.  The column-map constructor turns it off:
    // generated plans read one delta from several joins.
    bool empty_inner_zeroes_outer = true;

    RelationalJoin(Relation *inner_rel, RelationVersion inner_ver, Relation *outer_rel,
                   RelationVersion outer_ver, Relation *output_rel, tuple_generator_hook tp_gen,
                   tuple_predicate tp_pred, JoinDirection direction, int grid_size, int block_size,
                   float *detail_time)
        : inner_rel(inner_rel), inner_ver(inner_ver), outer_rel(outer_rel), outer_ver(outer_ver),
          output_rel(output_rel), tuple_generator(tp_gen), tuple_pred(tp_pred),
          direction(direction), grid_size(grid_size), block_size(block_size),
          detail_time(detail_time) {
        gen_map = gdlog::probe_generator(tp_gen, output_rel ? output_rel->arity : 0);
        pred_map = gdlog::probe_predicate(tp_pred, output_rel ? output_rel->arity : 0);
        hook_join = &gdlog::hook_exec_detail::launch_join;
    }

    // Column-map form (an extension for generated plans, gdlog/datalog.h): the
    // output columns and the predicate are given instead of probed from hooks.
    RelationalJoin(Relation *inner_rel, RelationVersion inner_ver, Relation *outer_rel,
                   RelationVersion outer_ver, Relation *output_rel, gdl_colmap map,
                   gdl_pred pred)
        : inner_rel(inner_rel), inner_ver(inner_ver), outer_rel(outer_rel), outer_ver(outer_ver),
          output_rel(output_rel), tuple_generator(nullptr), tuple_pred(nullptr),
          direction(LEFT), grid_size(0), block_size(0), detail_time(nullptr),
          empty_inner_zeroes_outer(false) {
        gen_map.ok = true;
        gen_map.map = map;
        pred_map.ok = true;
        pred_map.pred = pred;
    }

    void operator()();

    // Runs up to max_bodies iterations of a LIE whose only operator is this
    // join and whose only update relation is its outer (DELTA) and fused
    // output, with the inner a static FULL (TC's shape, test/tc.cu:80-149), as
    // one stream-ordered loop (gdl_fixpoint_linear).  The output's delta must
    // be the last segment of its FULL log.  new_tuples / raw_tuples get each
    // body's counts.  Returns false, doing nothing, when the join does not
    // have that shape.
    bool run_linear(int max_bodies, std::vector<uint64_t> *new_tuples,
                    std::vector<uint64_t> *raw_tuples);
};

struct RelationalCopy {
    Relation *src_rel;
    RelationVersion src_ver;
    Relation *dest_rel;
    tuple_copy_hook tuple_generator;
    tuple_predicate tuple_pred;  // stored, not applied (as in the reference, copy.cu)
    int grid_size;
    int block_size;
    bool copied = false;

    gdlog::HookMap copy_map;
    gdlog::HookCopyLaunch hook_copy = nullptr;  // a copy hook that is not a column selection
    // The reference's handling of an empty source: it zeroes the
    // destination's newt count (copy.cu:23-26), dropping what earlier
    // operators put there in that iteration.  Off in the column-map form.
    bool empty_source_zeroes_dest = true;
    // Column-map form only: copied tuples must pass this comparison (over the
    // destination's columns); the hook form ignores predicates, as the
    // reference's copy does.
    gdl_pred filter{GDL_PRED_NONE, 0, 0};

    RelationalCopy(Relation *src, RelationVersion src_ver, Relation *dest,
                   tuple_copy_hook tuple_generator, tuple_predicate tuple_pred, int grid_size,
                   int block_size)
        : src_rel(src), src_ver(src_ver), dest_rel(dest), tuple_generator(tuple_generator),
          tuple_pred(tuple_pred), grid_size(grid_size), block_size(block_size) {
        copy_map = gdlog::probe_copy(tuple_generator, dest ? dest->arity : 0);
        hook_copy = &gdlog::hook_exec_detail::launch_copy;
    }

    // Column-map form (source columns as GDL_COL_INNER0 / GDL_COL_INNER1),
    // with an optional filter on the copied tuples.
    RelationalCopy(Relation *src, RelationVersion src_ver, Relation *dest, gdl_colmap map,
                   gdl_pred filter = gdl_pred{GDL_PRED_NONE, 0, 0})
        : src_rel(src), src_ver(src_ver), dest_rel(dest), tuple_generator(nullptr),
          tuple_pred(nullptr), grid_size(0), block_size(0), empty_source_zeroes_dest(false),
          filter(filter) {
        copy_map.ok = true;
        copy_map.map = map;
    }

    void operator()();
};

struct RelationalACopy {
    Relation *src_rel;
    Relation *dest_rel;
    tuple_copy_hook tuple_generator;
    tuple_predicate tuple_pred;  // stored, not applied (as in the reference, acopy.cu)
    int grid_size;
    int block_size;

    gdlog::HookMap copy_map;
    gdlog::HookCopyLaunch hook_copy = nullptr;  // a copy hook that is not a column selection

    RelationalACopy(Relation *src, Relation *dest, tuple_copy_hook tuple_generator,
                    tuple_predicate tuple_pred, int grid_size, int block_size)
        : src_rel(src), dest_rel(dest), tuple_generator(tuple_generator), tuple_pred(tuple_pred),
          grid_size(grid_size), block_size(block_size) {
        copy_map = gdlog::probe_copy(tuple_generator, dest ? dest->arity : 0);
        hook_copy = &gdlog::hook_exec_detail::launch_copy;
    }

    // Column-map form (source columns as GDL_COL_INNER0 / GDL_COL_INNER1).
    RelationalACopy(Relation *src, Relation *dest, gdl_colmap map)
        : src_rel(src), dest_rel(dest), tuple_generator(nullptr), tuple_pred(nullptr),
          grid_size(0), block_size(0) {
        copy_map.ok = true;
        copy_map.map = map;
    }

    void operator()();
};

using ra_op = std::variant<RelationalJoin, RelationalCopy, RelationalACopy>;

enum RAtypes { JOIN, COPY, ACOPY };
