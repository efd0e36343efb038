/*
 * gdl_programs.h -- C ABI of the TC, SG and CSPA programs built on the gdlog
 * C++ API (the reference's benchmark drivers test/tc.cu, test/sg.cu and
 * test/cspa.cu, whose analysis_bench functions these replace), for callers
 * that cannot link C++: tests, benchmarks, foreign-function bindings.
 *
 * Edges are packed keys (src << 32) | dst.  With on_device = 0 they are host
 * arrays loaded through load_relation (the reference's host path); with
 * on_device = 1 they are device arrays of the given context.
 */
#ifndef GDL_PROGRAMS_H
#define GDL_PROGRAMS_H

#include <stdint.h>

#include "gdl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdl_trace_rec {
    int32_t iteration;
    int32_t relation; /* index in the LIE's update-relation order */
    uint64_t newt;    /* newt size before the set difference */
    uint64_t delta;   /* new tuples */
    uint64_t full;    /* full size after the merge */
} gdl_trace_rec;

typedef struct gdl_prog_result {
    int32_t iterations;      /* iteration counter of the (last) LIE */
    int32_t init_iterations; /* SG / CSPA init LIE */
    uint64_t derived;        /* sum of |delta| over update relations and iterations */
    uint64_t n_out;          /* size of the main output (path / sg / ValueAlias) */
    uint64_t sizes[4];       /* CSPA: ValueFlow, MemoryAlias, ValueAlias, ValueFlow mirror */
    double load_ms;          /* relation loading (sort, unique) */
    double fixpoint_ms;      /* fixpoint loop(s), device-synchronised */
    uint64_t n_trace;        /* records of the (last) LIE */
    uint64_t n_trace_init;   /* records of the init LIE */
    uint64_t checksum;       /* gdl_checksum_u64 of the main output */
} gdl_prog_result;

/* Optional output buffers (NULL / 0 capacity to skip).  A buffer too small
 * for its data yields GDL_ERR_OVERFLOW after the run (counts still set). */
typedef struct gdl_prog_buffers {
    uint64_t *out;           /* sorted keys of the main output */
    uint64_t out_cap;
    uint64_t *extra0;        /* CSPA: ValueFlow */
    uint64_t extra0_cap;
    uint64_t *extra1;        /* CSPA: MemoryAlias */
    uint64_t extra1_cap;
    gdl_trace_rec *trace;
    uint64_t trace_cap;
    gdl_trace_rec *trace_init;
    uint64_t trace_init_cap;
} gdl_prog_buffers;

/* max_iteration < 0: unbounded (INT_MAX, lie.cuh:30). */
gdl_status gdl_prog_tc(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                       int max_iteration, gdl_prog_result *res, const gdl_prog_buffers *bufs);
/* TC with device hooks that are not column selections (an extension that
 * exercises the reference's per-tuple hook calls, relation.cu:348-358 and
 * 390-403): variant 0 keeps test/tc.cu's generator and adds the predicate
 * t[1] != 42; variant 1 uses the generator (inner[1], (outer[1] + 1) mod
 * 1000); variant 2 seeds path with a RelationalCopy whose hook is
 * (src[0], (src[1] * 7) mod 1000), then runs test/tc.cu's join (its trace in
 * trace_init).  Outputs as gdl_prog_tc. */
gdl_status gdl_prog_tc_hooks(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                             int variant, gdl_prog_result *res, const gdl_prog_buffers *bufs);
gdl_status gdl_prog_sg(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                       gdl_prog_result *res, const gdl_prog_buffers *bufs);
/* All-pairs hop counts: the reference's path_length program
 * (test/path_length.cu:49-124, the PLEN executable), path(a, c, l + 1) :-
 * edge(a, b), path(b, c, l) with l a dependent column (edges load with l = 1,
 * a pair keeps the l of the iteration that first derives it).  res as for
 * TC, with sizes[0] the longest hop count.  keys / lens (host, cap entries,
 * either may be NULL): the pairs in sorted order and each pair's hop count;
 * GDL_ERR_OVERFLOW when cap < res->n_out. */
gdl_status gdl_prog_path_length(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                                gdl_prog_result *res, uint64_t *keys, uint32_t *lens,
                                uint64_t cap);
/* The same result from the reference's own encoding of path_length
 * (test/path_length.cu:49-124 unchanged: an arity-3 path relation with a
 * dependent column, load_relation(..., 3, ..., 1, 1, ...), and the generator
 * hook newt[2] = outer[2] + 1, which runs on the device for every produced
 * tuple).  Arguments and outputs as gdl_prog_path_length. */
gdl_status gdl_prog_path_length_hooks(gdl_ctx *ctx, const uint64_t *edges, uint64_t n,
                                      int on_device, gdl_prog_result *res, uint64_t *keys,
                                      uint32_t *lens, uint64_t cap);
gdl_status gdl_prog_cspa(gdl_ctx *ctx, const uint64_t *assign, uint64_t n_assign,
                         const uint64_t *dereference, uint64_t n_dereference, int on_device,
                         gdl_prog_result *res, const gdl_prog_buffers *bufs);
/* Message of the last failed gdl_prog_* call on this thread. */
const char *gdl_prog_last_error(void);

/* Checkpoint / resume of the programs' fixpoints on this thread (an
 * extension: the reference has none).  With dir != NULL, the k-th LIE that a
 * later gdl_prog_* call runs (k from 0 in each call: TC has one, SG and CSPA
 * two) writes a snapshot of its update relations' FULL and delta, its
 * iteration counter and trace to <dir>/lie<k>.ckpt every `every` iterations
 * (0: none in between), when max_iteration stops it, and at its fixpoint.
 * resume = 1: a LIE whose file exists starts from it, and the result and
 * trace are those of an uninterrupted run (the file must come from the same
 * program and inputs: GDL_ERR_INVALID_ARG otherwise).  stop_after >= 0: a LIE
 * that reaches that iteration stops right after its snapshot, and the call
 * returns GDL_ERR_BUSY (a preempted run; later LIEs of the call do not run).
 * dir = NULL turns all of this off. */
gdl_status gdl_prog_set_checkpoint(const char *dir, int every, int resume, int stop_after);

/* FULL maintenance of the fixpoint for later runs: 0 = auto (append log +
 * hash set for relations whose FULL is never a join inner), 1 = merge (sorted
 * merge every iteration, the reference's structure).  Results are identical;
 * the trace's newt field counts raw rows in append mode. */
gdl_status gdl_prog_set_full_mode(int mode);

#ifdef __cplusplus
}
#endif
#endif
