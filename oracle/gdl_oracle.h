/*
 * gdl_oracle.h -- CPU oracle for the semi-naive Datalog hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference engine's (harp-lab/gdlog) algorithm, used as the checker for the
 * HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * Representation: a tuple of arity <= 2 over u32 ids is a u64 key
 * (c0 << 32) | c1 (arity 1: c0 << 32).  Lexicographic order over the columns
 * (include/tuple.cuh:94-127, tuple_indexed_less) is then integer order, and
 * full-tuple equality (tuple.cuh:51-65, t_equal) is integer equality.
 *
 * Hooks follow the reference typedefs (include/tuple.cuh:23-25): they read
 * and write 64-bit column arrays, exactly as the reference programs' device
 * hooks do (test/tc.cu:61-78, test/sg.cu:62-91, test/cspa.cu:62-96).
 *
 * Pinning (what each function is checked against, see tests/test_oracle.py):
 *   orc_sort_unique / orc_merge_union / orc_set_difference: Python's
 *     sorted()/set operations on seeded random and adversarial inputs.
 *   orc_join: brute-force nested loops on small inputs.
 *   LIE loop + TC/SG/CSPA encodings: the reference's worked examples
 *     (data/tc.png, data/JOIN.pdf), BFS shortest-distance closed forms,
 *     cuDF SG result sizes (test/cuDF/sg.json), a scipy sparse-matrix SG
 *     fixpoint, and a naive Datalog evaluation of test/souffle/cspa.dl
 *     written in Python in the tests.
 *   The per-iteration CSPA trace (newt, delta, full of all five update
 *     relations in both LIEs) and the per-operator raw join sizes are pinned
 *     by a known-answer case derived by hand from src/lie.cu, src/join.cu,
 *     src/copy.cu and src/acopy.cu (tests/test_cspa_kat.py), which also pins
 *     the empty-source and empty-inner zeroing on LIEs built for them.
 */
#ifndef GDL_ORACLE_H
#define GDL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t orc_key;
typedef uint64_t orc_col; /* reference column_type is 8 bytes (tuple.cuh:6-8) */

typedef void (*orc_gen_hook)(const orc_col *inner, const orc_col *outer, orc_col *out);
typedef void (*orc_copy_hook)(const orc_col *src, orc_col *dst);
typedef int (*orc_pred_hook)(const orc_col *t);

enum { ORC_DELTA = 0, ORC_FULL = 1, ORC_NEWT = 2 }; /* relation.cuh:211 */

/* ---- primitives ---------------------------------------------------- */
size_t orc_sort_unique(orc_key *keys, size_t n);           /* in place; returns unique count */
void orc_sort(orc_key *keys, size_t n);                      /* in place, ascending */
size_t orc_merge_union(const orc_key *a, size_t na, const orc_key *b, size_t nb, orc_key *out);
size_t orc_set_difference(const orc_key *b, size_t nb, const orc_key *a, size_t na, orc_key *out);
/* Hash join of relation.cu:221-373: for each outer tuple, every inner tuple
 * whose first `join_cols` columns equal the outer's.  Emits gen(inner, outer)
 * filtered by pred.  out == NULL only counts. */
size_t orc_join(const orc_key *inner, size_t ni, int inner_arity,
                const orc_key *outer, size_t no, int outer_arity, int join_cols,
                orc_gen_hook gen, orc_pred_hook pred, int out_arity, orc_key *out);
size_t orc_project(const orc_key *src, size_t n, int src_arity, orc_copy_hook cp,
                   int out_arity, orc_key *out);

/* ---- LIE semi-naive loop (src/lie.cu:26-274) -------------------------- */
typedef struct orc_rel orc_rel;
typedef struct orc_lie orc_lie;

orc_rel *orc_rel_new(const char *name, int arity, int index_cols, int tmp_flag);
void orc_rel_free(orc_rel *r);
void orc_rel_set_index_flag(orc_rel *r, int flag);
/* load_relation (relation.cu:815-838): full = sort+unique(keys). */
void orc_rel_load(orc_rel *r, const orc_key *keys, size_t n);
size_t orc_rel_size(const orc_rel *r, int version);
const orc_key *orc_rel_data(const orc_rel *r, int version);

orc_lie *orc_lie_new(void);
void orc_lie_free(orc_lie *l);
void orc_lie_add_relation(orc_lie *l, orc_rel *r, int static_flag);
void orc_lie_add_tmp(orc_lie *l, orc_rel *r);
void orc_lie_add_join(orc_lie *l, orc_rel *inner, int inner_ver, orc_rel *outer,
                      int outer_ver, orc_rel *out, orc_gen_hook gen, orc_pred_hook pred);
void orc_lie_add_copy(orc_lie *l, orc_rel *src, int src_ver, orc_rel *dst, orc_copy_hook cp);
void orc_lie_add_acopy(orc_lie *l, orc_rel *src, orc_rel *dst, orc_copy_hook cp);
void orc_lie_set_max_iteration(orc_lie *l, int max_iteration);
void orc_lie_set_reload_full(orc_lie *l, int flag);
/* Runs the fixpoint; returns the iteration counter (lie.cu:208-215). */
int orc_lie_run(orc_lie *l);

/* Per-iteration trace, one record per (iteration, update relation) in the
 * order lie.cu:93-198 visits them. */
typedef struct {
    int iteration;
    int rel;             /* index in update-relation order */
    uint64_t newt;       /* newt size before set difference (0 if empty) */
    uint64_t delta;      /* |newt \ full| */
    uint64_t full;       /* |full| after the merge */
} orc_trace_rec;
size_t orc_lie_trace(const orc_lie *l, const orc_trace_rec **recs);

/* Per-operator log: one record per RA op per iteration, in execution order.
 * raw is the operator's output before deduplication -- for a join the
 * "join result size(non dedup)" the reference prints (join.cu:86-88), for a
 * Copy/ACopy the source count -- and 0 when the op returned early (empty
 * outer or inner, join.cu:38-45; empty source, copy.cu:23-26, acopy.cu:18-22).
 * ran is 0 for a FULL-source Copy skipped after its first run (lie.cu:69-77).
 * out_newt is the output relation's newt count after the op. */
typedef struct {
    int iteration;
    int op;              /* index in add_ra order */
    int ran;
    uint64_t raw;
    uint64_t out_newt;
} orc_op_rec;
size_t orc_lie_oplog(const orc_lie *l, const orc_op_rec **recs);

/* ---- program encodings (test/tc.cu, test/sg.cu, test/cspa.cu) ----------- */
typedef struct {
    int iterations;          /* iteration counter of the (last) LIE */
    int init_iterations;     /* SG/CSPA: iterations of the init LIE */
    uint64_t derived;        /* sum of |delta| over update relations and iterations */
    size_t n_out;            /* size of the main output relation */
    orc_key *out;            /* sorted unique keys of the main output (malloc'd) */
    size_t n_trace;
    orc_trace_rec *trace;    /* trace of the (last) LIE (malloc'd) */
    size_t n_trace_init;
    orc_trace_rec *trace_init; /* trace of the init LIE (malloc'd) */
    /* CSPA only: sizes of the other update relations (VF, MA, VA) */
    uint64_t sizes[4];
    orc_key *extra[3];       /* CSPA: VF, MA sorted (malloc'd), VA == out */
    size_t n_extra[3];
    size_t n_oplog;
    orc_op_rec *oplog;       /* operator log of the (last) LIE (malloc'd) */
} orc_result;

/* edges: n pairs (src,dst) as u64 keys (src << 32 | dst). */
int orc_tc(const orc_key *edges, size_t n, int max_iteration, orc_result *res);
int orc_sg(const orc_key *edges, size_t n, orc_result *res);
int orc_cspa(const orc_key *assign, size_t na, const orc_key *deref, size_t nd, orc_result *res);
void orc_result_free(orc_result *res);

/* Multi-core CPU baseline of the same TC loop (cpu_tc_omp.c): the fixpoint
 * split by the path's second column over `threads` OpenMP threads.  Returns
 * |path|, the derived count, the iteration count and the per-body |delta|
 * (trace_cap entries).  0 on success, -1 when memory ran out.  Checked
 * against orc_tc in tests/test_oracle.py. */
int cpu_tc_omp(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *derived, int *iterations, uint64_t *delta_trace,
               int trace_cap);

/* Multi-core CPU baselines restating Souffle's generated plans
 * (cpu_souffle.c): SG (test/souffle/sg.cpp:474-645) and TC with tc.dl's
 * non-linear rule (test/souffle/tc.cpp:379-489), split over `threads`
 * OpenMP threads.  Return |result|, its checksum (sum of fmix64 of the
 * keys), the number of recursive loop bodies run (the last one derives
 * nothing; at most max_iteration when >= 0) and the per-body |new| (SG: entry
 * 0 is the non-recursive rule's |sg|, entry i body i; TC: entry i-1 body i).
 * 0 on success, -1 when memory ran out.  Checked against orc_sg / orc_tc in
 * tests/test_oracle.py. */
int souffle_sg(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *checksum, int *iterations, uint64_t *delta_trace,
               int trace_cap);
int souffle_tc(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *checksum, int *iterations, uint64_t *delta_trace,
               int trace_cap);
/* CSPA (test/souffle/cspa.cpp:344-815, cspa.dl's ten rules): sizes and
 * checksums of ValueFlow, MemoryAlias, ValueAlias (in that order) and the
 * recursive loop bodies run.  Checked against orc_cspa. */
int souffle_cspa(const uint64_t *assign, uint64_t na, const uint64_t *deref, uint64_t nd,
                 int threads, int max_iteration, uint64_t *sizes, uint64_t *checksums,
                 int *iterations);

#ifdef __cplusplus
}
#endif
#endif
