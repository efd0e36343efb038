/*
 * gdl_datalog.h -- C ABI of the Datalog front end (gdlog/datalog.h): compile
 * a program over binary relations (a subset of Souffle's syntax), run it on a
 * context, read its relations back.  No counterpart in the reference, whose
 * programs are RA-operator lists written by hand (test/tc.cu, test/sg.cu,
 * test/cspa.cu encode test/souffle/{tc,sg,cspa}.dl); this is the entry point
 * a binding would use to run such a .dl program without writing C++.
 *
 * Tuples are packed keys (c0 << 32) | c1 in host memory; a tuple of a unary
 * relation is c0 << 32.
 */
#ifndef GDL_DATALOG_H
#define GDL_DATALOG_H

#include <stdint.h>

#include "gdl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdl_datalog gdl_datalog;

/* Parses and plans `source`.  GDL_ERR_INVALID_ARG: a syntax error or an
 * undeclared relation; GDL_ERR_UNSUPPORTED: a construct outside the subset
 * (constants in rules, negation, arity above 2, ...).  The message, with
 * its line, is in gdl_datalog_last_error(). */
gdl_status gdl_datalog_compile(const char *source, gdl_datalog **out);
/* Same for a Slog compiler's RA plan (the s-expression format of the
 * reference's test/souffle/cspa.slogc): its joins and copies are read back
 * into rules (index-maintenance acopies dropped, DELTA / FULL variants
 * merged), underived relations become inputs, derived ones not marked tmp are
 * kept, and Slog names become identifiers (value-flow -> value_flow). */
gdl_status gdl_datalog_compile_slog(const char *plan, gdl_datalog **out);
void gdl_datalog_destroy(gdl_datalog *prog);
/* The Datalog text the program was compiled from (for a Slog plan, the rules
 * read back); valid until the program is destroyed. */
const char *gdl_datalog_source(const gdl_datalog *prog);
/* The lowered plan (LIEs, their relations and RA operators), one per line;
 * valid until the program is destroyed. */
const char *gdl_datalog_plan(const gdl_datalog *prog);
/* Runs the program on ctx.  names[i], keys[i], counts[i]: the tuples of each
 * .input relation (any order, duplicates allowed); every .input relation must
 * be given.  Results of an earlier run are replaced. */
gdl_status gdl_datalog_run(gdl_ctx *ctx, gdl_datalog *prog, uint32_t n_inputs,
                           const char *const *names, const uint64_t *const *keys,
                           const uint64_t *counts);
/* Sorted, duplicate-free tuples of a kept relation (.output / .printsize;
 * every derived relation when the program names none) after a run.  *n gets
 * the count; keys = NULL only queries it; GDL_ERR_OVERFLOW when cap < *n. */
gdl_status gdl_datalog_result(const gdl_datalog *prog, const char *relation, uint64_t *keys,
                              uint64_t cap, uint64_t *n);
/* Name of the i-th kept relation, in declaration order; NULL past the last.
 * Valid until the program is destroyed. */
const char *gdl_datalog_kept(const gdl_datalog *prog, uint32_t i);
/* Declared arity of a relation: 1 or 2, 0 if the program has no such relation. */
int gdl_datalog_arity(const gdl_datalog *prog, const char *relation);
/* Statistics of the last run: new tuples over all LIEs and iterations, and
 * the iteration count of each LIE in plan order (*n_lies of them). */
gdl_status gdl_datalog_stats(const gdl_datalog *prog, uint64_t *derived, int32_t *iterations,
                             uint32_t cap, uint32_t *n_lies);
/* Message of this thread's last failed compile, run or result call. */
const char *gdl_datalog_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GDL_DATALOG_H */
