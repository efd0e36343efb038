/*
 * gdl_serve.h -- a query server for TC, SG and CSPA programs on one device.
 *
 * Queries are submitted with their input relations (host arrays, copied) and
 * admitted when their device-memory footprint estimate fits the budget next
 * to the queries already running:
 *   - TC queries join one continuous TC batch (include/gdl_batch.h) at its
 *     next body boundary, as long as it has a free slot;
 *   - SG and CSPA queries (and TC queries the batch cannot hold) run the
 *     program (include/gdl_programs.h) on one of `workers` worker threads,
 *     each with its own engine context and stream, so their kernels overlap
 *     on the device.
 * Relations stay device-resident for the whole query.  Admission counts the
 * estimates; a worker's context is capped at 4x its query's estimate (at
 * least an even share of the budget) so a runaway query cannot starve the
 * others, and a query that runs out of memory is rerun alone (no cap,
 * nothing else admitted beside it): an estimate that was too small costs
 * time, not correctness.  Each result is the program's
 * result alone: the same tuples, per-iteration |delta| and |full|, and
 * iteration counts.
 *
 * The reference has no server (one fixpoint per process: test/tc.cu,
 * test/sg.cu, test/cspa.cu); nothing here replaces a reference entry point.
 */
#ifndef GDL_SERVE_H
#define GDL_SERVE_H

#include "gdl_hip.h"
#include "gdl_programs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdl_server gdl_server;

enum { GDL_QUERY_TC = 0, GDL_QUERY_SG = 1, GDL_QUERY_CSPA = 2 };

typedef struct gdl_server_config {
    int device;
    int workers;          /* program worker threads (default 1: two or more heavy
                             queries sharing the device slowed each other down
                             several-fold on MI355X, see the design notes) */
    int batch_tc;         /* 1: TC queries join one continuous TC batch (default 1) */
    uint32_t batch_slots; /* instances of that batch (default 64) */
    uint32_t node_bits;   /* batch: ids per instance < 2^node_bits (default 18) */
    uint64_t mem_budget;  /* device bytes the admitted estimates may take
                             (0: 80% of the device memory free at create) */
    int keep_keys;        /* 1: results keep the sorted output tuples */
    int partition_cus;    /* must be 0: per-executor CU masks (gdl_set_cu_mask) are
                             not supported, and gdl_server_create returns
                             GDL_ERR_UNSUPPORTED for 1 */
    int take_turns;       /* 1: the TC batch and the workers take turns on the device
                             (one at a time, in the order they asked; the batch's
                             turn is a slice of about 20 ms of bodies), so no
                             executor's kernels share the device with another's;
                             0: they run concurrently (default 0: on the TC / SG /
                             CSPA mix concurrency gave the same p99 and a 4x lower
                             p50) */
} gdl_server_config;

typedef struct gdl_query_result {
    uint64_t qid;
    int32_t kind;
    int32_t status;          /* gdl_status of the run */
    uint64_t n_out;          /* |path|, |sg| or |ValueAlias| */
    uint64_t checksum;       /* sum of fmix64 over the output tuples, mod 2^64 */
    uint64_t derived;        /* sum of |delta| over iterations */
    int32_t iterations;      /* (last) fixpoint's iteration count */
    int32_t init_iterations; /* SG / CSPA init fixpoint */
    uint64_t est_bytes;      /* admission estimate */
    double submit_ms, admit_ms, done_ms; /* server clock */
    int32_t batched;         /* 1: ran in the TC batch */
    int32_t rerun_alone;     /* 1: outgrew its reservation and was rerun alone */
    uint64_t n_trace;        /* per-iteration records of the (last) fixpoint */
    uint64_t n_keys;         /* output tuples kept (keep_keys) */
} gdl_query_result;

void gdl_server_config_default(gdl_server_config *cfg);
gdl_status gdl_server_create(const gdl_server_config *cfg, gdl_server **out);
/* Waits for every submitted query, then stops the threads. */
gdl_status gdl_server_destroy(gdl_server *s);

/* Device bytes a query is expected to hold (the admission estimate). */
uint64_t gdl_query_estimate(int kind, const uint64_t *a, uint64_t na, const uint64_t *b,
                            uint64_t nb);

/* Queues a query (inputs copied; returns at once).  TC and SG: a = edges
 * (packed (src << 32) | dst keys); CSPA: a = assign, b = dereference. */
gdl_status gdl_server_submit(gdl_server *s, uint64_t qid, int kind, const uint64_t *a,
                             uint64_t na, const uint64_t *b, uint64_t nb);

/* Pops the oldest finished query into *out: 1, or 0 when none is ready.
 * gdl_server_wait blocks up to timeout_ms (< 0: no limit). */
int gdl_server_poll(gdl_server *s, gdl_query_result *out);
int gdl_server_wait(gdl_server *s, gdl_query_result *out, double timeout_ms);

/* Copies a popped query's kept tuples (n_keys) and trace (n_trace records)
 * to the host and drops them.  Either pointer may be NULL. */
gdl_status gdl_server_fetch(gdl_server *s, uint64_t qid, uint64_t *keys, gdl_trace_rec *trace);

/* Counters: submitted, finished, queries run in the batch, reruns alone,
 * bytes reserved now, peak reserved, the budget. */
gdl_status gdl_server_stats(gdl_server *s, uint64_t *submitted, uint64_t *finished,
                            uint64_t *batched, uint64_t *reruns, uint64_t *reserved,
                            uint64_t *peak_reserved, uint64_t *budget);

#ifdef __cplusplus
}
#endif

#endif /* GDL_SERVE_H */
