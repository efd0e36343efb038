/*
 * gdl_batch.h -- batched transitive-closure fixpoints on one device.
 *
 * Many independent TC instances (path(a, c) :- edge(a, b), path(b, c), the
 * semi-naive program of the reference's test/tc.cu:80-149) run as one
 * disjoint union: instance slot s owns node ids (s << node_bits) | v, so its
 * tuples are told apart by the top bits of their keys.  Every body of the
 * batch is one fused join + insert over the union (gdl_join_insert_seg),
 * with one hash set per instance, then one histogram of the new tuples by
 * slot.  Instances are admitted at body boundaries while slots and the
 * memory budget allow (continuous batching) and retire, result extracted,
 * the body their delta comes out empty.
 *
 * Per instance the bodies compute exactly what the program computes alone:
 * the same initial delta (the first |full| loaded rows, relation.cu:764-813
 * with lie.cu:37-45), the same per-body |delta| and |full|, and the same
 * iteration count (bodies including the final empty one, lie.cu:208-215).
 *
 * The reference runs one fixpoint per process and has no batching; nothing
 * here replaces a reference entry point.
 */
#ifndef GDL_BATCH_H
#define GDL_BATCH_H

#include "gdl_hip.h"
#include "gdl_programs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdl_tcb gdl_tcb;

typedef struct gdl_tcb_config {
    uint32_t max_slots;  /* instances resident at once, 1..256 (default 256) */
    uint32_t node_bits;  /* an instance's node ids are < 2^node_bits (default 16);
                            node_bits + bits(max_slots - 1) <= 32 */
    uint64_t mem_budget; /* device bytes the admitted instances may reserve, by
                            their footprint estimates (0: 80% of the device memory
                            free at create) */
    int keep_keys;       /* 1: a finished instance's sorted tuples are copied to the
                            host; 0: only its size and checksum */
    uint64_t solo_tuples; /* an instance whose |TC| estimate is at least this runs
                             in the solo lane (default 2^23; ~0: never) */
    uint32_t solo_threads; /* the solo lane's threads, each with a context and
                              stream of its own (1..8, default 2) */
} gdl_tcb_config;

typedef struct gdl_tcb_result {
    uint64_t qid;
    uint64_t n_out;     /* |path| */
    uint64_t checksum;  /* sum over tuples of fmix64((a << 32) | c), mod 2^64 */
    uint64_t derived;   /* sum of |delta| over bodies */
    int32_t iterations; /* bodies, the final empty one included */
    int32_t slot;
    uint64_t n_trace;   /* per-body records (== iterations) */
    double admit_ms;    /* admission time, ms on the batch clock */
    double done_ms;     /* result on the host, ms on the batch clock */
    int has_keys;
} gdl_tcb_result;

void gdl_tcb_config_default(gdl_tcb_config *cfg);
gdl_status gdl_tcb_create(gdl_ctx *ctx, const gdl_tcb_config *cfg, gdl_tcb **out);
gdl_status gdl_tcb_destroy(gdl_tcb *b);

/* Estimated |TC| of a graph (packed (a << 32) | b edge keys, host) from
 * breadth-first searches out of `samples` sources drawn with SplitMix64(seed):
 * mean reach x n_nodes (exact when samples >= n_nodes). */
uint64_t gdl_tc_estimate(const uint64_t *h_edges, uint64_t n, uint32_t n_nodes, uint32_t samples,
                         uint64_t seed);

/* Admits an instance: n raw edge rows (host, packed keys, ids < n_nodes, any
 * order, repeats allowed).  est_tuples: the footprint estimate of its |TC|
 * (0: gdl_tc_estimate with 8 samples).  GDL_ERR_BUSY when no slot is free
 * or the estimate does not fit the budget next to the instances admitted
 * (an instance is always admitted into an empty batch).  The instance's first
 * body is the batch's next one.  An instance estimated at cfg.solo_tuples or
 * more (one that fills the device by itself) instead joins the solo lane:
 * threads of the batch with a context (and stream) of their own run such
 * instances, beside the batch's bodies and each other, on the single-query
 * engine path (the linear loop, with bit-matrix bodies while the delta is
 * dense).  Their results (slot -1) come to gdl_tcb_poll with the same tuples,
 * checksum and trace; a failure there is reported by the next gdl_tcb_step.
 * gdl_tcb_destroy waits for the lane's queued instances. */
gdl_status gdl_tcb_admit(gdl_tcb *b, uint64_t qid, const uint64_t *h_edges, uint64_t n,
                         uint32_t n_nodes, uint64_t est_tuples);

/* Runs one body over every admitted instance and retires those whose delta
 * came out empty.  *n_active: instances still running after it (the solo
 * lane's not counted; with none running here and solo instances pending, it
 * waits up to 1 ms for a solo result). */
gdl_status gdl_tcb_step(gdl_tcb *b, uint32_t *n_active);

/* Oldest finished, unfetched instance: 1 and *out filled, or 0. */
int gdl_tcb_poll(gdl_tcb *b, gdl_tcb_result *out);

/* Copies a finished instance's sorted tuples (n_out keys, when kept) and its
 * per-body trace (n_trace records) to the host, then drops the result.
 * Either pointer may be NULL. */
gdl_status gdl_tcb_fetch(gdl_tcb *b, uint64_t qid, uint64_t *h_keys, gdl_trace_rec *h_trace);

/* Counters: bodies run, instances admitted / finished, bytes reserved. */
gdl_status gdl_tcb_info(gdl_tcb *b, uint64_t *bodies, uint64_t *admitted, uint64_t *finished,
                        uint64_t *reserved_bytes, uint64_t *budget_bytes);

#ifdef __cplusplus
}
#endif

#endif /* GDL_BATCH_H */
