/*
 * gdl_hip.h -- C ABI of the MI355X (gfx950) semi-naive Datalog engine.
 *
 * Plain pointers and sizes only.  Every call returns a gdl_status; nothing
 * calls exit().  A context owns one HIP device, one HIP stream (default
 * priority), a caching device allocator with an optional byte cap, and the
 * tuning parameters.  All work of a context is ordered on its stream; calls
 * that return a size to the host synchronise that stream.  One context per
 * host thread; a context is not re-entrant.
 *
 * Tuple encoding.  A tuple of arity 1 or 2 over u32 ids is one u64 key:
 * (c0 << 32) | c1, arity 1 as c0 << 32.  Integer order of keys is the
 * reference's lexicographic tuple order (include/tuple.cuh:94-127) and key
 * equality its full-tuple equality (tuple.cuh:51-65).  The join prefix of
 * `prefix_cols` columns is key >> 32 (1 column) or the whole key (2).
 *
 * The reference exposes a C++ API only (it has no FFI); each entry point
 * below names the reference function or Thrust call site it replaces.
 */
#ifndef GDL_HIP_H
#define GDL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Default load factor of the hash index and hash sets.  Lower than the
 * reference's 0.8 (relation.cu:603-604): an insert into an open-addressing
 * set pays one dependent atomic round trip per occupied slot it passes, and
 * a fixpoint body is bound by that chain's latency, so memory is traded for
 * shorter chains. */
#define GDL_DEFAULT_LOAD_FACTOR 0.35f

typedef enum gdl_status {
    GDL_OK = 0,
    GDL_ERR_INVALID_ARG = 1,   /* null pointer, bad arity, bad column map ... */
    GDL_ERR_OUT_OF_MEMORY = 2, /* device allocation failed or the byte cap was hit */
    GDL_ERR_HIP = 3,           /* a HIP runtime call failed (see gdl_last_error) */
    GDL_ERR_UNSUPPORTED = 4,   /* valid request outside what this engine implements */
    GDL_ERR_OVERFLOW = 5,      /* a size exceeds a documented limit */
    GDL_ERR_DEVICE = 6,        /* a kernel reported an internal error */
    GDL_ERR_BUSY = 7           /* admission refused: no free batch slot or the memory
                                  budget is taken; retry after running work retires */
} gdl_status;

typedef struct gdl_ctx gdl_ctx;

typedef struct gdl_config {
    int device;             /* HIP device ordinal */
    uint64_t mem_cap_bytes; /* 0: no cap.  Otherwise allocations that would take the
                               bytes in use above the cap fail (GDL_ERR_OUT_OF_MEMORY) */
    float load_factor;      /* hash index / hash set load factor, (0, 0.95];
                               default GDL_DEFAULT_LOAD_FACTOR */
    int collect_stats;      /* 1: time every kernel with hipEvents on the context stream */
    int verbose;            /* 1: print the reference-style per-op / per-iteration log */
} gdl_config;

/* Fills the defaults: device 0, no cap, GDL_DEFAULT_LOAD_FACTOR, no stats. */
void gdl_config_default(gdl_config *cfg);

/* Replaces the implicit device-0 / legacy-default-stream state of the
 * reference (src/print.cu:85-124, every <<<grid,block>>> launch). */
gdl_status gdl_ctx_create(const gdl_config *cfg, gdl_ctx **out);
gdl_status gdl_ctx_destroy(gdl_ctx *ctx);
const char *gdl_status_string(gdl_status s);
/* Detail of the last failure on this context (empty string if none). */
const char *gdl_last_error(gdl_ctx *ctx);
/* Waits for all work of the context; reports kernel-side errors. */
gdl_status gdl_sync(gdl_ctx *ctx);
/* The context's hipStream_t, as void*. */
void *gdl_stream(gdl_ctx *ctx);
int gdl_device(gdl_ctx *ctx);
gdl_status gdl_set_load_factor(gdl_ctx *ctx, float load_factor);

/* Launch and layout parameters of a context, as an autotuner picks them
 * (gdlog-multinode_amd/tune.py).  The reference hard-codes them: load factor
 * 0.8 (relation.cu:603-604), 512 threads x 32 blocks per SM (test/tc.cu:224-225)
 * and Thrust's sort. */
typedef struct gdl_tuning {
    float load_factor;      /* hash index / hash set load factor, (0, 0.95]
                               (hash sets are capped at 0.8);
                               default GDL_DEFAULT_LOAD_FACTOR */
    uint32_t blocks_per_cu; /* grid cap of the grid-stride kernels, 1..32; default 8 */
    uint32_t radix_bits;    /* LSD radix sort digit width, 4..8; default 8 */
    uint32_t join_block;    /* threads per block of the fused join: 64, 128, 256 or 512;
                               default 256 */
} gdl_tuning;
gdl_status gdl_get_tuning(gdl_ctx *ctx, gdl_tuning *out);
/* Upper bound of the bit length of every id (both columns of a key) that the
 * context's joins read and produce; 0 (the default) = unknown.  The fused
 * join (gdl_join_insert) then sizes its hash set and new-tuple buffer by the
 * keys the id domain still has room for, not only by the produced count
 * (which a join with duplicate-heavy output can exceed by orders of
 * magnitude).  A bitmap set always sizes by its own domain. */
gdl_status gdl_set_id_bits(gdl_ctx *ctx, int bits);
gdl_status gdl_set_tuning(gdl_ctx *ctx, const gdl_tuning *t);

/* ---- device memory (replaces the per-op cudaMalloc/cudaFree/cudaMemset of
 *      join.cu:49-62,109-114 and relation.cu:547-561,606-616) ---------- */
gdl_status gdl_alloc(gdl_ctx *ctx, size_t bytes, void **dptr);
gdl_status gdl_free(gdl_ctx *ctx, void *dptr);
gdl_status gdl_memcpy_h2d(gdl_ctx *ctx, void *dst, const void *src, size_t bytes);
gdl_status gdl_memcpy_d2h(gdl_ctx *ctx, void *dst, const void *src, size_t bytes);
gdl_status gdl_memcpy_d2d(gdl_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Free/total device memory (replaces get_free_memory / get_total_memory,
 * print.cu:98-124) and the allocator's own accounting. */
gdl_status gdl_mem_info(gdl_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes,
                        uint64_t *in_use, uint64_t *peak, uint64_t *cached);
/* Returns cached (freed) blocks to HIP. */
gdl_status gdl_release_cached(gdl_ctx *ctx);
/* Changes the byte cap (0 = none). */
gdl_status gdl_set_mem_cap(gdl_ctx *ctx, uint64_t cap_bytes);
/* CU masks (a stream restricted to a subset of the CUs) are not supported:
 * returns GDL_ERR_UNSUPPORTED for any context (GDL_ERR_INVALID_ARG for a null
 * one).  Kept in the ABI so callers probing for it get a status, not a
 * missing symbol. */
gdl_status gdl_set_cu_mask(gdl_ctx *ctx, const uint32_t *mask, uint32_t words);
/* Bytes of freed blocks the allocator keeps for reuse (0 = no limit); past
 * it the largest cached blocks go back to HIP.  Freed blocks are reused best
 * fit, up to twice the requested size. */
gdl_status gdl_set_cache_limit(gdl_ctx *ctx, uint64_t bytes);

/* ---- sort / unique ---------------------------------------------------
 * Replaces thrust::sort(tuple_indexed_less) + thrust::unique(t_equal)
 * (relation.cu:576-590).  LSD radix sort over 8-bit digits of the
 * order-preserving compressed key (c0 << lo_bits) | c1, where hi_bits and
 * lo_bits bound the bit length of c0 and c1 (0..32; pass 32/32 when unknown,
 * or -1 to measure them on the device).  d_in and d_out must not overlap.
 * At most 2^32 - 1 keys per call. */
gdl_status gdl_sort_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, int hi_bits,
                        int lo_bits, uint64_t *d_out);
/* Adjacent-duplicate removal of a sorted array (thrust::unique, t_equal). */
gdl_status gdl_unique_u64(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n,
                          uint64_t *d_out, uint64_t *n_out);
/* Bit lengths of the largest c0 and c1 over n keys (0 for an empty array). */
gdl_status gdl_key_bits(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, int *hi_bits,
                        int *lo_bits);
/* sort + unique; d_out holds n slots. */
gdl_status gdl_sort_unique_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, int hi_bits,
                               int lo_bits, uint64_t *d_out, uint64_t *n_out);

/* gdl_sort_unique_u64 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP-graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is one fixed device array per GPU (18 MiB, held until the process
 * ends), so n is limited to GDL_NOALLOC_MAX_KEYS (larger n: INVALID_ARG, and
 * nothing runs).  The caller still passes the input and output buffers.
 * Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (other threads, or other contexts' streams) race, and the caller must
 * serialise them.  Kernel timing events are not recorded for the call. */
#define GDL_NOALLOC_MAX_KEYS (1u << 20)
gdl_status gdl_sort_unique_u64_noalloc(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n,
                                       int hi_bits, int lo_bits, uint64_t *d_out,
                                       uint64_t *n_out);

/* ---- merge / set difference over sorted duplicate-free arrays ----------
 * gdl_merge_union_u64: thrust::merge + thrust::unique of newt accumulation
 *   (join.cu:181-195, copy.cu:63-72); d_out holds na + nb slots.
 * gdl_set_difference_u64: thrust::set_difference(newt, full) (lie.cu:131-138),
 *   d_out = b \ a, holds nb slots.
 * gdl_merge_diff_u64: set difference and Relation::flush_delta's
 *   thrust::merge(full, delta) (relation.cu:484-487) fused in one merge path:
 *   d_delta = newt \ full (nn slots), d_full_out = full U newt (nf + nn slots;
 *   may be NULL to skip the union). */
gdl_status gdl_merge_union_u64(gdl_ctx *ctx, const uint64_t *d_a, uint64_t na,
                               const uint64_t *d_b, uint64_t nb, uint64_t *d_out,
                               uint64_t *n_out);
gdl_status gdl_set_difference_u64(gdl_ctx *ctx, const uint64_t *d_b, uint64_t nb,
                                  const uint64_t *d_a, uint64_t na, uint64_t *d_out,
                                  uint64_t *n_out);
gdl_status gdl_merge_diff_u64(gdl_ctx *ctx, const uint64_t *d_full, uint64_t nf,
                              const uint64_t *d_newt, uint64_t nn, uint64_t *d_full_out,
                              uint64_t *n_full_out, uint64_t *d_delta, uint64_t *n_delta);

/* Stable compaction: d_out gets, in order, the d_in[i] with d_flags[i] != 0. */
gdl_status gdl_compact_flags_u64(gdl_ctx *ctx, const uint64_t *d_in, const uint8_t *d_flags,
                                 uint64_t n, uint64_t *d_out, uint64_t *n_out);
/* thrust::set_difference(newt, full, tuple_indexed_less(k, arity - dependent))
 * of the fixpoint's set difference when a relation has dependent columns
 * (lie.cu:131-138): d_b and d_a sorted by key; tuples compare by the prefix
 * key >> shift (the non-dependent columns).  Of the m elements of b sharing a
 * prefix that n elements of a have, the first min(m, n) are dropped, the
 * rest kept in order.  d_out holds nb slots. */
gdl_status gdl_set_difference_prefix_u64(gdl_ctx *ctx, const uint64_t *d_b, uint64_t nb,
                                         const uint64_t *d_a, uint64_t na, int shift,
                                         uint64_t *d_out, uint64_t *n_out);

/* ---- hash index over the distinct join prefixes of a sorted array ------
 * Replaces init_index_map + calculate_index_hash (relation.cu:12-80,124-137):
 * open addressing with linear probing; each distinct prefix maps to the
 * [begin, begin + count) run of the sorted array.  Slots are 16 bytes. */
typedef struct gdl_index {
    void *slots;        /* device; owned by the context allocator */
    uint64_t n_slots;
    uint64_t n_distinct; /* distinct prefixes */
    int prefix_cols;    /* 1 or 2 */
    uint64_t max_run;   /* longest run (tuples sharing one prefix) */
    int dense;          /* 1: slots is a direct-mapped array of n_slots run words
                           indexed by the prefix itself (prefix_cols = 1 and the
                           prefixes below n_slots, dense enough that the array is
                           no larger than the hash index); 0: the hash index */
} gdl_index;
gdl_status gdl_index_build(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n, int prefix_cols,
                           gdl_index *out);
gdl_status gdl_index_free(gdl_ctx *ctx, gdl_index *idx);
/* Looks up n prefixes (test/diagnostic entry): d_begin/d_count get the run
 * of each (count 0 when absent). */
gdl_status gdl_index_lookup(gdl_ctx *ctx, const gdl_index *idx, const uint64_t *d_sorted,
                            uint64_t n_sorted, const uint64_t *d_prefixes, uint64_t n,
                            uint64_t *d_begin, uint32_t *d_count);

/* ---- device hash set of keys (membership of an append-only FULL) --------
 * Replaces the set-difference role of thrust::set_difference(newt, full)
 * (lie.cu:131-138) for relations whose FULL is never a join inner, where
 * keeping FULL sorted (Relation::flush_delta, relation.cu:484-487) is not
 * needed until the fixpoint ends.  Load kept at or below the context's
 * load factor (at most 0.8), grown by rehashing. */
typedef struct gdl_hashset {
    void *slots;       /* device; owned by the context allocator */
    uint64_t capacity; /* slots, a power of two (bitmap: bits) */
    uint64_t size;     /* members */
    uint32_t key_bits; /* 0: hash table; k > 0: bitmap over keys whose two
                          columns are both below 2^k (bit (c0 << k) | c1) */
    uint32_t pad_;
} gdl_hashset;
gdl_status gdl_hashset_create(gdl_ctx *ctx, uint64_t expected, gdl_hashset *out);
/* A bitmap set for a dense id domain: every key (c0 << 32) | c1 with c0, c1 <
 * 2^key_bits (1..18; 2^(2 key_bits) bits, at most 8 GiB) has its own bit, so
 * an insert is one atomic OR on one word and the set never grows.  Same
 * operations as the hash table.  A key outside the domain is not inserted; it
 * sets a flag that gdl_hashset_check reports. */
gdl_status gdl_hashset_create_bitmap(gdl_ctx *ctx, uint32_t key_bits, gdl_hashset *out);
/* GDL_ERR_OVERFLOW when a key outside a bitmap set's domain was offered
 * (always GDL_OK for a hash table). */
gdl_status gdl_hashset_check(gdl_ctx *ctx, const gdl_hashset *hs);
gdl_status gdl_hashset_free(gdl_ctx *ctx, gdl_hashset *hs);
/* Inserts n keys (any order, repeats allowed); every key not yet a member is
 * written once to d_new (n slots), in no particular order; *n_new = count. */
gdl_status gdl_hashset_insert_new(gdl_ctx *ctx, gdl_hashset *hs, const uint64_t *d_keys,
                                  uint64_t n, uint64_t *d_new, uint64_t *n_new);
/* Grows the table (rehash) so that size + incoming members fit. */
gdl_status gdl_hashset_reserve(gdl_ctx *ctx, gdl_hashset *hs, uint64_t incoming);
/* Writes the members, in no particular order, to d_out (out_cap keys);
 * GDL_ERR_OVERFLOW (nothing written past out_cap) when they do not fit.
 * Replaces flatten_tuples_raw_data of a FULL version (relation.cu:375-388). */
gdl_status gdl_hashset_extract(gdl_ctx *ctx, const gdl_hashset *hs, uint64_t *d_out,
                               uint64_t out_cap, uint64_t *n_out);
/* The members in key order (same contract otherwise).  A bitmap set is read
 * in bit order, which is key order (two streaming passes: no sort); a hash
 * table's members are extracted and radix-sorted.  Gives the sorted FULL
 * that Relation::flush_delta keeps (relation.cu:484-487) when the fixpoint
 * ends. */
gdl_status gdl_hashset_extract_sorted(gdl_ctx *ctx, const gdl_hashset *hs, uint64_t *d_out,
                                      uint64_t out_cap, uint64_t *n_out);

/* ---- column maps and predicates (the build's form of the reference hooks,
 *      tuple_generator_hook / tuple_copy_hook / tuple_predicate,
 *      include/tuple.cuh:23-25) ------------------------------------------ */
enum {
    GDL_COL_INNER0 = 0, /* join: inner column 0; copy: source column 0 */
    GDL_COL_INNER1 = 1, /* join: inner column 1; copy: source column 1 */
    GDL_COL_OUTER0 = 2,
    GDL_COL_OUTER1 = 3
};
typedef struct gdl_colmap {
    int32_t arity;  /* output arity, 1 or 2 */
    int32_t src[2]; /* GDL_COL_* of output columns 0 and 1 */
} gdl_colmap;

enum { GDL_PRED_NONE = 0, GDL_PRED_EQ, GDL_PRED_NE, GDL_PRED_LT, GDL_PRED_LE, GDL_PRED_GT,
       GDL_PRED_GE, GDL_PRED_FALSE };
/* Flag on a comparison op: rhs is a constant id (its 32 bits), not a column. */
#define GDL_PRED_CONST 0x100
/* Keeps output tuple t iff t[lhs] <op> t[rhs] (columns of the generated tuple;
 * with op | GDL_PRED_CONST, iff t[lhs] <op> (uint32_t)rhs). */
typedef struct gdl_pred {
    int32_t op;
    int32_t lhs;
    int32_t rhs;
} gdl_pred;

/* ---- join (count, scan, materialize) ----------------------------------
 * Replaces get_join_result_size, the chunked thrust::reduce and
 * thrust::exclusive_scan, and get_join_result (relation.cu:221-373,
 * join.cu:68-118).  For each outer key, every inner key of the same
 * `prefix_cols`-column prefix (via the inner index) produces colmap(inner,
 * outer); the predicate is applied to the produced tuple.
 *   gdl_join_count: d_counts[i] = run length, d_begins[i] = run start.
 *   gdl_join_scan: d_offsets (n + 1 slots) = exclusive scan, *total to host.
 *   gdl_join_materialize: writes `total` unfiltered outputs (load-balanced). */
gdl_status gdl_join_count(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                          const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                          uint32_t *d_counts, uint64_t *d_begins);
gdl_status gdl_join_scan(gdl_ctx *ctx, const uint32_t *d_counts, uint64_t n,
                         uint64_t *d_offsets, uint64_t *total);
gdl_status gdl_join_materialize(gdl_ctx *ctx, const uint64_t *d_inner, const uint64_t *d_outer,
                                uint64_t n_outer, const uint64_t *d_offsets,
                                const uint64_t *d_begins, uint64_t total, gdl_colmap map,
                                uint64_t *d_out);
/* Whole join: allocates *d_out from the context allocator (free with
 * gdl_free), applies the predicate, returns the row count in *n_out. */
gdl_status gdl_join(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                    const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                    gdl_colmap map, gdl_pred pred, uint64_t **d_out, uint64_t *n_out);

/* Join fused with the set difference against an append-only FULL (the hash
 * set): each produced tuple colmap(inner, outer) that passes pred is inserted
 * into hs, and those that were not yet members are appended to the growable
 * device buffer *d_new (capacity *new_cap keys, reallocated from the context
 * allocator when needed, existing *n_new keys kept).  *n_raw gets the number
 * of produced tuples (the reference's "join result size (non dedup)").
 * Equivalent to RelationalJoin's newt accumulation followed by the
 * iteration's set difference when no other operator reads the output
 * relation's FULL or NEWT in the same iteration. */
gdl_status gdl_join_insert(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                           const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                           gdl_colmap map, gdl_pred pred, gdl_hashset *hs, uint64_t **d_new,
                           uint64_t *new_cap, uint64_t *n_new, uint64_t *n_raw);

/* The loop of a linear recursive rule, out := out U map(inner x delta(out)),
 * with out's FULL an append log *d_log (capacity *log_cap keys) plus the hash
 * set hs, and the delta the log's last segment [*delta_begin, *delta_end).
 * Runs up to max_bodies bodies -- each one gdl_join_insert whose new tuples
 * are appended to the log and become the next delta -- enqueued back to back
 * on the stream: the host waits once per 128 bodies (or when the set or the
 * log must grow), not once per body.  Stops after the first body with no new
 * tuple (the fixpoint; that body is counted).  h_new / h_raw (max_bodies
 * entries) get each body's new and produced tuple counts, *bodies the bodies
 * run; the log and the set are grown as needed.  Equivalent to max_bodies
 * iterations of LIE::fixpoint_loop (lie.cu:26-274) over a one-join LIE whose
 * output is the outer relation. */
gdl_status gdl_fixpoint_linear(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                               const gdl_index *inner_idx, gdl_colmap map, gdl_pred pred,
                               gdl_hashset *hs, uint64_t **d_log, uint64_t *log_cap,
                               uint64_t *delta_begin, uint64_t *delta_end, int max_bodies,
                               uint64_t *h_new, uint64_t *h_raw, int *bodies);

/* Bodies of a recursive rule run on bit matrices.  When FULL is a bitmap set
 * (key_bits 6..17) the delta can be held as bit rows too (row c0, bit c1),
 * and a body builds each output row from whole delta rows: coalesced row
 * streams instead of a random probe per produced tuple.  Two shapes, the
 * column maps of the reference's TC and SG encodings:
 *   GDL_DENSE_TC  out(x, z) :- r1(y, x), delta(y, z)         (test/tc.cu:80-149)
 *   GDL_DENSE_SG  out(x, y) :- r1(a, x), delta(a, b), r2(b, y) (test/sg.cu:95-147;
 *                 key_bits <= 16), the two joins through a deduplicated tmp.
 * r1 and r2 are the static relations' keys ((c0 << 32) | c1, any order).
 * gdl_fixpoint_linear uses the TC shape by itself when its map is TC's. */
enum { GDL_DENSE_TC = 1, GDL_DENSE_SG = 2 };
/* flags: GDL_DENSE_LOG_LAST appends only the last body's new tuples to the
 * log (the next delta); the other bodies' tuples are in full only, for a
 * caller that reads FULL from the set afterwards. */
enum { GDL_DENSE_LOG_LAST = 1 };
typedef struct gdl_dense_rule {
    int shape;
    const uint64_t *r1;
    uint64_t n1;
    const uint64_t *r2; /* GDL_DENSE_SG only */
    uint64_t n2;
    uint32_t flags;
    uint32_t pad_;
} gdl_dense_rule;
/* Runs bodies of `rule` from the delta d_delta[0, n_delta) (keys that are
 * members of full, duplicates allowed): each body's new tuples (the set
 * difference against full) are added to full and appended to the growable
 * log *d_log (capacity *log_cap keys, *log_end in use; reallocated from the
 * context allocator when needed), and become the next delta.  h_new / h_raw
 * (max_bodies entries) get each body's new and produced ("join result size
 * (non dedup)") tuple counts; *bodies the bodies run.  The context's dense
 * mode (gdl_set_dense_mode) decides whether bodies run at all: in mode 1
 * from a delta dense enough for row streams to beat tuple probes, and for
 * GDL_DENSE_TC only while it stays so (a body with fewer new tuples than that
 * ends the run; it is counted), GDL_DENSE_SG until the fixpoint or
 * max_bodies; in mode 2 until the fixpoint or max_bodies.  *bodies = 0 when the rule does
 * not apply: mode 0, a hash-table FULL, key_bits outside the range above, a
 * static relation with ids outside the domain, or (mode 1) a sparse delta.
 * Equivalent to *bodies iterations of LIE::fixpoint_loop (lie.cu:26-274)
 * over the rule's RA operators. */
gdl_status gdl_fixpoint_dense(gdl_ctx *ctx, const gdl_dense_rule *rule, gdl_hashset *full,
                              const uint64_t *d_delta, uint64_t n_delta, uint64_t **d_log,
                              uint64_t *log_cap, uint64_t *log_end, int max_bodies,
                              uint64_t *h_new, uint64_t *h_raw, int *bodies);
/* 0: never run dense bodies; 1 (the default, or the GDLOG_DENSE environment
 * variable at context creation): when a cost model prefers them; 2: whenever
 * they apply.  An SG body builds U = delta x r2 either bit by bit (a
 * successor-list probe per delta bit) or by columns (U^T = r2^T x delta^T:
 * two bit-matrix transposes around a row OR over r2's pred lists, a fixed
 * cost for domains of 2^9 ids or more); a cost model picks per body, or mode
 * 3 (dense whenever they apply, always by columns) and mode 4 (likewise,
 * never by columns) force one.  Results are the same in every mode. */
gdl_status gdl_set_dense_mode(gdl_ctx *ctx, int mode);

/* Batched form of gdl_join_insert: many fixpoint instances share the inner
 * and outer arrays, told apart by the top bits of their keys, and each has
 * its own hash set: a produced key k goes into d_segs[k >> seg_shift]
 * (a device array of n_segs <= 256 entries; slots == NULL: no set, the key
 * is dropped).  The caller sizes every set for its worst case and gives a
 * fixed output of out_cap keys; GDL_ERR_OVERFLOW reports an output bound
 * that was too small (nothing is written past it).  Set sizes are not
 * updated: count the new keys per instance with gdl_key_histogram_u64.
 * No counterpart in the reference, which runs one fixpoint per process. */
typedef struct gdl_hs_seg {
    void *slots;   /* a gdl_hashset's slots, or NULL */
    uint64_t mask; /* its capacity - 1 */
} gdl_hs_seg;
gdl_status gdl_join_insert_seg(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                               const gdl_index *inner_idx, const uint64_t *d_outer,
                               uint64_t n_outer, gdl_colmap map, gdl_pred pred,
                               const gdl_hs_seg *d_segs, uint32_t n_segs, int seg_shift,
                               uint64_t *d_out, uint64_t out_cap, uint64_t *n_out,
                               uint64_t *n_raw);

/* d_counts[k >> shift] += 1 for every key k with k >> shift < n_bins,
 * n_bins <= 2048 (device u64 counters, accumulated: clear them first). */
gdl_status gdl_key_histogram_u64(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, int shift,
                                 uint32_t n_bins, uint64_t *d_counts);
/* Order-independent checksum of a tuple set: the sum, mod 2^64, of the
 * MurmurHash3 fmix64 finalizer of every key.  Result-comparison aid (a
 * relation's identity without copying it to the host). */
gdl_status gdl_checksum_u64(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint64_t *out);
/* out = ((c0 & col_mask) << 32) | (c1 & col_mask): strips an instance
 * prefix from both columns (in == out allowed). */
gdl_status gdl_mask_columns_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n,
                                uint32_t col_mask, uint64_t *d_out);

/* ---- graph capture -----------------------------------------------------
 * Records the context's stream into a HIP graph and replays it (the build's
 * replacement for launching a fixed sequence of calls one by one; no
 * counterpart in the reference, whose calls synchronize after every kernel).
 * Between gdl_graph_begin and gdl_graph_end, only entry points that launch
 * kernels on caller-owned buffers may be called: gdl_project_u64,
 * gdl_mask_columns_u64, gdl_segment_labels_u32 (and async copies between
 * device buffers).  Any call that synchronizes, reads a size back or
 * allocates fails during the capture, and gdl_graph_end then reports the
 * error.  Kernel statistics must be off.  A graph replays on the context it
 * was captured on, ordered on its stream; the buffers it names must stay
 * allocated. */
typedef struct gdl_graph gdl_graph;
gdl_status gdl_graph_begin(gdl_ctx *ctx);
gdl_status gdl_graph_end(gdl_ctx *ctx, gdl_graph **out);
gdl_status gdl_graph_launch(gdl_ctx *ctx, const gdl_graph *graph);
gdl_status gdl_graph_destroy(gdl_graph *graph);

/* ---- labels of an append log ------------------------------------------
 * No counterpart in the reference's code; they produce the dependent column
 * of its path_length program (test/path_length.cu:49-124: the hop count
 * newt[2] = outer[2] + 1, left out of the set difference), which equals one
 * plus the log segment (iteration) that first derived the pair.
 * gdl_segment_labels_u32: d_out[i] = first + s for bounds[s] <= i <
 * bounds[s + 1] (d_bounds: n_seg + 1 nondecreasing device words, the last n).
 * gdl_scatter_by_key_u32: for each d_keys[i], its position p in the sorted,
 * duplicate-free d_sorted gets d_out[p] = d_vals[i]; the keys must be
 * distinct and all present (GDL_ERR_DEVICE otherwise). */
gdl_status gdl_segment_labels_u32(gdl_ctx *ctx, const uint64_t *d_bounds, uint32_t n_seg,
                                  uint32_t first, uint32_t *d_out, uint64_t n);
gdl_status gdl_scatter_by_key_u32(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n_sorted,
                                  const uint64_t *d_keys, const uint32_t *d_vals, uint64_t n,
                                  uint32_t *d_out);
/* Both steps at once for a log of distinct keys whose ids have at most
 * id_bits bits: d_keys gets the log sorted, d_labels[p] the segment label of
 * d_keys[p] (first + segment), the label carried through the sort in the low
 * bits of the key.  GDL_ERR_UNSUPPORTED when id_bits plus the label's bit
 * length exceed 32 (then use the two calls above). */
gdl_status gdl_segment_labels_sorted(gdl_ctx *ctx, const uint64_t *d_log, uint64_t n,
                                     const uint64_t *d_bounds, uint32_t n_seg, uint32_t first,
                                     int id_bits, uint64_t *d_keys, uint32_t *d_labels);

/* ---- projection / filter ----------------------------------------------
 * gdl_project_u64 replaces get_copy_result (relation.cu:390-403);
 * gdl_filter_u64 keeps keys whose tuple satisfies pred. */
gdl_status gdl_project_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, gdl_colmap map,
                           uint64_t *d_out);
gdl_status gdl_filter_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, gdl_pred pred,
                          uint64_t *d_out, uint64_t *n_out);

/* ---- phase statistics (replaces KernelTimer / detail_time[], timer.cuh:5-30,
 *      join.cu:74-258) ---------------------------------------------------- */
typedef struct gdl_kernel_stat {
    const char *name;
    uint64_t launches;
    double total_ms;       /* sum of hipEvent durations (collect_stats = 1) */
    double alg_bytes;      /* algorithmic HBM bytes of all launches */
} gdl_kernel_stat;
gdl_status gdl_stats_enable(gdl_ctx *ctx, int enable);
gdl_status gdl_stats_reset(gdl_ctx *ctx);
/* Resolves pending events (synchronises) and returns the number of kernel
 * classes; fills up to `cap` entries. */
gdl_status gdl_stats_get(gdl_ctx *ctx, gdl_kernel_stat *out, int cap, int *n_out);

#ifdef __cplusplus
}
#endif
#endif
