/*
 * gdl_oracle.c -- CPU oracle (test infrastructure only; see gdl_oracle.h).
 *
 * Plain-C restatement of harp-lab/gdlog's semi-naive engine.  Every function
 * cites the reference file:line whose behaviour it restates.  Tuple sets,
 * per-iteration delta sizes and iteration counts are the observable results;
 * hash-slot layout and pointer arrays of the reference are not observable and
 * are not restated.
 */
#include "gdl_oracle.h"

#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_COLS 20 /* reference scratch tuple width (relation.cu:269,349) */

static void *xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of host memory (%zu bytes)\n", bytes);
        abort();
    }
    return p;
}

static void *xrealloc(void *p, size_t bytes) {
    void *q = realloc(p, bytes ? bytes : 1);
    if (!q) {
        fprintf(stderr, "oracle: out of host memory (%zu bytes)\n", bytes);
        abort();
    }
    return q;
}

/* ---- tuple <-> key ---------------------------------------------------- */

static inline void key_to_cols(orc_key k, int arity, orc_col *cols) {
    cols[0] = k >> 32;
    if (arity > 1) cols[1] = k & 0xffffffffull;
}

static inline orc_key cols_to_key(const orc_col *cols, int arity) {
    assert(cols[0] <= 0xffffffffull);
    orc_key k = cols[0] << 32;
    if (arity > 1) {
        assert(cols[1] <= 0xffffffffull);
        k |= cols[1];
    }
    return k;
}

/* prefix of the first `cols` columns: tuple_eq over `cols` (tuple.cuh:41-49) */
static inline orc_key key_prefix(orc_key k, int cols) {
    return cols >= 2 ? k : (k >> 32);
}

/* ---- sort / unique / merge / set difference ----------------------------
 * thrust::sort(tuple_indexed_less) + thrust::unique(t_equal)
 * (relation.cu:576-590) restated as an LSD radix sort over 16-bit digits
 * followed by adjacent-duplicate removal. */

void orc_sort(orc_key *a, size_t n) {
    if (n < 2) return;
    if (n < 64) {
        for (size_t i = 1; i < n; i++) {
            orc_key v = a[i];
            size_t j = i;
            while (j > 0 && a[j - 1] > v) {
                a[j] = a[j - 1];
                j--;
            }
            a[j] = v;
        }
        return;
    }
    orc_key *tmp = (orc_key *)xmalloc(n * sizeof(orc_key));
    size_t *count = (size_t *)xmalloc(65536 * sizeof(size_t));
    orc_key *src = a, *dst = tmp;
    for (int shift = 0; shift < 64; shift += 16) {
        memset(count, 0, 65536 * sizeof(size_t));
        for (size_t i = 0; i < n; i++) count[(src[i] >> shift) & 0xffff]++;
        if (count[(src[0] >> shift) & 0xffff] == n) continue; /* constant digit */
        size_t sum = 0;
        for (size_t d = 0; d < 65536; d++) {
            size_t c = count[d];
            count[d] = sum;
            sum += c;
        }
        for (size_t i = 0; i < n; i++) dst[count[(src[i] >> shift) & 0xffff]++] = src[i];
        orc_key *t = src;
        src = dst;
        dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(orc_key));
    free(count);
    free(tmp);
}

size_t orc_sort_unique(orc_key *a, size_t n) {
    if (n == 0) return 0;
    orc_sort(a, n);
    size_t w = 1;
    for (size_t i = 1; i < n; i++)
        if (a[i] != a[w - 1]) a[w++] = a[i];
    return w;
}

/* thrust::merge(tuple_indexed_less) then thrust::unique (join.cu:181-195,
 * copy.cu:63-72): union of two sorted duplicate-free sequences. */
size_t orc_merge_union(const orc_key *a, size_t na, const orc_key *b, size_t nb, orc_key *out) {
    size_t i = 0, j = 0, w = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) out[w++] = a[i++];
        else if (b[j] < a[i]) out[w++] = b[j++];
        else { out[w++] = a[i++]; j++; }
    }
    while (i < na) out[w++] = a[i++];
    while (j < nb) out[w++] = b[j++];
    return w;
}

/* thrust::set_difference(newt, full) (lie.cu:131-138): elements of b not in a. */
size_t orc_set_difference(const orc_key *b, size_t nb, const orc_key *a, size_t na, orc_key *out) {
    size_t i = 0, j = 0, w = 0;
    while (j < nb) {
        while (i < na && a[i] < b[j]) i++;
        if (i < na && a[i] == b[j]) { j++; continue; }
        out[w++] = b[j++];
    }
    return w;
}

/* ---- join and projection ----------------------------------------------
 * get_join_result_size / get_join_result (relation.cu:221-373): for each
 * outer tuple, locate the inner run whose first `join_cols` columns equal the
 * outer's (the reference finds the run start through its prefix hash index;
 * the run itself is a contiguous range of the sorted inner), then emit
 * gen(inner, outer) for every tuple of the run, keeping it only if
 * pred(gen(...)) holds when a predicate is given. */

static size_t lower_bound_prefix(const orc_key *a, size_t n, orc_key p, int cols) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_prefix(a[mid], cols) < p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

size_t orc_join(const orc_key *inner, size_t ni, int inner_arity,
                const orc_key *outer, size_t no, int outer_arity, int join_cols,
                orc_gen_hook gen, orc_pred_hook pred, int out_arity, orc_key *out) {
    size_t w = 0;
    orc_col icols[ORC_MAX_COLS], ocols[ORC_MAX_COLS], res[ORC_MAX_COLS];
    for (size_t o = 0; o < no; o++) {
        orc_key p = key_prefix(outer[o], join_cols);
        size_t s = lower_bound_prefix(inner, ni, p, join_cols);
        key_to_cols(outer[o], outer_arity, ocols);
        for (size_t i = s; i < ni && key_prefix(inner[i], join_cols) == p; i++) {
            key_to_cols(inner[i], inner_arity, icols);
            memset(res, 0, sizeof(res));
            gen(icols, ocols, res);
            if (pred && !pred(res)) continue;
            if (out) out[w] = cols_to_key(res, out_arity);
            w++;
        }
    }
    return w;
}

/* get_copy_result (relation.cu:390-403): dst[i] = copy(src[i]). */
size_t orc_project(const orc_key *src, size_t n, int src_arity, orc_copy_hook cp,
                   int out_arity, orc_key *out) {
    orc_col in[ORC_MAX_COLS], res[ORC_MAX_COLS];
    for (size_t i = 0; i < n; i++) {
        key_to_cols(src[i], src_arity, in);
        memset(res, 0, sizeof(res));
        cp(in, res);
        out[i] = cols_to_key(res, out_arity);
    }
    return n;
}

/* ---- relations ---------------------------------------------------------
 * Relation {full, delta, newt} (relation.cuh:257-300).  `full_view` is the
 * full container's tuple_counts, which RelationalJoin may zero
 * (join.cu:38-45); `full_base` is Relation::current_full_size, the array
 * set_difference and flush_delta use (lie.cu:131-138, relation.cu:405-521). */

typedef struct {
    orc_key *data;
    size_t n;
    size_t cap;
} orc_vec;

struct orc_rel {
    char name[64];
    int arity;
    int index_cols;
    int tmp_flag;
    int index_flag;
    orc_vec raw;        /* full container data_raw: rows as loaded (unsorted) */
    orc_vec full;       /* data + full_base (n) */
    size_t full_view;   /* full container tuple_counts */
    orc_vec delta;
    orc_vec newt;
};

static void vec_set(orc_vec *v, const orc_key *src, size_t n) {
    if (v->cap < n) {
        v->data = (orc_key *)xrealloc(v->data, n * sizeof(orc_key));
        v->cap = n;
    }
    if (n) memcpy(v->data, src, n * sizeof(orc_key));
    v->n = n;
}

static void vec_take(orc_vec *v, orc_key *buf, size_t n, size_t cap) {
    free(v->data);
    v->data = buf;
    v->n = n;
    v->cap = cap;
}

static void vec_clear(orc_vec *v) { v->n = 0; }

orc_rel *orc_rel_new(const char *name, int arity, int index_cols, int tmp_flag) {
    orc_rel *r = (orc_rel *)calloc(1, sizeof(orc_rel));
    snprintf(r->name, sizeof(r->name), "%s", name ? name : "");
    assert(arity >= 1 && arity <= 2);
    r->arity = arity;
    r->index_cols = index_cols;
    r->tmp_flag = tmp_flag;
    r->index_flag = 1;
    return r;
}

void orc_rel_free(orc_rel *r) {
    if (!r) return;
    free(r->raw.data);
    free(r->full.data);
    free(r->delta.data);
    free(r->newt.data);
    free(r);
}

void orc_rel_set_index_flag(orc_rel *r, int flag) { r->index_flag = flag; }

void orc_rel_load(orc_rel *r, const orc_key *keys, size_t n) {
    vec_set(&r->raw, keys, n);
    vec_set(&r->full, keys, n);
    r->full.n = orc_sort_unique(r->full.data, n);
    r->full_view = r->full.n;
    vec_clear(&r->delta);
    vec_clear(&r->newt);
}

size_t orc_rel_size(const orc_rel *r, int version) {
    if (version == ORC_FULL) return r->full_view;
    if (version == ORC_DELTA) return r->delta.n;
    return r->newt.n;
}

const orc_key *orc_rel_data(const orc_rel *r, int version) {
    if (version == ORC_FULL) return r->full.data;
    if (version == ORC_DELTA) return r->delta.data;
    return r->newt.data;
}

/* ---- RA ops ------------------------------------------------------------ */

enum { OP_JOIN, OP_COPY, OP_ACOPY };

typedef struct {
    int kind;
    orc_rel *inner, *outer, *out;  /* join */
    int inner_ver, outer_ver;
    orc_gen_hook gen;
    orc_pred_hook pred;
    orc_rel *src, *dst;            /* copy / acopy */
    int src_ver;
    orc_copy_hook cp;
    int copied;                    /* RelationalCopy::copied (lie.cu:69-77) */
} orc_op;

struct orc_lie {
    orc_op *ops;
    size_t n_ops;
    orc_rel **update;
    size_t n_update;
    orc_rel **stat;
    size_t n_stat;
    orc_rel **tmp;
    size_t n_tmp;
    int max_iteration;
    int reload_full_flag; /* lie.cuh:29 */
    orc_trace_rec *trace;
    size_t n_trace, cap_trace;
    orc_op_rec *oplog;
    size_t n_oplog, cap_oplog;
};

orc_lie *orc_lie_new(void) {
    orc_lie *l = (orc_lie *)calloc(1, sizeof(orc_lie));
    l->max_iteration = 0x7fffffff; /* INT_MAX (lie.cuh:30) */
    l->reload_full_flag = 1;
    return l;
}

void orc_lie_free(orc_lie *l) {
    if (!l) return;
    free(l->ops);
    free(l->update);
    free(l->stat);
    free(l->tmp);
    free(l->trace);
    free(l->oplog);
    free(l);
}

static void push_rel(orc_rel ***arr, size_t *n, orc_rel *r) {
    *arr = (orc_rel **)xrealloc(*arr, (*n + 1) * sizeof(orc_rel *));
    (*arr)[(*n)++] = r;
}

static orc_op *push_op(orc_lie *l) {
    l->ops = (orc_op *)xrealloc(l->ops, (l->n_ops + 1) * sizeof(orc_op));
    orc_op *op = &l->ops[l->n_ops++];
    memset(op, 0, sizeof(*op));
    return op;
}

/* LIE::add_relations (lie.cu:15-22) */
void orc_lie_add_relation(orc_lie *l, orc_rel *r, int static_flag) {
    if (static_flag) push_rel(&l->stat, &l->n_stat, r);
    else push_rel(&l->update, &l->n_update, r);
}

void orc_lie_add_tmp(orc_lie *l, orc_rel *r) { push_rel(&l->tmp, &l->n_tmp, r); }

void orc_lie_add_join(orc_lie *l, orc_rel *inner, int inner_ver, orc_rel *outer,
                      int outer_ver, orc_rel *out, orc_gen_hook gen, orc_pred_hook pred) {
    orc_op *op = push_op(l);
    op->kind = OP_JOIN;
    op->inner = inner;
    op->inner_ver = inner_ver;
    op->outer = outer;
    op->outer_ver = outer_ver;
    op->out = out;
    op->gen = gen;
    op->pred = pred;
}

void orc_lie_add_copy(orc_lie *l, orc_rel *src, int src_ver, orc_rel *dst, orc_copy_hook cp) {
    orc_op *op = push_op(l);
    op->kind = OP_COPY;
    op->src = src;
    op->src_ver = src_ver;
    op->dst = dst;
    op->cp = cp;
}

void orc_lie_add_acopy(orc_lie *l, orc_rel *src, orc_rel *dst, orc_copy_hook cp) {
    orc_op *op = push_op(l);
    op->kind = OP_ACOPY;
    op->src = src;
    op->dst = dst;
    op->cp = cp;
}

void orc_lie_set_max_iteration(orc_lie *l, int max_iteration) { l->max_iteration = max_iteration; }
void orc_lie_set_reload_full(orc_lie *l, int flag) { l->reload_full_flag = flag; }

/* newt accumulation of a non-tmp output (join.cu:130-140,160-225;
 * copy.cu:42-93): newt := newt U sort+unique(fresh).  Takes ownership of
 * `fresh` (n elements, capacity cap). */
static void newt_accumulate_sorted(orc_rel *r, orc_key *fresh, size_t n, size_t cap) {
    n = orc_sort_unique(fresh, n);
    if (r->newt.n == 0) {
        vec_take(&r->newt, fresh, n, cap);
        return;
    }
    size_t tot = r->newt.n + n;
    orc_key *m = (orc_key *)xmalloc(tot * sizeof(orc_key));
    size_t w = orc_merge_union(r->newt.data, r->newt.n, fresh, n, m);
    free(fresh);
    vec_take(&r->newt, m, w, tot);
}

/* tmp output (join.cu:141-151,226-254): unsorted concatenation, no dedup. */
static void newt_append_raw(orc_rel *r, orc_key *fresh, size_t n, size_t cap) {
    if (r->newt.n == 0) {
        vec_take(&r->newt, fresh, n, cap);
        return;
    }
    size_t tot = r->newt.n + n;
    if (r->newt.cap < tot) {
        r->newt.data = (orc_key *)xrealloc(r->newt.data, tot * sizeof(orc_key));
        r->newt.cap = tot;
    }
    memcpy(r->newt.data + r->newt.n, fresh, n * sizeof(orc_key));
    r->newt.n = tot;
    free(fresh);
}

/* RelationalJoin::operator() (join.cu:12-270) */
static uint64_t run_join(orc_op *op) {
    orc_rel *in = op->inner, *ou = op->outer;
    /* version select (join.cu:15-29): inner DELTA, otherwise FULL */
    const orc_key *inner;
    size_t ni;
    if (op->inner_ver == ORC_DELTA) { inner = in->delta.data; ni = in->delta.n; }
    else { inner = in->full.data; ni = in->full_view; }
    size_t *outer_n;
    const orc_key *outer;
    if (op->outer_ver == ORC_DELTA) { outer = ou->delta.data; outer_n = &ou->delta.n; }
    else if (op->outer_ver == ORC_FULL) { outer = ou->full.data; outer_n = &ou->full_view; }
    else { outer = ou->newt.data; outer_n = &ou->newt.n; }
    /* early outs (join.cu:38-45): an empty inner zeroes the outer's count */
    if (*outer_n == 0) return 0;
    if (ni == 0) { *outer_n = 0; return 0; }
    size_t no = *outer_n;
    int k = ou->index_cols; /* join_column_counts = outer->index_column_size */
    size_t m = orc_join(inner, ni, in->arity, outer, no, ou->arity, k, op->gen, op->pred,
                        op->out->arity, NULL);
    orc_key *res = (orc_key *)xmalloc(m * sizeof(orc_key));
    orc_join(inner, ni, in->arity, outer, no, ou->arity, k, op->gen, op->pred, op->out->arity, res);
    if (op->out->tmp_flag) newt_append_raw(op->out, res, m, m);
    else newt_accumulate_sorted(op->out, res, m, m);
    return m;
}

/* RelationalCopy::operator() (copy.cu:12-95) */
static uint64_t run_copy(orc_op *op) {
    orc_rel *s = op->src;
    const orc_key *src;
    size_t n;
    if (op->src_ver == ORC_DELTA) { src = s->delta.data; n = s->delta.n; }
    else { src = s->full.data; n = s->full_view; }
    if (n == 0) { op->dst->newt.n = 0; return 0; } /* copy.cu:23-26 */
    orc_key *res = (orc_key *)xmalloc(n * sizeof(orc_key));
    orc_project(src, n, s->arity, op->cp, op->dst->arity, res);
    newt_accumulate_sorted(op->dst, res, n, n);
    return n;
}

/* RelationalACopy::operator() (acopy.cu:12-50): dest newt is overwritten. */
static uint64_t run_acopy(orc_op *op) {
    orc_rel *s = op->src;
    if (s->newt.n == 0) { op->dst->newt.n = 0; return 0; }
    size_t n = s->newt.n;
    orc_key *res = (orc_key *)xmalloc(n * sizeof(orc_key));
    orc_project(s->newt.data, n, s->arity, op->cp, op->dst->arity, res);
    size_t raw = n;
    n = orc_sort_unique(res, n);
    vec_take(&op->dst->newt, res, n, s->newt.n);
    return raw;
}

static void oplog_push(orc_lie *l, int it, int op, int ran, uint64_t raw, uint64_t newt) {
    if (l->n_oplog == l->cap_oplog) {
        l->cap_oplog = l->cap_oplog ? 2 * l->cap_oplog : 64;
        l->oplog = (orc_op_rec *)xrealloc(l->oplog, l->cap_oplog * sizeof(orc_op_rec));
    }
    orc_op_rec *o = &l->oplog[l->n_oplog++];
    o->iteration = it;
    o->op = op;
    o->ran = ran;
    o->raw = raw;
    o->out_newt = newt;
}

static void trace_push(orc_lie *l, int it, int rel, uint64_t newt, uint64_t delta, uint64_t full) {
    if (l->n_trace == l->cap_trace) {
        l->cap_trace = l->cap_trace ? 2 * l->cap_trace : 64;
        l->trace = (orc_trace_rec *)xrealloc(l->trace, l->cap_trace * sizeof(orc_trace_rec));
    }
    orc_trace_rec *t = &l->trace[l->n_trace++];
    t->iteration = it;
    t->rel = rel;
    t->newt = newt;
    t->delta = delta;
    t->full = full;
}

/* LIE::fixpoint_loop (lie.cu:26-274) */
int orc_lie_run(orc_lie *l) {
    l->n_trace = 0;
    l->n_oplog = 0;
    /* init (lie.cu:49-59): current_full_size = full count; delta is
     * copy_relation_container(delta, full) (relation.cu:764-795), which copies
     * the first tuple_counts rows of full's data_raw -- the rows as loaded,
     * unsorted -- and loads them with sorted_flag set, i.e. without sorting or
     * deduplicating.  For duplicate-free input this is the same set as full. */
    for (size_t r = 0; r < l->n_update; r++) {
        orc_rel *rel = l->update[r];
        rel->full.n = rel->full_view;
        size_t nd = rel->full_view < rel->raw.n ? rel->full_view : rel->raw.n;
        vec_set(&rel->delta, rel->raw.data, nd);
    }
    int iteration = 0;
    for (;;) {
        for (size_t i = 0; i < l->n_ops; i++) {
            orc_op *op = &l->ops[i];
            int ran = 1;
            uint64_t raw = 0;
            if (op->kind == OP_JOIN) raw = run_join(op);
            else if (op->kind == OP_ACOPY) raw = run_acopy(op);
            else if (op->src_ver == ORC_FULL) {
                if (!op->copied) { raw = run_copy(op); op->copied = 1; }
                else ran = 0;
            } else raw = run_copy(op);
            orc_rel *dst = op->kind == OP_JOIN ? op->out : op->dst;
            oplog_push(l, iteration, (int)i, ran, raw, dst->newt.n);
        }
        for (size_t t = 0; t < l->n_tmp; t++) vec_clear(&l->tmp[t]->newt); /* lie.cu:85-87 */
        int fixpoint = 1;
        for (size_t r = 0; r < l->n_update; r++) {
            orc_rel *rel = l->update[r];
            vec_clear(&rel->delta); /* lie.cu:98-111 */
            if (rel->newt.n == 0) { /* lie.cu:114-121 */
                trace_push(l, iteration, (int)r, 0, 0, rel->full.n);
                continue;
            }
            size_t nn = rel->newt.n;
            orc_key *d = (orc_key *)xmalloc(nn * sizeof(orc_key));
            size_t nd = orc_set_difference(rel->newt.data, nn, rel->full.data, rel->full.n, d);
            if (nd != 0) fixpoint = 0;
            vec_clear(&rel->newt); /* lie.cu:163 */
            vec_take(&rel->delta, d, nd, nn);
            if (nd > 0) { /* flush_delta (relation.cu:405-521) */
                size_t tot = rel->full.n + nd;
                orc_key *m = (orc_key *)xmalloc(tot * sizeof(orc_key));
                size_t w = orc_merge_union(rel->full.data, rel->full.n, d, nd, m);
                assert(w == tot);
                vec_take(&rel->full, m, w, tot);
                rel->full_view = w;
            }
            trace_push(l, iteration, (int)r, nn, nd, rel->full.n);
        }
        iteration++;
        if (fixpoint || iteration > l->max_iteration) break;
    }
    /* finish (lie.cu:218-260): with reload_full_flag the merged full is
     * flattened and reloaded, so data_raw becomes the sorted full; without it
     * data_raw keeps the rows as loaded. */
    for (size_t r = 0; r < l->n_update; r++) {
        orc_rel *rel = l->update[r];
        rel->full_view = rel->full.n;
        if (l->reload_full_flag) {
            vec_set(&rel->raw, rel->full.data, rel->full.n);
            vec_clear(&rel->delta);
            vec_clear(&rel->newt);
        }
    }
    return iteration;
}

size_t orc_lie_trace(const orc_lie *l, const orc_trace_rec **recs) {
    *recs = l->trace;
    return l->n_trace;
}

size_t orc_lie_oplog(const orc_lie *l, const orc_op_rec **recs) {
    *recs = l->oplog;
    return l->n_oplog;
}

/* ---- program encodings -------------------------------------------------- */

static orc_key *reverse_pairs(const orc_key *e, size_t n) {
    orc_key *r = (orc_key *)xmalloc(n * sizeof(orc_key));
    for (size_t i = 0; i < n; i++) r[i] = (e[i] << 32) | (e[i] >> 32);
    return r;
}

static orc_key *copy_keys(const orc_key *src, size_t n) {
    orc_key *r = (orc_key *)xmalloc(n * sizeof(orc_key));
    if (n) memcpy(r, src, n * sizeof(orc_key));
    return r;
}

static orc_trace_rec *copy_trace(const orc_lie *l, size_t *n) {
    const orc_trace_rec *t;
    *n = orc_lie_trace(l, &t);
    orc_trace_rec *r = (orc_trace_rec *)xmalloc(*n * sizeof(orc_trace_rec));
    if (*n) memcpy(r, t, *n * sizeof(orc_trace_rec));
    return r;
}

static orc_op_rec *copy_oplog(const orc_lie *l, size_t *n) {
    const orc_op_rec *t;
    *n = orc_lie_oplog(l, &t);
    orc_op_rec *r = (orc_op_rec *)xmalloc((*n ? *n : 1) * sizeof(orc_op_rec));
    if (*n) memcpy(r, t, *n * sizeof(orc_op_rec));
    return r;
}

static uint64_t sum_delta(const orc_lie *l) {
    uint64_t s = 0;
    for (size_t i = 0; i < l->n_trace; i++) s += l->trace[i].delta;
    return s;
}

/* test/tc.cu:61-65 */
static void reorder_path(const orc_col *inner, const orc_col *outer, orc_col *newt) {
    newt[0] = inner[1];
    newt[1] = outer[1];
}
/* test/sg.cu:67-71 */
static void reorder_path1(const orc_col *inner, const orc_col *outer, orc_col *newt) {
    newt[0] = outer[1];
    newt[1] = inner[1];
}
/* test/sg.cu:90 */
static int tuple_pred_eq_11(const orc_col *t) { return t[0] != t[1]; }

/* TC program (test/tc.cu:80-149): path(a,c) :- edge(a,b), path(b,c) with
 * edge_2__2_1 (reversed edges) FULL as inner and path DELTA as outer. */
int orc_tc(const orc_key *edges, size_t n, int max_iteration, orc_result *res) {
    memset(res, 0, sizeof(*res));
    orc_key *rev = reverse_pairs(edges, n);
    orc_rel *edge = orc_rel_new("edge_2__2_1", 2, 1, 0);
    orc_rel *path = orc_rel_new("path_2__1_2", 2, 1, 0);
    path->index_flag = 0;
    orc_rel_load(path, edges, n);
    orc_rel_load(edge, rev, n);
    free(rev);
    orc_lie *lie = orc_lie_new();
    orc_lie_add_relation(lie, edge, 1);
    orc_lie_add_relation(lie, path, 0);
    orc_lie_add_join(lie, edge, ORC_FULL, path, ORC_DELTA, path, reorder_path, NULL);
    orc_lie_set_reload_full(lie, 0); /* tc.cu:122 */
    if (max_iteration >= 0) orc_lie_set_max_iteration(lie, max_iteration);
    res->iterations = orc_lie_run(lie);
    res->derived = sum_delta(lie);
    res->n_out = path->full.n;
    res->out = copy_keys(path->full.data, path->full.n);
    res->trace = copy_trace(lie, &res->n_trace);
    res->oplog = copy_oplog(lie, &res->n_oplog);
    orc_lie_free(lie);
    orc_rel_free(edge);
    orc_rel_free(path);
    return 0;
}

/* SG program (test/sg.cu:93-191): init_scc then sg_lie with tmp relation. */
int orc_sg(const orc_key *edges, size_t n, orc_result *res) {
    memset(res, 0, sizeof(*res));
    orc_rel *edge = orc_rel_new("edge_2__1_2", 2, 1, 0);
    orc_rel *sg = orc_rel_new("sg_2__2_1", 2, 1, 0);
    sg->index_flag = 0;
    orc_rel_load(sg, NULL, 0);
    orc_rel_load(edge, edges, n);

    orc_lie *init = orc_lie_new();
    orc_lie_add_relation(init, edge, 1);
    orc_lie_add_relation(init, sg, 0);
    /* sg(x, y) :- edge(p, x), edge(p, y), x != y. */
    orc_lie_add_join(init, edge, ORC_FULL, edge, ORC_FULL, sg, reorder_path, tuple_pred_eq_11);
    res->init_iterations = orc_lie_run(init);
    res->trace_init = copy_trace(init, &res->n_trace_init);
    uint64_t derived = sum_delta(init);

    /* tmp is loaded with tmp_flag false (sg.cu:157-159), so its newt is
     * sorted and deduplicated; it is registered as a tmp relation. */
    orc_rel *tmp = orc_rel_new("tmp", 2, 1, 0);
    orc_rel_load(tmp, NULL, 0);
    tmp->index_flag = 0;
    orc_lie *lie = orc_lie_new();
    orc_lie_add_relation(lie, edge, 1);
    orc_lie_add_relation(lie, sg, 0);
    orc_lie_add_tmp(lie, tmp);
    /* tmp(b,x) :- edge(a, x), sg(a, b). */
    orc_lie_add_join(lie, edge, ORC_FULL, sg, ORC_DELTA, tmp, reorder_path1, NULL);
    /* sg(x, y) :- edge(b, y), tmp(b, x). */
    orc_lie_add_join(lie, edge, ORC_FULL, tmp, ORC_NEWT, sg, reorder_path1, NULL);
    res->iterations = orc_lie_run(lie);
    res->derived = derived + sum_delta(lie);
    res->n_out = sg->full.n;
    res->out = copy_keys(sg->full.data, sg->full.n);
    res->trace = copy_trace(lie, &res->n_trace);
    res->oplog = copy_oplog(lie, &res->n_oplog);
    orc_lie_free(init);
    orc_lie_free(lie);
    orc_rel_free(edge);
    orc_rel_free(sg);
    orc_rel_free(tmp);
    return 0;
}

/* test/cspa.cu:62-96 */
static void cp_2_1__1(const orc_col *in, orc_col *out) { out[0] = in[0]; out[1] = in[0]; }
static void cp_2_1__2(const orc_col *in, orc_col *out) { out[0] = in[1]; out[1] = in[1]; }
static void cp_2_1__1_2(const orc_col *in, orc_col *out) { out[0] = in[1]; out[1] = in[0]; }
static void join_10_11(const orc_col *inner, const orc_col *outer, orc_col *out) {
    out[1] = inner[1];
    out[0] = outer[1];
}
static void join_01_11(const orc_col *inner, const orc_col *outer, orc_col *out) {
    out[0] = inner[1];
    out[1] = outer[1];
}

/* CSPA program (test/cspa.cu:100-379). */
int orc_cspa(const orc_key *assign, size_t na, const orc_key *deref, size_t nd, orc_result *res) {
    memset(res, 0, sizeof(*res));
    orc_key *rassign = reverse_pairs(assign, na);
    orc_key *rderef = reverse_pairs(deref, nd);
    orc_rel *assign_2__2_1 = orc_rel_new("assign_2__2_1", 2, 1, 0);
    orc_rel_load(assign_2__2_1, rassign, na);
    orc_rel *dereference_2__1_2 = orc_rel_new("dereference_2__1_2", 2, 1, 0);
    orc_rel_load(dereference_2__1_2, deref, nd);
    orc_rel *dereference_2__2_1 = orc_rel_new("dereference_2__2_1", 2, 1, 0);
    orc_rel_load(dereference_2__2_1, rderef, nd);
    free(rassign);
    free(rderef);

    orc_rel *vf12 = orc_rel_new("value_flow_2__1_2", 2, 1, 0);
    orc_rel_load(vf12, NULL, 0);
    orc_rel *vf21 = orc_rel_new("value_flow_2__2_1", 2, 1, 0);
    orc_rel_load(vf21, NULL, 0);
    orc_rel *ma12 = orc_rel_new("memory_alias_2__1_2", 2, 1, 0);
    orc_rel_load(ma12, NULL, 0);
    orc_rel *ma21 = orc_rel_new("memory_alias_2__2_1", 2, 1, 0);
    orc_rel_load(ma21, NULL, 0);

    orc_lie *init = orc_lie_new();
    orc_lie_add_relation(init, vf12, 0);
    orc_lie_add_relation(init, vf21, 0);
    orc_lie_add_relation(init, ma12, 0);
    orc_lie_add_relation(init, ma21, 0);
    orc_lie_add_relation(init, assign_2__2_1, 1);
    orc_lie_add_copy(init, assign_2__2_1, ORC_FULL, vf12, cp_2_1__1);
    orc_lie_add_copy(init, assign_2__2_1, ORC_FULL, vf12, cp_2_1__2);
    orc_lie_add_copy(init, assign_2__2_1, ORC_FULL, vf12, cp_2_1__1_2);
    orc_lie_add_copy(init, assign_2__2_1, ORC_FULL, ma12, cp_2_1__1);
    orc_lie_add_copy(init, assign_2__2_1, ORC_FULL, ma12, cp_2_1__2);
    orc_lie_add_copy(init, vf12, ORC_DELTA, vf21, cp_2_1__1_2);
    orc_lie_add_copy(init, ma12, ORC_DELTA, ma21, cp_2_1__1_2);
    res->init_iterations = orc_lie_run(init);
    res->trace_init = copy_trace(init, &res->n_trace_init);
    uint64_t derived = sum_delta(init);

    orc_rel *vff12 = orc_rel_new("value_flow_forward_2__1_2", 2, 1, 0);
    orc_rel_load(vff12, NULL, 0);
    orc_rel *vff21 = orc_rel_new("value_flow_forward_2__2_1", 2, 1, 0);
    orc_rel_load(vff21, NULL, 0);
    orc_rel *va = orc_rel_new("value_alias_2__1_2", 2, 1, 0);
    va->index_flag = 0;
    orc_rel_load(va, NULL, 0);
    orc_rel *tmp_def = orc_rel_new("tmp_rel_def", 2, 1, 0);
    tmp_def->index_flag = 0;
    orc_rel_load(tmp_def, NULL, 0);
    orc_rel *tmp_ma1 = orc_rel_new("tmp_rel_ma1", 2, 1, 1);
    tmp_ma1->index_flag = 0;
    orc_rel_load(tmp_ma1, NULL, 0);
    orc_rel *tmp_ma2 = orc_rel_new("tmp_rel_ma2", 2, 1, 1);
    tmp_ma2->index_flag = 0;
    orc_rel_load(tmp_ma2, NULL, 0);

    orc_lie *a = orc_lie_new();
    orc_lie_add_relation(a, assign_2__2_1, 1);
    orc_lie_add_relation(a, dereference_2__1_2, 1);
    orc_lie_add_relation(a, dereference_2__2_1, 1);
    orc_lie_add_relation(a, vf12, 0);
    orc_lie_add_relation(a, vf21, 0);
    orc_lie_add_relation(a, ma12, 0);
    orc_lie_add_relation(a, ma21, 0);
    orc_lie_add_relation(a, va, 0);
    orc_lie_add_tmp(a, tmp_def);
    orc_lie_add_tmp(a, tmp_ma1);
    orc_lie_add_tmp(a, tmp_ma2);
    /* ValueFlow(x, y) :- ValueFlow(x, z), ValueFlow(z, y). */
    orc_lie_add_join(a, vf12, ORC_FULL, vf21, ORC_DELTA, vf12, join_10_11, NULL);
    orc_lie_add_join(a, vf21, ORC_FULL, vf12, ORC_DELTA, vf12, join_01_11, NULL);
    /* ValueAlias(x, y) :- ValueFlow(z, x), ValueFlow(z, y). */
    orc_lie_add_join(a, vf12, ORC_FULL, vf12, ORC_DELTA, va, join_01_11, NULL);
    orc_lie_add_join(a, vf12, ORC_FULL, vf12, ORC_DELTA, va, join_10_11, NULL);
    /* ValueFlow(x, y) :- Assign(x, z), MemoryAlias(z, y). */
    orc_lie_add_join(a, assign_2__2_1, ORC_FULL, ma12, ORC_DELTA, vf12, join_01_11, NULL);
    /* tmp_rel_def(z, x) :- Dereference(y, x), ValueAlias(y, z) */
    orc_lie_add_join(a, dereference_2__1_2, ORC_FULL, va, ORC_DELTA, tmp_def, join_10_11, NULL);
    /* MemoryAlias(x, w) :- Dereference(z, w), tmp_rel_def(z, x) */
    orc_lie_add_join(a, dereference_2__1_2, ORC_FULL, tmp_def, ORC_NEWT, ma12, join_10_11, NULL);
    /* ValueAlias(x,y) :- ValueFlow(z,x), MemoryAlias(z,w), ValueFlow(w,y). */
    orc_lie_add_join(a, ma12, ORC_FULL, vf12, ORC_DELTA, tmp_ma1, join_01_11, NULL);
    orc_lie_add_join(a, vf12, ORC_FULL, ma12, ORC_DELTA, tmp_ma1, join_10_11, NULL);
    orc_lie_add_join(a, vf12, ORC_FULL, tmp_ma1, ORC_NEWT, va, join_10_11, NULL);
    orc_lie_add_join(a, ma21, ORC_FULL, vf12, ORC_DELTA, tmp_ma2, join_01_11, NULL);
    orc_lie_add_join(a, vf12, ORC_FULL, tmp_ma2, ORC_NEWT, va, join_01_11, NULL);
    orc_lie_add_acopy(a, vf12, vf21, cp_2_1__1_2);
    orc_lie_add_acopy(a, ma12, ma21, cp_2_1__1_2);
    res->iterations = orc_lie_run(a);
    res->derived = derived + sum_delta(a);
    res->trace = copy_trace(a, &res->n_trace);
    res->oplog = copy_oplog(a, &res->n_oplog);
    res->n_out = va->full.n;
    res->out = copy_keys(va->full.data, va->full.n);
    res->extra[0] = copy_keys(vf12->full.data, vf12->full.n);
    res->n_extra[0] = vf12->full.n;
    res->extra[1] = copy_keys(ma12->full.data, ma12->full.n);
    res->n_extra[1] = ma12->full.n;
    res->sizes[0] = vf12->full.n;
    res->sizes[1] = ma12->full.n;
    res->sizes[2] = va->full.n;
    res->sizes[3] = vf21->full.n;

    orc_lie_free(init);
    orc_lie_free(a);
    orc_rel *rels[] = {assign_2__2_1, dereference_2__1_2, dereference_2__2_1, vf12, vf21, ma12,
                       ma21, vff12, vff21, va, tmp_def, tmp_ma1, tmp_ma2};
    for (size_t i = 0; i < sizeof(rels) / sizeof(rels[0]); i++) orc_rel_free(rels[i]);
    return 0;
}

void orc_result_free(orc_result *res) {
    free(res->out);
    free(res->trace);
    free(res->trace_init);
    free(res->oplog);
    for (int i = 0; i < 3; i++) free(res->extra[i]);
    memset(res, 0, sizeof(*res));
}
