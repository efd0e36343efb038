/*
 * cpu_souffle.c -- multi-core CPU baselines restating the loop plans that
 * Souffle generated for the reference's Datalog programs.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times
 * these and tests/test_oracle.py checks their results against the oracle; the
 * product path never links or calls them.
 *
 * Souffle itself is not available here (no binary, no runtime headers), so the
 * generated C++ (test/souffle/{tc,sg,cspa}.cpp) cannot be compiled.  These functions
 * run the same semi-naive plans -- the same rule versions, loop nests, range
 * lookups and contains() filters, with delta / new relations swapped at the end
 * of each iteration -- over OpenMP threads:
 *
 *   souffle_sg  test/souffle/sg.cpp:474-645:
 *     sg(x,y) :- edge(p,x), edge(p,y), x != y.        (once)
 *     sg(x,y) :- edge(a,x), sg(a,b), edge(b,y).        (recursive, delta on sg)
 *       for edge(a,x): for delta_sg(a,b): for edge(b,y):
 *         if !sg.contains(x,y): new_sg.insert(x,y)
 *   souffle_tc  test/souffle/tc.cpp:379-489 (tc.dl's non-linear rule):
 *     path(x,y) :- edge(x,y).
 *     path(x,z) :- path(x,y), path(y,z).  two versions per iteration:
 *       for delta(x,y): for path(y,z):
 *         if !path.contains(x,z) && !delta.contains(y,z): new.insert(x,z)
 *       for path(x,y): for delta(y,z):
 *         if !path.contains(x,z): new.insert(x,z)
 *   souffle_cspa  test/souffle/cspa.cpp:344-815 (cspa.dl's ten rules; the
 *     rule versions are listed above souffle_cspa below)
 *
 * Souffle's pfor splits the outer relation of each loop nest into chunks.  Here
 * the outer tuples are split by the first column of the tuple they derive (x
 * in every rule), so each thread owns every relation's tuples whose first
 * column x has x mod T == t: its contains() and insert() calls touch only its
 * own hash sets, and the range lookups read other threads' adjacency lists,
 * which change only between iterations.  Souffle's B-trees become
 * open-addressing hash sets (membership) plus per-node adjacency arrays
 * (ranges on the first column).
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gdl_oracle.h"

/* ---- per-thread hash set of u64 keys (EMPTY = ~0, which no pair of u32 ids
 * below 2^32 - 1 in both columns reaches; such a key is tracked by a flag) */
typedef struct {
    uint64_t *slots;
    uint64_t mask, size;
    int has_max;
} hset;

static uint64_t smix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static int hset_init(hset *s, uint64_t cap) {
    uint64_t c = 256;
    while (c < cap) c <<= 1;
    s->slots = (uint64_t *)malloc(c * sizeof(uint64_t));
    if (!s->slots) return -1;
    memset(s->slots, 0xff, c * sizeof(uint64_t));
    s->mask = c - 1;
    s->size = 0;
    s->has_max = 0;
    return 0;
}

static void hset_free(hset *s) {
    free(s->slots);
    s->slots = NULL;
}

static int hset_contains(const hset *s, uint64_t k) {
    if (k == ~0ull) return s->has_max;
    for (uint64_t i = smix(k) & s->mask;; i = (i + 1) & s->mask) {
        const uint64_t v = s->slots[i];
        if (v == k) return 1;
        if (v == ~0ull) return 0;
    }
}

static int hset_grow(hset *s);

/* 1 if inserted, 0 if present, -1 out of memory */
static int hset_insert(hset *s, uint64_t k) {
    if (k == ~0ull) {
        if (s->has_max) return 0;
        s->has_max = 1;
        return 1;
    }
    if (2 * (s->size + 1) > s->mask + 1 && hset_grow(s)) return -1;
    for (uint64_t i = smix(k) & s->mask;; i = (i + 1) & s->mask) {
        const uint64_t v = s->slots[i];
        if (v == k) return 0;
        if (v == ~0ull) {
            s->slots[i] = k;
            s->size++;
            return 1;
        }
    }
}

static int hset_grow(hset *s) {
    hset n;
    if (hset_init(&n, 2 * (s->mask + 1))) return -1;
    n.has_max = s->has_max;
    for (uint64_t i = 0; i <= s->mask; ++i) {
        const uint64_t k = s->slots[i];
        if (k == ~0ull) continue;
        uint64_t j = smix(k) & n.mask;
        while (n.slots[j] != ~0ull) j = (j + 1) & n.mask;
        n.slots[j] = k;
        n.size++;
    }
    free(s->slots);
    *s = n;
    return 0;
}

/* ---- growable u32 / u64 arrays ---------------------------------------- */
typedef struct {
    uint32_t *v;
    uint32_t n, cap;
} vec32;

static int vec32_push(vec32 *a, uint32_t x) {
    if (a->n == a->cap) {
        uint32_t c = a->cap ? 2 * a->cap : 4;
        uint32_t *p = (uint32_t *)realloc(a->v, (size_t)c * sizeof(uint32_t));
        if (!p) return -1;
        a->v = p;
        a->cap = c;
    }
    a->v[a->n++] = x;
    return 0;
}

typedef struct {
    uint64_t *v;
    uint64_t n, cap;
} vec64;

static int vec64_push(vec64 *a, uint64_t x) {
    if (a->n == a->cap) {
        uint64_t c = a->cap ? 2 * a->cap : 64;
        uint64_t *p = (uint64_t *)realloc(a->v, c * sizeof(uint64_t));
        if (!p) return -1;
        a->v = p;
        a->cap = c;
    }
    a->v[a->n++] = x;
    return 0;
}

/* Number of node ids (1 + the largest id) of an edge list. */
static uint32_t node_count(const uint64_t *e, uint64_t n) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t a = e[i] >> 32, b = e[i] & 0xffffffffull;
        if (a > m) m = a;
        if (b > m) m = b;
    }
    return n ? (uint32_t)(m + 1) : 0;
}

/* Static adjacency (CSR) of an edge list by its first column. */
typedef struct {
    uint64_t *off;
    uint32_t *dst;
} csr;

static int csr_build(csr *g, const uint64_t *e, uint64_t n, uint32_t nodes) {
    g->off = (uint64_t *)calloc((size_t)nodes + 1, sizeof(uint64_t));
    g->dst = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!g->off || !g->dst) return -1;
    for (uint64_t i = 0; i < n; ++i) g->off[(e[i] >> 32) + 1]++;
    for (uint32_t v = 0; v < nodes; ++v) g->off[v + 1] += g->off[v];
    uint64_t *pos = (uint64_t *)malloc(((size_t)nodes + 1) * sizeof(uint64_t));
    if (!pos) return -1;
    memcpy(pos, g->off, ((size_t)nodes + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) g->dst[pos[e[i] >> 32]++] = (uint32_t)(e[i] & 0xffffffffull);
    free(pos);
    return 0;
}

static void csr_free(csr *g) {
    free(g->off);
    free(g->dst);
}

static uint64_t pk(uint64_t a, uint64_t b) { return (a << 32) | b; }

/* Per-thread state: the thread's partition of the full relation and of the
 * iteration's new relation (hash sets), the new tuples in insertion order,
 * and the iteration's raw candidate count. */
typedef struct {
    hset full, fresh;
    vec64 fresh_list;
    uint64_t checksum;
    int oom;
} part;

static int parts_init(part *p, int T, uint64_t cap) {
    for (int t = 0; t < T; ++t) {
        memset(&p[t], 0, sizeof(part));
        if (hset_init(&p[t].full, cap) || hset_init(&p[t].fresh, 1024)) return -1;
    }
    return 0;
}

static void parts_free(part *p, int T) {
    for (int t = 0; t < T; ++t) {
        hset_free(&p[t].full);
        hset_free(&p[t].fresh);
        free(p[t].fresh_list.v);
    }
}

/* After an iteration: moves thread t's new tuples into its full set and into
 * the next delta's adjacency (delta_adj[x], x owned by t); returns |new|. */
static uint64_t commit_new(part *p, vec32 *delta_adj, vec32 *full_adj, uint32_t nodes, int T,
                           int *oom) {
    uint64_t total = 0;
#pragma omp parallel num_threads(T) reduction(+ : total)
    {
        const int t = omp_get_thread_num();
        part *me = &p[t];
        for (uint32_t x = (uint32_t)t; x < nodes; x += (uint32_t)T) delta_adj[x].n = 0;
        for (uint64_t i = 0; i < me->fresh_list.n; ++i) {
            const uint64_t k = me->fresh_list.v[i];
            const uint32_t x = (uint32_t)(k >> 32), y = (uint32_t)k;
            if (hset_insert(&me->full, k) < 0 || vec32_push(&delta_adj[x], y) ||
                (full_adj && vec32_push(&full_adj[x], y)))
                me->oom = 1;
            me->checksum += smix(k);
        }
        total += me->fresh_list.n;
        me->fresh_list.n = 0;
        hset_free(&me->fresh);
        if (hset_init(&me->fresh, 1024)) me->oom = 1;
    }
    for (int t = 0; t < T; ++t) *oom |= p[t].oom;
    return total;
}

static void vecs_free(vec32 *a, uint32_t n) {
    if (!a) return;
    for (uint32_t i = 0; i < n; ++i) free(a[i].v);
    free(a);
}

int souffle_sg(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *checksum, int *iterations, uint64_t *delta_trace,
               int trace_cap) {
    const int T = threads > 0 ? threads : 1;
    const uint32_t nodes = node_count(edges, n);
    csr g;
    part *p = (part *)calloc((size_t)T, sizeof(part));
    vec32 *delta = (vec32 *)calloc((size_t)nodes + 1, sizeof(vec32));
    int oom = 0;
    if (!p || !delta || csr_build(&g, edges, n, nodes) || parts_init(p, T, 1024)) {
        free(p);
        free(delta);
        return -1;
    }
    /* sg(x,y) :- edge(p,x), edge(p,y), x != y.  Outer edge(p,x) split by x. */
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        part *me = &p[t];
        for (uint32_t pp = 0; pp < nodes; ++pp)
            for (uint64_t i = g.off[pp]; i < g.off[pp + 1]; ++i) {
                const uint32_t x = g.dst[i];
                if ((int)(x % (uint32_t)T) != t) continue;
                for (uint64_t j = g.off[pp]; j < g.off[pp + 1]; ++j) {
                    const uint32_t y = g.dst[j];
                    if (x == y) continue;
                    const int r = hset_insert(&me->fresh, pk(x, y));
                    if (r < 0 || (r > 0 && vec64_push(&me->fresh_list, pk(x, y)))) me->oom = 1;
                }
            }
    }
    uint64_t size = commit_new(p, delta, NULL, nodes, T, &oom);
    int it = 0;
    if (trace_cap > 0) delta_trace[0] = size;
    /* sg(x,y) :- edge(a,x), sg(a,b), edge(b,y) over delta_sg. */
    while (!oom && size && (max_iteration < 0 || it < max_iteration)) {
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num();
            part *me = &p[t];
            for (uint32_t a = 0; a < nodes && !me->oom; ++a) {
                const vec32 *da = &delta[a];
                if (!da->n) continue;
                for (uint64_t i = g.off[a]; i < g.off[a + 1]; ++i) {
                    const uint32_t x = g.dst[i];
                    if ((int)(x % (uint32_t)T) != t) continue;
                    for (uint32_t j = 0; j < da->n; ++j) {
                        const uint32_t b = da->v[j];
                        for (uint64_t k = g.off[b]; k < g.off[b + 1]; ++k) {
                            const uint64_t key = pk(x, g.dst[k]);
                            if (hset_contains(&me->full, key)) continue;
                            const int r = hset_insert(&me->fresh, key);
                            if (r < 0 || (r > 0 && vec64_push(&me->fresh_list, key))) me->oom = 1;
                        }
                    }
                }
            }
        }
        const uint64_t nd = commit_new(p, delta, NULL, nodes, T, &oom);
        ++it;
        if (it < trace_cap) delta_trace[it] = nd;
        size += nd;
        if (!nd) break;
    }
    uint64_t cs = 0;
    for (int t = 0; t < T; ++t) cs += p[t].checksum;
    *n_out = size;
    *checksum = cs;
    *iterations = it;
    parts_free(p, T);
    free(p);
    vecs_free(delta, nodes + 1);
    csr_free(&g);
    return oom ? -1 : 0;
}

int souffle_tc(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *checksum, int *iterations, uint64_t *delta_trace,
               int trace_cap) {
    const int T = threads > 0 ? threads : 1;
    const uint32_t nodes = node_count(edges, n);
    part *p = (part *)calloc((size_t)T, sizeof(part));
    vec32 *delta = (vec32 *)calloc((size_t)nodes + 1, sizeof(vec32));
    vec32 *path = (vec32 *)calloc((size_t)nodes + 1, sizeof(vec32));
    hset *dset = (hset *)calloc((size_t)T, sizeof(hset));
    int oom = 0;
    if (!p || !delta || !path || !dset || parts_init(p, T, 1024)) {
        free(p);
        free(delta);
        free(path);
        free(dset);
        return -1;
    }
    /* path(x,y) :- edge(x,y); delta_path = path */
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        part *me = &p[t];
        for (uint64_t i = 0; i < n; ++i) {
            if ((int)((edges[i] >> 32) % (uint64_t)T) != t) continue;
            const int r = hset_insert(&me->fresh, edges[i]);
            if (r < 0 || (r > 0 && vec64_push(&me->fresh_list, edges[i]))) me->oom = 1;
        }
    }
    uint64_t size = 0;
    int it = 0;
    for (;;) {
        /* the new tuples become path and the next delta; delta's hash sets
         * (for the contains() of the first version) follow them */
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num();
            hset_free(&dset[t]);
            if (hset_init(&dset[t], 2 * p[t].fresh_list.n + 16)) p[t].oom = 1;
            for (uint64_t i = 0; i < p[t].fresh_list.n; ++i)
                if (hset_insert(&dset[t], p[t].fresh_list.v[i]) < 0) p[t].oom = 1;
        }
        const uint64_t nd = commit_new(p, delta, path, nodes, T, &oom);
        if (it > 0 && it - 1 < trace_cap) delta_trace[it - 1] = nd;
        size += nd;
        if (oom || !nd || (max_iteration >= 0 && it >= max_iteration)) break;
        ++it;
        /* path(x,z) :- path(x,y), path(y,z), both versions; outer split by x */
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num();
            part *me = &p[t];
            for (uint32_t x = (uint32_t)t; x < nodes && !me->oom; x += (uint32_t)T) {
                const vec32 *dx = &delta[x];
                for (uint32_t i = 0; i < dx->n; ++i) {  /* delta(x,y), path(y,z) */
                    const uint32_t y = dx->v[i];
                    const vec32 *py = &path[y];
                    const hset *dy = &dset[y % (uint32_t)T];
                    for (uint32_t j = 0; j < py->n; ++j) {
                        const uint64_t key = pk(x, py->v[j]);
                        if (hset_contains(&me->full, key) || hset_contains(dy, pk(y, py->v[j])))
                            continue;
                        const int r = hset_insert(&me->fresh, key);
                        if (r < 0 || (r > 0 && vec64_push(&me->fresh_list, key))) me->oom = 1;
                    }
                }
                const vec32 *px = &path[x];
                for (uint32_t i = 0; i < px->n; ++i) {  /* path(x,y), delta(y,z) */
                    const vec32 *dy = &delta[px->v[i]];
                    for (uint32_t j = 0; j < dy->n; ++j) {
                        const uint64_t key = pk(x, dy->v[j]);
                        if (hset_contains(&me->full, key)) continue;
                        const int r = hset_insert(&me->fresh, key);
                        if (r < 0 || (r > 0 && vec64_push(&me->fresh_list, key))) me->oom = 1;
                    }
                }
            }
        }
    }
    uint64_t cs = 0;
    for (int t = 0; t < T; ++t) {
        cs += p[t].checksum;
        hset_free(&dset[t]);
    }
    *n_out = size;
    *checksum = cs;
    *iterations = it;
    parts_free(p, T);
    free(p);
    free(dset);
    vecs_free(delta, nodes + 1);
    vecs_free(path, nodes + 1);
    return oom ? -1 : 0;
}

/* ---- CSPA: test/souffle/cspa.cpp:344-815 ---------------------------------
 * Non-recursive rules (cspa.dl:34-39) seed ValueFlow and MemoryAlias from
 * assign; then one loop over the mutually recursive ValueFlow / MemoryAlias /
 * ValueAlias with these rule versions (outer relation first), all reading the
 * previous iteration's deltas:
 *   MA(x,w)  :- deref(y,x), dVA(y,z), deref(z,w)             !MA(x,w)
 *   VA(x,y)  :- dVF(z,x), VF(z,y)                             !VA, !dVF(z,y)
 *   VA(x,y)  :- VF(z,x), dVF(z,y)                             !VA
 *   VA(x,y)  :- dVF(z,x), MA(z,w), VF(w,y)   !dMA(z,w)        !VA, !dVF(w,y)
 *   VA(x,y)  :- VF(z,x), dMA(z,w), VF(w,y)                    !VA, !dVF(w,y)
 *   VA(x,y)  :- VF(z,x), MA(z,w), dVF(w,y)                    !VA
 *   VF(x,y)  :- dVF(x,z), VF(z,y)                             !VF, !dVF(z,y)
 *   VF(x,y)  :- VF(x,z), dVF(z,y)                             !VF
 *   VF(x,y)  :- assign(x,z), dMA(z,y)                         !VF
 * Each thread derives the tuples whose first column x has x mod T == t. */

typedef struct {
    part *p;       /* per thread: full set, new set, new list */
    vec32 *full;   /* adjacency of the full relation by column 0 (NULL: not kept) */
    vec32 *delta;  /* adjacency of the delta by column 0 */
    hset *dset;    /* per thread: the delta's tuples (first column owned by t) */
    uint64_t size;
} crel;

static int crel_init(crel *r, int T, uint32_t nodes, int keep_full) {
    memset(r, 0, sizeof(*r));
    r->p = (part *)calloc((size_t)T, sizeof(part));
    r->delta = (vec32 *)calloc((size_t)nodes + 1, sizeof(vec32));
    r->dset = (hset *)calloc((size_t)T, sizeof(hset));
    if (keep_full) r->full = (vec32 *)calloc((size_t)nodes + 1, sizeof(vec32));
    if (!r->p || !r->delta || !r->dset || (keep_full && !r->full) || parts_init(r->p, T, 1024))
        return -1;
    for (int t = 0; t < T; ++t)
        if (hset_init(&r->dset[t], 16)) return -1;
    return 0;
}

static void crel_free(crel *r, int T, uint32_t nodes) {
    if (r->p) parts_free(r->p, T);
    free(r->p);
    if (r->dset)
        for (int t = 0; t < T; ++t) hset_free(&r->dset[t]);
    free(r->dset);
    vecs_free(r->delta, nodes + 1);
    vecs_free(r->full, nodes + 1);
}

/* thread t's candidate (x,y): kept when not in full and not yet new */
static void crel_offer(crel *r, int t, uint64_t key) {
    part *me = &r->p[t];
    if (hset_contains(&me->full, key)) return;
    const int ins = hset_insert(&me->fresh, key);
    if (ins < 0 || (ins > 0 && vec64_push(&me->fresh_list, key))) me->oom = 1;
}

/* end of an iteration: new -> full, delta, delta set; returns |new| */
static uint64_t crel_commit(crel *r, uint32_t nodes, int T, int *oom) {
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        hset_free(&r->dset[t]);
        if (hset_init(&r->dset[t], 2 * r->p[t].fresh_list.n + 16)) r->p[t].oom = 1;
        for (uint64_t i = 0; i < r->p[t].fresh_list.n; ++i)
            if (hset_insert(&r->dset[t], r->p[t].fresh_list.v[i]) < 0) r->p[t].oom = 1;
    }
    const uint64_t nd = commit_new(r->p, r->delta, r->full, nodes, T, oom);
    r->size += nd;
    return nd;
}

static int dcontains(const crel *r, int T, uint64_t a, uint64_t b) {
    return hset_contains(&r->dset[a % (uint64_t)T], pk(a, b));
}

int souffle_cspa(const uint64_t *assign, uint64_t na, const uint64_t *deref, uint64_t nd_,
                 int threads, int max_iteration, uint64_t *sizes, uint64_t *checksums,
                 int *iterations) {
    const int T = threads > 0 ? threads : 1;
    uint32_t nodes = node_count(assign, na), n2 = node_count(deref, nd_);
    if (n2 > nodes) nodes = n2;
    crel VF, MA, VA;
    csr dg;
    int oom = 0;
    if (crel_init(&VF, T, nodes, 1) || crel_init(&MA, T, nodes, 1) || crel_init(&VA, T, nodes, 0) ||
        csr_build(&dg, deref, nd_, nodes))
        return -1;
    /* cspa.dl:34-39 */
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        for (uint64_t i = 0; i < na; ++i) {
            const uint64_t a0 = assign[i] >> 32, a1 = assign[i] & 0xffffffffull;
            if ((int)(a1 % T) == t) crel_offer(&MA, t, pk(a1, a1));
            if ((int)(a0 % T) == t) {
                crel_offer(&MA, t, pk(a0, a0));
                crel_offer(&VF, t, pk(a0, a1));
                crel_offer(&VF, t, pk(a0, a0));
            }
            if ((int)(a1 % T) == t) crel_offer(&VF, t, pk(a1, a1));
        }
    }
    crel_commit(&VF, nodes, T, &oom);
    crel_commit(&MA, nodes, T, &oom);
    crel_commit(&VA, nodes, T, &oom);
    int it = 0;
    while (!oom && (max_iteration < 0 || it < max_iteration)) {
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num();
            /* MA(x,w) :- deref(y,x), dVA(y,z), deref(z,w) */
            for (uint32_t y = 0; y < nodes; ++y) {
                const vec32 *dv = &VA.delta[y];
                if (!dv->n) continue;
                for (uint64_t i = dg.off[y]; i < dg.off[y + 1]; ++i) {
                    const uint32_t x = dg.dst[i];
                    if ((int)(x % (uint32_t)T) != t) continue;
                    for (uint32_t j = 0; j < dv->n; ++j) {
                        const uint32_t z = dv->v[j];
                        for (uint64_t k = dg.off[z]; k < dg.off[z + 1]; ++k)
                            crel_offer(&MA, t, pk(x, dg.dst[k]));
                    }
                }
            }
            for (uint32_t z = 0; z < nodes; ++z) {
                const vec32 *dz = &VF.delta[z], *fz = &VF.full[z];
                const vec32 *mz = &MA.full[z], *dmz = &MA.delta[z];
                /* VA(x,y) :- dVF(z,x), VF(z,y) and VF(z,x), dVF(z,y) */
                for (uint32_t i = 0; i < dz->n; ++i) {
                    const uint32_t x = dz->v[i];
                    if ((int)(x % (uint32_t)T) != t) continue;
                    for (uint32_t j = 0; j < fz->n; ++j)
                        if (!dcontains(&VF, T, z, fz->v[j])) crel_offer(&VA, t, pk(x, fz->v[j]));
                    /* VA(x,y) :- dVF(z,x), MA(z,w), VF(w,y) */
                    for (uint32_t j = 0; j < mz->n; ++j) {
                        const uint32_t w = mz->v[j];
                        if (dcontains(&MA, T, z, w)) continue;
                        const vec32 *fw = &VF.full[w];
                        for (uint32_t k = 0; k < fw->n; ++k)
                            if (!dcontains(&VF, T, w, fw->v[k]))
                                crel_offer(&VA, t, pk(x, fw->v[k]));
                    }
                }
                for (uint32_t i = 0; i < fz->n; ++i) {
                    const uint32_t x = fz->v[i];
                    if ((int)(x % (uint32_t)T) != t) continue;
                    for (uint32_t j = 0; j < dz->n; ++j) crel_offer(&VA, t, pk(x, dz->v[j]));
                    /* VA(x,y) :- VF(z,x), dMA(z,w), VF(w,y) */
                    for (uint32_t j = 0; j < dmz->n; ++j) {
                        const vec32 *fw = &VF.full[dmz->v[j]];
                        for (uint32_t k = 0; k < fw->n; ++k)
                            if (!dcontains(&VF, T, dmz->v[j], fw->v[k]))
                                crel_offer(&VA, t, pk(x, fw->v[k]));
                    }
                    /* VA(x,y) :- VF(z,x), MA(z,w), dVF(w,y) */
                    for (uint32_t j = 0; j < mz->n; ++j) {
                        const vec32 *dw = &VF.delta[mz->v[j]];
                        for (uint32_t k = 0; k < dw->n; ++k) crel_offer(&VA, t, pk(x, dw->v[k]));
                    }
                }
            }
            /* VF(x,y) :- dVF(x,z), VF(z,y) and VF(x,z), dVF(z,y) */
            for (uint32_t x = (uint32_t)t; x < nodes; x += (uint32_t)T) {
                const vec32 *dx = &VF.delta[x], *fx = &VF.full[x];
                for (uint32_t i = 0; i < dx->n; ++i) {
                    const uint32_t z = dx->v[i];
                    const vec32 *fz = &VF.full[z];
                    for (uint32_t j = 0; j < fz->n; ++j)
                        if (!dcontains(&VF, T, z, fz->v[j])) crel_offer(&VF, t, pk(x, fz->v[j]));
                }
                for (uint32_t i = 0; i < fx->n; ++i) {
                    const vec32 *dz = &VF.delta[fx->v[i]];
                    for (uint32_t j = 0; j < dz->n; ++j) crel_offer(&VF, t, pk(x, dz->v[j]));
                }
            }
            /* VF(x,y) :- assign(x,z), dMA(z,y) */
            for (uint64_t i = 0; i < na; ++i) {
                const uint32_t x = (uint32_t)(assign[i] >> 32), z = (uint32_t)assign[i];
                if ((int)(x % (uint32_t)T) != t) continue;
                const vec32 *dz = &MA.delta[z];
                for (uint32_t j = 0; j < dz->n; ++j) crel_offer(&VF, t, pk(x, dz->v[j]));
            }
        }
        ++it;
        const uint64_t a = crel_commit(&MA, nodes, T, &oom);
        const uint64_t b = crel_commit(&VA, nodes, T, &oom);
        const uint64_t c = crel_commit(&VF, nodes, T, &oom);
        if (!a && !b && !c) break;
    }
    for (int r = 0; r < 3; ++r) {
        crel *R = r == 0 ? &VF : r == 1 ? &MA : &VA;
        uint64_t cs = 0;
        for (int t = 0; t < T; ++t) {
            cs += R->p[t].checksum;
            oom |= R->p[t].oom;
        }
        sizes[r] = R->size;
        checksums[r] = cs;
    }
    *iterations = it;
    crel_free(&VF, T, nodes);
    crel_free(&MA, T, nodes);
    crel_free(&VA, T, nodes);
    csr_free(&dg);
    return oom ? -1 : 0;
}
