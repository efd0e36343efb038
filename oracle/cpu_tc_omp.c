/*
 * cpu_tc_omp.c -- multi-core CPU baseline of the TC fixpoint.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times
 * it and tests/test_oracle.py checks it against the oracle; the product path
 * never links or calls it.
 *
 * The same semi-naive loop as the reference's TC program (test/tc.cu:80-149,
 * path(a, c) :- edge(a, b), path(b, c) over src/lie.cu:26-274): the same
 * initial delta (the first |full| raw rows, src/relation.cu:764-795), the
 * same per-body delta and the same iteration count.  It is parallel the way a
 * CPU does it best (the way Souffle's pfor splits outer relations): a new
 * tuple (a, c) keeps the second column c of the delta tuple (b, c) it comes
 * from, so partitioning path by c splits the whole fixpoint into independent
 * per-thread fixpoints with no shared state -- each thread owns the c with
 * c mod T == t, its FULL is a private open-addressing hash set, its delta a
 * private array.  The global |delta| of body i is the sum over threads.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gdl_oracle.h"

typedef struct {
    uint64_t *slots;
    uint64_t mask;
    uint64_t size;
    int has_max;  /* the key ~0 (EMPTY) is a member */
} cset;

static uint64_t cmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static int cset_init(cset *s, uint64_t cap) {
    uint64_t c = 1024;
    while (c < cap) c <<= 1;
    s->slots = (uint64_t *)malloc(c * sizeof(uint64_t));
    if (!s->slots) return -1;
    memset(s->slots, 0xff, c * sizeof(uint64_t));
    s->mask = c - 1;
    s->size = 0;
    s->has_max = 0;
    return 0;
}

static int cset_insert(cset *s, uint64_t k);

static int cset_grow(cset *s) {
    cset n;
    if (cset_init(&n, 2 * (s->mask + 1))) return -1;
    n.has_max = s->has_max;
    for (uint64_t i = 0; i <= s->mask; ++i)
        if (s->slots[i] != ~0ull) cset_insert(&n, s->slots[i]);
    free(s->slots);
    *s = n;
    return 0;
}

/* 1 if k was new, 0 if present, -1 on allocation failure */
static int cset_insert(cset *s, uint64_t k) {
    if (k == ~0ull) {
        if (s->has_max) return 0;
        s->has_max = 1;
        return 1;
    }
    if (2 * (s->size + 1) > s->mask + 1 && cset_grow(s)) return -1;
    uint64_t i = cmix(k) & s->mask;
    while (s->slots[i] != ~0ull) {
        if (s->slots[i] == k) return 0;
        i = (i + 1) & s->mask;
    }
    s->slots[i] = k;
    ++s->size;
    return 1;
}

typedef struct {
    uint64_t *v;
    uint64_t n, cap;
} kvec;

static int kvec_push(kvec *a, uint64_t k) {
    if (a->n == a->cap) {
        uint64_t nc = a->cap ? 2 * a->cap : 1024;
        uint64_t *nv = (uint64_t *)realloc(a->v, nc * sizeof(uint64_t));
        if (!nv) return -1;
        a->v = nv;
        a->cap = nc;
    }
    a->v[a->n++] = k;
    return 0;
}

static int cmp_u64(const void *x, const void *y) {
    const uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : a > b;
}

int cpu_tc_omp(const uint64_t *edges, uint64_t n, int threads, int max_iteration,
               uint64_t *n_out, uint64_t *derived, int *iterations, uint64_t *delta_trace,
               int trace_cap) {
    if (threads < 1) threads = 1;
    *n_out = *derived = 0;
    *iterations = 0;
    if (n == 0) {
        *iterations = 1;  /* one body, nothing derived */
        if (trace_cap > 0) delta_trace[0] = 0;
        return 0;
    }
    /* FULL0 = the distinct edges; delta0 = the first |FULL0| raw rows */
    uint64_t *u = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!u) return -1;
    memcpy(u, edges, n * sizeof(uint64_t));
    qsort(u, n, sizeof(uint64_t), cmp_u64);
    uint64_t nu = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (i == 0 || u[i] != u[i - 1]) u[nu++] = u[i];
    /* predecessors of b: the a with an edge (a, b), as CSR over b */
    uint64_t max_id = 0;
    for (uint64_t i = 0; i < nu; ++i) {
        const uint64_t a = u[i] >> 32, b = u[i] & 0xffffffffull;
        if (a > max_id) max_id = a;
        if (b > max_id) max_id = b;
    }
    const uint64_t nn = max_id + 1;
    uint64_t *off = (uint64_t *)calloc(nn + 1, sizeof(uint64_t));
    uint32_t *pred = (uint32_t *)malloc(nu * sizeof(uint32_t));
    if (!off || !pred) {
        free(u), free(off), free(pred);
        return -1;
    }
    for (uint64_t i = 0; i < nu; ++i) ++off[(u[i] & 0xffffffffull) + 1];
    for (uint64_t v = 0; v < nn; ++v) off[v + 1] += off[v];
    {
        uint64_t *at = (uint64_t *)malloc(nn * sizeof(uint64_t));
        if (!at) {
            free(u), free(off), free(pred);
            return -1;
        }
        memcpy(at, off, nn * sizeof(uint64_t));
        for (uint64_t i = 0; i < nu; ++i) pred[at[u[i] & 0xffffffffull]++] = (uint32_t)(u[i] >> 32);
        free(at);
    }
    const int T = threads;
    const int cap_bodies = trace_cap;
    uint64_t *tcount = (uint64_t *)calloc((size_t)T * (cap_bodies > 0 ? cap_bodies : 1),
                                          sizeof(uint64_t));
    int *tbodies = (int *)calloc(T, sizeof(int));
    uint64_t *tsize = (uint64_t *)calloc(T, sizeof(uint64_t));
    int fail = !tcount || !tbodies || !tsize;
    const uint64_t nd0 = nu; /* |FULL0| raw rows seed the delta */
    const int limit = max_iteration < 0 ? 0x7fffffff : max_iteration + 1;
#pragma omp parallel num_threads(T) reduction(| : fail)
    {
        const int t = omp_get_thread_num();
        cset full;
        kvec d = {0, 0, 0}, nd = {0, 0, 0};
        if (cset_init(&full, 1024)) fail = 1;
        for (uint64_t i = 0; i < nu && !fail; ++i)
            if ((int)((u[i] & 0xffffffffull) % (uint64_t)T) == t && cset_insert(&full, u[i]) < 0)
                fail = 1;
        for (uint64_t i = 0; i < nd0 && !fail; ++i)
            if ((int)((edges[i] & 0xffffffffull) % (uint64_t)T) == t && kvec_push(&d, edges[i]))
                fail = 1;
        int body = 0;
        while (!fail && body < limit) {
            nd.n = 0;
            for (uint64_t i = 0; i < d.n && !fail; ++i) {
                const uint64_t b = d.v[i] >> 32, c = d.v[i] & 0xffffffffull;
                if (b >= nn) continue;
                for (uint64_t e = off[b]; e < off[b + 1]; ++e) {
                    const uint64_t k = ((uint64_t)pred[e] << 32) | c;
                    const int r = cset_insert(&full, k);
                    if (r < 0 || (r == 1 && kvec_push(&nd, k))) {
                        fail = 1;
                        break;
                    }
                }
            }
            if (body < cap_bodies) tcount[(size_t)t * cap_bodies + body] = nd.n;
            ++body;
            kvec sw = d;
            d = nd;
            nd = sw;
            if (d.n == 0) break;
        }
        tbodies[t] = body;
        tsize[t] = full.size + (uint64_t)full.has_max;
        free(full.slots);
        free(d.v);
        free(nd.v);
    }
    int ret = fail ? -1 : 0;
    if (!fail) {
        /* the global loop runs until every partition's delta is empty */
        int bodies = 0;
        for (int t = 0; t < T; ++t)
            if (tbodies[t] > bodies) bodies = tbodies[t];
        uint64_t total = 0, der = 0;
        for (int t = 0; t < T; ++t) total += tsize[t];
        for (int b = 0; b < bodies && b < cap_bodies; ++b) {
            uint64_t s = 0;
            for (int t = 0; t < T; ++t) s += tcount[(size_t)t * cap_bodies + b];
            delta_trace[b] = s;
            der += s;
        }
        *n_out = total;
        *derived = der;
        *iterations = bodies;
    }
    free(u), free(off), free(pred), free(tcount), free(tbodies), free(tsize);
    return ret;
}
