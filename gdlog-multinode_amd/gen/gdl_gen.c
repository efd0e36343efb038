/*
 * gdl_gen.c -- seeded synthetic inputs for the TC / SG / CSPA workloads.
 *
 * PRNG: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014).  A uniform integer in [0, n) is the high 64 bits
 * of r * n; a uniform double is (r >> 11) * 2^-53.
 *
 * Every generator returns edges as u64 keys (src << 32) | dst in generation
 * order (the order a TSV file of the same edges would list them).
 */
#include "gdl_gen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t uniform_below(uint64_t *s, uint64_t n) {
    return (uint64_t)(((unsigned __int128)sm64(s) * n) >> 64);
}

static inline double uniform01(uint64_t *s) { return (double)(sm64(s) >> 11) * 0x1.0p-53; }

/* open-addressing set of u64 keys (never holds ~0) */
typedef struct {
    uint64_t *slots;
    size_t mask;
} keyset;

static int keyset_init(keyset *ks, size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    ks->slots = (uint64_t *)malloc(cap * sizeof(uint64_t));
    if (!ks->slots) return -1;
    memset(ks->slots, 0xff, cap * sizeof(uint64_t));
    ks->mask = cap - 1;
    return 0;
}

static inline uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

/* returns 1 if inserted, 0 if present */
static int keyset_insert(keyset *ks, uint64_t k) {
    size_t i = mix64(k) & ks->mask;
    for (;;) {
        if (ks->slots[i] == ~0ull) {
            ks->slots[i] = k;
            return 1;
        }
        if (ks->slots[i] == k) return 0;
        i = (i + 1) & ks->mask;
    }
}

static void keyset_free(keyset *ks) { free(ks->slots); }

size_t gdl_gen_er(uint64_t seed, uint32_t n_nodes, size_t m, uint64_t *out) {
    if (n_nodes < 2) return 0;
    uint64_t maxm = (uint64_t)n_nodes * (n_nodes - 1);
    if (m > maxm) m = (size_t)maxm;
    uint64_t s = seed;
    keyset ks;
    if (keyset_init(&ks, m)) return 0;
    size_t w = 0;
    while (w < m) {
        uint64_t u = uniform_below(&s, n_nodes);
        uint64_t v = uniform_below(&s, n_nodes);
        if (u == v) continue;
        uint64_t k = (u << 32) | v;
        if (keyset_insert(&ks, k)) out[w++] = k;
    }
    keyset_free(&ks);
    return w;
}

static size_t gen_tree_stream(uint64_t *s, uint32_t n_nodes, uint32_t window, uint64_t *out) {
    size_t w = 0;
    for (uint64_t i = 1; i < n_nodes; i++) {
        uint64_t lo = (window && i > window) ? i - window : 0;
        uint64_t parent = lo + uniform_below(s, i - lo);
        out[w++] = (i << 32) | parent;
    }
    return w;
}

size_t gdl_gen_tree(uint64_t seed, uint32_t n_nodes, uint32_t window, uint64_t *out) {
    uint64_t s = seed;
    return gen_tree_stream(&s, n_nodes, window, out);
}

size_t gdl_gen_rmat(uint64_t seed, int scale, size_t m, double a, double b, double c,
                    uint64_t *out) {
    uint64_t s = seed;
    keyset ks;
    if (keyset_init(&ks, m)) return 0;
    double ab = a + b, abc = a + b + c;
    size_t w = 0;
    uint64_t attempts = 0, max_attempts = 1000ull * (m + 1000);
    while (w < m && attempts++ < max_attempts) {
        uint64_t u = 0, v = 0;
        for (int l = 0; l < scale; l++) {
            double r = uniform01(&s);
            uint64_t bu, bv;
            if (r < a) { bu = 0; bv = 0; }
            else if (r < ab) { bu = 0; bv = 1; }
            else if (r < abc) { bu = 1; bv = 0; }
            else { bu = 1; bv = 1; }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u == v) continue;
        uint64_t k = (u << 32) | v;
        if (keyset_insert(&ks, k)) out[w++] = k;
    }
    keyset_free(&ks);
    return w;
}

void gdl_gen_batched_shape(uint64_t seed, size_t min_edges, size_t max_edges, int *is_tree,
                           size_t *m, uint32_t *n_nodes) {
    uint64_t s = seed;
    double lmin = log((double)min_edges), lmax = log((double)max_edges);
    double u = uniform01(&s);
    size_t mm = (size_t)llround(exp(lmin + u * (lmax - lmin)));
    if (mm < min_edges) mm = min_edges;
    if (mm > max_edges) mm = max_edges;
    int tree = uniform01(&s) >= 0.5;
    *is_tree = tree;
    *m = mm;
    *n_nodes = tree ? (uint32_t)(mm + 1) : (uint32_t)((mm * 2 + 4) / 5); /* ceil(m / 2.5) */
}

size_t gdl_gen_batched_instance(uint64_t seed, size_t min_edges, size_t max_edges,
                                uint32_t tree_window, uint64_t *out) {
    int tree;
    size_t m;
    uint32_t n;
    gdl_gen_batched_shape(seed, min_edges, max_edges, &tree, &m, &n);
    /* edges come from a stream seeded independently of the shape draws */
    uint64_t edge_seed = seed ^ 0xD1B54A32D192ED03ull;
    if (tree) return gdl_gen_tree(edge_seed, n, tree_window, out);
    return gdl_gen_er(edge_seed, n, m, out);
}

size_t gdl_gen_cspa(uint64_t seed, uint32_t n_ids, double zipf_s, size_t n_assign,
                    size_t n_deref, uint64_t *assign_out, uint64_t *deref_out,
                    size_t *n_deref_out) {
    uint64_t s = seed;
    double *cdf = (double *)malloc((size_t)n_ids * sizeof(double));
    if (!cdf) return 0;
    double acc = 0.0;
    for (uint32_t r = 1; r <= n_ids; r++) {
        acc += 1.0 / pow((double)r, zipf_s);
        cdf[r - 1] = acc;
    }
    for (uint32_t r = 0; r < n_ids; r++) cdf[r] /= acc;
    keyset ka, kd;
    if (keyset_init(&ka, n_assign) || keyset_init(&kd, n_deref)) {
        free(cdf);
        return 0;
    }
    size_t counts[2] = {0, 0};
    size_t want[2] = {n_assign, n_deref};
    uint64_t *outs[2] = {assign_out, deref_out};
    keyset *sets[2] = {&ka, &kd};
    for (int rel = 0; rel < 2; rel++) {
        uint64_t attempts = 0, max_attempts = 1000ull * (want[rel] + 1000);
        while (counts[rel] < want[rel] && attempts++ < max_attempts) {
            double u = uniform01(&s);
            /* inverse CDF: smallest rank r with cdf[r-1] >= u */
            size_t lo = 0, hi = n_ids - 1;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (cdf[mid] < u) lo = mid + 1;
                else hi = mid;
            }
            uint64_t src = lo; /* rank lo+1 -> id lo */
            uint64_t dst = uniform_below(&s, n_ids);
            uint64_t k = (src << 32) | dst;
            if (keyset_insert(sets[rel], k)) outs[rel][counts[rel]++] = k;
        }
    }
    keyset_free(&ka);
    keyset_free(&kd);
    free(cdf);
    if (n_deref_out) *n_deref_out = counts[1];
    return counts[0];
}
