/*
 * gdl_gen.h -- seeded synthetic graph / points-to inputs (SplitMix64).
 * Edges are u64 keys (src << 32) | dst in generation order.
 */
#ifndef GDL_GEN_H
#define GDL_GEN_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Directed Erdos-Renyi: draw (u, v) uniformly until m distinct pairs with u != v. */
size_t gdl_gen_er(uint64_t seed, uint32_t n_nodes, size_t m, uint64_t *out);
/* child -> parent tree: node i (1..n-1) gets parent uniform in
 * [max(0, i - window), i - 1]; window 0 means [0, i - 1].  n - 1 edges. */
size_t gdl_gen_tree(uint64_t seed, uint32_t n_nodes, uint32_t window, uint64_t *out);
/* R-MAT: `scale` levels per edge, quadrant probabilities a, b, c (d = 1-a-b-c);
 * self-loops and duplicates rejected until m edges. */
size_t gdl_gen_rmat(uint64_t seed, int scale, size_t m, double a, double b, double c,
                    uint64_t *out);
/* Batched TC instance: edge count log-uniform in [min_edges, max_edges];
 * ER (nodes = ceil(m / 2.5)) or tree (nodes = m + 1) with probability 1/2. */
void gdl_gen_batched_shape(uint64_t seed, size_t min_edges, size_t max_edges, int *is_tree,
                           size_t *m, uint32_t *n_nodes);
size_t gdl_gen_batched_instance(uint64_t seed, size_t min_edges, size_t max_edges,
                                uint32_t tree_window, uint64_t *out);
/* CSPA facts: sources Zipf(s) over ranks 1..n_ids (rank r -> id r-1), targets
 * uniform; distinct assign pairs first, then distinct dereference pairs.
 * Returns the assign count; *n_deref_out gets the dereference count. */
size_t gdl_gen_cspa(uint64_t seed, uint32_t n_ids, double zipf_s, size_t n_assign,
                    size_t n_deref, uint64_t *assign_out, uint64_t *deref_out,
                    size_t *n_deref_out);

#ifdef __cplusplus
}
#endif
#endif
