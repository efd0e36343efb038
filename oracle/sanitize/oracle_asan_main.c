/*
 * oracle_asan_main.c -- drives the CPU oracle and the CPU baselines over
 * small seeded inputs, for the AddressSanitizer / UndefinedBehaviorSanitizer
 * build (`make sanitize`).  Test infrastructure only.  Exits non-zero when a
 * result disagrees between the oracle and the baselines; the sanitizers abort
 * on any memory or undefined-behaviour error.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../gdl_oracle.h"

static uint64_t sm_state;
static uint64_t sm_next(void) {
    uint64_t z = (sm_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* m random pairs, then sorted and deduplicated (*m shrinks): with repeated
 * rows the reference's TC starts from the first |FULL| raw rows, which the
 * Souffle plan does not, so the comparison uses duplicate-free inputs. */
static uint64_t *random_edges(uint64_t seed, uint32_t nodes, size_t *m) {
    uint64_t *e = (uint64_t *)malloc(*m * sizeof(uint64_t));
    sm_state = seed;
    for (size_t i = 0; i < *m; ++i) {
        const uint64_t a = sm_next() % nodes, b = sm_next() % nodes;
        e[i] = (a << 32) | b;
    }
    *m = orc_sort_unique(e, *m);
    return e;
}

static int check(int ok, const char *what) {
    if (!ok) fprintf(stderr, "MISMATCH: %s\n", what);
    return ok ? 0 : 1;
}

int main(void) {
    int bad = 0;
    for (uint64_t seed = 0; seed < 6; ++seed) {
        const uint32_t nodes = 40 + 30 * (uint32_t)seed;
        size_t m = 60 + 50 * seed, md = m / 2;
        uint64_t *e = random_edges(seed, nodes, &m);
        orc_result tc, sg, cs;
        orc_tc(e, m, -1, &tc);
        orc_sg(e, m, &sg);
        uint64_t *d = random_edges(seed + 100, nodes, &md);
        orc_cspa(e, m / 2, d, md, &cs);
        for (int threads = 1; threads <= 3; threads += 2) {
            uint64_t n_out, der, cksum, trace[256];
            int it;
            if (cpu_tc_omp(e, m, threads, -1, &n_out, &der, &it, trace, 256) != 0) bad++;
            bad += check(n_out == tc.n_out && it == tc.iterations, "cpu_tc_omp vs orc_tc");
            if (souffle_tc(e, m, threads, -1, &n_out, &cksum, &it, trace, 256) != 0) bad++;
            bad += check(n_out == tc.n_out, "souffle_tc vs orc_tc");
            if (souffle_sg(e, m, threads, -1, &n_out, &cksum, &it, trace, 256) != 0) bad++;
            bad += check(n_out == sg.n_out, "souffle_sg vs orc_sg");
            uint64_t sz[3], cks[3];
            if (souffle_cspa(e, m / 2, d, md, threads, -1, sz, cks, &it) != 0) bad++;
            bad += check(sz[2] == cs.n_out && sz[0] == cs.sizes[0], "souffle_cspa vs orc_cspa");
        }
        orc_result_free(&tc);
        orc_result_free(&sg);
        orc_result_free(&cs);
        free(e);
        free(d);
    }
    /* empty inputs */
    {
        orc_result tc;
        orc_tc(NULL, 0, -1, &tc);
        bad += check(tc.n_out == 0, "empty TC");
        orc_result_free(&tc);
    }
    printf("oracle sanitizer run: %s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
