// gdlog/hook_exec.h -- runs device hooks that are not column selections.
//
// The reference calls its tuple_generator_hook / tuple_copy_hook /
// tuple_predicate device functions once per produced tuple
// (relation.cu:221-373 for joins: the predicate on the generated tuple, then
// the generator into the output row; relation.cu:390-403 for copies).  Hooks
// that gdlog/hook_probe.h cannot turn into a column map or a column
// comparison -- arithmetic on columns, outputs of arity 3 or 4, predicates on
// constants -- run through the kernels below.  They are compiled into the
// translation unit that constructs the RA operator, the code object that also
// holds the hooks, so every indirect call stays inside one module.  Each
// kernel decodes its input keys into 8-byte columns (KeyLayout of the
// relation), calls the hooks, and packs the output under the output
// relation's layout; a value that does not fit its field is reported, never
// truncated.  The library reaches the kernels through the launch functions an
// operator records when it is constructed.
//
// This header must be compiled with hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gdlog/tuple.h"

namespace gdlog {

// Plain copy of a relation's KeyLayout, passed to kernels by value.
struct HookLayout {
    int arity;
    int bits[4];
    int shift[4];
};

// One fused pass of a join whose generator or predicate is a device hook:
// output position p (0 <= p < total) belongs to the outer tuple o with
// offsets[o] <= p < offsets[o + 1] and to the inner tuple begins[o] + p -
// offsets[o]; out[p] gets the packed generated tuple and keep[p] whether it
// passed the predicate.
struct HookJoinArgs {
    hipStream_t stream;
    tuple_generator_hook gen;
    tuple_predicate pred;  // may be null
    const uint64_t *inner;
    HookLayout inner_layout;
    const uint64_t *outer;
    HookLayout outer_layout;
    uint64_t n_outer;
    const uint64_t *offsets;  // n_outer + 1 entries
    const uint64_t *begins;   // n_outer entries
    uint64_t total;
    HookLayout out_layout;
    uint64_t *out;
    uint8_t *keep;
    // [0] OR of the kept keys' high halves, [1] of their low halves (bounds
    // for the radix sort), [2] non-zero when a value did not fit its field;
    // cleared by the caller
    unsigned long long *stats;
};

// A copy whose hook is not a column selection: out[i] = cp(src[i]) packed;
// keep[i] = 0 when a value did not fit (also flagged in stats[2]).
struct HookCopyArgs {
    hipStream_t stream;
    tuple_copy_hook cp;
    const uint64_t *src;
    HookLayout src_layout;
    uint64_t n;
    HookLayout out_layout;
    uint64_t *out;
    uint8_t *keep;
    unsigned long long *stats;
};

using HookJoinLaunch = hipError_t (*)(const HookJoinArgs *);
using HookCopyLaunch = hipError_t (*)(const HookCopyArgs *);

namespace hook_exec_detail {

constexpr int kW = 20;  // scratch tuple width of the reference (relation.cu:269,349)
constexpr int kB = 256;

__device__ inline void decode(uint64_t k, const HookLayout &L, column_type *c) {
    for (int i = 0; i < L.arity; ++i)
        c[i] = (k >> L.shift[i]) & ((L.bits[i] >= 64) ? ~0ull : ((1ull << L.bits[i]) - 1));
}

__device__ inline bool encode(const column_type *c, const HookLayout &L, uint64_t *k) {
    uint64_t v = 0;
    for (int i = 0; i < L.arity; ++i) {
        if (L.bits[i] < 64 && (c[i] >> L.bits[i]) != 0) return false;
        v |= (uint64_t)c[i] << L.shift[i];
    }
    *k = v;
    return true;
}

__device__ inline void publish(unsigned long long *stats, uint64_t hi, uint64_t lo, bool bad) {
    for (int d = 32; d >= 1; d >>= 1) {
        hi |= __shfl_xor(hi, d, 64);
        lo |= __shfl_xor(lo, d, 64);
    }
    const bool any_bad = __ballot(bad) != 0;
    if ((threadIdx.x & 63) == 0) {
        if (hi) atomicOr(&stats[0], (unsigned long long)hi);
        if (lo) atomicOr(&stats[1], (unsigned long long)lo);
        if (any_bad) atomicOr(&stats[2], 1ull);
    }
}

static __global__ __launch_bounds__(kB) void join_kernel(HookJoinArgs a) {
    column_type ic[kW], oc[kW], nc[kW];
    uint64_t hi = 0, lo = 0;
    bool bad = false;
    // every lane of a wave runs the same rounds (the shuffles in publish)
    for (uint64_t base = (uint64_t)blockIdx.x * kB; base < a.total;
         base += (uint64_t)gridDim.x * kB) {
        const uint64_t p = base + threadIdx.x;
        if (p >= a.total) continue;
        uint64_t l = 0, h = a.n_outer - 1;  // last o with offsets[o] <= p
        while (l < h) {
            const uint64_t mid = (l + h + 1) >> 1;
            if (a.offsets[mid] <= p) l = mid;
            else h = mid - 1;
        }
        const uint64_t j = a.begins[l] + (p - a.offsets[l]);
        for (int c = 0; c < kW; ++c) ic[c] = oc[c] = nc[c] = 0;
        decode(a.inner[j], a.inner_layout, ic);
        decode(a.outer[l], a.outer_layout, oc);
        (*a.gen)(ic, oc, nc);
        bool keep = a.pred ? (*a.pred)(nc) : true;
        uint64_t key = 0;
        if (keep && !encode(nc, a.out_layout, &key)) {
            bad = true;
            keep = false;
        }
        a.out[p] = key;
        a.keep[p] = keep ? 1 : 0;
        if (keep) {
            hi |= key >> 32;
            lo |= key & 0xffffffffull;
        }
    }
    publish(a.stats, hi, lo, bad);
}

static __global__ __launch_bounds__(kB) void copy_kernel(HookCopyArgs a) {
    column_type sc[kW], nc[kW];
    uint64_t hi = 0, lo = 0;
    bool bad = false;
    for (uint64_t base = (uint64_t)blockIdx.x * kB; base < a.n; base += (uint64_t)gridDim.x * kB) {
        const uint64_t i = base + threadIdx.x;
        if (i >= a.n) continue;
        for (int c = 0; c < kW; ++c) sc[c] = nc[c] = 0;
        decode(a.src[i], a.src_layout, sc);
        (*a.cp)(sc, nc);
        uint64_t key = 0;
        const bool ok = encode(nc, a.out_layout, &key);
        if (!ok) bad = true;
        a.out[i] = key;
        a.keep[i] = ok ? 1 : 0;
        if (ok) {
            hi |= key >> 32;
            lo |= key & 0xffffffffull;
        }
    }
    publish(a.stats, hi, lo, bad);
}

static inline unsigned grid_of(uint64_t n) {
    const uint64_t g = (n + kB - 1) / kB;
    return (unsigned)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

static hipError_t launch_join(const HookJoinArgs *a) {
    if (a->total == 0) return hipSuccess;
    hipLaunchKernelGGL(join_kernel, dim3(grid_of(a->total)), dim3(kB), 0, a->stream, *a);
    return hipGetLastError();
}

static hipError_t launch_copy(const HookCopyArgs *a) {
    if (a->n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(grid_of(a->n)), dim3(kB), 0, a->stream, *a);
    return hipGetLastError();
}

}  // namespace hook_exec_detail
}  // namespace gdlog
