// hashset.hip -- device hash set of tuple keys: the membership side of an
// append-only FULL version.
//
// Open addressing, linear probing, power-of-two capacity kept at load <= the
// context's load factor, at most 0.8 (grown by rehashing before an insert
// would pass it).  EMPTY = ~0; the key
// ~0 itself (ids 2^32 - 1 in both columns) is tracked by a flag word after
// the slots.  insert_new inserts a batch of keys (any order, repeats allowed)
// and emits each key that was not yet present exactly once: the thread whose
// CAS claims the slot emits it, wave-aggregated through one atomic counter.
#include <cmath>

#include "gdl_internal.hpp"

namespace gdl {

constexpr u64 kEmpty = ~0ull;
constexpr int HB = 256;

__device__ __forceinline__ u64 shfl64(u64 v, int src) {
    const int lo = __shfl((int)(u32)v, src, 64), hi = __shfl((int)(u32)(v >> 32), src, 64);
    return ((u64)(u32)hi << 32) | (u64)(u32)lo;
}

// Bitmap insert of one key per lane (valid lanes), aggregated over the wave:
// a run of consecutive lanes on one word ORs its bits with one atomic (read
// first: no atomic when they are all set), whose old value the run's lanes
// share; a lane whose bit an earlier lane of its run also sets is not new.
// Sorted keys put up to 64 lanes on one word, whose atomics would otherwise
// serialise.  Returns whether the lane's key is new.
__device__ __forceinline__ bool bm_insert_wave(u64 *words, int kb, bool valid, u64 k) {
    const int lane = threadIdx.x & 63;
    const bool dom = valid && bm_in_domain(k, kb);
    if (valid && !dom) atomicOr((unsigned long long *)&words[bm_flag_word(kb)], 1ull);
    const u64 b = dom ? bm_bit(k, kb) : 0;
    const u64 wi = dom ? (b >> 6) : ~0ull;
    const u64 m = dom ? (1ull << (b & 63)) : 0ull;
    const u64 pwi = shfl64(wi, lane > 0 ? lane - 1 : 0);
    const bool head = lane == 0 || pwi != wi;
    const u64 heads = __ballot(head);
    const u64 le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int rid = __popcll(heads & le);  // run number, 1-based
    // within runs: the OR of the lanes from this one on (the head's is the
    // run's) and of the lanes up to this one
    u64 acc = m, pre = m;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int up = lane + d < 64 ? lane + d : lane, dn = lane >= d ? lane - d : lane;
        const u64 x = shfl64(acc, up), y = shfl64(pre, dn);
        const int ru = __shfl(rid, up, 64), rd = __shfl(rid, dn, 64);
        if (lane + d < 64 && ru == rid) acc |= x;
        if (lane >= d && rd == rid) pre |= y;
    }
    const u64 pp = shfl64(pre, lane > 0 ? lane - 1 : 0);
    const bool dup = dom && !head && (pp & m) != 0;
    u64 old = 0;
    if (head && dom) {
        old = words[wi];
        if ((old & acc) != acc)
            old = atomicOr((unsigned long long *)&words[wi], (unsigned long long)acc);
    }
    const int h = 63 - __clzll((long long)(heads & le));  // this lane's run head
    old = shfl64(old, h);
    return dom && !dup && !(old & m);
}

// flag word at slots[cap]: 1 when the key ~0 is a member.  A block takes
// kHsPer * HB consecutive keys a round (thread t keys t, t + HB, ...) and
// reserves the round's output with one atomic: one counter atomic per wave,
// all on one address, serialised at a few ns each.
constexpr int kHsPer = 4;
__global__ __launch_bounds__(HB) void k_hs_insert_new(u64 *slots, u64 mask, const u64 *keys, u64 n,
                                                      u64 *out, unsigned long long *counter) {
    __shared__ u32 s_cnt[HB / 64 + 1];
    __shared__ unsigned long long s_base;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u64 lt = lane_mask_lt();
    const u64 per_round = (u64)HB * kHsPer;
    // every thread of a block runs the same rounds (barriers below)
    for (u64 base = (u64)blockIdx.x * per_round; base < n; base += (u64)gridDim.x * per_round) {
        u64 k[kHsPer], bal[kHsPer];
        bool fresh[kHsPer];
        u32 c = 0;
#pragma unroll
        for (int j = 0; j < kHsPer; ++j) {
            const u64 i = base + (u64)j * HB + threadIdx.x;
            k[j] = 0;
            fresh[j] = false;
            if (mask & kHsBitmap) {
                if (i < n) k[j] = keys[i];
                fresh[j] = bm_insert_wave(slots, (int)(mask & 63), i < n, k[j]);
            } else if (i < n) {
                k[j] = keys[i];
                fresh[j] = hs_insert_key(slots, mask, k[j]);
            }
            bal[j] = __ballot(fresh[j]);
            c += (u32)__popcll(bal[j]);
        }
        if (lane == 0) s_cnt[wv] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            u32 t = 0;
            for (int w = 0; w < HB / 64; ++w) {
                const u32 x = s_cnt[w];
                s_cnt[w] = t;
                t += x;
            }
            s_base = t ? atomicAdd(counter, (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        u64 pos = s_base + s_cnt[wv];
#pragma unroll
        for (int j = 0; j < kHsPer; ++j) {
            if (fresh[j]) out[pos + __popcll(bal[j] & lt)] = k[j];
            pos += (u64)__popcll(bal[j]);
        }
        __syncthreads();  // s_cnt and s_base are rewritten next round
    }
}

// Re-inserts every member of an old table into a new one.
__global__ __launch_bounds__(HB) void k_hs_rehash(const u64 *old_slots, u64 old_cap, u64 *slots,
                                                  u64 mask) {
    for (u64 i = (u64)blockIdx.x * HB + threadIdx.x; i < old_cap; i += (u64)gridDim.x * HB) {
        const u64 k = old_slots[i];
        if (k != kEmpty) (void)hs_insert_key(slots, mask, k);
    }
}

// Compacts the members of a table (the slots of cap words plus the ~0 flag
// word after them) into out, one atomic per wave.
__global__ __launch_bounds__(HB) void k_hs_extract(const u64 *slots, u64 cap, u64 *out,
                                                   u64 out_cap, unsigned long long *counter) {
    const u64 lt = lane_mask_lt();
    for (u64 base = (u64)blockIdx.x * HB; base <= cap; base += (u64)gridDim.x * HB) {
        const u64 i = base + threadIdx.x;
        bool keep = false;
        u64 k = 0;
        if (i < cap) {
            k = slots[i];
            keep = k != kEmpty;
        } else if (i == cap) {
            k = kEmpty;
            keep = slots[cap] != 0ull;
        }
        const u64 bal = __ballot(keep);
        if (bal) {
            unsigned long long pos = 0;
            const int leader = __ffsll((unsigned long long)bal) - 1;
            if ((int)(threadIdx.x & 63) == leader)
                pos = atomicAdd(counter, (unsigned long long)__popcll(bal));
            pos = __shfl(pos, leader, 64);
            const u64 at = pos + __popcll(bal & lt);
            if (keep && at < out_cap) out[at] = k;
        }
    }
}

__global__ __launch_bounds__(HB) void k_hs_fill(u64 *slots, u64 cap) {
    for (u64 i = (u64)blockIdx.x * HB + threadIdx.x; i < cap; i += (u64)gridDim.x * HB)
        slots[i] = i < cap - 1 ? kEmpty : 0ull;  // last word: the ~0 flag (cleared)
}

__global__ __launch_bounds__(HB) void k_bm_clear(u64 *words, u64 n) {
    for (u64 i = (u64)blockIdx.x * HB + threadIdx.x; i < n; i += (u64)gridDim.x * HB)
        words[i] = 0ull;
}

// Members of a bitmap set: each lane takes one word and emits the key of
// every set bit, one atomic per wave for the positions.
__global__ __launch_bounds__(HB) void k_bm_extract(const u64 *words, u64 n_words, int kb, u64 *out,
                                                   u64 out_cap, unsigned long long *counter) {
    const u64 lo_mask = (1ull << kb) - 1;
    for (u64 base = (u64)blockIdx.x * HB; base < n_words; base += (u64)gridDim.x * HB) {
        const u64 i = base + threadIdx.x;
        u64 w = i < n_words ? words[i] : 0ull;
        const u32 c = (u32)__popcll(w);
        // wave prefix sum of the per-lane bit counts
        u64 incl = c;
        for (int d = 1; d < 64; d <<= 1) {
            const u64 v = __shfl_up(incl, d, 64);
            if ((int)(threadIdx.x & 63) >= d) incl += v;
        }
        const u64 total = __shfl(incl, 63, 64);
        unsigned long long pos = 0;
        if ((threadIdx.x & 63) == 63 && total) pos = atomicAdd(counter, (unsigned long long)total);
        pos = __shfl(pos, 63, 64);
        u64 at = pos + incl - c;
        while (w) {
            const int b = __ffsll((unsigned long long)w) - 1;
            w &= w - 1;
            const u64 bit = i * 64 + (u64)b;
            if (at < out_cap) out[at] = ((bit >> kb) << 32) | (bit & lo_mask);
            ++at;
        }
    }
}

static unsigned hs_grid(gdl_ctx *ctx, u64 n) {
    const u64 cap = grid_cap(ctx, 2);
    u64 g = (n + HB - 1) / HB;
    return (unsigned)std::max<u64>(1, std::min(g, cap));
}

static gdl_status hs_alloc(gdl_ctx *ctx, u64 cap, u64 **slots) {
    void *p = nullptr;
    std::string err;
    gdl_status st = ctx->alloc->alloc((cap + 1) * sizeof(u64), &p, &err);
    if (st != GDL_OK) return fail(ctx, st, err);
    KTimer kt(ctx, K_FILL, 8.0 * (cap + 1));
    hipLaunchKernelGGL(k_hs_fill, dim3(hs_grid(ctx, cap + 1)), dim3(HB), 0, ctx->stream, (u64 *)p,
                       cap + 1);
    gdl_status ls = check_launch(ctx, "k_hs_fill");
    if (ls != GDL_OK) {
        ctx->alloc->free(p, nullptr);
        return ls;
    }
    *slots = (u64 *)p;
    return GDL_OK;
}

static u64 pow2_at_least(u64 x) {
    u64 c = 1024;
    while (c < x) c <<= 1;
    return c;
}

// Grows the table so that size + incoming stays at load <= the context's
// load factor.
gdl_status hashset_reserve(gdl_ctx *ctx, gdl_hashset *hs, u64 incoming) {
    return hashset_reserve2(ctx, hs, incoming, incoming);
}

// Grows the table so that size + expected new members stay at load <= the
// load factor (`load`, or the context's when 0; at most 0.8) and size + worst
// (an upper bound of the new members) at load <= kHsHardLoad: every insert
// still finds a free slot when the worst case happens, and the table is not
// sized by a loose bound.  A table that has to grow is made large enough for
// size + grow_to members at the load factor as well, so a set that keeps
// growing is rehashed every few multiples of its size, not at every doubling.
gdl_status hashset_reserve2(gdl_ctx *ctx, gdl_hashset *hs, u64 expected, u64 worst, double load,
                            u64 grow_to) {
    if (hs->key_bits) return GDL_OK;  // a bitmap holds its whole domain
    if (!(load > 0.0)) load = ctx->cfg.load_factor > 0.f ? (double)ctx->cfg.load_factor : 0.5;
    const double lf = std::min(0.8, load);
    expected = std::min(expected, worst);
    u64 need = pow2_at_least(
        std::max((u64)std::ceil((double)(hs->size + expected) / lf),
                 (u64)std::ceil((double)(hs->size + worst) / kHsHardLoad) + 1));
    if (hs->slots && hs->capacity >= need) return GDL_OK;
    if (grow_to) need = std::max(need, pow2_at_least((u64)std::ceil((double)(hs->size + grow_to) / lf)));
    u64 *ns = nullptr;
    GDL_TRY(hs_alloc(ctx, need, &ns));
    if (hs->slots) {
        KTimer kt(ctx, K_HS_REHASH, 16.0 * hs->capacity);
        hipLaunchKernelGGL(k_hs_rehash, dim3(hs_grid(ctx, hs->capacity)), dim3(HB), 0, ctx->stream,
                           (const u64 *)hs->slots, hs->capacity, ns, need - 1);
        GDL_TRY(check_launch(ctx, "k_hs_rehash"));
        // carry the ~0 flag word
        GDL_HIP(ctx, hipMemcpyAsync(ns + need, (u64 *)hs->slots + hs->capacity, sizeof(u64),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        ctx->alloc->free(hs->slots, nullptr);
    }
    hs->slots = ns;
    hs->capacity = need;
    return GDL_OK;
}

gdl_status hashset_insert_new(gdl_ctx *ctx, gdl_hashset *hs, const u64 *keys, u64 n, u64 *d_new,
                              u64 *n_new) {
    *n_new = 0;
    if (n == 0) return GDL_OK;
    GDL_TRY(hashset_reserve(ctx, hs, n));
    Scratch scr(ctx);
    u64 *counter;
    GDL_TRY(scr.get(1, &counter));
    GDL_HIP(ctx, hipMemsetAsync(counter, 0, sizeof(u64), ctx->stream));
    {
        KTimer kt(ctx, K_HS_INSERT, 24.0 * n);
        hipLaunchKernelGGL(k_hs_insert_new, dim3(hs_grid(ctx, (n + kHsPer - 1) / kHsPer)),
                           dim3(HB), 0, ctx->stream,
                           (u64 *)hs->slots, hs_arg(hs), keys, n, d_new,
                           (unsigned long long *)counter);
        GDL_TRY(check_launch(ctx, "k_hs_insert_new"));
    }
    GDL_TRY(read_u64(ctx, counter, n_new));
    hs->size += *n_new;
    return GDL_OK;
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_hashset_create(gdl_ctx *ctx, uint64_t expected, gdl_hashset *out) {
    if (!ctx || !out) return GDL_ERR_INVALID_ARG;
    out->slots = nullptr;
    out->capacity = 0;
    out->size = 0;
    out->key_bits = 0;
    out->pad_ = 0;
    return hashset_reserve(ctx, out, expected);
}

gdl_status gdl_hashset_create_bitmap(gdl_ctx *ctx, uint32_t key_bits, gdl_hashset *out) {
    if (!ctx || !out) return GDL_ERR_INVALID_ARG;
    if (key_bits < 1 || key_bits > 18)
        return fail(ctx, GDL_ERR_INVALID_ARG, "bitmap set: key_bits must be 1..18");
    out->slots = nullptr;
    out->capacity = 0;
    out->size = 0;
    out->key_bits = key_bits;
    out->pad_ = 0;
    const u64 bits = 1ull << (2 * key_bits);
    const u64 words = (bits + 63) / 64 + 1;  // + the out-of-domain flag word
    void *p = nullptr;
    std::string err;
    gdl_status st = ctx->alloc->alloc(words * sizeof(u64), &p, &err);
    if (st != GDL_OK) return fail(ctx, st, err);
    {
        KTimer kt(ctx, K_FILL, 8.0 * words);
        hipLaunchKernelGGL(k_bm_clear, dim3(hs_grid(ctx, words)), dim3(HB), 0, ctx->stream,
                           (u64 *)p, words);
        gdl_status ls = check_launch(ctx, "k_bm_clear");
        if (ls != GDL_OK) {
            ctx->alloc->free(p, nullptr);
            return ls;
        }
    }
    out->slots = p;
    out->capacity = bits;
    return GDL_OK;
}

gdl_status gdl_hashset_check(gdl_ctx *ctx, const gdl_hashset *hs) {
    if (!ctx || !hs) return GDL_ERR_INVALID_ARG;
    if (!hs->key_bits || !hs->slots) return GDL_OK;
    const u64 flag_at = hs->key_bits >= 3 ? (1ull << (2 * hs->key_bits - 6)) : 1ull;
    u64 flag = 0;
    GDL_TRY(read_u64(ctx, (const u64 *)hs->slots + flag_at, &flag));
    if (flag)
        return fail(ctx, GDL_ERR_OVERFLOW, "bitmap set: a key outside its " +
                                               std::to_string(hs->key_bits) +
                                               "-bit id domain was offered");
    return GDL_OK;
}

gdl_status gdl_hashset_free(gdl_ctx *ctx, gdl_hashset *hs) {
    if (!ctx || !hs) return GDL_ERR_INVALID_ARG;
    if (hs->slots) {
        std::string err;
        gdl_status st = ctx->alloc->free(hs->slots, &err);
        if (st != GDL_OK) return fail(ctx, st, err);
    }
    hs->slots = nullptr;
    hs->capacity = 0;
    hs->size = 0;
    hs->key_bits = 0;
    return GDL_OK;
}

gdl_status gdl_hashset_insert_new(gdl_ctx *ctx, gdl_hashset *hs, const uint64_t *d_keys,
                                  uint64_t n, uint64_t *d_new, uint64_t *n_new) {
    if (!ctx || !hs || !n_new || (n && (!d_keys || !d_new))) return GDL_ERR_INVALID_ARG;
    return hashset_insert_new(ctx, hs, d_keys, n, d_new, n_new);
}

gdl_status gdl_hashset_reserve(gdl_ctx *ctx, gdl_hashset *hs, uint64_t incoming) {
    if (!ctx || !hs) return GDL_ERR_INVALID_ARG;
    return hashset_reserve(ctx, hs, incoming);
}

gdl_status gdl_hashset_extract(gdl_ctx *ctx, const gdl_hashset *hs, uint64_t *d_out,
                               uint64_t out_cap, uint64_t *n_out) {
    if (!ctx || !hs || !n_out || (out_cap && !d_out)) return GDL_ERR_INVALID_ARG;
    *n_out = 0;
    if (!hs->slots || hs->capacity == 0) return GDL_OK;
    Scratch scr(ctx);
    u64 *counter;
    GDL_TRY(scr.get(1, &counter));
    GDL_HIP(ctx, hipMemsetAsync(counter, 0, sizeof(u64), ctx->stream));
    if (hs->key_bits) {
        const u64 words = (hs->capacity + 63) / 64;
        KTimer kt(ctx, K_HS_EXTRACT, 8.0 * words + 8.0 * hs->size);
        hipLaunchKernelGGL(k_bm_extract, dim3(hs_grid(ctx, words)), dim3(HB), 0, ctx->stream,
                           (const u64 *)hs->slots, words, (int)hs->key_bits, d_out, out_cap,
                           (unsigned long long *)counter);
        GDL_TRY(check_launch(ctx, "k_bm_extract"));
    } else {
        KTimer kt(ctx, K_HS_EXTRACT, 8.0 * hs->capacity + 8.0 * hs->size);
        hipLaunchKernelGGL(k_hs_extract, dim3(hs_grid(ctx, hs->capacity + 1)), dim3(HB), 0,
                           ctx->stream, (const u64 *)hs->slots, hs->capacity, d_out, out_cap,
                           (unsigned long long *)counter);
        GDL_TRY(check_launch(ctx, "k_hs_extract"));
    }
    GDL_TRY(read_u64(ctx, counter, n_out));
    if (*n_out > out_cap)
        return fail(ctx, GDL_ERR_OVERFLOW, "hash set extract: " + std::to_string(*n_out) +
                                               " members, output holds " +
                                               std::to_string(out_cap));
    return GDL_OK;
}

}  // extern "C"
