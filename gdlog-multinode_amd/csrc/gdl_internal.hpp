// gdl_internal.hpp -- context, allocator, statistics and wave64 device helpers
// shared by the engine's translation units.  Written for gfx950 (CDNA4,
// 64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "gdl_hip.h"

namespace gdl {

using u64 = uint64_t;
using u32 = uint32_t;

constexpr int kWave = 64;

// ---------------------------------------------------------------- allocator
// Size-classed caching allocator: classes are 512 B, then 8 steps per power
// of two (<= 12.5 % slack).  Freed blocks are kept for reuse; reuse is safe
// because every kernel of a context runs on its one stream.
class Allocator {
  public:
    explicit Allocator(int device) : device_(device) {}
    ~Allocator();
    gdl_status alloc(size_t bytes, void **out, std::string *err);
    gdl_status free(void *p, std::string *err);
    void release_cached();
    uint64_t in_use() const { return in_use_; }
    uint64_t peak() const { return peak_; }
    uint64_t cached() const { return cached_; }
    void set_cap(uint64_t cap) { cap_ = cap; }
    uint64_t cap() const { return cap_; }
    void set_cache_limit(uint64_t bytes) { cache_limit_ = bytes; }
    static size_t size_class(size_t bytes);

  private:
    int device_ [[maybe_unused]];
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> live_;
    uint64_t in_use_ = 0, peak_ = 0, cached_ = 0, cap_ = 0, cache_limit_ = 0;
};

// ------------------------------------------------------------------ stats
enum KernelId {
    K_RS_UPSWEEP = 0,
    K_RS_ROWSCAN,
    K_RS_DOWNSWEEP,
    K_RS_ONEBLOCK,
    K_KEY_BITS,
    K_SCAN_REDUCE,
    K_SCAN_TOP,
    K_SCAN_DOWN,
    K_COMPACT_COUNT,
    K_COMPACT_WRITE,
    K_MERGE_COUNT,
    K_MERGE_WRITE,
    K_IDX_COUNT,
    K_IDX_BUILD,
    K_IDX_LOOKUP,
    K_JOIN_COUNT,
    K_JOIN_MAT,
    K_PROJECT,
    K_FILL,
    K_HS_INSERT,
    K_HS_REHASH,
    K_JOIN_INSERT,
    K_HS_EXTRACT,
    K_SEG_HIST,
    K_MASK_COLS,
    K_LINEAR,
    K_LABEL,
    K_SCATTER,
    K_DENSE_SETUP,
    K_DENSE_EXPAND,
    K_DENSE_COMBINE,
    K_DENSE_APPEND,
    K_BM_EXTRACT,
    K_DENSE_TRANSPOSE,
    K_DENSE_ORROWS,
    K_JOIN_DENSE,
    K_NUM
};
extern const char *kKernelNames[K_NUM];

struct PendingTiming {
    int kid;
    hipEvent_t start, stop;
};

struct KernelStat {
    uint64_t launches = 0;
    double total_ms = 0.0;
    double alg_bytes = 0.0;
};

}  // namespace gdl

struct gdl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    gdl_config cfg{};
    gdl::Allocator *alloc = nullptr;
    uint64_t *h_pinned = nullptr;  // pinned staging for scalars read back
    char *h_stage = nullptr;       // pinned staging for bulk copies (gdl::Staging)
    size_t h_stage_cap = 0;
    uint32_t *d_err = nullptr;     // kernel-side error word
    int cu_count = 0;
    gdl_tuning tune{GDL_DEFAULT_LOAD_FACTOR, 8, 8, 256};  // load_factor mirrors cfg.load_factor
    std::string last_error;
    bool stats_on = false;
    int id_bits = 0;  // gdl_set_id_bits: bound of every id's bit length, 0 = unknown
    bool capturing = false;  // between gdl_graph_begin and gdl_graph_end
    int dense_mode = 1;      // gdl_set_dense_mode
    gdl::KernelStat stats[gdl::K_NUM];
    std::vector<gdl::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    // Fixed scratch arena: while set, Scratch carves from it instead of the
    // allocator (the allocation-free entry points set it for one call).
    char *arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
};

namespace gdl {

// A comparison op (optionally against a constant) over columns < arity.
inline bool valid_pred(const gdl_pred &p, int arity) {
    const int op = p.op & ~GDL_PRED_CONST;
    if (op < GDL_PRED_NONE || op > GDL_PRED_FALSE) return false;
    if (op == GDL_PRED_NONE || op == GDL_PRED_FALSE) return !(p.op & GDL_PRED_CONST);
    if (p.lhs < 0 || p.lhs >= arity) return false;
    return (p.op & GDL_PRED_CONST) || (p.rhs >= 0 && p.rhs < arity);
}

// Grid cap of a grid-stride kernel: the tuned blocks per CU (x mult).
inline uint64_t grid_cap(const gdl_ctx *ctx, uint64_t mult = 1) {
    return (uint64_t)(ctx->cu_count > 0 ? ctx->cu_count : 256) * ctx->tune.blocks_per_cu * mult;
}

// Records the status + message on the context and returns the status.
gdl_status fail(gdl_ctx *ctx, gdl_status st, const std::string &msg);
gdl_status hip_fail(gdl_ctx *ctx, hipError_t e, const char *what);

#define GDL_HIP(ctx, call)                                         \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::gdl::hip_fail(ctx, _e, #call); \
    } while (0)

#define GDL_TRY(expr)                         \
    do {                                      \
        gdl_status _s = (expr);               \
        if (_s != GDL_OK) return _s;          \
    } while (0)

// Device scratch owned by the context allocator, freed on scope exit; or,
// while the context has a fixed arena, 256-byte aligned pieces of the arena
// (bump allocation, released when the arena is cleared).
class Scratch {
  public:
    Scratch(gdl_ctx *ctx) : ctx_(ctx) {}
    ~Scratch() {
        for (void *p : ptrs_) ctx_->alloc->free(p, nullptr);
    }
    template <typename T>
    gdl_status get(size_t count, T **out) {
        void *p = nullptr;
        if (ctx_->arena) {
            const size_t at = (ctx_->arena_used + 255) & ~size_t(255);
            const size_t bytes = count * sizeof(T) + 16;
            if (at > ctx_->arena_size || bytes > ctx_->arena_size - at)
                return fail(ctx_, GDL_ERR_OUT_OF_MEMORY, "scratch arena exhausted");
            ctx_->arena_used = at + bytes;
            *out = reinterpret_cast<T *>(ctx_->arena + at);
            return GDL_OK;
        }
        std::string err;
        gdl_status st = ctx_->alloc->alloc(count * sizeof(T) + 16, &p, &err);
        if (st != GDL_OK) return fail(ctx_, st, err);
        ptrs_.push_back(p);
        *out = static_cast<T *>(p);
        return GDL_OK;
    }

  private:
    gdl_ctx *ctx_;
    std::vector<void *> ptrs_;
};

// Kernel timing scope: with statistics on, brackets a launch with hipEvents
// on the context stream; always counts launches and algorithmic bytes.
class KTimer {
  public:
    KTimer(gdl_ctx *ctx, int kid, double alg_bytes);
    ~KTimer();

  private:
    gdl_ctx *ctx_;
    int kid_;
    hipEvent_t start_ = nullptr;
};

gdl_status read_u64(gdl_ctx *ctx, const u64 *d, u64 *h, int count = 1);
// Looks up this device's fixed scratch array (once per device; the device
// must be current).  Called by gdl_ctx_create.
gdl_status noalloc_setup(gdl_ctx *ctx);
gdl_status check_launch(gdl_ctx *ctx, const char *what);
gdl_status check_device_error(gdl_ctx *ctx);
// The status for a device error word read with other values (resets it).
gdl_status device_error_status(gdl_ctx *ctx, u32 err);

// Bulk host <-> device copies through the context's pinned staging buffer: a
// copy between the device and pageable host memory makes the runtime pin or
// stage that memory on every call (hundreds of microseconds for a few
// hundred KB).  up() copies into the buffer and sends it asynchronously on
// the context's stream (the caller's host data may change as soon as it
// returns); down() waits for the stream.  The buffer is reused after each
// wait; the destructor waits for uploads still in flight.  One thread per
// context, as for every call on it.
class Staging {
  public:
    explicit Staging(gdl_ctx *ctx) : ctx_(ctx) {}
    ~Staging();
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;
    gdl_status up(void *d_dst, const void *h_src, size_t n);
    gdl_status down(void *h_dst, const void *d_src, size_t n);
    gdl_status sync();
    // A copy to the host that lands in h_dst at the next land(), land_at()
    // or sync(): several reads for one wait.
    gdl_status later(void *h_dst, const void *d_src, size_t n);
    gdl_status land();
    // Waits for `ev` only (recorded after the later() copies): the stream
    // keeps running the work enqueued after it.
    gdl_status land_at(hipEvent_t ev);

  private:
    struct Pending {
        void *h_dst;
        size_t at, n;
    };
    gdl_status room(size_t n, char **at);
    void land_pending();
    gdl_ctx *ctx_;
    size_t used_ = 0;
    std::vector<Pending> pending_;
};

// ---- internal entry points shared between translation units
gdl_status exclusive_scan_u32_to_u64(gdl_ctx *ctx, const u32 *d_in, u64 n, u64 *d_out,
                                     u64 *total_host /* may be null: no sync */);
gdl_status exclusive_scan_u64(gdl_ctx *ctx, const u64 *d_in, u64 n, u64 *d_out);
// In-place exclusive scan of v[0..n); v[n] receives the total (v holds n + 1).
gdl_status scan_inplace(gdl_ctx *ctx, u64 *v, u64 n);
gdl_status sort_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, int hi_bits, int lo_bits, u64 *d_out);
gdl_status unique_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, u64 *d_out, u64 *n_out);
gdl_status sort_unique_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, int hi_bits, int lo_bits,
                            u64 *d_out, u64 *n_out);
gdl_status filter_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, gdl_pred pred, u64 *d_out,
                       u64 *n_out);
gdl_status hashset_insert_new(gdl_ctx *ctx, gdl_hashset *hs, const u64 *keys, u64 n, u64 *d_new,
                              u64 *n_new);
// Grows hs so that `incoming` more keys keep its load at or below the
// context's load factor.
gdl_status hashset_reserve(gdl_ctx *ctx, gdl_hashset *hs, u64 incoming);
constexpr double kHsHardLoad = 0.9;
gdl_status hashset_reserve2(gdl_ctx *ctx, gdl_hashset *hs, u64 expected, u64 worst,
                            double load = 0.0, u64 grow_to = 0);
gdl_status index_build(gdl_ctx *ctx, const u64 *keys, u64 n, int prefix_cols, gdl_index *out);
gdl_status join_count(gdl_ctx *ctx, const u64 *inner, u64 n_inner, const gdl_index *idx,
                      const u64 *outer, u64 n_outer, u32 *counts, u64 *begins);
// Per-instance hash sets of a segmented fused join (one per batch slot).
constexpr u32 kMaxSegs = 256;
// Bodies on bit matrices (dense.hip): the smallest delta for which the
// context's dense mode runs them (~0: never), and the run itself
// (gdl_fixpoint_dense's contract).
u64 dense_threshold_for(const gdl_ctx *ctx, const gdl_hashset *hs, int shape, u64 n1, u64 n2);
// Whether a join into the bitmap set hs producing `produced` tuples would run
// on bit matrices (the conditions of dense_join_insert).
bool dense_join_pays(const gdl_ctx *ctx, const gdl_hashset *hs, gdl_colmap map, gdl_pred pred,
                     int prefix_cols, u64 n_inner, u64 n_outer, u64 produced);
// A join whose exact produced count is `produced`, into the bitmap set hs,
// run on bit matrices when that is much cheaper than producing its tuples
// one by one: returns false (nothing done) when the join does not qualify.
// Otherwise *status is the outcome, with gdl_join_insert's results (new keys
// appended to *d_new in key order, *n_new, *n_raw = produced).
bool dense_join_insert(gdl_ctx *ctx, const u64 *d_inner, u64 n_inner, const u64 *d_outer,
                       u64 n_outer, gdl_colmap map, gdl_pred pred, int prefix_cols,
                       gdl_hashset *hs, u64 produced, u64 **d_new, u64 *new_cap, u64 *n_new,
                       u64 *n_raw, gdl_status *status);
gdl_status dense_run(gdl_ctx *ctx, const gdl_dense_rule *rule, gdl_hashset *hs,
                     const u64 *d_delta, u64 n_delta, u64 **d_log, u64 *log_cap, u64 *log_end,
                     int max_bodies, u64 *h_new, u64 *h_raw, int *bodies);
// Exclusive scans of two arrays in place in one launch (v[n] = total each).
gdl_status scan_inplace2(gdl_ctx *ctx, u64 *v0, u64 *v1, u64 n, u64 *totals /* 2 */);

inline unsigned grid_for(u64 n, u64 per_block) {
    u64 g = (n + per_block - 1) / per_block;
    return (unsigned)(g ? g : 1);
}

}  // namespace gdl

// ----------------------------------------------------------- device helpers
namespace gdl {

// Orders a wave's LDS accesses: what one lane wrote is visible to the others
// after it.
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ u64 lane_mask_lt() {
    return (1ull << (threadIdx.x & 63)) - 1ull;
}

// Inclusive scan of a u64 across the 64 lanes of a wave.
__device__ __forceinline__ u64 wave_inclusive_scan(u64 v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        u64 o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Block-wide exclusive scan of one u64 per thread.  `s_wave` holds
// BLOCK/64 + 1 u64.  Returns the exclusive prefix; *total gets the block sum.
template <int BLOCK>
__device__ __forceinline__ u64 block_exclusive_scan(u64 v, u64 *s_wave, u64 *total) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    u64 inc = wave_inclusive_scan(v);
    if (lane == 63) s_wave[w] = inc;
    __syncthreads();
    if (w == 0) {
        u64 x = (lane < NW) ? s_wave[lane] : 0;
        u64 xi = wave_inclusive_scan(x);
        if (lane < NW) s_wave[lane] = xi - x;
        if (lane == NW - 1) s_wave[NW] = xi;
    }
    __syncthreads();
    u64 res = s_wave[w] + inc - v;
    *total = s_wave[NW];
    __syncthreads();
    return res;
}

__device__ __forceinline__ u32 key_col(u64 k, int c) {
    return c ? (u32)(k & 0xffffffffull) : (u32)(k >> 32);
}

// Column map: src in {0: inner c0, 1: inner c1, 2: outer c0, 3: outer c1}.
__device__ __forceinline__ u32 map_col(u64 ik, u64 ok, int src) {
    u64 v = (src & 2) ? ok : ik;
    return (src & 1) ? (u32)(v & 0xffffffffull) : (u32)(v >> 32);
}

__device__ __forceinline__ u64 apply_map(u64 ik, u64 ok, int arity, int s0, int s1) {
    u64 hi = (u64)map_col(ik, ok, s0) << 32;
    return arity == 2 ? (hi | (u64)map_col(ik, ok, s1)) : hi;
}

__device__ __forceinline__ bool eval_pred(u64 k, int op, int lhs, int rhs) {
    const u32 a = key_col(k, lhs);
    const u32 b = (op & GDL_PRED_CONST) ? (u32)rhs : key_col(k, rhs);
    switch (op & ~GDL_PRED_CONST) {
        case GDL_PRED_EQ: return a == b;
        case GDL_PRED_NE: return a != b;
        case GDL_PRED_LT: return a < b;
        case GDL_PRED_LE: return a <= b;
        case GDL_PRED_GT: return a > b;
        case GDL_PRED_GE: return a >= b;
        case GDL_PRED_FALSE: return false;
        default: return true;
    }
}

// 64-bit finalizer (MurmurHash3 fmix64) for the prefix hash index.
__host__ __device__ __forceinline__ u64 mix64(u64 k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ u64 mulhi64(u64 a, u64 b) { return __umul64hi(a, b); }

struct alignas(16) IdxSlot {
    u64 key;  // prefix
    u64 val;  // (begin << 24) | min(count, 0xFFFFFF); 0 = empty
};
constexpr u64 kCountBits = 24;
constexpr u64 kCountSat = (1ull << kCountBits) - 1;

// First index >= from whose key prefix differs from p (keys sorted).
__device__ __forceinline__ u64 run_end(const u64 *keys, u64 n, u64 from, u64 p, int shift) {
    u64 lo = from, step = 1, hi = from + 1;
    while (hi < n && (keys[hi] >> shift) == p) {
        lo = hi;
        step <<= 1;
        hi = from + step;
    }
    if (hi > n) hi = n;
    // keys[lo] has prefix p; first mismatch in (lo, hi]
    u64 a = lo + 1, b = hi;
    while (a < b) {
        u64 mid = a + ((b - a) >> 1);
        if ((keys[mid] >> shift) == p) a = mid + 1;
        else b = mid;
    }
    return a;
}

// Hash set of keys: open addressing over a power-of-two table, EMPTY = ~0,
// the key ~0 tracked by the flag word slots[mask + 1].  Returns true when k
// was not a member (this call inserted it).
constexpr u64 kHsEmpty = ~0ull;

constexpr int kHsScan = 8;  // slots read together while probing (one 64-byte line)

// Probe from slot s: read the next kHsScan slots with independent plain loads
// (one round trip), then CAS only the first that read EMPTY, so a run of
// occupied slots costs about two round trips instead of one atomic per slot.
// Slots are written once, so a read can only be stale by showing EMPTY for a
// slot filled since; the CAS catches that and the probe goes on after it.
__device__ __forceinline__ bool hs_probe_from(u64 *slots, u64 mask, u64 k, u64 s) {
    for (u64 seen = 0; seen <= mask;) {
        u64 v[kHsScan];
#pragma unroll
        for (int j = 0; j < kHsScan; ++j) v[j] = slots[(s + j) & mask];
        int pos = kHsScan;
        bool is_k = false;
#pragma unroll
        for (int j = 0; j < kHsScan; ++j) {
            if (pos == kHsScan && (v[j] == k || v[j] == kHsEmpty)) {
                pos = j;
                is_k = v[j] == k;
            }
        }
        if (is_k) return false;
        if (pos == kHsScan) {
            s = (s + kHsScan) & mask;
            seen += kHsScan;
            continue;
        }
        const u64 t = (s + pos) & mask;
        const unsigned long long old = atomicCAS(
            (unsigned long long *)&slots[t], (unsigned long long)kHsEmpty, (unsigned long long)k);
        if (old == (unsigned long long)kHsEmpty) return true;
        if (old == (unsigned long long)k) return false;
        s = (t + 1) & mask;
        seen += (u64)pos + 1;
    }
    return false;  // unreachable: the table is never full
}

// A set argument with this bit set is a bitmap over a dense key domain: the
// low 6 bits are key_bits k, bit (c0 << k) | c1 of the word array stands for
// the key (c0 << 32) | c1, and the word after the 2^(2k) bits is a flag set
// when a key outside the domain was offered (it is not inserted).
constexpr u64 kHsBitmap = 1ull << 63;

__device__ __forceinline__ u64 bm_bit(u64 k, int kb) {
    return ((k >> 32) << kb) | (k & 0xffffffffull);
}

__device__ __forceinline__ bool bm_in_domain(u64 k, int kb) {
    return ((k >> 32) >> kb) == 0 && ((k & 0xffffffffull) >> kb) == 0;
}

__device__ __forceinline__ u64 bm_flag_word(int kb) {
    return kb >= 3 ? (1ull << (2 * kb - 6)) : 1ull;
}

// Inserts k into a bitmap set; true when k was not a member.  Bits are only
// ever set while a set lives, so a plain load that shows the bit set proves
// membership without an atomic (a load showing it clear may be stale; the
// atomic OR then decides).  Where most produced keys are already members
// (SG: ~20 produced per new key) this keeps most of the join's traffic in
// loads, which caches serve, instead of atomics, which leave L2.
__device__ __forceinline__ bool bm_insert_key(u64 *words, int kb, u64 k) {
    if (!bm_in_domain(k, kb)) {
        atomicOr((unsigned long long *)&words[bm_flag_word(kb)], 1ull);
        return false;
    }
    const u64 b = bm_bit(k, kb);
    const u64 m = 1ull << (b & 63);
    if (words[b >> 6] & m) return false;
    return (atomicOr((unsigned long long *)&words[b >> 6], (unsigned long long)m) & m) == 0;
}

// CAS the home slot first: one round trip decides most inserts (reading
// first instead costs a second round trip for every key that is new, and
// was slower even where most keys are already members); past the home slot
// the probe reads kHsScan slots at a time.
__device__ __forceinline__ bool hs_insert_key(u64 *slots, u64 mask, u64 k) {
    if (mask & kHsBitmap) return bm_insert_key(slots, (int)(mask & 63), k);
    if (k == kHsEmpty) return atomicExch((unsigned int *)&slots[mask + 1], 1u) == 0u;
    const u64 s = mix64(k) & mask;
    const unsigned long long old = atomicCAS(
        (unsigned long long *)&slots[s], (unsigned long long)kHsEmpty, (unsigned long long)k);
    if (old == (unsigned long long)kHsEmpty) return true;
    if (old == (unsigned long long)k) return false;
    return hs_probe_from(slots, mask, k, (s + 1) & mask);
}

// The index is an array of buckets of kBucket slots (one 64-byte line);
// a key lives in the first bucket of its probe sequence that had a free slot
// when it was inserted, so a bucket with a free slot ends every probe.
constexpr u64 kBucket = 4;

// Kernels get an index's n_slots with this bit set when it is direct-mapped:
// slots is then an array of run words ((begin << 24) | count, 0 = no run)
// indexed by the prefix, one load per lookup whether the prefix is present
// or not (a hash probe for an absent prefix walks to a bucket with room).
constexpr u64 kIdxDense = 1ull << 63;
inline u64 idx_arg(const gdl_index *idx) { return idx->n_slots | (idx->dense ? kIdxDense : 0); }

// The set argument of a kernel: capacity - 1 of a hash table, or the bitmap
// flag with key_bits.
inline u64 hs_arg(const gdl_hashset *hs) {
    return hs->key_bits ? (kHsBitmap | hs->key_bits) : hs->capacity - 1;
}

__device__ __forceinline__ bool idx_probe(const IdxSlot *slots, u64 n_slots, u64 p, u64 *begin,
                                          u64 *count) {
    if (n_slots & kIdxDense) {
        if (p >= (n_slots & ~kIdxDense)) return false;
        const u64 v = ((const u64 *)slots)[p];
        if (v == 0) return false;
        *begin = v >> kCountBits;
        *count = v & kCountSat;
        return true;
    }
    const u64 nbk = n_slots / kBucket;
    u64 b = mulhi64(mix64(p), nbk);
    for (u64 step = 0; step < nbk; ++step) {
        const IdxSlot *bk = slots + b * kBucket;
        bool any_empty = false;
#pragma unroll
        for (u64 j = 0; j < kBucket; ++j) {
            const IdxSlot e = bk[j];
            if (e.val == 0) {
                any_empty = true;
            } else if (e.key == p) {
                *begin = e.val >> kCountBits;
                *count = e.val & kCountSat;
                return true;
            }
        }
        if (any_empty) return false;
        if (++b == nbk) b = 0;
    }
    return false;
}

}  // namespace gdl
