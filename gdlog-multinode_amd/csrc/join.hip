// join.hip -- prefix hash index, hash join (count, scan, materialize),
// projection.
//
// Index: open addressing with linear probing over the distinct join prefixes
// of a sorted key array; each slot maps a prefix to its run [begin,
// begin + count).  Slot = 16 bytes, probed with one 128-bit load.  The run
// length is stored saturated at 2^24 - 1 and recomputed by galloping search
// when saturated.
//
// Materialize is load-balanced: each thread produces 8 consecutive outputs,
// finds its outer tuple by binary search over the output offsets, walks the
// inner run, and the block stores its 2048 outputs from LDS so stores are
// contiguous across lanes.
#include "gdl_internal.hpp"

namespace gdl {

constexpr int JB = 256;
constexpr int JM_I = 8;
constexpr int JM_TILE = JB * JM_I;

__device__ __forceinline__ int prefix_shift(int prefix_cols) { return prefix_cols >= 2 ? 0 : 32; }

// out[0] += distinct prefixes, out[1] = max(longest run); one atomic pair per
// block; out[2] = the largest prefix (the last key's: keys are sorted).
__global__ __launch_bounds__(JB) void k_idx_count(const u64 *keys, u64 n, int shift, u64 *out) {
    __shared__ u64 s_sum[JB / 64], s_max[JB / 64];
    if (blockIdx.x == 0 && threadIdx.x == 0) out[2] = keys[n - 1] >> shift;
    u64 acc = 0, mx = 0;
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 p = keys[i] >> shift;
        if (i == 0 || (keys[i - 1] >> shift) != p) {
            ++acc;
            const u64 len = run_end(keys, n, i, p, shift) - i;
            mx = len > mx ? len : mx;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        acc += __shfl_xor(acc, d, 64);
        const u64 o = __shfl_xor(mx, d, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        s_sum[threadIdx.x >> 6] = acc;
        s_max[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 a = 0, m = 0;
        for (int w = 0; w < JB / 64; ++w) {
            a += s_sum[w];
            m = s_max[w] > m ? s_max[w] : m;
        }
        if (a) atomicAdd((unsigned long long *)&out[0], (unsigned long long)a);
        if (m) atomicMax((unsigned long long *)&out[1], (unsigned long long)m);
    }
}

__global__ __launch_bounds__(JB) void k_idx_build(const u64 *keys, u64 n, int shift,
                                                  IdxSlot *slots, u64 n_slots, u32 *err) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 p = keys[i] >> shift;
        if (i != 0 && (keys[i - 1] >> shift) == p) continue;
        const u64 end = run_end(keys, n, i, p, shift);
        const u64 cnt = end - i;
        const u64 val = (i << kCountBits) | (cnt < kCountSat ? cnt : kCountSat);
        const u64 nbk = n_slots / kBucket;
        u64 b = mulhi64(mix64(p), nbk);
        bool placed = false;
        for (u64 step = 0; step < nbk && !placed; ++step) {
            for (u64 j = 0; j < kBucket; ++j) {
                IdxSlot *e = slots + b * kBucket + j;
                if (e->val != 0) continue;
                const unsigned long long old =
                    atomicCAS((unsigned long long *)&e->val, 0ull, (unsigned long long)val);
                if (old == 0ull) {
                    e->key = p;
                    placed = true;
                    break;
                }
            }
            if (++b == nbk) b = 0;
        }
        if (!placed) atomicOr(err, 1u);
    }
}

// Direct-mapped index: the run word of each prefix (run starts only).
__global__ __launch_bounds__(JB) void k_idx_build_dense(const u64 *keys, u64 n, int shift,
                                                        u64 *words, u64 range) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 p = keys[i] >> shift;
        if (i != 0 && (keys[i - 1] >> shift) == p) continue;
        const u64 cnt = run_end(keys, n, i, p, shift) - i;
        if (p < range) words[p] = (i << kCountBits) | (cnt < kCountSat ? cnt : kCountSat);
    }
}

__device__ __forceinline__ bool lookup_run(const IdxSlot *slots, u64 n_slots, const u64 *keys,
                                           u64 n, int shift, u64 p, u64 *begin, u64 *count) {
    if (!idx_probe(slots, n_slots, p, begin, count)) return false;
    if (*begin >= n) return false;  // an index never points past its array
    if (*count == kCountSat) *count = run_end(keys, n, *begin, p, shift) - *begin;
    if (*count > n - *begin) *count = n - *begin;
    return true;
}

__global__ __launch_bounds__(JB) void k_idx_lookup(const IdxSlot *slots, u64 n_slots,
                                                   const u64 *keys, u64 n, int shift,
                                                   const u64 *q, u64 nq, u64 *begin, u32 *count) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < nq; i += (u64)gridDim.x * JB) {
        u64 b = 0, c = 0;
        if (!lookup_run(slots, n_slots, keys, n, shift, q[i], &b, &c)) c = 0;
        begin[i] = b;
        count[i] = (u32)c;
    }
}

__global__ __launch_bounds__(JB) void k_join_count(const IdxSlot *slots, u64 n_slots,
                                                   const u64 *inner, u64 n_inner, int shift,
                                                   const u64 *outer, u64 n_outer, u32 *counts,
                                                   u64 *begins) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n_outer; i += (u64)gridDim.x * JB) {
        const u64 p = outer[i] >> shift;
        u64 b = 0, c = 0;
        if (!lookup_run(slots, n_slots, inner, n_inner, shift, p, &b, &c)) c = 0;
        counts[i] = (u32)c;
        begins[i] = b;
    }
}

// offsets: n_outer + 1 entries (offsets[n_outer] == total).
__global__ __launch_bounds__(JB) void k_join_materialize(const u64 *inner, const u64 *outer,
                                                         u64 n_outer, const u64 *offsets,
                                                         const u64 *begins, u64 total, int arity,
                                                         int s0, int s1, u64 *out) {
    __shared__ u64 s_out[JM_TILE];
    const u64 tile0 = (u64)blockIdx.x * JM_TILE;
    const u64 p0 = tile0 + (u64)threadIdx.x * JM_I;
    if (p0 < total) {
        // last outer o with offsets[o] <= p0
        u64 lo = 0, hi = n_outer - 1;
        while (lo < hi) {
            const u64 mid = (lo + hi + 1) >> 1;
            if (offsets[mid] <= p0) lo = mid;
            else hi = mid - 1;
        }
        u64 o = lo;
        u64 ofs = offsets[o];
        u64 next = offsets[o + 1];
        u64 ok = outer[o];
        u64 b = begins[o];
#pragma unroll
        for (int r = 0; r < JM_I; ++r) {
            const u64 p = p0 + r;
            if (p >= total) break;
            while (next <= p && o + 1 < n_outer) {
                ++o;
                ofs = next;
                next = offsets[o + 1];
                ok = outer[o];
                b = begins[o];
            }
            const u64 ik = inner[b + (p - ofs)];
            s_out[threadIdx.x * JM_I + r] = apply_map(ik, ok, arity, s0, s1);
        }
    }
    __syncthreads();
    const u64 rem = total - tile0;
    const u32 n_tile = rem < (u64)JM_TILE ? (u32)rem : (u32)JM_TILE;
    for (u32 x = threadIdx.x; x < n_tile; x += JB) out[tile0 + x] = s_out[x];
}

// Join fused with insert-if-absent into a hash set.  Each wave takes 64
// outer tuples, probes their runs, and flattens the (outer, run position)
// pairs with a wave prefix sum, so every lane handles one produced tuple per
// round whatever the run lengths (a hub run is spread over the wave).  New
// keys are staged per wave in LDS and flushed with one atomic per kStage keys
// (counters[0] is the output cursor); counters[1] counts produced tuples.
constexpr u32 kStage = 256;
constexpr u64 kHubRun = 32;  // longer inner runs may use the load-balanced variant

// Writes stay below out_cap; a cursor past it tells the host the caller's
// output bound was wrong.
__device__ __forceinline__ void flush_stage(const u64 *stage, u32 n, u64 *out, u64 out_cap,
                                            unsigned long long *cursor, int lane) {
    unsigned long long pos = 0;
    if (lane == 0) pos = atomicAdd(cursor, (unsigned long long)n);
    pos = __shfl(pos, 0, 64);
    for (u32 x = lane; x < n; x += 64)
        if (pos + x < out_cap) out[pos + x] = stage[x];
}

// End of a fused-join block: what the waves still stage goes out with one
// cursor atomic for the whole block, and the produced count with one more
// (instead of one of each per wave, all on the same two addresses).  Every
// thread of the block calls it.
template <int B>
__device__ __forceinline__ void block_final_flush(u64 (*s_stage)[kStage], u32 staged,
                                                  u64 produced, u64 *out, u64 out_cap,
                                                  unsigned long long *counters) {
    __shared__ u32 s_cnt[B / 64];
    __shared__ u64 s_prod[B / 64];
    __shared__ unsigned long long s_base;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = 32; d >= 1; d >>= 1) produced += __shfl_xor(produced, d, 64);
    if (lane == 0) {
        s_cnt[w] = staged;
        s_prod[w] = produced;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 tot = 0, prod = 0;
        for (int v = 0; v < B / 64; ++v) {
            tot += s_cnt[v];
            prod += s_prod[v];
        }
        s_base = tot ? atomicAdd(&counters[0], (unsigned long long)tot) : 0ull;
        if (prod) atomicAdd(&counters[1], (unsigned long long)prod);
        // release: both counter updates are ordered (device scope) before
        // whatever this thread does next, so a caller that counts finished
        // blocks after this function and reads the counters after an acquire
        // fence sees them
        __threadfence();
    }
    __syncthreads();
    u64 off = s_base;
    for (int v = 0; v < w; ++v) off += s_cnt[v];
    const u64 *stage = s_stage[w];
    for (u32 x = lane; x < staged; x += 64)
        if (off + x < out_cap) out[off + x] = stage[x];
}

// Target of a fused join's inserts: one hash set, or one set per fixpoint
// instance of a batch, chosen by the key's top bits (key >> seg_shift).
struct HsSel {
    u64 *slots;
    u64 mask;
    const gdl_hs_seg *segs;
    u32 n_segs;
    int seg_shift;
    // > 0: the caller sized the set and the output by this id bound; a
    // produced key with an id at or above 2^dom_bits is not inserted and
    // sets counters[2]
    int dom_bits;
};

// The set a key belongs to (slots == nullptr: a segment with no set).
template <bool SEG>
__device__ __forceinline__ gdl_hs_seg sel_set(const HsSel &h, const gdl_hs_seg *s_segs, u64 key) {
    if constexpr (SEG) {
        const u64 s = key >> h.seg_shift;
        if (s >= h.n_segs) return gdl_hs_seg{nullptr, 0};
        return s_segs[s];
    } else {
        (void)s_segs;
        return gdl_hs_seg{h.slots, h.mask};
    }
}

template <bool SEG>
__device__ __forceinline__ void load_segs(const HsSel &h, gdl_hs_seg *s_segs) {
    if constexpr (SEG) {
        for (u32 i = threadIdx.x; i < h.n_segs; i += blockDim.x) s_segs[i] = h.segs[i];
        __syncthreads();
    } else {
        (void)h;
        (void)s_segs;
    }
}

// One block's share of a fused join + insert (the blocks of the grid stride
// over the outer tuples); counters[0] is the output cursor, counters[1] the
// produced-tuple count.
template <int B, bool SEG>
__device__ __forceinline__ void join_insert_blocks(
    const IdxSlot *slots, u64 n_slots, const u64 *inner, u64 n_inner, int shift, const u64 *outer,
    u64 n_outer, int arity, int s0, int s1, int pop, int plhs, int prhs, HsSel sel, u64 *out,
    u64 out_cap, unsigned long long *counters) {
    __shared__ u64 s_excl[B], s_b[B], s_ok[B];
    __shared__ u64 s_stage[B / 64][kStage];
    __shared__ gdl_hs_seg s_segs[SEG ? kMaxSegs : 1];
    load_segs<SEG>(sel, s_segs);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const u64 lt = lane_mask_lt();
    const u64 *wx = s_excl + w * 64;
    u64 *stage = s_stage[w];
    u32 staged = 0;  // wave-uniform
    u64 produced = 0;
    for (u64 base = (u64)blockIdx.x * B; base < n_outer; base += (u64)gridDim.x * B) {
        const u64 i = base + threadIdx.x;
        u64 b = 0, c = 0, ok = 0;
        if (i < n_outer) {
            ok = outer[i];
            if (!lookup_run(slots, n_slots, inner, n_inner, shift, ok >> shift, &b, &c)) c = 0;
        }
        const u64 incl = wave_inclusive_scan(c);
        const u64 total = __shfl(incl, 63, 64);
        s_excl[threadIdx.x] = incl - c;
        s_b[threadIdx.x] = b;
        s_ok[threadIdx.x] = ok;
        wave_sync_lds();
        for (u64 q = 0; q < total; q += 64) {
            const u64 t = q + lane;
            bool fresh = false;
            u64 key = 0;
            if (t < total) {
                // the last lane whose run starts at or before t owns position t
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (wx[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const u64 j = t - wx[lo];
                key = apply_map(inner[s_b[w * 64 + lo] + j], s_ok[w * 64 + lo], arity, s0, s1);
                if (eval_pred(key, pop, plhs, prhs)) {
                    ++produced;
                    const gdl_hs_seg hs = sel_set<SEG>(sel, s_segs, key);
                    if (sel.dom_bits && !bm_in_domain(key, sel.dom_bits))
                        atomicOr(&counters[2], 1ull);
                    else if (hs.slots)
                        fresh = hs_insert_key((u64 *)hs.slots, hs.mask, key);
                }
            }
            const u64 bal = __ballot(fresh);
            const u32 cnt = (u32)__popcll(bal);
            if (staged + cnt > kStage) {
                wave_sync_lds();
                flush_stage(stage, staged, out, out_cap, &counters[0], lane);
                staged = 0;
                wave_sync_lds();
            }
            if (fresh) stage[staged + __popcll(bal & lt)] = key;
            staged += cnt;
        }
        wave_sync_lds();  // s_excl / s_b / s_ok are rewritten next round
    }
    block_final_flush<B>(s_stage, staged, produced, out, out_cap, counters);
}

template <int B, bool SEG>
__global__ __launch_bounds__(B) void k_join_insert(
    const IdxSlot *slots, u64 n_slots, const u64 *inner, u64 n_inner, int shift, const u64 *outer,
    u64 n_outer, int arity, int s0, int s1, int pop, int plhs, int prhs, HsSel sel, u64 *out,
    u64 out_cap, unsigned long long *counters) {
    join_insert_blocks<B, SEG>(slots, n_slots, inner, n_inner, shift, outer, n_outer, arity, s0,
                               s1, pop, plhs, prhs, sel, out, out_cap, counters);
}

// State of a stream-ordered linear fixpoint (gdl_fixpoint_linear) in device
// memory.  cursor and produced come first: they are a body's counters.
struct LinState {
    unsigned long long cursor;    // log end while a body runs
    unsigned long long produced;  // tuples the body produced
    u64 delta_begin, delta_end;   // the body's delta: log[delta_begin, delta_end)
    u64 size;                     // hash set members
    u64 hs_limit;                 // members the set may reach (a fraction of its capacity)
    u64 log_cap;                  // log capacity, keys
    u64 max_run;                  // longest inner run: new tuples <= |delta| x max_run
    u64 delta_limit;              // a larger delta ends the chunk (the grid was sized smaller)
    u32 stop;                     // 0 run, 1 fixpoint, 2 grow first, 3 body limit, 4 overflow
    u32 body;                     // bodies finished in this call
    u32 blocks_done;              // blocks of the running body that finished
    u32 max_bodies;
};
enum : u32 { kLinRun = 0, kLinFixpoint = 1, kLinGrow = 2, kLinLimit = 3, kLinOverflow = 4 };

// End of a body (thread 0 of the last block to finish): record its counts,
// move the delta window to the tuples it appended, and decide whether the
// next enqueued body runs.  Every block releases its cursor / produced
// updates (a device-scope fence) before it counts itself finished, and the
// last block fences again before reading them, so the atomic reads below see
// the final values; the plain stores reach the next body at its launch.
__device__ void linear_epilogue(LinState *st, u64 *counts, u64 d1) {
    const u64 end = atomicAdd(&st->cursor, 0ull);
    const u64 produced = atomicAdd(&st->produced, 0ull);
    const u32 b = st->body;
    const u64 nd = end > d1 ? end - d1 : 0;
    counts[2 * b] = nd;
    counts[2 * b + 1] = produced;
    const u64 size = st->size + nd;
    st->size = size;
    st->delta_begin = d1;
    st->delta_end = end;
    st->produced = 0;
    st->blocks_done = 0;
    st->body = b + 1;
    u32 stop = kLinRun;
    if (end > st->log_cap) stop = kLinOverflow;
    else if (nd == 0) stop = kLinFixpoint;
    else if (b + 1 >= st->max_bodies || nd > st->delta_limit) stop = kLinLimit;
    else if (size + nd * st->max_run > st->hs_limit || end + nd * st->max_run > st->log_cap)
        stop = kLinGrow;
    st->stop = stop;
}

// One body of out := out U map(inner x delta) with FULL = hash set + append
// log: the delta is the log's last segment, new tuples are appended after it.
// A stop code set by an earlier body makes it return at once.
template <int B>
__global__ __launch_bounds__(B) void k_linear_body(const IdxSlot *slots, u64 n_slots,
                                                   const u64 *inner, u64 n_inner, int shift,
                                                   int arity, int s0, int s1, int pop, int plhs,
                                                   int prhs, u64 *hs_slots, u64 hs_mask, u64 *log,
                                                   LinState *st, u64 *counts) {
    __shared__ bool s_last;
    if (st->stop != kLinRun) return;  // the same for every block
    const u64 d0 = st->delta_begin, d1 = st->delta_end;
    join_insert_blocks<B, false>(slots, n_slots, inner, n_inner, shift, log + d0, d1 - d0, arity,
                                 s0, s1, pop, plhs, prhs, HsSel{hs_slots, hs_mask, nullptr, 0, 0},
                                 log, st->log_cap, &st->cursor);
    // join_insert_blocks ends with this block's counter updates acknowledged
    // (thread 0, before a barrier)
    if (threadIdx.x == 0) s_last = atomicAdd(&st->blocks_done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();  // acquire: every block's counter updates precede its blocks_done add
        linear_epilogue(st, counts, d1);
    }
}

// Load-balanced variant for inner relations with long runs: outputs are
// numbered by the exclusive scan of run lengths (offsets), each thread takes
// JM_I consecutive output positions of its block's tile, computes their keys
// by walking the offsets, then issues the home-slot CASes of all of them
// back to back before resolving collisions.
template <bool SEG>
__global__ __launch_bounds__(JB) void k_join_insert_lb(
    const u64 *inner, const u64 *outer, u64 n_outer, const u64 *offsets, const u64 *begins,
    u64 total, int arity, int s0, int s1, int pop, int plhs, int prhs, HsSel sel, u64 *out,
    u64 out_cap, unsigned long long *counters) {
    __shared__ u64 s_stage[JB / 64][kStage];
    __shared__ gdl_hs_seg s_segs[SEG ? kMaxSegs : 1];
    load_segs<SEG>(sel, s_segs);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const u64 lt = lane_mask_lt();
    u64 *stage = s_stage[w];
    u32 staged = 0;
    u64 produced = 0;
    for (u64 tile0 = (u64)blockIdx.x * JM_TILE; tile0 < total; tile0 += (u64)gridDim.x * JM_TILE) {
        const u64 p0 = tile0 + (u64)threadIdx.x * JM_I;
        u64 keys[JM_I];
        u32 live = 0;  // bit r: output r exists and passes the predicate
        if (p0 < total) {
            u64 lo = 0, hi = n_outer - 1;
            while (lo < hi) {
                const u64 mid = (lo + hi + 1) >> 1;
                if (offsets[mid] <= p0) lo = mid;
                else hi = mid - 1;
            }
            u64 o = lo, ofs = offsets[o], next = offsets[o + 1], ok = outer[o], b = begins[o];
#pragma unroll
            for (int r = 0; r < JM_I; ++r) {
                const u64 p = p0 + r;
                keys[r] = 0;
                if (p >= total) continue;
                while (next <= p && o + 1 < n_outer) {
                    ++o;
                    ofs = next;
                    next = offsets[o + 1];
                    ok = outer[o];
                    b = begins[o];
                }
                keys[r] = apply_map(inner[b + (p - ofs)], ok, arity, s0, s1);
                if (eval_pred(keys[r], pop, plhs, prhs)) live |= 1u << r;
            }
        }
        produced += __popc(live);
        // home-slot CAS (bitmap: the OR of the key's bit) of every key first,
        // independent round trips in flight
        unsigned long long old[JM_I];
        u32 bm = 0;  // bit r: key r went to a bitmap set (old[r]: its bit before)
#pragma unroll
        for (int r = 0; r < JM_I; ++r) {
            old[r] = 0;
            if ((live >> r) & 1u) {
                const gdl_hs_seg hs = sel_set<SEG>(sel, s_segs, keys[r]);
                if (sel.dom_bits && !bm_in_domain(keys[r], sel.dom_bits)) {
                    atomicOr(&counters[2], 1ull);
                    live &= ~(1u << r);
                } else if (!hs.slots) {
                    live &= ~(1u << r);
                } else if (hs.mask & kHsBitmap) {
                    const int kb = (int)(hs.mask & 63);
                    if (bm_in_domain(keys[r], kb)) {
                        // a plain load first: a set bit proves membership
                        const u64 b = bm_bit(keys[r], kb);
                        old[r] = ((const u64 *)hs.slots)[b >> 6] & (1ull << (b & 63));
                        bm |= 1u << r;
                    }
                } else if (keys[r] != kHsEmpty) {
                    old[r] = atomicCAS(
                        (unsigned long long *)((u64 *)hs.slots + (mix64(keys[r]) & hs.mask)),
                        (unsigned long long)kHsEmpty, (unsigned long long)keys[r]);
                }
            }
        }
        // bitmap keys whose bit read clear: the atomic OR decides, all in flight
#pragma unroll
        for (int r = 0; r < JM_I; ++r) {
            if (((bm >> r) & 1u) && old[r] == 0) {
                const gdl_hs_seg hs = sel_set<SEG>(sel, s_segs, keys[r]);
                const u64 b = bm_bit(keys[r], (int)(hs.mask & 63));
                const u64 m = 1ull << (b & 63);
                old[r] = atomicOr((unsigned long long *)((u64 *)hs.slots + (b >> 6)),
                                  (unsigned long long)m) & m;
            }
        }
#pragma unroll
        for (int r = 0; r < JM_I; ++r) {
            bool fresh = false;
            if ((bm >> r) & 1u) {
                fresh = old[r] == 0;
            } else if ((live >> r) & 1u) {
                if (keys[r] == kHsEmpty || (old[r] != (unsigned long long)kHsEmpty &&
                                            old[r] != (unsigned long long)keys[r]) ||
                    (sel_set<SEG>(sel, s_segs, keys[r]).mask & kHsBitmap)) {
                    // collision (or the key ~0): the full probe
                    const gdl_hs_seg hs = sel_set<SEG>(sel, s_segs, keys[r]);
                    fresh = hs_insert_key((u64 *)hs.slots, hs.mask, keys[r]);
                } else {
                    fresh = old[r] == (unsigned long long)kHsEmpty;
                }
            }
            const u64 bal = __ballot(fresh);
            const u32 cnt = (u32)__popcll(bal);
            if (staged + cnt > kStage) {
                wave_sync_lds();
                flush_stage(stage, staged, out, out_cap, &counters[0], lane);
                staged = 0;
                wave_sync_lds();
            }
            if (fresh) stage[staged + __popcll(bal & lt)] = keys[r];
            staged += cnt;
        }
    }
    block_final_flush<JB>(s_stage, staged, produced, out, out_cap, counters);
}

__global__ void k_set2(u64 *p, u64 a, u64 b) {
    p[0] = a;
    p[1] = b;
}

__global__ void k_set3(u64 *p, u64 a, u64 b, u64 c) {
    p[0] = a;
    p[1] = b;
    p[2] = c;
}

__global__ __launch_bounds__(JB) void k_project(const u64 *in, u64 n, int arity, int s0, int s1,
                                                u64 *out) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 k = in[i];
        out[i] = apply_map(k, k, arity, s0, s1);
    }
}

static unsigned stride_grid(gdl_ctx *ctx, u64 n) {
    const u64 cap = grid_cap(ctx);
    u64 g = (n + JB - 1) / JB;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

static bool valid_map(const gdl_colmap &m, bool join) {
    if (m.arity < 1 || m.arity > 2) return false;
    const int maxsrc = join ? 3 : 1;
    for (int c = 0; c < m.arity; ++c)
        if (m.src[c] < 0 || m.src[c] > maxsrc) return false;
    return true;
}

gdl_status index_build(gdl_ctx *ctx, const u64 *keys, u64 n, int prefix_cols, gdl_index *out) {
    out->slots = nullptr;
    out->n_slots = 0;
    out->n_distinct = 0;
    out->prefix_cols = prefix_cols;
    out->max_run = 0;
    out->dense = 0;
    if (n == 0) return GDL_OK;
    const int shift = prefix_cols >= 2 ? 0 : 32;
    u64 distinct = 0, max_run = 0, max_prefix = 0;
    {
        Scratch scr(ctx);
        u64 *cnt;
        GDL_TRY(scr.get(3, &cnt));
        GDL_HIP(ctx, hipMemsetAsync(cnt, 0, 3 * sizeof(u64), ctx->stream));
        KTimer kt(ctx, K_IDX_COUNT, 8.0 * n);
        const unsigned g = std::min<unsigned>(stride_grid(ctx, n), 512u);
        hipLaunchKernelGGL(k_idx_count, dim3(g), dim3(JB), 0, ctx->stream, keys, n, shift, cnt);
        GDL_TRY(check_launch(ctx, "k_idx_count"));
        u64 h[3];
        GDL_TRY(read_u64(ctx, cnt, h, 3));
        distinct = h[0];
        max_run = h[1];
        max_prefix = h[2];
    }
    out->n_distinct = distinct;
    out->max_run = max_run;
    // one column of ids that are dense enough: a direct-mapped array of run
    // words no larger than the hash index (or at most 1 MB)
    const u64 hash_slots = (u64)((double)distinct / (double)ctx->cfg.load_factor) + 1;
    if (prefix_cols == 1 && max_prefix < std::max<u64>(2 * hash_slots, 1ull << 17)) {
        const u64 range = max_prefix + 1;
        void *p = nullptr;
        std::string err;
        gdl_status st = ctx->alloc->alloc(range * sizeof(u64), &p, &err);
        if (st != GDL_OK) return fail(ctx, st, err);
        GDL_HIP(ctx, hipMemsetAsync(p, 0, range * sizeof(u64), ctx->stream));
        KTimer kt(ctx, K_IDX_BUILD, 8.0 * n + 8.0 * (double)distinct);
        hipLaunchKernelGGL(k_idx_build_dense, dim3(stride_grid(ctx, n)), dim3(JB), 0, ctx->stream,
                           keys, n, shift, (u64 *)p, range);
        GDL_TRY(check_launch(ctx, "k_idx_build_dense"));
        out->slots = p;
        out->n_slots = range;
        out->dense = 1;
        return GDL_OK;
    }
    u64 n_slots = (u64)((double)distinct / (double)ctx->cfg.load_factor) + 1;
    if (n_slots <= distinct) n_slots = distinct + 1;
    n_slots = (n_slots + kBucket - 1) / kBucket * kBucket;  // whole buckets
    void *p = nullptr;
    std::string err;
    gdl_status st = ctx->alloc->alloc(n_slots * sizeof(IdxSlot), &p, &err);
    if (st != GDL_OK) return fail(ctx, st, err);
    GDL_HIP(ctx, hipMemsetAsync(p, 0, n_slots * sizeof(IdxSlot), ctx->stream));
    {
        KTimer kt(ctx, K_IDX_BUILD, 8.0 * n + 16.0 * (double)distinct);
        hipLaunchKernelGGL(k_idx_build, dim3(stride_grid(ctx, n)), dim3(JB), 0, ctx->stream, keys,
                           n, shift, (IdxSlot *)p, n_slots, ctx->d_err);
        GDL_TRY(check_launch(ctx, "k_idx_build"));
    }
    out->slots = p;
    out->n_slots = n_slots;
    out->n_distinct = distinct;
    out->max_run = max_run;
    return GDL_OK;
}

gdl_status join_count(gdl_ctx *ctx, const u64 *inner, u64 n_inner, const gdl_index *idx,
                      const u64 *outer, u64 n_outer, u32 *counts, u64 *begins) {
    if (n_outer == 0) return GDL_OK;
    const int shift = idx->prefix_cols >= 2 ? 0 : 32;
    if (n_inner == 0 || idx->n_slots == 0) {
        GDL_HIP(ctx, hipMemsetAsync(counts, 0, n_outer * sizeof(u32), ctx->stream));
        GDL_HIP(ctx, hipMemsetAsync(begins, 0, n_outer * sizeof(u64), ctx->stream));
        return GDL_OK;
    }
    KTimer kt(ctx, K_JOIN_COUNT, 8.0 * n_outer + 16.0 * n_outer + 12.0 * n_outer);
    hipLaunchKernelGGL(k_join_count, dim3(stride_grid(ctx, n_outer)), dim3(JB), 0, ctx->stream,
                       (const IdxSlot *)idx->slots, idx_arg(idx), inner, n_inner, shift, outer,
                       n_outer, counts, begins);
    return check_launch(ctx, "k_join_count");
}

gdl_status join_materialize(gdl_ctx *ctx, const u64 *inner, const u64 *outer, u64 n_outer,
                            const u64 *offsets, const u64 *begins, u64 total, gdl_colmap map,
                            u64 *out) {
    if (total == 0) return GDL_OK;
    const u64 nt = (total + JM_TILE - 1) / JM_TILE;
    if (nt > 0x7fffffffull) return fail(ctx, GDL_ERR_OVERFLOW, "join: output too large");
    KTimer kt(ctx, K_JOIN_MAT, 16.0 * total);
    hipLaunchKernelGGL(k_join_materialize, dim3((unsigned)nt), dim3(JB), 0, ctx->stream, inner,
                       outer, n_outer, offsets, begins, total, map.arity, map.src[0],
                       map.arity == 2 ? map.src[1] : 0, out);
    return check_launch(ctx, "k_join_materialize");
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_index_build(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n, int prefix_cols,
                           gdl_index *out) {
    if (!ctx || !out || (n && !d_sorted) || prefix_cols < 1 || prefix_cols > 2)
        return GDL_ERR_INVALID_ARG;
    return index_build(ctx, d_sorted, n, prefix_cols, out);
}

gdl_status gdl_index_free(gdl_ctx *ctx, gdl_index *idx) {
    if (!ctx || !idx) return GDL_ERR_INVALID_ARG;
    if (idx->slots) {
        std::string err;
        gdl_status st = ctx->alloc->free(idx->slots, &err);
        if (st != GDL_OK) return fail(ctx, st, err);
    }
    idx->slots = nullptr;
    idx->n_slots = 0;
    idx->n_distinct = 0;
    idx->dense = 0;
    return GDL_OK;
}

gdl_status gdl_index_lookup(gdl_ctx *ctx, const gdl_index *idx, const uint64_t *d_sorted,
                            uint64_t n_sorted, const uint64_t *d_prefixes, uint64_t n,
                            uint64_t *d_begin, uint32_t *d_count) {
    if (!ctx || !idx || (n && (!d_prefixes || !d_begin || !d_count))) return GDL_ERR_INVALID_ARG;
    if (n == 0) return GDL_OK;
    if (idx->n_slots == 0) {
        GDL_HIP(ctx, hipMemsetAsync(d_count, 0, n * sizeof(u32), ctx->stream));
        GDL_HIP(ctx, hipMemsetAsync(d_begin, 0, n * sizeof(u64), ctx->stream));
        return GDL_OK;
    }
    const int shift = idx->prefix_cols >= 2 ? 0 : 32;
    KTimer kt(ctx, K_IDX_LOOKUP, 28.0 * n);
    hipLaunchKernelGGL(k_idx_lookup, dim3(stride_grid(ctx, n)), dim3(JB), 0, ctx->stream,
                       (const IdxSlot *)idx->slots, idx_arg(idx), d_sorted, n_sorted, shift,
                       d_prefixes, n, d_begin, d_count);
    return check_launch(ctx, "k_idx_lookup");
}

gdl_status gdl_join_count(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                          const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                          uint32_t *d_counts, uint64_t *d_begins) {
    if (!ctx || !inner_idx || (n_outer && (!d_outer || !d_counts || !d_begins)) ||
        (n_inner && !d_inner))
        return GDL_ERR_INVALID_ARG;
    return join_count(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, d_counts, d_begins);
}

gdl_status gdl_join_materialize(gdl_ctx *ctx, const uint64_t *d_inner, const uint64_t *d_outer,
                                uint64_t n_outer, const uint64_t *d_offsets,
                                const uint64_t *d_begins, uint64_t total, gdl_colmap map,
                                uint64_t *d_out) {
    if (!ctx || (total && (!d_inner || !d_outer || !d_offsets || !d_begins || !d_out)))
        return GDL_ERR_INVALID_ARG;
    if (!valid_map(map, true)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad join column map");
    return join_materialize(ctx, d_inner, d_outer, n_outer, d_offsets, d_begins, total, map, d_out);
}

gdl_status gdl_join(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                    const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                    gdl_colmap map, gdl_pred pred, uint64_t **d_out, uint64_t *n_out) {
    if (!ctx || !inner_idx || !d_out || !n_out || (n_outer && !d_outer) || (n_inner && !d_inner))
        return GDL_ERR_INVALID_ARG;
    if (!valid_map(map, true)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad join column map");
    if (!valid_pred(pred, map.arity))
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate");
    *d_out = nullptr;
    *n_out = 0;
    if (n_outer == 0 || n_inner == 0) return GDL_OK;
    Scratch scr(ctx);
    u32 *counts;
    u64 *begins, *offsets;
    GDL_TRY(scr.get(n_outer, &counts));
    GDL_TRY(scr.get(n_outer, &begins));
    GDL_TRY(scr.get(n_outer + 1, &offsets));
    GDL_TRY(join_count(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, counts, begins));
    u64 total = 0;
    GDL_TRY(exclusive_scan_u32_to_u64(ctx, counts, n_outer, offsets, &total));
    if (total == 0) return GDL_OK;
    void *buf = nullptr;
    std::string err;
    gdl_status st = ctx->alloc->alloc(total * sizeof(u64), &buf, &err);
    if (st != GDL_OK) return fail(ctx, st, err);
    u64 *res = (u64 *)buf;
    st = join_materialize(ctx, d_inner, d_outer, n_outer, offsets, begins, total, map, res);
    if (st != GDL_OK) {
        ctx->alloc->free(buf, nullptr);
        return st;
    }
    if (pred.op != GDL_PRED_NONE) {
        void *fbuf = nullptr;
        st = ctx->alloc->alloc(total * sizeof(u64), &fbuf, &err);
        if (st != GDL_OK) {
            ctx->alloc->free(buf, nullptr);
            return fail(ctx, st, err);
        }
        u64 kept = 0;
        st = filter_keys(ctx, res, total, pred, (u64 *)fbuf, &kept);
        ctx->alloc->free(buf, nullptr);
        if (st != GDL_OK) {
            ctx->alloc->free(fbuf, nullptr);
            return st;
        }
        *d_out = (u64 *)fbuf;
        *n_out = kept;
        return GDL_OK;
    }
    *d_out = res;
    *n_out = total;
    return GDL_OK;
}

}  // extern "C"

// Fused join + insert.  SEG = false: into one set `hs`, growing it and the
// output (*d_new, *new_cap) to the bound.  SEG = true: into the per-instance
// sets of `sel`, which the caller has sized, and into the caller's output of
// *new_cap keys (a fixed buffer; an overrun is reported, never written).
template <bool SEG>
static gdl_status join_insert_impl(gdl_ctx *ctx, const u64 *d_inner, u64 n_inner,
                                   const gdl_index *inner_idx, const u64 *d_outer, u64 n_outer,
                                   gdl_colmap map, gdl_pred pred, gdl_hashset *hs, HsSel sel,
                                   u64 **d_new, u64 *new_cap, u64 *n_new, u64 *n_raw) {
    *n_raw = 0;
    if (n_outer == 0 || n_inner == 0 || inner_idx->n_slots == 0) return GDL_OK;
    const int shift = inner_idx->prefix_cols >= 2 ? 0 : 32;
    // upper bound of produced tuples: |outer| x longest inner run, or an
    // exact count when that bound is loose
    const u64 mr = inner_idx->max_run ? inner_idx->max_run : n_inner;
    u64 bound = (n_outer > ~0ull / mr) ? ~0ull : n_outer * mr;
    const u64 loose_limit = std::max<u64>(1ull << 24, 4 * n_outer);
    const u64 cap_blocks = grid_cap(ctx);
    const int fb = (int)ctx->tune.join_block;  // block of the flattened kernel
    // the count sizes the output set and buffer; with per-instance sets (SEG)
    // the caller has sized both already (the output capacity is checked after
    // the kernel), so only the load-balanced kernel below counts, for its
    // offsets
    // (and a join that would run on bit matrices if its bound were its count)
    bool counted = !SEG && bound > loose_limit;
    if constexpr (!SEG)
        counted = counted || dense_join_pays(ctx, hs, map, pred, inner_idx->prefix_cols, n_inner,
                                             n_outer, bound);
    Scratch lb(ctx);
    u32 *counts = nullptr;
    u64 *begins = nullptr, *offsets = nullptr;
    if (counted) {
        GDL_TRY(lb.get(n_outer, &counts));
        GDL_TRY(lb.get(n_outer, &begins));
        GDL_TRY(lb.get(n_outer + 1, &offsets));
        GDL_TRY(join_count(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, counts, begins));
        GDL_TRY(exclusive_scan_u32_to_u64(ctx, counts, n_outer, offsets, &bound));
    }
    if (bound == 0) return GDL_OK;
    if constexpr (!SEG) {
        // many tuples per new one over a small id domain: bit matrices
        gdl_status ds = GDL_OK;
        if (counted && dense_join_insert(ctx, d_inner, n_inner, d_outer, n_outer, map, pred,
                                         inner_idx->prefix_cols, hs, bound, d_new, new_cap,
                                         n_new, n_raw, &ds))
            return ds;
    }
    // The wave-flattened kernel is balanced inside a wave but not across
    // waves: a wave whose 64 outer tuples hit a hub walks the whole hub run,
    // one step (64 outputs) per hub tuple.  Number the outputs and spread
    // them over the grid instead when that is long against the expected
    // steps per wave.
    const u64 waves = std::min<u64>((n_outer + fb - 1) / fb, cap_blocks) * (fb / 64);
    const u64 avg_run = std::max<u64>(1, n_inner / std::max<u64>(1, inner_idx->n_distinct));
    const u64 est_total = counted ? bound : n_outer * avg_run;
    // (or when there are too few waves for the work: each would walk many
    // 64-output steps, each a chain of dependent loads)
    const bool balanced =
        mr > kHubRun && (mr * 16 > est_total / waves || est_total / (waves * 64) > 32);
    if (ctx->cfg.verbose)
        fprintf(stderr,
                "gdl: join_insert outer %llu inner %llu max_run %llu produced %s%llu set %s%d -> %s\n",
                (unsigned long long)n_outer, (unsigned long long)n_inner, (unsigned long long)mr,
                counted ? "" : "<= ", (unsigned long long)bound,
                hs && hs->key_bits ? "bitmap " : "hash ", hs ? (int)hs->key_bits : 0,
                balanced ? "balanced" : "wave-flattened");
    if (balanced && !counted) {
        GDL_TRY(lb.get(n_outer, &counts));
        GDL_TRY(lb.get(n_outer, &begins));
        GDL_TRY(lb.get(n_outer + 1, &offsets));
        GDL_TRY(join_count(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, counts, begins));
        GDL_TRY(exclusive_scan_u32_to_u64(ctx, counts, n_outer, offsets, &bound));
        if (bound == 0) return GDL_OK;
    }
    if constexpr (!SEG) {
        // new members are distinct keys of the id domain the set does not
        // hold yet: at most the domain's free keys, however many tuples the
        // join produces
        u64 room = ~0ull;
        if (hs->key_bits) {
            room = hs->capacity - std::min(hs->capacity, hs->size);
        } else if (ctx->id_bits > 0 && ctx->id_bits < 32) {
            const u64 dom = 1ull << (2 * ctx->id_bits);
            room = dom - std::min(dom, hs->size);
            sel.dom_bits = ctx->id_bits;  // the kernel checks every key against it
        }
        const u64 new_bound = std::min(bound, room);
        // size for up to 2 |outer| new members at the load factor; the bound
        // only has to fit
        GDL_TRY(hashset_reserve2(ctx, hs, std::max<u64>(2 * n_outer, 1024), new_bound));
        sel.slots = (u64 *)hs->slots;
        sel.mask = hs_arg(hs);
        if (*n_new + new_bound > *new_cap || !*d_new) {
            const u64 cap = *n_new + new_bound + (*n_new + new_bound) / 4 + 1024;
            void *p = nullptr;
            std::string err;
            gdl_status st = ctx->alloc->alloc(cap * sizeof(u64), &p, &err);
            if (st != GDL_OK) return fail(ctx, st, err);
            if (*n_new)
                GDL_HIP(ctx, hipMemcpyAsync(p, *d_new, *n_new * sizeof(u64),
                                            hipMemcpyDeviceToDevice, ctx->stream));
            if (*d_new) ctx->alloc->free(*d_new, nullptr);
            *d_new = (u64 *)p;
            *new_cap = cap;
        }
    }
    Scratch scr(ctx);
    u64 *counters;
    GDL_TRY(scr.get(3, &counters));
    hipLaunchKernelGGL(k_set3, dim3(1), dim3(1), 0, ctx->stream, counters, (u64)*n_new, (u64)0,
                       (u64)0);
    GDL_TRY(check_launch(ctx, "k_set3"));
    if (balanced) {
        KTimer kt(ctx, K_JOIN_INSERT, 24.0 * n_outer);
        const unsigned g = (unsigned)std::max<u64>(
            1, std::min<u64>((bound + JM_TILE - 1) / JM_TILE, cap_blocks));
        hipLaunchKernelGGL(k_join_insert_lb<SEG>, dim3(g), dim3(JB), 0, ctx->stream, d_inner,
                           d_outer, n_outer, offsets, begins, bound, map.arity, map.src[0],
                           map.arity == 2 ? map.src[1] : 0, pred.op, pred.lhs, pred.rhs, sel,
                           *d_new, *new_cap, (unsigned long long *)counters);
        GDL_TRY(check_launch(ctx, "k_join_insert_lb"));
    } else {
        KTimer kt(ctx, K_JOIN_INSERT, 8.0 * n_outer + 64.0 * n_outer);
        const unsigned g = (unsigned)std::max<u64>(1, std::min<u64>((n_outer + fb - 1) / fb,
                                                                    cap_blocks));
        auto launch = [&](auto kern, int block) {
            hipLaunchKernelGGL(kern, dim3(g), dim3(block), 0, ctx->stream,
                               (const IdxSlot *)inner_idx->slots, idx_arg(inner_idx), d_inner,
                               n_inner, shift, d_outer, n_outer, map.arity, map.src[0],
                               map.arity == 2 ? map.src[1] : 0, pred.op, pred.lhs, pred.rhs, sel,
                               *d_new, *new_cap, (unsigned long long *)counters);
        };
        switch (fb) {
            case 64: launch(k_join_insert<64, SEG>, 64); break;
            case 128: launch(k_join_insert<128, SEG>, 128); break;
            case 512: launch(k_join_insert<512, SEG>, 512); break;
            default: launch(k_join_insert<256, SEG>, 256); break;
        }
        GDL_TRY(check_launch(ctx, "k_join_insert"));
    }
    u64 h[3];
    GDL_TRY(read_u64(ctx, counters, h, 3));
    if (h[2])
        return fail(ctx, GDL_ERR_OVERFLOW,
                    "fused join: a produced key has an id at or above 2^" +
                        std::to_string(sel.dom_bits) + ", the context's id bound (gdl_set_id_bits)");
    if (h[0] > *new_cap)
        return fail(ctx, GDL_ERR_OVERFLOW, "fused join: new tuples exceed the output capacity (" +
                                               std::to_string(h[0]) + " > " +
                                               std::to_string(*new_cap) + ")");
    const u64 added = h[0] - *n_new;
    // inner run reads + hash probes per produced tuple, new keys written
    ctx->stats[K_JOIN_INSERT].alg_bytes += 16.0 * h[1] + 8.0 * added;
    if constexpr (!SEG) hs->size += added;
    *n_new = h[0];
    *n_raw = h[1];
    return GDL_OK;
}

extern "C" {

gdl_status gdl_join_insert(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                           const gdl_index *inner_idx, const uint64_t *d_outer, uint64_t n_outer,
                           gdl_colmap map, gdl_pred pred, gdl_hashset *hs, uint64_t **d_new,
                           uint64_t *new_cap, uint64_t *n_new, uint64_t *n_raw) {
    if (!ctx || !inner_idx || !hs || !d_new || !new_cap || !n_new || !n_raw ||
        (n_outer && !d_outer) || (n_inner && !d_inner))
        return GDL_ERR_INVALID_ARG;
    if (!valid_map(map, true)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad join column map");
    if (!valid_pred(pred, map.arity))
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate");
    return join_insert_impl<false>(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, map, pred,
                                   hs, HsSel{nullptr, 0, nullptr, 0, 0}, d_new, new_cap, n_new,
                                   n_raw);
}

gdl_status gdl_join_insert_seg(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                               const gdl_index *inner_idx, const uint64_t *d_outer,
                               uint64_t n_outer, gdl_colmap map, gdl_pred pred,
                               const gdl_hs_seg *d_segs, uint32_t n_segs, int seg_shift,
                               uint64_t *d_out, uint64_t out_cap, uint64_t *n_out,
                               uint64_t *n_raw) {
    if (!ctx || !inner_idx || !d_segs || !n_out || !n_raw || (n_outer && !d_outer) ||
        (n_inner && !d_inner) || (out_cap && !d_out))
        return GDL_ERR_INVALID_ARG;
    if (n_segs == 0 || n_segs > kMaxSegs || seg_shift < 1 || seg_shift > 63)
        return fail(ctx, GDL_ERR_INVALID_ARG, "segmented join insert: 1..256 sets, shift 1..63");
    if (!valid_map(map, true)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad join column map");
    if (!valid_pred(pred, map.arity))
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate");
    *n_out = 0;
    u64 *out = d_out;
    u64 cap = out_cap;
    return join_insert_impl<true>(ctx, d_inner, n_inner, inner_idx, d_outer, n_outer, map, pred,
                                  nullptr, HsSel{nullptr, 0, d_segs, n_segs, seg_shift}, &out,
                                  &cap, n_out, n_raw);
}

gdl_status gdl_fixpoint_linear(gdl_ctx *ctx, const uint64_t *d_inner, uint64_t n_inner,
                               const gdl_index *inner_idx, gdl_colmap map, gdl_pred pred,
                               gdl_hashset *hs, uint64_t **d_log, uint64_t *log_cap,
                               uint64_t *delta_begin, uint64_t *delta_end, int max_bodies,
                               uint64_t *h_new, uint64_t *h_raw, int *bodies) {
    if (!ctx || !inner_idx || !hs || !d_log || !log_cap || !delta_begin || !delta_end ||
        !h_new || !h_raw || !bodies || (n_inner && !d_inner) || *delta_begin > *delta_end ||
        (*delta_end && !*d_log) || *delta_end > *log_cap)
        return GDL_ERR_INVALID_ARG;
    if (!valid_map(map, true)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad join column map");
    if (!valid_pred(pred, map.arity))
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate");
    *bodies = 0;
    if (max_bodies <= 0) return GDL_OK;
    if (n_inner == 0 || inner_idx->n_slots == 0 || *delta_end == *delta_begin) {
        h_new[0] = h_raw[0] = 0;  // one body, nothing produced: the fixpoint
        *delta_begin = *delta_end;
        *bodies = 1;
        return GDL_OK;
    }
    // at most 128 bodies per host wait (on the 200k-node tree 128 beat 32 by
    // 2.5 %; 64 did not; a chunk that stops early leaves no-op launches)
    constexpr int kChunk = 128;
    // a body is bound by the latency of each tuple's chain of dependent
    // accesses, so every outer tuple gets its own lane (more chains in
    // flight) in blocks of 512 (fewer blocks: fewer barriers and end-of-block
    // counter atomics on the shared state); on the 200k-node tree both beat
    // 256-thread blocks with two tuples per thread
    constexpr int LB = 512;
    const int shift = inner_idx->prefix_cols >= 2 ? 0 : 32;
    const u64 mr = std::max<u64>(inner_idx->max_run, 1);
    Scratch scr(ctx);
    u64 *st_words, *counts;
    GDL_TRY(scr.get((sizeof(LinState) + 7) / 8, &st_words));
    GDL_TRY(scr.get(2 * kChunk, &counts));
    LinState *st = (LinState *)st_words;
    u64 d0 = *delta_begin, d1 = *delta_end;
    // TC's map over a bitmap FULL: bodies whose delta is dense run on bit
    // matrices (dense.hip); the delta size where they start
    const bool tc_map = map.arity == 2 && map.src[0] == GDL_COL_INNER1 &&
                        map.src[1] == GDL_COL_OUTER1 && pred.op == GDL_PRED_NONE &&
                        inner_idx->prefix_cols == 1;
    const gdl_dense_rule rule{GDL_DENSE_TC, d_inner, n_inner, nullptr, 0, 0, 0};
    u64 dense_at = tc_map ? dense_threshold_for(ctx, hs, GDL_DENSE_TC, n_inner, 0) : ~0ull;
    int done = 0, ramp = 4;
    while (done < max_bodies) {
        const u64 nd = d1 - d0;
        if (nd >= dense_at) {
            int nb = 0;
            u64 end = d1;
            GDL_TRY(dense_run(ctx, &rule, hs, *d_log + d0, nd, d_log, log_cap, &end,
                              max_bodies - done, h_new + done, h_raw + done, &nb));
            if (nb > 0) {
                done += nb;
                const u64 last = h_new[done - 1];
                d0 = end - last;
                d1 = end;
                if (last == 0) break;  // the fixpoint
                continue;
            }
            dense_at = ~0ull;  // declined (ids beyond the set's domain): tuple bodies only
        }
        const u64 worst = nd > (~0ull >> 20) / mr ? (~0ull >> 20) : nd * mr;
        // the set: the next chunk's bodies (about |delta| each) at the load
        // factor and the worst body at 0.9; when it has to grow, it grows to
        // room for 4x its size more (each growth is a rehash of every member,
        // so a set that keeps growing is rehashed every ~5x, not every 2x);
        // the log: the worst body after the delta, grown 2x at a time (each
        // growth copies it)
        const u64 next = std::max<u64>(nd, 1024) * (u64)std::min(ramp, kChunk);
        GDL_TRY(hashset_reserve2(ctx, hs, next, std::max(worst, next), 0.0,
                                 std::max<u64>(next, 4 * hs->size)));
        u64 need_log = d1 + std::max<u64>(worst, 4096);
        if (*log_cap < need_log) {
            need_log = std::max<u64>(need_log, 2 * *log_cap);  // geometric
            void *p = nullptr;
            std::string err;
            gdl_status s = ctx->alloc->alloc(need_log * sizeof(u64), &p, &err);
            if (s != GDL_OK) return fail(ctx, s, err);
            if (d1)
                GDL_HIP(ctx, hipMemcpyAsync(p, *d_log, d1 * sizeof(u64), hipMemcpyDeviceToDevice,
                                            ctx->stream));
            if (*d_log) ctx->alloc->free(*d_log, nullptr);
            *d_log = (u64 *)p;
            *log_cap = need_log;
        }
        const int k = std::min({ramp, kChunk, max_bodies - done});
        LinState h{};
        h.cursor = d1;
        h.delta_begin = d0;
        h.delta_end = d1;
        h.size = hs->size;
        h.hs_limit = hs->key_bits ? ~0ull : (u64)(kHsHardLoad * (double)hs->capacity);
        h.log_cap = *log_cap;
        h.max_run = mr;
        h.max_bodies = (u32)k;
        // a grid for this delta (one outer tuple per thread); a body whose
        // delta outgrows it 8x ends the chunk, so the next chunk is launched
        // wider
        const u64 want = (nd + LB - 1) / LB;
        const unsigned g = (unsigned)std::min<u64>(grid_cap(ctx), std::max<u64>(want, 16));
        h.delta_limit = g >= grid_cap(ctx) ? ~0ull : (u64)g * LB * 8;
        // a delta that turns dense ends the chunk too (bit-matrix bodies next)
        if (dense_at != ~0ull) h.delta_limit = std::min<u64>(h.delta_limit, dense_at - 1);
        Staging stg(ctx);
        GDL_TRY(stg.up(st, &h, sizeof(h)));
        for (int i = 0; i < k; ++i) {
            KTimer kt(ctx, K_LINEAR, 0.0);
            hipLaunchKernelGGL(k_linear_body<LB>, dim3(g), dim3(LB), 0, ctx->stream,
                               (const IdxSlot *)inner_idx->slots, idx_arg(inner_idx), d_inner,
                               n_inner, shift, map.arity, map.src[0],
                               map.arity == 2 ? map.src[1] : 0, pred.op, pred.lhs, pred.rhs,
                               (u64 *)hs->slots, hs_arg(hs), *d_log, st, counts);
            GDL_TRY(check_launch(ctx, "k_linear_body"));
        }
        LinState r{};
        u64 c[2 * kChunk];
        u32 dev_err = 0;
        GDL_TRY(stg.later(&r, st, sizeof(r)));
        GDL_TRY(stg.later(c, counts, sizeof(c)));
        GDL_TRY(stg.later(&dev_err, ctx->d_err, sizeof(u32)));
        GDL_TRY(stg.land());  // one wait for the state, the counts and the error word
        GDL_TRY(device_error_status(ctx, dev_err));
        if (r.stop == kLinOverflow)
            return fail(ctx, GDL_ERR_OVERFLOW, "linear fixpoint: a body outgrew its log bound");
        u64 outer_n = nd;
        for (u32 b = 0; b < r.body && done < max_bodies; ++b, ++done) {
            h_new[done] = c[2 * b];
            h_raw[done] = c[2 * b + 1];
            // probe + outer read per outer tuple, inner read + hash probe per
            // produced tuple, the new keys written
            ctx->stats[K_LINEAR].alg_bytes += 72.0 * outer_n + 16.0 * c[2 * b + 1] + 8.0 * c[2 * b];
            outer_n = c[2 * b];
        }
        hs->size = r.size;
        d0 = r.delta_begin;
        d1 = r.delta_end;
        if (r.stop == kLinFixpoint || r.body == 0) break;
        // chunks double while every enqueued body runs (a stopped chunk's
        // remaining launches return at once, but still cost a launch)
        ramp = (int)r.body == k ? std::min(2 * ramp, kChunk) : std::max(4, (int)r.body);
    }
    *bodies = done;
    *delta_begin = d0;
    *delta_end = d1;
    return GDL_OK;
}

gdl_status gdl_project_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, gdl_colmap map,
                           uint64_t *d_out) {
    if (!ctx || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    if (!valid_map(map, false)) return fail(ctx, GDL_ERR_INVALID_ARG, "bad copy column map");
    if (n == 0) return GDL_OK;
    KTimer kt(ctx, K_PROJECT, 16.0 * n);
    hipLaunchKernelGGL(k_project, dim3(stride_grid(ctx, n)), dim3(JB), 0, ctx->stream, d_in, n,
                       map.arity, map.src[0], map.arity == 2 ? map.src[1] : 0, d_out);
    return check_launch(ctx, "k_project");
}

}  // extern "C"
