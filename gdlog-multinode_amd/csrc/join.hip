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

__global__ __launch_bounds__(JB) void k_idx_count(const u64 *keys, u64 n, int shift, u64 *out) {
    u64 acc = 0;
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 p = keys[i] >> shift;
        if (i == 0 || (keys[i - 1] >> shift) != p) ++acc;
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long *)out, (unsigned long long)acc);
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

__global__ __launch_bounds__(JB) void k_project(const u64 *in, u64 n, int arity, int s0, int s1,
                                                u64 *out) {
    for (u64 i = (u64)blockIdx.x * JB + threadIdx.x; i < n; i += (u64)gridDim.x * JB) {
        const u64 k = in[i];
        out[i] = apply_map(k, k, arity, s0, s1);
    }
}

static unsigned stride_grid(gdl_ctx *ctx, u64 n) {
    const u64 cap = (u64)(ctx->cu_count > 0 ? ctx->cu_count : 256) * 8;
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
    if (n == 0) return GDL_OK;
    const int shift = prefix_cols >= 2 ? 0 : 32;
    u64 distinct = 0;
    {
        Scratch scr(ctx);
        u64 *cnt;
        GDL_TRY(scr.get(1, &cnt));
        GDL_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(u64), ctx->stream));
        KTimer kt(ctx, K_IDX_COUNT, 8.0 * n);
        hipLaunchKernelGGL(k_idx_count, dim3(stride_grid(ctx, n)), dim3(JB), 0, ctx->stream, keys,
                           n, shift, cnt);
        GDL_TRY(check_launch(ctx, "k_idx_count"));
        GDL_TRY(read_u64(ctx, cnt, &distinct));
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
                       (const IdxSlot *)idx->slots, idx->n_slots, inner, n_inner, shift, outer,
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
                       (const IdxSlot *)idx->slots, idx->n_slots, d_sorted, n_sorted, shift,
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
    if (pred.op < GDL_PRED_NONE || pred.op > GDL_PRED_FALSE)
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate");
    if (pred.op != GDL_PRED_NONE && pred.op != GDL_PRED_FALSE &&
        (pred.lhs < 0 || pred.lhs >= map.arity || pred.rhs < 0 || pred.rhs >= map.arity))
        return fail(ctx, GDL_ERR_INVALID_ARG, "bad join predicate columns");
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
