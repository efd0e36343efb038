// scan.hip -- device-wide exclusive scans and stream compaction (unique,
// filter).  Tiles of 4096 items: 256 threads, each wave owns 1024 consecutive
// items processed in 16 rounds of 64 lanes, so every round's loads and the
// compacted stores are contiguous across the wave.
#include "gdl_internal.hpp"

namespace gdl {

constexpr int TB = 256;
constexpr int TI = 16;
constexpr int TILE = TB * TI;
constexpr int WSPAN = 64 * TI;

__device__ __forceinline__ u64 tile_item(u64 base, int w, int r, int lane) {
    return base + (u64)w * WSPAN + (u64)r * 64 + lane;
}

__device__ __forceinline__ u64 wave_bcast_last(u64 v) { return __shfl(v, 63, 64); }

// Sum of a tile -> block_sums[blockIdx.x]
template <typename T>
__global__ __launch_bounds__(TB) void k_scan_reduce(const T *in, u64 n, u64 *block_sums) {
    __shared__ u64 s_w[TB / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const u64 base = (u64)blockIdx.x * TILE;
    u64 acc = 0;
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        u64 i = tile_item(base, w, r, lane);
        if (i < n) acc += (u64)in[i];
    }
    u64 tot;
    block_exclusive_scan<TB>(acc, s_w, &tot);
    if (t == 0) block_sums[blockIdx.x] = tot;
}

// Exclusive scan in place by one block per array (block 0: v0, block 1: v1);
// v[n] = total, and block 0 also stores it to *total_out.  Each thread scans
// 4 consecutive elements per round of 4 * BLOCK.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_scan_top(u64 *v0, u64 *v1, u64 n, u64 *total_out) {
    __shared__ u64 s_w[BLOCK / 64 + 1];
    u64 *v = blockIdx.x ? v1 : v0;
    const int t = threadIdx.x;
    u64 carry = 0;
    for (u64 b = 0; b < n; b += 4 * BLOCK) {
        const u64 i0 = b + 4 * (u64)t;
        u64 x[4];
        u64 sum = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x[j] = i0 + j < n ? v[i0 + j] : 0;
            sum += x[j];
        }
        u64 tot;
        u64 ex = carry + block_exclusive_scan<BLOCK>(sum, s_w, &tot);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i0 + j < n) v[i0 + j] = ex;
            ex += x[j];
        }
        carry += tot;
    }
    if (t == 0) {
        v[n] = carry;
        if (total_out) total_out[blockIdx.x] = carry;
    }
}

// Exclusive scan of a tile with the block offset from block_sums.
template <typename T>
__global__ __launch_bounds__(TB) void k_scan_down(const T *in, u64 n, const u64 *block_sums,
                                                  u64 *out) {
    __shared__ u64 s_w[TB / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const u64 base = (u64)blockIdx.x * TILE;
    u64 ex[TI];
    u64 carry = 0;
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        u64 i = tile_item(base, w, r, lane);
        u64 v = i < n ? (u64)in[i] : 0;
        u64 inc = wave_inclusive_scan(v);
        ex[r] = carry + inc - v;
        carry += wave_bcast_last(inc);
    }
    if (lane == 0) s_w[w] = carry;
    __syncthreads();
    u64 woff = 0;
    for (int k = 0; k < w; ++k) woff += s_w[k];
    const u64 off = block_sums[blockIdx.x] + woff;
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        u64 i = tile_item(base, w, r, lane);
        if (i < n) out[i] = off + ex[r];
    }
}

// ---- compaction: mode 0 = unique (adjacent-different), 1 = predicate filter
struct UniqueFlag {
    __device__ __forceinline__ bool operator()(const u64 *in, u64 i, u64 k) const {
        return i == 0 || in[i - 1] != k;
    }
};
struct PredFlag {
    int op, lhs, rhs;
    __device__ __forceinline__ bool operator()(const u64 *, u64, u64 k) const {
        return eval_pred(k, op, lhs, rhs);
    }
};
// keeps in[i] where flags[i] != 0
struct ByteFlag {
    const uint8_t *flags;
    __device__ __forceinline__ bool operator()(const u64 *, u64 i, u64) const {
        return flags[i] != 0;
    }
};

// First index in [lo, hi) of sorted keys whose prefix (key >> shift) is >= p.
__device__ __forceinline__ u64 prefix_lower_bound(const u64 *keys, u64 lo, u64 hi, u64 p,
                                                  int shift) {
    while (lo < hi) {
        const u64 mid = lo + ((hi - lo) >> 1);
        if ((keys[mid] >> shift) < p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// std::set_difference(b, a) under the order of the prefix key >> shift, over
// arrays sorted by key (hence by prefix): of the m elements of b sharing a
// prefix that n elements of a have, the first min(m, n) are dropped and the
// rest kept, in order.
struct PrefixDiffFlag {
    const u64 *a;
    u64 na;
    int shift;
    __device__ __forceinline__ bool operator()(const u64 *b, u64 i, u64 k) const {
        const u64 p = k >> shift;
        const u64 first_b = prefix_lower_bound(b, 0, i, p, shift);
        const u64 lo = prefix_lower_bound(a, 0, na, p, shift);
        const u64 n = prefix_lower_bound(a, lo, na, p + 1, shift) - lo;
        return i - first_b >= n;
    }
};

template <typename F>
__global__ __launch_bounds__(TB) void k_compact_count(const u64 *in, u64 n, F flag,
                                                      u64 *block_counts) {
    __shared__ u64 s_w[TB / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const u64 base = (u64)blockIdx.x * TILE;
    u64 acc = 0;
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        u64 i = tile_item(base, w, r, lane);
        if (i < n && flag(in, i, in[i])) acc++;
    }
    u64 tot;
    block_exclusive_scan<TB>(acc, s_w, &tot);
    if (t == 0) block_counts[blockIdx.x] = tot;
}

template <typename F>
__global__ __launch_bounds__(TB) void k_compact_write(const u64 *in, u64 n, F flag,
                                                      const u64 *block_offsets, u64 *out) {
    __shared__ u64 s_w[TB / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const u64 base = (u64)blockIdx.x * TILE;
    u64 keys[TI];
    u64 bits = 0;  // flag of round r in bit r
    u64 carry = 0;
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        u64 i = tile_item(base, w, r, lane);
        u64 k = i < n ? in[i] : 0;
        keys[r] = k;
        bool f = i < n && flag(in, i, k);
        bits |= (u64)f << r;
        carry += __popcll(__ballot(f));
    }
    if (lane == 0) s_w[w] = carry;
    __syncthreads();
    u64 off = block_offsets[blockIdx.x];
    for (int k = 0; k < w; ++k) off += s_w[k];
    const u64 lt = lane_mask_lt();
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        bool f = (bits >> r) & 1;
        u64 bal = __ballot(f);
        if (f) out[off + __popcll(bal & lt)] = keys[r];
        off += __popcll(bal);
    }
}

// ----------------------------------------------------------------- host side

static gdl_status scan_top(gdl_ctx *ctx, u64 *v, u64 n, u64 *total_dev) {
    KTimer kt(ctx, K_SCAN_TOP, 16.0 * n);
    hipLaunchKernelGGL(k_scan_top<1024>, dim3(1), dim3(1024), 0, ctx->stream, v, v, n,
                       total_dev);
    return check_launch(ctx, "k_scan_top");
}

gdl_status scan_inplace2(gdl_ctx *ctx, u64 *v0, u64 *v1, u64 n, u64 *totals) {
    if (n > 262144) {
        GDL_TRY(scan_inplace(ctx, v0, n));
        GDL_TRY(scan_inplace(ctx, v1, n));
        GDL_HIP(ctx, hipMemcpyAsync(totals, v0 + n, sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        GDL_HIP(ctx, hipMemcpyAsync(totals + 1, v1 + n, sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        return GDL_OK;
    }
    KTimer kt(ctx, K_SCAN_TOP, 32.0 * n);
    hipLaunchKernelGGL(k_scan_top<1024>, dim3(2), dim3(1024), 0, ctx->stream, v0, v1, n, totals);
    return check_launch(ctx, "k_scan_top");
}

template <typename T>
static gdl_status exclusive_scan_impl(gdl_ctx *ctx, const T *d_in, u64 n, u64 *d_out,
                                      u64 *total_host) {
    if (n == 0) {
        GDL_HIP(ctx, hipMemsetAsync(d_out, 0, sizeof(u64), ctx->stream));
        if (total_host) {
            GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
            *total_host = 0;
        }
        return GDL_OK;
    }
    const u64 nb = (n + TILE - 1) / TILE;
    Scratch scr(ctx);
    u64 *bs;
    GDL_TRY(scr.get(nb + 1, &bs));
    {
        KTimer kt(ctx, K_SCAN_REDUCE, (double)sizeof(T) * n);
        hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(TB), 0, ctx->stream, d_in,
                           n, bs);
        GDL_TRY(check_launch(ctx, "k_scan_reduce"));
    }
    GDL_TRY(scan_top(ctx, bs, nb, d_out + n));
    {
        KTimer kt(ctx, K_SCAN_DOWN, (double)(sizeof(T) + 8) * n);
        hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned)nb), dim3(TB), 0, ctx->stream, d_in, n,
                           bs, d_out);
        GDL_TRY(check_launch(ctx, "k_scan_down"));
    }
    if (total_host) GDL_TRY(read_u64(ctx, d_out + n, total_host));
    return GDL_OK;
}

gdl_status exclusive_scan_u32_to_u64(gdl_ctx *ctx, const u32 *d_in, u64 n, u64 *d_out,
                                     u64 *total_host) {
    return exclusive_scan_impl<u32>(ctx, d_in, n, d_out, total_host);
}

gdl_status exclusive_scan_u64(gdl_ctx *ctx, const u64 *d_in, u64 n, u64 *d_out) {
    return exclusive_scan_impl<u64>(ctx, d_in, n, d_out, nullptr);
}

gdl_status scan_inplace(gdl_ctx *ctx, u64 *v, u64 n) {
    // one block up to 2^15 entries (a few microseconds); a 1024-thread block
    // over 2^18 of them takes ~150 us, the three-pass scan a few launches
    if (n <= 32768) return scan_top(ctx, v, n, nullptr);
    Scratch scr(ctx);
    u64 *tmp;
    GDL_TRY(scr.get(n + 1, &tmp));
    GDL_TRY(exclusive_scan_impl<u64>(ctx, v, n, tmp, nullptr));
    GDL_HIP(ctx, hipMemcpyAsync(v, tmp, (n + 1) * sizeof(u64), hipMemcpyDeviceToDevice,
                                ctx->stream));
    return GDL_OK;
}

template <typename F>
static gdl_status compact_impl(gdl_ctx *ctx, const u64 *d_in, u64 n, F flag, u64 *d_out,
                               u64 *n_out) {
    if (n == 0) {
        *n_out = 0;
        return GDL_OK;
    }
    const u64 nb = (n + TILE - 1) / TILE;
    Scratch scr(ctx);
    u64 *bc;
    GDL_TRY(scr.get(nb + 1, &bc));
    {
        KTimer kt(ctx, K_COMPACT_COUNT, 8.0 * n);
        hipLaunchKernelGGL(k_compact_count<F>, dim3((unsigned)nb), dim3(TB), 0, ctx->stream, d_in,
                           n, flag, bc);
        GDL_TRY(check_launch(ctx, "k_compact_count"));
    }
    GDL_TRY(scan_top(ctx, bc, nb, nullptr));
    {
        KTimer kt(ctx, K_COMPACT_WRITE, 16.0 * n);
        hipLaunchKernelGGL(k_compact_write<F>, dim3((unsigned)nb), dim3(TB), 0, ctx->stream, d_in,
                           n, flag, bc, d_out);
        GDL_TRY(check_launch(ctx, "k_compact_write"));
    }
    return read_u64(ctx, bc + nb, n_out);
}

gdl_status unique_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, u64 *d_out, u64 *n_out) {
    return compact_impl(ctx, d_in, n, UniqueFlag{}, d_out, n_out);
}

gdl_status compact_flags(gdl_ctx *ctx, const u64 *d_in, const uint8_t *flags, u64 n, u64 *d_out,
                         u64 *n_out) {
    return compact_impl(ctx, d_in, n, ByteFlag{flags}, d_out, n_out);
}

gdl_status filter_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, gdl_pred pred, u64 *d_out,
                       u64 *n_out) {
    return compact_impl(ctx, d_in, n, PredFlag{pred.op, pred.lhs, pred.rhs}, d_out, n_out);
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_unique_u64(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n, uint64_t *d_out,
                          uint64_t *n_out) {
    if (!ctx || !n_out || (n && (!d_sorted || !d_out))) return GDL_ERR_INVALID_ARG;
    return unique_keys(ctx, d_sorted, n, d_out, n_out);
}

gdl_status gdl_filter_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, gdl_pred pred,
                          uint64_t *d_out, uint64_t *n_out) {
    if (!ctx || !n_out || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    if (!valid_pred(pred, 2)) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_filter_u64: bad predicate");
    return filter_keys(ctx, d_in, n, pred, d_out, n_out);
}

gdl_status gdl_compact_flags_u64(gdl_ctx *ctx, const uint64_t *d_in, const uint8_t *d_flags,
                                 uint64_t n, uint64_t *d_out, uint64_t *n_out) {
    if (!ctx || !n_out || (n && (!d_in || !d_flags || !d_out))) return GDL_ERR_INVALID_ARG;
    return compact_flags(ctx, d_in, d_flags, n, d_out, n_out);
}

gdl_status gdl_set_difference_prefix_u64(gdl_ctx *ctx, const uint64_t *d_b, uint64_t nb,
                                         const uint64_t *d_a, uint64_t na, int shift,
                                         uint64_t *d_out, uint64_t *n_out) {
    if (!ctx || !n_out || (nb && (!d_b || !d_out)) || (na && !d_a)) return GDL_ERR_INVALID_ARG;
    if (shift < 0 || shift > 63)
        return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_set_difference_prefix_u64: shift 0..63");
    return compact_impl(ctx, d_b, nb, PrefixDiffFlag{d_a, na, shift}, d_out, n_out);
}

gdl_status gdl_join_scan(gdl_ctx *ctx, const uint32_t *d_counts, uint64_t n, uint64_t *d_offsets,
                         uint64_t *total) {
    if (!ctx || !d_offsets || (n && !d_counts)) return GDL_ERR_INVALID_ARG;
    u64 tot = 0;
    GDL_TRY(exclusive_scan_u32_to_u64(ctx, d_counts, n, d_offsets, &tot));
    if (total) *total = tot;
    return GDL_OK;
}

}  // extern "C"
