// label.hip -- per-tuple labels of an append log, carried into sorted order.
//
// A relation kept as an append log holds its iterations' deltas one after
// another, so the iteration that first derived a tuple follows from its log
// position.  The reference's path_length program (test/path_length.cu:49-124)
// carries that value as a dependent column (newt[2] = outer[2] + 1, compared
// out of the set difference): a hop count equal to one plus the log segment
// of the pair.  These two passes produce it without a third tuple column:
//   k_segment_labels  -- label[i] = first + (segment of log position i), by a
//                        binary search over the segment bounds (L2-resident)
//   k_scatter_by_key  -- for each log key, its position in the sorted FULL
//                        (binary search), and the label written there.
// Both are latency-bound binary searches; per tuple they read 8 B of key and
// write 4 B of label (the scatter also reads 4 B of label).  Where the ids
// leave room, the label instead rides in the low bits of the key through the
// sort (gdl_segment_labels_sorted): k_segment_fold writes
// (c0 << 32) | (c1 << L) | label, the radix sort orders it (the keys are
// distinct, so by key), and k_unfold splits it; no binary search over the
// sorted keys.
#include "gdl_internal.hpp"

namespace gdl {

constexpr int LB_B = 256;

static unsigned label_grid(gdl_ctx *ctx, u64 n) {
    const u64 want = (n + LB_B - 1) / LB_B;
    return (unsigned)std::max<u64>(1, std::min<u64>(want, grid_cap(ctx, 4)));
}

// bounds[0..n_seg]: segment s covers log positions [bounds[s], bounds[s+1]).
__global__ __launch_bounds__(LB_B) void k_segment_labels(const u64 *bounds, u32 n_seg, u32 first,
                                                        u32 *out, u64 n) {
    for (u64 i = (u64)blockIdx.x * LB_B + threadIdx.x; i < n; i += (u64)gridDim.x * LB_B) {
        u32 lo = 0, hi = n_seg;  // last s with bounds[s] <= i, in [0, n_seg)
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (bounds[mid] <= i) lo = mid;
            else hi = mid;
        }
        out[i] = first + lo;
    }
}

__global__ __launch_bounds__(LB_B) void k_scatter_by_key(const u64 *sorted, u64 n_sorted,
                                                        const u64 *keys, const u32 *vals, u64 n,
                                                        u32 *out, u32 *err) {
    for (u64 i = (u64)blockIdx.x * LB_B + threadIdx.x; i < n; i += (u64)gridDim.x * LB_B) {
        const u64 k = keys[i];
        u64 lo = 0, hi = n_sorted;  // first position with sorted[p] >= k
        while (lo < hi) {
            const u64 mid = (lo + hi) >> 1;
            if (sorted[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        if (lo < n_sorted && sorted[lo] == k) out[lo] = vals[i];
        else atomicOr(err, 8u);  // a key missing from the sorted array
    }
}

// out[i] = (c0 << 32) | (c1 << lbits) | label of log position i.
__global__ __launch_bounds__(LB_B) void k_segment_fold(const u64 *log, const u64 *bounds, u32 n_seg,
                                                      u32 first, int lbits, u64 *out, u64 n) {
    for (u64 i = (u64)blockIdx.x * LB_B + threadIdx.x; i < n; i += (u64)gridDim.x * LB_B) {
        u32 lo = 0, hi = n_seg;
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (bounds[mid] <= i) lo = mid;
            else hi = mid;
        }
        const u64 k = log[i];
        out[i] = (k & 0xffffffff00000000ull) | ((k & 0xffffffffull) << lbits) | (u64)(first + lo);
    }
}

__global__ __launch_bounds__(LB_B) void k_unfold(const u64 *in, int lbits, u64 *keys, u32 *labels,
                                                u64 n) {
    const u64 lmask = (1ull << lbits) - 1ull;
    for (u64 i = (u64)blockIdx.x * LB_B + threadIdx.x; i < n; i += (u64)gridDim.x * LB_B) {
        const u64 k = in[i];
        keys[i] = (k & 0xffffffff00000000ull) | ((k & 0xffffffffull) >> lbits);
        labels[i] = (u32)(k & lmask);
    }
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_segment_labels_u32(gdl_ctx *ctx, const uint64_t *d_bounds, uint32_t n_seg,
                                  uint32_t first, uint32_t *d_out, uint64_t n) {
    if (!ctx || (n && (!d_bounds || !d_out || n_seg == 0))) return GDL_ERR_INVALID_ARG;
    if (n == 0) return GDL_OK;
    KTimer kt(ctx, K_LABEL, 12.0 * n);
    hipLaunchKernelGGL(k_segment_labels, dim3(label_grid(ctx, n)), dim3(LB_B), 0, ctx->stream,
                       d_bounds, n_seg, first, d_out, n);
    return check_launch(ctx, "k_segment_labels");
}

gdl_status gdl_segment_labels_sorted(gdl_ctx *ctx, const uint64_t *d_log, uint64_t n,
                                     const uint64_t *d_bounds, uint32_t n_seg, uint32_t first,
                                     int id_bits, uint64_t *d_keys, uint32_t *d_labels) {
    if (!ctx || (n && (!d_log || !d_bounds || !d_keys || !d_labels || n_seg == 0)) ||
        id_bits < 0 || id_bits > 32)
        return GDL_ERR_INVALID_ARG;
    const u64 last = (u64)first + n_seg - 1;  // the largest label
    if (last > 0xffffffffull) return GDL_ERR_INVALID_ARG;
    int lbits = 1;
    while (lbits < 32 && (last >> lbits)) ++lbits;
    if (id_bits + lbits > 32)
        return fail(ctx, GDL_ERR_UNSUPPORTED, "segment labels do not fit beside the ids");
    if (n == 0) return GDL_OK;
    Scratch scr(ctx);
    u64 *folded, *sorted;
    GDL_TRY(scr.get(n, &folded));
    GDL_TRY(scr.get(n, &sorted));
    {
        KTimer kt(ctx, K_LABEL, 16.0 * n);
        hipLaunchKernelGGL(k_segment_fold, dim3(label_grid(ctx, n)), dim3(LB_B), 0, ctx->stream,
                           d_log, d_bounds, n_seg, first, lbits, folded, n);
        GDL_TRY(check_launch(ctx, "k_segment_fold"));
    }
    GDL_TRY(sort_keys(ctx, folded, n, id_bits, id_bits + lbits, sorted));
    KTimer kt(ctx, K_LABEL, 20.0 * n);
    hipLaunchKernelGGL(k_unfold, dim3(label_grid(ctx, n)), dim3(LB_B), 0, ctx->stream, sorted,
                       lbits, d_keys, d_labels, n);
    return check_launch(ctx, "k_unfold");
}

gdl_status gdl_scatter_by_key_u32(gdl_ctx *ctx, const uint64_t *d_sorted, uint64_t n_sorted,
                                  const uint64_t *d_keys, const uint32_t *d_vals, uint64_t n,
                                  uint32_t *d_out) {
    if (!ctx || (n && (!d_sorted || !d_keys || !d_vals || !d_out))) return GDL_ERR_INVALID_ARG;
    if (n == 0) return GDL_OK;
    KTimer kt(ctx, K_SCATTER, 16.0 * n);
    hipLaunchKernelGGL(k_scatter_by_key, dim3(label_grid(ctx, n)), dim3(LB_B), 0, ctx->stream,
                       d_sorted, n_sorted, d_keys, d_vals, n, d_out, ctx->d_err);
    GDL_TRY(check_launch(ctx, "k_scatter_by_key"));
    return check_device_error(ctx);
}

}  // extern "C"
