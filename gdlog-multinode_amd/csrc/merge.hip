// merge.hip -- merge path over two sorted duplicate-free key arrays
// A (full) and B (newt) with A-first tie order.  One pass yields
//   U = A u B   (flush_delta's merge, newt accumulation) and
//   D = B \ A   (the set difference of the semi-naive step).
// A B-key is a duplicate exactly when the A-key consumed just before it is
// equal (equal keys are adjacent in merge order and A goes first).
//
// Two launches per merge plus a scan of per-tile counts: count (per 2048-key
// tile of the merged sequence: |U|, |D|) and write (LDS-staged, coalesced).
#include "gdl_internal.hpp"

namespace gdl {

constexpr int MG_B = 256;
constexpr int MG_I = 8;
constexpr int MG_TILE = MG_B * MG_I;

// Number of A-keys among the first `diag` keys of the merge (A first on ties).
__device__ __forceinline__ u64 merge_path(const u64 *A, u64 na, const u64 *B, u64 nb, u64 diag) {
    u64 lo = diag > nb ? diag - nb : 0;
    u64 hi = diag < na ? diag : na;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (A[mid] <= B[diag - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

struct TileSplit {
    u64 i0, j0;  // start of the tile in A and B
    u32 la, lb;  // lengths in A and B
};

// Loads the tile's A and B ranges into s (A at [0, la), B at [la, la + lb)),
// and the A-key before the tile into *s_prev (s_has_prev = 0 if none).
__device__ __forceinline__ TileSplit load_tile(const u64 *A, u64 na, const u64 *B, u64 nb, u64 *s,
                                               u64 *s_split, u64 *s_prev, u32 *s_has_prev,
                                               u32 *err) {
    const int t = threadIdx.x;
    const u64 d0 = (u64)blockIdx.x * MG_TILE;
    const u64 tot = na + nb;
    const u64 d1 = d0 + MG_TILE < tot ? d0 + MG_TILE : tot;
    if (t == 0) s_split[0] = merge_path(A, na, B, nb, d0);
    if (t == 64) s_split[1] = merge_path(A, na, B, nb, d1);
    __syncthreads();
    TileSplit sp;
    sp.i0 = s_split[0];
    const u64 i1 = s_split[1];
    sp.j0 = d0 - sp.i0;
    // sorted inputs give i0 <= i1 <= i0 + (d1 - d0); anything else is treated
    // as an empty tile so no LDS or global index leaves its buffer
    if (i1 < sp.i0 || i1 - sp.i0 > d1 - d0) {
        if (t == 0) atomicOr(err, 4u);
        sp.la = 0;
        sp.lb = 0;
        __syncthreads();
        return sp;
    }
    sp.la = (u32)(i1 - sp.i0);
    sp.lb = (u32)((d1 - i1) - sp.j0);
    for (u32 x = t; x < sp.la; x += MG_B) s[x] = A[sp.i0 + x];
    for (u32 x = t; x < sp.lb; x += MG_B) s[sp.la + x] = B[sp.j0 + x];
    if (t == 0) {
        *s_has_prev = sp.i0 > 0 ? 1u : 0u;
        *s_prev = sp.i0 > 0 ? A[sp.i0 - 1] : 0;
    }
    __syncthreads();
    return sp;
}

// Serial merge of this thread's MG_I merge steps inside the tile.  Bit r of
// *is_u / *is_d says output r is a union / difference element; vals[r] holds it.
__device__ __forceinline__ void thread_merge(const u64 *s, const TileSplit &sp, u64 prev,
                                             bool has_prev, u64 *vals, u32 *is_u, u32 *is_d,
                                             u32 *nu, u32 *nd) {
    const u32 t = threadIdx.x;
    const u32 tile_len = sp.la + sp.lb;
    const u32 k0 = t * MG_I;
    *is_u = 0;
    *is_d = 0;
    *nu = 0;
    *nd = 0;
    if (k0 >= tile_len) return;
    const u32 steps = tile_len - k0 < (u32)MG_I ? tile_len - k0 : (u32)MG_I;
    const u64 *sA = s;
    const u64 *sB = s + sp.la;
    // local merge path
    u32 lo = k0 > sp.lb ? k0 - sp.lb : 0;
    u32 hi = k0 < sp.la ? k0 : sp.la;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (sA[mid] <= sB[k0 - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    u32 ia = lo, ib = k0 - lo;
    u64 last_a = ia > 0 ? sA[ia - 1] : prev;
    bool has_last = ia > 0 || has_prev;
#pragma unroll
    for (u32 r = 0; r < (u32)MG_I; ++r) {
        if (r >= steps) break;
        const bool take_a = ib >= sp.lb || (ia < sp.la && sA[ia] <= sB[ib]);
        if (take_a) {
            last_a = sA[ia++];
            has_last = true;
            vals[r] = last_a;
            *is_u |= 1u << r;
            ++*nu;
        } else {
            const u64 v = sB[ib++];
            vals[r] = v;
            if (!(has_last && last_a == v)) {
                *is_u |= 1u << r;
                *is_d |= 1u << r;
                ++*nu;
                ++*nd;
            }
        }
    }
}

__global__ __launch_bounds__(MG_B) void k_merge_count(const u64 *A, u64 na, const u64 *B, u64 nb,
                                                      u64 *tile_u, u64 *tile_d, u32 *err) {
    __shared__ u64 s[MG_TILE];
    __shared__ u64 s_split[2];
    __shared__ u64 s_prev;
    __shared__ u32 s_has_prev;
    __shared__ u64 s_w[MG_B / 64 + 1];
    const TileSplit sp = load_tile(A, na, B, nb, s, s_split, &s_prev, &s_has_prev, err);
    u64 vals[MG_I];
    u32 iu, id, nu, nd;
    thread_merge(s, sp, s_prev, s_has_prev != 0, vals, &iu, &id, &nu, &nd);
    u64 tu, td;
    block_exclusive_scan<MG_B>(nu, s_w, &tu);
    block_exclusive_scan<MG_B>(nd, s_w, &td);
    if (threadIdx.x == 0) {
        tile_u[blockIdx.x] = tu;
        tile_d[blockIdx.x] = td;
    }
}

// off_u / off_d: exclusive scans of the tile counts.  U or D may be null.
__global__ __launch_bounds__(MG_B) void k_merge_write(const u64 *A, u64 na, const u64 *B, u64 nb,
                                                      const u64 *off_u, const u64 *off_d, u64 *U,
                                                      u64 *D, u32 *err) {
    __shared__ u64 s[MG_TILE];
    __shared__ u64 s_split[2];
    __shared__ u64 s_prev;
    __shared__ u32 s_has_prev;
    __shared__ u64 s_w[MG_B / 64 + 1];
    const TileSplit sp = load_tile(A, na, B, nb, s, s_split, &s_prev, &s_has_prev, err);
    u64 vals[MG_I];
    u32 iu, id, nu, nd;
    thread_merge(s, sp, s_prev, s_has_prev != 0, vals, &iu, &id, &nu, &nd);
    u64 tu, td;
    const u64 pu = block_exclusive_scan<MG_B>(nu, s_w, &tu);
    const u64 pd = block_exclusive_scan<MG_B>(nd, s_w, &td);
    // s is free again (block_exclusive_scan ended with a barrier)
    if (U) {
        u32 k = 0;
#pragma unroll
        for (int r = 0; r < MG_I; ++r)
            if ((iu >> r) & 1u) s[pu + k++] = vals[r];
        __syncthreads();
        const u64 base = off_u[blockIdx.x];
        for (u32 x = threadIdx.x; x < (u32)tu; x += MG_B) U[base + x] = s[x];
        __syncthreads();
    }
    if (D) {
        u32 k = 0;
#pragma unroll
        for (int r = 0; r < MG_I; ++r)
            if ((id >> r) & 1u) s[pd + k++] = vals[r];
        __syncthreads();
        const u64 base = off_d[blockIdx.x];
        for (u32 x = threadIdx.x; x < (u32)td; x += MG_B) D[base + x] = s[x];
    }
}

static gdl_status merge_impl(gdl_ctx *ctx, const u64 *A, u64 na, const u64 *B, u64 nb, u64 *U,
                             u64 *n_u, u64 *D, u64 *n_d) {
    const u64 tot = na + nb;
    if (tot == 0) {
        if (n_u) *n_u = 0;
        if (n_d) *n_d = 0;
        return GDL_OK;
    }
    const u64 nt = (tot + MG_TILE - 1) / MG_TILE;
    if (nt > 0xffffffffull) return fail(ctx, GDL_ERR_OVERFLOW, "merge: too many tiles");
    Scratch scr(ctx);
    u64 *ou, *od, *totals;
    GDL_TRY(scr.get(nt + 1, &ou));
    GDL_TRY(scr.get(nt + 1, &od));
    GDL_TRY(scr.get(2, &totals));
    {
        KTimer kt(ctx, K_MERGE_COUNT, 8.0 * tot);
        hipLaunchKernelGGL(k_merge_count, dim3((unsigned)nt), dim3(MG_B), 0, ctx->stream, A, na, B,
                           nb, ou, od, ctx->d_err);
        GDL_TRY(check_launch(ctx, "k_merge_count"));
    }
    GDL_TRY(scan_inplace2(ctx, ou, od, nt, totals));
    {
        // algorithmic bytes: read A and B here; the U / D writes are added
        // below once their sizes are known
        KTimer kt(ctx, K_MERGE_WRITE, 8.0 * tot);
        hipLaunchKernelGGL(k_merge_write, dim3((unsigned)nt), dim3(MG_B), 0, ctx->stream, A, na, B,
                           nb, ou, od, U, D, ctx->d_err);
        GDL_TRY(check_launch(ctx, "k_merge_write"));
    }
    u64 h[2];
    GDL_TRY(read_u64(ctx, totals, h, 2));
    ctx->stats[K_MERGE_WRITE].alg_bytes += (U ? 8.0 * h[0] : 0.0) + (D ? 8.0 * h[1] : 0.0);
    if (n_u) *n_u = h[0];
    if (n_d) *n_d = h[1];
    return GDL_OK;
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_merge_union_u64(gdl_ctx *ctx, const uint64_t *d_a, uint64_t na, const uint64_t *d_b,
                               uint64_t nb, uint64_t *d_out, uint64_t *n_out) {
    if (!ctx || !n_out || (na && !d_a) || (nb && !d_b) || ((na + nb) && !d_out))
        return GDL_ERR_INVALID_ARG;
    return merge_impl(ctx, d_a, na, d_b, nb, d_out, n_out, nullptr, nullptr);
}

gdl_status gdl_set_difference_u64(gdl_ctx *ctx, const uint64_t *d_b, uint64_t nb,
                                  const uint64_t *d_a, uint64_t na, uint64_t *d_out,
                                  uint64_t *n_out) {
    if (!ctx || !n_out || (na && !d_a) || (nb && !d_b) || (nb && !d_out))
        return GDL_ERR_INVALID_ARG;
    if (nb == 0) {
        *n_out = 0;
        return GDL_OK;
    }
    return merge_impl(ctx, d_a, na, d_b, nb, nullptr, nullptr, d_out, n_out);
}

gdl_status gdl_merge_diff_u64(gdl_ctx *ctx, const uint64_t *d_full, uint64_t nf,
                              const uint64_t *d_newt, uint64_t nn, uint64_t *d_full_out,
                              uint64_t *n_full_out, uint64_t *d_delta, uint64_t *n_delta) {
    if (!ctx || !n_delta || (nf && !d_full) || (nn && (!d_newt || !d_delta)))
        return GDL_ERR_INVALID_ARG;
    if (d_full_out && !n_full_out) return GDL_ERR_INVALID_ARG;
    return merge_impl(ctx, d_full, nf, d_newt, nn, d_full_out, n_full_out, d_delta, n_delta);
}

}  // extern "C"
