// sort.hip -- LSD radix sort of u64 tuple keys for gfx950.
//
// The key (c0 << 32) | c1 is first compressed order-preservingly to
// (c0 << lo_bits) | c1 (lo_bits = bit length bound of c1), so the number of
// digit passes is ceil((hi_bits + lo_bits) / r) instead of 64 / r, for the
// context's radix width r (4..8 bits, default 8; digits narrower than 8 bits
// leave the upper histogram rows empty).  With r = 8, 9-bit digits (512
// bins, two per thread) are used where they save a pass: 36-bit keys (two
// 18-bit ids) sort in 4 passes instead of 5.
//
// Per pass over n > 4096 keys, three launches:
//   upsweep   -- per-tile digit histogram (LDS, one sub-histogram per wave),
//                written as counts[digit][tile];
//   rowscan   -- one block per digit scans its row over tiles;
//   downsweep -- stable rank of every key inside its 4096-key tile (wave64
//                ballot match per 64-key round plus per-wave running LDS
//                histograms), keys placed in LDS in digit order, then written
//                out tile-run by tile-run so consecutive lanes store
//                consecutive addresses.
// Arrays of at most 4096 keys are sorted by one launch that runs every pass in
// LDS.
#include "gdl_internal.hpp"

namespace gdl {

constexpr int RS_B = 256;           // threads per block
constexpr int RS_I = 16;            // keys per thread
constexpr int RS_TILE = RS_B * RS_I; // 4096
constexpr int RS_WSPAN = 64 * RS_I;  // keys per wave
constexpr int RADIX = 256;          // 8-bit digits (the one-block sort)

__device__ __forceinline__ u64 compress_key(u64 k, int lo_bits) {
    return ((k >> 32) << lo_bits) | (k & 0xffffffffull);
}
__device__ __forceinline__ u64 expand_key(u64 k, int lo_bits) {
    return ((k >> lo_bits) << 32) | (k & ((1ull << lo_bits) - 1ull));
}
// A pass's digit: `pass` packs the shift (bits 0..7) and the digit width
// (bits 8..11, 4..8 bits; the tuned radix width).
__device__ __forceinline__ u32 digit_of(u64 k, int pass) {
    return (u32)(k >> (pass & 0xff)) & ((1u << (pass >> 8)) - 1u);
}
__host__ __device__ __forceinline__ int pass_code(int shift, int bits) {
    return shift | (bits << 8);
}

__device__ __forceinline__ u64 tile_pos(int w, int r, int lane) {
    return (u64)w * RS_WSPAN + (u64)r * 64 + lane;
}

// Lanes of this wave that hold the same RB-bit digit (among `valid` lanes).
template <int RB>
__device__ __forceinline__ u64 match_digit(u32 d, bool valid) {
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const bool bit = (d >> b) & 1u;
        const u64 bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Ranks the 16 keys a thread holds (wave-contiguous layout) within the tile,
// for digits of up to RB bits (2^RB bins, 2^RB / 256 per thread): on return
// s_hist[w][d] = exclusive offset of wave w's keys of digit d inside the
// tile's digit-d run, s_start[d] = start of digit d's run in the tile, and
// rank[r] = rank of key r among its wave's keys of the same digit.
template <int RB>
__device__ __forceinline__ void rank_tile(const u64 *keys, const bool *valid, u32 *rank, int shift,
                                          u32 (*s_hist)[1 << RB], u32 *s_start, u64 *s_scan) {
    constexpr int R = 1 << RB, DPT = R / RS_B;
    const int t = threadIdx.x, w = t >> 6;
    const u64 lt = lane_mask_lt();
    for (int x = t; x < 4 * R; x += RS_B) (&s_hist[0][0])[x] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const u32 d = digit_of(keys[r], shift);
        const u64 peers = match_digit<RB>(d, valid[r]);
        const u32 before = s_hist[w][d];
        const u32 within = (u32)__popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid[r] && within == 0) s_hist[w][d] = before + (u32)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[r] = before + within;
    }
    __syncthreads();
    // thread t owns digits t * DPT .. t * DPT + DPT - 1 (consecutive)
    u32 c[DPT][4];
    u32 sum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const int d = t * DPT + j;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            c[j][v] = s_hist[v][d];
            sum += c[j][v];
        }
    }
    u64 tot;
    u32 start = (u32)block_exclusive_scan<RS_B>((u64)sum, s_scan, &tot);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const int d = t * DPT + j;
        s_start[d] = start;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            s_hist[v][d] = start;
            start += c[j][v];
        }
    }
    __syncthreads();
}

template <bool IN_X, int RB>
__global__ __launch_bounds__(RS_B) void k_rs_upsweep(const u64 *in, u64 n, int shift, int lo_bits,
                                                     u32 *counts, u32 nb) {
    constexpr int R = 1 << RB;
    __shared__ u32 s_h[4][R];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int x = t; x < 4 * R; x += RS_B) (&s_h[0][0])[x] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const u64 i = base + tile_pos(w, r, lane);
        if (i < n) {
            u64 k = in[i];
            if (IN_X) k = compress_key(k, lo_bits);
            atomicAdd(&s_h[w][digit_of(k, shift)], 1u);
        }
    }
    __syncthreads();
    for (int d = t; d < R; d += RS_B)
        counts[(u64)d * nb + blockIdx.x] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
}

// One block per digit: exclusive scan of counts[d][0..nb) in place, totals[d].
__global__ __launch_bounds__(1024) void k_rs_rowscan(u32 *counts, u32 nb, u32 *totals) {
    __shared__ u64 s_w[1024 / 64 + 1];
    u32 *row = counts + (u64)blockIdx.x * nb;
    u64 carry = 0;
    for (u32 b = 0; b < nb; b += 1024) {
        const u32 i = b + threadIdx.x;
        const u64 x = i < nb ? row[i] : 0;
        u64 tot;
        const u64 ex = block_exclusive_scan<1024>(x, s_w, &tot);
        if (i < nb) row[i] = (u32)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = (u32)carry;
}

template <bool IN_X, bool OUT_X, int RB>
__global__ __launch_bounds__(RS_B) void k_rs_downsweep(const u64 *in, u64 *out, u64 n, int shift,
                                                       int lo_bits, const u32 *counts,
                                                       const u32 *totals, u32 nb, u32 *err) {
    constexpr int R = 1 << RB, DPT = R / RS_B;
    __shared__ u64 s_keys[RS_TILE];
    __shared__ u32 s_hist[4][R];
    __shared__ u32 s_start[R];
    __shared__ u32 s_gbase[R];
    __shared__ u64 s_scan[RS_B / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const u64 base = (u64)blockIdx.x * RS_TILE;

    // global start of each digit (thread t: digits t * DPT ..) plus this
    // tile's offset inside it
    u32 tl[DPT];
    u64 tsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        tl[j] = totals[t * DPT + j];
        tsum += tl[j];
    }
    u64 gtot;
    u64 gex = block_exclusive_scan<RS_B>(tsum, s_scan, &gtot);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const int d = t * DPT + j;
        s_gbase[d] = (u32)gex + counts[(u64)d * nb + blockIdx.x];
        gex += tl[j];
    }

    u64 keys[RS_I];
    bool valid[RS_I];
    u32 rank[RS_I];
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const u64 i = base + tile_pos(w, r, lane);
        valid[r] = i < n;
        u64 k = valid[r] ? in[i] : 0;
        if (IN_X) k = compress_key(k, lo_bits);
        keys[r] = k;
    }
    rank_tile<RB>(keys, valid, rank, shift, s_hist, s_start, s_scan);
#pragma unroll
    for (int r = 0; r < RS_I; ++r)
        if (valid[r]) {
            const u32 pos = s_hist[w][digit_of(keys[r], shift)] + rank[r];
            if (pos < (u32)RS_TILE) s_keys[pos] = keys[r];
        }
    __syncthreads();
    const u64 rem = n - base;
    const u32 tile_n = rem < (u64)RS_TILE ? (u32)rem : (u32)RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_I; ++j) {
        const u32 p = (u32)j * RS_B + t;
        if (p < tile_n) {
            u64 k = s_keys[p];
            const u32 d = digit_of(k, shift);
            const u32 dst = s_gbase[d] + (p - s_start[d]);
            if (OUT_X) k = expand_key(k, lo_bits);
            if (dst < n) out[dst] = k;
            else atomicOr(err, 2u);  // inconsistent digit counts: never write out of bounds
        }
    }
}

// Whole sort of n <= 4096 keys in one block: every pass runs in LDS.
// With UNIQUE, adjacent duplicates are dropped and the count is stored to
// *n_out.
template <bool XF, bool UNIQUE>
__global__ __launch_bounds__(RS_B) void k_rs_oneblock(const u64 *in, u64 *out, u32 n, int passes,
                                                      int lo_bits, u64 *n_out) {
    __shared__ u64 s_keys[RS_TILE];
    __shared__ u32 s_hist[4][RADIX];
    __shared__ u32 s_start[RADIX];
    __shared__ u64 s_scan[RS_B / 64 + 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    u64 keys[RS_I];
    bool valid[RS_I];
    u32 rank[RS_I];
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const u64 p = tile_pos(w, r, lane);
        valid[r] = p < n;
        u64 k = valid[r] ? in[p] : 0;
        if (XF) k = compress_key(k, lo_bits);
        keys[r] = k;
    }
    // passes packs the pass count (bits 0..7) and the digit width (8..11)
    const int n_pass = passes & 0xff, dbits = passes >> 8;
    for (int pass = 0; pass < n_pass; ++pass) {
        const int shift = pass_code(pass * dbits, dbits);
        rank_tile<8>(keys, valid, rank, shift, s_hist, s_start, s_scan);
#pragma unroll
        for (int r = 0; r < RS_I; ++r)
            if (valid[r]) {
                const u32 pos = s_hist[w][digit_of(keys[r], shift)] + rank[r];
                if (pos < (u32)RS_TILE) s_keys[pos] = keys[r];
            }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_I; ++r)
            if (valid[r]) keys[r] = s_keys[tile_pos(w, r, lane)];
        __syncthreads();
    }
    if (!UNIQUE) {
#pragma unroll
        for (int r = 0; r < RS_I; ++r) {
            if (valid[r]) out[tile_pos(w, r, lane)] = XF ? expand_key(keys[r], lo_bits) : keys[r];
        }
        return;
    }
    // keys[r] sits at tile position tile_pos(w, r, lane); s_keys holds the
    // sorted tile when passes > 0, otherwise reload it.
    if (n_pass == 0) {
#pragma unroll
        for (int r = 0; r < RS_I; ++r)
            if (valid[r]) s_keys[tile_pos(w, r, lane)] = keys[r];
        __syncthreads();
    }
    const u64 lt = lane_mask_lt();
    u64 carry = 0;
    u32 bits = 0;
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const u64 p = tile_pos(w, r, lane);
        const bool f = valid[r] && (p == 0 || s_keys[p - 1] != keys[r]);
        bits |= (u32)f << r;
        carry += __popcll(__ballot(f));
    }
    if (lane == 0) s_scan[w] = carry;
    __syncthreads();
    u64 off = 0;
    for (int k = 0; k < w; ++k) off += s_scan[k];
#pragma unroll
    for (int r = 0; r < RS_I; ++r) {
        const bool f = (bits >> r) & 1u;
        const u64 bal = __ballot(f);
        if (f) out[off + __popcll(bal & lt)] = XF ? expand_key(keys[r], lo_bits) : keys[r];
        off += __popcll(bal);
    }
    if (t == RS_B - 1) *n_out = off;  // wave 3's running offset ends at the total
}

// Bit lengths of the two 32-bit halves: out[0] |= c0 of every key, out[1] |= c1.
__global__ __launch_bounds__(256) void k_key_bits(const u64 *in, u64 n, u64 *out) {
    u64 hi = 0, lo = 0;
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < n; i += (u64)gridDim.x * 256) {
        const u64 k = in[i];
        hi |= k >> 32;
        lo |= k & 0xffffffffull;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        hi |= __shfl_xor(hi, d, 64);
        lo |= __shfl_xor(lo, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicOr((unsigned long long *)&out[0], (unsigned long long)hi);
        atomicOr((unsigned long long *)&out[1], (unsigned long long)lo);
    }
}

static int bit_length(u64 v) { return v ? 64 - __builtin_clzll(v) : 0; }

static gdl_status measure_bits(gdl_ctx *ctx, const u64 *d_in, u64 n, int *hi_bits, int *lo_bits) {
    Scratch scr(ctx);
    u64 *acc;
    GDL_TRY(scr.get(2, &acc));
    GDL_HIP(ctx, hipMemsetAsync(acc, 0, 2 * sizeof(u64), ctx->stream));
    unsigned grid = (unsigned)std::min<u64>(grid_for(n, 256 * 8), 2048);
    {
        KTimer kt(ctx, K_KEY_BITS, 8.0 * n);
        hipLaunchKernelGGL(k_key_bits, dim3(grid), dim3(256), 0, ctx->stream, d_in, n, acc);
        GDL_TRY(check_launch(ctx, "k_key_bits"));
    }
    u64 h[2];
    GDL_TRY(read_u64(ctx, acc, h, 2));
    *hi_bits = bit_length(h[0]);
    *lo_bits = bit_length(h[1]);
    return GDL_OK;
}

// The multi-tile LSD passes over n > 4096 keys with digits of dbits <= RB
// bits (2^RB histogram bins); the last pass lands in d_out.
template <int RB>
static gdl_status tile_passes(gdl_ctx *ctx, const u64 *d_in, u64 n, int lo_bits, bool xf,
                              int dbits, int passes, u64 *d_out) {
    constexpr int R = 1 << RB;
    const u32 nb = (u32)((n + RS_TILE - 1) / RS_TILE);
    Scratch scr(ctx);
    u32 *counts, *totals;
    u64 *tmp = nullptr;
    GDL_TRY(scr.get((u64)R * nb, &counts));
    GDL_TRY(scr.get(R, &totals));
    if (passes > 1) GDL_TRY(scr.get(n, &tmp));
    const u64 *src = d_in;
    for (int p = 0; p < passes; ++p) {
        u64 *dst = ((passes - 1 - p) % 2 == 0) ? d_out : tmp;
        const int shift = pass_code(p * dbits, dbits);
        const bool in_x = xf && p == 0;
        const bool out_x = xf && p == passes - 1;
        {
            KTimer kt(ctx, K_RS_UPSWEEP, 8.0 * n);
            if (in_x)
                hipLaunchKernelGGL((k_rs_upsweep<true, RB>), dim3(nb), dim3(RS_B), 0, ctx->stream,
                                   src, n, shift, lo_bits, counts, nb);
            else
                hipLaunchKernelGGL((k_rs_upsweep<false, RB>), dim3(nb), dim3(RS_B), 0, ctx->stream,
                                   src, n, shift, lo_bits, counts, nb);
            GDL_TRY(check_launch(ctx, "k_rs_upsweep"));
        }
        {
            KTimer kt(ctx, K_RS_ROWSCAN, 8.0 * R * nb);
            hipLaunchKernelGGL(k_rs_rowscan, dim3(R), dim3(1024), 0, ctx->stream, counts, nb,
                               totals);
            GDL_TRY(check_launch(ctx, "k_rs_rowscan"));
        }
        {
            KTimer kt(ctx, K_RS_DOWNSWEEP, 16.0 * n);
            if (in_x && out_x)
                hipLaunchKernelGGL((k_rs_downsweep<true, true, RB>), dim3(nb), dim3(RS_B), 0,
                                   ctx->stream, src, dst, n, shift, lo_bits, counts, totals, nb,
                                   ctx->d_err);
            else if (in_x)
                hipLaunchKernelGGL((k_rs_downsweep<true, false, RB>), dim3(nb), dim3(RS_B), 0,
                                   ctx->stream, src, dst, n, shift, lo_bits, counts, totals, nb,
                                   ctx->d_err);
            else if (out_x)
                hipLaunchKernelGGL((k_rs_downsweep<false, true, RB>), dim3(nb), dim3(RS_B), 0,
                                   ctx->stream, src, dst, n, shift, lo_bits, counts, totals, nb,
                                   ctx->d_err);
            else
                hipLaunchKernelGGL((k_rs_downsweep<false, false, RB>), dim3(nb), dim3(RS_B), 0,
                                   ctx->stream, src, dst, n, shift, lo_bits, counts, totals, nb,
                                   ctx->d_err);
            GDL_TRY(check_launch(ctx, "k_rs_downsweep"));
        }
        src = dst;
    }
    return GDL_OK;
}

// Shared driver of sort and sort+unique.  For n <= 4096 with `unique`,
// *n_unique receives the unique count (the caller then skips unique_keys).
static gdl_status sort_driver(gdl_ctx *ctx, const u64 *d_in, u64 n, int hi_bits, int lo_bits,
                              u64 *d_out, bool unique, u64 *n_unique, bool *uniqued) {
    *uniqued = false;
    if (n == 0) return GDL_OK;
    if (n >= 0xffffffffull) return fail(ctx, GDL_ERR_OVERFLOW, "sort: more than 2^32-1 keys");
    if (hi_bits < 0 || lo_bits < 0) GDL_TRY(measure_bits(ctx, d_in, n, &hi_bits, &lo_bits));
    if (hi_bits > 32 || lo_bits > 32)
        return fail(ctx, GDL_ERR_INVALID_ARG, "sort: key bit bounds above 32");
    const bool xf = lo_bits < 32;
    const int total_bits = hi_bits + (xf ? lo_bits : 32);
    const int dbits = ctx->tune.radix_bits;  // 4..8
    const int passes = (total_bits + dbits - 1) / dbits;
    const int pcode = passes | (dbits << 8);

    if (n <= (u64)RS_TILE) {
        KTimer kt(ctx, K_RS_ONEBLOCK, 16.0 * n);
        u64 *cnt = nullptr;
        Scratch scr(ctx);
        if (unique) GDL_TRY(scr.get(1, &cnt));
        if (xf) {
            if (unique)
                hipLaunchKernelGGL((k_rs_oneblock<true, true>), dim3(1), dim3(RS_B), 0, ctx->stream,
                                   d_in, d_out, (u32)n, pcode, lo_bits, cnt);
            else
                hipLaunchKernelGGL((k_rs_oneblock<true, false>), dim3(1), dim3(RS_B), 0,
                                   ctx->stream, d_in, d_out, (u32)n, pcode, lo_bits, cnt);
        } else {
            if (unique)
                hipLaunchKernelGGL((k_rs_oneblock<false, true>), dim3(1), dim3(RS_B), 0,
                                   ctx->stream, d_in, d_out, (u32)n, pcode, lo_bits, cnt);
            else
                hipLaunchKernelGGL((k_rs_oneblock<false, false>), dim3(1), dim3(RS_B), 0,
                                   ctx->stream, d_in, d_out, (u32)n, pcode, lo_bits, cnt);
        }
        GDL_TRY(check_launch(ctx, "k_rs_oneblock"));
        if (unique) {
            GDL_TRY(read_u64(ctx, cnt, n_unique));
            *uniqued = true;
        }
        return GDL_OK;
    }

    if (passes == 0) {  // every key equal
        GDL_HIP(ctx, hipMemcpyAsync(d_out, d_in, n * sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        return GDL_OK;
    }
    // nine-bit digits where they save a pass (36-bit keys: 4 passes, not 5;
    // 18-bit keys: 2, not 3)
    if (dbits == 8 && (total_bits + 8) / 9 < passes)
        return tile_passes<9>(ctx, d_in, n, lo_bits, xf, 9, (total_bits + 8) / 9, d_out);
    return tile_passes<8>(ctx, d_in, n, lo_bits, xf, dbits, passes, d_out);
}

gdl_status sort_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, int hi_bits, int lo_bits, u64 *d_out) {
    u64 dummy;
    bool uniqued;
    return sort_driver(ctx, d_in, n, hi_bits, lo_bits, d_out, false, &dummy, &uniqued);
}

gdl_status sort_unique_keys(gdl_ctx *ctx, const u64 *d_in, u64 n, int hi_bits, int lo_bits,
                            u64 *d_out, u64 *n_out) {
    if (n == 0) {
        *n_out = 0;
        return GDL_OK;
    }
    bool uniqued = false;
    if (n <= (u64)RS_TILE)
        return sort_driver(ctx, d_in, n, hi_bits, lo_bits, d_out, true, n_out, &uniqued);
    Scratch scr(ctx);
    u64 *tmp;
    GDL_TRY(scr.get(n, &tmp));
    GDL_TRY(sort_driver(ctx, d_in, n, hi_bits, lo_bits, tmp, false, n_out, &uniqued));
    return unique_keys(ctx, tmp, n, d_out, n_out);
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_sort_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, int hi_bits, int lo_bits,
                        uint64_t *d_out) {
    if (!ctx || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    if (n && d_in == d_out) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_sort_u64: d_in == d_out");
    return sort_keys(ctx, d_in, n, hi_bits, lo_bits, d_out);
}

gdl_status gdl_key_bits(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, int *hi_bits,
                        int *lo_bits) {
    if (!ctx || !hi_bits || !lo_bits || (n && !d_keys)) return GDL_ERR_INVALID_ARG;
    if (n == 0) {
        *hi_bits = *lo_bits = 0;
        return GDL_OK;
    }
    return measure_bits(ctx, d_keys, n, hi_bits, lo_bits);
}

gdl_status gdl_sort_unique_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, int hi_bits,
                               int lo_bits, uint64_t *d_out, uint64_t *n_out) {
    if (!ctx || !n_out || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    if (n && d_in == d_out)
        return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_sort_unique_u64: d_in == d_out");
    return sort_unique_keys(ctx, d_in, n, hi_bits, lo_bits, d_out, n_out);
}

}  // extern "C"
