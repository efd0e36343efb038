// dense.hip -- recursive bodies over a small id domain run on bit matrices,
// and key-ordered extraction of bitmap sets.
//
// A relation whose ids fit kb bits keeps its FULL as a bitmap set
// (hashset.hip): bit (c0 << kb) | c1, so row c0 is 2^kb consecutive bits
// (rw = 2^(kb-6) words) and bit order is key order.  When a body's delta
// is dense in that domain, the body runs on rows instead of tuples: the delta
// is a second bit matrix D (double-buffered, with a per-row bit count pop;
// a row with pop 0 is never read), and one wave builds each output row x,
// takes the set difference against FULL's row, writes the difference as the
// next delta row and appends its bits to FULL's append log as keys.  Every
// matrix access is a coalesced 512-byte row segment, so a body streams rows
// at HBM bandwidth instead of waiting on a random probe per produced tuple.
//
// Two rule shapes, the column maps of the reference's TC and SG encodings
// (test/tc.cu:80-149, test/sg.cu:95-147):
//   TC  out(x, z) :- r1(y, x), delta(y, z)
//       row x = OR of D[y] over the y with r1(y, x);
//   SG  out(x, y) :- r1(a, x), delta(a, b), r2(b, y)
//       T = OR of D[a] over the a with r1(a, x) (the deduplicated tmp row of
//       the first join, in registers), then row x = union of r2's successor
//       lists of the b in T (a bit row in LDS).
// New tuples equal the tuple-at-a-time bodies' (join, then set difference
// against FULL).  The produced count ("join result size (non dedup)",
// join.cu:230) is the same sum: for TC, the delta tuples (y, z) weighted by
// the r1 tuples keyed y (delta as given, duplicates included); for SG, the
// second join's produced tuples over the deduplicated tmp.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "gdl_internal.hpp"

namespace gdl {

constexpr int DB = 256;  // threads per block: 4 waves
constexpr int DWV = DB / 64;

// ---------------------------------------------------------------------------
// Key-ordered extraction.  Tiles of DB * 16 words: pass 1 counts each tile's
// bits, an exclusive scan of the counts gives each tile's first output
// position, pass 2 writes each tile's keys in bit order, 256 consecutive words
// a round with a block-wide scan of their bit counts.  With `minus`, rows whose
// minus_pop entry is non-zero have minus's bits removed (FULL minus a delta).
constexpr int XT_ROUNDS = 16;
constexpr u64 XT_WORDS = (u64)DB * XT_ROUNDS;

__device__ __forceinline__ u64 xt_word(const u64 *words, const u64 *minus, const u32 *minus_pop,
                                       int rs, u64 i) {
    u64 w = words[i];
    if (minus && minus_pop[i >> rs]) w &= ~minus[i];
    return w;
}

__global__ __launch_bounds__(DB) void k_xt_count(const u64 *words, u64 n_words, const u64 *minus,
                                                 const u32 *minus_pop, int rs, u64 *tile_counts) {
    __shared__ u64 s_w[DWV + 1];
    const u64 t0 = (u64)blockIdx.x * XT_WORDS;
    u64 c = 0;
#pragma unroll 4
    for (int r = 0; r < XT_ROUNDS; ++r) {
        const u64 i = t0 + (u64)r * DB + threadIdx.x;
        if (i < n_words) c += (u64)__popcll(xt_word(words, minus, minus_pop, rs, i));
    }
    u64 tot;
    block_exclusive_scan<DB>(c, s_w, &tot);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(DB) void k_xt_write(const u64 *words, u64 n_words, const u64 *minus,
                                                 const u32 *minus_pop, int rs, int kb,
                                                 const u64 *tile_offsets, u64 *out, u64 out_cap) {
    __shared__ u64 s_w[DWV + 1];
    const u64 t0 = (u64)blockIdx.x * XT_WORDS;
    const u64 lo_mask = (1ull << kb) - 1;
    u64 base = tile_offsets[blockIdx.x];
    for (int r = 0; r < XT_ROUNDS; ++r) {
        if (t0 + (u64)r * DB >= n_words) break;  // the same for the whole block
        const u64 i = t0 + (u64)r * DB + threadIdx.x;
        u64 w = i < n_words ? xt_word(words, minus, minus_pop, rs, i) : 0ull;
        u64 tot;
        u64 at = base + block_exclusive_scan<DB>((u64)__popcll(w), s_w, &tot);
        while (w) {
            const int b = __ffsll((unsigned long long)w) - 1;
            w &= w - 1;
            const u64 bit = i * 64 + (u64)b;
            if (at < out_cap) out[at] = ((bit >> kb) << 32) | (bit & lo_mask);
            ++at;
        }
        base += tot;
    }
}

static u64 bitmap_words(int kb) { return kb >= 3 ? (1ull << (2 * kb - 6)) : 1ull; }

// Keys of a kb-bit bitmap (minus the rows of `minus` flagged in minus_pop) in
// key order into d_out; *n_out gets the member count (a host wait).
static gdl_status bitmap_extract_sorted(gdl_ctx *ctx, const u64 *words, int kb, const u64 *minus,
                                        const u32 *minus_pop, u64 *d_out, u64 out_cap,
                                        u64 *n_out) {
    const u64 nw = bitmap_words(kb);
    const u64 tiles = (nw + XT_WORDS - 1) / XT_WORDS;
    const int rs = kb >= 6 ? kb - 6 : 0;
    Scratch scr(ctx);
    u64 *tc;
    GDL_TRY(scr.get(tiles + 1, &tc));
    {
        KTimer kt(ctx, K_BM_EXTRACT, 8.0 * nw);
        hipLaunchKernelGGL(k_xt_count, dim3((unsigned)tiles), dim3(DB), 0, ctx->stream, words, nw,
                           minus, minus_pop, rs, tc);
        GDL_TRY(check_launch(ctx, "k_xt_count"));
    }
    GDL_TRY(scan_inplace(ctx, tc, tiles));
    u64 total = 0;
    GDL_TRY(read_u64(ctx, tc + tiles, &total));
    *n_out = total;
    if (total > out_cap)
        return fail(ctx, GDL_ERR_OVERFLOW, "bitmap extract: " + std::to_string(total) +
                                               " members, output holds " + std::to_string(out_cap));
    if (total == 0) return GDL_OK;
    KTimer kt(ctx, K_BM_EXTRACT, 8.0 * nw + 8.0 * total);
    hipLaunchKernelGGL(k_xt_write, dim3((unsigned)tiles), dim3(DB), 0, ctx->stream, words, nw,
                       minus, minus_pop, rs, kb, tc, d_out, out_cap);
    return check_launch(ctx, "k_xt_write");
}

// ---------------------------------------------------------------------------
// Bodies on bit matrices.

// Device state of a run of dense bodies.  The counters come first.
struct DenseState {
    unsigned long long cursor;    // log end
    unsigned long long n_new;     // the running body's new tuples
    unsigned long long raw;       // its produced tuples
    unsigned long long next_raw;  // TC: the next body's produced tuples (accumulating)
    unsigned long long rows_io;   // row reads + writes of the running body
    u64 log_cap;
    u64 dom_left;      // free bits of the active rows (a bound on new tuples)
    u64 sparse_below;  // a body with fewer new tuples ends the run
    u64 next_bound;    // TC: a bound on the next body's new tuples
    u64 append_base;   // SG: the log end before the last body's append
    u32 stop, body, blocks_done, max_bodies;
    u32 appended;       // SG: bodies whose new tuples are in the log
    u32 append_blocks;  // SG: blocks of the running append that finished
    u32 bad;            // a delta key outside the domain
    u32 pad_;
};
enum : u32 {
    kDnRun = 0,
    kDnFixpoint = 1,
    kDnGrow = 2,
    kDnLimit = 3,
    kDnSparse = 4,
    kDnOverflow = 5
};

// The rule's static relations as row lists.  pred: r1 grouped by its second
// column (for row x, the y with r1(y, x)); act: the rows with a non-empty
// pred list (the only rows a body can write); deg1: r1 tuples per first
// column (TC's produced count of a delta row); succ: r2 grouped by its first
// column (SG).
struct DenseArgs {
    const u64 *pred_off;
    const u32 *pred_idx;
    const u32 *act;
    u64 n_act;
    const u32 *deg1;
    const u64 *succ_off;
    const u32 *succ_idx;
    u64 rw;
    u64 *full;  // FULL's bitmap words
    u64 *log;
};

__device__ __forceinline__ u64 wave_sum(u64 v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// End of a body (thread 0 of the last block): record its counts and decide
// whether the next enqueued body runs.
__device__ void dense_epilogue(DenseState *st, u64 *counts, bool tc) {
    const u64 nn = atomicAdd(&st->n_new, 0ull);
    const u64 raw = atomicAdd(&st->raw, 0ull);
    const u64 nr = atomicAdd(&st->next_raw, 0ull);
    const u64 io = atomicAdd(&st->rows_io, 0ull);
    const u64 end = atomicAdd(&st->cursor, 0ull);
    const u32 b = st->body;
    counts[3 * b] = nn;
    counts[3 * b + 1] = raw;
    counts[3 * b + 2] = io;
    st->n_new = 0;
    st->raw = 0;
    st->next_raw = 0;
    st->rows_io = 0;
    st->blocks_done = 0;
    st->append_base = end;
    const u64 left = st->dom_left > nn ? st->dom_left - nn : 0;
    st->dom_left = left;
    // the next body's new tuples: at most its produced count (TC) and the
    // active rows' free bits
    const u64 bound = tc ? (nr < left ? nr : left) : left;
    st->next_bound = bound;
    u32 stop = kDnRun;
    if (tc && end > st->log_cap) stop = kDnOverflow;
    else if (nn == 0) stop = kDnFixpoint;
    else if (b + 1 >= st->max_bodies) stop = kDnLimit;
    else if (nn < st->sparse_below) stop = kDnSparse;
    else if (tc && end + bound > st->log_cap) stop = kDnGrow;
    st->stop = stop;
    st->body = b + 1;  // last: an append launched next reads it
}

// Block totals -> state; the last block runs the epilogue.
__device__ __forceinline__ void dense_block_end(DenseState *st, u64 *counts, bool tc, u64 nn,
                                                u64 raw, u64 nr, u64 io) {
    __shared__ u64 s_v[DWV][4];
    __shared__ bool s_last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_v[wv][0] = nn;
        s_v[wv][1] = raw;
        s_v[wv][2] = nr;
        s_v[wv][3] = io;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t[4] = {0, 0, 0, 0};
        for (int v = 0; v < DWV; ++v)
            for (int j = 0; j < 4; ++j) t[j] += s_v[v][j];
        if (t[0]) atomicAdd(&st->n_new, (unsigned long long)t[0]);
        if (t[1]) atomicAdd(&st->raw, (unsigned long long)t[1]);
        if (t[2]) atomicAdd(&st->next_raw, (unsigned long long)t[2]);
        if (t[3]) atomicAdd(&st->rows_io, (unsigned long long)t[3]);
        __threadfence();  // release the counter updates before counting this block finished
        s_last = atomicAdd(&st->blocks_done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();  // acquire: every block's updates precede its blocks_done add
        dense_epilogue(st, counts, tc);
    }
}

// acc[k] |= OR of D_in rows over x's pred list entries whose pop is non-zero;
// returns (wave-uniform) whether any was; *praw gets the lane's share of the
// preds' pop sum, *pflag the number of flagged preds.
template <int K>
__device__ __forceinline__ bool or_pred_rows(const DenseArgs &a, u32 x, const u32 *pop_in,
                                             const u64 *D_in, u64 *acc, u64 *praw,
                                             u64 *pflag) {
    const int lane = threadIdx.x & 63;
    const u64 pb = a.pred_off[x], pe = a.pred_off[x + 1];
    bool any = false;
    for (u64 base = pb; base < pe; base += 64) {
        const u64 j = base + lane;
        u32 y = 0, p = 0;
        if (j < pe) {
            y = a.pred_idx[j];
            p = pop_in[y];
        }
        *praw += p;
        u64 m = __ballot(p != 0);
        if (m) any = true;
        *pflag += (u64)__popcll(m);
        while (m) {
            const int l = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            const u32 yy = (u32)__shfl((int)y, l, 64);
            const u64 *row = D_in + (u64)yy * a.rw;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                if (K > 1 || w < a.rw) acc[k] |= row[w];
            }
        }
    }
    return any;
}

// nw[k] = src[k] & ~FULL row; FULL |= nw; D_out row = nw.  Returns the
// lane's new-bit count.
template <int K>
__device__ __forceinline__ u64 diff_row(const DenseArgs &a, u32 x, u64 *nw, u64 *D_out) {
    const int lane = threadIdx.x & 63;
    u64 *frow = a.full + (u64)x * a.rw;
    u64 *drow = D_out + (u64)x * a.rw;
    u64 cnt = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const u64 w = (u64)lane + 64u * k;
        if (K > 1 || w < a.rw) {
            u64 v = nw[k];
            if (v) {
                const u64 f = frow[w];
                v &= ~f;
                if (v) frow[w] = f | v;
            }
            drow[w] = v;
            nw[k] = v;
            cnt += (u64)__popcll(v);
        } else {
            nw[k] = 0;
        }
    }
    return cnt;
}

// Appends row x's new bits to the log (positions from one cursor atomic per
// row).
template <int K>
__device__ __forceinline__ void append_row(const DenseArgs &a, u32 x, const u64 *nw, u64 cnt,
                                           u64 total, DenseState *st) {
    const int lane = threadIdx.x & 63;
    unsigned long long pos = 0;
    if (lane == 0) pos = atomicAdd(&st->cursor, (unsigned long long)total);
    pos = __shfl(pos, 0, 64);
    u64 at = pos + wave_inclusive_scan(cnt) - cnt;
    const u64 cap = st->log_cap;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        u64 v = nw[k];
        const u64 w = (u64)lane + 64u * k;
        while (v) {
            const int b = __ffsll((unsigned long long)v) - 1;
            v &= v - 1;
            if (at < cap) a.log[at] = ((u64)x << 32) | (w * 64 + (u64)b);
            ++at;
        }
    }
}

// One TC body: a wave per active row, rows strided over the grid's waves.
template <int K>
__global__ __launch_bounds__(DB) void k_dense_tc(DenseArgs a, const u32 *pop_in, const u64 *D_in,
                                                 u32 *pop_out, u64 *D_out, DenseState *st,
                                                 u64 *counts) {
    if (st->stop != kDnRun) return;  // the same for every block
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    u64 w_new = 0, w_raw = 0, w_next = 0, w_io = 0;
    for (u64 i = wave; i < a.n_act; i += n_waves) {
        const u32 x = a.act[i];
        u64 acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0;
        u64 raw = 0, flagged = 0;
        const bool any = or_pred_rows<K>(a, x, pop_in, D_in, acc, &raw, &flagged);
        w_raw += raw;
        u64 total = 0;
        if (any) {
            const u64 cnt = diff_row<K>(a, x, acc, D_out);
            total = wave_sum(cnt);
            if (total) append_row<K>(a, x, acc, cnt, total, st);
            w_io += flagged + 2;
        }
        if (lane == 0) pop_out[x] = (u32)total;
        w_new += total;
        w_next += total * a.deg1[x];
    }
    dense_block_end(st, counts, true, w_new, wave_sum(w_raw), w_next, w_io);
}

// One SG body: a wave per active row; T in registers, the output row in LDS
// (dynamic, rw words per wave).  Successor lists longer than kLight go to a
// per-wave queue that the whole wave then walks.
constexpr u32 kLight = 16;
constexpr u32 kQueue = 64;

template <int K>
__global__ __launch_bounds__(DB) void k_dense_sg(DenseArgs a, const u32 *pop_in, const u64 *D_in,
                                                 u32 *pop_out, u64 *D_out, DenseState *st,
                                                 u64 *counts) {
    extern __shared__ u64 s_rows[];
    __shared__ u32 s_q[DWV][kQueue];
    __shared__ u32 s_qn[DWV];
    if (st->stop != kDnRun) return;  // the same for every block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    u64 *s_new = s_rows + (u64)wv * a.rw;
    const u64 wave = (u64)blockIdx.x * DWV + wv;
    const u64 n_waves = (u64)gridDim.x * DWV;
    u64 w_new = 0, w_raw = 0, w_io = 0;
    for (u64 i = wave; i < a.n_act; i += n_waves) {
        const u32 x = a.act[i];
        u64 acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0;
        u64 unused = 0, flagged = 0;
        const bool any = or_pred_rows<K>(a, x, pop_in, D_in, acc, &unused, &flagged);
        u64 total = 0;
        if (any) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                if (K > 1 || w < a.rw) s_new[w] = 0;
            }
            if (lane == 0) s_qn[wv] = 0;
            wave_sync_lds();
            u64 raw = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                u64 bits = acc[k];
                const u64 w = (u64)lane + 64u * k;
                while (bits) {
                    const int bit = __ffsll((unsigned long long)bits) - 1;
                    bits &= bits - 1;
                    const u32 b = (u32)(w * 64 + (u64)bit);
                    const u64 so = a.succ_off[b], se = a.succ_off[b + 1];
                    raw += se - so;
                    if (se - so > kLight) {
                        const u32 q = atomicAdd(&s_qn[wv], 1u);
                        if (q < kQueue) {
                            s_q[wv][q] = b;
                            continue;
                        }
                    }
                    for (u64 j = so; j < se; ++j) {
                        const u32 y = a.succ_idx[j];
                        atomicOr((unsigned long long *)&s_new[y >> 6], 1ull << (y & 63));
                    }
                }
            }
            wave_sync_lds();
            const u32 nq = min(s_qn[wv], kQueue);
            for (u32 q = 0; q < nq; ++q) {
                const u32 b = s_q[wv][q];
                const u64 so = a.succ_off[b], se = a.succ_off[b + 1];
                for (u64 j = so + lane; j < se; j += 64) {
                    const u32 y = a.succ_idx[j];
                    atomicOr((unsigned long long *)&s_new[y >> 6], 1ull << (y & 63));
                }
            }
            wave_sync_lds();
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                acc[k] = (K > 1 || w < a.rw) ? s_new[w] : 0ull;
            }
            wave_sync_lds();  // s_new is rewritten for the next row
            w_raw += raw;
            const u64 cnt = diff_row<K>(a, x, acc, D_out);
            total = wave_sum(cnt);
            w_io += flagged + 2;
        }
        if (lane == 0) pop_out[x] = (u32)total;
        w_new += total;
    }
    dense_block_end(st, counts, false, w_new, wave_sum(w_raw), 0, w_io);
}

// SG: appends the last body's new tuples (the D rows with a non-zero pop) to
// the log once their count is known, or asks the host to grow the log first.
// Nothing to do when the body launched before it did not run (appended ==
// body).  Every block reads the same state: only the last block to finish
// updates `appended`, after all of them have read it.
template <int K>
__global__ __launch_bounds__(DB) void k_dense_append(DenseArgs a, const u32 *pop, const u64 *D,
                                                     DenseState *st, const u64 *counts) {
    __shared__ bool s_last;
    const u32 body = st->body;
    if (st->appended >= body) return;
    const u64 nn = counts[3 * (body - 1)];
    const u64 base = st->append_base;
    if (nn > st->log_cap || base > st->log_cap - nn) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && st->stop == kDnRun) st->stop = kDnGrow;
        return;
    }
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    for (u64 i = wave; i < a.n_act; i += n_waves) {
        const u32 x = a.act[i];
        if (pop[x] == 0) continue;  // the same for the whole wave
        u64 v[K];
        u64 cnt = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const u64 w = (u64)lane + 64u * k;
            v[k] = (K > 1 || w < a.rw) ? D[(u64)x * a.rw + w] : 0ull;
            cnt += (u64)__popcll(v[k]);
        }
        const u64 total = wave_sum(cnt);
        unsigned long long pos = 0;
        if (lane == 0) pos = atomicAdd(&st->cursor, (unsigned long long)total);
        pos = __shfl(pos, 0, 64);
        u64 at = pos + wave_inclusive_scan(cnt) - cnt;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            u64 bits = v[k];
            const u64 w = (u64)lane + 64u * k;
            while (bits) {
                const int b = __ffsll((unsigned long long)bits) - 1;
                bits &= bits - 1;
                a.log[at++] = ((u64)x << 32) | (w * 64 + (u64)b);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&st->append_blocks, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();
        st->append_blocks = 0;
        st->appended = body;
    }
}

// ---- setup kernels
__global__ __launch_bounds__(DB) void k_dense_count_col(const u64 *keys, u64 n, int col, int kb,
                                                        u32 *cnt, u32 *bad) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) {
        const u64 k = keys[i];
        const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
        if ((c0 >> kb) || (c1 >> kb)) {
            atomicOr(bad, 1u);
            continue;
        }
        atomicAdd(&cnt[col ? c1 : c0], 1u);
    }
}

__global__ __launch_bounds__(DB) void k_dense_fill_col(const u64 *keys, u64 n, int col, int kb,
                                                       u64 *cursor, u32 *idx) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) {
        const u64 k = keys[i];
        const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
        if ((c0 >> kb) || (c1 >> kb)) continue;
        const u64 row = col ? c1 : c0;
        const u64 at = atomicAdd((unsigned long long *)&cursor[row], 1ull);
        idx[at] = (u32)(col ? c0 : c1);
    }
}

__global__ __launch_bounds__(DB) void k_u32_to_u64(const u32 *in, u64 n, u64 *out) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB)
        out[i] = in[i];
}

__global__ __launch_bounds__(DB) void k_dense_act(const u64 *off, u64 rows, u32 *act,
                                                  unsigned long long *n_act) {
    for (u64 x = (u64)blockIdx.x * DB + threadIdx.x; x < rows; x += (u64)gridDim.x * DB) {
        const bool has = off[x + 1] > off[x];
        const u64 bal = __ballot(has);
        unsigned long long base = 0;
        if ((threadIdx.x & 63) == 0 && bal) base = atomicAdd(n_act, (unsigned long long)__popcll(bal));
        base = __shfl(base, 0, 64);
        if (has) act[base + __popcll(bal & lane_mask_lt())] = (u32)x;
    }
}

// Delta keys -> pop counts (duplicates counted) and TC's first produced count.
__global__ __launch_bounds__(DB) void k_dense_mark(const u64 *keys, u64 n, int kb, u32 *pop,
                                                   const u32 *deg1, DenseState *st) {
    u64 nr = 0;
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) {
        const u64 k = keys[i];
        const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
        if ((c0 >> kb) || (c1 >> kb)) {
            atomicOr(&st->bad, 1u);
            continue;
        }
        atomicAdd(&pop[c0], 1u);
        if (deg1) nr += deg1[c0];
    }
    nr = wave_sum(nr);
    if ((threadIdx.x & 63) == 0 && nr) atomicAdd(&st->next_raw, (unsigned long long)nr);
}

__global__ __launch_bounds__(DB) void k_dense_zero_rows(const u32 *pop, u64 rows, u64 rw, u64 *D) {
    const u64 n = rows * rw;
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB)
        if (pop[i / rw]) D[i] = 0ull;
}

__global__ __launch_bounds__(DB) void k_dense_set_bits(const u64 *keys, u64 n, int kb, u64 rw,
                                                       u64 *D) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) {
        const u64 k = keys[i];
        const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
        if ((c0 >> kb) || (c1 >> kb)) continue;
        atomicOr((unsigned long long *)&D[c0 * rw + (c1 >> 6)], 1ull << (c1 & 63));
    }
}

__global__ __launch_bounds__(DB) void k_fill_u32(u32 *p, u64 n, u32 v) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) p[i] = v;
}

// ---------------------------------------------------------------- host side

static unsigned dgrid(gdl_ctx *ctx, u64 n) {
    const u64 g = (n + DB - 1) / DB;
    return (unsigned)std::max<u64>(1, std::min<u64>(g, grid_cap(ctx)));
}

// Device memory owned by the context allocator for the lifetime of a run.
struct DevBufs {
    gdl_ctx *ctx;
    std::vector<void *> ptrs;
    explicit DevBufs(gdl_ctx *c) : ctx(c) {}
    ~DevBufs() {
        for (void *p : ptrs) ctx->alloc->free(p, nullptr);
    }
    template <typename T>
    gdl_status get(u64 count, T **out) {
        void *p = nullptr;
        std::string err;
        gdl_status st = ctx->alloc->alloc(std::max<u64>(count, 1) * sizeof(T), &p, &err);
        if (st != GDL_OK) return fail(ctx, st, err);
        ptrs.push_back(p);
        *out = (T *)p;
        return GDL_OK;
    }
};

// Rows of keys grouped by column `col` (the other column as the values):
// off[rows + 1], idx[n].  *bad: a key outside the kb-bit domain.
static gdl_status group_by_col(gdl_ctx *ctx, DevBufs &mem, const u64 *keys, u64 n, int col,
                               int kb, u64 rows, u32 *cnt_scratch, u64 **off, u32 **idx,
                               u32 *bad) {
    GDL_TRY(mem.get(rows + 1, off));
    GDL_TRY(mem.get(n, idx));
    hipLaunchKernelGGL(k_fill_u32, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream, cnt_scratch,
                       rows, 0u);
    GDL_TRY(check_launch(ctx, "k_fill_u32"));
    if (n) {
        hipLaunchKernelGGL(k_dense_count_col, dim3(dgrid(ctx, n)), dim3(DB), 0, ctx->stream, keys,
                           n, col, kb, cnt_scratch, bad);
        GDL_TRY(check_launch(ctx, "k_dense_count_col"));
    }
    hipLaunchKernelGGL(k_u32_to_u64, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream,
                       cnt_scratch, rows, *off);
    GDL_TRY(check_launch(ctx, "k_u32_to_u64"));
    GDL_TRY(scan_inplace(ctx, *off, rows));
    if (n) {
        Scratch scr(ctx);
        u64 *cursor;
        GDL_TRY(scr.get(rows, &cursor));
        GDL_HIP(ctx, hipMemcpyAsync(cursor, *off, rows * sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        hipLaunchKernelGGL(k_dense_fill_col, dim3(dgrid(ctx, n)), dim3(DB), 0, ctx->stream, keys,
                           n, col, kb, cursor, *idx);
        GDL_TRY(check_launch(ctx, "k_dense_fill_col"));
    }
    return GDL_OK;
}

// The cost model behind ctx->dense_mode 1 (0: never dense, 2: dense whenever
// the rule and the set allow it).
static double dense_bytes_per_tuple() {
    const char *e = std::getenv("GDLOG_DENSE_BYTES_PER_TUPLE");
    const double v = e ? std::atof(e) : 0.0;
    return v > 0.0 ? v : 256.0;
}

// Smallest delta for which a dense body is expected to beat a tuple-at-a-time
// one: a dense body streams at most (|r1| + 3 rows) rows of rw words; a sparse
// body costs about dense_bytes_per_tuple() bytes of streaming time per
// produced tuple (random probes).
static u64 dense_threshold(int shape, int kb, u64 n1, u64 n2) {
    const double rows = (double)std::min<u64>(1ull << kb, std::max<u64>(n1, 1));
    const double rw = (double)(kb >= 6 ? (1ull << (kb - 6)) : 1);
    const double bytes = ((double)n1 + 3.0 * rows) * rw * 8.0;
    double fan = (double)n1 / rows;
    if (shape == GDL_DENSE_SG) {
        const double rows2 = (double)std::min<u64>(1ull << kb, std::max<u64>(n2, 1));
        fan *= 1.0 + (double)n2 / rows2;
    }
    const double t = bytes / (std::max(fan, 1e-3) * dense_bytes_per_tuple());
    return (u64)std::max(1.0, t);
}

static bool dense_rule_ok(const gdl_hashset *hs, int shape) {
    if (!hs || !hs->slots || hs->key_bits < 6 || hs->key_bits > 17) return false;
    if (shape == GDL_DENSE_SG && hs->key_bits > 16) return false;  // LDS rows of <= 8 KB
    return shape == GDL_DENSE_TC || shape == GDL_DENSE_SG;
}

template <int K>
static gdl_status launch_body(gdl_ctx *ctx, int shape, const DenseArgs &a, const u32 *pop_in,
                              const u64 *D_in, u32 *pop_out, u64 *D_out, DenseState *st,
                              u64 *counts, unsigned grid) {
    if (shape == GDL_DENSE_TC) {
        hipLaunchKernelGGL(k_dense_tc<K>, dim3(grid), dim3(DB), 0, ctx->stream, a, pop_in, D_in,
                           pop_out, D_out, st, counts);
        return check_launch(ctx, "k_dense_tc");
    }
    const size_t lds = (size_t)DWV * a.rw * sizeof(u64);
    hipLaunchKernelGGL(k_dense_sg<K>, dim3(grid), dim3(DB), lds, ctx->stream, a, pop_in, D_in,
                       pop_out, D_out, st, counts);
    GDL_TRY(check_launch(ctx, "k_dense_sg"));
    hipLaunchKernelGGL(k_dense_append<K>, dim3(grid), dim3(DB), 0, ctx->stream, a, pop_out, D_out,
                       st, counts);
    return check_launch(ctx, "k_dense_append");
}

static gdl_status launch_append(gdl_ctx *ctx, int kb, const DenseArgs &a, const u32 *pop,
                                const u64 *D, DenseState *st, const u64 *counts, unsigned grid) {
#define GDL_DENSE_APPEND(K)                                                                     \
    hipLaunchKernelGGL(k_dense_append<K>, dim3(grid), dim3(DB), 0, ctx->stream, a, pop, D, st, \
                       counts);                                                                 \
    return check_launch(ctx, "k_dense_append")
    switch (kb) {
    case 17: GDL_DENSE_APPEND(32);
    case 16: GDL_DENSE_APPEND(16);
    case 15: GDL_DENSE_APPEND(8);
    case 14: GDL_DENSE_APPEND(4);
    case 13: GDL_DENSE_APPEND(2);
    default: GDL_DENSE_APPEND(1);
    }
#undef GDL_DENSE_APPEND
}

static gdl_status dispatch_body(gdl_ctx *ctx, int shape, int kb, const DenseArgs &a,
                                const u32 *pop_in, const u64 *D_in, u32 *pop_out, u64 *D_out,
                                DenseState *st, u64 *counts, unsigned grid) {
    switch (kb) {
    case 17: return launch_body<32>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    case 16: return launch_body<16>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    case 15: return launch_body<8>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    case 14: return launch_body<4>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    case 13: return launch_body<2>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    default: return launch_body<1>(ctx, shape, a, pop_in, D_in, pop_out, D_out, st, counts, grid);
    }
}

// Grows the log to hold `need` keys (geometrically), keeping its first
// `keep` keys.
static gdl_status grow_log(gdl_ctx *ctx, u64 **d_log, u64 *log_cap, u64 keep, u64 need) {
    if (*log_cap >= need) return GDL_OK;
    need = std::max<u64>(need, 2 * *log_cap);
    void *p = nullptr;
    std::string err;
    gdl_status s = ctx->alloc->alloc(need * sizeof(u64), &p, &err);
    if (s != GDL_OK) return fail(ctx, s, err);
    if (keep)
        GDL_HIP(ctx, hipMemcpyAsync(p, *d_log, keep * sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
    if (*d_log) ctx->alloc->free(*d_log, nullptr);
    *d_log = (u64 *)p;
    *log_cap = need;
    return GDL_OK;
}

// Smallest delta for which ctx's policy runs dense bodies (~0: never; 0:
// whenever the rule and the set allow it).
u64 dense_threshold_for(const gdl_ctx *ctx, const gdl_hashset *hs, int shape, u64 n1, u64 n2) {
    if (ctx->dense_mode == 0 || !dense_rule_ok(hs, shape)) return ~0ull;
    if (ctx->dense_mode == 2) return 0;
    return dense_threshold(shape, (int)hs->key_bits, n1, n2);
}

// Runs dense bodies from the delta d_delta[0, n_delta) (members of FULL =
// the bitmap set hs, whose append log holds *log_end keys); see
// gdl_fixpoint_dense.
gdl_status dense_run(gdl_ctx *ctx, const gdl_dense_rule *rule, gdl_hashset *hs,
                     const u64 *d_delta, u64 n_delta, u64 **d_log, u64 *log_cap, u64 *log_end,
                     int max_bodies, u64 *h_new, u64 *h_raw, int *bodies) {
    *bodies = 0;
    const int shape = rule->shape;
    const u64 threshold = dense_threshold_for(ctx, hs, shape, rule->n1, rule->n2);
    if (max_bodies <= 0 || n_delta == 0 || threshold == ~0ull || n_delta < threshold)
        return GDL_OK;
    const u64 sparse_below = threshold / 2;
    const int kb = (int)hs->key_bits;
    const u64 rows = 1ull << kb, rw = 1ull << (kb - 6);
    DevBufs mem(ctx);
    Scratch scr(ctx);
    u32 *cnt, *bad;
    GDL_TRY(scr.get(rows, &cnt));
    GDL_TRY(scr.get(1, &bad));
    GDL_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(u32), ctx->stream));
    DenseArgs a{};
    a.rw = rw;
    a.full = (u64 *)hs->slots;
    u64 *pred_off;
    u32 *pred_idx;
    {
        KTimer kt(ctx, K_DENSE_SETUP,
                  24.0 * rule->n1 + (shape == GDL_DENSE_SG ? 24.0 * rule->n2 : 0.0));
        GDL_TRY(group_by_col(ctx, mem, rule->r1, rule->n1, 1, kb, rows, cnt, &pred_off, &pred_idx,
                             bad));
    }
    a.pred_off = pred_off;
    a.pred_idx = pred_idx;
    u32 *act;
    GDL_TRY(mem.get(rows, &act));
    u64 *n_act_d;
    GDL_TRY(scr.get(1, &n_act_d));
    GDL_HIP(ctx, hipMemsetAsync(n_act_d, 0, sizeof(u64), ctx->stream));
    hipLaunchKernelGGL(k_dense_act, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream, pred_off,
                       rows, act, (unsigned long long *)n_act_d);
    GDL_TRY(check_launch(ctx, "k_dense_act"));
    a.act = act;
    if (shape == GDL_DENSE_TC) {
        u32 *deg1;
        GDL_TRY(mem.get(rows, &deg1));
        hipLaunchKernelGGL(k_fill_u32, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream, deg1,
                           rows, 0u);
        GDL_TRY(check_launch(ctx, "k_fill_u32"));
        if (rule->n1) {
            hipLaunchKernelGGL(k_dense_count_col, dim3(dgrid(ctx, rule->n1)), dim3(DB), 0,
                               ctx->stream, rule->r1, rule->n1, 0, kb, deg1, bad);
            GDL_TRY(check_launch(ctx, "k_dense_count_col"));
        }
        a.deg1 = deg1;
    } else {
        u64 *succ_off;
        u32 *succ_idx;
        GDL_TRY(group_by_col(ctx, mem, rule->r2, rule->n2, 0, kb, rows, cnt, &succ_off, &succ_idx,
                             bad));
        a.succ_off = succ_off;
        a.succ_idx = succ_idx;
    }
    u64 setup[2] = {0, 0};
    GDL_HIP(ctx, hipMemcpyAsync(&setup[0], n_act_d, sizeof(u64), hipMemcpyDeviceToHost,
                                ctx->stream));
    GDL_HIP(ctx, hipMemcpyAsync(&setup[1], bad, sizeof(u32), hipMemcpyDeviceToHost, ctx->stream));
    GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    a.n_act = setup[0];
    // ids beyond the set's domain in a static relation: the tuple-at-a-time
    // bodies handle (and report) them
    if (setup[1] || a.n_act == 0) return GDL_OK;

    // the delta as bit rows (buffer 0)
    u64 *D[2];
    u32 *pop[2];
    GDL_TRY(mem.get(rows * rw, &D[0]));
    GDL_TRY(mem.get(rows * rw, &D[1]));
    GDL_TRY(mem.get(rows, &pop[0]));
    GDL_TRY(mem.get(rows, &pop[1]));
    for (int p = 0; p < 2; ++p) {
        hipLaunchKernelGGL(k_fill_u32, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream, pop[p],
                           rows, 0u);
        GDL_TRY(check_launch(ctx, "k_fill_u32"));
    }
    constexpr int kChunk = 64;
    u64 *st_words, *counts;
    GDL_TRY(scr.get((sizeof(DenseState) + 7) / 8, &st_words));
    GDL_TRY(scr.get(3 * kChunk, &counts));
    DenseState *st = (DenseState *)st_words;
    GDL_HIP(ctx, hipMemsetAsync(st, 0, sizeof(DenseState), ctx->stream));
    {
        KTimer kt(ctx, K_DENSE_SETUP, 24.0 * n_delta + 8.0 * rows * rw);
        hipLaunchKernelGGL(k_dense_mark, dim3(dgrid(ctx, n_delta)), dim3(DB), 0, ctx->stream,
                           d_delta, n_delta, kb, pop[0],
                           shape == GDL_DENSE_TC ? a.deg1 : (const u32 *)nullptr, st);
        GDL_TRY(check_launch(ctx, "k_dense_mark"));
        hipLaunchKernelGGL(k_dense_zero_rows, dim3(dgrid(ctx, rows * rw)), dim3(DB), 0,
                           ctx->stream, pop[0], rows, rw, D[0]);
        GDL_TRY(check_launch(ctx, "k_dense_zero_rows"));
        hipLaunchKernelGGL(k_dense_set_bits, dim3(dgrid(ctx, n_delta)), dim3(DB), 0, ctx->stream,
                           d_delta, n_delta, kb, rw, D[0]);
        GDL_TRY(check_launch(ctx, "k_dense_set_bits"));
    }
    DenseState r{};
    GDL_HIP(ctx, hipMemcpyAsync(&r, st, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
    GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (r.bad)
        return fail(ctx, GDL_ERR_INVALID_ARG, "dense bodies: a delta key outside the set's domain");

    u64 end = *log_end;
    // new tuples land in active rows only
    u64 dom_left = a.n_act * rows;
    u64 bound = std::min<u64>(r.next_raw, dom_left);
    const unsigned grid =
        (unsigned)std::max<u64>(1, std::min<u64>((a.n_act + DWV - 1) / DWV, grid_cap(ctx)));
    int done = 0, ramp = 4;
    while (done < max_bodies) {
        if (shape == GDL_DENSE_TC)  // room for the next body's new tuples
            GDL_TRY(grow_log(ctx, d_log, log_cap, end, end + std::max<u64>(bound, 4096)));
        a.log = *d_log;
        const int k = std::min({ramp, kChunk, max_bodies - done});
        DenseState s{};
        s.cursor = end;
        s.log_cap = *log_cap;
        s.dom_left = dom_left;
        s.sparse_below = sparse_below;
        s.max_bodies = (u32)k;
        GDL_HIP(ctx, hipMemcpyAsync(st, &s, sizeof(s), hipMemcpyHostToDevice, ctx->stream));
        for (int i = 0; i < k; ++i) {
            const int p = (done + i) & 1;
            KTimer kt(ctx, K_DENSE_BODY, 0.0);
            GDL_TRY(dispatch_body(ctx, shape, kb, a, pop[p], D[p], pop[p ^ 1], D[p ^ 1], st,
                                  counts, grid));
        }
        u64 c[3 * kChunk];
        GDL_HIP(ctx, hipMemcpyAsync(&r, st, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
        GDL_HIP(ctx, hipMemcpyAsync(c, counts, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
        GDL_TRY(check_device_error(ctx));
        if (r.stop == kDnOverflow)
            return fail(ctx, GDL_ERR_OVERFLOW, "dense bodies: a body outgrew its log bound");
        for (u32 b = 0; b < r.body; ++b, ++done) {
            h_new[done] = c[3 * b];
            h_raw[done] = c[3 * b + 1];
            ctx->stats[K_DENSE_BODY].alg_bytes +=
                8.0 * (double)rw * (double)c[3 * b + 2] + 16.0 * (double)c[3 * b];
            hs->size += c[3 * b];
        }
        if (shape == GDL_DENSE_SG && r.appended < r.body) {
            // the last body's append found the log full: grow it and append
            const u64 nn = c[3 * (r.body - 1)];
            GDL_TRY(grow_log(ctx, d_log, log_cap, r.cursor, r.cursor + nn));
            a.log = *d_log;
            const u32 stop = r.stop;
            r.stop = kDnRun;
            r.log_cap = *log_cap;
            GDL_HIP(ctx, hipMemcpyAsync(st, &r, sizeof(r), hipMemcpyHostToDevice, ctx->stream));
            GDL_TRY(launch_append(ctx, kb, a, pop[done & 1], D[done & 1], st, counts, grid));
            GDL_HIP(ctx, hipMemcpyAsync(&r, st, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
            GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
            GDL_TRY(check_device_error(ctx));
            if (r.appended != r.body)
                return fail(ctx, GDL_ERR_OVERFLOW, "dense bodies: the log did not grow");
            r.stop = stop == kDnGrow ? kDnLimit : stop;
        }
        end = r.cursor;
        dom_left = r.dom_left;
        bound = r.next_bound;
        if (r.body == 0 || r.stop == kDnFixpoint || r.stop == kDnSparse) break;
        ramp = (int)r.body == k ? std::min(2 * ramp, kChunk) : std::max(4, (int)r.body);
    }
    *bodies = done;
    *log_end = end;
    return GDL_OK;
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_hashset_extract_sorted(gdl_ctx *ctx, const gdl_hashset *hs, uint64_t *d_out,
                                      uint64_t out_cap, uint64_t *n_out) {
    if (!ctx || !hs || !n_out || (out_cap && !d_out)) return GDL_ERR_INVALID_ARG;
    *n_out = 0;
    if (!hs->slots || hs->capacity == 0) return GDL_OK;
    if (hs->key_bits)
        return bitmap_extract_sorted(ctx, (const u64 *)hs->slots, (int)hs->key_bits, nullptr,
                                     nullptr, d_out, out_cap, n_out);
    // a hash set: its members, then the radix sort
    Scratch scr(ctx);
    u64 *tmp;
    GDL_TRY(scr.get(std::max<u64>(hs->size, 1), &tmp));
    u64 n = 0;
    GDL_TRY(gdl_hashset_extract(ctx, hs, tmp, hs->size, &n));
    *n_out = n;
    if (n > out_cap)
        return fail(ctx, GDL_ERR_OVERFLOW, "hash set extract: " + std::to_string(n) +
                                               " members, output holds " + std::to_string(out_cap));
    return n ? sort_keys(ctx, tmp, n, 32, 32, d_out) : GDL_OK;
}

gdl_status gdl_set_dense_mode(gdl_ctx *ctx, int mode) {
    if (!ctx || mode < 0 || mode > 2) return GDL_ERR_INVALID_ARG;
    ctx->dense_mode = mode;
    return GDL_OK;
}

gdl_status gdl_fixpoint_dense(gdl_ctx *ctx, const gdl_dense_rule *rule, gdl_hashset *full,
                              const uint64_t *d_delta, uint64_t n_delta, uint64_t **d_log,
                              uint64_t *log_cap, uint64_t *log_end, int max_bodies,
                              uint64_t *h_new, uint64_t *h_raw, int *bodies) {
    if (!ctx || !rule || !full || !d_log || !log_cap || !log_end || !h_new || !h_raw || !bodies ||
        (n_delta && !d_delta) || (rule->n1 && !rule->r1) ||
        (rule->shape == GDL_DENSE_SG && rule->n2 && !rule->r2) || *log_end > *log_cap ||
        (*log_end && !*d_log))
        return GDL_ERR_INVALID_ARG;
    if (rule->shape != GDL_DENSE_TC && rule->shape != GDL_DENSE_SG)
        return fail(ctx, GDL_ERR_INVALID_ARG, "dense rule: unknown shape");
    return dense_run(ctx, rule, full, d_delta, n_delta, d_log, log_cap, log_end, max_bodies,
                     h_new, h_raw, bodies);
}

}  // extern "C"
