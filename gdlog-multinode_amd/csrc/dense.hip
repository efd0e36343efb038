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
#include <vector>

#include "gdl_internal.hpp"

namespace gdl {

constexpr int DB = 256;  // threads per block: 4 waves
constexpr int DWV = DB / 64;

__device__ __forceinline__ u64 wave_sum(u64 v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Position of the j-th set bit (0-based) of w (j < popcount(w)).
__device__ __forceinline__ int select_bit(u64 w, int j) {
    int pos = 0;
    u32 x = (u32)w;
    int c = __popc(x);
    if (j >= c) {
        j -= c;
        pos = 32;
        x = (u32)(w >> 32);
    }
    c = __popc(x & 0xffffu);
    if (j >= c) { j -= c; pos += 16; x >>= 16; }
    c = __popc(x & 0xffu);
    if (j >= c) { j -= c; pos += 8; x >>= 8; }
    c = __popc(x & 0xfu);
    if (j >= c) { j -= c; pos += 4; x >>= 4; }
    c = __popc(x & 0x3u);
    if (j >= c) { j -= c; pos += 2; x >>= 2; }
    if (j >= (int)(x & 1u)) pos += 1;
    return pos;
}

// Writes the keys of one round of 64 consecutive words, one per lane (lane
// l: word wd, bit i -> key of bit number bitbase + 64 l + i, (row << 32) |
// column with kb-bit columns) to out[base ..), in bit order.  Each lane puts
// the round-local numbers of its bits in the wave's LDS list (buf, 4096
// entries) from its exclusive count on; the wave then stores the keys with
// consecutive lanes on consecutive positions.  Returns the round's total
// (the same in every lane).
__device__ __forceinline__ u64 stage_write(u64 wd, u64 bitbase, int kb, u64 *out, u64 base,
                                           u64 out_cap, uint16_t *buf) {
    const int lane = threadIdx.x & 63;
    const u64 c = (u64)__popcll(wd);
    const u64 incl = wave_inclusive_scan(c);
    const u64 total = __shfl(incl, 63, 64);
    if (total == 0) return 0;
    u32 at = (u32)(incl - c);
    const u32 l64 = (u32)lane << 6;
    while (wd) {
        buf[at++] = (uint16_t)(l64 | (u32)(__ffsll((unsigned long long)wd) - 1));
        wd &= wd - 1;
    }
    wave_sync_lds();
    const u64 lo_mask = (1ull << kb) - 1;
    for (u64 t = lane; t < total; t += 64) {
        const u64 bit = bitbase + buf[t];
        if (base + t < out_cap) out[base + t] = ((bit >> kb) << 32) | (bit & lo_mask);
    }
    wave_sync_lds();  // the list is rewritten by the next round
    return total;
}

// ---------------------------------------------------------------------------
// Key-ordered extraction.  Tiles of up to 1024 words, one per wave (rounds
// of 64 consecutive words): pass 1 counts each tile's bits, an exclusive scan
// of the counts gives each tile's first output position, pass 2 writes each
// tile's keys in bit order with a wave prefix sum per round (consecutive
// words, consecutive output ranges; no block barrier).  With `minus`, rows
// whose minus_pop entry is non-zero have minus's bits removed.
// rounds per tile: 16, fewer when that would leave fewer than 2^16 tiles
static int xt_rounds(u64 n_words) {
    // small tiles: a clustered bitmap has tiles far denser than the rest,
    // and many small ones let the scheduler spread them
    int r = 16;
    while (r > 1 && n_words / (64ull * r) < (1ull << 16)) r >>= 1;
    return r;
}

__device__ __forceinline__ u64 xt_word(const u64 *words, const u64 *minus, const u32 *minus_pop,
                                       int rs, u64 i) {
    u64 w = words[i];
    if (minus && minus_pop[i >> rs]) w &= ~minus[i];
    return w;
}

// R rounds per tile: every round's words are loaded before the first is
// processed (R loads in flight per lane).
template <int R>
__global__ __launch_bounds__(DB) void k_xt_count(const u64 *words, u64 n_words, u64 n_tiles,
                                                 const u64 *minus, const u32 *minus_pop, int rs,
                                                 u64 *tile_counts) {
    const u64 tile = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;  // the same for the whole wave
    const int lane = threadIdx.x & 63;
    const u64 t0 = tile * 64ull * R;
    u64 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const u64 i = t0 + (u64)r * 64 + lane;
        w[r] = i < n_words ? xt_word(words, minus, minus_pop, rs, i) : 0ull;
    }
    u64 c = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) c += (u64)__popcll(w[r]);
    c = wave_sum(c);
    if (lane == 0) tile_counts[tile] = c;
}

template <int R>
__global__ __launch_bounds__(DB) void k_xt_write(const u64 *words, u64 n_words, u64 n_tiles,
                                                 const u64 *minus, const u32 *minus_pop, int rs,
                                                 int kb, const u64 *tile_offsets, u64 *out,
                                                 u64 out_cap) {
    __shared__ uint16_t s_buf[DWV][4096];
    const u64 tile = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;  // the same for the whole wave
    const int lane = threadIdx.x & 63;
    const u64 t0 = tile * 64ull * R;
    u64 w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const u64 i = t0 + (u64)r * 64 + lane;
        w[r] = i < n_words ? xt_word(words, minus, minus_pop, rs, i) : 0ull;
    }
    u64 base = tile_offsets[tile];
#pragma unroll
    for (int r = 0; r < R; ++r)
        base += stage_write(w[r], (t0 + (u64)r * 64) * 64, kb, out, base, out_cap,
                            s_buf[threadIdx.x >> 6]);
}

// Pass 1 and pass 2 of the extraction over `tiles` tiles of 64 * rounds words.
static gdl_status launch_xt_count(gdl_ctx *ctx, int rounds, const u64 *words, u64 nw, u64 tiles,
                                  const u64 *minus, const u32 *minus_pop, int rs, u64 *tc) {
    const unsigned blocks = (unsigned)((tiles + DWV - 1) / DWV);
#define GDL_XT_COUNT(R)                                                                        \
    hipLaunchKernelGGL(k_xt_count<R>, dim3(blocks), dim3(DB), 0, ctx->stream, words, nw, tiles, \
                       minus, minus_pop, rs, tc)
    switch (rounds) {
    case 1: GDL_XT_COUNT(1); break;
    case 2: GDL_XT_COUNT(2); break;
    case 4: GDL_XT_COUNT(4); break;
    case 8: GDL_XT_COUNT(8); break;
    default: GDL_XT_COUNT(16); break;
    }
#undef GDL_XT_COUNT
    return check_launch(ctx, "k_xt_count");
}

static gdl_status launch_xt_write(gdl_ctx *ctx, int rounds, const u64 *words, u64 nw, u64 tiles,
                                  const u64 *minus, const u32 *minus_pop, int rs, int kb,
                                  const u64 *tc, u64 *out, u64 out_cap) {
    const unsigned blocks = (unsigned)((tiles + DWV - 1) / DWV);
#define GDL_XT_WRITE(R)                                                                        \
    hipLaunchKernelGGL(k_xt_write<R>, dim3(blocks), dim3(DB), 0, ctx->stream, words, nw, tiles, \
                       minus, minus_pop, rs, kb, tc, out, out_cap)
    switch (rounds) {
    case 1: GDL_XT_WRITE(1); break;
    case 2: GDL_XT_WRITE(2); break;
    case 4: GDL_XT_WRITE(4); break;
    case 8: GDL_XT_WRITE(8); break;
    default: GDL_XT_WRITE(16); break;
    }
#undef GDL_XT_WRITE
    return check_launch(ctx, "k_xt_write");
}

static u64 bitmap_words(int kb) { return kb >= 3 ? (1ull << (2 * kb - 6)) : 1ull; }

// Keys of a kb-bit bitmap (minus the rows of `minus` flagged in minus_pop) in
// key order into d_out; *n_out gets the member count (a host wait).
static gdl_status bitmap_extract_sorted(gdl_ctx *ctx, const u64 *words, int kb, const u64 *minus,
                                        const u32 *minus_pop, u64 *d_out, u64 out_cap,
                                        u64 *n_out) {
    const u64 nw = bitmap_words(kb);
    const int rounds = xt_rounds(nw);
    const u64 tiles = (nw + 64ull * rounds - 1) / (64ull * rounds);
    const int rs = kb >= 6 ? kb - 6 : 0;
    Scratch scr(ctx);
    u64 *tc;
    GDL_TRY(scr.get(tiles + 1, &tc));
    {
        KTimer kt(ctx, K_BM_EXTRACT, 8.0 * nw);
        GDL_TRY(launch_xt_count(ctx, rounds, words, nw, tiles, minus, minus_pop, rs, tc));
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
    return launch_xt_write(ctx, rounds, words, nw, tiles, minus, minus_pop, rs, kb, tc, d_out,
                           out_cap);
}

// ---------------------------------------------------------------------------
// Bodies on bit matrices.

// Device state of a run of dense bodies.  The counters come first.
struct DenseState {
    unsigned long long cursor;   // log end
    unsigned long long n_new;    // the running body's new tuples
    unsigned long long raw;      // its produced tuples
    unsigned long long rows_io;   // the combine's row reads + writes of the running body
    unsigned long long exp_io;    // the expansion's row reads + writes
    unsigned long long rows_new;  // rows that got new tuples
    unsigned long long tr_io;     // SG column expansion: 64-byte segments the transposes moved
    unsigned long long or_io;     // and the row reads + writes of its row OR
    u64 col_above;     // SG: a body whose delta has at least this many tuples expands by columns
    u64 log_cap;
    u64 sparse_below;  // a body with fewer new tuples ends the run
    u64 append_base;   // the log end before the last body's append
    u32 stop, body, blocks_done, max_bodies;
    u32 appended;       // bodies whose new tuples are in the log
    u32 append_blocks;  // blocks of the running append that finished
    u32 bad;            // a delta key outside the domain
    u32 col;            // SG: the running body expands by columns
    u32 asym;           // set-up: FULL or the first delta differs from its transpose
    u32 d0_out;         // set-up: the first delta has a tuple FULL lacks
    u32 d0_loose;       // set-up: its keys are not strictly increasing
};
enum : u32 { kDnRun = 0, kDnFixpoint = 1, kDnGrow = 2, kDnLimit = 3, kDnSparse = 4 };
// per body: new tuples, produced tuples, combine rows, expansion rows, rows
// with new tuples, column expansion: transposed segments, OR rows
constexpr int kCounts = 7;

// The rule's static relations as row lists.  pred: r1 grouped by its second
// column (for row x, the y with r1(y, x)); act: the rows with a non-empty
// pred list (the only rows a body can write); items: the combine's work,
// (x, first pred, end pred, split index or ~0), rows with more than kSlice
// preds split into slices (most preds first); per split row the slice count,
// a finished-slice counter and accumulators (NEW, and SG's T).  SG: succ: r2
// grouped by its first column, and succ_tab its per-row entries; planes:
// bit-planes of r2's row lengths (bit b of plane j = bit j of |r2 row b|).
constexpr u32 kSlice = 32;

struct DenseArgs {
    const u64 *pred_off;
    const u32 *pred_idx;
    const u32 *act;
    u64 n_act;
    const uint4 *items;
    u64 n_items;
    const u32 *n_slices;
    u32 *slice_done;
    u64 *GU, *GT;
    u64 n_exp;  // SG: the rows U is built for (0 .. n_exp - 1, those with a delta row)
    const u64 *succ_off;
    const u32 *succ_idx;
    const uint4 *succ_tab;  // {deg, successors} (deg <= 3) or {deg, offset lo, hi, bit row}
    const u64 *heavy;       // bit rows of the lists longer than kHeavy
    const u64 *planes;
    int n_planes;
    // SG with r1 = r2 and a symmetric start (see dense_run): every delta is
    // symmetric, D^T = D, so the column expansion skips its first transpose
    // and its row OR's output V is the tmp T (the combine reads T[x] = V[x])
    int sym;
    const u64 *V;
    const u32 *vpop;
    u64 rw;
    u64 *full;  // FULL's bitmap words
    u64 *log;
};

// End of a body (thread 0 of the last block): record its counts and decide
// whether the next enqueued body runs.
__device__ void dense_epilogue(DenseState *st, u64 *counts) {
    const u64 nn = atomicAdd(&st->n_new, 0ull);
    const u64 raw = atomicAdd(&st->raw, 0ull);
    const u64 io = atomicAdd(&st->rows_io, 0ull);
    const u64 eio = atomicAdd(&st->exp_io, 0ull);
    const u64 rn = atomicAdd(&st->rows_new, 0ull);
    const u64 tio = atomicAdd(&st->tr_io, 0ull);
    const u64 oio = atomicAdd(&st->or_io, 0ull);
    const u32 b = st->body;
    u64 *c = counts + kCounts * b;
    c[0] = nn;
    c[1] = raw;
    c[2] = io;
    c[3] = eio;
    c[4] = rn;
    c[5] = tio;
    c[6] = oio;
    st->n_new = 0;
    st->raw = 0;
    st->rows_io = 0;
    st->exp_io = 0;
    st->rows_new = 0;
    st->tr_io = 0;
    st->or_io = 0;
    st->col = nn >= st->col_above ? 1u : 0u;  // how the next body expands
    st->blocks_done = 0;
    st->append_base = st->cursor;
    u32 stop = kDnRun;
    if (nn == 0) stop = kDnFixpoint;
    else if (b + 1 >= st->max_bodies) stop = kDnLimit;
    else if (nn < st->sparse_below) stop = kDnSparse;
    st->stop = stop;
    st->body = b + 1;  // last: the append launched next reads it
}

// Block totals -> state; the last block runs the epilogue.
__device__ __forceinline__ void dense_block_end(DenseState *st, u64 *counts, u64 nn, u64 raw,
                                                u64 io, u64 rn) {
    __shared__ u64 s_v[DWV][4];
    __shared__ bool s_last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_v[wv][0] = nn;
        s_v[wv][1] = raw;
        s_v[wv][2] = io;
        s_v[wv][3] = rn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t[4] = {0, 0, 0, 0};
        for (int v = 0; v < DWV; ++v)
            for (int j = 0; j < 4; ++j) t[j] += s_v[v][j];
        if (t[0]) atomicAdd(&st->n_new, (unsigned long long)t[0]);
        if (t[1]) atomicAdd(&st->raw, (unsigned long long)t[1]);
        if (t[2]) atomicAdd(&st->rows_io, (unsigned long long)t[2]);
        if (t[3]) atomicAdd(&st->rows_new, (unsigned long long)t[3]);
        __threadfence();  // release the counter updates before counting this block finished
        s_last = atomicAdd(&st->blocks_done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();  // acquire: every block's updates precede its blocks_done add
        dense_epilogue(st, counts);
    }
}

// ORs rows src[y] (rw words) into acc over the entries y of flag mask m of
// the wave's 64 pred entries (lane l holds entry y); two rows in flight at a
// time (m is the same in every lane).
template <int K>
__device__ __forceinline__ void or_rows(const u64 *src, u64 rw, u64 m, u32 y, u64 *acc) {
    const int lane = threadIdx.x & 63;
    while (m) {
        const int l1 = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const bool two = m != 0;
        int l2 = l1;
        if (two) {
            l2 = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
        }
        const u64 *r1 = src + (u64)(u32)__shfl((int)y, l1, 64) * rw;
        const u64 *r2 = src + (u64)(u32)__shfl((int)y, l2, 64) * rw;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const u64 w = (u64)lane + 64u * k;
            if (K > 1 || w < rw) {
                const u64 v1 = r1[w];
                const u64 v2 = two ? r2[w] : 0ull;
                acc[k] |= v1 | v2;
            }
        }
    }
}

// nw[k] &= ~FULL row; FULL |= nw; D_out row = nw.  Returns the lane's
// new-bit count.
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

// SG's produced count of a tmp row T: the sum of r2's row lengths over T's
// bits, as sum_j 2^j popcount(T & plane_j).  Returns the lane's share.
template <int K>
__device__ __forceinline__ u64 plane_dot(const DenseArgs &a, const u64 *t) {
    const int lane = threadIdx.x & 63;
    u64 sum = 0;
    for (int j = 0; j < a.n_planes; ++j) {
        const u64 *pl = a.planes + (u64)j * a.rw;
        u64 c = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const u64 w = (u64)lane + 64u * k;
            if (K > 1 || w < a.rw) c += (u64)__popcll(t[k] & pl[w]);
        }
        sum += c << j;
    }
    return sum;
}

// SG, a symmetric run's column bodies: the tmp row T[x] is V[x], the row
// OR's output (rows with vpop 0 are empty).  plane_dot of it.
template <int K>
__device__ __forceinline__ u64 v_dot(const DenseArgs &a, u32 x) {
    const int lane = threadIdx.x & 63;
    const bool live = a.vpop[x] != 0;  // the same for the whole wave
    u64 t[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const u64 w = (u64)lane + 64u * k;
        t[k] = (live && (K > 1 || w < a.rw)) ? a.V[(u64)x * a.rw + w] : 0ull;
    }
    return live ? plane_dot<K>(a, t) : 0ull;
}

// SG: U[a] = the union of r2's successor lists over the bits b of D[a] (the
// second join applied to one delta row), for every row a that is some row's
// pred; upop[a] = |U[a]|.  A wave per row, U built in LDS (rw words per
// wave).  Each lane expands its words' bits 4 at a time (4 independent entry
// loads in flight): lists of at most kInline successors are in the entry;
// longer ones are listed per wave and walked by the whole wave after the
// round, 64 successors a step (a prefix sum over the listed lengths and a
// binary search per lane); the lists of more than kHeavy successors are bit
// rows (Eh), listed too and ORed in by the whole wave, two rows in flight.
constexpr u32 kInline = 3;   // successor lists this short live in the entry itself
constexpr u32 kHeavy = 64;   // longer lists are bit rows
constexpr u32 kHeavyMax = 4096;  // bit rows kept at most
constexpr u32 kHList = 256;  // heavy bits collected per wave and round
constexpr u32 kMList = 128;  // walked lists collected per wave and round
constexpr int kBatch = 4;
constexpr u32 kNoBit = 0xffffffffu;

// Sets bit y of an LDS bit row: a 32-bit OR (a 64-bit one spans two banks,
// so lanes setting random bits collide on banks about twice as often).
__device__ __forceinline__ void lds_set(u64 *row, u32 y) {
    atomicOr(reinterpret_cast<u32 *>(row) + (y >> 5), 1u << (y & 31));
}

template <int K>
__global__ __launch_bounds__(DB) void k_dense_expand(DenseArgs a, const u32 *pop_in,
                                                     const u64 *D_in, u32 *upop, u64 *U,
                                                     DenseState *st) {
    // a block per row: thread t expands the bits of words t, t + DB, ...
    // into the block's LDS row; each wave lists its own walks and bit rows
    constexpr int KB = K >= 4 ? K / 4 : 1;  // words per thread
    extern __shared__ u64 s_new[];
    __shared__ u32 s_h[DWV][kHList];
    __shared__ u32 s_hn[DWV];
    __shared__ u64 s_mo[DWV][kMList];      // listed walks: CSR offset
    __shared__ u32 s_mx[DWV][kMList + 1];  // and the exclusive sums of their lengths
    __shared__ u32 s_mn[DWV];
    __shared__ u64 s_cnt[DWV + 1];
    if (st->stop != kDnRun || st->col) return;  // the same for every block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    u64 io = 0;
    for (u64 i = blockIdx.x; i < a.n_exp; i += gridDim.x) {
        const u32 r = (u32)i;
        if (pop_in[r] == 0) {  // the same for the whole block
            if (threadIdx.x == 0) upop[r] = 0;
            continue;
        }
        u64 v[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const u64 w = (u64)threadIdx.x + (u64)DB * k;
            v[k] = w < a.rw ? D_in[(u64)r * a.rw + w] : 0ull;
        }
        for (u64 w = threadIdx.x; w < a.rw; w += DB) s_new[w] = 0;
        __syncthreads();
        // rounds (per wave): every lane expands its light bits; walks and
        // bit rows are listed (up to the lists' sizes a round) and done by
        // the whole wave after the round
        bool more = true;
        while (more) {
            if (lane == 0) {
                s_hn[wv] = 0;
                s_mn[wv] = 0;
            }
            wave_sync_lds();
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const u32 w64 = (u32)(((u64)threadIdx.x + (u64)DB * k) * 64);
                while (v[k]) {
                    u32 bb[kBatch];
                    u64 rest = v[k];
#pragma unroll
                    for (int j = 0; j < kBatch; ++j) {
                        bb[j] = kNoBit;
                        if (rest) {
                            bb[j] = w64 + (u32)(__ffsll((unsigned long long)rest) - 1);
                            rest &= rest - 1;
                        }
                    }
                    uint4 e[kBatch];
#pragma unroll
                    for (int j = 0; j < kBatch; ++j)
                        e[j] = bb[j] != kNoBit ? a.succ_tab[bb[j]] : make_uint4(0, 0, 0, 0);
                    bool full = false;
#pragma unroll
                    for (int j = 0; j < kBatch; ++j) {
                        if (bb[j] == kNoBit || full) continue;
                        const u32 deg = e[j].x;
                        if (deg <= kInline) {
                            if (deg > 0) lds_set(s_new, e[j].y);
                            if (deg > 1) lds_set(s_new, e[j].z);
                            if (deg > 2) lds_set(s_new, e[j].w);
                        } else if (e[j].w != kNoBit) {  // a bit row: listed for the wave
                            const u32 q = atomicAdd(&s_hn[wv], 1u);
                            if (q >= kHList) {
                                full = true;  // this bit and the rest wait for the next round
                                continue;
                            }
                            s_h[wv][q] = e[j].w;
                        } else {  // a walk: listed for the wave
                            const u32 q = atomicAdd(&s_mn[wv], 1u);
                            if (q >= kMList) {
                                full = true;
                                continue;
                            }
                            s_mo[wv][q] = (u64)e[j].y | ((u64)e[j].z << 32);
                            s_mx[wv][q] = deg;
                        }
                        v[k] &= ~(1ull << (bb[j] - w64));  // done with this bit
                    }
                    if (full) break;
                }
            }
            wave_sync_lds();
            // the listed walks, flattened: 64 successors a step
            const u32 nm = min(s_mn[wv], kMList);
            if (nm) {
                const u32 c0 = lane < (int)nm ? s_mx[wv][lane] : 0u;
                const u32 c1 = lane + 64 < (int)nm ? s_mx[wv][lane + 64] : 0u;
                const u64 i0 = wave_inclusive_scan(c0);
                const u64 tot0 = __shfl(i0, 63, 64);
                const u64 i1 = wave_inclusive_scan(c1) + tot0;
                const u64 tot = __shfl(i1, 63, 64);
                wave_sync_lds();
                if (lane < (int)nm) s_mx[wv][lane] = (u32)(i0 - c0);
                if (lane + 64 < (int)nm) s_mx[wv][lane + 64] = (u32)(i1 - c1);
                wave_sync_lds();
                for (u64 t0 = 0; t0 < tot; t0 += 128) {
                    u32 y[2];
#pragma unroll
                    for (int z = 0; z < 2; ++z) {
                        const u64 t = t0 + 64u * z + lane;
                        y[z] = kNoBit;
                        if (t < tot) {
                            u32 lo = 0, hi = nm - 1;  // the last list starting at or before t
                            while (lo < hi) {
                                const u32 mid = (lo + hi + 1) >> 1;
                                if (s_mx[wv][mid] <= t) lo = mid;
                                else hi = mid - 1;
                            }
                            y[z] = a.succ_idx[s_mo[wv][lo] + (t - s_mx[wv][lo])];
                        }
                    }
#pragma unroll
                    for (int z = 0; z < 2; ++z)
                        if (y[z] != kNoBit) lds_set(s_new, y[z]);
                }
            }
            // the listed bit rows, ORed into the whole LDS row
            const u32 nh = min(s_hn[wv], kHList);
            for (u32 h = 0; h < nh; h += 2) {
                const u64 *r1 = a.heavy + (u64)s_h[wv][h] * a.rw;
                const bool two = h + 1 < nh;
                const u64 *r2 = a.heavy + (u64)s_h[wv][two ? h + 1 : h] * a.rw;
                for (u64 w = lane; w < a.rw; w += 64) {
                    const u64 x = r1[w] | (two ? r2[w] : 0ull);
                    if (x) atomicOr((unsigned long long *)&s_new[w], (unsigned long long)x);
                }
            }
            u64 left = 0;
#pragma unroll
            for (int k = 0; k < KB; ++k) left |= v[k];
            more = __ballot(left != 0) != 0;
        }
        __syncthreads();
        u64 cnt = 0;
        for (u64 w = threadIdx.x; w < a.rw; w += DB) {
            const u64 u = s_new[w];
            U[(u64)r * a.rw + w] = u;
            cnt += (u64)__popcll(u);
        }
        u64 tot;
        block_exclusive_scan<DB>(cnt, s_cnt, &tot);  // ends with a barrier: s_new is free again
        if (threadIdx.x == 0) upop[r] = (u32)tot;
        io += 2;
    }
    if (threadIdx.x == 0 && io) atomicAdd(&st->exp_io, (unsigned long long)io);
}

// One body's output rows, a wave per work item (row x, a slice of its pred
// list), items strided over the grid's waves.  TC: NEW = OR of D[y] over
// x's preds.  SG: T = OR of D[a] (the tmp row, for the produced count) and
// NEW = OR of U[a] over x's preds.  A row with one item takes the set
// difference at once; the slices of a split row OR into the row's
// accumulators and the last slice to finish takes it.
template <int K, bool SG>
__global__ __launch_bounds__(DB) void k_dense_combine(DenseArgs a, const u32 *pop_in,
                                                      const u64 *D_in, const u32 *upop,
                                                      const u64 *U, u32 *pop_out, u64 *D_out,
                                                      DenseState *st, u64 *counts) {
    if (st->stop != kDnRun) return;  // the same for every block
    // T = V in a symmetric run's column bodies (read before any block can
    // finish: the epilogue sets the next body's flag)
    const bool vt = SG && a.sym && st->col != 0;
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    u64 w_new = 0, w_raw = 0, w_io = 0, w_rn = 0;
    for (u64 i = wave; i < a.n_items; i += n_waves) {
        const uint4 it = a.items[i];
        const u32 x = it.x, g = it.w;
        const u64 pb = a.pred_off[x] + it.y, pe = a.pred_off[x] + it.z;
        u64 acc[K], acct[SG ? K : 1];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0;
#pragma unroll
        for (int k = 0; k < (SG ? K : 1); ++k) acct[k] = 0;
        u64 raw = 0, flagged = 0;
        bool any = false;
        for (u64 base = pb; base < pe; base += 64) {
            const u64 j = base + lane;
            u32 y = 0, p = 0, q = 0;
            if (j < pe) {
                y = a.pred_idx[j];
                p = pop_in[y];
                if (SG && p) q = upop[y];
            }
            if (!SG) raw += p;
            const u64 m = __ballot(p != 0);
            if (m) any = true;
            if (!vt) flagged += (u64)__popcll(m);
            if constexpr (SG) {
                if (!vt) or_rows<K>(D_in, a.rw, m, y, acct);
                const u64 mu = __ballot(q != 0);
                flagged += (u64)__popcll(mu);
                or_rows<K>(U, a.rw, mu, y, acc);
            } else {
                or_rows<K>(D_in, a.rw, m, y, acc);
            }
        }
        if (g == 0xffffffffu) {  // the whole row
            u64 total = 0;
            if (any) {
                if constexpr (SG) raw += vt ? v_dot<K>(a, x) : plane_dot<K>(a, acct);
                total = wave_sum(diff_row<K>(a, x, acc, D_out));
                w_io += flagged + (vt ? 3 : 2);
            }
            if (lane == 0) pop_out[x] = (u32)total;
            w_new += total;
            w_rn += total ? 1 : 0;
            w_raw += raw;
            continue;
        }
        // a slice of a split row
        u64 *gu = a.GU + (u64)g * a.rw;
        u64 *gt = SG ? a.GT + (u64)g * a.rw : nullptr;
        if (any) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                if (K > 1 || w < a.rw) {
                    if (acc[k]) atomicOr((unsigned long long *)&gu[w], (unsigned long long)acc[k]);
                    if constexpr (SG)
                        if (!vt && acct[k])
                            atomicOr((unsigned long long *)&gt[w], (unsigned long long)acct[k]);
                }
            }
            w_io += flagged;
        }
        if (!SG) w_raw += raw;
        __threadfence();  // release this slice's ORs before counting it finished
        u32 old = 0;
        if (lane == 0) old = atomicAdd(&a.slice_done[g], 1u);
        old = (u32)__shfl((int)old, 0, 64);
        if (old + 1 != a.n_slices[g]) continue;  // the same for the whole wave
        __threadfence();  // acquire: every slice's ORs precede its count
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const u64 w = (u64)lane + 64u * k;
            acc[k] = 0;
            if (K > 1 || w < a.rw) {
                acc[k] = __hip_atomic_load(&gu[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&gu[w], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if constexpr (SG) {
                    if (!vt) {
                        acct[k] = __hip_atomic_load(&gt[w], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(&gt[w], 0ull, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        if constexpr (SG) w_raw += vt ? v_dot<K>(a, x) : plane_dot<K>(a, acct);
        const u64 total = wave_sum(diff_row<K>(a, x, acc, D_out));
        if (vt) w_io += 1;
        if (lane == 0) {
            pop_out[x] = (u32)total;
            a.slice_done[g] = 0;
        }
        w_new += total;
        w_rn += total ? 1 : 0;
        w_io += 2;
    }
    dense_block_end(st, counts, w_new, wave_sum(w_raw), w_io, w_rn);
}

// Appends the last body's new tuples (the D rows with a non-zero pop) to the
// log once their count is known, or asks the host to grow the log first.
// Work items are 64-word chunks of the active rows (a dense row is spread
// over waves); each block takes a contiguous range of them and each wave a
// contiguous part of it: a counting pass, one cursor atomic per block, then
// positions from a block scan of the waves' sums and a running sum.
// Nothing to do when the body launched before it did not run (appended ==
// body).  Every block reads the same state: only the last block to finish
// updates `appended`, after all of them have read it.
template <int K>
__global__ __launch_bounds__(DB) void k_dense_append(DenseArgs a, const u32 *pop, const u64 *D,
                                                     DenseState *st, const u64 *counts) {
    __shared__ u64 s_w[DWV + 1];
    __shared__ unsigned long long s_base;
    __shared__ bool s_last;
    __shared__ uint16_t s_buf[DWV][4096];
    const u32 body = st->body;
    if (st->appended >= body) return;
    const u64 nn = counts[kCounts * (body - 1)];
    const u64 base0 = st->append_base;
    if (nn > st->log_cap || base0 > st->log_cap - nn) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && st->stop == kDnRun) st->stop = kDnGrow;
        return;
    }
    // items: (active row, chunk of 64 words), in row order; a block takes a
    // contiguous range of them, each wave a contiguous part of it
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kb = 6 + __ffsll((unsigned long long)a.rw) - 1;
    const u64 cpr = a.rw >= 64 ? a.rw / 64 : 1;  // chunks per row
    const u64 n_items = a.n_act * cpr;
    const u64 per_block = (n_items + gridDim.x - 1) / gridDim.x;
    const u64 r0 = min(n_items, (u64)blockIdx.x * per_block);
    const u64 r1 = min(n_items, r0 + per_block);
    const u64 per_wave = (r1 - r0 + DWV - 1) / DWV;
    const u64 w0 = min(r1, r0 + (u64)wv * per_wave), w1 = min(r1, w0 + per_wave);
    u64 sum = 0;
    for (u64 i = w0; i < w1; ++i) {
        const u32 x = a.act[i / cpr];
        if (pop[x] == 0) continue;  // the same for the whole wave
        const u64 w = (i % cpr) * 64 + lane;
        if (w < a.rw) sum += (u64)__popcll(D[(u64)x * a.rw + w]);
    }
    sum = wave_sum(sum);
    u64 tot;
    const u64 ex = block_exclusive_scan<DB>(lane == 0 ? sum : 0ull, s_w, &tot);
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(&st->cursor, (unsigned long long)tot) : 0ull;
    __syncthreads();
    u64 at = s_base + (u64)__shfl((long long)ex, 0, 64);
    for (u64 i = w0; i < w1; ++i) {
        const u32 x = a.act[i / cpr];
        if (pop[x] == 0) continue;
        const u64 w0 = (i % cpr) * 64, w = w0 + lane;
        const u64 v = w < a.rw ? D[(u64)x * a.rw + w] : 0ull;
        at += stage_write(v, ((u64)x << kb) + w0 * 64, kb, a.log, at, ~0ull, s_buf[wv]);
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

// ---- SG's column expansion.  U = D R2 (R2 the bit matrix of r2) is also
// (R2^T D^T)^T: transpose D, OR the rows of D^T over each column's r2 pred
// list, transpose back.  That is a fixed few passes over the bit matrix plus
// |r2| row reads, whatever the delta's size, where the expansion above
// costs a successor-table probe per delta bit and an LDS OR per successor:
// large deltas expand by columns (DenseState::col, set by the previous
// body's epilogue from its new-tuple count), small ones bit by bit.

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int m) {
    const int lo = __shfl_xor((int)(u32)v, m, 64);
    const int hi = __shfl_xor((int)(u32)(v >> 32), m, 64);
    return ((u64)(u32)hi << 32) | (u64)(u32)lo;
}

// Transposes the 64 x 64 bit block held a row per lane (bit c of lane r =
// element (r, c)): afterwards lane c holds column c (bit r = element (r, c)).
// Six block swaps (the recursive transpose of Hacker's Delight, 7-3): at
// step s the elements whose row and column differ in bit s trade places.
__device__ __forceinline__ u64 transpose64(u64 v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int s = 32 >> i;
        const u64 m = i == 0   ? 0x00000000ffffffffull
                      : i == 1 ? 0x0000ffff0000ffffull
                      : i == 2 ? 0x00ff00ff00ff00ffull
                      : i == 3 ? 0x0f0f0f0f0f0f0f0full
                      : i == 4 ? 0x3333333333333333ull
                               : 0x5555555555555555ull;  // the columns with bit s clear
        const u64 x = shfl_xor_u64(v, s);
        v = (lane & s) ? (((x >> s) & m) | (v & ~m)) : ((v & m) | ((x & m) << s));
    }
    return v;
}

constexpr int TB = 512;  // transpose block: 8 waves
constexpr int kTW = 8;   // a tile is 64 kTW rows x kTW words (one 64-bit block per wave and word)

// dst = src^T for the 2^kb x 2^kb bit matrix src (rows of rw words), a tile
// per block (grid (rw / kTW)^2).  Wave w loads rows 64 w .. 64 w + 63 of the
// tile, 64 contiguous bytes a lane, and transposes its kTW bit blocks in
// registers; LDS regroups them so that each output row's kTW words (64
// contiguous bytes) are stored by one lane.  Rows with src_pop 0 read as
// zero (their words may be stale).  dst_pop[r] += the bits stored in row r
// (dst_pop starts at zero); zero_pop[0, n_zero) is cleared on the way.  With
// st, it runs only in a column body of a running dense loop.
__global__ __launch_bounds__(TB) void k_bt_transpose(const u64 *src, const u32 *src_pop, u64 rw,
                                                     u64 *dst, u32 *dst_pop, u32 *zero_pop,
                                                     u64 n_zero, DenseState *st) {
    __shared__ u64 s_t[kTW][64][kTW + 1];
    __shared__ u32 s_live[TB / 64];
    if (st && (st->stop != kDnRun || !st->col)) return;  // the same for every block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (u64 i = (u64)blockIdx.x * TB + threadIdx.x; i < n_zero; i += (u64)gridDim.x * TB)
        zero_pop[i] = 0;
    const u64 tpd = rw / kTW;
    const u64 ti = blockIdx.x / tpd, tj = blockIdx.x % tpd;
    const u64 row = ti * (64 * kTW) + (u64)wv * 64 + lane;
    u64 v[kTW];
    const bool live = src_pop == nullptr || src_pop[row] != 0;
    if (live) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(src + row * rw + tj * kTW);
#pragma unroll
        for (int j = 0; j < kTW / 2; ++j) {
            const ulonglong2 q = p[j];
            v[2 * j] = q.x;
            v[2 * j + 1] = q.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kTW; ++j) v[j] = 0;
    }
    const u64 lm = __ballot(live);
    if (lane == 0) s_live[wv] = (u32)__popcll(lm);
#pragma unroll
    for (int j = 0; j < kTW; ++j)
        s_t[j][lane][wv] = __ballot(v[j] != 0) ? transpose64(v[j]) : 0ull;
    __syncthreads();
    // wave w stores bit block w of every row tile: output rows 64 w + lane
    const u64 orow = tj * (64 * kTW) + (u64)wv * 64 + lane;
    u64 o[kTW];
    u32 c = 0;
#pragma unroll
    for (int t = 0; t < kTW; ++t) {
        o[t] = s_t[wv][lane][t];
        c += (u32)__popcll(o[t]);
    }
    ulonglong2 *q = reinterpret_cast<ulonglong2 *>(dst + orow * rw + ti * kTW);
#pragma unroll
    for (int j = 0; j < kTW / 2; ++j) q[j] = make_ulonglong2(o[2 * j], o[2 * j + 1]);
    if (dst_pop && c) atomicAdd(&dst_pop[orow], c);
    if (st && threadIdx.x == 0) {
        u64 segs = TB;  // the stored ones
        for (int w = 0; w < TB / 64; ++w) segs += s_live[w];
        atomicAdd(&st->tr_io, (unsigned long long)segs);
    }
}

// Sets *asym when the bit matrix M (rows of rw words; with pop, rows whose
// pop is 0 read as zero) differs from its transpose: each tile is
// transposed as k_bt_transpose does and compared with the tile it would be
// stored over.  Grid (rw / kTW)^2.  Returns at once when `same` (given
// only when M was set from as many keys as the checked bitmap has tuples)
// records that those keys were distinct and all in the checked bitmap.
__global__ __launch_bounds__(TB) void k_bt_asym(const u64 *M, const u32 *pop, u64 rw,
                                                u32 *asym, const DenseState *same) {
    __shared__ u64 s_t[kTW][64][kTW + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u64 tpd = rw / kTW;
    const u64 ti = blockIdx.x / tpd, tj = blockIdx.x % tpd;
    // tile (ti, tj) against (tj, ti) is the same test as (tj, ti) against
    // (ti, tj): the upper triangle suffices
    if (ti > tj) return;
    // M holds exactly the tuples of a matrix checked already: as many
    // strictly increasing (so distinct) keys as that matrix has tuples, all
    // of them in it
    if (same && same->d0_loose == 0 && same->d0_out == 0) return;
    auto load = [&](u64 row, u64 w0, u64 *v) {
        if (pop == nullptr || pop[row] != 0) {
            const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(M + row * rw + w0);
#pragma unroll
            for (int j = 0; j < kTW / 2; ++j) {
                const ulonglong2 q = p[j];
                v[2 * j] = q.x;
                v[2 * j + 1] = q.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kTW; ++j) v[j] = 0;
        }
    };
    u64 v[kTW];
    load(ti * (64 * kTW) + (u64)wv * 64 + lane, tj * kTW, v);
#pragma unroll
    for (int j = 0; j < kTW; ++j)
        s_t[j][lane][wv] = __ballot(v[j] != 0) ? transpose64(v[j]) : 0ull;
    __syncthreads();
    load(tj * (64 * kTW) + (u64)wv * 64 + lane, ti * kTW, v);
    bool diff = false;
#pragma unroll
    for (int t = 0; t < kTW; ++t) diff |= s_t[wv][lane][t] != v[t];
    if (__ballot(diff) && lane == 0) atomicOr(asym, 1u);
}

// One static relation's rows as OR work: row x's sources are idx[off[x],
// off[x + 1]); items as the combine's (x, first, end, split index or ~0),
// with the split rows' slice counts, finished-slice counters and
// accumulators G.
struct OrRows {
    const u64 *off;
    const u32 *idx;
    const uint4 *items;
    u64 n_items;
    const u32 *n_slices;
    u32 *slice_done;
    u64 *G;
};

// dst[x] = OR of src[b] over x's sources b (sources with src_pop 0 skipped),
// dst_pop[x] = |dst[x]|, a wave per item (rows written only when some source
// is not empty: a reader skips rows with pop 0).
template <int K>
__global__ __launch_bounds__(DB) void k_dense_or_rows(OrRows o, u64 rw, const u32 *src_pop,
                                                      const u64 *src, u32 *dst_pop, u64 *dst,
                                                      u32 *zero_pop, u64 n_zero,
                                                      DenseState *st) {
    __shared__ u64 s_io[DWV];
    if (st->stop != kDnRun || !st->col) return;  // the same for every block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n_zero; i += (u64)gridDim.x * DB)
        zero_pop[i] = 0;
    const u64 wave = (u64)blockIdx.x * DWV + wv;
    const u64 n_waves = (u64)gridDim.x * DWV;
    u64 w_io = 0;
    for (u64 i = wave; i < o.n_items; i += n_waves) {
        const uint4 it = o.items[i];
        const u32 x = it.x, g = it.w;
        const u64 pb = o.off[x] + it.y, pe = o.off[x] + it.z;
        u64 acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0;
        bool any = false;
        for (u64 base = pb; base < pe; base += 64) {
            const u64 j = base + lane;
            u32 y = 0, p = 0;
            if (j < pe) {
                y = o.idx[j];
                p = src_pop[y];
            }
            const u64 m = __ballot(p != 0);
            if (m) any = true;
            w_io += (u64)__popcll(m);
            or_rows<K>(src, rw, m, y, acc);
        }
        u64 *drow = dst + (u64)x * rw;
        if (g != 0xffffffffu) {  // a slice of a split row
            u64 *gr = o.G + (u64)g * rw;
            if (any) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const u64 w = (u64)lane + 64u * k;
                    if ((K > 1 || w < rw) && acc[k])
                        atomicOr((unsigned long long *)&gr[w], (unsigned long long)acc[k]);
                }
            }
            __threadfence();  // release this slice's ORs before counting it finished
            u32 old = 0;
            if (lane == 0) old = atomicAdd(&o.slice_done[g], 1u);
            old = (u32)__shfl((int)old, 0, 64);
            if (old + 1 != o.n_slices[g]) continue;  // the same for the whole wave
            __threadfence();  // acquire: every slice's ORs precede its count
            any = false;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                acc[k] = 0;
                if (K > 1 || w < rw) {
                    acc[k] = __hip_atomic_load(&gr[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&gr[w], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (acc[k]) any = true;
            }
            any = __ballot(any) != 0;
            if (lane == 0) o.slice_done[g] = 0;
        }
        u64 c = 0;
        if (any) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const u64 w = (u64)lane + 64u * k;
                if (K > 1 || w < rw) {
                    drow[w] = acc[k];
                    c += (u64)__popcll(acc[k]);
                }
            }
            w_io += 1;
        }
        c = wave_sum(c);
        if (lane == 0) dst_pop[x] = (u32)c;
    }
    if (lane == 0) s_io[wv] = w_io;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int w = 0; w < DWV; ++w) t += s_io[w];
        if (t) atomicAdd(&st->or_io, (unsigned long long)t);
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

// Bit-planes of r2's row lengths: bit b of plane j is bit j of off[b+1] -
// off[b] (a wave per 64-bit word: lane t holds row 64 w + t, one ballot per
// plane).
__global__ __launch_bounds__(DB) void k_dense_planes(const u64 *off, u64 rw, int n_planes,
                                                     u64 *planes) {
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    for (u64 w = wave; w < rw; w += n_waves) {
        const u64 b = w * 64 + lane;
        const u64 d = off[b + 1] - off[b];
        for (int j = 0; j < n_planes; ++j) {
            const u64 m = __ballot((d >> j) & 1ull);
            if (lane == 0) planes[(u64)j * rw + w] = m;
        }
    }
}

// Delta keys -> per-row counts (duplicates counted).
__global__ __launch_bounds__(DB) void k_dense_mark(const u64 *keys, u64 n, int kb, u32 *pop,
                                                   DenseState *st) {
    // one add per run of equal rows within a wave's 64 keys (delta keys
    // usually come sorted: a row's keys would all add to one counter)
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    for (u64 base = wave * 64; base < n; base += n_waves * 64) {
        const u64 i = base + lane;
        bool valid = false;
        u64 row = ~0ull - (u64)lane;  // distinct per lane when not a key
        if (i < n) {
            const u64 k = keys[i];
            const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
            if ((c0 >> kb) || (c1 >> kb)) {
                atomicOr(&st->bad, 1u);
            } else {
                valid = true;
                row = c0;
            }
        }
        const u64 prev = __shfl_up(row, 1, 64);
        const bool head = lane == 0 || row != prev;
        const u64 heads = __ballot(head);
        const u64 above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
        const int next = above ? __ffsll((unsigned long long)above) - 1 : 64;
        if (head && valid) atomicAdd(&pop[row], (u32)(next - lane));
    }
}

__global__ __launch_bounds__(DB) void k_dense_zero_rows(const u32 *pop, u64 rows, u64 rw, u64 *D) {
    const u64 n = rows * rw;
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB)
        if (pop[i / rw]) D[i] = 0ull;
}

// Sets the delta's bits in D (atomics whose result is not needed: nothing
// waits on them).  With `full` (a bitmap of the same layout), flags
// st->d0_out when a key is not in `full` and st->d0_loose when the keys are
// not strictly increasing (then they may repeat).
__global__ __launch_bounds__(DB) void k_dense_set_bits(const u64 *keys, u64 n, int kb, u64 rw,
                                                       u64 *D, const u64 *full,
                                                       DenseState *st) {
    const int lane = threadIdx.x & 63;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    bool out = false, loose = false;
    for (u64 base = wave * 64; base < n; base += n_waves * 64) {
        const u64 i = base + lane;
        u64 k = 0;
        if (i < n) {
            k = keys[i];
            const u64 c0 = k >> 32, c1 = k & 0xffffffffull;
            if (!(c0 >> kb) && !(c1 >> kb)) {
                const u64 wi = c0 * rw + (c1 >> 6), bit = 1ull << (c1 & 63);
                atomicOr((unsigned long long *)&D[wi], bit);
                if (full && !(full[wi] & bit)) out = true;
            }
        }
        if (full) {
            u64 prev = __shfl_up(k, 1, 64);
            if (lane == 0 && i > 0 && i < n) prev = keys[i - 1];
            if (i > 0 && i < n && k <= prev) loose = true;
        }
    }
    if (!full) return;
    const bool any_out = __ballot(out) != 0, any_loose = __ballot(loose) != 0;
    if (lane == 0) {
        if (any_out) atomicOr(&st->d0_out, 1u);
        if (any_loose) atomicOr(&st->d0_loose, 1u);
    }
}

// SG's successor entries: one 16-byte load gives a row's successors when it
// has at most kInline of them, else its length, CSR offset and bit-row index
// (hidx[b], kNoBit when the list is walked).
__global__ __launch_bounds__(DB) void k_succ_tab(const u64 *off, const u32 *idx, const u32 *hidx,
                                                 u64 rows, uint4 *tab) {
    for (u64 b = (u64)blockIdx.x * DB + threadIdx.x; b < rows; b += (u64)gridDim.x * DB) {
        const u64 so = off[b], deg = off[b + 1] - so;
        uint4 e = make_uint4((u32)deg, 0, 0, 0);
        if (deg <= kInline) {
            if (deg > 0) e.y = idx[so];
            if (deg > 1) e.z = idx[so + 1];
            if (deg > 2) e.w = idx[so + 2];
        } else {
            e.y = (u32)so;
            e.z = (u32)(so >> 32);
            e.w = hidx[b];
        }
        tab[b] = e;
    }
}

// Bit rows of the heavy lists: row h gets the successors of heavy[h].
__global__ __launch_bounds__(DB) void k_heavy_rows(const u64 *off, const u32 *idx,
                                                   const u32 *heavy_b, u64 n_heavy, u64 rw,
                                                   u64 *rows_out) {
    for (u64 h = blockIdx.x; h < n_heavy; h += gridDim.x) {
        const u32 b = heavy_b[h];
        for (u64 t = off[b] + threadIdx.x; t < off[b + 1]; t += DB) {
            const u32 y = idx[t];
            atomicOr((unsigned long long *)&rows_out[h * rw + (y >> 6)], 1ull << (y & 63));
        }
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

__global__ __launch_bounds__(DB) void k_fill_u64(u64 *p, u64 n, u64 v) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n; i += (u64)gridDim.x * DB) p[i] = v;
}

// ---- joins on bit matrices.  A column-map join of two binary relations on
// their column 0 into a bitmap set, when the id domain is small: with M the
// bit matrix of one side, every tuple (b, d) of the other ORs row M[b] into
// row d of the result R (out(a, c) :- outer(b, a), inner(b, c) is R = O^T I:
// M = I, the outer tuples scatter; with the columns swapped, M = O and the
// inner tuples scatter).  NEW = R minus the set; the set takes NEW; NEW's
// keys are appended in key order.  The join's produced count is the count
// pass's exact one.

// Bit c of row r of M per key (r, c).  Lanes on one word (a run of
// neighbouring lanes: a sorted relation puts up to 64 keys of a wave on a
// few words) OR their bits together first, and the run's first lane makes
// the one atomic.
__global__ __launch_bounds__(DB) void k_dj_set(const u64 *keys, u64 n, u64 rw, u64 *M) {
    const int lane = threadIdx.x & 63;
    const u64 stride = (u64)gridDim.x * DB;
    // every lane of a wave runs the same rounds (shuffles below)
    for (u64 base = (u64)blockIdx.x * DB + (threadIdx.x & ~63u); base < n; base += stride) {
        const u64 i = base + lane;
        u64 wi = ~0ull, m = 0;
        if (i < n) {
            const u64 k = keys[i];
            const u64 r = k >> 32, c = k & 0xffffffffull;
            wi = r * rw + (c >> 6);
            m = 1ull << (c & 63);
        }
        const int prev = lane > 0 ? lane - 1 : 0;
        const u64 pwi = ((u64)(u32)__shfl((int)(wi >> 32), prev, 64) << 32) |
                        (u64)(u32)__shfl((int)(u32)wi, prev, 64);
        const bool head = lane == 0 || pwi != wi;
        const u64 heads = __ballot(head);
        const u64 le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        const int rid = __popcll(heads & le);
        u64 acc = m;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int up = lane + d < 64 ? lane + d : lane;
            // every lane shuffles (a lane outside the branch would not)
            const u64 x = ((u64)(u32)__shfl((int)(acc >> 32), up, 64) << 32) |
                          (u64)(u32)__shfl((int)(u32)acc, up, 64);
            const int ru = __shfl(rid, up, 64);
            if (lane + d < 64 && ru == rid) acc |= x;
        }
        if (head && wi != ~0ull && acc)
            atomicOr((unsigned long long *)&M[wi], (unsigned long long)acc);
    }
}

// R[d] |= M[b] for every tuple (b, d) of keys; L = min(64, rw) lanes a tuple.
__global__ __launch_bounds__(DB) void k_dj_scatter(const u64 *keys, u64 n, u64 rw, const u64 *M,
                                                   u64 *R) {
    const int lane = threadIdx.x & 63;
    const int L = rw >= 64 ? 64 : (int)rw;
    const int per = 64 / L;
    const u64 wave = (u64)blockIdx.x * DWV + (threadIdx.x >> 6);
    const u64 n_waves = (u64)gridDim.x * DWV;
    for (u64 t0 = wave * per; t0 < n; t0 += n_waves * per) {
        const u64 t = t0 + (u64)(lane / L);
        if (t >= n) continue;
        const u64 k = keys[t];
        const u64 *src = M + (k >> 32) * rw;
        u64 *dst = R + (k & 0xffffffffull) * rw;
        for (u64 w = (u64)(lane % L); w < rw; w += L) {
            const u64 v = src[w];
            if (v) atomicOr((unsigned long long *)&dst[w], (unsigned long long)v);
        }
    }
}

// The same two steps for a matrix that fits in LDS (a small id domain, where
// global atomics from every tuple would pile up on a few thousand words):
// block g builds its own copy of the matrix in LDS from its contiguous share
// of the tuples (SCATTER: row M[b] ORed into row d per tuple (b, d); else bit
// c of row r per tuple (r, c)) and writes it to part[g]; k_dj_fold ORs the
// copies.
template <bool SCATTER>
__global__ __launch_bounds__(DB) void k_dj_lds(const u64 *keys, u64 n, u64 rw, const u64 *M,
                                               u64 *part, u64 words) {
    extern __shared__ u64 s_m[];
    for (u64 i = threadIdx.x; i < words; i += DB) s_m[i] = 0;
    __syncthreads();
    const u64 chunk = (n + gridDim.x - 1) / gridDim.x;
    const u64 t0 = (u64)blockIdx.x * chunk, t1 = min(n, t0 + chunk);
    if constexpr (SCATTER) {
        const u32 L = rw >= 64 ? 64u : (u32)rw;
        for (u64 t = t0 + threadIdx.x / L; t < t1; t += DB / L) {
            const u64 k = keys[t];
            const u64 *src = M + (k >> 32) * rw;
            u64 *dst = s_m + (k & 0xffffffffull) * rw;
            for (u64 w = threadIdx.x % L; w < rw; w += L) {
                const u64 v = src[w];
                if (v) atomicOr((unsigned long long *)&dst[w], (unsigned long long)v);
            }
        }
    } else {
        for (u64 t = t0 + threadIdx.x; t < t1; t += DB) {
            const u64 k = keys[t];
            const u64 r = k >> 32, c = k & 0xffffffffull;
            atomicOr((unsigned long long *)&s_m[r * rw + (c >> 6)], 1ull << (c & 63));
        }
    }
    __syncthreads();
    u64 *out = part + (u64)blockIdx.x * words;
    for (u64 i = threadIdx.x; i < words; i += DB) out[i] = s_m[i];
}

// out[i] |= OR of the G copies' word i (out starts at zero): kDjFold threads a
// word, each ORing every kDjFold-th copy, one atomic each.
constexpr u32 kDjFold = 16;
__global__ __launch_bounds__(DB) void k_dj_fold(const u64 *part, u32 G, u64 words, u64 *out) {
    const u64 n = words * kDjFold;
    for (u64 t = (u64)blockIdx.x * DB + threadIdx.x; t < n; t += (u64)gridDim.x * DB) {
        const u64 i = t % words;
        u64 v = 0;
        for (u32 g = (u32)(t / words); g < G; g += kDjFold) v |= part[(u64)g * words + i];
        if (v) atomicOr((unsigned long long *)&out[i], (unsigned long long)v);
    }
}

// R := R & ~S, S |= R (one thread a word).
__global__ __launch_bounds__(DB) void k_dj_diff(u64 *R, u64 *S, u64 n_words) {
    for (u64 i = (u64)blockIdx.x * DB + threadIdx.x; i < n_words; i += (u64)gridDim.x * DB) {
        const u64 r = R[i];
        if (!r) continue;
        const u64 s = S[i];
        const u64 nw = r & ~s;
        R[i] = nw;
        if (nw) S[i] = s | nw;
    }
}

constexpr int kDenseJoinMaxBits = 13;  // matrices of at most 8 MiB
constexpr u64 kDjLds = 64 * 1024;      // a matrix this small is built in LDS copies

bool dense_join_pays(const gdl_ctx *ctx, const gdl_hashset *hs, gdl_colmap map, gdl_pred pred,
                     int prefix_cols, u64 n_inner, u64 n_outer, u64 produced) {
    const int kb = (int)hs->key_bits;
    // a bitmap set over a small domain that holds every loaded id, a one-column
    // join, a map of two non-key columns, no predicate
    if (kb < 6 || kb > kDenseJoinMaxBits || ctx->id_bits <= 0 || ctx->id_bits > kb) return false;
    if (ctx->dense_mode == 0 || prefix_cols != 1 || map.arity != 2 || pred.op != GDL_PRED_NONE)
        return false;
    const bool o_i = map.src[0] == GDL_COL_OUTER1 && map.src[1] == GDL_COL_INNER1;
    const bool i_o = map.src[0] == GDL_COL_INNER1 && map.src[1] == GDL_COL_OUTER1;
    if (!o_i && !i_o) return false;
    if (ctx->dense_mode >= 2) return true;
    const u64 rows = 1ull << kb, rw = 1ull << (kb - 6), words = rows * rw;
    // the scattered side: the smaller one where the result can be transposed
    const u64 n_scatter = kb >= 9 ? std::min(n_outer, n_inner) : (o_i ? n_outer : n_inner);
    // bytes moved: the matrices cleared, built, differenced and read (~6 passes)
    // and a row read + OR per scattered tuple, against ~16 bytes per produced
    // tuple one by one (inner read, set probe), with a margin for the passes'
    // launches and host wait
    const double dense_bytes = 8.0 * (6.0 * (double)words + (double)n_scatter * (double)rw);
    return 16.0 * (double)produced > 4.0 * dense_bytes + 4.0e6;
}

bool dense_join_insert(gdl_ctx *ctx, const u64 *d_inner, u64 n_inner, const u64 *d_outer,
                       u64 n_outer, gdl_colmap map, gdl_pred pred, int prefix_cols,
                       gdl_hashset *hs, u64 produced, u64 **d_new, u64 *new_cap, u64 *n_new,
                       u64 *n_raw, gdl_status *status) {
    if (!dense_join_pays(ctx, hs, map, pred, prefix_cols, n_inner, n_outer, produced))
        return false;
    const int kb = (int)hs->key_bits;
    const bool o_i = map.src[0] == GDL_COL_OUTER1;
    const u64 rows = 1ull << kb, rw = 1ull << (kb - 6), words = rows * rw;
    // scatter over the outer tuples with M = I (row a of O^T I per outer tuple
    // (b, a)) or over the inner ones with M = O (row c of I^T O): whichever side
    // is smaller, transposing when that gives the map's transpose
    const bool scatter_outer = kb >= 9 ? n_outer <= n_inner : o_i;
    const bool transpose = scatter_outer != o_i;
    const u64 n_scatter = scatter_outer ? n_outer : n_inner;
    auto run = [&]() -> gdl_status {
        Scratch scr(ctx);
        u64 *M, *R;
        GDL_TRY(scr.get(words, &M));
        GDL_TRY(scr.get(words, &R));
        const unsigned gw = dgrid(ctx, words);
        hipLaunchKernelGGL(k_fill_u64, dim3(gw), dim3(DB), 0, ctx->stream, M, words, 0ull);
        GDL_TRY(check_launch(ctx, "k_fill_u64"));
        hipLaunchKernelGGL(k_fill_u64, dim3(gw), dim3(DB), 0, ctx->stream, R, words, 0ull);
        GDL_TRY(check_launch(ctx, "k_fill_u64"));
        const u64 *mk = scatter_outer ? d_inner : d_outer;
        const u64 nm = scatter_outer ? n_inner : n_outer;
        const u64 *sk = scatter_outer ? d_outer : d_inner;
        {
            KTimer kt(ctx, K_JOIN_DENSE, 8.0 * (3.0 * words + nm + n_scatter) +
                                             8.0 * (double)n_scatter * (double)rw);
            if (words * sizeof(u64) <= kDjLds) {
                // block-private copies in LDS, folded
                const u64 cus = (u64)std::max(ctx->cu_count, 1);
                const u32 G = (u32)std::max<u64>(
                    1, std::min<u64>(cus, (std::max(nm, n_scatter) + 4095) / 4096));
                u64 *part;
                GDL_TRY(scr.get((u64)G * words, &part));
                const size_t lds = words * sizeof(u64);
                const unsigned gf = dgrid(ctx, words * kDjFold);
                if (nm) {
                    hipLaunchKernelGGL(k_dj_lds<false>, dim3(G), dim3(DB), lds, ctx->stream, mk,
                                       nm, rw, (const u64 *)nullptr, part, words);
                    GDL_TRY(check_launch(ctx, "k_dj_lds"));
                    hipLaunchKernelGGL(k_dj_fold, dim3(gf), dim3(DB), 0, ctx->stream,
                                       (const u64 *)part, G, words, M);
                    GDL_TRY(check_launch(ctx, "k_dj_fold"));
                }
                hipLaunchKernelGGL(k_dj_lds<true>, dim3(G), dim3(DB), lds, ctx->stream, sk,
                                   n_scatter, rw, (const u64 *)M, part, words);
                GDL_TRY(check_launch(ctx, "k_dj_lds"));
                hipLaunchKernelGGL(k_dj_fold, dim3(gf), dim3(DB), 0, ctx->stream,
                                   (const u64 *)part, G, words, R);
                GDL_TRY(check_launch(ctx, "k_dj_fold"));
            } else {
                if (nm) {
                    hipLaunchKernelGGL(k_dj_set, dim3(dgrid(ctx, nm)), dim3(DB), 0, ctx->stream,
                                       mk, nm, rw, M);
                    GDL_TRY(check_launch(ctx, "k_dj_set"));
                }
                const u64 per = rw >= 64 ? 1 : 64 / rw;
                const u64 sg = std::max<u64>(
                    1, std::min<u64>((n_scatter + per * DWV - 1) / (per * DWV), grid_cap(ctx)));
                hipLaunchKernelGGL(k_dj_scatter, dim3((unsigned)sg), dim3(DB), 0, ctx->stream, sk,
                                   n_scatter, rw, (const u64 *)M, R);
                GDL_TRY(check_launch(ctx, "k_dj_scatter"));
            }
            if (transpose) {  // R^T into M (no longer needed), then swap the names
                hipLaunchKernelGGL(k_bt_transpose, dim3((unsigned)((rw / kTW) * (rw / kTW))),
                                   dim3(TB), 0, ctx->stream, (const u64 *)R, (const u32 *)nullptr,
                                   rw, M, (u32 *)nullptr, (u32 *)nullptr, (u64)0,
                                   (DenseState *)nullptr);
                GDL_TRY(check_launch(ctx, "k_bt_transpose"));
                std::swap(M, R);
            }
            hipLaunchKernelGGL(k_dj_diff, dim3(gw), dim3(DB), 0, ctx->stream, R, (u64 *)hs->slots,
                               words);
            GDL_TRY(check_launch(ctx, "k_dj_diff"));
        }
        // NEW's keys, in key order, after the keys already in *d_new
        const int rounds = xt_rounds(words);
        const u64 tiles = (words + 64ull * rounds - 1) / (64ull * rounds);
        u64 *tc;
        GDL_TRY(scr.get(tiles + 1, &tc));
        {
            KTimer kt(ctx, K_BM_EXTRACT, 8.0 * words);
            GDL_TRY(launch_xt_count(ctx, rounds, (const u64 *)R, words, tiles, nullptr, nullptr, 0,
                                    tc));
        }
        GDL_TRY(scan_inplace(ctx, tc, tiles));
        u64 total = 0;
        GDL_TRY(read_u64(ctx, tc + tiles, &total));
        if (*n_new + total > *new_cap || !*d_new) {
            const u64 cap = *n_new + total + (*n_new + total) / 4 + 1024;
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
        if (total) {
            KTimer kt(ctx, K_BM_EXTRACT, 8.0 * words + 8.0 * total);
            GDL_TRY(launch_xt_write(ctx, rounds, (const u64 *)R, words, tiles, nullptr, nullptr,
                                    0, kb, (const u64 *)tc, *d_new + *n_new, total));
        }
        hs->size += total;
        *n_new += total;
        *n_raw = produced;
        return GDL_OK;
    };
    *status = run();
    return true;
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

// SG: the smallest delta that expands by columns.  Expanding by columns
// moves about (4 rows + |r2|) rows of rw words (two transposes, the row OR);
// expanding bit by bit costs about col_bytes_per_bit() bytes of streaming
// time per delta bit (a successor-table probe, LDS ORs per successor).
static double col_bytes_per_bit() {
    const char *e = std::getenv("GDLOG_DENSE_COL_BYTES_PER_BIT");
    const double v = e ? std::atof(e) : 0.0;
    return v > 0.0 ? v : 1024.0;
}

static u64 col_threshold(int mode, int kb, u64 n2) {
    if (mode == 4 || kb < 9) return ~0ull;  // a transpose tile is 512 rows x 8 words
    if (mode == 3) return 0;
    if (kb < 13) return ~0ull;  // small matrices: the fixed passes do not pay
    const double rw = (double)(1ull << (kb - 6));
    const double bytes = (4.0 * (double)(1ull << kb) + (double)n2) * rw * 8.0;
    return (u64)std::max(1.0, bytes / col_bytes_per_bit());
}

static bool dense_rule_ok(const gdl_hashset *hs, int shape) {
    if (!hs || !hs->slots || hs->key_bits < 6 || hs->key_bits > 17) return false;
    if (shape == GDL_DENSE_SG && hs->key_bits > 16) return false;  // LDS rows of <= 8 KB
    return shape == GDL_DENSE_TC || shape == GDL_DENSE_SG;
}

// The buffers of one body: delta in (D, pop), out, and SG's U.
struct BodyBufs {
    const u32 *pop_in;
    const u64 *D_in;
    u32 *pop_out;
    u64 *D_out;
    u32 *upop;
    u64 *U;
};

// SG's column expansion: r2 grouped by its second column as OR work, the
// OR's output V (= U^T) with its pops, the pops of D^T (stored in U's
// buffer: U is built after D^T is no longer needed).
struct ColBufs {
    bool on;
    OrRows o;
    u64 *V;
    u32 *vpop, *dtpop;
    u64 rows;
    unsigned grid_or, grid_tr;
};

template <int K>
static gdl_status launch_col(gdl_ctx *ctx, const DenseArgs &a, const BodyBufs &bb,
                             const ColBufs &cb, DenseState *st) {
    if (!a.sym) {
        KTimer kt(ctx, K_DENSE_TRANSPOSE, 0.0);  // D^T into U's buffer
        hipLaunchKernelGGL(k_bt_transpose, dim3(cb.grid_tr), dim3(TB), 0, ctx->stream, bb.D_in,
                           bb.pop_in, a.rw, bb.U, cb.dtpop, bb.upop, cb.rows, st);
        GDL_TRY(check_launch(ctx, "k_bt_transpose"));
    }
    {
        // V = R2^T D^T (D^T = D when the run is symmetric: then it clears U's
        // pops, which the transpose below accumulates)
        KTimer kt(ctx, K_DENSE_ORROWS, 0.0);
        const u32 *src_pop = a.sym ? bb.pop_in : (const u32 *)cb.dtpop;
        const u64 *src = a.sym ? bb.D_in : (const u64 *)bb.U;
        hipLaunchKernelGGL(k_dense_or_rows<K>, dim3(cb.grid_or), dim3(DB), 0, ctx->stream, cb.o,
                           a.rw, src_pop, src, cb.vpop, cb.V, a.sym ? bb.upop : (u32 *)nullptr,
                           a.sym ? cb.rows : (u64)0, st);
        GDL_TRY(check_launch(ctx, "k_dense_or_rows"));
    }
    KTimer kt(ctx, K_DENSE_TRANSPOSE, 0.0);  // U = V^T
    hipLaunchKernelGGL(k_bt_transpose, dim3(cb.grid_tr), dim3(TB), 0, ctx->stream, cb.V,
                       cb.vpop, a.rw, bb.U, bb.upop, cb.dtpop, cb.rows, st);
    return check_launch(ctx, "k_bt_transpose");
}

template <int K>
static gdl_status launch_body(gdl_ctx *ctx, int shape, const DenseArgs &a, const BodyBufs &bb,
                              const ColBufs &cb, DenseState *st, u64 *counts, unsigned grid,
                              unsigned grid_exp, bool append) {
    if (shape == GDL_DENSE_TC) {
        KTimer kt(ctx, K_DENSE_COMBINE, 0.0);
        hipLaunchKernelGGL((k_dense_combine<K, false>), dim3(grid), dim3(DB), 0, ctx->stream, a,
                           bb.pop_in, bb.D_in, (const u32 *)nullptr, (const u64 *)nullptr,
                           bb.pop_out, bb.D_out, st, counts);
        GDL_TRY(check_launch(ctx, "k_dense_combine"));
    } else if constexpr (K <= 16) {  // SG rows of at most 2^16 bits (LDS)
        if (cb.on) GDL_TRY(launch_col<K>(ctx, a, bb, cb, st));
        {
            KTimer kt(ctx, K_DENSE_EXPAND, 0.0);
            const size_t lds = (size_t)a.rw * sizeof(u64);
            hipLaunchKernelGGL(k_dense_expand<K>, dim3(grid_exp), dim3(DB), lds, ctx->stream, a,
                               bb.pop_in, bb.D_in, bb.upop, bb.U, st);
            GDL_TRY(check_launch(ctx, "k_dense_expand"));
        }
        KTimer kt(ctx, K_DENSE_COMBINE, 0.0);
        hipLaunchKernelGGL((k_dense_combine<K, true>), dim3(grid), dim3(DB), 0, ctx->stream, a,
                           bb.pop_in, bb.D_in, (const u32 *)bb.upop, (const u64 *)bb.U,
                           bb.pop_out, bb.D_out, st, counts);
        GDL_TRY(check_launch(ctx, "k_dense_combine"));
    } else {
        return fail(ctx, GDL_ERR_UNSUPPORTED, "dense SG bodies: ids of at most 16 bits");
    }
    if (!append) return GDL_OK;
    KTimer kt(ctx, K_DENSE_APPEND, 0.0);
    hipLaunchKernelGGL(k_dense_append<K>, dim3(grid), dim3(DB), 0, ctx->stream, a, bb.pop_out,
                       bb.D_out, st, counts);
    return check_launch(ctx, "k_dense_append");
}

static gdl_status launch_append(gdl_ctx *ctx, int kb, const DenseArgs &a, const u32 *pop,
                                const u64 *D, DenseState *st, const u64 *counts, unsigned grid) {
    KTimer kt(ctx, K_DENSE_APPEND, 0.0);
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
                                const BodyBufs &bb, const ColBufs &cb, DenseState *st,
                                u64 *counts, unsigned grid, unsigned grid_exp, bool append) {
    switch (kb) {
    case 17: return launch_body<32>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    case 16: return launch_body<16>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    case 15: return launch_body<8>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    case 14: return launch_body<4>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    case 13: return launch_body<2>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    default: return launch_body<1>(ctx, shape, a, bb, cb, st, counts, grid, grid_exp, append);
    }
}

// The set-up's host copies of row lists, kept per thread between calls: a
// fresh vector of a few hundred KB faults in its pages on every call.
struct DenseHostScratch {
    std::vector<u64> soff, off, off2;
    std::vector<u32> hb, hidx, act, slices, act2, slices2;
    std::vector<uint4> items, items2;
};

// An event the host waits on, destroyed with the scope.
struct HostEvent {
    hipEvent_t e = nullptr;
    HostEvent() = default;
    HostEvent(const HostEvent &) = delete;
    HostEvent &operator=(const HostEvent &) = delete;
    ~HostEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};

static DenseHostScratch &dense_host_scratch() {
    static thread_local DenseHostScratch h;
    return h;
}

// Work items of rows grouped in CSR form (h_off): the rows with a non-empty
// list, longest first by the bit length of their list length (rows in order
// within one length class: a linear bucket pass, where a comparison sort of
// a few ten thousand rows takes hundreds of microseconds), those of more
// than kSlice entries split into slices; slices gets each split row's slice
// count.  act is reordered likewise.
static void build_items(const std::vector<u64> &h_off, std::vector<u32> &act,
                        std::vector<uint4> &items, std::vector<u32> &slices) {
    auto cls = [&](u32 x) { return 64 - __builtin_clzll((unsigned long long)(h_off[x + 1] - h_off[x]) | 1ull); };
    u64 start[66] = {};
    for (u32 x : act) ++start[65 - cls(x)];
    for (int c = 1; c < 66; ++c) start[c] += start[c - 1];
    static thread_local std::vector<u32> sorted;
    sorted.resize(act.size());
    for (u32 x : act) sorted[start[64 - cls(x)]++] = x;
    act.swap(sorted);
    items.clear();
    slices.clear();
    items.reserve(act.size());
    for (u32 x : act) {
        const u64 c = h_off[x + 1] - h_off[x];
        if (c <= kSlice) {
            items.push_back(make_uint4(x, 0, (u32)c, 0xffffffffu));
            continue;
        }
        const u32 g = (u32)slices.size();
        const u32 ns = (u32)((c + kSlice - 1) / kSlice);
        slices.push_back(ns);
        for (u32 j = 0; j < ns; ++j)
            items.push_back(
                make_uint4(x, j * kSlice, (u32)std::min<u64>(c, (u64)(j + 1) * kSlice), g));
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
    if (ctx->dense_mode >= 2) return 0;
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
    // TC's tuple-at-a-time bodies run stream-ordered too, so a TC run hands a
    // delta that turned sparse back to them; SG's run one iteration per host
    // round trip (hundreds of microseconds), against a few short launches for
    // a dense body on a small delta: an SG run goes on to the fixpoint
    const u64 sparse_below = shape == GDL_DENSE_SG ? 0 : threshold / 2;
    const int kb = (int)hs->key_bits;
    const u64 rows = 1ull << kb, rw = 1ull << (kb - 6);
    DevBufs mem(ctx);
    Scratch scr(ctx);
    Staging stg(ctx);  // the set-up's host copies
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
    u64 *succ_off = nullptr;
    u32 *succ_idx = nullptr;
    if (shape == GDL_DENSE_SG) {
        GDL_TRY(group_by_col(ctx, mem, rule->r2, rule->n2, 0, kb, rows, cnt, &succ_off, &succ_idx,
                             bad));
        a.succ_off = succ_off;
        a.succ_idx = succ_idx;
    }
    // SG's column expansion: r2 grouped by its second column, as OR work
    ColBufs col{};
    const u64 col_above =
        shape == GDL_DENSE_SG ? col_threshold(ctx->dense_mode, kb, rule->n2) : ~0ull;
    u64 *off2 = nullptr;
    u32 *idx2 = nullptr;
    if (col_above != ~0ull)
        GDL_TRY(group_by_col(ctx, mem, rule->r2, rule->n2, 1, kb, rows, cnt, &off2, &idx2, bad));

    // the row offsets and the active rows to the host in one wait, which the
    // first delta's bit rows (enqueued next) run behind
    DenseHostScratch &H = dense_host_scratch();
    std::vector<u64> &h_soff = H.soff;  // SG: r2's row offsets on the host
    std::vector<u64> &h_off = H.off;
    std::vector<u64> &h_off2 = H.off2;
    std::vector<u32> &h_act = H.act;
    u64 n_act = 0;
    u32 h_bad = 0;
    h_off.resize(rows + 1);
    h_act.resize(rows);
    GDL_TRY(stg.later(&n_act, n_act_d, sizeof(u64)));
    GDL_TRY(stg.later(&h_bad, bad, sizeof(u32)));
    GDL_TRY(stg.later(h_off.data(), pred_off, (rows + 1) * sizeof(u64)));
    GDL_TRY(stg.later(h_act.data(), act, rows * sizeof(u32)));
    if (succ_off) {
        h_soff.resize(rows + 1);
        GDL_TRY(stg.later(h_soff.data(), succ_off, (rows + 1) * sizeof(u64)));
    }
    if (off2) {
        h_off2.resize(rows + 1);
        GDL_TRY(stg.later(h_off2.data(), off2, (rows + 1) * sizeof(u64)));
    }
    HostEvent grouped;
    GDL_HIP(ctx, hipEventCreateWithFlags(&grouped.e, hipEventDisableTiming));
    GDL_HIP(ctx, hipEventRecord(grouped.e, ctx->stream));

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
    GDL_TRY(scr.get(kCounts * kChunk, &counts));
    DenseState *st = (DenseState *)st_words;
    GDL_HIP(ctx, hipMemsetAsync(st, 0, sizeof(DenseState), ctx->stream));
    // SG with r1 = r2 (one relation, as in the reference's program): when FULL
    // and the first delta are symmetric, every later delta is too (NEW =
    // R1^T D R1 minus FULL is symmetric when D and FULL are, and FULL grows
    // by NEW), so the column expansion needs no transpose of D.  Checked
    // whenever the column expansion may run.
    const bool sym_cand = col_above != ~0ull && rule->r1 == rule->r2 && rule->n1 == rule->n2;
    {
        KTimer kt(ctx, K_DENSE_SETUP, 24.0 * n_delta + 8.0 * rows * rw);
        hipLaunchKernelGGL(k_dense_mark, dim3(dgrid(ctx, n_delta)), dim3(DB), 0, ctx->stream,
                           d_delta, n_delta, kb, pop[0], st);
        GDL_TRY(check_launch(ctx, "k_dense_mark"));
        hipLaunchKernelGGL(k_dense_zero_rows, dim3(dgrid(ctx, rows * rw)), dim3(DB), 0,
                           ctx->stream, pop[0], rows, rw, D[0]);
        GDL_TRY(check_launch(ctx, "k_dense_zero_rows"));
        hipLaunchKernelGGL(k_dense_set_bits, dim3(dgrid(ctx, n_delta)), dim3(DB), 0, ctx->stream,
                           d_delta, n_delta, kb, rw, D[0],
                           sym_cand ? (const u64 *)hs->slots : (const u64 *)nullptr, st);
        GDL_TRY(check_launch(ctx, "k_dense_set_bits"));
    }
    if (sym_cand) {
        // FULL's words read once; the first delta's check, when it runs,
        // is not counted
        KTimer kt(ctx, K_DENSE_SETUP, 8.0 * rows * rw);
        const unsigned grid_tr = (unsigned)((rw / kTW) * (rw / kTW));
        hipLaunchKernelGGL(k_bt_asym, dim3(grid_tr), dim3(TB), 0, ctx->stream,
                           (const u64 *)hs->slots, (const u32 *)nullptr, rw, &st->asym,
                           (const DenseState *)nullptr);
        GDL_TRY(check_launch(ctx, "k_bt_asym"));
        // skipped on the device when the first delta is all of FULL
        hipLaunchKernelGGL(k_bt_asym, dim3(grid_tr), dim3(TB), 0, ctx->stream,
                           (const u64 *)D[0], (const u32 *)pop[0], rw, &st->asym,
                           n_delta == hs->size ? (const DenseState *)st
                                               : (const DenseState *)nullptr);
        GDL_TRY(check_launch(ctx, "k_bt_asym"));
    }

    GDL_TRY(stg.land_at(grouped.e));
    a.n_act = n_act;
    // ids beyond the set's domain in a static relation: the tuple-at-a-time
    // bodies handle (and report) them
    if (h_bad || a.n_act == 0) return GDL_OK;
    h_act.resize(a.n_act);

    if (shape == GDL_DENSE_SG) {
        // the longest lists (more than kHeavy successors) become bit rows
        std::vector<u32> &hb = H.hb;
        hb.clear();
        for (u64 r = 0; r < rows; ++r)
            if (h_soff[r + 1] - h_soff[r] > kHeavy) hb.push_back((u32)r);
        if (hb.size() > kHeavyMax) {
            std::sort(hb.begin(), hb.end(), [&](u32 p, u32 q) {
                return h_soff[p + 1] - h_soff[p] > h_soff[q + 1] - h_soff[q];
            });
            hb.resize(kHeavyMax);
        }
        std::vector<u32> &h_hidx = H.hidx;
        h_hidx.assign(rows, kNoBit);
        for (u32 h = 0; h < (u32)hb.size(); ++h) h_hidx[hb[h]] = h;
        u32 *hidx, *heavy_b;
        u64 *heavy;
        GDL_TRY(mem.get(rows, &hidx));
        GDL_TRY(mem.get(std::max<size_t>(hb.size(), 1), &heavy_b));
        GDL_TRY(mem.get(std::max<u64>(hb.size() * rw, 1), &heavy));
        GDL_TRY(stg.up(hidx, h_hidx.data(), rows * sizeof(u32)));
        if (!hb.empty()) {
            GDL_TRY(stg.up(heavy_b, hb.data(), hb.size() * sizeof(u32)));
            GDL_HIP(ctx, hipMemsetAsync(heavy, 0, hb.size() * rw * sizeof(u64), ctx->stream));
            hipLaunchKernelGGL(k_heavy_rows, dim3((unsigned)std::min<size_t>(hb.size(), 4096)),
                               dim3(DB), 0, ctx->stream, succ_off, succ_idx, heavy_b,
                               (u64)hb.size(), rw, heavy);
            GDL_TRY(check_launch(ctx, "k_heavy_rows"));
        }
        a.heavy = heavy;
        uint4 *tab;
        GDL_TRY(mem.get(rows, &tab));
        hipLaunchKernelGGL(k_succ_tab, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream,
                           succ_off, succ_idx, hidx, rows, tab);
        GDL_TRY(check_launch(ctx, "k_succ_tab"));
        a.succ_tab = tab;
    }

    // work items: the active rows, most preds first, split into slices of at
    // most kSlice preds (on the host: a few hundred thousand rows at most)
    std::vector<uint4> &h_items = H.items;
    std::vector<u32> &h_slices = H.slices;
    build_items(h_off, h_act, h_items, h_slices);
    uint4 *items;
    GDL_TRY(mem.get(h_items.size(), &items));
    GDL_TRY(stg.up(act, h_act.data(), a.n_act * sizeof(u32)));
    GDL_TRY(stg.up(items, h_items.data(), h_items.size() * sizeof(uint4)));
    a.items = items;
    a.n_items = h_items.size();
    if (!h_slices.empty()) {
        const u64 ns = h_slices.size();
        u32 *n_slices, *done;
        u64 *gu, *gt = nullptr;
        GDL_TRY(mem.get(ns, &n_slices));
        GDL_TRY(mem.get(ns, &done));
        GDL_TRY(mem.get(ns * rw, &gu));
        GDL_TRY(stg.up(n_slices, h_slices.data(), ns * sizeof(u32)));
        GDL_HIP(ctx, hipMemsetAsync(done, 0, ns * sizeof(u32), ctx->stream));
        GDL_HIP(ctx, hipMemsetAsync(gu, 0, ns * rw * sizeof(u64), ctx->stream));
        if (shape == GDL_DENSE_SG) {
            GDL_TRY(mem.get(ns * rw, &gt));
            GDL_HIP(ctx, hipMemsetAsync(gt, 0, ns * rw * sizeof(u64), ctx->stream));
        }
        a.n_slices = n_slices;
        a.slice_done = done;
        a.GU = gu;
        a.GT = gt;
    }
    u32 *upop = nullptr;
    u64 *U = nullptr;
    unsigned grid_exp = 1;
    if (shape == GDL_DENSE_SG) {
        // U is built for every delta row (a row that is no row's pred is
        // built and never read: cheaper than listing the pred rows on the
        // host); the planes of r2's row lengths
        a.n_exp = rows;
        u64 max_deg = 0;
        for (u64 r = 0; r < rows; ++r) max_deg = std::max(max_deg, h_soff[r + 1] - h_soff[r]);
        int n_planes = 0;
        while (n_planes < 64 && (max_deg >> n_planes)) ++n_planes;
        if (n_planes > 24)
            return fail(ctx, GDL_ERR_UNSUPPORTED, "dense SG bodies: a row of 2^24 or more");
        u64 *planes;
        GDL_TRY(mem.get(std::max<u64>((u64)n_planes * rw, 1), &planes));
        if (n_planes) {
            hipLaunchKernelGGL(k_dense_planes, dim3(dgrid(ctx, rw * 64)), dim3(DB), 0,
                               ctx->stream, a.succ_off, rw, n_planes, planes);
            GDL_TRY(check_launch(ctx, "k_dense_planes"));
        }
        a.planes = planes;
        a.n_planes = n_planes;
        GDL_TRY(mem.get(rows * rw, &U));
        GDL_TRY(mem.get(rows, &upop));
        grid_exp = (unsigned)std::max<u64>(1, std::min<u64>(a.n_exp, grid_cap(ctx)));
    }
    if (off2) {
        std::vector<u32> &h_act2 = H.act2;
        h_act2.clear();
        for (u64 r = 0; r < rows; ++r)
            if (h_off2[r + 1] > h_off2[r]) h_act2.push_back((u32)r);
        std::vector<uint4> &h_items2 = H.items2;
        std::vector<u32> &h_slices2 = H.slices2;
        build_items(h_off2, h_act2, h_items2, h_slices2);
        col.on = !h_items2.empty();
        col.o.off = off2;
        col.o.idx = idx2;
        col.rows = rows;
        if (col.on) {
            uint4 *items2;
            GDL_TRY(mem.get(h_items2.size(), &items2));
            GDL_TRY(stg.up(items2, h_items2.data(), h_items2.size() * sizeof(uint4)));
            col.o.items = items2;
            col.o.n_items = h_items2.size();
            if (!h_slices2.empty()) {
                const u64 ns = h_slices2.size();
                u32 *n_slices, *done;
                u64 *g;
                GDL_TRY(mem.get(ns, &n_slices));
                GDL_TRY(mem.get(ns, &done));
                GDL_TRY(mem.get(ns * rw, &g));
                GDL_TRY(stg.up(n_slices, h_slices2.data(), ns * sizeof(u32)));
                GDL_HIP(ctx, hipMemsetAsync(done, 0, ns * sizeof(u32), ctx->stream));
                GDL_HIP(ctx, hipMemsetAsync(g, 0, ns * rw * sizeof(u64), ctx->stream));
                col.o.n_slices = n_slices;
                col.o.slice_done = done;
                col.o.G = g;
            }
            GDL_TRY(mem.get(rows * rw, &col.V));
            GDL_TRY(mem.get(rows, &col.vpop));
            GDL_TRY(mem.get(rows, &col.dtpop));
            for (u32 *p : {col.vpop, col.dtpop}) {
                hipLaunchKernelGGL(k_fill_u32, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream,
                                   p, rows, 0u);
                GDL_TRY(check_launch(ctx, "k_fill_u32"));
            }
            col.grid_or = (unsigned)std::max<u64>(
                1, std::min<u64>((col.o.n_items + DWV - 1) / DWV, grid_cap(ctx)));
            col.grid_tr = (unsigned)((rw / kTW) * (rw / kTW));
        }
    }

    const bool sym_rule = sym_cand && col.on;
    DenseState r{};
    GDL_TRY(stg.down(&r, st, sizeof(r)));
    if (r.bad)
        return fail(ctx, GDL_ERR_INVALID_ARG, "dense bodies: a delta key outside the set's domain");
    a.sym = sym_rule && r.asym == 0 ? 1 : 0;
    a.V = col.V;
    a.vpop = col.vpop;

    u64 end = *log_end;
    const bool log_last = (rule->flags & GDL_DENSE_LOG_LAST) != 0;
    // room for the next bodies at about 4x the delta (a dense delta tends to
    // grow; a body that does not fit grows the log first, at the cost of a host
    // wait and a copy of the log)
    GDL_TRY(grow_log(ctx, d_log, log_cap, end, end + std::max<u64>(4 * n_delta, 4096)));
    const unsigned grid =
        (unsigned)std::max<u64>(1, std::min<u64>((a.n_items + DWV - 1) / DWV, grid_cap(ctx)));
    int done = 0, ramp = 4;
    u64 in_delta = n_delta;  // the delta entering the next body
    u64 last_rows = 0;       // rows with new tuples in the last body
    while (done < max_bodies) {
        a.log = *d_log;
        const int k = std::min({ramp, kChunk, max_bodies - done});
        DenseState s{};
        s.cursor = end;
        s.log_cap = *log_cap;
        s.sparse_below = sparse_below;
        s.max_bodies = (u32)k;
        s.col_above = col.on ? col_above : ~0ull;
        s.col = in_delta >= s.col_above ? 1u : 0u;
        GDL_TRY(stg.up(st, &s, sizeof(s)));
        for (int i = 0; i < k; ++i) {
            const int p = (done + i) & 1;
            const BodyBufs bb{pop[p], D[p], pop[p ^ 1], D[p ^ 1], upop, U};
            GDL_TRY(dispatch_body(ctx, shape, kb, a, bb, col, st, counts, grid, grid_exp,
                                  !log_last));
            if (done + i == 0) {
                // the first delta's pops: a later combine writes the pops of
                // the rows with preds only, so the rest must read as empty
                hipLaunchKernelGGL(k_fill_u32, dim3(dgrid(ctx, rows)), dim3(DB), 0, ctx->stream,
                                   pop[0], rows, 0u);
                GDL_TRY(check_launch(ctx, "k_fill_u32"));
            }
        }
        u64 c[kCounts * kChunk];
        u32 dev_err = 0;
        GDL_TRY(stg.later(&r, st, sizeof(r)));
        GDL_TRY(stg.later(c, counts, sizeof(c)));
        GDL_TRY(stg.later(&dev_err, ctx->d_err, sizeof(u32)));
        GDL_TRY(stg.land());
        GDL_TRY(device_error_status(ctx, dev_err));
        for (u32 b = 0; b < r.body; ++b, ++done) {
            const u64 *cb = c + kCounts * b;
            h_new[done] = cb[0];
            h_raw[done] = cb[1];
            const double row = 8.0 * (double)rw;
            // combine: the pred rows read (D, and SG's U), FULL's row read,
            // the new delta row written, new words set in FULL
            ctx->stats[K_DENSE_COMBINE].alg_bytes += row * (double)cb[2] + 8.0 * (double)cb[0];
            // expansion bit by bit: the delta row read, U's row written, one
            // 16-byte entry per delta bit; by columns: the 64-byte segments
            // the transposes read and wrote, the OR's row reads and writes
            if (cb[5] == 0)
                ctx->stats[K_DENSE_EXPAND].alg_bytes +=
                    row * (double)cb[3] + 16.0 * (double)in_delta;
            ctx->stats[K_DENSE_TRANSPOSE].alg_bytes += 64.0 * (double)cb[5];
            ctx->stats[K_DENSE_ORROWS].alg_bytes += row * (double)cb[6];
            // append: the rows with new tuples read, their keys written
            if (!log_last)
                ctx->stats[K_DENSE_APPEND].alg_bytes += row * (double)cb[4] + 8.0 * (double)cb[0];
            in_delta = cb[0];
            last_rows = cb[4];
            hs->size += cb[0];
        }
        if (!log_last && r.appended < r.body) {
            // the last body's append found the log full: grow it, with room for
            // two more bodies as large (a dense delta tends to stay large for
            // a few bodies), then append
            const u64 nn = c[kCounts * (r.body - 1)];
            GDL_TRY(grow_log(ctx, d_log, log_cap, r.cursor, r.cursor + 3 * nn));
            a.log = *d_log;
            const u32 stop = r.stop;
            r.stop = kDnRun;
            r.log_cap = *log_cap;
            GDL_TRY(stg.up(st, &r, sizeof(r)));
            GDL_TRY(launch_append(ctx, kb, a, pop[done & 1], D[done & 1], st, counts, grid));
            GDL_TRY(stg.later(&r, st, sizeof(r)));
            GDL_TRY(stg.later(&dev_err, ctx->d_err, sizeof(u32)));
            GDL_TRY(stg.land());
            GDL_TRY(device_error_status(ctx, dev_err));
            if (r.appended != r.body)
                return fail(ctx, GDL_ERR_OVERFLOW, "dense bodies: the log did not grow");
            r.stop = stop == kDnGrow ? kDnLimit : stop;
        }
        end = r.cursor;
        if (r.body == 0 || r.stop == kDnFixpoint || r.stop == kDnSparse) break;
        ramp = (int)r.body == k ? std::min(2 * ramp, kChunk) : std::max(4, (int)r.body);
    }
    if (log_last && done > 0 && h_new[done - 1] > 0) {
        // the last body's new tuples (its D rows) go to the log: the next delta
        const u64 nn = h_new[done - 1];
        GDL_TRY(grow_log(ctx, d_log, log_cap, end, end + nn));
        a.log = *d_log;
        DenseState s{};
        s.cursor = end;
        s.append_base = end;
        s.log_cap = *log_cap;
        s.body = 1;
        s.max_bodies = 1;
        s.col_above = ~0ull;
        GDL_TRY(stg.up(st, &s, sizeof(s)));
        GDL_TRY(stg.up(counts, &nn, sizeof(nn)));
        GDL_TRY(launch_append(ctx, kb, a, pop[done & 1], D[done & 1], st, counts, grid));
        u32 dev_err = 0;
        GDL_TRY(stg.later(&r, st, sizeof(r)));
        GDL_TRY(stg.later(&dev_err, ctx->d_err, sizeof(u32)));
        GDL_TRY(stg.land());
        GDL_TRY(device_error_status(ctx, dev_err));
        if (r.appended != 1 || r.cursor != end + nn)
            return fail(ctx, GDL_ERR_DEVICE, "dense bodies: the last body's append failed");
        end = r.cursor;
        const double row = 8.0 * (double)rw;
        ctx->stats[K_DENSE_APPEND].alg_bytes += row * (double)last_rows + 8.0 * (double)nn;
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
    if (!ctx || mode < 0 || mode > 4) return GDL_ERR_INVALID_ARG;
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
