// batch.hip -- batched TC fixpoints: many instances in one disjoint union.
//
// Device layout of a batch with slots of `node_bits`-bit node ids:
//   E      the union of the admitted instances' reversed edges (b, a) with
//          ids (slot << node_bits) | v, sorted (slot order is key order, so
//          the union is the concatenation of per-instance sorted arrays), and
//          its prefix index on b;
//   delta  the union's current delta, unsorted;
//   hs[s]  one hash set per instance: its FULL (membership and result);
//   segs   the device table {hs[s].slots, mask} the fused join selects by
//          key >> (32 + node_bits).
// A body is: reserve every instance's set for its worst case, one fused join
// + insert of E x delta into the sets (new keys form the next delta), one
// histogram of the next delta by slot (its per-instance |delta|).  Only the
// histogram and two counters come back to the host.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gdl_batch.h"
#include "gdl_internal.hpp"

namespace gdl {

constexpr int BB = 256;
constexpr u32 kMaxHistBins = 2048;

// counts[k >> shift] += weight (1, or weights[i]) for bins below n_bins.
// Bins in LDS; a wave whose live lanes share one bin adds once.
__global__ __launch_bounds__(BB) void k_seg_hist(const u64 *keys, const u32 *weights, u64 n,
                                                 int shift, u32 n_bins,
                                                 unsigned long long *counts) {
    __shared__ unsigned long long s_bins[kMaxHistBins];
    for (u32 i = threadIdx.x; i < n_bins; i += BB) s_bins[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (u64 base = (u64)blockIdx.x * BB; base < n; base += (u64)gridDim.x * BB) {
        const u64 i = base + threadIdx.x;
        const u64 bin = i < n ? keys[i] >> shift : ~0ull;
        const bool live = bin < n_bins;
        u64 wt = live ? (weights ? (u64)weights[i] : 1ull) : 0ull;
        const u64 first = __shfl(bin, 0, 64);
        const u64 same = __ballot(live && bin == first);
        const u64 any = __ballot(live);
        if (same == any) {  // one bin (or none) in this wave
            for (int d = 32; d >= 1; d >>= 1) wt += __shfl_xor(wt, d, 64);
            if (lane == 0 && any && wt) atomicAdd(&s_bins[first], (unsigned long long)wt);
        } else if (live && wt) {
            atomicAdd(&s_bins[bin], (unsigned long long)wt);
        }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < n_bins; i += BB)
        if (s_bins[i]) atomicAdd(&counts[i], s_bins[i]);
}

__global__ __launch_bounds__(BB) void k_mask_cols(const u64 *in, u64 n, u64 m, u64 *out) {
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < n; i += (u64)gridDim.x * BB) {
        const u64 k = in[i];
        out[i] = (((k >> 32) & m) << 32) | (k & m);
    }
}

// Sum of fmix64 of the members of a table with the instance prefix
// stripped (order independent).
__global__ __launch_bounds__(BB) void k_hs_checksum(const u64 *slots, u64 cap, u64 m,
                                                    unsigned long long *sum) {
    __shared__ u64 s_part[BB / 64];
    u64 acc = 0;
    for (u64 i = (u64)blockIdx.x * BB + threadIdx.x; i < cap; i += (u64)gridDim.x * BB) {
        const u64 k = slots[i];
        if (k != kHsEmpty) acc += mix64((((k >> 32) & m) << 32) | (k & m));
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int w = 0; w < BB / 64; ++w) t += s_part[w];
        if (t) atomicAdd(sum, (unsigned long long)t);
    }
}

// Sum of fmix64 over an array of keys (order independent).
__global__ __launch_bounds__(BB) void k_checksum(const u64 *keys, u64 n, unsigned long long *sum) {
    __shared__ u64 s_part[BB / 64];
    u64 acc = 0;
    const u64 stride = (u64)gridDim.x * BB;
    u64 i = (u64)blockIdx.x * BB + threadIdx.x;
    // four independent loads in flight per lane, then the remainder
    for (; i + 3 * stride < n; i += 4 * stride) {
        const u64 a = keys[i], b = keys[i + stride], c = keys[i + 2 * stride],
                  d = keys[i + 3 * stride];
        acc += (mix64(a) + mix64(b)) + (mix64(c) + mix64(d));
    }
    for (; i < n; i += stride) acc += mix64(keys[i]);
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int w = 0; w < BB / 64; ++w) t += s_part[w];
        if (t) atomicAdd(sum, (unsigned long long)t);
    }
}

static unsigned batch_grid(gdl_ctx *ctx, u64 n) {
    const u64 cap = grid_cap(ctx);
    return (unsigned)std::max<u64>(1, std::min<u64>((n + BB - 1) / BB, cap));
}

gdl_status key_histogram(gdl_ctx *ctx, const u64 *keys, const u32 *weights, u64 n, int shift,
                         u32 n_bins, u64 *counts) {
    if (n == 0) return GDL_OK;
    KTimer kt(ctx, K_SEG_HIST, (weights ? 12.0 : 8.0) * n);
    hipLaunchKernelGGL(k_seg_hist, dim3(batch_grid(ctx, n)), dim3(BB), 0, ctx->stream, keys,
                       weights, n, shift, n_bins, (unsigned long long *)counts);
    return check_launch(ctx, "k_seg_hist");
}

gdl_status mask_columns(gdl_ctx *ctx, const u64 *in, u64 n, u32 col_mask, u64 *out) {
    if (n == 0) return GDL_OK;
    KTimer kt(ctx, K_MASK_COLS, 16.0 * n);
    hipLaunchKernelGGL(k_mask_cols, dim3(batch_grid(ctx, n)), dim3(BB), 0, ctx->stream, in, n,
                       (u64)col_mask, out);
    return check_launch(ctx, "k_mask_cols");
}

// ------------------------------------------------------------ TC batch

static double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static u64 splitmix_next(u64 *s) {
    u64 z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct TcbInst {
    u64 qid = 0;
    u32 slot = 0;
    u32 n_nodes = 0;
    u64 *edges = nullptr;  // sorted unique reversed edges, prefixed
    u64 n_edges = 0;
    u64 max_run = 0;       // largest in-degree (a run of E)
    gdl_hashset hs{};
    u64 delta = 0;         // |delta| entering the next body
    u64 derived = 0;
    u64 est = 0;
    u64 reserved = 0;      // bytes of the budget held
    int bodies = 0;
    std::vector<gdl_trace_rec> trace;
    double admit_ms = 0;
};

struct TcbDone {
    gdl_tcb_result r{};
    std::vector<u64> keys;
    std::vector<gdl_trace_rec> trace;
};

}  // namespace gdl

using namespace gdl;

struct gdl_tcb {
    gdl_ctx *ctx = nullptr;
    gdl_tcb_config cfg{};
    int seg_shift = 48;
    u64 col_mask = 0xffff;
    std::vector<std::unique_ptr<TcbInst>> slots;
    u32 n_active = 0;
    // E and its index
    u64 *E = nullptr;
    u64 n_E = 0, cap_E = 0;
    gdl_index idx{};
    bool e_dirty = false;
    // deltas
    u64 *delta = nullptr, *next = nullptr;
    u64 n_delta = 0, cap_delta = 0, cap_next = 0;
    // set table and per-slot counters
    gdl_hs_seg *d_segs = nullptr, *h_segs = nullptr;  // h_segs: pinned
    u64 *d_hist = nullptr, *h_hist = nullptr;          // h_hist: pinned
    bool segs_dirty = true;
    std::deque<TcbDone> done;  // guarded by mu
    u64 reserved = 0, budget = 0;
    u64 bodies = 0, admitted = 0, finished = 0;  // finished: guarded by mu
    double t0 = 0;
    // the solo lane: a thread with a context (and stream) of its own runs the
    // queued instances one after another, beside the batch's bodies
    struct SoloJob {
        u64 qid;
        std::vector<u64> edges;
        u32 n_nodes;
        double admit_ms;
    };
    std::mutex mu;
    std::condition_variable cv;        // a job queued, or stop
    std::condition_variable cv_done;   // a result ready
    std::deque<SoloJob> solo_q;        // guarded by mu
    u32 solo_running = 0;              // guarded by mu
    gdl_status solo_status = GDL_OK;   // guarded by mu: the first failure
    std::string solo_error;            // guarded by mu
    bool stop = false;                 // guarded by mu
    std::vector<gdl_ctx *> solo_ctx;
    std::vector<std::thread> solo_threads;
};

namespace gdl {

static gdl_status dev_alloc(gdl_ctx *ctx, u64 n, u64 **p) {
    void *q = nullptr;
    std::string err;
    gdl_status st = ctx->alloc->alloc(std::max<u64>(n, 1) * sizeof(u64), &q, &err);
    if (st != GDL_OK) return fail(ctx, st, err);
    *p = (u64 *)q;
    return GDL_OK;
}

static void dev_free(gdl_ctx *ctx, u64 **p) {
    if (*p) ctx->alloc->free(*p, nullptr);
    *p = nullptr;
}

// Grows *buf to hold `need` keys, keeping the first `keep`.
static gdl_status ensure_cap(gdl_ctx *ctx, u64 **buf, u64 *cap, u64 need, u64 keep) {
    if (*buf && *cap >= need) return GDL_OK;
    const u64 nc = std::max<u64>(need + need / 4, 4096);
    u64 *nb = nullptr;
    GDL_TRY(dev_alloc(ctx, nc, &nb));
    if (keep)
        GDL_HIP(ctx, hipMemcpyAsync(nb, *buf, keep * sizeof(u64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
    dev_free(ctx, buf);
    *buf = nb;
    *cap = nc;
    return GDL_OK;
}

static u64 pow2_ceil(u64 x) {
    u64 c = 1024;
    while (c < x) c <<= 1;
    return c;
}

// Device bytes an instance is expected to hold: its hash set at the load
// factor (power-of-two capacity), its share of the two delta buffers, its
// edges and the transient arrays of admission.
// Load factor of the instances' hash sets: the context's, but at least 1/2.
// A batch holds many instances' sets at once, and they are sized for each
// body's worst case, so memory (not probe length) is what limits it.
static double batch_load(const gdl_ctx *ctx) {
    return std::min(0.8, std::max(0.5, (double)ctx->cfg.load_factor));
}

static u64 footprint_bytes(const gdl_tcb *b, u64 est, u64 n_edges) {
    const double lf = batch_load(b->ctx);
    const u64 hs = 8 * (pow2_ceil((u64)std::ceil((double)est / lf)) + 1);
    return hs + 16 * est / 2 + 32 * n_edges + (1ull << 16);
}

static gdl_status rebuild_edges(gdl_tcb *b) {
    gdl_ctx *ctx = b->ctx;
    u64 total = 0;
    for (auto &in : b->slots)
        if (in) total += in->n_edges;
    GDL_TRY(ensure_cap(ctx, &b->E, &b->cap_E, total, 0));
    u64 at = 0;
    for (auto &in : b->slots) {
        if (!in || !in->n_edges) continue;
        GDL_HIP(ctx, hipMemcpyAsync(b->E + at, in->edges, in->n_edges * sizeof(u64),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        at += in->n_edges;
    }
    b->n_E = total;
    if (b->idx.slots) GDL_TRY(gdl_index_free(ctx, &b->idx));
    std::memset(&b->idx, 0, sizeof(b->idx));
    if (total) GDL_TRY(index_build(ctx, b->E, total, 1, &b->idx));
    b->e_dirty = false;
    return GDL_OK;
}

static gdl_status retire(gdl_tcb *b, TcbInst *in) {
    gdl_ctx *ctx = b->ctx;
    TcbDone d;
    d.r.qid = in->qid;
    d.r.n_out = in->hs.size;
    d.r.derived = in->derived;
    d.r.iterations = in->bodies;
    d.r.slot = (int32_t)in->slot;
    d.r.n_trace = in->trace.size();
    d.r.admit_ms = in->admit_ms;
    {
        Scratch scr(ctx);
        u64 *sum;
        GDL_TRY(scr.get(1, &sum));
        GDL_HIP(ctx, hipMemsetAsync(sum, 0, sizeof(u64), ctx->stream));
        hipLaunchKernelGGL(k_hs_checksum, dim3(batch_grid(ctx, in->hs.capacity)), dim3(BB), 0,
                           ctx->stream, (const u64 *)in->hs.slots, in->hs.capacity, b->col_mask,
                           (unsigned long long *)sum);
        GDL_TRY(check_launch(ctx, "k_hs_checksum"));
        GDL_TRY(read_u64(ctx, sum, &d.r.checksum));
    }
    // key ~0 is never an instance tuple: slots < 2^(32 - node_bits) - 1 keep it out
    if (b->cfg.keep_keys && in->hs.size) {
        Scratch scr(ctx);
        u64 *raw, *sorted;
        GDL_TRY(scr.get(in->hs.size, &raw));
        GDL_TRY(scr.get(in->hs.size, &sorted));
        u64 n = 0;
        GDL_TRY(gdl_hashset_extract(ctx, &in->hs, raw, in->hs.size, &n));
        if (n != in->hs.size)
            return fail(ctx, GDL_ERR_DEVICE, "batch: extracted " + std::to_string(n) +
                                                 " tuples, expected " +
                                                 std::to_string(in->hs.size));
        GDL_TRY(mask_columns(ctx, raw, n, (u32)b->col_mask, raw));
        GDL_TRY(sort_keys(ctx, raw, n, (int)b->cfg.node_bits, (int)b->cfg.node_bits, sorted));
        d.keys.resize(n);
        GDL_HIP(ctx, hipMemcpyAsync(d.keys.data(), sorted, n * sizeof(u64),
                                    hipMemcpyDeviceToHost, ctx->stream));
        GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
        d.r.has_keys = 1;
    }
    d.trace = std::move(in->trace);
    GDL_TRY(gdl_hashset_free(ctx, &in->hs));
    dev_free(ctx, &in->edges);
    b->h_segs[in->slot] = gdl_hs_seg{nullptr, 0};
    b->segs_dirty = true;
    b->reserved -= std::min(b->reserved, in->reserved);
    d.r.done_ms = wall_ms() - b->t0;
    {
        std::lock_guard<std::mutex> lock(b->mu);
        b->done.push_back(std::move(d));
        ++b->finished;
    }
    b->cv_done.notify_all();
    b->slots[in->slot].reset();
    --b->n_active;
    return GDL_OK;
}

static gdl_status hip_ok(gdl_ctx *ctx, hipError_t e) {
    return e == hipSuccess ? GDL_OK : hip_fail(ctx, e, "batch solo");
}

// The solo lane: an instance whose closure is estimated at cfg.solo_tuples
// or more fills the device on its own, so sharing bodies with others gains
// nothing.  It runs to its fixpoint at admission, on the engine path of a
// single query: FULL = its distinct rows (a bitmap set when its ids fit 17
// bits), the first body from the first |FULL| raw rows, then the stream-
// ordered linear loop (gdl_fixpoint_linear, with bit-matrix bodies while the
// delta is dense).  Its result and trace are those of a batched instance.
static gdl_status solo_run(gdl_ctx *ctx, u64 qid, const u64 *h_edges, u64 n, u32 n_nodes,
                           int keep_keys, double t_admit, double t0, TcbDone *out) {
    int kb = 1;
    while (kb < 32 && (1ull << kb) < (u64)n_nodes) ++kb;
    gdl_hashset hs{};
    Scratch scr(ctx);
    u64 *d_raw, *d_rev, *inner;
    GDL_TRY(scr.get(n, &d_raw));
    GDL_TRY(scr.get(n, &d_rev));
    GDL_TRY(scr.get(n, &inner));
    GDL_HIP(ctx, hipMemcpyAsync(d_raw, h_edges, n * sizeof(u64), hipMemcpyHostToDevice,
                                ctx->stream));
    if (kb <= 17) GDL_TRY(gdl_hashset_create_bitmap(ctx, (u32)kb, &hs));
    else GDL_TRY(gdl_hashset_create(ctx, n + 1024, &hs));
    u64 *log = nullptr, cap = 0;
    std::vector<gdl_trace_rec> trace;
    gdl_index idx{};
    gdl_status st = GDL_OK;
    do {
        // FULL: the distinct rows, also the log's start
        cap = std::max<u64>(2 * n, 4096);
        st = dev_alloc(ctx, cap, &log);
        if (st != GDL_OK) break;
        u64 n_full = 0;
        st = hashset_insert_new(ctx, &hs, d_raw, n, log, &n_full);
        if (st != GDL_OK) break;
        hs.size = n_full;
        // the inner: reversed edges (b, a), sorted and unique, indexed on b
        st = gdl_project_u64(ctx, d_raw, n, gdl_colmap{2, {GDL_COL_INNER1, GDL_COL_INNER0}},
                             d_rev);
        if (st != GDL_OK) break;
        u64 n_inner = 0;
        st = sort_unique_keys(ctx, d_rev, n, kb, kb, inner, &n_inner);
        if (st != GDL_OK) break;
        st = index_build(ctx, inner, n_inner, 1, &idx);
        if (st != GDL_OK) break;
        const gdl_colmap map{2, {GDL_COL_INNER1, GDL_COL_OUTER1}};
        const gdl_pred none{GDL_PRED_NONE, 0, 0};
        // body 0: the first |FULL| raw rows as the delta
        u64 *d_new = nullptr, new_cap = 0, nn = 0, raw = 0;
        st = gdl_join_insert(ctx, inner, n_inner, &idx, d_raw, n_full, map, none, &hs, &d_new,
                             &new_cap, &nn, &raw);
        if (st == GDL_OK) {
            hs.size = n_full + nn;
            trace.push_back(gdl_trace_rec{0, 0, nn, nn, hs.size});
            st = ensure_cap(ctx, &log, &cap, n_full + nn, n_full);
            if (st == GDL_OK && nn)
                st = hip_ok(ctx, hipMemcpyAsync(log + n_full, d_new, nn * sizeof(u64),
                                                hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (d_new) ctx->alloc->free(d_new, nullptr);
        if (st != GDL_OK) break;
        // the later bodies: the linear loop over the log's last segment
        u64 d0 = n_full, d1 = n_full + nn;
        if (nn) {
            constexpr int kSlice = 4096;
            std::vector<u64> hn(kSlice), hr(kSlice);
            for (;;) {
                int bodies = 0;
                st = gdl_fixpoint_linear(ctx, inner, n_inner, &idx, map, none, &hs, &log, &cap,
                                         &d0, &d1, kSlice, hn.data(), hr.data(), &bodies);
                if (st != GDL_OK) break;
                u64 full = trace.back().full;
                for (int k = 0; k < bodies; ++k) {
                    full += hn[k];
                    trace.push_back(gdl_trace_rec{(int32_t)trace.size(), 0, hn[k], hn[k], full});
                }
                if (bodies == 0 || hn[bodies - 1] == 0) break;
            }
            if (st != GDL_OK) break;
        }
        TcbDone d;
        d.r.qid = qid;
        d.r.n_out = d1;
        d.r.iterations = (int32_t)trace.size();
        d.r.slot = -1;
        d.r.n_trace = trace.size();
        d.r.admit_ms = t_admit;
        for (const auto &t : trace) d.r.derived += t.delta;
        st = gdl_checksum_u64(ctx, log, d1, &d.r.checksum);
        if (st != GDL_OK) break;
        if (keep_keys && d1) {
            u64 *sorted;
            st = scr.get(d1, &sorted);
            if (st != GDL_OK) break;
            u64 got = 0;
            st = gdl_hashset_extract_sorted(ctx, &hs, sorted, d1, &got);
            if (st == GDL_OK && got != d1)
                st = fail(ctx, GDL_ERR_DEVICE, "batch solo: set holds " + std::to_string(got) +
                                                   " tuples, log " + std::to_string(d1));
            if (st != GDL_OK) break;
            d.keys.resize(d1);
            st = hip_ok(ctx, hipMemcpyAsync(d.keys.data(), sorted, d1 * sizeof(u64),
                                            hipMemcpyDeviceToHost, ctx->stream));
            if (st == GDL_OK) st = hip_ok(ctx, hipStreamSynchronize(ctx->stream));
            if (st != GDL_OK) break;
            d.r.has_keys = 1;
        }
        st = gdl_hashset_check(ctx, &hs);
        if (st != GDL_OK) break;
        d.trace = std::move(trace);
        d.r.done_ms = wall_ms() - t0;
        *out = std::move(d);
    } while (false);
    if (idx.slots) gdl_index_free(ctx, &idx);
    if (log) ctx->alloc->free(log, nullptr);
    gdl_hashset_free(ctx, &hs);
    return st;
}

// The solo lane's thread: its own context on the batch's device (the
// batch's tuning), jobs in admission order, results into the batch's done
// queue.  The first failure stops the lane; gdl_tcb_step reports it.
static void solo_main(gdl_tcb *b, gdl_ctx *ctx) {
    (void)hipSetDevice(ctx->device);  // this thread's calls go to the batch's device
    for (;;) {
        gdl_tcb::SoloJob job;
        {
            std::unique_lock<std::mutex> lock(b->mu);
            b->cv.wait(lock, [&] { return b->stop || !b->solo_q.empty(); });
            if (b->stop && b->solo_q.empty()) return;
            job = std::move(b->solo_q.front());
            b->solo_q.pop_front();
        }
        TcbDone d;
        gdl_status st = solo_run(ctx, job.qid, job.edges.data(), job.edges.size(), job.n_nodes,
                                 b->cfg.keep_keys, job.admit_ms, b->t0, &d);
        {
            std::lock_guard<std::mutex> lock(b->mu);
            --b->solo_running;
            if (st == GDL_OK) {
                b->done.push_back(std::move(d));
                ++b->finished;
            } else if (b->solo_status == GDL_OK) {
                b->solo_status = st;
                b->solo_error = ctx->last_error;
            }
        }
        b->cv_done.notify_all();
    }
}

static gdl_status solo_enqueue(gdl_tcb *b, u64 qid, const u64 *h_edges, u64 n, u32 n_nodes) {
    if (b->solo_threads.empty()) {
        // the lane's threads, each with a context of its own on the batch's
        // device (its tuning and dense mode): two big instances at once fill
        // the device through each other's host waits and latency-bound bodies
        gdl_config cfg = b->ctx->cfg;
        cfg.mem_cap_bytes = 0;
        for (u32 t = 0; t < b->cfg.solo_threads; ++t) {
            gdl_ctx *c = nullptr;
            GDL_TRY(gdl_ctx_create(&cfg, &c));
            c->tune = b->ctx->tune;
            c->dense_mode = b->ctx->dense_mode;
            b->solo_ctx.push_back(c);
        }
        for (gdl_ctx *c : b->solo_ctx) b->solo_threads.emplace_back(solo_main, b, c);
    }
    gdl_tcb::SoloJob job{qid, std::vector<u64>(h_edges, h_edges + n), n_nodes,
                         wall_ms() - b->t0};
    {
        std::lock_guard<std::mutex> lock(b->mu);
        b->solo_q.push_back(std::move(job));
        ++b->solo_running;
        ++b->admitted;
    }
    b->cv.notify_one();
    return GDL_OK;
}

static gdl_status tcb_admit(gdl_tcb *b, u64 qid, const u64 *h_edges, u64 n, u32 n_nodes,
                            u64 est) {
    gdl_ctx *ctx = b->ctx;
    if (n_nodes == 0 || (u64)n_nodes > (1ull << b->cfg.node_bits))
        return fail(ctx, GDL_ERR_INVALID_ARG, "batch: n_nodes outside 1..2^node_bits");
    for (u64 i = 0; i < n; ++i)
        if ((h_edges[i] >> 32) >= n_nodes || (h_edges[i] & 0xffffffffull) >= n_nodes)
            return fail(ctx, GDL_ERR_INVALID_ARG, "batch: edge id >= n_nodes");
    // 8 sampled sources: admission and the solo decision need the order of
    // magnitude; each sample is a breadth-first search on the host (32 of them
    // cost the 256-instance workload about as much as its small instances' GPU
    // work)
    if (est == 0) est = gdl_tc_estimate(h_edges, n, n_nodes, 8, qid);
    if (n && est >= b->cfg.solo_tuples) return solo_enqueue(b, qid, h_edges, n, n_nodes);
    u32 slot = 0;
    while (slot < b->cfg.max_slots && b->slots[slot]) ++slot;
    if (slot == b->cfg.max_slots) return GDL_ERR_BUSY;
    const u64 need = footprint_bytes(b, est, n);
    if (b->n_active && b->reserved + need > b->budget) return GDL_ERR_BUSY;

    auto in = std::make_unique<TcbInst>();
    in->qid = qid;
    in->slot = slot;
    in->n_nodes = n_nodes;
    in->est = est;
    in->reserved = need;
    in->admit_ms = wall_ms() - b->t0;
    // prefixed raw rows and the largest in-degree (a bound of E's runs)
    const u64 pre = (u64)slot << b->cfg.node_bits;
    std::vector<u64> raw(n);
    std::vector<u32> indeg(n_nodes, 0);
    for (u64 i = 0; i < n; ++i) {
        const u64 a = h_edges[i] >> 32, c = h_edges[i] & 0xffffffffull;
        raw[i] = ((pre | a) << 32) | (pre | c);
        in->max_run = std::max<u64>(in->max_run, ++indeg[c]);
    }
    if (n) {
        Scratch scr(ctx);
        u64 *d_raw, *d_tmp, *d_rev;
        GDL_TRY(scr.get(n, &d_raw));
        GDL_TRY(scr.get(n, &d_tmp));
        GDL_TRY(scr.get(n, &d_rev));
        GDL_HIP(ctx, hipMemcpyAsync(d_raw, raw.data(), n * sizeof(u64), hipMemcpyHostToDevice,
                                    ctx->stream));
        // FULL = the distinct rows; the first |FULL| raw rows are the delta
        in->hs = gdl_hashset{};
        GDL_TRY(hashset_reserve2(ctx, &in->hs, est, est, batch_load(ctx)));
        u64 n_full = 0;
        GDL_TRY(hashset_insert_new(ctx, &in->hs, d_raw, n, d_tmp, &n_full));
        GDL_TRY(ensure_cap(ctx, &b->delta, &b->cap_delta, b->n_delta + n_full, b->n_delta));
        GDL_HIP(ctx, hipMemcpyAsync(b->delta + b->n_delta, d_raw, n_full * sizeof(u64),
                                    hipMemcpyDeviceToDevice, ctx->stream));
        b->n_delta += n_full;
        in->delta = n_full;
        // reversed edges (b, a), sorted and unique: this instance's part of E
        GDL_TRY(gdl_project_u64(ctx, d_raw, n, gdl_colmap{2, {GDL_COL_INNER1, GDL_COL_INNER0}},
                                d_rev));
        GDL_TRY(dev_alloc(ctx, n, &in->edges));
        GDL_TRY(sort_unique_keys(ctx, d_rev, n, -1, -1, in->edges, &in->n_edges));
        GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        GDL_TRY(gdl_hashset_create(ctx, 0, &in->hs));
    }
    b->h_segs[slot] = gdl_hs_seg{in->hs.slots, in->hs.capacity - 1};
    b->segs_dirty = true;
    b->e_dirty = true;
    b->reserved += need;
    b->slots[slot] = std::move(in);
    ++b->n_active;
    ++b->admitted;
    // an instance with no edges finishes in its first body (delta empty)
    return GDL_OK;
}

static gdl_status tcb_step(gdl_tcb *b, u32 *n_active_out) {
    gdl_ctx *ctx = b->ctx;
    {
        std::unique_lock<std::mutex> lock(b->mu);
        if (b->solo_status != GDL_OK)
            return fail(ctx, b->solo_status, "batch solo lane: " + b->solo_error);
        // nothing to run here: wait (at most 1 ms) for a solo result instead
        // of returning to a caller that polls in a loop
        if (b->n_active == 0 && b->solo_running > 0 && b->done.empty())
            b->cv_done.wait_for(lock, std::chrono::milliseconds(1));
    }
    if (b->n_active == 0) {
        *n_active_out = 0;
        return GDL_OK;
    }
    if (b->e_dirty) GDL_TRY(rebuild_edges(b));
    // every instance's worst case: |delta| x longest run, at most the pairs
    // not yet derived; when that is loose, the exact produced count per
    // instance (a count pass and a weighted histogram by slot)
    u64 loose = 0;
    for (auto &p : b->slots) {
        if (!p) continue;
        const u64 mr = std::max<u64>(p->max_run, 1);
        loose += p->delta > (~0ull >> 8) / mr ? (~0ull >> 8) : p->delta * mr;
    }
    const bool exact = b->n_delta && b->n_E && loose > std::max<u64>(1ull << 24, 4 * b->n_delta);
    if (exact) {
        Scratch scr(ctx);
        u32 *counts;
        u64 *begins;
        GDL_TRY(scr.get(b->n_delta, &counts));
        GDL_TRY(scr.get(b->n_delta, &begins));
        GDL_TRY(join_count(ctx, b->E, b->n_E, &b->idx, b->delta, b->n_delta, counts, begins));
        GDL_HIP(ctx, hipMemsetAsync(b->d_hist, 0, b->cfg.max_slots * sizeof(u64), ctx->stream));
        GDL_TRY(key_histogram(ctx, b->delta, counts, b->n_delta, b->seg_shift, b->cfg.max_slots,
                              b->d_hist));
        GDL_HIP(ctx, hipMemcpyAsync(b->h_hist, b->d_hist, b->cfg.max_slots * sizeof(u64),
                                    hipMemcpyDeviceToHost, ctx->stream));
        GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    u64 total_bound = 0;
    for (auto &p : b->slots) {
        if (!p) continue;
        TcbInst *in = p.get();
        const u64 nn = (u64)in->n_nodes * in->n_nodes;
        const u64 room = nn > in->hs.size ? nn - in->hs.size : 0;
        const u64 mr = std::max<u64>(in->max_run, 1);
        u64 bound = in->delta > room / mr ? room : in->delta * mr;
        if (exact) bound = std::min(bound, b->h_hist[in->slot]);
        const u64 cap0 = in->hs.capacity;
        const void *s0 = in->hs.slots;
        GDL_TRY(hashset_reserve2(ctx, &in->hs, std::max<u64>(in->delta, 1024), bound,
                                 batch_load(ctx)));
        if (in->hs.capacity != cap0 || in->hs.slots != s0) {
            b->h_segs[in->slot] = gdl_hs_seg{in->hs.slots, in->hs.capacity - 1};
            b->segs_dirty = true;
        }
        total_bound += bound;
    }
    GDL_TRY(ensure_cap(ctx, &b->next, &b->cap_next, total_bound, 0));
    if (b->segs_dirty) {
        GDL_HIP(ctx, hipMemcpyAsync(b->d_segs, b->h_segs, b->cfg.max_slots * sizeof(gdl_hs_seg),
                                    hipMemcpyHostToDevice, ctx->stream));
        b->segs_dirty = false;
    }
    u64 n_next = 0, n_raw = 0;
    if (b->n_delta && b->n_E)
        GDL_TRY(gdl_join_insert_seg(ctx, b->E, b->n_E, &b->idx, b->delta, b->n_delta,
                                    gdl_colmap{2, {GDL_COL_INNER1, GDL_COL_OUTER1}},
                                    gdl_pred{GDL_PRED_NONE, 0, 0}, b->d_segs, b->cfg.max_slots,
                                    b->seg_shift, b->next, b->cap_next, &n_next, &n_raw));
    GDL_HIP(ctx, hipMemsetAsync(b->d_hist, 0, b->cfg.max_slots * sizeof(u64), ctx->stream));
    GDL_TRY(key_histogram(ctx, b->next, nullptr, n_next, b->seg_shift, b->cfg.max_slots,
                          b->d_hist));
    GDL_HIP(ctx, hipMemcpyAsync(b->h_hist, b->d_hist, b->cfg.max_slots * sizeof(u64),
                                hipMemcpyDeviceToHost, ctx->stream));
    // the device error word in the same wait (both host buffers are pinned)
    GDL_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->d_err, sizeof(u32), hipMemcpyDeviceToHost,
                                ctx->stream));
    GDL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    u32 dev_err = 0;
    memcpy(&dev_err, ctx->h_pinned, sizeof(u32));
    GDL_TRY(device_error_status(ctx, dev_err));
    u64 check = 0;
    for (auto &p : b->slots) {
        if (!p) continue;
        TcbInst *in = p.get();
        const u64 nd = b->h_hist[in->slot];
        check += nd;
        in->hs.size += nd;
        in->derived += nd;
        in->trace.push_back(gdl_trace_rec{in->bodies, 0, nd, nd, in->hs.size});
        ++in->bodies;
        in->delta = nd;
    }
    if (check != n_next)
        return fail(ctx, GDL_ERR_DEVICE, "batch: per-instance new tuples " +
                                             std::to_string(check) + " != " +
                                             std::to_string(n_next));
    std::swap(b->delta, b->next);
    std::swap(b->cap_delta, b->cap_next);
    b->n_delta = n_next;
    ++b->bodies;
    for (auto &p : b->slots)
        if (p && p->delta == 0) GDL_TRY(retire(b, p.get()));
    *n_active_out = b->n_active;
    return GDL_OK;
}

}  // namespace gdl

extern "C" {

gdl_status gdl_key_histogram_u64(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, int shift,
                                 uint32_t n_bins, uint64_t *d_counts) {
    if (!ctx || (n && !d_keys) || !d_counts || shift < 0 || shift > 63 || n_bins == 0 ||
        n_bins > kMaxHistBins)
        return GDL_ERR_INVALID_ARG;
    return key_histogram(ctx, d_keys, nullptr, n, shift, n_bins, d_counts);
}

gdl_status gdl_checksum_u64(gdl_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint64_t *out) {
    if (!ctx || !out || (n && !d_keys)) return GDL_ERR_INVALID_ARG;
    *out = 0;
    if (n == 0) return GDL_OK;
    Scratch scr(ctx);
    u64 *sum;
    GDL_TRY(scr.get(1, &sum));
    GDL_HIP(ctx, hipMemsetAsync(sum, 0, sizeof(u64), ctx->stream));
    hipLaunchKernelGGL(k_checksum, dim3(batch_grid(ctx, n)), dim3(BB), 0, ctx->stream, d_keys, n,
                       (unsigned long long *)sum);
    GDL_TRY(check_launch(ctx, "k_checksum"));
    return read_u64(ctx, sum, out);
}

gdl_status gdl_mask_columns_u64(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n, uint32_t col_mask,
                                uint64_t *d_out) {
    if (!ctx || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    return mask_columns(ctx, d_in, n, col_mask, d_out);
}

void gdl_tcb_config_default(gdl_tcb_config *cfg) {
    if (!cfg) return;
    cfg->max_slots = 256;
    cfg->node_bits = 16;
    cfg->mem_budget = 0;
    cfg->keep_keys = 1;
    cfg->solo_tuples = 1ull << 23;
    cfg->solo_threads = 2;
}

uint64_t gdl_tc_estimate(const uint64_t *h_edges, uint64_t n, uint32_t n_nodes, uint32_t samples,
                         uint64_t seed) {
    if (!h_edges || n == 0 || n_nodes == 0) return 0;
    // forward adjacency (CSR)
    std::vector<u64> off((size_t)n_nodes + 1, 0);
    for (u64 i = 0; i < n; ++i) {
        const u64 a = h_edges[i] >> 32;
        if (a < n_nodes && (h_edges[i] & 0xffffffffull) < n_nodes) ++off[a + 1];
    }
    for (u32 v = 0; v < n_nodes; ++v) off[v + 1] += off[v];
    std::vector<u32> adj(off[n_nodes]);
    std::vector<u64> at(off.begin(), off.end() - 1);
    for (u64 i = 0; i < n; ++i) {
        const u64 a = h_edges[i] >> 32, c = h_edges[i] & 0xffffffffull;
        if (a < n_nodes && c < n_nodes) adj[at[a]++] = (u32)c;
    }
    const bool exact = samples >= n_nodes;
    const u32 k = exact ? n_nodes : std::max<u32>(samples, 1);
    std::vector<u32> mark(n_nodes, 0), queue;
    queue.reserve(n_nodes);
    u64 s = seed, reach = 0;
    for (u32 t = 0; t < k; ++t) {
        const u32 src = exact ? t : (u32)(((unsigned __int128)splitmix_next(&s) * n_nodes) >> 64);
        const u32 stamp = t + 1;
        queue.clear();
        for (u64 e = off[src]; e < off[src + 1]; ++e)
            if (mark[adj[e]] != stamp) {
                mark[adj[e]] = stamp;
                queue.push_back(adj[e]);
            }
        for (size_t h = 0; h < queue.size(); ++h) {
            const u32 v = queue[h];
            for (u64 e = off[v]; e < off[v + 1]; ++e)
                if (mark[adj[e]] != stamp) {
                    mark[adj[e]] = stamp;
                    queue.push_back(adj[e]);
                }
        }
        reach += queue.size();
    }
    if (exact) return reach;
    return (u64)std::ceil((double)reach / k * n_nodes);
}

gdl_status gdl_tcb_create(gdl_ctx *ctx, const gdl_tcb_config *cfg_in, gdl_tcb **out) {
    if (!ctx || !out) return GDL_ERR_INVALID_ARG;
    *out = nullptr;
    gdl_tcb_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else gdl_tcb_config_default(&cfg);
    if (cfg.max_slots == 0 || cfg.max_slots > kMaxSegs || cfg.node_bits < 1 ||
        cfg.node_bits > 30 || cfg.solo_threads < 1 || cfg.solo_threads > 8)
        return fail(ctx, GDL_ERR_INVALID_ARG,
                    "batch: max_slots 1..256, node_bits 1..30, solo_threads 1..8");
    // slot ids must fit the bits above node_bits, and the all-ones slot is
    // kept free so no tuple key is ~0 (the hash sets' empty marker)
    const u64 slot_room = 1ull << (32 - cfg.node_bits);
    if ((u64)cfg.max_slots >= slot_room)
        return fail(ctx, GDL_ERR_INVALID_ARG, "batch: max_slots needs more than 32 - node_bits bits");
    auto b = std::make_unique<gdl_tcb>();
    b->ctx = ctx;
    b->cfg = cfg;
    b->seg_shift = 32 + (int)cfg.node_bits;
    b->col_mask = (1ull << cfg.node_bits) - 1;
    b->slots.resize(cfg.max_slots);
    if (cfg.mem_budget) {
        b->budget = cfg.mem_budget;
    } else {
        size_t fr = 0, tot = 0;
        GDL_HIP(ctx, hipMemGetInfo(&fr, &tot));
        b->budget = (u64)(0.8 * (double)fr);
    }
    GDL_HIP(ctx, hipHostMalloc((void **)&b->h_segs, cfg.max_slots * sizeof(gdl_hs_seg),
                               hipHostMallocDefault));
    GDL_HIP(ctx, hipHostMalloc((void **)&b->h_hist, cfg.max_slots * sizeof(u64),
                               hipHostMallocDefault));
    for (u32 i = 0; i < cfg.max_slots; ++i) b->h_segs[i] = gdl_hs_seg{nullptr, 0};
    u64 *p = nullptr;
    GDL_TRY(dev_alloc(ctx, 2 * cfg.max_slots, &p));
    b->d_segs = (gdl_hs_seg *)p;
    GDL_TRY(dev_alloc(ctx, cfg.max_slots, &b->d_hist));
    b->t0 = wall_ms();
    *out = b.release();
    return GDL_OK;
}

gdl_status gdl_tcb_destroy(gdl_tcb *b) {
    if (!b) return GDL_ERR_INVALID_ARG;
    gdl_ctx *ctx = b->ctx;
    if (!b->solo_threads.empty()) {
        {
            std::lock_guard<std::mutex> lock(b->mu);
            b->stop = true;
        }
        b->cv.notify_all();
        for (auto &t : b->solo_threads) t.join();  // the queued solo instances finish first
    }
    for (gdl_ctx *c : b->solo_ctx) gdl_ctx_destroy(c);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : b->slots) {
        if (!p) continue;
        gdl_hashset_free(ctx, &p->hs);
        dev_free(ctx, &p->edges);
    }
    if (b->idx.slots) gdl_index_free(ctx, &b->idx);
    dev_free(ctx, &b->E);
    dev_free(ctx, &b->delta);
    dev_free(ctx, &b->next);
    u64 *segs = (u64 *)b->d_segs;
    dev_free(ctx, &segs);
    dev_free(ctx, &b->d_hist);
    if (b->h_segs) (void)hipHostFree(b->h_segs);
    if (b->h_hist) (void)hipHostFree(b->h_hist);
    delete b;
    return GDL_OK;
}

gdl_status gdl_tcb_admit(gdl_tcb *b, uint64_t qid, const uint64_t *h_edges, uint64_t n,
                         uint32_t n_nodes, uint64_t est_tuples) {
    if (!b || (n && !h_edges)) return GDL_ERR_INVALID_ARG;
    return tcb_admit(b, qid, h_edges, n, n_nodes, est_tuples);
}

gdl_status gdl_tcb_step(gdl_tcb *b, uint32_t *n_active) {
    if (!b || !n_active) return GDL_ERR_INVALID_ARG;
    return tcb_step(b, n_active);
}

int gdl_tcb_poll(gdl_tcb *b, gdl_tcb_result *out) {
    if (!b || !out) return 0;
    std::lock_guard<std::mutex> lock(b->mu);
    if (b->done.empty()) return 0;
    *out = b->done.front().r;
    return 1;
}

gdl_status gdl_tcb_fetch(gdl_tcb *b, uint64_t qid, uint64_t *h_keys, gdl_trace_rec *h_trace) {
    if (!b) return GDL_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(b->mu);
    for (auto it = b->done.begin(); it != b->done.end(); ++it) {
        if (it->r.qid != qid) continue;
        if (h_keys && !it->keys.empty())
            std::memcpy(h_keys, it->keys.data(), it->keys.size() * sizeof(u64));
        if (h_trace && !it->trace.empty())
            std::memcpy(h_trace, it->trace.data(), it->trace.size() * sizeof(gdl_trace_rec));
        b->done.erase(it);
        return GDL_OK;
    }
    return fail(b->ctx, GDL_ERR_INVALID_ARG, "batch: no finished instance " + std::to_string(qid));
}

gdl_status gdl_tcb_info(gdl_tcb *b, uint64_t *bodies, uint64_t *admitted, uint64_t *finished,
                        uint64_t *reserved_bytes, uint64_t *budget_bytes) {
    if (!b) return GDL_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(b->mu);
    if (bodies) *bodies = b->bodies;
    if (admitted) *admitted = b->admitted;
    if (finished) *finished = b->finished;
    if (reserved_bytes) *reserved_bytes = b->reserved;
    if (budget_bytes) *budget_bytes = b->budget;
    return GDL_OK;
}

}  // extern "C"
