// Allocation-free sort + unique for callers that cannot allocate device
// memory (code under HIP-graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its scratch -- the radix sort's ping-pong buffer and
// digit histograms, the sorted copy ahead of the unique pass and the
// compaction's tile counts -- lives in one file-scope device array of fixed
// size instead of the context allocator.
//
// Each GPU that loads this library holds the array until the process ends.
// Its device address is looked up once per device, with that device current,
// when the first context on the device is created (gdl_ctx_create calls
// noalloc_setup); kernels then receive that address, never the array's name.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind: calls that overlap on one GPU (two host threads, or two streams of
// different contexts) race and must be serialised by the caller.
#include <mutex>

#include "gdl_internal.hpp"

namespace gdl {

static constexpr uint64_t kNoallocMaxKeys = GDL_NOALLOC_MAX_KEYS;
// Worst case for kNoallocMaxKeys keys: the sorted copy (8n) and the radix
// ping-pong buffer (8n), 2^9-bin digit counts per 4096-key tile, and the
// compaction's tile counts, each piece 256-byte aligned.
static constexpr size_t kScratchBytes = 18u << 20;
static_assert(kScratchBytes >= 16 * GDL_NOALLOC_MAX_KEYS + 4 * 512 * (GDL_NOALLOC_MAX_KEYS / 4096) +
                                   (64u << 10),
              "scratch array too small for the key limit");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

gdl_status noalloc_setup(gdl_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, GDL_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return GDL_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                        "hipGetSymbolAddress(scratch)");
    }
    g_scratch_ptr[ctx->device] = p;
    return GDL_OK;
}

static void *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace gdl

using namespace gdl;

extern "C" {

gdl_status gdl_sort_unique_u64_noalloc(gdl_ctx *ctx, const uint64_t *d_in, uint64_t n,
                                       int hi_bits, int lo_bits, uint64_t *d_out,
                                       uint64_t *n_out) {
    if (!ctx || !n_out || (n && (!d_in || !d_out))) return GDL_ERR_INVALID_ARG;
    if (n && d_in == d_out)
        return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_sort_unique_u64_noalloc: d_in == d_out");
    if (n > kNoallocMaxKeys)
        return fail(ctx, GDL_ERR_INVALID_ARG,
                    "gdl_sort_unique_u64_noalloc: " + std::to_string(n) +
                        " keys exceed the fixed scratch limit of " +
                        std::to_string(kNoallocMaxKeys));
    if (ctx->arena) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_sort_unique_u64_noalloc: re-entered");
    void *base = scratch_of(ctx->device);
    if (!base) return fail(ctx, GDL_ERR_HIP, "gdl_sort_unique_u64_noalloc: scratch not set up");
    // Statistics keep their launch and byte counts; per-launch event timing
    // is off for the call (an empty event pool would otherwise create events).
    const bool stats_on = ctx->stats_on;
    ctx->stats_on = false;
    ctx->arena = static_cast<char *>(base);
    ctx->arena_size = kScratchBytes;
    ctx->arena_used = 0;
    gdl_status st = sort_unique_keys(ctx, d_in, n, hi_bits, lo_bits, d_out, n_out);
    ctx->arena = nullptr;
    ctx->arena_size = ctx->arena_used = 0;
    ctx->stats_on = stats_on;
    return st;
}

}  // extern "C"
