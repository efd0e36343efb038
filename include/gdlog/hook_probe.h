// gdlog/hook_probe.h -- turns the reference's device hooks into column maps.
//
// The reference calls every hook through a device function pointer once per
// produced tuple (relation.cu:348-358, 390-403).  Here each hook is called a
// few times when the RA operator is constructed, on tagged sentinel tuples,
// by a one-thread kernel compiled into the caller's translation unit -- the
// same code object as the hook, so the indirect call stays inside one module.
// A hook whose output columns are copies of input columns (every hook of the
// TC, SG and CSPA programs, test/tc.cu:61-78, test/sg.cu:62-91,
// test/cspa.cu:62-96) becomes a gdl_colmap; a predicate that agrees with a
// comparison of two output columns on every probe becomes a gdl_pred.  Any
// other hook is recorded as unsupported and the operator fails with
// GDL_ERR_UNSUPPORTED when it runs.
//
// This header must be compiled with hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "gdl_hip.h"
#include "gdlog/tuple.h"

namespace gdlog {

struct HookMap {
    bool ok = false;
    gdl_colmap map{};
};

struct PredMap {
    bool ok = true;
    gdl_pred pred{GDL_PRED_NONE, 0, 0};
};

namespace probe_detail {

constexpr int kW = 20;  // scratch tuple width of the reference (relation.cu:269,349)
constexpr int kProbeTuples = 64;

static __global__ void probe_gen_kernel(tuple_generator_hook h, column_type *buf) {
    (*h)(buf, buf + kW, buf + 2 * kW);
}

static __global__ void probe_copy_kernel(tuple_copy_hook h, column_type *buf) {
    (*h)(buf, buf + 2 * kW);
}

static __global__ void probe_pred_kernel(tuple_predicate p, column_type *tuples, int *out, int n) {
    for (int i = 0; i < n; ++i) out[i] = (*p)(tuples + i * kW) ? 1 : 0;
}

// Runs one probe; returns false on a HIP error.
template <typename Launch>
static bool run_probe(column_type *host, size_t words, Launch launch) {
    column_type *dev = nullptr;
    if (hipMalloc(&dev, words * sizeof(column_type)) != hipSuccess) return false;
    bool ok = hipMemcpy(dev, host, words * sizeof(column_type), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        launch(dev);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
             hipMemcpy(host, dev, words * sizeof(column_type), hipMemcpyDeviceToHost) ==
                 hipSuccess;
    }
    (void)hipFree(dev);
    return ok;
}

// Decodes out[0..arity) against the inner/outer tags.
static bool decode(const column_type *out, int arity, column_type tag_i, column_type tag_o,
                   bool join, gdl_colmap *m) {
    m->arity = arity;
    m->src[0] = m->src[1] = 0;
    for (int c = 0; c < arity; ++c) {
        const column_type v = out[c];
        if (v >= tag_i && v < tag_i + 2) m->src[c] = (int)(v - tag_i);  // inner / source col
        else if (join && v >= tag_o && v < tag_o + 2) m->src[c] = 2 + (int)(v - tag_o);
        else return false;
    }
    return true;
}

}  // namespace probe_detail

// Column map of a tuple_generator_hook producing `out_arity` columns from
// (inner, outer) tuples of arity <= 2.
static inline HookMap probe_generator_now(tuple_generator_hook h, int out_arity) {
    using namespace probe_detail;
    HookMap r;
    if (!h || out_arity < 1 || out_arity > 2) return r;
    gdl_colmap maps[2];
    const column_type tags[2][2] = {{0x1000000000ul, 0x2000000000ul},
                                    {0x5000000000ul, 0x6000000000ul}};
    for (int round = 0; round < 2; ++round) {
        column_type buf[3 * kW];
        for (int j = 0; j < kW; ++j) {
            buf[j] = tags[round][0] + j;
            buf[kW + j] = tags[round][1] + j;
            buf[2 * kW + j] = 0x3000000000ul + j;
        }
        if (!run_probe(buf, 3 * kW, [&](column_type *d) {
                hipLaunchKernelGGL(probe_gen_kernel, dim3(1), dim3(1), 0, 0, h, d);
            }))
            return r;
        if (!decode(buf + 2 * kW, out_arity, tags[round][0], tags[round][1], true, &maps[round]))
            return r;
    }
    if (memcmp(&maps[0], &maps[1], sizeof(gdl_colmap)) != 0) return r;
    r.ok = true;
    r.map = maps[0];
    return r;
}

// Column map of a tuple_copy_hook.
static inline HookMap probe_copy_now(tuple_copy_hook h, int out_arity) {
    using namespace probe_detail;
    HookMap r;
    if (!h || out_arity < 1 || out_arity > 2) return r;
    gdl_colmap maps[2];
    const column_type tags[2] = {0x1000000000ul, 0x5000000000ul};
    for (int round = 0; round < 2; ++round) {
        column_type buf[3 * kW];
        for (int j = 0; j < kW; ++j) {
            buf[j] = tags[round] + j;
            buf[kW + j] = 0;
            buf[2 * kW + j] = 0x3000000000ul + j;
        }
        if (!run_probe(buf, 3 * kW, [&](column_type *d) {
                hipLaunchKernelGGL(probe_copy_kernel, dim3(1), dim3(1), 0, 0, h, d);
            }))
            return r;
        if (!decode(buf + 2 * kW, out_arity, tags[round], 0, false, &maps[round])) return r;
    }
    if (memcmp(&maps[0], &maps[1], sizeof(gdl_colmap)) != 0) return r;
    r.ok = true;
    r.map = maps[0];
    return r;
}

// Comparison form of a tuple_predicate over a generated tuple of `arity`
// columns: classified on (a,a), (a,b), (b,a) with a < b, then checked on
// pseudo-random tuples including 0 and 2^32 - 1.
static inline PredMap probe_predicate_now(tuple_predicate p, int arity) {
    using namespace probe_detail;
    PredMap r;
    if (!p) return r;  // no predicate: keep everything
    r.ok = false;
    column_type buf[kProbeTuples * kW + kProbeTuples];
    std::memset(buf, 0, sizeof(buf));
    auto put = [&](int i, column_type c0, column_type c1) {
        buf[i * kW] = c0;
        buf[i * kW + 1] = c1;
    };
    put(0, 5, 5);
    put(1, 5, 9);
    put(2, 9, 5);
    uint64_t s = 0x243F6A8885A308D3ull;
    for (int i = 3; i < kProbeTuples; ++i) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        column_type a = (column_type)(s >> 32);
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        column_type b = (i % 4 == 0) ? a : (column_type)(s >> 32);
        if (i == 3) { a = 0; b = 0xffffffffu; }
        if (i == 4) { a = 0xffffffffu; b = 0; }
        if (i >= 8 && i < 32) {  // small ids, where constants in hooks tend to be
            a = (column_type)(i - 8);
            b = (column_type)((i * 7) % 24);
        }
        put(i, a, arity == 2 ? b : 0);
    }
    int *dout = nullptr;
    column_type *dt = nullptr;
    int res[kProbeTuples];
    bool ok = hipMalloc(&dout, sizeof(res)) == hipSuccess &&
              hipMalloc(&dt, kProbeTuples * kW * sizeof(column_type)) == hipSuccess &&
              hipMemcpy(dt, buf, kProbeTuples * kW * sizeof(column_type), hipMemcpyHostToDevice) ==
                  hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(probe_pred_kernel, dim3(1), dim3(1), 0, 0, p, dt, dout, kProbeTuples);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
             hipMemcpy(res, dout, sizeof(res), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (dout) (void)hipFree(dout);
    if (dt) (void)hipFree(dt);
    if (!ok) return r;
    gdl_pred pr{GDL_PRED_NONE, 0, arity == 2 ? 1 : 0};
    if (arity == 2) {
        const int code = res[0] * 4 + res[1] * 2 + res[2];  // (eq, lt, gt) outcomes
        switch (code) {
            case 4: pr.op = GDL_PRED_EQ; break;
            case 3: pr.op = GDL_PRED_NE; break;
            case 2: pr.op = GDL_PRED_LT; break;
            case 6: pr.op = GDL_PRED_LE; break;
            case 1: pr.op = GDL_PRED_GT; break;
            case 5: pr.op = GDL_PRED_GE; break;
            case 7: pr.op = GDL_PRED_NONE; break;
            default: pr.op = GDL_PRED_FALSE; break;
        }
    } else {
        pr.op = res[0] ? GDL_PRED_NONE : GDL_PRED_FALSE;
    }
    for (int i = 0; i < kProbeTuples; ++i) {
        const column_type a = buf[i * kW], b = buf[i * kW + (arity == 2 ? 1 : 0)];
        bool want;
        switch (pr.op) {
            case GDL_PRED_EQ: want = a == b; break;
            case GDL_PRED_NE: want = a != b; break;
            case GDL_PRED_LT: want = a < b; break;
            case GDL_PRED_LE: want = a <= b; break;
            case GDL_PRED_GT: want = a > b; break;
            case GDL_PRED_GE: want = a >= b; break;
            case GDL_PRED_FALSE: want = false; break;
            default: want = true; break;
        }
        if ((res[i] != 0) != want) return r;
    }
    // A non-null predicate that kept every probe is not a comparison of two
    // columns (t[1] != 42 is true on all of them): it is not dropped.
    if (pr.op == GDL_PRED_NONE) return r;
    r.ok = true;
    r.pred = pr;
    return r;
}

namespace probe_detail {
// A hook's meaning is fixed for the process, so each (hook, arity) is probed
// once: later operators with the same hook (every run of a program) skip the
// probe kernels, their allocations and device-wide synchronisation.
template <typename R, typename F>
static R cached(const void *hook, int arity, F compute) {
    static std::mutex mu;
    static std::map<std::pair<uintptr_t, int>, R> memo;
    const auto key = std::make_pair((uintptr_t)hook, arity);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    R r = compute();
    std::lock_guard<std::mutex> g(mu);
    memo.emplace(key, r);
    return r;
}
}  // namespace probe_detail

static inline HookMap probe_generator(tuple_generator_hook h, int out_arity) {
    return probe_detail::cached<HookMap>((const void *)h, out_arity,
                                         [&] { return probe_generator_now(h, out_arity); });
}

static inline HookMap probe_copy(tuple_copy_hook h, int out_arity) {
    return probe_detail::cached<HookMap>((const void *)h, out_arity,
                                         [&] { return probe_copy_now(h, out_arity); });
}

static inline PredMap probe_predicate(tuple_predicate p, int arity) {
    return probe_detail::cached<PredMap>((const void *)p, arity,
                                         [&] { return probe_predicate_now(p, arity); });
}

}  // namespace gdlog
