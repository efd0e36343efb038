// ra_ops.cpp -- RelationalJoin / RelationalCopy / RelationalACopy
// (reference: src/join.cu, src/copy.cu, src/acopy.cu), on the C ABI.
#include <chrono>
#include <iostream>

#include "gdlog/exception.h"
#include "gdlog/relational_algebra.h"

using namespace gdlog;

namespace {

// Wall-clock phase timer into detail_time[slot] (seconds); active only with
// gdlog::set_phase_timing(true), since it synchronises the stream.
class Phase {
  public:
    Phase(float *detail, int slot) : detail_(phase_timing() ? detail : nullptr), slot_(slot) {
        if (detail_) {
            check(gdl_sync(context()), "gdl_sync");
            t0_ = std::chrono::steady_clock::now();
        }
    }
    ~Phase() {
        if (!detail_) return;
        (void)gdl_sync(context());
        detail_[slot_] +=
            std::chrono::duration<float>(std::chrono::steady_clock::now() - t0_).count();
    }

  private:
    float *detail_;
    int slot_;
    std::chrono::steady_clock::time_point t0_;
};

GHashRelContainer *version_of(Relation *r, RelationVersion v, bool inner) {
    if (v == DELTA) return r->delta;
    if (v == FULL || inner) return r->full;  // NEWT as inner reads FULL (join.cu:15-20)
    return r->newt;
}

// newt := newt U sort+unique(fresh); takes ownership of `fresh`.
void accumulate_sorted(GHashRelContainer *newt, uint64_t *fresh, tuple_size_t n, float *detail) {
    uint64_t *su = nullptr;
    tuple_size_t m;
    {
        Phase ph(detail, 3);
        m = sort_unique(fresh, n, &su);
    }
    device_free(fresh);
    if (!newt->keys || newt->tuple_counts == 0) {
        container_take(newt, su, m, n, true);
        return;
    }
    container_normalize(newt);
    Phase ph(detail, 6);
    const tuple_size_t na = newt->tuple_counts;
    uint64_t *merged = device_alloc_keys(na + m);
    uint64_t nm = 0;
    gdl_status st = gdl_merge_union_u64(context(), newt->keys, na, su, m, merged, &nm);
    device_free(su);
    if (st != GDL_OK) device_free(merged);
    check(st, "newt merge");
    container_take(newt, merged, nm, na + m, true);
}

// tmp newt: raw concatenation, no dedup (join.cu:141-151,226-254).
// Also the newt of an append-mode relation, deduplicated only when it meets
// FULL's hash set at the end of the iteration.
void append_raw(GHashRelContainer *newt, uint64_t *fresh, tuple_size_t n) {
    if (!newt->keys || newt->tuple_counts == 0) {
        container_take(newt, fresh, n, n, false);
        return;
    }
    container_append(newt, fresh, n);
    device_free(fresh);
}

gdlog::HookLayout hook_layout(const Relation *r) {
    gdlog::HookLayout h{};
    h.arity = r->layout.arity;
    for (int c = 0; c < 4; ++c) {
        h.bits[c] = r->layout.bits[c];
        h.shift[c] = r->layout.shift[c];
    }
    return h;
}

int bits_of(uint64_t v) { return v ? 64 - __builtin_clzll((unsigned long long)v) : 0; }

// After a hook kernel: the overflow flag, and the produced keys' bit bound
// for the radix sorts that follow.
void finish_hook_stats(gdl_ctx *ctx, const uint64_t *d_stats, const std::string &what) {
    uint64_t st[3] = {0, 0, 0};
    check(gdl_memcpy_d2h(ctx, st, d_stats, sizeof(st)), "hook stats");
    if (st[2])
        throw Error(GDL_ERR_OVERFLOW, what + ": a hook produced a value that does not fit its "
                                             "column of the output relation's key layout");
    note_id_bits(bits_of(st[0] | st[1]));
}

// Stable compaction of the kept outputs; frees out and keep.
tuple_size_t compact_kept(gdl_ctx *ctx, uint64_t *out, uint64_t *keep, tuple_size_t n,
                          uint64_t **kept) {
    uint64_t *k = device_alloc_keys(n);
    uint64_t m = 0;
    gdl_status st = gdl_compact_flags_u64(ctx, out, (const uint8_t *)keep, n, k, &m);
    device_free(out);
    device_free(keep);
    if (st != GDL_OK) device_free(k);
    check(st, "gdl_compact_flags_u64");
    if (m == 0) {
        device_free(k);
        k = nullptr;
    }
    *kept = k;
    return m;
}

uint64_t *zeroed_stats(gdl_ctx *ctx) {
    uint64_t *d = device_alloc_keys(3);
    const uint64_t z[3] = {0, 0, 0};
    check(gdl_memcpy_h2d(ctx, d, z, sizeof(z)), "hook stats");
    return d;
}

}  // namespace

// The join through its device hooks (gdlog/hook_exec.h): the runs are
// counted and scanned on the engine, the hook kernel of the translation unit
// that built the operator generates and filters every produced tuple, and the
// kept ones come back in production order.
static tuple_size_t run_hook_join(RelationalJoin &op, GHashRelContainer *inner,
                              GHashRelContainer *outer, const gdl_index *idx, uint64_t **res) {
    gdl_ctx *ctx = context();
    *res = nullptr;
    if (!op.hook_join)
        throw Error(GDL_ERR_UNSUPPORTED, "RelationalJoin into " + op.output_rel->name +
                                             ": no hook kernel for a non-column-map operator");
    const tuple_size_t no = outer->tuple_counts;
    uint64_t *counts = device_alloc_keys((no + 1) / 2);
    uint64_t *begins = device_alloc_keys(no);
    uint64_t *offsets = device_alloc_keys(no + 1);
    uint64_t total = 0;
    gdl_status st = gdl_join_count(ctx, inner->keys, inner->tuple_counts, idx, outer->keys, no,
                                   (uint32_t *)counts, begins);
    if (st == GDL_OK) st = gdl_join_scan(ctx, (const uint32_t *)counts, no, offsets, &total);
    device_free(counts);
    if (st != GDL_OK || total == 0) {
        device_free(begins);
        device_free(offsets);
        check(st, "hook join count");
        return 0;
    }
    uint64_t *out = device_alloc_keys(total);
    uint64_t *keep = device_alloc_keys((total + 7) / 8);
    uint64_t *stats = zeroed_stats(ctx);
    gdlog::HookJoinArgs a{};
    a.stream = (hipStream_t)gdl_stream(ctx);
    a.gen = op.tuple_generator;
    a.pred = op.tuple_pred;
    a.inner = inner->keys;
    a.inner_layout = hook_layout(op.inner_rel);
    a.outer = outer->keys;
    a.outer_layout = hook_layout(op.outer_rel);
    a.n_outer = no;
    a.offsets = offsets;
    a.begins = begins;
    a.total = total;
    a.out_layout = hook_layout(op.output_rel);
    a.out = out;
    a.keep = (uint8_t *)keep;
    a.stats = (unsigned long long *)stats;
    const hipError_t e = op.hook_join(&a);
    device_free(begins);
    device_free(offsets);
    if (e != hipSuccess) {
        device_free(out);
        device_free(keep);
        device_free(stats);
        throw Error(GDL_ERR_HIP, "RelationalJoin into " + op.output_rel->name +
                                     ": hook kernel launch: " + hipGetErrorString(e));
    }
    try {
        finish_hook_stats(ctx, stats, "RelationalJoin into " + op.output_rel->name);
    } catch (...) {
        device_free(out);
        device_free(keep);
        device_free(stats);
        throw;
    }
    device_free(stats);
    return compact_kept(ctx, out, keep, total, res);
}

// A copy through its device hook: one output per source tuple.
static tuple_size_t run_hook_copy(gdlog::HookCopyLaunch launch, tuple_copy_hook cp,
                              const Relation *src_rel, const GHashRelContainer *src,
                              const Relation *dest_rel, uint64_t **res) {
    gdl_ctx *ctx = context();
    *res = nullptr;
    if (!launch)
        throw Error(GDL_ERR_UNSUPPORTED, "copy into " + dest_rel->name +
                                             ": no hook kernel for a non-column-map operator");
    const tuple_size_t n = src->tuple_counts;
    uint64_t *out = device_alloc_keys(n);
    uint64_t *keep = device_alloc_keys((n + 7) / 8);
    uint64_t *stats = zeroed_stats(ctx);
    gdlog::HookCopyArgs a{};
    a.stream = (hipStream_t)gdl_stream(ctx);
    a.cp = cp;
    a.src = src->keys;
    a.src_layout = hook_layout(src_rel);
    a.n = n;
    a.out_layout = hook_layout(dest_rel);
    a.out = out;
    a.keep = (uint8_t *)keep;
    a.stats = (unsigned long long *)stats;
    const hipError_t e = launch(&a);
    if (e != hipSuccess) {
        device_free(out);
        device_free(keep);
        device_free(stats);
        throw Error(GDL_ERR_HIP, "copy into " + dest_rel->name + ": hook kernel launch: " +
                                     hipGetErrorString(e));
    }
    try {
        finish_hook_stats(ctx, stats, "copy into " + dest_rel->name);
    } catch (...) {
        device_free(out);
        device_free(keep);
        device_free(stats);
        throw;
    }
    device_free(stats);
    return compact_kept(ctx, out, keep, n, res);
}

void RelationalJoin::operator()() {
    const bool output_is_tmp = output_rel->tmp_flag;
    GHashRelContainer *inner = version_of(inner_rel, inner_ver, true);
    GHashRelContainer *outer = version_of(outer_rel, outer_ver, false);
    if (!outer->keys || outer->tuple_counts == 0) {
        outer->tuple_counts = 0;
        return;
    }
    if (!inner->keys || inner->tuple_counts == 0) {
        // join.cu:42-45: an empty inner zeroes the outer's count
        if (empty_inner_zeroes_outer) outer->tuple_counts = 0;
        return;
    }
    const int k = (int)outer->index_column_size;
    if (k < 1 || k > inner->arity || k > outer->arity)
        throw Error(GDL_ERR_INVALID_ARG, "RelationalJoin: bad join column count");
    const gdl_index *idx;
    {
        Phase ph(detail_time, 5);
        idx = container_index(inner, k);
    }
    if (uses_hooks() || inner_rel->arity > 2 || outer_rel->arity > 2 || output_rel->arity > 2) {
        if (k != 1)
            throw Error(GDL_ERR_UNSUPPORTED, "RelationalJoin into " + output_rel->name +
                                                 ": a hook join matches on one column");
        uint64_t *res = nullptr;
        tuple_size_t m;
        {
            Phase ph(detail_time, 2);
            m = run_hook_join(*this, inner, outer, idx, &res);
        }
        if (verbose())
            std::cout << output_rel->name << "   " << k << " join result size(non dedup) " << m
                      << std::endl;
        if (disable_load || m == 0) {
            device_free(res);
            return;
        }
        if (output_is_tmp || output_rel->append_mode) append_raw(output_rel->newt, res, m);
        else accumulate_sorted(output_rel->newt, res, m, detail_time);
        return;
    }
    if (output_rel->tmp_dedup && !disable_load) {
        // a tmp relation's newt is the set of this iteration's outputs: the
        // join inserts into the iteration's set and appends what is new
        gdl_ctx *ctx = context();
        if (!output_rel->tmp_set.slots) {
            const int b = id_bits();
            if (b >= 1 && b <= 17)
                check(gdl_hashset_create_bitmap(ctx, (uint32_t)b, &output_rel->tmp_set),
                      "gdl_hashset_create_bitmap");
            else
                check(gdl_hashset_create(ctx, outer->tuple_counts + 1024, &output_rel->tmp_set),
                      "gdl_hashset_create");
        }
        GHashRelContainer *nw = output_rel->newt;
        uint64_t n_new = nw->keys ? nw->tuple_counts : 0, cap = nw->capacity, m = 0;
        {
            Phase ph(detail_time, 2);
            check(gdl_join_insert(ctx, inner->keys, inner->tuple_counts, idx, outer->keys,
                                  outer->tuple_counts, gen_map.map, pred_map.pred,
                                  &output_rel->tmp_set, &nw->keys, &cap, &n_new, &m),
                  "gdl_join_insert");
        }
        nw->capacity = cap;
        nw->tuple_counts = n_new;
        nw->data_raw_row_size = n_new;
        nw->sorted = n_new <= 1;
        if (verbose())
            std::cout << output_rel->name << "   " << k << " join result size(non dedup) " << m
                      << std::endl;
        return;
    }
    if (output_rel->fused_output && !output_is_tmp && !disable_load) {
        uint64_t m = 0;
        {
            Phase ph(detail_time, 2);
            check(gdl_join_insert(context(), inner->keys, inner->tuple_counts, idx, outer->keys,
                                  outer->tuple_counts, gen_map.map, pred_map.pred,
                                  &output_rel->full_set, &output_rel->pend_keys,
                                  &output_rel->pend_cap, &output_rel->pend_n, &m),
                  "gdl_join_insert");
        }
        output_rel->pend_raw += m;
        if (verbose())
            std::cout << output_rel->name << "   " << k << " join result size(non dedup) " << m
                      << std::endl;
        return;
    }
    uint64_t *res = nullptr;
    uint64_t m = 0;
    {
        Phase ph(detail_time, 2);
        check(gdl_join(context(), inner->keys, inner->tuple_counts, idx, outer->keys,
                       outer->tuple_counts, gen_map.map, pred_map.pred, &res, &m),
              "gdl_join");
    }
    if (verbose())
        std::cout << output_rel->name << "   " << k << " join result size(non dedup) " << m
                  << std::endl;
    if (disable_load || m == 0) {
        device_free(res);
        return;
    }
    if (output_is_tmp || output_rel->append_mode) append_raw(output_rel->newt, res, m);
    else accumulate_sorted(output_rel->newt, res, m, detail_time);
}

bool RelationalJoin::run_linear(int max_bodies, std::vector<uint64_t> *new_tuples,
                                std::vector<uint64_t> *raw_tuples) {
    Relation *out = output_rel;
    if (max_bodies <= 0 || !out || !out->fused_output || out->tmp_flag || disable_load ||
        outer_rel != out || outer_ver != DELTA || inner_rel == out || inner_ver != FULL ||
        !gen_map.ok || !pred_map.ok)
        return false;
    GHashRelContainer *inner = inner_rel->full;
    GHashRelContainer *delta = out->delta;
    GHashRelContainer *full = out->full;
    if (!inner->keys || inner->tuple_counts == 0) return false;  // the quirk path
    const int k = (int)delta->index_column_size;
    if (k < 1 || k > inner->arity || k > delta->arity) return false;
    // the delta is the log's tail (the fused output appends it to FULL); a
    // log that does not end with the delta (a FULL restored some other way)
    // leaves the bodies to the per-iteration path
    const tuple_size_t nd = delta->tuple_counts, nf = out->current_full_size;
    if (nd == 0 || nd > nf || !full->keys || !delta->keys) return false;
    {
        uint64_t sum_delta = 0, sum_tail = 0;
        check(gdl_checksum_u64(context(), delta->keys, nd, &sum_delta), "linear delta checksum");
        check(gdl_checksum_u64(context(), full->keys + (nf - nd), nd, &sum_tail),
              "linear log checksum");
        if (sum_delta != sum_tail) return false;
    }
    const gdl_index *idx = container_index(inner, k);
    if (idx->max_run > 256) return false;  // hub runs: the per-body path balances them
    uint64_t *log = full->keys;
    uint64_t cap = full->capacity, d0 = nf - nd, d1 = nf;
    new_tuples->clear();
    raw_tuples->clear();
    container_drop_index(full);
    // in slices of at most kSlice bodies (the counts' buffers), to the
    // fixpoint or max_bodies
    constexpr int kSlice = 4096;
    std::vector<uint64_t> nb(kSlice), rb(kSlice);
    gdl_status st = GDL_OK;
    int left = max_bodies;
    while (left > 0) {
        int bodies = 0;
        st = gdl_fixpoint_linear(context(), inner->keys, inner->tuple_counts, idx, gen_map.map,
                                 pred_map.pred, &out->full_set, &log, &cap, &d0, &d1,
                                 std::min(left, kSlice), nb.data(), rb.data(), &bodies);
        if (st != GDL_OK) break;
        new_tuples->insert(new_tuples->end(), nb.begin(), nb.begin() + bodies);
        raw_tuples->insert(raw_tuples->end(), rb.begin(), rb.begin() + bodies);
        left -= bodies;
        if (bodies == 0 || nb[bodies - 1] == 0) break;  // the fixpoint
    }
    full->keys = log;
    full->capacity = cap;
    full->tuple_counts = d1;
    full->data_raw_row_size = d1;
    full->sorted = false;
    out->current_full_size = d1;
    check(st, "gdl_fixpoint_linear");
    // the delta after the last body: its new tuples (none at the fixpoint)
    container_free(delta);
    if (d1 > d0) {
        uint64_t *dk = device_alloc_keys(d1 - d0);
        check(gdl_memcpy_d2d(context(), dk, log + d0, (d1 - d0) * sizeof(uint64_t)),
              "linear delta");
        container_take(delta, dk, d1 - d0, d1 - d0, false);
    }
    return true;
}

void RelationalCopy::operator()() {
    GHashRelContainer *src = src_ver == DELTA ? src_rel->delta : src_rel->full;
    if (verbose()) std::cout << "Copy " << src_rel->name << " to " << dest_rel->name << std::endl;
    if (src->tuple_counts == 0) {
        if (empty_source_zeroes_dest) dest_rel->newt->tuple_counts = 0;  // copy.cu:23-26
        return;
    }
    tuple_size_t n = src->tuple_counts;
    uint64_t *copied = nullptr;
    if (!copy_map.ok || src_rel->arity > 2 || dest_rel->arity > 2) {
        n = run_hook_copy(hook_copy, tuple_generator, src_rel, src, dest_rel, &copied);
        if (n == 0) return;
    } else {
        copied = device_alloc_keys(n);
        check(gdl_project_u64(context(), src->keys, n, copy_map.map, copied), "gdl_project_u64");
    }
    if (filter.op != GDL_PRED_NONE) {
        uint64_t *kept = device_alloc_keys(n);
        uint64_t m = 0;
        gdl_status st = gdl_filter_u64(context(), copied, n, filter, kept, &m);
        device_free(copied);
        if (st != GDL_OK) device_free(kept);
        check(st, "gdl_filter_u64");
        copied = kept;
        n = m;
        if (n == 0) {
            device_free(copied);
            return;
        }
    }
    if (dest_rel->append_mode) append_raw(dest_rel->newt, copied, n);
    else accumulate_sorted(dest_rel->newt, copied, n, nullptr);
}

void RelationalACopy::operator()() {
    GHashRelContainer *src = src_rel->newt;
    GHashRelContainer *dest = dest_rel->newt;
    if (verbose()) std::cout << "ACopy " << src_rel->name << " to " << dest_rel->name << std::endl;
    if (src->tuple_counts == 0) {
        container_free(dest);  // acopy.cu:19-23
        return;
    }
    tuple_size_t n = src->tuple_counts;
    uint64_t *copied = nullptr;
    if (!copy_map.ok || src_rel->arity > 2 || dest_rel->arity > 2) {
        n = run_hook_copy(hook_copy, tuple_generator, src_rel, src, dest_rel, &copied);
        if (n == 0) {
            container_free(dest);
            return;
        }
    } else {
        copied = device_alloc_keys(n);
        check(gdl_project_u64(context(), src->keys, n, copy_map.map, copied), "gdl_project_u64");
    }
    if (dest_rel->append_mode) {
        container_take(dest, copied, n, n, false);
        return;
    }
    uint64_t *su = nullptr;
    tuple_size_t m = sort_unique(copied, n, &su);
    device_free(copied);
    container_take(dest, su, m, n, true);
}
