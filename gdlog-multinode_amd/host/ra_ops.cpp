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

}  // namespace

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
    if (!gen_map.ok)
        throw Error(GDL_ERR_UNSUPPORTED, "RelationalJoin into " + output_rel->name +
                                             ": the generator hook is not a column selection");
    if (!pred_map.ok)
        throw Error(GDL_ERR_UNSUPPORTED, "RelationalJoin into " + output_rel->name +
                                             ": the predicate is not a column comparison");
    const int k = (int)outer->index_column_size;
    if (k < 1 || k > inner->arity || k > outer->arity)
        throw Error(GDL_ERR_INVALID_ARG, "RelationalJoin: bad join column count");
    const gdl_index *idx;
    {
        Phase ph(detail_time, 5);
        idx = container_index(inner, k);
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
    if (!copy_map.ok)
        throw Error(GDL_ERR_UNSUPPORTED, "RelationalCopy into " + dest_rel->name +
                                             ": the copy hook is not a column selection");
    tuple_size_t n = src->tuple_counts;
    uint64_t *copied = device_alloc_keys(n);
    check(gdl_project_u64(context(), src->keys, n, copy_map.map, copied), "gdl_project_u64");
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
    if (!copy_map.ok)
        throw Error(GDL_ERR_UNSUPPORTED, "RelationalACopy into " + dest_rel->name +
                                             ": the copy hook is not a column selection");
    const tuple_size_t n = src->tuple_counts;
    uint64_t *copied = device_alloc_keys(n);
    check(gdl_project_u64(context(), src->keys, n, copy_map.map, copied), "gdl_project_u64");
    if (dest_rel->append_mode) {
        container_take(dest, copied, n, n, false);
        return;
    }
    uint64_t *su = nullptr;
    tuple_size_t m = sort_unique(copied, n, &su);
    device_free(copied);
    container_take(dest, su, m, n, true);
}
