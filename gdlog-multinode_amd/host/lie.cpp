// lie.cpp -- LIE::fixpoint_loop, the semi-naive loop of one SCC
// (reference: src/lie.cu:26-274), on the C ABI.
//
// Per iteration: run the RA operators in insertion order (a RelationalCopy
// from FULL only on the first iteration), drop the tmp relations' newt, then
// for each update relation compute delta = newt \ full and full = full U newt
// in one merge-path pass (gdl_merge_diff_u64), which is the reference's
// thrust::set_difference followed by Relation::flush_delta's thrust::merge.
#include <algorithm>
#include <iostream>

#include "gdlog/exception.h"
#include "gdlog/lie.h"

using namespace gdlog;

void LIE::add_ra(ra_op op) { ra_ops.push_back(op); }

void LIE::add_relations(Relation *rel, bool static_flag) {
    if (static_flag) static_relations.push_back(rel);
    else update_relations.push_back(rel);
}

void LIE::add_tmp_relation(Relation *rel) { tmp_relations.push_back(rel); }

namespace {

// delta := first full->tuple_counts raw rows, unsorted (copy_relation_container
// of lie.cu:56 copies full's data_raw rows and loads them without sorting).
void seed_delta(Relation *rel) {
    container_free(rel->delta);
    const tuple_size_t n = std::min(rel->full->tuple_counts, rel->raw_rows);
    if (n == 0) return;
    uint64_t *k = device_alloc_keys(n);
    check(gdl_memcpy_d2d(context(), k, rel->raw_keys, n * sizeof(uint64_t)), "seed delta");
    container_take(rel->delta, k, n, n, rel->raw_sorted);
}

// True when some join of the LIE reads rel's FULL as its inner side (NEWT as
// inner reads FULL too, join.cu:15-20).
bool full_is_inner(const std::vector<ra_op> &ops, const Relation *rel) {
    for (const auto &op : ops)
        if (const auto *j = std::get_if<RelationalJoin>(&op))
            if (j->inner_rel == rel && j->inner_ver != DELTA) return true;
    return false;
}

// True when only joins write rel and no operator reads rel's FULL or NEWT
// (reading DELTA is fine: the pending delta is a separate buffer).
bool output_fusable(const std::vector<ra_op> &ops, const Relation *rel) {
    for (const auto &op : ops) {
        if (const auto *j = std::get_if<RelationalJoin>(&op)) {
            if (j->inner_rel == rel && j->inner_ver != DELTA) return false;
            if (j->outer_rel == rel && j->outer_ver != DELTA) return false;
            if (j->output_rel == rel && (rel->tmp_flag || j->disable_load)) return false;
        } else if (const auto *c = std::get_if<RelationalCopy>(&op)) {
            if (c->dest_rel == rel) return false;
            if (c->src_rel == rel && c->src_ver != DELTA) return false;
        } else if (const auto *a = std::get_if<RelationalACopy>(&op)) {
            if (a->dest_rel == rel || a->src_rel == rel) return false;
        }
    }
    return true;
}

void enter_append_mode(Relation *rel) {
    gdl_ctx *ctx = context();
    const tuple_size_t n = rel->current_full_size;
    check(gdl_hashset_create(ctx, n + 1024, &rel->full_set), "gdl_hashset_create");
    if (n) {
        uint64_t *scratch = device_alloc_keys(n);
        uint64_t added = 0;
        gdl_status st = gdl_hashset_insert_new(ctx, &rel->full_set, rel->full->keys, n, scratch,
                                               &added);
        device_free(scratch);
        check(st, "gdl_hashset_insert_new");
    }
    rel->append_mode = true;
}

// FULL back to one sorted array; drops the hash set.
void leave_append_mode(Relation *rel) {
    gdl_ctx *ctx = context();
    const tuple_size_t n = rel->current_full_size;
    if (n && !rel->full->sorted) {
        uint64_t *sorted = device_alloc_keys(n);
        const int b = id_bits();
        gdl_status st = gdl_sort_u64(ctx, rel->full->keys, n, b, b, sorted);
        if (st != GDL_OK) device_free(sorted);
        check(st, "gdl_sort_u64");
        container_take(rel->full, sorted, n, n, true);
    }
    check(gdl_hashset_free(ctx, &rel->full_set), "gdl_hashset_free");
    device_free(rel->pend_keys);
    rel->pend_keys = nullptr;
    rel->pend_cap = rel->pend_n = rel->pend_raw = 0;
    rel->fused_output = false;
    rel->append_mode = false;
}

}  // namespace

void LIE::fixpoint_loop() {
    trace.clear();
    gdl_ctx *ctx = context();
    if (verbose()) std::cout << "start lie .... " << std::endl;
    for (Relation *rel : update_relations) {
        rel->current_full_size = rel->full->tuple_counts;
        seed_delta(rel);
        if (full_mode() == FullMode::Auto && !full_is_inner(ra_ops, rel)) {
            enter_append_mode(rel);
            rel->fused_output = output_fusable(ra_ops, rel);
        }
    }
    // A one-join LIE of TC's shape runs its bodies after the first as one
    // stream-ordered loop (RelationalJoin::run_linear): the same bodies, no
    // host synchronisation between them.
    RelationalJoin *linear = nullptr;
    if (ra_ops.size() == 1 && update_relations.size() == 1 && tmp_relations.empty()) {
        auto *j = std::get_if<RelationalJoin>(&ra_ops[0]);
        if (j && update_relations[0]->fused_output &&
            std::find(static_relations.begin(), static_relations.end(), j->inner_rel) !=
                static_relations.end())
            linear = j;
    }
    int iteration_counter = 0;
    while (true) {
        if (linear && iteration_counter > 0) {
            Relation *rel = update_relations[0];
            const long long left = (long long)max_iteration + 1 - iteration_counter;
            const int bodies = (int)std::min<long long>(left, 1 << 30);
            std::vector<uint64_t> nn, rr;
            tuple_size_t full = rel->current_full_size;
            if (bodies > 0 && linear->run_linear(bodies, &nn, &rr)) {
                bool fixpoint = false;
                for (size_t b = 0; b < nn.size(); ++b) {
                    full += nn[b];
                    trace.push_back({iteration_counter, 0, rr[b], nn[b], full});
                    if (verbose())
                        std::cout << "iteration " << iteration_counter << " relation "
                                  << rel->name << " finish dedup new tuples : " << nn[b]
                                  << " delta tuple size: " << nn[b] << " full counts " << full
                                  << std::endl;
                    iteration_counter++;
                    if (nn[b] == 0) {
                        fixpoint = true;
                        break;
                    }
                }
                if (fixpoint || iteration_counter > max_iteration) break;
                continue;
            }
            linear = nullptr;  // not that shape any more: one body at a time
        }
        for (auto &op : ra_ops) {
            if (auto *j = std::get_if<RelationalJoin>(&op)) {
                (*j)();
            } else if (auto *a = std::get_if<RelationalACopy>(&op)) {
                (*a)();
            } else if (auto *c = std::get_if<RelationalCopy>(&op)) {
                if (c->src_ver == FULL) {
                    if (!c->copied) {
                        (*c)();
                        c->copied = true;
                    }
                } else {
                    (*c)();
                }
            }
        }
        for (Relation *rel : tmp_relations) container_free(rel->newt);

        if (verbose())
            std::cout << "Iteration " << iteration_counter << " popluating new tuple" << std::endl;
        bool fixpoint_flag = true;
        int rel_idx = -1;
        for (Relation *rel : update_relations) {
            ++rel_idx;
            if (verbose()) std::cout << rel->name << std::endl;
            container_free(rel->delta);
            if (rel->fused_output) {
                // the set difference already happened inside gdl_join_insert
                const tuple_size_t nd = rel->pend_n, raw = rel->pend_raw;
                if (nd != 0) fixpoint_flag = false;
                if (nd > 0) {
                    rel->full->tuple_counts = rel->current_full_size;
                    container_append(rel->full, rel->pend_keys, nd);
                    rel->current_full_size += nd;
                    container_take(rel->delta, rel->pend_keys, nd, rel->pend_cap, false);
                    rel->pend_keys = nullptr;
                    rel->pend_cap = 0;
                }
                rel->pend_n = rel->pend_raw = 0;
                trace.push_back({iteration_counter, rel_idx, raw, nd, rel->current_full_size});
                if (verbose())
                    std::cout << "iteration " << iteration_counter << " relation " << rel->name
                              << " finish dedup new tuples : " << nd
                              << " delta tuple size: " << nd << " full counts "
                              << rel->current_full_size << std::endl;
                continue;
            }
            GHashRelContainer *newt = rel->newt;
            if (!newt->keys || newt->tuple_counts == 0) {
                container_free(newt);
                trace.push_back({iteration_counter, rel_idx, 0, 0, rel->current_full_size});
                if (verbose())
                    std::cout << "iteration " << iteration_counter << " relation " << rel->name
                              << " no new tuple added" << std::endl;
                continue;
            }
            if (rel->append_mode) {
                // delta = newt \ full through the hash set; full grows by delta
                const tuple_size_t nn = newt->tuple_counts;
                uint64_t *delta = device_alloc_keys(nn);
                uint64_t nd = 0;
                gdl_status st =
                    gdl_hashset_insert_new(ctx, &rel->full_set, newt->keys, nn, delta, &nd);
                if (st != GDL_OK) device_free(delta);
                check(st, "gdl_hashset_insert_new");
                if (nd != 0) fixpoint_flag = false;
                container_free(newt);
                if (nd > 0) {
                    // flush_delta restores full's count before growing it
                    // (relation.cu:514-515); a join may have zeroed it
                    rel->full->tuple_counts = rel->current_full_size;
                    container_append(rel->full, delta, nd);
                    rel->current_full_size += nd;
                }
                container_take(rel->delta, delta, nd, nn, false);
                trace.push_back({iteration_counter, rel_idx, nn, nd, rel->current_full_size});
                if (verbose())
                    std::cout << "iteration " << iteration_counter << " relation " << rel->name
                              << " finish dedup new tuples : " << nd
                              << " delta tuple size: " << nd << " full counts "
                              << rel->current_full_size << std::endl;
                continue;
            }
            container_normalize(newt);
            const tuple_size_t nn = newt->tuple_counts;
            const tuple_size_t nf = rel->current_full_size;
            uint64_t *new_full = device_alloc_keys(nf + nn);
            uint64_t *delta = device_alloc_keys(nn);
            uint64_t nu = 0, nd = 0;
            gdl_status st = gdl_merge_diff_u64(ctx, rel->full->keys, nf, newt->keys, nn, new_full,
                                               &nu, delta, &nd);
            if (st != GDL_OK) {
                device_free(new_full);
                device_free(delta);
                check(st, "gdl_merge_diff_u64");
            }
            if (nd != 0) fixpoint_flag = false;
            container_free(newt);
            container_take(rel->delta, delta, nd, nn, true);
            if (nd > 0) {
                container_take(rel->full, new_full, nu, nf + nn, true);
                rel->current_full_size = nu;
            } else {
                device_free(new_full);
            }
            trace.push_back({iteration_counter, rel_idx, nn, nd, rel->current_full_size});
            if (verbose())
                std::cout << "iteration " << iteration_counter << " relation " << rel->name
                          << " finish dedup new tuples : " << nd
                          << " delta tuple size: " << rel->delta->tuple_counts << " full counts "
                          << rel->current_full_size << std::endl;
        }
        if (verbose())
            std::cout << "Iteration " << iteration_counter << " finish populating" << std::endl;
        iteration_counter++;
        if (fixpoint_flag || iteration_counter > max_iteration) break;
    }
    iteration_count = iteration_counter;

    for (Relation *rel : update_relations) {
        if (rel->append_mode) leave_append_mode(rel);
        if (reload_full_flag) {
            // lie.cu:219-254: full reloaded from the merged tuples; its rows
            // (data_raw) become the sorted full.
            rel->full->tuple_counts = rel->current_full_size;
            device_free(rel->raw_keys);
            rel->raw_keys = nullptr;
            rel->raw_rows = 0;
            const tuple_size_t n = rel->current_full_size;
            if (n) {
                rel->raw_keys = device_alloc_keys(n);
                check(gdl_memcpy_d2d(ctx, rel->raw_keys, rel->full->keys, n * sizeof(uint64_t)),
                      "reload full");
                rel->raw_rows = n;
            }
            rel->raw_sorted = true;
            container_free(rel->delta);
            container_free(rel->newt);
        }
        if (verbose())
            std::cout << "Finished! " << rel->name << " has " << rel->full->tuple_counts
                      << std::endl;
    }
}
