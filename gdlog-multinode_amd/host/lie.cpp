// lie.cpp -- LIE::fixpoint_loop, the semi-naive loop of one SCC
// (reference: src/lie.cu:26-274), on the C ABI.
//
// Per iteration: run the RA operators in insertion order (a RelationalCopy
// from FULL only on the first iteration), drop the tmp relations' newt, then
// for each update relation compute delta = newt \ full and full = full U newt
// in one merge-path pass (gdl_merge_diff_u64), which is the reference's
// thrust::set_difference followed by Relation::flush_delta's thrust::merge.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>

#include "gdlog/exception.h"
#include "gdlog/lie.h"

#include <cstdlib>
#include <string>

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
    check(gdl_memcpy_d2d(context(), k, raw_rows_ptr(rel), n * sizeof(uint64_t)), "seed delta");
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
// True when an operator runs its device hooks (not a column map) or moves
// tuples of arity 3 or 4.
bool op_uses_hooks(const ra_op &op) {
    if (const auto *j = std::get_if<RelationalJoin>(&op))
        return j->uses_hooks() || j->inner_rel->arity > 2 || j->outer_rel->arity > 2 ||
               j->output_rel->arity > 2;
    if (const auto *c = std::get_if<RelationalCopy>(&op))
        return !c->copy_map.ok || c->src_rel->arity > 2 || c->dest_rel->arity > 2;
    if (const auto *a = std::get_if<RelationalACopy>(&op))
        return !a->copy_map.ok || a->src_rel->arity > 2 || a->dest_rel->arity > 2;
    return false;
}

// A relation whose FULL stays one sorted array merged every iteration (the
// reference's structure): dependent columns (a prefix set difference), keys
// of arity 3 or 4, or an output of a hook-running operator.
bool needs_sorted_full(const std::vector<ra_op> &ops, const Relation *rel) {
    if (rel->dependent_column_size > 0 || rel->arity > 2) return true;
    for (const auto &op : ops) {
        if (!op_uses_hooks(op)) continue;
        if (const auto *j = std::get_if<RelationalJoin>(&op)) {
            if (j->output_rel == rel) return true;
        } else if (const auto *c = std::get_if<RelationalCopy>(&op)) {
            if (c->dest_rel == rel) return true;
        } else if (const auto *a = std::get_if<RelationalACopy>(&op)) {
            if (a->dest_rel == rel) return true;
        }
    }
    return false;
}

// A tmp relation whose newt can be deduplicated by a set as the joins produce
// it: not a raw (tmp_flag) relation, written only by column-map joins, never
// read as a join inner (which would need it sorted).
bool tmp_dedup_ok(const std::vector<ra_op> &ops, const Relation *rel) {
    if (rel->tmp_flag || rel->arity > 2) return false;
    bool written = false;
    for (const auto &op : ops) {
        if (const auto *j = std::get_if<RelationalJoin>(&op)) {
            if (j->inner_rel == rel) return false;
            if (j->output_rel == rel) {
                if (op_uses_hooks(op) || j->disable_load) return false;
                written = true;
            }
        } else if (const auto *c = std::get_if<RelationalCopy>(&op)) {
            if (c->dest_rel == rel || c->src_rel == rel) return false;
        } else if (const auto *a = std::get_if<RelationalACopy>(&op)) {
            if (a->dest_rel == rel || a->src_rel == rel) return false;
        }
    }
    return written;
}

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

// A snapshot written in merge mode holds FULL sorted, with the delta inside
// it at sorted positions; the append log needs the delta as its tail (the
// linear loop reads the delta there).  FULL := (FULL \ delta) then delta.
// A log kept for its segment order (keep_full_log) cannot be rebuilt from a
// sorted FULL.
void log_with_delta_tail(Relation *rel, bool keep_log) {
    if (keep_log)
        throw Error(GDL_ERR_UNSUPPORTED, "resume of " + rel->name +
                                             ": the snapshot's FULL is sorted (merge mode), "
                                             "not the append log this LIE keeps");
    container_normalize(rel->delta);
    const tuple_size_t nf = rel->current_full_size, nd = rel->delta->tuple_counts;
    uint64_t *log = device_alloc_keys(nf + nd);
    uint64_t rest = 0;
    gdl_status st = gdl_set_difference_u64(context(), rel->full->keys, nf, rel->delta->keys, nd,
                                           log, &rest);
    if (st == GDL_OK && rest + nd != nf)
        st = GDL_ERR_INVALID_ARG;  // the delta must be part of FULL
    if (st == GDL_OK)
        st = gdl_memcpy_d2d(context(), log + rest, rel->delta->keys, nd * sizeof(uint64_t));
    if (st != GDL_OK) device_free(log);
    check(st, "resume: append log from a sorted FULL");
    container_take(rel->full, log, nf, nf + nd, false);
}

// FULL's membership set: a bitmap when every id of the program fits 17 bits
// (one bit per possible pair, at most 2 GiB, the size of a hash table of 90M
// members; an insert is one atomic OR and the set never grows), else the hash
// table.  GDLOG_FULL_SET=hash forces the hash table.
static bool bitmap_full_set(bool hooks) {
    const int b = id_bits();
    // hooks may produce ids beyond those loaded so far
    if (hooks || b < 1 || b > 17) return false;
    const char *e = std::getenv("GDLOG_FULL_SET");
    return !(e && std::string(e) == "hash");
}

void enter_append_mode(Relation *rel, bool hooks) {
    gdl_ctx *ctx = context();
    const tuple_size_t n = rel->current_full_size;
    if (bitmap_full_set(hooks))
        check(gdl_hashset_create_bitmap(ctx, (uint32_t)id_bits(), &rel->full_set),
              "gdl_hashset_create_bitmap");
    else
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

// FULL's log rebuilt from its set (in key order) after bit-matrix bodies that
// logged only their last body's tuples.
void materialize_log(Relation *rel) {
    const tuple_size_t n = rel->current_full_size;
    uint64_t *keys = n ? device_alloc_keys(n) : nullptr;
    uint64_t got = 0;
    gdl_status st = n ? gdl_hashset_extract_sorted(context(), &rel->full_set, keys, n, &got)
                      : GDL_OK;
    if (st == GDL_OK && got != n) st = GDL_ERR_DEVICE;
    if (st != GDL_OK) device_free(keys);
    check(st, "FULL's log from its set");
    container_take(rel->full, keys, n, n, true);
    rel->log_stale = false;
}

// FULL back to one sorted array (unless keep_log: FULL stays the append log,
// its iterations' deltas in order); drops the hash set.
void leave_append_mode(Relation *rel, bool keep_log) {
    gdl_ctx *ctx = context();
    const tuple_size_t n = rel->current_full_size;
    if (n && !rel->full->sorted && !keep_log) {
        uint64_t *sorted = device_alloc_keys(n);
        const int kb = (int)rel->full_set.key_bits;
        gdl_status st;
        if (kb && (rel->log_stale || (1ull << (2 * kb)) / 8 <= 32ull * n)) {
            // a bitmap set is FULL in key order: two streaming passes over it
            // instead of a sort of the log (when the bitmap is not much larger
            // than the keys)
            uint64_t got = 0;
            st = gdl_hashset_extract_sorted(ctx, &rel->full_set, sorted, n, &got);
            if (st == GDL_OK && got != n) {
                device_free(sorted);
                throw Error(GDL_ERR_DEVICE, rel->name + ": the FULL set holds " +
                                                std::to_string(got) + " tuples, its log " +
                                                std::to_string(n));
            }
        } else {
            const int b = id_bits();
            st = gdl_sort_u64(ctx, rel->full->keys, n, b, b, sorted);
        }
        if (st != GDL_OK) device_free(sorted);
        check(st, "sorted FULL");
        container_take(rel->full, sorted, n, n, true);
    }
    if (rel->log_stale && n && !rel->full->sorted)
        throw Error(GDL_ERR_DEVICE, rel->name + ": FULL's log is stale");
    rel->log_stale = false;
    check(gdl_hashset_check(ctx, &rel->full_set), "gdl_hashset_check");
    check(gdl_hashset_free(ctx, &rel->full_set), "gdl_hashset_free");
    device_free(rel->pend_keys);
    rel->pend_keys = nullptr;
    rel->pend_cap = rel->pend_n = rel->pend_raw = 0;
    rel->fused_output = false;
    rel->append_mode = false;
}

// ---- checkpoint / resume ---------------------------------------------------
// Snapshot file (host byte order):
//   "GDLCKPT1", u32 update relations, u32 static relations, i32 iteration
//   counter, i32 finished, u64 trace records, the records (i32, i32, u64 x 3);
//   per static relation: name, u64 tuples, u64 checksum (a fingerprint: a
//   resumed run must have loaded the same inputs);
//   per update relation: name, i32 arity, u8 FULL sorted, u8 delta sorted,
//   u64 |FULL|, u64 |delta|, the FULL keys, the delta keys.
// A name is u32 length + bytes.
constexpr char kCkptMagic[8] = {'G', 'D', 'L', 'C', 'K', 'P', 'T', '1'};

struct CkptDefaults {
    std::string dir;
    int every = 0;
    bool resume = false;
    int stop_after = -1;
    int next = 0;
};
thread_local CkptDefaults g_ckpt;

class CkptFile {
  public:
    CkptFile(const std::string &path, const char *mode) : f_(std::fopen(path.c_str(), mode)) {}
    ~CkptFile() {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr; }
    void put(const void *p, size_t n) {
        if (n && std::fwrite(p, 1, n, f_) != n) throw Error(GDL_ERR_INVALID_ARG, "checkpoint write failed");
    }
    void get(void *p, size_t n) {
        if (n && std::fread(p, 1, n, f_) != n)
            throw Error(GDL_ERR_INVALID_ARG, "checkpoint file is truncated");
    }
    template <typename T>
    void put_v(T v) { put(&v, sizeof(v)); }
    template <typename T>
    T get_v() {
        T v;
        get(&v, sizeof(v));
        return v;
    }
    void put_name(const std::string &s) {
        put_v<uint32_t>((uint32_t)s.size());
        put(s.data(), s.size());
    }
    std::string get_name() {
        const uint32_t n = get_v<uint32_t>();
        if (n > (1u << 20)) throw Error(GDL_ERR_INVALID_ARG, "checkpoint file is corrupt");
        std::string s(n, '\0');
        get(&s[0], n);
        return s;
    }
    void close() {
        if (f_ && std::fclose(f_) != 0) {
            f_ = nullptr;
            throw Error(GDL_ERR_INVALID_ARG, "checkpoint write failed");
        }
        f_ = nullptr;
    }

  private:
    FILE *f_;
};

void put_keys(CkptFile &f, const uint64_t *d_keys, tuple_size_t n) {
    if (n == 0) return;
    std::vector<uint64_t> h(n);
    check(gdl_memcpy_d2h(context(), h.data(), d_keys, n * sizeof(uint64_t)), "checkpoint D2H");
    f.put(h.data(), n * sizeof(uint64_t));
}

uint64_t *get_keys(CkptFile &f, tuple_size_t n) {
    if (n == 0) return nullptr;
    std::vector<uint64_t> h(n);
    f.get(h.data(), n * sizeof(uint64_t));
    uint64_t hi = 0, lo = 0;
    for (uint64_t k : h) {
        hi |= k >> 32;
        lo |= k & 0xffffffffull;
    }
    note_id_bits(std::max(hi ? 64 - __builtin_clzll(hi) : 0, lo ? 64 - __builtin_clzll(lo) : 0));
    uint64_t *d = device_alloc_keys(n);
    check(gdl_memcpy_h2d(context(), d, h.data(), n * sizeof(uint64_t)), "checkpoint H2D");
    return d;
}

// The state at an iteration boundary: every update relation's newt has been
// consumed, its delta is the next iteration's outer, and its FULL holds
// current_full_size keys (a sorted array, or the append log).
void write_snapshot(const LIE &lie, const std::string &path, int iteration, bool finished) {
    const std::string tmp = path + ".tmp";
    {
        CkptFile f(tmp, "wb");
        if (!f.ok()) throw Error(GDL_ERR_INVALID_ARG, "cannot write checkpoint " + tmp);
        f.put(kCkptMagic, sizeof(kCkptMagic));
        f.put_v<uint32_t>((uint32_t)lie.update_relations.size());
        f.put_v<uint32_t>((uint32_t)lie.static_relations.size());
        f.put_v<int32_t>(iteration);
        f.put_v<int32_t>(finished ? 1 : 0);
        f.put_v<uint64_t>(lie.trace.size());
        for (const auto &t : lie.trace) {
            f.put_v<int32_t>(t.iteration);
            f.put_v<int32_t>(t.relation);
            f.put_v<uint64_t>(t.newt);
            f.put_v<uint64_t>(t.delta);
            f.put_v<uint64_t>(t.full);
        }
        for (Relation *rel : lie.static_relations) {
            f.put_name(rel->name);
            f.put_v<uint64_t>(rel->full->tuple_counts);
            f.put_v<uint64_t>(relation_checksum(rel->full));
        }
        for (Relation *rel : lie.update_relations) {
            const tuple_size_t nf = rel->current_full_size;
            const tuple_size_t nd = rel->delta->keys ? rel->delta->tuple_counts : 0;
            f.put_name(rel->name);
            f.put_v<int32_t>(rel->arity);
            f.put_v<uint8_t>(rel->full->sorted ? 1 : 0);
            f.put_v<uint8_t>(rel->delta->sorted ? 1 : 0);
            f.put_v<uint64_t>(nf);
            f.put_v<uint64_t>(nd);
            put_keys(f, rel->full->keys, nf);
            put_keys(f, rel->delta->keys, nd);
        }
        f.close();
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0)
        throw Error(GDL_ERR_INVALID_ARG, "cannot move checkpoint into place: " + path);
}

// Restores a snapshot into the LIE's update relations; returns false when
// the file does not exist.  The static relations must match the snapshot's.
bool read_snapshot(LIE &lie, const std::string &path, int *iteration, bool *finished) {
    CkptFile f(path, "rb");
    if (!f.ok()) return false;
    char magic[8];
    f.get(magic, sizeof(magic));
    if (std::memcmp(magic, kCkptMagic, sizeof(magic)) != 0)
        throw Error(GDL_ERR_INVALID_ARG, "not a checkpoint file: " + path);
    const uint32_t nu = f.get_v<uint32_t>(), ns = f.get_v<uint32_t>();
    if (nu != lie.update_relations.size() || ns != lie.static_relations.size())
        throw Error(GDL_ERR_INVALID_ARG, "checkpoint " + path + " is of another LIE");
    *iteration = f.get_v<int32_t>();
    *finished = f.get_v<int32_t>() != 0;
    const uint64_t nt = f.get_v<uint64_t>();
    if (nt > (1ull << 32)) throw Error(GDL_ERR_INVALID_ARG, "checkpoint file is corrupt");
    lie.trace.clear();
    for (uint64_t i = 0; i < nt; ++i) {
        IterationRecord t;
        t.iteration = f.get_v<int32_t>();
        t.relation = f.get_v<int32_t>();
        t.newt = f.get_v<uint64_t>();
        t.delta = f.get_v<uint64_t>();
        t.full = f.get_v<uint64_t>();
        lie.trace.push_back(t);
    }
    for (Relation *rel : lie.static_relations) {
        const std::string name = f.get_name();
        const uint64_t n = f.get_v<uint64_t>(), sum = f.get_v<uint64_t>();
        if (name != rel->name || n != rel->full->tuple_counts || sum != relation_checksum(rel->full))
            throw Error(GDL_ERR_INVALID_ARG, "checkpoint " + path + ": static relation " +
                                                 rel->name + " differs from the snapshot's");
    }
    for (Relation *rel : lie.update_relations) {
        const std::string name = f.get_name();
        const int arity = f.get_v<int32_t>();
        const bool full_sorted = f.get_v<uint8_t>() != 0, delta_sorted = f.get_v<uint8_t>() != 0;
        const uint64_t nf = f.get_v<uint64_t>(), nd = f.get_v<uint64_t>();
        if (name != rel->name || arity != rel->arity)
            throw Error(GDL_ERR_INVALID_ARG, "checkpoint " + path + ": relation " + rel->name +
                                                 " is not the snapshot's " + name);
        materialize_raw(rel);  // FULL is replaced below
        uint64_t *full = get_keys(f, nf);
        container_take(rel->full, full, nf, nf, full_sorted);
        rel->current_full_size = nf;
        uint64_t *delta = get_keys(f, nd);
        container_free(rel->delta);
        if (nd) container_take(rel->delta, delta, nd, nd, delta_sorted);
        container_free(rel->newt);
    }
    return true;
}

// An SG-shaped LIE (gdl_fixpoint_dense's GDL_DENSE_SG): out(x, y) :-
// r1(a, x), delta(a, b), r2(b, y) as a join of a static r1 with out's delta
// into a deduplicated tmp (b, x), then a join of a static r2 with the tmp
// into out, both with the (outer column 1, inner column 1) map and no
// predicate.  Null when the LIE has another shape.
struct Chain {
    RelationalJoin *j1 = nullptr, *j2 = nullptr;
};

Chain sg_chain(std::vector<ra_op> &ops, const std::vector<Relation *> &updates,
               const std::vector<Relation *> &tmps, const std::vector<Relation *> &statics) {
    Chain c;
    if (ops.size() != 2 || updates.size() != 1 || tmps.size() != 1) return c;
    auto *j1 = std::get_if<RelationalJoin>(&ops[0]);
    auto *j2 = std::get_if<RelationalJoin>(&ops[1]);
    Relation *out = updates[0], *tmp = tmps[0];
    if (!j1 || !j2 || !out->fused_output || !out->full_set.key_bits || !tmp->tmp_dedup) return c;
    auto is_static = [&](const Relation *r) {
        return std::find(statics.begin(), statics.end(), r) != statics.end();
    };
    auto swap_map = [](const RelationalJoin *j) {
        return j->gen_map.ok && j->pred_map.ok && j->pred_map.pred.op == GDL_PRED_NONE &&
               j->gen_map.map.arity == 2 && j->gen_map.map.src[0] == GDL_COL_OUTER1 &&
               j->gen_map.map.src[1] == GDL_COL_INNER1 && !j->disable_load;
    };
    auto binary = [](const Relation *r) { return r->arity == 2 && r->index_column_size == 1; };
    if (swap_map(j1) && swap_map(j2) && j1->output_rel == tmp && j1->outer_rel == out &&
        j1->outer_ver == DELTA && j1->inner_ver == FULL && is_static(j1->inner_rel) &&
        j2->outer_rel == tmp && j2->outer_ver == NEWT && j2->inner_ver == FULL &&
        is_static(j2->inner_rel) && j2->output_rel == out && binary(out) && binary(tmp) &&
        binary(j1->inner_rel) && binary(j2->inner_rel)) {
        c.j1 = j1;
        c.j2 = j2;
    }
    return c;
}

// Runs up to max_bodies bodies of the chain on bit matrices
// (gdl_fixpoint_dense); the set and the delta are left as the per-iteration
// bodies would leave them, and FULL's log too unless log_last: then the log
// gets only the last body's tuples (the delta) and is marked stale (the set
// holds FULL).  Returns false when no body ran.
bool run_dense_chain(const Chain &c, Relation *rel, int max_bodies, bool log_last,
                     std::vector<uint64_t> *new_tuples, std::vector<uint64_t> *raw_tuples) {
    GHashRelContainer *delta = rel->delta, *full = rel->full;
    GHashRelContainer *r1 = c.j1->inner_rel->full, *r2 = c.j2->inner_rel->full;
    new_tuples->clear();
    raw_tuples->clear();
    if (max_bodies <= 0 || !delta->keys || delta->tuple_counts == 0 || !r1->keys ||
        !r1->tuple_counts || !r2->keys || !r2->tuple_counts)
        return false;
    const gdl_dense_rule rule{GDL_DENSE_SG, r1->keys, r1->tuple_counts, r2->keys,
                              r2->tuple_counts, log_last ? (uint32_t)GDL_DENSE_LOG_LAST : 0u, 0};
    uint64_t *log = full->keys;
    // the log's length: FULL's count, or less when it is stale already
    uint64_t cap = full->capacity, end = rel->log_stale ? full->tuple_counts : rel->current_full_size;
    const uint64_t start = rel->current_full_size;
    const uint64_t *d = delta->keys;
    uint64_t nd = delta->tuple_counts;
    constexpr int kSlice = 4096;
    std::vector<uint64_t> nb(kSlice), rb(kSlice);
    int left = max_bodies;
    gdl_status st = GDL_OK;
    while (left > 0) {
        int bodies = 0;
        st = gdl_fixpoint_dense(context(), &rule, &rel->full_set, d, nd, &log, &cap, &end,
                                std::min(left, kSlice), nb.data(), rb.data(), &bodies);
        if (st != GDL_OK || bodies == 0) break;
        new_tuples->insert(new_tuples->end(), nb.begin(), nb.begin() + bodies);
        raw_tuples->insert(raw_tuples->end(), rb.begin(), rb.begin() + bodies);
        left -= bodies;
        nd = nb[bodies - 1];
        d = log + (end - nd);
        if (nd == 0 || bodies < std::min(left + bodies, kSlice)) break;  // fixpoint or sparse
    }
    uint64_t added = 0;
    for (uint64_t v : *new_tuples) added += v;
    if (!new_tuples->empty() || log != full->keys) {
        full->keys = log;
        full->capacity = cap;
        full->tuple_counts = end;
        full->data_raw_row_size = end;
        full->sorted = false;
        rel->current_full_size = start + added;
        rel->log_stale = end != rel->current_full_size;
    }
    check(st, "gdl_fixpoint_dense");
    if (new_tuples->empty()) return false;
    container_free(delta);
    const uint64_t last = new_tuples->back();
    if (last) {
        uint64_t *dk = device_alloc_keys(last);
        check(gdl_memcpy_d2d(context(), dk, log + (end - last), last * sizeof(uint64_t)),
              "dense delta");
        container_take(delta, dk, last, last, false);
    }
    return true;
}

}  // namespace

namespace gdlog {
void set_checkpoint_defaults(const char *dir, int every, bool resume, int stop_after) {
    g_ckpt.dir = dir ? dir : "";
    g_ckpt.every = every > 0 ? every : 0;
    g_ckpt.resume = resume;
    g_ckpt.stop_after = stop_after;
    g_ckpt.next = 0;
}
void reset_checkpoint_sequence() { g_ckpt.next = 0; }
}  // namespace gdlog

void LIE::fixpoint_loop() {
    trace.clear();
    gdl_ctx *ctx = context();
    struct Scope {
        Scope() { enter_lie_scope(); }
        ~Scope() { leave_lie_scope(); }
    } id_scope;
    if (verbose()) std::cout << "start lie .... " << std::endl;
    // checkpoint files: the LIE's own, else the thread's defaults
    std::string ck_file = checkpoint_file, rs_file = resume_file;
    int ck_every = checkpoint_every, ck_stop = stop_after;
    if (ck_file.empty() && rs_file.empty() && !g_ckpt.dir.empty()) {
        ck_file = g_ckpt.dir + "/lie" + std::to_string(g_ckpt.next++) + ".ckpt";
        ck_every = g_ckpt.every;
        ck_stop = g_ckpt.stop_after;
        if (g_ckpt.resume) rs_file = ck_file;
    }
    int iteration_counter = 0;
    bool finished = false;
    const bool resumed =
        !rs_file.empty() && read_snapshot(*this, rs_file, &iteration_counter, &finished);
    if (resumed) {
        // a snapshot is taken after iteration 0, which ran the FULL copies
        for (auto &op : ra_ops)
            if (auto *c = std::get_if<RelationalCopy>(&op))
                if (c->src_ver == FULL) c->copied = true;
    }
    // keep_full_log needs the append log whatever the thread's FULL mode
    const bool auto_mode = keep_full_log || full_mode() == FullMode::Auto;
    bool lie_uses_hooks = false;
    for (const auto &op : ra_ops) lie_uses_hooks = lie_uses_hooks || op_uses_hooks(op);
    // tmp relations whose newt is a set: deduplicated by a per-iteration set
    for (Relation *rel : tmp_relations)
        rel->tmp_dedup = auto_mode && tmp_dedup_ok(ra_ops, rel);
    for (Relation *rel : update_relations) {
        if (!resumed) {
            rel->current_full_size = rel->full->tuple_counts;
            seed_delta(rel);
        }
        // the raw rows outlive this fixpoint's changes to FULL unless its
        // end reloads them
        if (!reload_full_flag) materialize_raw(rel);
        if (auto_mode && !full_is_inner(ra_ops, rel) && !needs_sorted_full(ra_ops, rel)) {
            if (resumed && rel->full->sorted && rel->delta->keys && rel->delta->tuple_counts)
                log_with_delta_tail(rel, keep_full_log);
            enter_append_mode(rel, lie_uses_hooks);
            rel->fused_output = output_fusable(ra_ops, rel);
        } else if (resumed && !rel->full->sorted) {
            container_normalize(rel->full);  // the merge needs a sorted FULL
            rel->current_full_size = rel->full->tuple_counts;
        }
    }
    // an iteration boundary: the snapshot when one is due; stop_after ends
    // the run right after its snapshot
    auto boundary = [&](bool fixpoint) {
        if (ck_file.empty()) return;
        const bool stop = !fixpoint && ck_stop >= 0 && iteration_counter >= ck_stop;
        if (fixpoint || stop || iteration_counter > max_iteration ||
            (ck_every > 0 && iteration_counter % ck_every == 0))
            write_snapshot(*this, ck_file, iteration_counter, fixpoint);
        if (stop)
            throw Error(GDL_ERR_BUSY, "fixpoint stopped after its checkpoint at iteration " +
                                          std::to_string(iteration_counter));
    };
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
    // A two-join LIE of SG's shape runs its bodies on bit matrices while its
    // delta is dense.
    const Chain chain = linear ? Chain{} : sg_chain(ra_ops, update_relations, tmp_relations,
                                                    static_relations);
    // An LIE whose operators are all joins of two static relations computes at
    // iteration 1 exactly what it computed at iteration 0 (the same newt, all
    // of it already in FULL): iteration 1 is recorded from iteration 0 -- the
    // same newt counts, no new tuple, the fixpoint -- without running it.
    bool static_joins = !ra_ops.empty();
    for (const auto &op : ra_ops) {
        const auto *j = std::get_if<RelationalJoin>(&op);
        static_joins = static_joins && j &&
                       std::find(static_relations.begin(), static_relations.end(),
                                 j->inner_rel) != static_relations.end() &&
                       std::find(static_relations.begin(), static_relations.end(),
                                 j->outer_rel) != static_relations.end();
    }
    while (!finished) {
        if (static_joins && iteration_counter == 1 && trace.size() >= update_relations.size()) {
            const size_t first = trace.size() - update_relations.size();
            for (size_t r = 0; r < update_relations.size(); ++r) {
                const IterationRecord t0 = trace[first + r];
                trace.push_back({1, t0.relation, t0.newt, 0, update_relations[r]->current_full_size});
                if (verbose())
                    std::cout << "iteration 1 relation " << update_relations[r]->name
                              << " no new tuple added" << std::endl;
            }
            iteration_counter = 2;
            boundary(true);
            break;
        }
        if (chain.j1) {
            Relation *rel = update_relations[0];
            long long left = (long long)max_iteration + 1 - iteration_counter;
            if (!ck_file.empty()) {  // land on the next snapshot
                if (ck_every > 0) left = std::min<long long>(left, ck_every - iteration_counter % ck_every);
                if (ck_stop > iteration_counter)
                    left = std::min<long long>(left, ck_stop - iteration_counter);
            }
            std::vector<uint64_t> nn, rr;
            // without snapshots or a kept log, FULL's log is read only at the
            // end of the fixpoint, where the set gives FULL: the bodies log
            // only their last body's tuples (the delta)
            const bool log_last = ck_file.empty() && !keep_full_log;
            if (left > 0 && run_dense_chain(chain, rel, (int)std::min<long long>(left, 1 << 30),
                                            log_last, &nn, &rr)) {
                bool fixpoint = false;
                tuple_size_t full = rel->current_full_size;
                for (uint64_t v : nn) full -= v;
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
                boundary(fixpoint);
                if (fixpoint || iteration_counter > max_iteration) break;
                if (rel->log_stale) materialize_log(rel);  // tuple-at-a-time bodies next
                continue;
            }
        }
        if (linear && iteration_counter > 0) {
            Relation *rel = update_relations[0];
            long long left = (long long)max_iteration + 1 - iteration_counter;
            if (!ck_file.empty()) {  // land on the next snapshot
                if (ck_every > 0) left = std::min<long long>(left, ck_every - iteration_counter % ck_every);
                if (ck_stop > iteration_counter)
                    left = std::min<long long>(left, ck_stop - iteration_counter);
            }
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
                boundary(fixpoint);
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
        for (Relation *rel : tmp_relations) {
            container_free(rel->newt);
            if (rel->tmp_set.slots) check(gdl_hashset_free(ctx, &rel->tmp_set), "gdl_hashset_free");
        }

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
            gdl_status st;
            if (rel->dependent_column_size > 0) {
                // set difference over the first arity - dependent columns
                // (lie.cu:131-138), then the merge of the delta into FULL
                const int last = rel->arity - rel->dependent_column_size - 1;
                st = gdl_set_difference_prefix_u64(ctx, newt->keys, nn, rel->full->keys, nf,
                                                   rel->layout.shift[last], delta, &nd);
                if (st == GDL_OK)
                    st = gdl_merge_union_u64(ctx, rel->full->keys, nf, delta, nd, new_full, &nu);
            } else {
                st = gdl_merge_diff_u64(ctx, rel->full->keys, nf, newt->keys, nn, new_full, &nu,
                                        delta, &nd);
            }
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
        boundary(fixpoint_flag);
        if (fixpoint_flag || iteration_counter > max_iteration) break;
    }
    iteration_count = iteration_counter;
    for (Relation *rel : tmp_relations) {
        rel->tmp_dedup = false;
        if (rel->tmp_set.slots) check(gdl_hashset_free(ctx, &rel->tmp_set), "gdl_hashset_free");
    }

    for (Relation *rel : update_relations) {
        if (rel->append_mode) leave_append_mode(rel, keep_full_log);
        if (reload_full_flag) {
            // lie.cu:219-254: full reloaded from the merged tuples; its rows
            // (data_raw) become the sorted full.
            rel->full->tuple_counts = rel->current_full_size;
            raw_rows_from_full(rel);  // the rows are FULL's keys until FULL changes
            container_free(rel->delta);
            container_free(rel->newt);
        }
        if (verbose())
            std::cout << "Finished! " << rel->name << " has " << rel->full->tuple_counts
                      << std::endl;
    }
}
