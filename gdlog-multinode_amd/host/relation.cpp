// relation.cpp -- Relation / GHashRelContainer storage of the gdlog C++ API
// (reference: include/relation.cuh, src/relation.cu).  Calls the engine only
// through the C ABI of include/gdl_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "gdlog/exception.h"
#include "gdlog/relation.h"

namespace gdlog {

namespace {
// One engine context per host thread (the engine's contexts are not
// re-entrant); programs on different threads run on their own streams.
thread_local gdl_ctx *g_ctx = nullptr;
thread_local int g_id_bits = 0;
bool g_phase_timing = false;
int g_verbose = -1;
}  // namespace

void check(gdl_status st, const char *what) {
    if (st == GDL_OK) return;
    std::string msg = std::string(what) + ": " + gdl_status_string(st);
    const char *detail = g_ctx ? gdl_last_error(g_ctx) : "";
    if (detail && *detail) msg += std::string(" (") + detail + ")";
    throw Error(st, msg);
}

gdl_ctx *context() {
    if (!g_ctx) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        gdl_config cfg;
        gdl_config_default(&cfg);
        cfg.device = dev;
        cfg.verbose = verbose() ? 1 : 0;
        check(gdl_ctx_create(&cfg, &g_ctx), "gdl_ctx_create");
    }
    return g_ctx;
}

gdl_ctx *context_or_null() { return g_ctx; }

void set_context(gdl_ctx *ctx) { g_ctx = ctx; }

int id_bits() { return g_id_bits; }
// The engine context's id bound (gdl_set_id_bits) is set only while a
// fixpoint of this thread runs, and follows note_id_bits meanwhile: a direct
// engine call on the same context outside a fixpoint is never sized by it.
thread_local int g_lie_depth = 0;
void note_id_bits(int bits) {
    g_id_bits = std::max(g_id_bits, std::min(bits, 32));
    if (g_ctx && g_lie_depth > 0) (void)gdl_set_id_bits(g_ctx, g_id_bits);
}
void reset_id_bits() { g_id_bits = 0; }
void enter_lie_scope() {
    if (g_lie_depth++ == 0 && g_ctx) (void)gdl_set_id_bits(g_ctx, g_id_bits);
}
void leave_lie_scope() {
    if (--g_lie_depth == 0 && g_ctx) (void)gdl_set_id_bits(g_ctx, 0);
}

bool verbose() {
    if (g_verbose < 0) {
        const char *e = std::getenv("GDLOG_VERBOSE");
        g_verbose = (e && *e && *e != '0') ? 1 : 0;
    }
    return g_verbose == 1;
}

void set_phase_timing(bool on) { g_phase_timing = on; }
bool phase_timing() { return g_phase_timing; }

namespace {
// Per thread, like the context: a program on one thread never switches the
// mode of programs running on others.
thread_local FullMode g_full_mode = [] {
    const char *e = std::getenv("GDLOG_FULL_MODE");
    return (e && std::string(e) == "merge") ? FullMode::Merge : FullMode::Auto;
}();
}  // namespace

void set_full_mode(FullMode m) { g_full_mode = m; }
FullMode full_mode() { return g_full_mode; }

uint64_t *device_alloc_keys(tuple_size_t n) {
    void *p = nullptr;
    check(gdl_alloc(context(), (n ? n : 1) * sizeof(uint64_t), &p), "gdl_alloc");
    return static_cast<uint64_t *>(p);
}

void device_free(void *p) {
    if (p) check(gdl_free(context(), p), "gdl_free");
}

void container_drop_index(GHashRelContainer *c) {
    if (c->index.slots) check(gdl_index_free(context(), &c->index), "gdl_index_free");
    c->index = gdl_index{};
    c->index_prefix_cols = 0;
    c->index_map_size = 0;
}

void container_free(GHashRelContainer *c) {
    container_drop_index(c);
    device_free(c->keys);
    c->keys = nullptr;
    c->capacity = 0;
    c->tuple_counts = 0;
    c->data_raw_row_size = 0;
    c->sorted = true;
}

void container_take(GHashRelContainer *c, uint64_t *keys, tuple_size_t n, tuple_size_t cap,
                    bool sorted) {
    container_free(c);
    c->keys = keys;
    c->capacity = cap;
    c->tuple_counts = n;
    c->data_raw_row_size = n;
    c->sorted = sorted;
}

void container_append(GHashRelContainer *c, const uint64_t *keys, tuple_size_t n) {
    if (n == 0) return;
    const tuple_size_t have = c->keys ? c->tuple_counts : 0;
    container_drop_index(c);
    if (!c->keys || c->capacity < have + n) {
        const tuple_size_t cap = (have + n) + (have + n) / 2 + 1024;
        uint64_t *buf = device_alloc_keys(cap);
        if (have)
            check(gdl_memcpy_d2d(context(), buf, c->keys, have * sizeof(uint64_t)), "append");
        device_free(c->keys);
        c->keys = buf;
        c->capacity = cap;
    }
    check(gdl_memcpy_d2d(context(), c->keys + have, keys, n * sizeof(uint64_t)), "append");
    c->tuple_counts = have + n;
    c->data_raw_row_size = have + n;
    c->sorted = false;
}

tuple_size_t sort_unique(const uint64_t *d_in, tuple_size_t n, uint64_t **d_out) {
    *d_out = nullptr;
    if (n == 0) return 0;
    uint64_t *out = device_alloc_keys(n);
    uint64_t m = 0;
    const int b = id_bits();
    gdl_status st = gdl_sort_unique_u64(context(), d_in, n, b, b, out, &m);
    if (st != GDL_OK) {
        device_free(out);
        check(st, "gdl_sort_unique_u64");
    }
    *d_out = out;
    return m;
}

void container_normalize(GHashRelContainer *c) {
    if (c->sorted || c->tuple_counts == 0) {
        c->sorted = true;
        return;
    }
    uint64_t *out = nullptr;
    tuple_size_t m = sort_unique(c->keys, c->tuple_counts, &out);
    container_take(c, out, m, c->tuple_counts, true);
}

const gdl_index *container_index(GHashRelContainer *c, int prefix_cols) {
    if (c->index.slots && c->index_prefix_cols == prefix_cols) return &c->index;
    container_drop_index(c);
    container_normalize(c);
    check(gdl_index_build(context(), c->keys, c->tuple_counts, prefix_cols, &c->index),
          "gdl_index_build");
    c->index_prefix_cols = prefix_cols;
    c->index_map_size = c->index.n_slots;
    return &c->index;
}

}  // namespace gdlog

using namespace gdlog;

namespace {

int bit_length(uint64_t v) { return v ? 64 - __builtin_clzll((unsigned long long)v) : 0; }

KeyLayout layout_from_bits(int arity, const int *low_bits) {
    KeyLayout L;
    L.arity = arity;
    for (int c = 0; c < 4; ++c) L.bits[c] = L.shift[c] = 0;
    L.bits[0] = 32;
    L.shift[0] = 32;
    if (arity == 2) L.bits[1] = 32;
    if (arity >= 3) {
        int at = 32;
        for (int c = 1; c < arity; ++c) {
            L.bits[c] = low_bits[c - 1];
            at -= L.bits[c];
            L.shift[c] = at;
        }
    }
    return L;
}

// Key layout of a relation being loaded with `rows` rows of host columns.
KeyLayout choose_layout(const Relation *r, const column_type *data, tuple_size_t rows) {
    const int a = r->arity;
    if (a <= 2) {
        KeyLayout L;
        L.arity = a;
        if (a == 1) L.bits[1] = 0;
        return L;
    }
    int low[3] = {0, 0, 0};
    if (!r->column_bits.empty()) {
        if ((int)r->column_bits.size() != a - 1)
            throw Error(GDL_ERR_INVALID_ARG, "relation " + r->name +
                                                 ": column_bits needs one width per column 1.." +
                                                 std::to_string(a - 1));
        int sum = 0;
        for (int c = 0; c < a - 1; ++c) {
            low[c] = r->column_bits[c];
            if (low[c] < 1) throw Error(GDL_ERR_INVALID_ARG, "relation " + r->name + ": width < 1");
            sum += low[c];
        }
        if (sum != 32)
            throw Error(GDL_ERR_INVALID_ARG,
                        "relation " + r->name + ": column_bits must sum to 32");
        return layout_from_bits(a, low);
    }
    uint64_t mx[4] = {0, 0, 0, 0};
    for (tuple_size_t i = 0; i < rows; ++i)
        for (int c = 1; c < a; ++c) mx[c] |= data[i * a + c];
    int used = 0;
    for (int c = 1; c < a - 1; ++c) {
        low[c - 1] = std::max({bit_length(mx[c]), id_bits(), 1});
        used += low[c - 1];
    }
    low[a - 2] = 32 - used;
    if (low[a - 2] < std::max(bit_length(mx[a - 1]), 1))
        throw Error(GDL_ERR_UNSUPPORTED,
                    "relation " + r->name + ": columns 1.." + std::to_string(a - 1) +
                        " do not fit the 32 bits beside column 0 (set column_bits)");
    return layout_from_bits(a, low);
}

// Packs host rows of 8-byte columns into keys under the relation's layout,
// checking every value against its field.
std::vector<uint64_t> pack_rows(const column_type *data, tuple_size_t rows, const Relation *r) {
    const KeyLayout &L = r->layout;
    const int arity = L.arity;
    std::vector<uint64_t> keys(rows);
    column_type mx = 0;
    for (tuple_size_t i = 0; i < rows; ++i) {
        uint64_t k;
        if (!pack_key(L, data + i * arity, &k))
            throw Error(GDL_ERR_OVERFLOW, "load_relation " + r->name + ": row " +
                                              std::to_string(i) + " does not fit its key layout");
        keys[i] = k;
        mx |= (k >> 32) | (k & 0xffffffffull);
    }
    note_id_bits(bit_length(mx));
    return keys;
}

void check_shape(const std::string &name, int arity, tuple_size_t index_column_size,
                 int dependent_column_size) {
    if (arity < 1 || arity > 4)
        throw Error(GDL_ERR_UNSUPPORTED,
                    "relation " + name + ": arity " + std::to_string(arity) +
                        " (this engine stores arity 1 to 4)");
    if (index_column_size < 1 || index_column_size > (tuple_size_t)arity)
        throw Error(GDL_ERR_INVALID_ARG, "relation " + name + ": bad index_column_size");
    if (arity >= 3 && index_column_size != 1)
        throw Error(GDL_ERR_UNSUPPORTED,
                    "relation " + name + ": arity 3 and 4 are indexed on column 0 only");
    if (dependent_column_size < 0 || dependent_column_size >= arity ||
        (tuple_size_t)(arity - dependent_column_size) < index_column_size)
        throw Error(GDL_ERR_INVALID_ARG, "relation " + name + ": bad dependent_column_size");
}

void init_relation(Relation *target, const std::string &name, int arity,
                   tuple_size_t index_column_size, int dependent_column_size, bool tmp_flag) {
    check_shape(name, arity, index_column_size, dependent_column_size);
    target->name = name;
    target->arity = arity;
    target->index_column_size = (int)index_column_size;
    target->dependent_column_size = dependent_column_size;
    target->tmp_flag = tmp_flag;
    target->full = new GHashRelContainer(arity, (int)index_column_size, dependent_column_size);
    target->delta = new GHashRelContainer(arity, (int)index_column_size, dependent_column_size);
    target->newt = new GHashRelContainer(arity, (int)index_column_size, dependent_column_size);
    target->raw_keys = nullptr;
    target->raw_rows = 0;
    target->raw_sorted = false;
    target->raw_is_full = false;
    target->layout = choose_layout(target, nullptr, 0);
}

// full := sort + unique(raw)
void build_full_from_raw(Relation *target) {
    if (target->raw_rows == 0) return;
    uint64_t *out = nullptr;
    tuple_size_t m = sort_unique(target->raw_keys, target->raw_rows, &out);
    container_take(target->full, out, m, target->raw_rows, true);
}

}  // namespace

void load_relation(Relation *target, std::string name, int arity, column_type *data,
                   tuple_size_t data_row_size, tuple_size_t index_column_size,
                   int dependent_column_size, int grid_size, int block_size, bool tmp_flag) {
    (void)grid_size;
    (void)block_size;
    init_relation(target, name, arity, index_column_size, dependent_column_size, tmp_flag);
    if (data_row_size == 0 || !data) return;
    target->layout = choose_layout(target, data, data_row_size);
    std::vector<uint64_t> keys = pack_rows(data, data_row_size, target);
    target->raw_keys = device_alloc_keys(data_row_size);
    target->raw_rows = data_row_size;
    check(gdl_memcpy_h2d(context(), target->raw_keys, keys.data(), keys.size() * sizeof(uint64_t)),
          "load_relation H2D");
    build_full_from_raw(target);
}

void load_relation_device(Relation *target, std::string name, int arity, const uint64_t *d_keys,
                          tuple_size_t data_row_size, tuple_size_t index_column_size,
                          int dependent_column_size, int grid_size, int block_size,
                          bool tmp_flag) {
    (void)grid_size;
    (void)block_size;
    init_relation(target, name, arity, index_column_size, dependent_column_size, tmp_flag);
    if (data_row_size == 0 || !d_keys) return;
    int hi = 0, lo = 0;
    check(gdl_key_bits(context(), d_keys, data_row_size, &hi, &lo), "gdl_key_bits");
    note_id_bits(std::max(hi, lo));
    target->raw_keys = device_alloc_keys(data_row_size);
    target->raw_rows = data_row_size;
    check(gdl_memcpy_d2d(context(), target->raw_keys, d_keys, data_row_size * sizeof(uint64_t)),
          "load_relation_device D2D");
    build_full_from_raw(target);
}

void load_relation_container(GHashRelContainer *target, int arity, column_type *data,
                             tuple_size_t data_row_size, tuple_size_t index_column_size,
                             int dependent_column_size, float index_map_load_factor,
                             int grid_size, int block_size, float *detail_time,
                             bool gpu_data_flag, bool sorted_flag, bool build_index_flag,
                             bool tuples_array_flag) {
    (void)grid_size;
    (void)block_size;
    (void)detail_time;
    (void)tuples_array_flag;
    check_shape("container", arity, index_column_size, dependent_column_size);
    container_free(target);
    target->arity = arity;
    target->index_column_size = index_column_size;
    target->dependent_column_size = dependent_column_size;
    target->index_map_load_factor = index_map_load_factor;
    if (data_row_size == 0 || !data) return;
    uint64_t *raw = device_alloc_keys(data_row_size);
    if (gpu_data_flag) {
        // device data: packed u64 keys
        check(gdl_memcpy_d2d(context(), raw, data, data_row_size * sizeof(uint64_t)),
              "load_relation_container D2D");
        int hi = 0, lo = 0;
        check(gdl_key_bits(context(), raw, data_row_size, &hi, &lo), "gdl_key_bits");
        note_id_bits(std::max(hi, lo));
    } else {
        // a container has no relation: arity 1-2 as (c0 << 32) | c1, arity
        // 3-4 with the layout of an unloaded relation of that arity
        Relation shape;
        shape.name = "container";
        shape.arity = arity;
        shape.layout = choose_layout(&shape, data, data_row_size);
        std::vector<uint64_t> keys = pack_rows(data, data_row_size, &shape);
        check(gdl_memcpy_h2d(context(), raw, keys.data(), keys.size() * sizeof(uint64_t)),
              "load_relation_container H2D");
    }
    if (sorted_flag) {
        container_take(target, raw, data_row_size, data_row_size, true);
    } else {
        uint64_t *out = nullptr;
        tuple_size_t m = sort_unique(raw, data_row_size, &out);
        device_free(raw);
        container_take(target, out, m, data_row_size, true);
    }
    if (build_index_flag) {
        float old = index_map_load_factor;
        if (old > 0.f && old <= 0.95f) check(gdl_set_load_factor(context(), old), "load factor");
        container_index(target, (int)index_column_size);
    }
}

void free_relation_container(GHashRelContainer *target) { container_free(target); }

void copy_relation_container(GHashRelContainer *dst, GHashRelContainer *src, int grid_size,
                             int block_size) {
    (void)grid_size;
    (void)block_size;
    container_free(dst);
    dst->arity = src->arity;
    dst->index_column_size = src->index_column_size;
    dst->dependent_column_size = src->dependent_column_size;
    if (src->tuple_counts == 0) return;
    uint64_t *k = device_alloc_keys(src->tuple_counts);
    check(gdl_memcpy_d2d(context(), k, src->keys, src->tuple_counts * sizeof(uint64_t)),
          "copy_relation_container");
    container_take(dst, k, src->tuple_counts, src->tuple_counts, src->sorted);
}

void free_relation(Relation *target) {
    for (GHashRelContainer *c : {target->full, target->delta, target->newt}) {
        if (!c) continue;
        container_free(c);
        delete c;
    }
    target->full = target->delta = target->newt = nullptr;
    device_free(target->raw_keys);
    target->raw_keys = nullptr;
    target->raw_rows = 0;
    target->raw_is_full = false;
}


namespace gdlog {
void raw_rows_from_full(Relation *r) {
    device_free(r->raw_keys);
    r->raw_keys = nullptr;
    r->raw_rows = r->full->keys ? r->full->tuple_counts : 0;
    r->raw_sorted = r->full->sorted;
    r->raw_is_full = r->raw_rows > 0;
}

void materialize_raw(Relation *r) {
    if (!r->raw_is_full) return;
    const tuple_size_t n = r->raw_rows;
    uint64_t *k = device_alloc_keys(n);
    gdl_status st = gdl_memcpy_d2d(context(), k, r->full->keys, n * sizeof(uint64_t));
    if (st != GDL_OK) device_free(k);
    check(st, "raw rows from FULL");
    r->raw_keys = k;
    r->raw_is_full = false;
}

const uint64_t *raw_rows_ptr(const Relation *r) {
    return r->raw_is_full ? r->full->keys : r->raw_keys;
}
}  // namespace gdlog

void Relation::flush_delta(int grid_size, int block_size, float *detail_time) {
    (void)grid_size;
    (void)block_size;
    (void)detail_time;
    if (delta->tuple_counts == 0) return;
    materialize_raw(this);  // FULL changes below
    container_normalize(delta);
    const tuple_size_t nf = current_full_size, nd = delta->tuple_counts;
    uint64_t *merged = device_alloc_keys(nf + nd);
    uint64_t nm = 0;
    check(gdl_merge_union_u64(context(), full->keys, nf, delta->keys, nd, merged, &nm),
          "flush_delta merge");
    container_take(full, merged, nm, nf + nd, true);
    current_full_size = nm;
}

uint64_t relation_checksum(GHashRelContainer *c) {
    uint64_t s = 0;
    check(gdl_checksum_u64(context(), c->keys, c->tuple_counts, &s), "relation checksum");
    return s;
}

void unpack_key(const KeyLayout &L, uint64_t key, column_type *cols) {
    for (int c = 0; c < L.arity; ++c)
        cols[c] = (key >> L.shift[c]) & (L.bits[c] >= 64 ? ~0ull : ((1ull << L.bits[c]) - 1));
}

bool pack_key(const KeyLayout &L, const column_type *cols, uint64_t *key) {
    uint64_t k = 0;
    for (int c = 0; c < L.arity; ++c) {
        if (L.bits[c] < 64 && (cols[c] >> L.bits[c]) != 0) return false;
        k |= (uint64_t)cols[c] << L.shift[c];
    }
    *key = k;
    return true;
}

std::vector<uint64_t> relation_keys(GHashRelContainer *c) {
    std::vector<uint64_t> out(c->tuple_counts);
    if (c->tuple_counts)
        check(gdl_memcpy_d2h(context(), out.data(), c->keys, c->tuple_counts * sizeof(uint64_t)),
              "relation_keys D2H");
    if (!c->sorted) std::sort(out.begin(), out.end());
    return out;
}
