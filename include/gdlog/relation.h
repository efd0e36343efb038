// gdlog/relation.h -- relation storage of the gdlog C++ API.
//
// Mirrors include/relation.cuh:211-318 of the reference: RelationVersion,
// GHashRelContainer (one version of a relation), Relation {full, delta, newt}
// and load_relation.  The storage is MI355X-native: a version is one
// contiguous device array of u64 keys (sorted and duplicate-free except for
// the raw versions noted below) plus, when the version is used as the inner
// side of a join, an open-addressing hash index over its distinct join
// prefixes, built lazily.  There are no tuple-pointer arrays.
#pragma once

#include <string>
#include <vector>

#include "gdl_hip.h"
#include "gdlog/tuple.h"

enum RelationVersion { DELTA, FULL, NEWT };

// Bit layout of a relation's u64 tuple keys.  Column 0 holds the top 32 bits;
// columns 1..arity-1 share the low 32 bits (arity 2: column 1 alone), each in
// its own field in column order, so integer order is the reference's
// lexicographic order over all columns (tuple.cuh:94-127) and the one-column
// join prefix is key >> 32 for every arity.  Arity 3 and 4 (up to 4 columns;
// the reference allows 20, tuple.cuh:5-10) split the low 32 bits by the
// widths chosen when the relation is loaded.
struct KeyLayout {
    int arity = 2;
    int bits[4] = {32, 32, 0, 0};   // field widths
    int shift[4] = {32, 0, 0, 0};   // field offsets
};
enum JoinDirection { LEFT, RIGHT };
enum MonotonicOrder { DESC, ASC, UNSPEC };

struct GHashRelContainer {
    // fields read by programs, same names as the reference (relation.cuh:241-271)
    tuple_size_t index_map_size = 0;   // slots of the current index (0 if none)
    float index_map_load_factor = 0.8f;
    tuple_size_t index_column_size;
    int dependent_column_size = 0;
    tuple_size_t tuple_counts = 0;     // logical size of this version
    tuple_size_t data_raw_row_size = 0;
    int arity;
    bool tmp_flag = false;

    // device storage
    uint64_t *keys = nullptr;          // device array of tuple keys
    tuple_size_t capacity = 0;         // keys allocated
    bool sorted = true;                // false: rows in production order, may repeat
    gdl_index index{};                 // valid when index.slots != nullptr
    int index_prefix_cols = 0;         // prefix the index was built on

    GHashRelContainer(int arity, int indexed_column_size, int dependent_column_size,
                      bool tmp_flag = false)
        : index_column_size(indexed_column_size), dependent_column_size(dependent_column_size),
          arity(arity), tmp_flag(tmp_flag) {}
};

struct Relation {
    int arity = 0;
    int index_column_size = 1;
    std::string name;
    int dependent_column_size = 0;
    // Key layout (set by load_relation).  For arity 3 and 4, column_bits may
    // be set before loading to choose the widths of columns 1..arity-1 (they
    // must sum to 32); otherwise every column but the last gets the bit
    // length of the largest id loaded so far (at least that of its own loaded
    // values) and the last column the remaining bits.
    KeyLayout layout;
    std::vector<int> column_bits;
    bool index_flag = true;
    bool tmp_flag = false;

    GHashRelContainer *delta = nullptr;
    GHashRelContainer *newt = nullptr;
    GHashRelContainer *full = nullptr;

    // Relation::current_full_size (relation.cuh:275): the set-difference base.
    tuple_size_t current_full_size = 0;

    // Rows as loaded, in input order (the reference's full->data_raw).  The
    // fixpoint seeds delta from its first full->tuple_counts rows, as
    // copy_relation_container does (relation.cu:764-795).
    uint64_t *raw_keys = nullptr;
    tuple_size_t raw_rows = 0;
    bool raw_sorted = false;
    // The raw rows are FULL's keys (raw_keys unused): set where the rows are
    // reloaded from FULL (lie.cu:219-254) instead of copying them; whatever
    // is about to change FULL first makes the copy (materialize_raw).
    bool raw_is_full = false;

    MonotonicOrder monotonic_order = MonotonicOrder::DESC;

    // Set by LIE::fixpoint_loop for an update relation whose FULL is never a
    // join inner in that LIE: FULL is then an append log (sorted at the end of
    // the fixpoint) plus a device hash set, and newt stays unsorted.
    bool append_mode = false;
    gdl_hashset full_set{};
    // Append-mode relations that no operator of the LIE reads as FULL or NEWT
    // and that only joins write: join results go straight into full_set
    // (gdl_join_insert) and the new tuples collect in pend_keys, which becomes
    // the next delta.
    bool fused_output = false;
    // A LIE tmp relation loaded without tmp_flag (its newt is a set, dropped
    // after every iteration) whose writers are all column-map joins: joins
    // insert into tmp_set (a bitmap or hash set, cleared every iteration) and
    // append the new keys to newt, instead of sorting and merging.
    bool tmp_dedup = false;
    gdl_hashset tmp_set{};
    uint64_t *pend_keys = nullptr;
    uint64_t pend_cap = 0, pend_n = 0, pend_raw = 0;
    // Append mode after bit-matrix bodies that logged only their last body's
    // tuples: FULL's log holds part of FULL (current_full_size counts all of
    // it) and full_set all of it; the end of the fixpoint reads FULL from the
    // set.
    bool log_stale = false;

    // Merges DELTA into FULL (relation.cu:405-521).
    void flush_delta(int grid_size, int block_size, float *detail_time);
};

// load_relation (relation.cu:815-838): creates full/delta/newt and loads
// `data_row_size` rows of `arity` host columns into FULL (sort + unique).
// grid_size / block_size are accepted for API compatibility; launch shapes are
// chosen by the engine.
void load_relation(Relation *target, std::string name, int arity, column_type *data,
                   tuple_size_t data_row_size, tuple_size_t index_column_size,
                   int dependent_column_size, int grid_size, int block_size,
                   bool tmp_flag = false);

// Same, from device-resident packed keys ((c0 << 32) | c1 per row).
void load_relation_device(Relation *target, std::string name, int arity, const uint64_t *d_keys,
                          tuple_size_t data_row_size, tuple_size_t index_column_size,
                          int dependent_column_size, int grid_size, int block_size,
                          bool tmp_flag = false);

// load_relation_container (relation.cu:523-632) for host rows: loads `rows`
// rows into `target` (sort + unique unless sorted_flag), building the index
// when build_index_flag.
void load_relation_container(GHashRelContainer *target, int arity, column_type *data,
                             tuple_size_t data_row_size, tuple_size_t index_column_size,
                             int dependent_column_size, float index_map_load_factor,
                             int grid_size, int block_size, float *detail_time,
                             bool gpu_data_flag = false, bool sorted_flag = false,
                             bool build_index_flag = true, bool tuples_array_flag = true);

// free_relation_container (relation.cu:797-813).
void free_relation_container(GHashRelContainer *target);

// copy_relation_container (relation.cu:764-795): dst := copy of src.
void copy_relation_container(GHashRelContainer *dst, GHashRelContainer *src, int grid_size,
                             int block_size);

// Frees a Relation's three versions and its raw rows (the reference leaks them).
void free_relation(Relation *target);

// Copies a version's tuples to the host as sorted keys ((c0 << 32) | c1;
// arity 3 and 4: the relation's KeyLayout).
std::vector<uint64_t> relation_keys(GHashRelContainer *c);
// Columns of a key under a layout (cols has layout.arity entries), and back
// (false, with *key untouched, when a value does not fit its field).
void unpack_key(const KeyLayout &layout, uint64_t key, column_type *cols);
bool pack_key(const KeyLayout &layout, const column_type *cols, uint64_t *key);
// Order-independent checksum of a version's tuples (gdl_checksum_u64).
uint64_t relation_checksum(GHashRelContainer *c);

namespace gdlog {

// Engine context used by the C++ API (created on first use on the current
// HIP device unless set explicitly).
gdl_ctx *context();
gdl_ctx *context_or_null();
void set_context(gdl_ctx *ctx);
// Upper bound of the bit length of every id loaded so far (sort key bounds).
int id_bits();
void note_id_bits(int bits);
void reset_id_bits();
// Bracket a running fixpoint: the engine context's id bound
// (gdl_set_id_bits) holds id_bits() inside and 0 outside.
void enter_lie_scope();
void leave_lie_scope();

// Internal helpers shared by the RA operators and the fixpoint loop.
void container_free(GHashRelContainer *c);
void container_take(GHashRelContainer *c, uint64_t *keys, tuple_size_t n, tuple_size_t cap,
                    bool sorted);
// Makes `c` sorted and duplicate-free (no-op if already).
void container_normalize(GHashRelContainer *c);
// Builds (or reuses) the hash index of `c` on its first `prefix_cols` columns.
const gdl_index *container_index(GHashRelContainer *c, int prefix_cols);
void container_drop_index(GHashRelContainer *c);
uint64_t *device_alloc_keys(tuple_size_t n);
void device_free(void *p);
// The raw rows become FULL's keys (see Relation::raw_is_full).
void raw_rows_from_full(Relation *r);
// Gives the raw rows their own copy when they are FULL's keys.
void materialize_raw(Relation *r);
// The raw rows (FULL's keys when raw_is_full).
const uint64_t *raw_rows_ptr(const Relation *r);
// Sorted duplicate-free copy of `n` keys (does not consume `d_in`).
tuple_size_t sort_unique(const uint64_t *d_in, tuple_size_t n, uint64_t **d_out);

// Verbose reference-style log lines (gdl_config.verbose or GDLOG_VERBOSE=1).
bool verbose();

// Phase timing accumulated into detail_time[] (off by default: it
// synchronises the stream around every phase).
void set_phase_timing(bool on);
bool phase_timing();

// How the fixpoint keeps FULL of update relations.  Auto: append log + hash
// set for relations whose FULL is not a join inner in the LIE, sorted merge
// for the others.  Merge: sorted merge for every relation (the reference's
// structure: set_difference then flush_delta's merge every iteration).
enum class FullMode { Auto = 0, Merge = 1 };
void set_full_mode(FullMode m);
FullMode full_mode();

// Appends n device keys to a version's array (growing it), unsorted.
void container_append(GHashRelContainer *c, const uint64_t *keys, tuple_size_t n);

}  // namespace gdlog
