// plen.hip -- all-pairs hop counts (the reference's path_length program,
// test/path_length.cu:49-124) on the gdlog C++ API.
//
// The reference keeps path(from, to, len) with len a dependent column:
//   path(a, c, l + 1) :- edge(a, b), path(b, c, l)      edges load with l = 1
// joined like TC (edge_2__2_1 FULL inner, path DELTA outer) and deduplicated
// on (from, to) only, so a pair keeps the len of the iteration that first
// derives it.  Every tuple of one iteration's delta has the same len (the
// delta of iteration i holds len i + 1), so len = 1 + the append-log segment
// of the pair: the loaded edges, then each iteration's delta.  The program
// runs the TC encoding with FULL kept as the append log
// (LIE::keep_full_log), labels the log by segment on the device and sorts it
// with the labels carried in the ids' spare low bits
// (gdl_segment_labels_sorted), or, when the ids leave too few bits, sorts the
// keys and carries each label to its sorted position by binary search
// (gdl_segment_labels_u32 / gdl_scatter_by_key_u32).
#include "gdlog/exception.h"
#include "programs.h"

using namespace gdlog;

void path_length_program(const EdgeInput &graph, int grid_size, int block_size,
                         ProgramOutput *out, std::vector<uint32_t> *lens) {
    double t0 = now_ms();
    Relation *edge_2__2_1 = new Relation();
    Relation *path_2__1_2 = new Relation();
    path_2__1_2->index_flag = false;
    graph.load(path_2__1_2, "path_3__1_2_3", false, grid_size, block_size);
    graph.load(edge_2__2_1, "edge_2__2_1", true, grid_size, block_size);
    const tuple_size_t loaded = path_2__1_2->full->tuple_counts;

    LIE plen_scc(grid_size, block_size);
    plen_scc.reload_full_flag = false;
    plen_scc.keep_full_log = true;
    plen_scc.add_relations(edge_2__2_1, true);
    plen_scc.add_relations(path_2__1_2, false);
    // newt = (inner[1], outer[1]); the hop count is the log segment
    RelationalJoin join(edge_2__2_1, FULL, path_2__1_2, DELTA, path_2__1_2,
                        gdl_colmap{2, {GDL_COL_INNER1, GDL_COL_OUTER1}},
                        gdl_pred{GDL_PRED_NONE, 0, 0});
    join.empty_inner_zeroes_outer = true;  // the reference's hooks keep it
    plen_scc.add_ra(join);
    double t1 = now_ms();
    plen_scc.fixpoint_loop();

    // segment bounds of the log: the loaded edges, then each iteration's delta
    GHashRelContainer *full = path_2__1_2->full;
    const tuple_size_t n = full->tuple_counts;
    std::vector<uint64_t> bounds{0, loaded};
    for (const IterationRecord &r : plen_scc.trace)
        if (r.delta) bounds.push_back(bounds.back() + r.delta);
    if (bounds.back() != n)
        throw Error(GDL_ERR_DEVICE, "path_length: the log holds " + std::to_string(n) +
                                        " tuples, its segments " +
                                        std::to_string(bounds.back()));
    gdl_ctx *ctx = context();
    uint64_t *d_bounds = device_alloc_keys(bounds.size());
    uint64_t *sorted = nullptr;
    uint32_t *labels = nullptr, *sorted_labels = nullptr;
    std::vector<uint32_t> host_lens;
    try {
        check(gdl_memcpy_h2d(ctx, d_bounds, bounds.data(), bounds.size() * sizeof(uint64_t)),
              "segment bounds");
        if (n) {
            const uint32_t n_seg = (uint32_t)(bounds.size() - 1);
            sorted_labels = reinterpret_cast<uint32_t *>(device_alloc_keys((n + 1) / 2));
            sorted = device_alloc_keys(n);
            // the labels ride through the sort in the ids' spare bits when
            // there are enough; otherwise label, sort, and binary-search each
            // key's sorted position
            gdl_status st = gdl_segment_labels_sorted(ctx, full->keys, n, d_bounds, n_seg, 1,
                                                      id_bits(), sorted, sorted_labels);
            if (st == GDL_ERR_UNSUPPORTED) {
                labels = reinterpret_cast<uint32_t *>(device_alloc_keys((n + 1) / 2));
                check(gdl_segment_labels_u32(ctx, d_bounds, n_seg, 1, labels, n),
                      "gdl_segment_labels_u32");
                const int b = id_bits();
                check(gdl_sort_u64(ctx, full->keys, n, b, b, sorted), "gdl_sort_u64");
                check(gdl_scatter_by_key_u32(ctx, sorted, n, full->keys, labels, n, sorted_labels),
                      "gdl_scatter_by_key_u32");
            } else {
                check(st, "gdl_segment_labels_sorted");
            }
            if (lens) {
                host_lens.resize(n);
                check(gdl_memcpy_d2h(ctx, host_lens.data(), sorted_labels, n * sizeof(uint32_t)),
                      "lens D2H");
            }
            container_take(full, sorted, n, n, true);
            sorted = nullptr;
        }
    } catch (...) {
        device_free(d_bounds);
        device_free(labels);
        device_free(sorted_labels);
        device_free(sorted);
        throw;
    }
    device_free(d_bounds);
    device_free(labels);
    device_free(sorted_labels);
    double t2 = now_ms();

    out->load_ms = t1 - t0;
    out->fixpoint_ms = t2 - t1;
    out->iterations = plen_scc.iteration_count;
    out->trace = plen_scc.trace;
    out->derived = sum_delta(plen_scc.trace);
    out->n_main = n;
    out->checksum = relation_checksum(full);
    out->sizes[0] = bounds.size() - 1;  // distinct hop counts: the longest is this many
    if (out->want_keys) out->main_keys = relation_keys(full);
    if (lens) *lens = std::move(host_lens);
    free_relation(edge_2__2_1);
    free_relation(path_2__1_2);
    delete edge_2__2_1;
    delete path_2__1_2;
}
