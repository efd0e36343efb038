// runner.cpp -- C ABI of the programs (include/gdl_programs.h) and the
// shared input/output plumbing.
#include <chrono>
#include <cstring>
#include <string>

#include "gdl_programs.h"
#include "gdlog/exception.h"
#include "programs.h"

using namespace gdlog;

double now_ms() {
    check(gdl_sync(context()), "gdl_sync");
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint64_t sum_delta(const std::vector<IterationRecord> &t) {
    uint64_t s = 0;
    for (const auto &r : t) s += r.delta;
    return s;
}

void EdgeInput::load(Relation *r, const std::string &name, bool reversed, int grid_size,
                     int block_size, bool tmp_flag) const {
    if (dev) {
        if (!reversed) {
            load_relation_device(r, name, 2, dev, rows, 1, 0, grid_size, block_size, tmp_flag);
            return;
        }
        uint64_t *rev = device_alloc_keys(rows);
        gdl_colmap m{2, {GDL_COL_INNER1, GDL_COL_INNER0}};
        check(gdl_project_u64(context(), dev, rows, m, rev), "reverse edges");
        load_relation_device(r, name, 2, rev, rows, 1, 0, grid_size, block_size, tmp_flag);
        device_free(rev);
        return;
    }
    if (!reversed) {
        load_relation(r, name, 2, const_cast<column_type *>(host), rows, 1, 0, grid_size,
                      block_size, tmp_flag);
        return;
    }
    std::vector<column_type> rev(rows * 2);
    for (tuple_size_t i = 0; i < rows; ++i) {
        rev[i * 2 + 1] = host[i * 2];
        rev[i * 2] = host[i * 2 + 1];
    }
    load_relation(r, name, 2, rev.data(), rows, 1, 0, grid_size, block_size, tmp_flag);
}

namespace {

thread_local std::string g_last_error;

// launch shape the reference programs pass (test/tc.cu:224-225); the engine
// chooses its own, these are only carried through the API.
constexpr int kBlock = 512;
int grid_for_ctx(gdl_ctx *) { return 32 * 256; }

struct HostRows {
    std::vector<column_type> rows;
    EdgeInput input(const uint64_t *keys, uint64_t n, int on_device) {
        EdgeInput in;
        in.rows = n;
        if (on_device) {
            in.dev = keys;
            return in;
        }
        rows.resize(n * 2);
        for (uint64_t i = 0; i < n; ++i) {
            rows[i * 2] = (column_type)(keys[i] >> 32);
            rows[i * 2 + 1] = (column_type)(keys[i] & 0xffffffffull);
        }
        in.host = rows.data();
        return in;
    }
};

gdl_status copy_out(const std::vector<uint64_t> &src, uint64_t *dst, uint64_t cap) {
    if (!dst) return GDL_OK;
    if (src.size() > cap) return GDL_ERR_OVERFLOW;
    if (!src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(uint64_t));
    return GDL_OK;
}

gdl_status copy_trace(const std::vector<IterationRecord> &src, gdl_trace_rec *dst, uint64_t cap) {
    if (!dst) return GDL_OK;
    if (src.size() > cap) return GDL_ERR_OVERFLOW;
    for (size_t i = 0; i < src.size(); ++i)
        dst[i] = {src[i].iteration, src[i].relation, src[i].newt, src[i].delta, src[i].full};
    return GDL_OK;
}

template <typename Body>
gdl_status run_program(gdl_ctx *ctx, gdl_prog_result *res, const gdl_prog_buffers *bufs,
                       Body body) {
    if (!ctx || !res) return GDL_ERR_INVALID_ARG;
    g_last_error.clear();
    gdl_ctx *prev = context_or_null();
    set_context(ctx);
    reset_id_bits();
    reset_checkpoint_sequence();
    ProgramOutput out;
    out.want_keys = bufs && (bufs->out || bufs->extra0 || bufs->extra1);
    gdl_status st = GDL_OK;
    try {
        body(&out);
    } catch (const Error &e) {
        g_last_error = e.what();
        st = e.status;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        st = GDL_ERR_INVALID_ARG;
    }
    set_context(prev);
    if (st != GDL_OK) return st;
    std::memset(res, 0, sizeof(*res));
    res->iterations = out.iterations;
    res->init_iterations = out.init_iterations;
    res->derived = out.derived;
    res->n_out = out.n_main;
    for (int i = 0; i < 4; ++i) res->sizes[i] = out.sizes[i];
    res->load_ms = out.load_ms;
    res->fixpoint_ms = out.fixpoint_ms;
    res->n_trace = out.trace.size();
    res->n_trace_init = out.trace_init.size();
    res->checksum = out.checksum;
    if (bufs) {
        gdl_status s[5] = {copy_out(out.main_keys, bufs->out, bufs->out_cap),
                           copy_out(out.extra_keys[0], bufs->extra0, bufs->extra0_cap),
                           copy_out(out.extra_keys[1], bufs->extra1, bufs->extra1_cap),
                           copy_trace(out.trace, bufs->trace, bufs->trace_cap),
                           copy_trace(out.trace_init, bufs->trace_init, bufs->trace_init_cap)};
        for (gdl_status x : s)
            if (x != GDL_OK) {
                g_last_error = "an output buffer is too small";
                return x;
            }
    }
    return GDL_OK;
}

}  // namespace

gdl_status run_program_host(gdl_ctx *ctx, int kind, const uint64_t *a, uint64_t na,
                            const uint64_t *b, uint64_t nb, ProgramOutput *out,
                            std::string *err) {
    if (!ctx || !out || (na && !a) || (nb && !b)) return GDL_ERR_INVALID_ARG;
    gdl_ctx *prev = context_or_null();
    set_context(ctx);
    reset_id_bits();
    reset_checkpoint_sequence();
    gdl_status st = GDL_OK;
    try {
        HostRows ha, hb;
        EdgeInput ia = ha.input(a, na, 0);
        if (kind == 0) {
            tc_program(ia, -1, grid_for_ctx(ctx), kBlock, out);
        } else if (kind == 1) {
            sg_program(ia, grid_for_ctx(ctx), kBlock, out);
        } else if (kind == 2) {
            EdgeInput ib = hb.input(b, nb, 0);
            cspa_program(ia, ib, grid_for_ctx(ctx), kBlock, out);
        } else {
            st = GDL_ERR_INVALID_ARG;
        }
    } catch (const Error &e) {
        if (err) *err = e.what();
        st = e.status;
    } catch (const std::exception &e) {
        if (err) *err = e.what();
        st = GDL_ERR_INVALID_ARG;
    }
    set_context(prev);
    return st;
}

extern "C" {

gdl_status gdl_prog_tc(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                       int max_iteration, gdl_prog_result *res, const gdl_prog_buffers *bufs) {
    if (n && !edges) return GDL_ERR_INVALID_ARG;
    HostRows hr;
    EdgeInput in = hr.input(edges, n, on_device);
    return run_program(ctx, res, bufs, [&](ProgramOutput *out) {
        tc_program(in, max_iteration, grid_for_ctx(ctx), kBlock, out);
    });
}

gdl_status gdl_prog_tc_hooks(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                             int variant, gdl_prog_result *res, const gdl_prog_buffers *bufs) {
    if ((n && !edges) || variant < 0 || variant > 2) return GDL_ERR_INVALID_ARG;
    HostRows hr;
    EdgeInput in = hr.input(edges, n, on_device);
    return run_program(ctx, res, bufs, [&](ProgramOutput *out) {
        tc_hooks_program(in, variant, grid_for_ctx(ctx), kBlock, out);
    });
}

gdl_status gdl_prog_sg(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                       gdl_prog_result *res, const gdl_prog_buffers *bufs) {
    if (n && !edges) return GDL_ERR_INVALID_ARG;
    HostRows hr;
    EdgeInput in = hr.input(edges, n, on_device);
    return run_program(ctx, res, bufs, [&](ProgramOutput *out) {
        sg_program(in, grid_for_ctx(ctx), kBlock, out);
    });
}

gdl_status gdl_prog_cspa(gdl_ctx *ctx, const uint64_t *assign, uint64_t n_assign,
                         const uint64_t *dereference, uint64_t n_dereference, int on_device,
                         gdl_prog_result *res, const gdl_prog_buffers *bufs) {
    if ((n_assign && !assign) || (n_dereference && !dereference)) return GDL_ERR_INVALID_ARG;
    HostRows ha, hd;
    EdgeInput ia = ha.input(assign, n_assign, on_device);
    EdgeInput id = hd.input(dereference, n_dereference, on_device);
    return run_program(ctx, res, bufs, [&](ProgramOutput *out) {
        cspa_program(ia, id, grid_for_ctx(ctx), kBlock, out);
    });
}

gdl_status gdl_prog_path_length(gdl_ctx *ctx, const uint64_t *edges, uint64_t n, int on_device,
                                gdl_prog_result *res, uint64_t *keys, uint32_t *lens,
                                uint64_t cap) {
    if (n && !edges) return GDL_ERR_INVALID_ARG;
    HostRows hr;
    EdgeInput in = hr.input(edges, n, on_device);
    std::vector<uint32_t> l;
    std::vector<uint64_t> k;
    gdl_status st = run_program(ctx, res, nullptr, [&](ProgramOutput *out) {
        out->want_keys = keys != nullptr;
        path_length_program(in, grid_for_ctx(ctx), kBlock, out, lens ? &l : nullptr);
        k = std::move(out->main_keys);
    });
    if (st != GDL_OK) return st;
    if ((keys || lens) && res->n_out > cap) {
        g_last_error = "an output buffer is too small";
        return GDL_ERR_OVERFLOW;
    }
    if (keys && !k.empty()) std::memcpy(keys, k.data(), k.size() * sizeof(uint64_t));
    if (lens && !l.empty()) std::memcpy(lens, l.data(), l.size() * sizeof(uint32_t));
    return GDL_OK;
}

gdl_status gdl_prog_path_length_hooks(gdl_ctx *ctx, const uint64_t *edges, uint64_t n,
                                      int on_device, gdl_prog_result *res, uint64_t *keys,
                                      uint32_t *lens, uint64_t cap) {
    if (n && !edges) return GDL_ERR_INVALID_ARG;
    HostRows hr;
    EdgeInput in = hr.input(edges, n, on_device);
    std::vector<uint32_t> l;
    std::vector<uint64_t> k;
    gdl_status st = run_program(ctx, res, nullptr, [&](ProgramOutput *out) {
        out->want_keys = keys != nullptr;
        path_length_hooks_program(in, grid_for_ctx(ctx), kBlock, out, lens ? &l : nullptr);
        k = std::move(out->main_keys);
    });
    if (st != GDL_OK) return st;
    if ((keys || lens) && res->n_out > cap) {
        g_last_error = "an output buffer is too small";
        return GDL_ERR_OVERFLOW;
    }
    if (keys && !k.empty()) std::memcpy(keys, k.data(), k.size() * sizeof(uint64_t));
    if (lens && !l.empty()) std::memcpy(lens, l.data(), l.size() * sizeof(uint32_t));
    return GDL_OK;
}

const char *gdl_prog_last_error(void) { return g_last_error.c_str(); }

gdl_status gdl_prog_set_checkpoint(const char *dir, int every, int resume, int stop_after) {
    if (every < 0 || (resume != 0 && resume != 1)) return GDL_ERR_INVALID_ARG;
    set_checkpoint_defaults(dir, every, resume != 0, stop_after);
    return GDL_OK;
}

gdl_status gdl_prog_set_full_mode(int mode) {
    if (mode != 0 && mode != 1) return GDL_ERR_INVALID_ARG;
    set_full_mode(mode ? FullMode::Merge : FullMode::Auto);
    return GDL_OK;
}

}  // extern "C"
