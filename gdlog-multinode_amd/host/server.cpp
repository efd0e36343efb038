// server.cpp -- the query server of include/gdl_serve.h.
//
// One mutex guards the queues and the memory account.  Admission runs
// whenever a query is submitted or finishes: pending queries are taken in
// submission order when their estimate fits the budget beside the running
// ones and a place is free (a batch slot for TC, else an idle worker); a
// query that does not fit may be passed by later ones at most kMaxPasses
// times, then it blocks the queue until it fits.  A query that ran out of
// memory (its worker's cap, or the device) is rerun alone: it waits until
// nothing runs and nothing is admitted beside it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gdl_batch.h"
#include "gdl_serve.h"
#include "../programs/programs.h"

namespace {

constexpr int kMaxPasses = 8;
constexpr uint64_t kMinReservation = 64ull << 20;
constexpr uint64_t kBatchMaxTuples = 8ull << 20;  // larger TC queries run on a worker
constexpr double kBatchSliceMs = 20.0;             // take_turns: the batch's turn

// Executors of a server take turns on the device (take_turns): one at a time,
// in the order they asked (tickets), so neither the batch nor a worker can
// hold it while the other waits for more than its own turn.
class DeviceTurn {
  public:
    void acquire() {
        std::unique_lock<std::mutex> lk(m_);
        const uint64_t t = next_++;
        cv_.wait(lk, [&] { return serving_ == t; });
    }
    void release() {
        {
            std::lock_guard<std::mutex> lk(m_);
            ++serving_;
        }
        cv_.notify_all();
    }

  private:
    std::mutex m_;
    std::condition_variable cv_;
    uint64_t next_ = 0, serving_ = 0;
};

double clock_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

uint64_t max_id(const uint64_t *k, uint64_t n) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i)
        m = std::max<uint64_t>({m, k[i] >> 32, k[i] & 0xffffffffull});
    return m;
}

uint64_t pow2_ceil(uint64_t x) {
    uint64_t c = 1024;
    while (c < x) c <<= 1;
    return c;
}

// Bytes of a hash set holding `tuples` keys at the default load factor
// (power-of-two slots plus the flag word).
uint64_t hash_set_bytes(uint64_t tuples) {
    return 8 * (pow2_ceil((uint64_t)std::ceil((double)tuples / GDL_DEFAULT_LOAD_FACTOR)) + 1);
}

// Tuple count of a query's output relations, estimated from its inputs.
uint64_t estimate_tuples(int kind, const uint64_t *a, uint64_t na, const uint64_t *b,
                         uint64_t nb, uint32_t *n_nodes) {
    if (na == 0) {
        *n_nodes = 1;
        return 0;
    }
    const uint64_t ids = std::max(max_id(a, na), nb ? max_id(b, nb) : 0) + 1;
    *n_nodes = (uint32_t)std::min<uint64_t>(ids, 0xffffffffull);
    const double nn = (double)ids * (double)ids;
    if (kind == GDL_QUERY_TC) return gdl_tc_estimate(a, na, (uint32_t)ids, 8, 0x5eed);
    if (kind == GDL_QUERY_SG) {
        // generation pairs of the first rule (children of one parent), grown
        std::map<uint64_t, uint64_t> deg;
        for (uint64_t i = 0; i < na; ++i) ++deg[a[i] >> 32];
        double pairs = 0;
        for (auto &d : deg) pairs += (double)d.second * (double)d.second;
        return (uint64_t)std::min(nn, 40.0 * pairs + (double)na);
    }
    // CSPA: ValueFlow, MemoryAlias and ValueAlias, each at most ids^2
    return (uint64_t)std::min(3.0 * nn, 300.0 * (double)(na + nb) * 3.0);
}

// Device bytes for a query of `tuples` output tuples: the append log and
// hash set of the output relations, the delta / newt buffers and the join
// outputs of an iteration.
uint64_t estimate_bytes(int kind, uint64_t tuples, uint64_t na, uint64_t nb) {
    const uint64_t hs = hash_set_bytes(tuples);
    const uint64_t per = kind == GDL_QUERY_TC ? 24 : 64;  // SG / CSPA: materialized joins
    return hs + per * tuples + 64 * (na + nb) + kMinReservation;
}

struct Query {
    uint64_t qid = 0;
    int kind = 0;
    std::vector<uint64_t> a, b;
    uint32_t n_nodes = 0;
    uint64_t est_tuples = 0;
    uint64_t est = 0;
    bool alone = false;
    int passes = 0;
    bool to_batch = false;
    gdl_query_result r{};
    std::vector<uint64_t> keys;
    std::vector<gdl_trace_rec> trace;
};

using QueryPtr = std::unique_ptr<Query>;

}  // namespace

struct gdl_server {
    gdl_server_config cfg{};
    std::mutex mu;
    DeviceTurn turn;  // take_turns: one executor on the device at a time
    std::condition_variable cv_work, cv_batch, cv_done;
    std::deque<QueryPtr> pending, work, batch_in, done;
    std::map<uint64_t, QueryPtr> popped;
    int idle_workers = 0;
    int running = 0;
    uint32_t batch_count = 0;  // queries handed to the batch and not finished
    bool alone_running = false;
    uint64_t reserved = 0, peak = 0, budget = 0;
    uint64_t submitted = 0, finished = 0, n_batched = 0, reruns = 0;
    bool stop = false;
    std::vector<std::thread> threads;
    double t0 = 0;
    gdl_status init_status = GDL_OK;
    int ready_threads = 0;
    std::condition_variable cv_ready;

    double now() const { return clock_ms() - t0; }

    // Caller holds mu.
    void admit_locked() {
        for (auto it = pending.begin(); it != pending.end();) {
            Query *q = it->get();
            if (alone_running) return;
            if (q->alone) {
                if (running > 0 || idle_workers == 0) return;  // waits for an empty device
                alone_running = true;
                q->r.admit_ms = now();
                ++running;
                --idle_workers;
                work.push_back(std::move(*it));
                pending.erase(it);
                cv_work.notify_one();
                return;
            }
            const bool fits = running == 0 || reserved + q->est <= budget;
            const bool batch_ok = q->to_batch && batch_count < cfg.batch_slots;
            const bool worker_ok = !q->to_batch && idle_workers > 0;
            if (!fits || !(batch_ok || worker_ok)) {
                if (++q->passes > kMaxPasses) return;  // nothing passes it any more
                ++it;
                continue;
            }
            q->r.admit_ms = now();
            reserved += q->est;
            peak = std::max(peak, reserved);
            ++running;
            QueryPtr p = std::move(*it);
            it = pending.erase(it);
            if (batch_ok) {
                ++batch_count;
                ++n_batched;
                p->r.batched = 1;
                batch_in.push_back(std::move(p));
                cv_batch.notify_one();
            } else {
                --idle_workers;
                work.push_back(std::move(p));
                cv_work.notify_one();
            }
        }
    }

    // Caller holds mu.
    void finish_locked(QueryPtr q, bool was_alone, bool from_batch) {
        if (was_alone) alone_running = false;
        else reserved -= std::min(reserved, q->est);
        --running;
        if (from_batch) --batch_count;
        else ++idle_workers;
        q->r.done_ms = now();
        q->r.n_keys = q->keys.size();
        q->r.n_trace = q->trace.size();
        done.push_back(std::move(q));
        ++finished;
        admit_locked();
        cv_done.notify_all();
    }

    void worker_main(int e) {
        gdl_config c;
        gdl_config_default(&c);
        c.device = cfg.device;
        gdl_ctx *ctx = nullptr;
        gdl_status st = gdl_ctx_create(&c, &ctx);
        if (st == GDL_OK) {
            // keep freed blocks for the next query, up to this worker's share
            // of the budget: a block returned to HIP synchronises the whole
            // device, and allocating a multi-GB block again costs tens of ms,
            // so a cache that keeps evicting the blocks the next query needs
            // stalls every executor (on MI355X a 1,000-query mix spent a
            // third of its time in that churn at half this limit)
            const uint64_t share = budget / (uint64_t)(cfg.workers + (cfg.batch_tc ? 1 : 0));
            gdl_set_cache_limit(ctx, share);
        }
        {
            std::lock_guard<std::mutex> g(mu);
            if (st != GDL_OK) init_status = st;
            ++ready_threads;
            cv_ready.notify_all();
        }
        if (st != GDL_OK) {
            if (ctx) gdl_ctx_destroy(ctx);
            return;
        }
        for (;;) {
            QueryPtr q;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !work.empty(); });
                if (work.empty()) break;
                q = std::move(work.front());
                work.pop_front();
            }
            const bool alone = q->alone;
            // admission counts the estimate; the cap only stops a runaway
            // query (4x its estimate, at least an even share of the budget)
            const uint64_t share = budget / (uint64_t)(cfg.workers + (cfg.batch_tc ? 1 : 0));
            gdl_set_mem_cap(ctx, alone ? 0 : std::max({4 * q->est, share, kMinReservation}));
            ProgramOutput out;
            out.want_keys = cfg.keep_keys != 0;
            std::string err;
            if (cfg.take_turns) turn.acquire();
            gdl_status rs = run_program_host(ctx, q->kind, q->a.data(), q->a.size(),
                                             q->b.data(), q->b.size(), &out, &err);
            if (cfg.take_turns) turn.release();
            gdl_set_mem_cap(ctx, 0);
            if (rs == GDL_ERR_OUT_OF_MEMORY && !alone) {
                // outgrew its reservation: rerun alone, first in line
                std::lock_guard<std::mutex> g(mu);
                reserved -= std::min(reserved, q->est);
                --running;
                ++idle_workers;
                ++reruns;
                q->alone = true;
                q->r.rerun_alone = 1;
                pending.push_front(std::move(q));
                admit_locked();
                continue;
            }
            q->r.status = rs;
            if (rs == GDL_OK) {
                q->r.n_out = out.n_main;
                q->r.derived = out.derived;
                q->r.iterations = out.iterations;
                q->r.init_iterations = out.init_iterations;
                q->r.checksum = out.checksum;
                if (cfg.keep_keys) q->keys = std::move(out.main_keys);
                q->trace.reserve(out.trace.size());
                for (const auto &t : out.trace)
                    q->trace.push_back({t.iteration, t.relation, t.newt, t.delta, t.full});
            }
            std::lock_guard<std::mutex> g(mu);
            finish_locked(std::move(q), alone, false);
        }
        gdl_ctx_destroy(ctx);
    }

    void batch_main() {
        gdl_config c;
        gdl_config_default(&c);
        c.device = cfg.device;
        gdl_ctx *ctx = nullptr;
        gdl_tcb *b = nullptr;
        gdl_status st = gdl_ctx_create(&c, &ctx);
        if (st == GDL_OK) {
            // the same cache limit as a worker: instances of every size come
            // and go, and an unbounded cache of their freed sets fills the
            // device until every allocation miss (in any thread) starts
            // returning blocks to HIP, each return synchronising the device
            const uint64_t share = budget / (uint64_t)(cfg.workers + 1);
            gdl_set_cache_limit(ctx, share);
            gdl_tcb_config bc;
            gdl_tcb_config_default(&bc);
            bc.max_slots = cfg.batch_slots;
            bc.node_bits = cfg.node_bits;
            bc.mem_budget = ~0ull >> 1;  // the server accounts for memory
            bc.keep_keys = cfg.keep_keys;
            bc.solo_tuples = ~0ull;  // large queries go to the server's workers
            st = gdl_tcb_create(ctx, &bc, &b);
        }
        {
            std::lock_guard<std::mutex> g(mu);
            if (st != GDL_OK) init_status = st;
            ++ready_threads;
            cv_ready.notify_all();
        }
        if (st != GDL_OK) {
            if (ctx) gdl_ctx_destroy(ctx);
            return;
        }
        std::map<uint64_t, QueryPtr> inflight;
        // take_turns: the batch holds the device for a slice of bodies, then
        // lets the worker whose ticket is next run its query
        bool holding = false;
        double t_turn = 0.0;
        for (;;) {
            std::deque<QueryPtr> fresh;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_batch.wait(lk, [&] { return stop || !batch_in.empty() || !inflight.empty(); });
                if (stop && batch_in.empty() && inflight.empty()) break;
                fresh.swap(batch_in);
            }
            if (cfg.take_turns && !holding && (!fresh.empty() || !inflight.empty())) {
                turn.acquire();
                holding = true;
                t_turn = clock_ms();
            }
            for (auto &q : fresh) {
                gdl_status as = gdl_tcb_admit(b, q->qid, q->a.data(), q->a.size(), q->n_nodes,
                                              std::max<uint64_t>(q->est_tuples, 1));
                if (as != GDL_OK) {
                    q->r.status = as;
                    std::lock_guard<std::mutex> g(mu);
                    finish_locked(std::move(q), false, true);
                    continue;
                }
                const uint64_t id = q->qid;
                inflight[id] = std::move(q);
            }
            if (inflight.empty()) {
                if (holding) turn.release();
                holding = false;
                continue;
            }
            uint32_t n_active = 0;
            gdl_status ss = gdl_tcb_step(b, &n_active);
            if (ss != GDL_OK) {
                // the batch is unusable: fail what it holds
                std::lock_guard<std::mutex> g(mu);
                for (auto &kv : inflight) {
                    kv.second->r.status = ss;
                    finish_locked(std::move(kv.second), false, true);
                }
                inflight.clear();
                gdl_tcb_destroy(b);
                b = nullptr;
                gdl_tcb_config bc;
                gdl_tcb_config_default(&bc);
                bc.max_slots = cfg.batch_slots;
                bc.node_bits = cfg.node_bits;
                bc.mem_budget = ~0ull >> 1;
                bc.keep_keys = cfg.keep_keys;
                bc.solo_tuples = ~0ull;
                if (gdl_tcb_create(ctx, &bc, &b) != GDL_OK) break;
                continue;
            }
            gdl_tcb_result tr;
            while (gdl_tcb_poll(b, &tr)) {
                auto it = inflight.find(tr.qid);
                if (it == inflight.end()) {
                    gdl_tcb_fetch(b, tr.qid, nullptr, nullptr);
                    continue;
                }
                QueryPtr q = std::move(it->second);
                inflight.erase(it);
                q->r.status = GDL_OK;
                q->r.n_out = tr.n_out;
                q->r.checksum = tr.checksum;
                q->r.derived = tr.derived;
                q->r.iterations = tr.iterations;
                if (tr.has_keys) q->keys.resize(tr.n_out);
                q->trace.resize(tr.n_trace);
                gdl_tcb_fetch(b, tr.qid, tr.has_keys ? q->keys.data() : nullptr,
                              q->trace.data());
                std::lock_guard<std::mutex> g(mu);
                finish_locked(std::move(q), false, true);
            }
            if (holding && (inflight.empty() || clock_ms() - t_turn > kBatchSliceMs)) {
                turn.release();
                holding = false;
            }
        }
        if (holding) turn.release();
        if (b) gdl_tcb_destroy(b);
        gdl_ctx_destroy(ctx);
    }
};

extern "C" {

void gdl_server_config_default(gdl_server_config *cfg) {
    if (!cfg) return;
    cfg->device = 0;
    cfg->workers = 1;
    cfg->batch_tc = 1;
    cfg->batch_slots = 64;
    cfg->node_bits = 18;
    cfg->mem_budget = 0;
    cfg->keep_keys = 0;
    cfg->partition_cus = 0;
    cfg->take_turns = 0;
}

uint64_t gdl_query_estimate(int kind, const uint64_t *a, uint64_t na, const uint64_t *b,
                            uint64_t nb) {
    if ((na && !a) || (nb && !b) || kind < 0 || kind > 2) return 0;
    uint32_t nodes = 0;
    return estimate_bytes(kind, estimate_tuples(kind, a, na, b, nb, &nodes), na, nb);
}

gdl_status gdl_server_create(const gdl_server_config *cfg_in, gdl_server **out) {
    if (!out) return GDL_ERR_INVALID_ARG;
    *out = nullptr;
    gdl_server_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else gdl_server_config_default(&cfg);
    if (cfg.workers < 1 || cfg.workers > 16 || cfg.batch_slots < 1 || cfg.batch_slots > 256 ||
        cfg.node_bits < 8 || cfg.node_bits > 24)
        return GDL_ERR_INVALID_ARG;
    if (cfg.partition_cus) return GDL_ERR_UNSUPPORTED;  // no CU-masked streams
    auto s = std::make_unique<gdl_server>();
    s->cfg = cfg;
    s->t0 = clock_ms();
    if (cfg.mem_budget) {
        s->budget = cfg.mem_budget;
    } else {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(cfg.device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess)
            return GDL_ERR_HIP;
        s->budget = (uint64_t)(0.8 * (double)fr);
    }
    const int n_threads = cfg.workers + (cfg.batch_tc ? 1 : 0);
    s->idle_workers = cfg.workers;
    for (int i = 0; i < cfg.workers; ++i)
        s->threads.emplace_back([p = s.get(), i] { p->worker_main(i); });
    if (cfg.batch_tc) s->threads.emplace_back([p = s.get()] { p->batch_main(); });
    gdl_status st;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_ready.wait(lk, [&] { return s->ready_threads == n_threads; });
        st = s->init_status;
    }
    if (st != GDL_OK) {
        gdl_server_destroy(s.release());
        return st;
    }
    *out = s.release();
    return GDL_OK;
}

gdl_status gdl_server_destroy(gdl_server *s) {
    if (!s) return GDL_ERR_INVALID_ARG;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] {
            return s->init_status != GDL_OK || s->finished == s->submitted;
        });
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_batch.notify_all();
    for (auto &t : s->threads) t.join();
    delete s;
    return GDL_OK;
}

gdl_status gdl_server_submit(gdl_server *s, uint64_t qid, int kind, const uint64_t *a,
                             uint64_t na, const uint64_t *b, uint64_t nb) {
    if (!s || kind < 0 || kind > 2 || (na && !a) || (nb && !b)) return GDL_ERR_INVALID_ARG;
    auto q = std::make_unique<Query>();
    q->qid = qid;
    q->kind = kind;
    q->a.assign(a, a + na);
    if (kind == GDL_QUERY_CSPA) q->b.assign(b, b + nb);
    q->est_tuples = estimate_tuples(kind, a, na, b, nb, &q->n_nodes);
    // the batch is for many small instances; a large one (a deep tree) runs
    // the stream-ordered linear loop on a worker instead of holding the
    // batch's host-driven bodies for hundreds of iterations
    q->to_batch = kind == GDL_QUERY_TC && s->cfg.batch_tc &&
                  (uint64_t)q->n_nodes <= (1ull << s->cfg.node_bits) &&
                  q->est_tuples <= kBatchMaxTuples;
    if (q->to_batch) {
        // the batch's own accounting: hash set, delta share, edges
        q->est = hash_set_bytes(q->est_tuples) + 8 * q->est_tuples + 32 * na + (1ull << 16);
    } else {
        q->est = estimate_bytes(kind, q->est_tuples, na, nb);
    }
    q->r.qid = qid;
    q->r.kind = kind;
    q->r.est_bytes = q->est;
    std::lock_guard<std::mutex> g(s->mu);
    q->r.submit_ms = s->now();
    s->pending.push_back(std::move(q));
    ++s->submitted;
    s->admit_locked();
    return GDL_OK;
}

int gdl_server_wait(gdl_server *s, gdl_query_result *out, double timeout_ms) {
    if (!s || !out) return 0;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done.empty(); };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return 0;
    QueryPtr q = std::move(s->done.front());
    s->done.pop_front();
    *out = q->r;
    const uint64_t id = q->qid;
    s->popped[id] = std::move(q);
    return 1;
}

int gdl_server_poll(gdl_server *s, gdl_query_result *out) { return gdl_server_wait(s, out, 0.0); }

gdl_status gdl_server_fetch(gdl_server *s, uint64_t qid, uint64_t *keys, gdl_trace_rec *trace) {
    if (!s) return GDL_ERR_INVALID_ARG;
    QueryPtr q;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->popped.find(qid);
        if (it == s->popped.end()) return GDL_ERR_INVALID_ARG;
        q = std::move(it->second);
        s->popped.erase(it);
    }
    if (keys && !q->keys.empty()) std::memcpy(keys, q->keys.data(), q->keys.size() * 8);
    if (trace && !q->trace.empty())
        std::memcpy(trace, q->trace.data(), q->trace.size() * sizeof(gdl_trace_rec));
    return GDL_OK;
}

gdl_status gdl_server_stats(gdl_server *s, uint64_t *submitted, uint64_t *finished,
                            uint64_t *batched, uint64_t *reruns, uint64_t *reserved,
                            uint64_t *peak_reserved, uint64_t *budget) {
    if (!s) return GDL_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (submitted) *submitted = s->submitted;
    if (finished) *finished = s->finished;
    if (batched) *batched = s->n_batched;
    if (reruns) *reruns = s->reruns;
    if (reserved) *reserved = s->reserved;
    if (peak_reserved) *peak_reserved = s->peak;
    if (budget) *budget = s->budget;
    return GDL_OK;
}

}  // extern "C"
