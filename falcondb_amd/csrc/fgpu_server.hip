// fgpu_server.hip — the serving form: a request server with continuous
// batching over device-resident tables (host code; the kernels are the plain
// operators' and the shared COUNT(*) scan in fgpu_filter.hip).
//
// Reference shape: Database is Sync and the REST server runs one sql call per
// worker thread (src/db.rs:80-91, :242; src/bin/server.rs:181-203); each call
// runs its operators on the shared rayon pool.  Here each slot is a context
// with its own HIP stream and a worker thread, so admitted requests run side
// by side on the device, and admission is by HBM: a request starts when a
// slot is free and its scratch memory fits what the running requests leave of
// the budget.
//
// Invariant: the bytes all slots hold (an idle slot its arena, a running one
// its reservation) never exceed the budget.  A reservation is the larger of
// the request's learned need and the slot's fair share of the budget.  A
// running slot's arena is capped at its reservation (the runtime refuses to
// grow past it), so a request that
// needs more than its reservation fails inside the call and is re-queued with
// a larger one; what a request of each shape needs is learned from its runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "falcon_gpu.h"
#include "fgpu_host.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr int MAX_SHARED = 32;    // requests per shared COUNT(*) scan
constexpr int LOOKAHEAD = 64;     // queued requests a free slot looks at
constexpr int MAX_RETRIES = 4;    // budget re-queues of one request
constexpr int GATE_DEPTH = 2;     // waiting requests that open the backlog gate

struct Req {
    uint64_t ticket = 0;
    fgpu_request r{};
    // deep copies the request's pointers refer to
    std::vector<fgpu_pred_node> nodes;
    std::deque<std::string> strs;  // deque: element addresses stay put
    fgpu_host::PredCopies pred_copies;
    fgpu_pred pred{};
    std::vector<fgpu_agg> aggs;
    fgpu_groupby_spec gs{};
    fgpu_join_spec js{};
    uint64_t shape = 0;
    int64_t need = 0;       // scratch bytes to reserve
    int64_t rows_in = 0;
    int share_col = -1;     // column of a shareable COUNT(*)
    Clock::time_point t_submit, t_admit, t_done;
    int status = FGPU_OK;
    std::string err;
    fgpu_result *res = nullptr;
    int slot = -1, shared = 1, retries = 0;
    int64_t reserved = 0;
    bool done = false;
};

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

}  // namespace

struct fgpu_server {
    int device = 0;
    int n_slots = 0;
    int64_t budget = 0;
    bool share = true;
    bool gated = true;  // backlog gate on concurrent starts (pick)
    std::vector<fgpu_ctx *> ctx;
    std::vector<int64_t> held;  // bytes each slot holds: arena when idle, reservation when running
    std::vector<char> busy;
    std::deque<Req *> queue;
    std::unordered_map<uint64_t, Req *> reqs;       // ticket -> request (until waited for)
    std::unordered_map<uint64_t, int64_t> learned;  // request shape -> scratch bytes it needed
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    bool stop = false;
    bool paused = false;
    uint64_t next_ticket = 1;
    fgpu_server_stats st{};
    int running = 0;

    int64_t held_total() const {
        int64_t s = 0;
        for (int64_t h : held) s += h;
        return s;
    }
    void worker(int s);
    bool pick(int s, std::vector<Req *> &batch, std::vector<int> &trim);
    int run(int s, std::vector<Req *> &batch);
};

namespace {

// Copies what the request points at so the caller's memory is free at once.
int copy_request(Req &q, const fgpu_request &in) {
    q.r = in;
    auto keep = [&](const char *p) -> const char * {
        if (!p) return nullptr;
        q.strs.emplace_back(p);
        return q.strs.back().c_str();
    };
    q.r.alias = keep(in.alias);
    if (in.pred && in.pred->n_nodes > 0) {
        if (!in.pred->nodes || in.pred->n_nodes > FGPU_MAX_PRED_NODES)
            return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "request predicate: bad node array");
        q.nodes.assign(in.pred->nodes, in.pred->nodes + in.pred->n_nodes);
        for (auto &n : q.nodes) q.pred_copies.keep(n);
        q.pred.n_nodes = in.pred->n_nodes;
        q.pred.nodes = q.nodes.data();
        q.r.pred = &q.pred;
    } else {
        q.r.pred = nullptr;
    }
    auto copy_aggs = [&](const fgpu_agg *a, int n) -> int {
        if (n < 0 || n > FGPU_MAX_AGGS || (n > 0 && !a))
            return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "request aggregates: bad array");
        q.aggs.assign(a, a + n);
        for (auto &x : q.aggs) x.alias = keep(x.alias);
        return FGPU_OK;
    };
    switch (in.kind) {
        case FGPU_REQ_COUNT:
        case FGPU_REQ_TOPK:
            break;
        case FGPU_REQ_GROUPBY:
            if (!in.groupby) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "GROUPBY request without a spec");
            if (int rc = copy_aggs(in.groupby->aggs, in.groupby->n_aggs)) return rc;
            q.gs = *in.groupby;
            q.gs.aggs = q.aggs.data();
            q.r.groupby = &q.gs;
            break;
        case FGPU_REQ_JOIN:
            if (!in.join || !in.build || !in.index)
                return fgpu_host::set_error(FGPU_E_INVALID_QUERY,
                                            "JOIN request needs a build table, a join index and a spec");
            if (int rc = copy_aggs(in.join->aggs, in.join->n_aggs)) return rc;
            q.js = *in.join;
            q.js.aggs = q.aggs.data();
            q.r.join = &q.js;
            break;
        default:
            return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "unknown request kind");
    }
    return FGPU_OK;
}

// What determines a request's scratch size: the operator, its tables and the
// columns it aggregates (not the predicate literals).
uint64_t shape_of(const Req &q) {
    uint64_t h = mix(0, (uint64_t)q.r.kind);
    h = mix(h, (uint64_t)(uintptr_t)q.r.table);
    h = mix(h, (uint64_t)(uintptr_t)q.r.build);
    h = mix(h, q.r.pred ? (uint64_t)q.r.pred->n_nodes : 0);
    if (q.r.kind == FGPU_REQ_GROUPBY) h = mix(h, (uint64_t)(uint32_t)q.gs.key_col);
    if (q.r.kind == FGPU_REQ_JOIN) {
        h = mix(h, (uint64_t)q.js.type);
        h = mix(h, (uint64_t)(uint32_t)q.js.probe_key);
        h = mix(h, (uint64_t)(uint32_t)q.js.build_key);
    }
    if (q.r.kind == FGPU_REQ_TOPK) {
        h = mix(h, (uint64_t)(uint32_t)q.r.topk_col);
        h = mix(h, (uint64_t)(uint32_t)q.r.topk_k);
    }
    for (const auto &a : q.aggs) h = mix(mix(mix(h, (uint64_t)a.func), (uint64_t)(uint32_t)a.col), (uint64_t)a.side);
    return h;
}

// First guess of a request's scratch need before its shape has run: bounds
// from the row counts (partition cells hold up to one record per row).
int64_t first_guess(const Req &q) {
    const int64_t rows = fgpu_table_rows(q.r.table);
    switch (q.r.kind) {
        case FGPU_REQ_COUNT: return 8192;
        case FGPU_REQ_TOPK: return 4 << 20;
        case FGPU_REQ_GROUPBY: return rows * 8 * (q.gs.n_aggs + 2) * 13 / 10 + (64 << 20);
        default: return rows * 8 * 3 * 13 / 10 + fgpu_table_rows(q.r.build) * 64 + (64 << 20);
    }
}

}  // namespace

// Chooses what slot s runs next (called with the lock held).  Fills `batch`
// and returns true, or fills `trim` with idle slots whose arenas must be
// released before the oldest request fits, or returns false (wait).
bool fgpu_server::pick(int s, std::vector<Req *> &batch, std::vector<int> &trim) {
    batch.clear();
    trim.clear();
    // busy[s] while idle: another worker is releasing this slot's arena
    if (busy[s] || queue.empty() || (paused && !stop)) return false;
    // Backlog gate: beside a running request a new one starts only once
    // GATE_DEPTH requests wait.  The operators are HBM-bound, so two that
    // share the device both finish late; run one after the other, the first
    // finishes early and the second no later.  Overlap pays by hiding each
    // request's host work (launches, read-backs, result copies) behind the
    // other slots' kernels, which matters once a queue has formed.
    if (gated && running > 0 && (int)queue.size() < GATE_DEPTH) return false;
    int64_t others = held_total() - held[s];
    const int n = (int)std::min<size_t>(queue.size(), LOOKAHEAD);
    for (int i = 0; i < n; ++i) {
        Req *r = queue[i];
        if (i > 0 && r->need > budget / 16) continue;  // only small requests pass a waiting one
        if (others + std::max(held[s], r->need) > budget) continue;
        batch.push_back(r);
        queue.erase(queue.begin() + i);
        if (share && r->r.kind == FGPU_REQ_COUNT && r->share_col >= 0) {
            for (size_t j = 0; j < queue.size() && batch.size() < MAX_SHARED;) {
                Req *o = queue[j];
                if (o->r.kind == FGPU_REQ_COUNT && o->r.table == r->r.table && o->share_col == r->share_col) {
                    batch.push_back(o);
                    queue.erase(queue.begin() + (long)j);
                } else {
                    ++j;
                }
            }
        }
        return true;
    }
    // The oldest request does not fit: would it, with the idle arenas released?
    int64_t busy_held = 0;
    for (int i = 0; i < n_slots; ++i)
        if (i != s && busy[i]) busy_held += held[i];
    if (busy_held + queue.front()->need <= budget) {
        for (int i = 0; i < n_slots; ++i)
            if (!busy[i] && held[i] > 0) trim.push_back(i);
    }
    return false;
}

int fgpu_server::run(int s, std::vector<Req *> &batch) {
    fgpu_ctx *c = ctx[s];
    Req *r = batch[0];
    if (batch.size() > 1) {
        std::vector<const fgpu_pred *> preds;
        for (Req *q : batch) preds.push_back(q->r.pred);
        std::vector<int64_t> counts(batch.size());
        const int rc = fgpu_host::filter_count_shared(c, r->r.table, r->share_col, preds.data(),
                                                      (int)batch.size(), counts.data());
        if (rc) return rc;
        for (size_t i = 0; i < batch.size(); ++i)
            batch[i]->res = fgpu_host::result_int64(batch[i]->r.alias, counts[i]);
        return FGPU_OK;
    }
    switch (r->r.kind) {
        case FGPU_REQ_COUNT: {
            int64_t n = 0;
            if (int rc = fgpu_filter_count(c, r->r.table, r->r.pred, &n)) return rc;
            r->res = fgpu_host::result_int64(r->r.alias, n);
            return FGPU_OK;
        }
        case FGPU_REQ_GROUPBY:
            return fgpu_groupby(c, r->r.table, r->r.pred, r->r.groupby, &r->res);
        case FGPU_REQ_JOIN:
            return fgpu_join_agg(c, r->r.table, r->r.pred, r->r.build, r->r.index, r->r.join, &r->res);
        default: {
            const int k = std::max(0, r->r.topk_k);
            std::vector<int64_t> rows((size_t)std::max(k, 1)), vals((size_t)std::max(k, 1));
            int64_t n = 0;
            if (int rc = fgpu_topk(c, r->r.table, r->r.pred, r->r.topk_col, r->r.topk_desc, r->r.topk_k,
                                   rows.data(), vals.data(), &n))
                return rc;
            r->res = fgpu_host::result_topk(rows.data(), vals.data(), n,
                                            fgpu_host::column_is_f64(r->r.table, r->r.topk_col), nullptr);
            return FGPU_OK;
        }
    }
}

void fgpu_server::worker(int s) {
    std::vector<Req *> batch;
    std::vector<int> trim;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (!pick(s, batch, trim)) {
            if (!trim.empty()) {  // release idle arenas, then look again
                for (int i : trim) busy[i] = 1;
                lk.unlock();
                for (int i : trim) (void)fgpu_host::ctx_trim(ctx[i]);
                lk.lock();
                for (int i : trim) {
                    busy[i] = 0;
                    held[i] = fgpu_scratch_bytes(ctx[i]);
                }
                st.trims += (int64_t)trim.size();
                cv_work.notify_all();
                continue;
            }
            if (stop && queue.empty()) return;
            cv_work.wait(lk);
            continue;
        }
        // Beyond its learned need, a slot may keep a fair share of the budget
        // (budget / slots, less what the other slots hold beyond theirs): the
        // request kinds use different buffers and their sizes vary with the
        // predicate, and capping a slot at each request's own need would free
        // and re-allocate those buffers on every call (each hipFree also
        // waits for the other slots' streams).
        const int64_t need = batch[0]->need;
        const int64_t others = held_total() - held[s];
        const int64_t share = std::min(budget / n_slots, budget - others);
        const int64_t mine = std::max({held[s], need, share});
        held[s] = mine;
        busy[s] = 1;
        ++running;
        st.peak_running = std::max<int32_t>(st.peak_running, running);
        st.peak_held_bytes = std::max(st.peak_held_bytes, held_total());
        const auto t_admit = Clock::now();
        for (Req *q : batch) {
            q->t_admit = t_admit;
            q->slot = s;
            q->reserved = mine;
            q->shared = (int)batch.size();
        }
        if (batch.size() > 1) {
            ++st.shared_launches;
            st.shared_requests += (int64_t)batch.size();
        }
        lk.unlock();

        fgpu_host::ctx_budget_begin(ctx[s], mine);
        const int rc = run(s, batch);
        const std::string err = rc ? std::string(fgpu_last_error()) : std::string();
        const int64_t used = fgpu_host::ctx_call_need(ctx[s]);
        const bool over = rc && fgpu_host::ctx_limit_hit(ctx[s]);
        const int64_t arena = fgpu_scratch_bytes(ctx[s]);
        const auto t_done = Clock::now();

        lk.lock();
        held[s] = arena;
        busy[s] = 0;
        --running;
        Req *r = batch[0];
        if (over && batch.size() == 1 && r->retries < MAX_RETRIES) {
            // outgrew its reservation: queue it again, first in line, with more
            // doubling past the budget: one more try with the whole budget,
            // unless what the call already asked for is over it
            int64_t want = std::max(2 * r->need, used + used / 4);
            if (want > budget && used <= budget && r->need < budget) want = budget;
            learned[r->shape] = std::max(learned[r->shape], want);
            r->need = want;
            ++r->retries;
            ++st.retries;
            if (want > budget) {
                r->status = FGPU_E_OUT_OF_MEMORY;
                r->err = "request needs " + std::to_string(want) + " scratch bytes, more than the HBM budget " +
                         std::to_string(budget);
            } else {
                queue.push_front(r);
                cv_work.notify_all();
                continue;
            }
        } else {
            if (!rc) learned[r->shape] = std::max(learned[r->shape], used);
            for (Req *q : batch) {
                q->status = rc;
                q->err = err;
            }
        }
        for (Req *q : batch) {
            q->t_done = t_done;
            q->done = true;
            if (q->status) {
                ++st.failed;
                if (q->res) {
                    fgpu_result_free(q->res);
                    q->res = nullptr;
                }
            } else {
                ++st.completed;
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int fgpu_server_open(int device, int n_slots, int64_t hbm_budget_bytes, int flags, fgpu_server **out) {
    if (!out) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "out is NULL");
    if (n_slots < 1 || n_slots > 64) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "n_slots must be in [1, 64]");
    if (hbm_budget_bytes < (1 << 20))
        return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "HBM budget below 1 MiB");
    auto *s = new fgpu_server();
    s->device = device;
    s->n_slots = n_slots;
    s->budget = hbm_budget_bytes;
    if (flags & ~(FGPU_SERVE_SHARE_COUNTS | FGPU_SERVE_OVERLAP_ALWAYS)) {
        delete s;
        return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "unknown server flags");
    }
    s->share = (flags & FGPU_SERVE_SHARE_COUNTS) != 0;
    s->gated = (flags & FGPU_SERVE_OVERLAP_ALWAYS) == 0;
    for (int i = 0; i < n_slots; ++i) {
        fgpu_ctx *c = nullptr;
        if (int rc = fgpu_open(device, &c)) {
            for (fgpu_ctx *o : s->ctx) fgpu_close(o);
            delete s;
            return rc;
        }
        s->ctx.push_back(c);
    }
    s->held.assign((size_t)n_slots, 0);
    s->busy.assign((size_t)n_slots, 0);
    s->st.n_slots = n_slots;
    s->st.budget_bytes = hbm_budget_bytes;
    for (int i = 0; i < n_slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return FGPU_OK;
}

int fgpu_server_submit(fgpu_server *s, const fgpu_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket || !req->table) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "NULL argument");
    auto *q = new Req();
    if (int rc = copy_request(*q, *req)) {
        delete q;
        return rc;
    }
    q->rows_in = fgpu_table_rows(q->r.table) + (q->r.kind == FGPU_REQ_JOIN ? fgpu_table_rows(q->r.build) : 0);
    q->shape = shape_of(*q);
    q->share_col = q->r.kind == FGPU_REQ_COUNT ? fgpu_host::shared_count_column(q->r.table, q->r.pred) : -1;
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        delete q;
        return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "server is closing");
    }
    auto it = s->learned.find(q->shape);
    q->need = it != s->learned.end() ? it->second : std::min(first_guess(*q), s->budget);
    if (q->need > s->budget) {
        const int64_t need = q->need;
        delete q;
        return fgpu_host::set_error(FGPU_E_OUT_OF_MEMORY, ("request needs " + std::to_string(need) +
                                                           " scratch bytes, more than the HBM budget")
                                                              .c_str());
    }
    q->ticket = s->next_ticket++;
    s->reqs[q->ticket] = q;
    s->queue.push_back(q);
    ++s->st.submitted;
    *ticket = q->ticket;
    s->cv_work.notify_one();
    return FGPU_OK;
}

int fgpu_server_wait(fgpu_server *s, uint64_t ticket, fgpu_result **out, fgpu_request_stats *stats) {
    if (!s || !out) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "NULL argument");
    *out = nullptr;
    Req *q;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "unknown or already waited ticket");
        q = it->second;
        s->cv_done.wait(lk, [q] { return q->done; });
        s->reqs.erase(it);
    }
    if (stats) {
        stats->queue_ms = ms_between(q->t_submit, q->t_admit);
        stats->run_ms = ms_between(q->t_admit, q->t_done);
        stats->latency_ms = ms_between(q->t_submit, q->t_done);
        stats->rows_in = q->rows_in;
        stats->reserved_bytes = q->reserved;
        stats->slot = q->slot;
        stats->shared = q->shared;
        stats->retries = q->retries;
    }
    const int rc = q->status;
    if (rc) fgpu_host::set_error(rc, q->err.c_str());
    *out = q->res;
    delete q;
    return rc;
}

int fgpu_server_pause(fgpu_server *s, int on) {
    if (!s) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "NULL argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = on != 0;
    }
    s->cv_work.notify_all();
    return FGPU_OK;
}

int fgpu_server_get_stats(fgpu_server *s, fgpu_server_stats *out) {
    if (!s || !out) return fgpu_host::set_error(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return FGPU_OK;
}

void fgpu_server_close(fgpu_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &w : s->workers) w.join();
    for (auto &kv : s->reqs) {
        if (kv.second->res) fgpu_result_free(kv.second->res);
        delete kv.second;
    }
    for (fgpu_ctx *c : s->ctx) fgpu_close(c);
    delete s;
}

}  // extern "C"
