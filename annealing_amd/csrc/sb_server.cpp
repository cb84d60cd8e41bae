// sb_server.cpp -- continuous-batching request server.
//
// A worker thread owns one Engine (device-resident J and state per request).
// With sparse_density > 0, an instance with at most that fraction of nonzeros
// runs the sparse family (its CSR form is built on the device at admission).
// It is off by default: inside the persistent launch (two workgroups per CU)
// a sparse work item is bound by the latency of its dependent gathers, and
// in a mixed load where 1 % of the requests went sparse the served rate fell
// threefold.
// Between chunks of `chunk_steps` steps it retires finished requests (energies
// and spins read back, device memory freed) and admits queued requests while
// the resident set stays within two bounds: the HBM budget (J + state of every
// resident request; a request that does not fit yet lets smaller ones pass)
// and a tile budget of about two work items per workgroup of the persistent
// grid, which is enough to keep every CU busy.  Admitting more than that would
// not raise throughput, only stretch every resident request's steps
// (processor sharing).  The queue is ordered by total work (N^2 R steps,
// shortest first; ties in arrival order) or by arrival (SB_SERVE_ARRIVAL).
// All resident requests advance together in one replayed chunk graph; a
// request's arithmetic depends only on its own (N, R, params), so its result
// is the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "sb_engine.h"

namespace {

using clk = std::chrono::steady_clock;

enum ReqState { Q_QUEUED, Q_ACTIVE, Q_DONE, Q_CANCELLED, Q_FAILED };

struct Request {
    uint64_t ticket = 0;
    sb_params params{};
    int N = 0;
    // Source of J: host copy, or a generator spec.
    std::vector<float> J;
    int gen_kind = -1;
    double gen_p = 0.0;
    uint64_t gen_seed = 0;
    size_t bytes = 0;  // device bytes while resident
    size_t sparse_bytes = 0;  // CSR + spin-major state when admitted as sparse
    double work = 0.0; // N^2 R steps (queue order of SB_SERVE_SHORTEST)
    long long tiles = 0;  // work items per step (estimate, for the tile budget)
    sb_problem *prob = nullptr;
    int slot = -1;
    ReqState state = Q_QUEUED;
    int status = SB_OK;
    bool cancel = false;
    // Results.
    std::vector<double> H;
    std::vector<int64_t> Hi;
    std::vector<double> cut;
    std::vector<int8_t> S;
    int best = 0, exact = 0;
    double c0 = 0.0;
    clk::time_point t_submit, t_done;
};

// Work items per step of a request, from the batched tile shapes: MFMA tiles
// of about 128 spins x 64 replicas, or thin tiles of 16 spins for R <= 4.
long long tile_estimate(int N, int R)
{
    if (R <= sbk::kThinMaxR) return (N + 15) / 16;
    return (long long)((N + 127) / 128) * ((R + 63) / 64);
}

size_t problem_bytes(int N)
{
    const size_t e = (size_t)N * sbk::round_up(N, sbk::kSpinPad);
    return e * (4 + 2 + 1);  // FP32 + BF16 + FP8 forms of J
}

}  // namespace

struct sb_server {
    sb_ctx *ctx = nullptr;
    sb_server_opts opts{};
    std::unique_ptr<sbh::Engine> engine;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_ticket = 1;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> active;
    size_t bytes_in_use = 0;
    long long active_tiles = 0, tile_budget = 0;
    // Instances with at most this fraction of nonzeros run the sparse family
    // (0: never).
    double sparse_density = 0.0;
    sb_server_stats stats{};

    void run();
    int admit_one(Request *r);
    void finish(Request *r, int status);
};

void sb_server::finish(Request *r, int status)
{
    if (r->slot >= 0) {
        engine->remove(r->slot);
        r->slot = -1;
    }
    if (r->prob) {
        sb_problem_destroy(r->prob);
        r->prob = nullptr;
    }
    r->J.clear();
    r->J.shrink_to_fit();
    r->status = status;
    r->t_done = clk::now();
    r->state = status == SB_OK ? Q_DONE : (status == SB_E_CANCELLED ? Q_CANCELLED : Q_FAILED);
}

// Called by the worker without the server lock held (GPU work).
int sb_server::admit_one(Request *r)
{
    int rc;
    if (r->gen_kind >= 0) rc = sb_problem_generate(ctx, r->gen_kind, r->N, r->gen_p, r->gen_seed, &r->prob);
    else rc = sb_problem_create(ctx, r->N, r->J.data(), r->N, &r->prob);
    if (rc) return rc;
    r->J.clear();
    r->J.shrink_to_fit();
    rc = sbh::check_params(&r->params, r->N, r->prob->sigma, &r->c0);
    if (rc) return rc;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    // Low-density instances run the sparse family (a gather per nonzero
    // instead of the dense product).
    if (sparse_density > 0.0) {
        rc = sbh::build_csr(r->prob, sparse_density);
        if (rc) return rc;
        if (r->prob->nnz > 0)
            r->sparse_bytes = (size_t)r->prob->nnz * 8 + ((size_t)r->N + 1) * 4 +
                              sbh::Engine::sparse_state_bytes(r->N, r->params.replicas);
    }
    return engine->add(r->prob, r->params, r->c0, &r->slot);
}

void sb_server::run()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Request *> to_admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
            if (stop) break;
            // Admission: arrival order, backfilling requests that fit.
            for (auto it = queue.begin(); it != queue.end();) {
                Request *r = reqs[*it].get();
                if (r->cancel) {
                    finish(r, SB_E_CANCELLED);
                    ++stats.cancelled;
                    it = queue.erase(it);
                    cv_done.notify_all();
                    continue;
                }
                const int cap = opts.max_active > 0 ? opts.max_active : 1 << 30;
                if ((int)(active.size() + to_admit.size()) >= cap) break;
                // Tile budget (the first request is always admitted).
                if (active_tiles > 0 && active_tiles + r->tiles > tile_budget) break;
                if (bytes_in_use + r->bytes > opts.hbm_budget_bytes) {
                    ++it;
                    continue;
                }
                bytes_in_use += r->bytes;
                active_tiles += r->tiles;
                to_admit.push_back(r);
                it = queue.erase(it);
            }
        }
        bool changed = false;
        for (Request *r : to_admit) {
            int rc = admit_one(r);
            std::lock_guard<std::mutex> lk(mu);
            if (rc) {
                bytes_in_use -= r->bytes;
                active_tiles -= r->tiles;
                finish(r, rc);
                cv_done.notify_all();
            } else {
                // The sparse form's extra bytes, known only after generation.
                r->bytes += r->sparse_bytes;
                bytes_in_use += r->sparse_bytes;
                r->state = Q_ACTIVE;
                active.push_back(r->ticket);
                changed = true;
            }
        }
        {
            // Retire cancelled active requests before running.
            std::lock_guard<std::mutex> lk(mu);
            for (auto it = active.begin(); it != active.end();) {
                Request *r = reqs[*it].get();
                if (r->cancel) {
                    std::lock_guard<std::recursive_mutex> g(ctx->mu);
                    finish(r, SB_E_CANCELLED);
                    bytes_in_use -= r->bytes;
                    active_tiles -= r->tiles;
                    ++stats.cancelled;
                    it = active.erase(it);
                    changed = true;
                    cv_done.notify_all();
                } else {
                    ++it;
                }
            }
            stats.max_active_seen = std::max<int64_t>(stats.max_active_seen, (int64_t)active.size());
        }
        if (active.empty()) continue;
        int rc = SB_OK;
        {
            std::lock_guard<std::recursive_mutex> g(ctx->mu);
            if (changed) rc = engine->rebuild();
            if (rc == SB_OK) rc = engine->run_chunk(std::max(1, opts.chunk_steps));
        }
        std::vector<Request *> done;
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.busy_ms += engine->last_chunk_ms();
            ++stats.chunks;
            for (uint64_t t : active) {
                Request *r = reqs[t].get();
                if (rc != SB_OK || engine->finished(r->slot)) done.push_back(r);
            }
        }
        for (Request *r : done) {
            int st = rc;
            if (st == SB_OK) {
                const int R = r->params.replicas;
                r->H.assign(R, 0.0);
                r->Hi.assign(R, 0);
                r->S.assign((size_t)R * r->N, 0);
                r->cut.assign(R, 0.0);
                sb_result res;
                std::memset(&res, 0, sizeof(res));
                res.energy_f64 = r->H.data();
                res.energy_i64 = r->Hi.data();
                res.cut = r->cut.data();
                res.spins = r->S.data();
                std::lock_guard<std::recursive_mutex> g(ctx->mu);
                st = engine->read(r->slot, &res);
                r->best = res.best_replica;
                r->exact = res.energy_exact;
            }
            std::lock_guard<std::mutex> lk(mu);
            {
                std::lock_guard<std::recursive_mutex> g(ctx->mu);
                finish(r, st);
            }
            bytes_in_use -= r->bytes;
            active_tiles -= r->tiles;
            active.erase(std::find(active.begin(), active.end(), r->ticket));
            ++stats.completed;
            cv_done.notify_all();
        }
        if (!done.empty()) {
            std::lock_guard<std::recursive_mutex> g(ctx->mu);
            (void)engine->rebuild();
        }
    }
}

extern "C" SB_API int sb_server_create(sb_ctx *ctx, const sb_server_opts *opts, sb_server **out)
{
    if (!ctx || !opts || !out || opts->hbm_budget_bytes == 0 || opts->chunk_steps < 1) return SB_E_INVALID;
    if (opts->precision != SB_PREC_FP32 && opts->precision != SB_PREC_SPLIT) return SB_E_INVALID;
    if (opts->policy != SB_SERVE_SHORTEST && opts->policy != SB_SERVE_ARRIVAL) return SB_E_INVALID;
    if (!(opts->sparse_density >= 0.0 && opts->sparse_density <= 1.0)) return SB_E_INVALID;
    sb_server *s = new sb_server();
    s->ctx = ctx;
    s->opts = *opts;
    s->sparse_density = opts->sparse_density;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        if (hipSetDevice(ctx->device) != hipSuccess) {
            delete s;
            return SB_E_HIP;
        }
        s->engine.reset(new sbh::Engine(ctx));
        if (s->engine->grid() <= 0) {
            delete s;
            return SB_E_HIP;
        }
        s->tile_budget = 2LL * s->engine->grid();
        if (const char *tb = std::getenv("SB_SERVE_TILE_BUDGET")) s->tile_budget = std::max(1LL, std::atoll(tb));
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return SB_OK;
}

static int submit_common(sb_server *srv, std::unique_ptr<Request> r, uint64_t *ticket)
{
    const sb_params &p = r->params;
    if (p.precision != srv->opts.precision) {
        sbh::set_error("sb_server_submit: params.precision must match the server's");
        return SB_E_INVALID;
    }
    double c0;
    int rc = sbh::check_params(&p, r->N, 1.0, &c0);
    if (rc) return rc;
    r->bytes = problem_bytes(r->N) + sbh::Engine::state_bytes(r->N, p.replicas);
    r->work = (double)r->N * r->N * p.replicas * p.steps;
    r->tiles = tile_estimate(r->N, p.replicas);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return SB_E_SHUTDOWN;
    ++srv->stats.submitted;
    if (r->bytes > srv->opts.hbm_budget_bytes) {
        ++srv->stats.rejected;
        sbh::set_error("request needs more device memory than the server budget");
        return SB_E_BUDGET;
    }
    r->ticket = srv->next_ticket++;
    r->t_submit = clk::now();
    *ticket = r->ticket;
    if (srv->opts.policy == SB_SERVE_ARRIVAL) {
        srv->queue.push_back(r->ticket);
    } else {
        // Shortest total work first; equal work keeps arrival order.
        const double w = r->work;
        auto pos = std::upper_bound(srv->queue.begin(), srv->queue.end(), w,
                                    [&](double v, uint64_t t) { return v < srv->reqs[t]->work; });
        srv->queue.insert(pos, r->ticket);
    }
    srv->reqs[r->ticket] = std::move(r);
    srv->cv_work.notify_all();
    return SB_OK;
}

extern "C" SB_API int sb_server_submit(sb_server *srv, int32_t N, const float *J, int32_t ldj,
                                       const sb_params *params, uint64_t *ticket)
{
    if (!srv || !J || !params || !ticket || N < 1 || ldj < N) return SB_E_INVALID;
    std::unique_ptr<Request> r(new Request());
    r->params = *params;
    r->N = N;
    r->J.resize((size_t)N * N);
    for (int i = 0; i < N; ++i) std::memcpy(&r->J[(size_t)i * N], J + (size_t)i * ldj, sizeof(float) * N);
    // Validate the instance now so a bad J fails at submission.
    double sigma;
    int32_t integer;
    double mx;
    int rc = sb_instance_stats(N, r->J.data(), N, &sigma, &integer, &mx);
    if (rc) return rc;
    return submit_common(srv, std::move(r), ticket);
}

extern "C" SB_API int sb_server_submit_generated(sb_server *srv, int32_t kind, int32_t N, double p, uint64_t seed,
                                                 const sb_params *params, uint64_t *ticket)
{
    if (!srv || !params || !ticket || N < 1 || kind < 0 || kind > 3) return SB_E_INVALID;
    if ((kind == 0 || kind == 3) && !(p >= 0.0 && p <= 1.0)) return SB_E_INVALID;
    std::unique_ptr<Request> r(new Request());
    r->params = *params;
    r->N = N;
    r->gen_kind = kind;
    r->gen_p = p;
    r->gen_seed = seed;
    return submit_common(srv, std::move(r), ticket);
}

extern "C" SB_API int sb_server_wait(sb_server *srv, uint64_t ticket, sb_result *res, double timeout_ms)
{
    if (!srv || !res) return SB_E_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return SB_E_NOTFOUND;
    Request *r = it->second.get();
    auto ready = [&] { return r->state == Q_DONE || r->state == Q_CANCELLED || r->state == Q_FAILED; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        return SB_E_TIMEOUT;
    }
    int st = r->status;
    if (st == SB_OK) {
        const int R = r->params.replicas;
        if (res->energy_f64) std::memcpy(res->energy_f64, r->H.data(), sizeof(double) * R);
        if (res->energy_i64) std::memcpy(res->energy_i64, r->Hi.data(), sizeof(int64_t) * R);
        if (res->spins) std::memcpy(res->spins, r->S.data(), r->S.size());
        if (res->cut) std::memcpy(res->cut, r->cut.data(), sizeof(double) * R);
        res->best_replica = r->best;
        res->energy_exact = r->exact;
        res->c0_used = r->c0;
        res->total_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        res->step_ms = 0.0;
    }
    srv->reqs.erase(it);
    return st;
}

extern "C" SB_API int sb_server_cancel(sb_server *srv, uint64_t ticket)
{
    if (!srv) return SB_E_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return SB_E_NOTFOUND;
    Request *r = it->second.get();
    if (r->state == Q_QUEUED || r->state == Q_ACTIVE) {
        r->cancel = true;
        srv->cv_work.notify_all();
    }
    return SB_OK;
}

extern "C" SB_API int sb_server_stats_get(sb_server *srv, sb_server_stats *out)
{
    if (!srv || !out) return SB_E_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    out->active = (int64_t)srv->active.size();
    out->queued = (int64_t)srv->queue.size();
    out->bytes_in_use = srv->bytes_in_use;
    return SB_OK;
}

extern "C" SB_API int sb_server_destroy(sb_server *srv)
{
    if (!srv) return SB_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    {
        std::lock_guard<std::recursive_mutex> g(srv->ctx->mu);
        for (auto &kv : srv->reqs) {
            Request *r = kv.second.get();
            if (r->slot >= 0) srv->engine->remove(r->slot);
            r->slot = -1;
            if (r->prob) sb_problem_destroy(r->prob);
            r->prob = nullptr;
        }
        srv->engine.reset();
    }
    srv->cv_done.notify_all();
    delete srv;
    return SB_OK;
}
