// sb_engine.cpp -- multi-instance step engine (see sb_engine.h).
#include "sb_engine.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <queue>

using namespace sbk;

#define HIP_TRY(call)                                          \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return sbh::hip_fail(_e, #call); \
    } while (0)

namespace sbh {

int engine_family(int precision, int variant, const sb_problem *prob, int R)
{
    if (precision != SB_PREC_FP32 && precision != SB_PREC_SPLIT) return -1;
    // A problem with a CSR form runs the sparse family, as sb_solve does.
    if (prob->nnz > 0) return FAM_SPARSE;
    if (R <= kThinMaxR) {
        // Same family set as a solo solve; FP8 and BF16 thin fields are equal.
        if (precision == SB_PREC_FP32) return FAM_THIN_F32;
        if (precision == SB_PREC_SPLIT) return prob->Jfp8 ? FAM_THIN_FP8 : FAM_THIN_BF16;
        return -1;
    }
    if (precision == SB_PREC_FP32) return FAM_FP32;
    if (precision == SB_PREC_SPLIT) {
        // Batched instances stream J from HBM every step: the FP8 form when
        // exact halves the bytes of the BF16 form.
        if (prob->Jfp8) return variant == SB_DSB ? FAM_DSB_FP8 : FAM_SPLIT_FP8;
        return FAM_SPLIT_BF16;
    }
    return -1;
}

Engine::Engine(sb_ctx *ctx) : ctx_(ctx)
{
    grid_ = batch_grid_size(ctx->device);
    const char *g = std::getenv("SB_GRAPH");
    if (g && g[0] == '0') use_graph_ = false;
}

Engine::~Engine()
{
    (void)hipStreamSynchronize(ctx_->stream);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (size_t i = 0; i < slots_.size(); ++i)
        if (slots_[i].used) remove((int)i);
    (void)hipFree(d_inst_);
    (void)hipFree(d_items_);
    (void)hipFree(d_wg_);
    if (h_inst_) (void)hipHostFree(h_inst_);
}

size_t Engine::state_bytes(int N, int R)
{
    const size_t e = (size_t)R * round_up(N, kSpinPad);
    return e * sizeof(float) * 3 + e * plane_bytes(BK_SIGN_BF16) * 2;  // X[2], Y, sign planes[2] (<= 2 B)
}

size_t Engine::sparse_state_bytes(int N, int R)
{
    return (size_t)N * round_up(R, 64) * sizeof(float) * 3;  // XT[2], YT
}

int Engine::add(sb_problem *prob, const sb_params &p, double c0, int *slot_out, int force_bshape)
{
    const int N = prob->N, ld = prob->ld, R = p.replicas;
    const int family = engine_family(p.precision, p.variant, prob, R);
    if (family < 0) {
        set_error("batched solve supports SB_PREC_FP32 and SB_PREC_SPLIT");
        return SB_E_INVALID;
    }
    int bshape = 0, rc = SB_OK;
    if (family != FAM_SPARSE) {
        if (force_bshape >= 0 && force_bshape < batch_shape_count(family))
            bshape = force_bshape;
        else
            rc = choose_batch_shape(ctx_, prob, family, p.variant, R, &bshape);
    }
    if (rc) return rc;
    int idx = -1;
    for (size_t i = 0; i < slots_.size(); ++i)
        if (!slots_[i].used) {
            idx = (int)i;
            break;
        }
    if (idx < 0) {
        slots_.emplace_back();
        idx = (int)slots_.size() - 1;
    }
    EngineSlot &s = slots_[idx];
    s = EngineSlot();
    s.prob = prob;
    s.params = p;
    s.c0 = c0;
    s.family = family;
    s.bshape = bshape;
    s.bkind = plane_kind(family, p.variant);
    const size_t e = (size_t)R * ld;
    s.bytes = state_bytes(N, R);
    hipError_t err = hipSuccess;
    for (int b = 0; b < 2 && err == hipSuccess; ++b) err = hipMalloc((void **)&s.X[b], e * sizeof(float));
    if (err == hipSuccess) err = hipMalloc((void **)&s.Y, e * sizeof(float));
    if (err == hipSuccess && s.bkind != BK_NONE)
        for (int b = 0; b < 2 && err == hipSuccess; ++b) err = hipMalloc((void **)&s.P[b], e * plane_bytes(s.bkind));
    if (family == FAM_SPARSE) {
        s.ldr = round_up(R, 64);
        s.bytes += sparse_state_bytes(N, R);
        const size_t et = (size_t)N * s.ldr;
        for (int b = 0; b < 2 && err == hipSuccess; ++b) err = hipMalloc((void **)&s.XT[b], et * sizeof(float));
        if (err == hipSuccess) err = hipMalloc((void **)&s.YT, et * sizeof(float));
    }
    if (err != hipSuccess) {
        s.used = true;
        remove(idx);
        return hip_fail(err, "hipMalloc(engine slot)");
    }
    s.used = true;
    hipStream_t st = ctx_->stream;
    HIP_TRY(launch_init(p.init, p.seed, N, R, ld, s.X[0], s.Y, s.X[1], st));
    if (has_prologue(p.variant)) HIP_TRY(launch_asb_local(prologue_coeffs(p), N, R, ld, s.X[0], s.Y, st));
    // Both plane buffers start complete (the padding columns are read by the
    // K loop; the epilogue rewrites only groups holding real spins).
    if (s.bkind != BK_NONE) {
        HIP_TRY(launch_make_planes(s.bkind, s.X[0], R, ld, s.P[0], st));
        HIP_TRY(launch_make_planes(s.bkind, s.X[0], R, ld, s.P[1], st));
    }
    if (family == FAM_SPARSE) {
        // Spin-major copy of the initial state, padding replicas zero (as
        // sb_solve's sparse path).
        const size_t tb = sizeof(float) * (size_t)N * s.ldr;
        HIP_TRY(hipMemsetAsync(s.XT[0], 0, tb, st));
        HIP_TRY(hipMemsetAsync(s.XT[1], 0, tb, st));
        HIP_TRY(hipMemsetAsync(s.YT, 0, tb, st));
        HIP_TRY(launch_to_spin_major(s.X[0], ld, N, R, s.XT[0], s.ldr, st));
        HIP_TRY(launch_to_spin_major(s.Y, ld, N, R, s.YT, s.ldr, st));
    }
    *slot_out = idx;
    return SB_OK;
}

void Engine::remove(int slot)
{
    EngineSlot &s = slots_[slot];
    if (!s.used) return;
    (void)hipStreamSynchronize(ctx_->stream);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(s.X[b]);
        (void)hipFree(s.P[b]);
        (void)hipFree(s.XT[b]);
    }
    (void)hipFree(s.Y);
    (void)hipFree(s.YT);
    s = EngineSlot();
}

int Engine::active_count() const
{
    int n = 0;
    for (const auto &s : slots_)
        if (s.used && s.steps_done < s.params.steps) ++n;
    return n;
}

int Engine::ensure_tables(size_t n_slots, size_t n_items)
{
    bool realloc = false;
    if (cap_inst_ < n_slots || !d_inst_) {
        size_t cap = std::max<size_t>(64, std::max(n_slots, cap_inst_ * 2));
        (void)hipStreamSynchronize(ctx_->stream);
        (void)hipFree(d_inst_);
        if (h_inst_) (void)hipHostFree(h_inst_);
        d_inst_ = nullptr;
        h_inst_ = nullptr;
        HIP_TRY(hipMalloc((void **)&d_inst_, cap * sizeof(BatchInstance)));
        HIP_TRY(hipHostMalloc((void **)&h_inst_, cap * sizeof(BatchInstance), hipHostMallocDefault));
        std::memset(h_inst_, 0, cap * sizeof(BatchInstance));
        cap_inst_ = cap;
        realloc = true;
    }
    if (cap_items_ < n_items || !d_items_) {
        size_t cap = std::max<size_t>(1024, std::max(n_items, cap_items_ * 2));
        (void)hipStreamSynchronize(ctx_->stream);
        (void)hipFree(d_items_);
        d_items_ = nullptr;
        HIP_TRY(hipMalloc((void **)&d_items_, cap * sizeof(BatchItem)));
        cap_items_ = cap;
        realloc = true;
    }
    if (!d_wg_) {
        HIP_TRY(hipMalloc((void **)&d_wg_, (size_t)(grid_ + 1) * sizeof(int)));
        realloc = true;
    }
    if (realloc && exec_) {
        // Captured launches point at the old tables.
        (void)hipGraphExecDestroy(exec_);
        (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
        exec_C_ = 0;
    }
    return SB_OK;
}

// Rebuilds the instance table and the per-workgroup item lists (see the
// placement rules inside).
int Engine::rebuild()
{
    if (grid_ <= 0) {
        set_error("batched kernel does not fit on the device");
        return SB_E_HIP;
    }
    const size_t n = slots_.size();
    int rc = ensure_tables(std::max<size_t>(n, 1), 1);
    if (rc) return rc;
    struct Job {
        double cost;
        int slot, tile;
    };
    std::vector<std::vector<Job>> per_slot(n);
    std::vector<double> slot_cost(n, 0.0);
    size_t total_items = 0;
    for (size_t i = 0; i < n; ++i) {
        EngineSlot &s = slots_[i];
        BatchInstance &I = h_inst_[i];
        std::memset(&I, 0, sizeof(I));
        if (!s.used) continue;
        const TileShape sh = batch_shape_at(s.family, s.bshape);
        const int N = s.prob->N, R = s.params.replicas;
        I.J = family_J(s.family, s.prob);
        I.h = s.prob->h;
        I.X[0] = s.X[0];
        I.X[1] = s.X[1];
        I.Y = s.Y;
        I.P[0] = s.P[0];
        I.P[1] = s.P[1];
        I.bkind = s.bkind;
        I.ldx = s.prob->ld;
        I.N = N;
        I.R = R;
        I.family = s.family;
        I.bshape = s.bshape;
        if (s.family == FAM_SPARSE) {
            I.csr_ptr = s.prob->csr_ptr;
            I.csr_col = s.prob->csr_col;
            I.csr_w = s.prob->csr_w;
            I.XT[0] = s.XT[0];
            I.XT[1] = s.XT[1];
            I.YT = s.YT;
            I.ldr = s.ldr;
            I.tiles_m = (N + 3) / 4;
            I.tiles_n = s.ldr / 64;
        } else {
            I.tiles_m = (N + sh.mt() - 1) / sh.mt();
            I.tiles_n = (R + sh.nt() - 1) / sh.nt();
        }
        I.steps = s.params.steps;
        I.step_base = s.steps_done;
        I.active = s.steps_done < s.params.steps ? 1 : 0;
        I.variant = s.params.variant;
        I.M = s.params.M;
        I.dt = s.params.dt;
        I.a0 = s.params.a0;
        I.c0 = s.c0;
        I.K = s.params.K;
        if (!I.active) continue;
        // Cost model: MFMA tiles by their area times N; a thin tile streams its
        // rows of J on the VALU at roughly the cost of a 64-replica MFMA row
        // block.  A sparse tile is 4 waves, each a chain of dependent gathers
        // over one row's nonzeros (a few loads in flight per wave): estimated
        // at ~50 ns per nonzero, i.e. ~5e4 area-N units of a split-precision
        // MFMA tile (~1024 units per ns per workgroup at two workgroups per
        // CU) and ~8x fewer units of the FP32 MFMA rate.
        const double gather_units = s.params.precision == SB_PREC_FP32 ? 6400.0 : 51200.0;
        const double tile_cost =
            s.family == FAM_SPARSE
                ? gather_units * std::max(1.0, (double)s.prob->nnz / N)
                : (double)sh.mt() * (sh.rows ? 64 : sh.nt()) * (double)round_up(N, kSpinPad);
        for (int t = 0; t < I.tiles_m * I.tiles_n; ++t) per_slot[i].push_back({tile_cost, (int)i, t});
        slot_cost[i] = tile_cost * I.tiles_m * I.tiles_n;
        total_items += per_slot[i].size();
    }
    rc = ensure_tables(std::max<size_t>(n, 1), std::max<size_t>(total_items, 1));
    if (rc) return rc;

    // Instances keep all their tiles on one XCD (largest first onto the least
    // loaded XCD), so J and state stay in that XCD's L2.  An instance too
    // large for that (more than 1/16 of the step's work, and a J beyond 2 MB)
    // is split over the XCDs in bands of row tiles (XCD x streams band x of
    // J's rows): a few large instances then still occupy the whole device.
    // Splitting costs state traffic (every XCD reads the instance's whole X),
    // so many mid-sized instances stay whole: for 512 instances of N up to
    // 4096 at R = 64, splitting every J beyond 2 MB took 1255-1277 us per
    // batched step in the split form and 3529-3590 us in FP32, keeping them
    // whole 1139 and 3104 us.
    // Each XCD's tiles go, largest first, to its least loaded workgroup
    // (workgroup w runs on XCD w % 8).
    const int nx = 8;
    std::vector<int> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return slot_cost[a] > slot_cost[b]; });
    double total_cost = 0.0;
    for (size_t i = 0; i < n; ++i) total_cost += slot_cost[i];
    auto spread = [&](int i) {
        const BatchInstance &I = h_inst_[i];
        const double bj = I.family == FAM_FP32 || I.family == FAM_THIN_F32 ? 4.0
                          : I.family == FAM_SPLIT_BF16 || I.family == FAM_THIN_BF16 ? 2.0 : 1.0;
        const double jbytes = I.family == FAM_SPARSE ? 8.0 * (double)slots_[i].prob->nnz : (double)I.N * I.ldx * bj;
        return I.tiles_m >= nx && jbytes > 2.0 * (1 << 20) && slot_cost[i] > total_cost / (2 * nx);
    };
    std::vector<double> xcd_load(nx, 0.0);
    std::vector<std::vector<Job>> xcd_jobs(nx);
    for (int i : order) {
        if (per_slot[i].empty() || !spread(i)) continue;
        const int tiles_m = h_inst_[i].tiles_m;
        for (const Job &j : per_slot[i]) {
            const int x = (int)((long long)(j.tile % tiles_m) * nx / tiles_m);
            xcd_jobs[x].push_back(j);
            xcd_load[x] += j.cost;
        }
    }
    for (int i : order) {
        if (per_slot[i].empty() || spread(i)) continue;
        int x = (int)(std::min_element(xcd_load.begin(), xcd_load.end()) - xcd_load.begin());
        xcd_load[x] += slot_cost[i];
        xcd_jobs[x].insert(xcd_jobs[x].end(), per_slot[i].begin(), per_slot[i].end());
    }
    std::vector<std::vector<BatchItem>> wg_items(grid_);
    for (int x = 0; x < nx; ++x) {
        std::vector<int> wgs;
        for (int w = x; w < grid_; w += nx) wgs.push_back(w);
        if (wgs.empty()) wgs.push_back(x % grid_);
        std::vector<Job> &jobs = xcd_jobs[x];
        std::stable_sort(jobs.begin(), jobs.end(), [](const Job &a, const Job &b) { return a.cost > b.cost; });
        typedef std::pair<double, int> LW;
        std::priority_queue<LW, std::vector<LW>, std::greater<LW>> heap;
        for (int w : wgs) heap.push({0.0, w});
        for (const Job &j : jobs) {
            LW top = heap.top();
            heap.pop();
            wg_items[top.second].push_back({j.slot, j.tile});
            heap.push({top.first + j.cost, top.second});
        }
    }
    std::vector<BatchItem> flat;
    flat.reserve(total_items);
    std::vector<int> begin(grid_ + 1, 0);
    for (int w = 0; w < grid_; ++w) {
        begin[w] = (int)flat.size();
        flat.insert(flat.end(), wg_items[w].begin(), wg_items[w].end());
    }
    begin[grid_] = (int)flat.size();
    n_items_ = (int)flat.size();
    hipStream_t st = ctx_->stream;
    if (!flat.empty())
        HIP_TRY(hipMemcpyAsync(d_items_, flat.data(), flat.size() * sizeof(BatchItem), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_wg_, begin.data(), begin.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_inst_, h_inst_, std::max<size_t>(n, 1) * sizeof(BatchInstance), hipMemcpyHostToDevice,
                           st));
    // The pageable sources above must stay valid until the copies complete.
    HIP_TRY(hipStreamSynchronize(st));
    return SB_OK;
}

int Engine::capture(int C)
{
    if (exec_ && exec_C_ == C) return SB_OK;
    if (exec_) {
        (void)hipGraphExecDestroy(exec_);
        (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
    }
    hipStream_t st = ctx_->stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int c = 0; c < C && e == hipSuccess; ++c) e = launch_batch_step(grid_, d_inst_, d_items_, d_wg_, c, st);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e != hipSuccess ? e : e2, "graph capture");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    graph_ = g;
    exec_ = x;
    exec_C_ = C;
    return SB_OK;
}

void Engine::upload_bases(int C)
{
    (void)C;
    for (size_t i = 0; i < slots_.size(); ++i) {
        const EngineSlot &s = slots_[i];
        h_inst_[i].step_base = s.steps_done;
        h_inst_[i].active = (s.used && s.steps_done < s.params.steps) ? 1 : 0;
    }
}

int Engine::run_chunk(int C)
{
    if (C < 1) return SB_E_INVALID;
    hipStream_t st = ctx_->stream;
    upload_bases(C);
    const size_t n = std::max<size_t>(slots_.size(), 1);
    HIP_TRY(hipMemcpyAsync(d_inst_, h_inst_, n * sizeof(BatchInstance), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ctx_->ev0, st));
    graph_used_ = false;
    if (use_graph_) {
        int rc = capture(C);
        if (rc == SB_OK) {
            HIP_TRY(hipGraphLaunch(exec_, st));
            graph_used_ = true;
        }
    }
    if (!graph_used_)
        for (int c = 0; c < C; ++c) HIP_TRY(launch_batch_step(grid_, d_inst_, d_items_, d_wg_, c, st));
    HIP_TRY(hipEventRecord(ctx_->ev1, st));
    HIP_TRY(hipEventSynchronize(ctx_->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx_->ev0, ctx_->ev1));
    last_ms_ = ms;
    for (auto &s : slots_)
        if (s.used) s.steps_done = std::min(s.params.steps, s.steps_done + C);
    return SB_OK;
}

int Engine::read(int slot, sb_result *res)
{
    EngineSlot &s = slots_[slot];
    if (!s.used) return SB_E_INVALID;
    const int R = s.params.replicas, N = s.prob->N;
    int rc = ensure_workspace(ctx_, 0, R, N);
    if (rc) return rc;
    const int k = s.steps_done;
    if (s.family == FAM_SPARSE) {
        HIP_TRY(launch_to_replica_major(s.XT[k & 1], s.ldr, N, R, s.X[k & 1], s.prob->ld, ctx_->stream));
        HIP_TRY(launch_to_replica_major(s.YT, s.ldr, N, R, s.Y, s.prob->ld, ctx_->stream));
    }
    rc = read_results(ctx_, s.prob, s.X[k & 1], s.Y, R, res, ctx_->stream);
    if (rc) return rc;
    res->c0_used = s.c0;
    return SB_OK;
}

// Timed batch-shape choice.  A trial batch holds k copies of the instance
// (J shared, each copy its own state), enough work items to give the
// persistent grid about four items per workgroup, as a loaded batch has (at
// most 128 copies and 1 GiB of state); each candidate shape runs the same trial batch for a replayed
// chunk of steps after a warm-up chunk, and the fastest is kept.  A shape
// never changes a result, only its speed.
int tune_batch_shape(sb_ctx *ctx, sb_problem *prob, int family, int variant, int R, const std::vector<int> &cands,
                     int model_pick, int *best)
{
    const int N = prob->N;
    sb_params p;
    sb_default_params(&p);
    p.variant = variant;
    p.precision = family == FAM_FP32 || family == FAM_THIN_F32 ? SB_PREC_FP32 : SB_PREC_SPLIT;
    p.replicas = R;
    p.steps = 1 << 30;  // never finishes during the trial
    p.init = SB_INIT_UNIFORM;
    p.seed = 1;
    p.dt = 1.0;
    const TileShape ms = batch_shape_at(family, model_pick);
    const long items = (long)((N + ms.mt() - 1) / ms.mt()) * (is_thin(family) ? 1 : (R + ms.nt() - 1) / ms.nt());
    const int grid = batch_grid_size(ctx->device);
    long k = std::max(1L, (4L * grid + items - 1) / std::max(1L, items));
    k = std::min(k, 128L);
    const size_t per_copy = Engine::state_bytes(N, R);
    k = std::max(1L, std::min(k, (long)((size_t)1 << 30) / (long)std::max<size_t>(per_copy, 1)));
    const int C = 4;
    double best_ms = -1.0;
    *best = model_pick;
    for (int b : cands) {
        Engine eng(ctx);
        int rc = SB_OK;
        for (long i = 0; i < k && rc == SB_OK; ++i) {
            int slot = -1;
            rc = eng.add(prob, p, 0.1, &slot, b);
        }
        if (rc == SB_OK) rc = eng.rebuild();
        if (rc == SB_OK) rc = eng.run_chunk(C);  // warm-up (and graph capture)
        if (rc == SB_OK) rc = eng.run_chunk(C);
        if (rc) return rc;
        const double t = eng.last_chunk_ms();
        if (ctx->verbose)
            std::fprintf(stderr, "sb: batch shape family %d shape %d N=%d R=%d x%ld: %.3f ms per %d steps\n", family, b,
                         N, R, k, t, C);
        if (best_ms < 0.0 || t < best_ms) {
            best_ms = t;
            *best = b;
        }
    }
    return SB_OK;
}

}  // namespace sbh
