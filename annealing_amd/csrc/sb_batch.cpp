// sb_batch.cpp -- sb_batch_solve: many independent instances (different N,
// R, steps, variant) advanced together, one persistent launch per step, the
// step loop replayed from a captured hipGraph.  Replaces running the
// reference's TIMES loop (SB/CUDA_GPU/m2_opt.cu:152-202) once per instance.
#include <algorithm>
#include <chrono>
#include <vector>

#include "sb_engine.h"

extern "C" SB_API int sb_batch_solve(sb_ctx *ctx, int32_t n, sb_problem *const *probs, const sb_params *params,
                                     sb_result *results, sb_batch_stats *stats)
{
    if (!ctx || n < 1 || !probs || !params || !results) return SB_E_INVALID;
    auto t0 = std::chrono::steady_clock::now();
    const int prec = params[0].precision;
    std::vector<double> c0(n);
    int max_steps = 0;
    for (int i = 0; i < n; ++i) {
        if (!probs[i] || probs[i]->ctx != ctx) return SB_E_INVALID;
        int rc = sbh::check_params(&params[i], probs[i]->N, probs[i]->sigma, &c0[i]);
        if (rc) return rc;
        if (params[i].precision != prec || prec == SB_PREC_ORDERED) {
            sbh::set_error("sb_batch_solve: all problems need the same precision (FP32 or SPLIT)");
            return SB_E_INVALID;
        }
        max_steps = std::max(max_steps, params[i].steps);
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return SB_E_HIP;
    sbh::Engine eng(ctx);
    std::vector<int> slot(n);
    for (int i = 0; i < n; ++i) {
        int rc = eng.add(probs[i], params[i], c0[i], &slot[i]);
        if (rc) return rc;
    }
    int rc = eng.rebuild();
    if (rc) return rc;
    // Chunks of up to 512 steps; launches past an instance's own step count
    // skip its tiles.
    const int C = std::min(max_steps, 512);
    double step_ms = 0.0;
    for (int done = 0; done < max_steps; done += C) {
        rc = eng.run_chunk(C);
        if (rc) return rc;
        step_ms += eng.last_chunk_ms();
    }
    for (int i = 0; i < n; ++i) {
        rc = eng.read(slot[i], &results[i]);
        if (rc) return rc;
        results[i].step_ms = step_ms;
    }
    double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) results[i].total_ms = total;
    if (stats) {
        stats->work_items = eng.work_items();
        stats->workgroups = eng.grid();
        stats->graph_used = eng.graph_used() ? 1 : 0;
        stats->step_ms = step_ms;
        stats->total_ms = total;
    }
    return SB_OK;
}
