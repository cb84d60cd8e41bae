// sb_engine.h -- device-resident multi-instance step engine, shared by
// sb_batch_solve and the request server.
//
// Each resident instance ("slot") owns its state buffers (X ping-pong, Y,
// B-operand planes) and a kernel family + batch tile shape chosen by the
// autotuner for its (N, R) bucket.  A chunk of C steps is one captured
// hipGraph of C persistent launches of k_batch_step; between replays the host
// moves each slot's step_base and may add or retire slots (tables are device
// arrays at stable addresses, so the graph stays valid).
#pragma once

#include <vector>

#include "sb_host.h"

namespace sbh {

struct EngineSlot {
    sb_problem *prob = nullptr;  // J on the device (not owned)
    sb_params params{};
    double c0 = 0.0;
    int family = -1, bshape = -1, bkind = 0;
    float *X[2] = {nullptr, nullptr};
    float *Y = nullptr;
    unsigned char *P[2] = {nullptr, nullptr};
    // Sparse family: spin-major state [N][ldr] (X and Y above stay
    // replica-major for init and readout).
    float *XT[2] = {nullptr, nullptr};
    float *YT = nullptr;
    int ldr = 0;
    size_t bytes = 0;   // device bytes of this slot's state
    int steps_done = 0;
    bool used = false;  // slot holds an instance
};

class Engine {
  public:
    explicit Engine(sb_ctx *ctx);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Device bytes add() would allocate for this instance (dense families;
    // a sparse one adds sparse_state_bytes).
    static size_t state_bytes(int N, int R);
    static size_t sparse_state_bytes(int N, int R);
    // Choose family/shape, allocate state, run init (+ aSB prologue, planes).
    // force_bshape >= 0 skips the choice (the tuner's own trial batches).
    int add(sb_problem *prob, const sb_params &p, double c0, int *slot, int force_bshape = -1);
    // Free the slot's state.
    void remove(int slot);
    // Rebuild work lists and device tables after add/remove.
    int rebuild();
    // Advance every unfinished slot by up to C steps (graph replay).
    int run_chunk(int C);
    // Energies / spins of a slot (any time; normally once finished).
    int read(int slot, sb_result *res);
    bool finished(int slot) const { return slots_[slot].steps_done >= slots_[slot].params.steps; }
    int active_count() const;
    int work_items() const { return n_items_; }
    int grid() const { return grid_; }
    bool graph_used() const { return graph_used_; }
    double last_chunk_ms() const { return last_ms_; }
    const EngineSlot &slot(int i) const { return slots_[i]; }

  private:
    int ensure_tables(size_t n_slots, size_t n_items);
    int capture(int C);
    void upload_bases(int C);

    sb_ctx *ctx_;
    std::vector<EngineSlot> slots_;
    int grid_ = 0;
    int n_items_ = 0;
    // Device tables (stable addresses while capacities hold).
    sbk::BatchInstance *d_inst_ = nullptr;
    sbk::BatchItem *d_items_ = nullptr;
    int *d_wg_ = nullptr;
    size_t cap_inst_ = 0, cap_items_ = 0;
    // Host mirror of the instance table (pinned, for async upload).
    sbk::BatchInstance *h_inst_ = nullptr;
    // Captured chunk.
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    int exec_C_ = 0;
    bool graph_used_ = false;
    bool use_graph_ = true;
    double last_ms_ = 0.0;
};

// Kernel family for a slot (batched kernels only; not SB_PREC_ORDERED).
int engine_family(int precision, int variant, const sb_problem *prob, int R);
// Autotuned batch shape of a family for (N, R), cached per context.
int choose_batch_shape(sb_ctx *ctx, const sb_problem *prob, int family, int variant, int R, int *bshape);
// Times the candidate batch shapes on trial batches of copies of prob and
// returns the fastest in *best (choose_batch_shape's timed path).
int tune_batch_shape(sb_ctx *ctx, sb_problem *prob, int family, int variant, int R, const std::vector<int> &cands,
                     int model_pick, int *best);

}  // namespace sbh
