// sb_host.h -- host-side state shared by the library's translation units.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "sb_internal.h"

// Everything the launches of a solo step loop depend on (the key of the
// context's cache of instantiated step-loop graphs).  Compared bytewise.
struct SoloGraphKey {
    int family, shape, bkind, N, R, ld, steps, variant, M, pad;
    const void *J, *h, *X0, *X1, *Y, *P0, *P1;
    double dt, a0, c0, K;
};
constexpr size_t kSoloGraphs = 8;

struct sb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Second stream for solves run as two concurrent replica halves, and the
    // events that fork it from and join it back to `stream`.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Smallest R of a solve run as two replica halves: dSB on the FP8 family,
    // and the FP32 family (any variant).  Split bSB/aSB always run whole.
    int halves_min_r_fp8 = 1536, halves_min_r_fp32 = 768;
    std::recursive_mutex mu;
    // This device's static scratch (sb_energy_noalloc); null if the lookup failed.
    double *scratch = nullptr;
    // Solve workspace (grown on demand, reused across calls).
    float *X0 = nullptr, *X1 = nullptr, *Y = nullptr;
    size_t state_cap = 0;  // floats per buffer
    // B-operand sign planes of X0 / X1 (dSB; at most 2 bytes per state element).
    unsigned char *P0 = nullptr, *P1 = nullptr;
    size_t plane_cap = 0;  // bytes per buffer
    // Energies: H holds 3 R doubles (coupling part, field part, int64 exact
    // coupling part); part holds the per-row-tile partials (FP64 or int64).
    double *H = nullptr, *part = nullptr;
    size_t H_cap = 0, part_cap = 0;
    int8_t *S = nullptr;
    size_t S_cap = 0;
    float *tmp = nullptr;
    size_t tmp_cap = 0;
    // Spin-major state of the sparse family ([N][round_up(R, 64)] each).
    float *XT0 = nullptr, *XT1 = nullptr, *YT = nullptr;
    size_t t_cap = 0;
    // Autotuned (family, shape) per (family set, N bucket, R bucket).
    std::map<std::tuple<int, int, int, int>, std::pair<int, int>> tune;
    bool autotune = true;
    // Launch-bound solves replay their step loop from a cached graph.
    bool solo_graphs = true;
    std::vector<std::pair<SoloGraphKey, hipGraphExec_t>> solo_graph_cache;
    bool verbose = false;   // SB_VERBOSE=1: log autotune timings to stderr
    int live_problems = 0;  // guarded by mu
};

struct sb_problem {
    sb_ctx *ctx = nullptr;
    int N = 0, ld = 0;
    bool integer = false;
    double sigma = 0.0, max_abs = 0.0;
    double sum_upper = 0.0;   // sum_{i<j} J_ij (FP64; exact for integer J)
    float *J32 = nullptr;
    uint16_t *Jbf16 = nullptr;
    uint8_t *Jfp8 = nullptr;  // only when integer and |J| <= 16
    float *h = nullptr;       // external field [ld] (zero padded), or null
    bool h_integer = true;
    // CSR form of J (sb_problem_set_sparse): solves then use the sparse family.
    int *csr_ptr = nullptr, *csr_col = nullptr;
    float *csr_w = nullptr;
    long long nnz = 0;
};

namespace sbh {

// Thread-local detail of the last failure (sb_last_error_message).
void set_error(const std::string &msg);

int hip_fail(hipError_t e, const char *what);

// Validate params against a problem; fills the derived c0.
int check_params(const sb_params *p, int N, double sigma, double *c0_out);
// aSB forms: the coefficients of step 0's local update, applied before the
// first fused step (aSB: M sub-steps, m2_opt.cu:168 with s = 0; aSB-m1: the
// x-step, Bifurcation_m1.cu.c:166).  has_prologue is false for bSB/dSB.
bool has_prologue(int variant);
// Sparse family helpers (spin-major workspace, kernel arguments of a problem).
int ensure_sparse_ws(sb_ctx *ctx, size_t elems);
// CSR form of a problem's J built on the device; nothing is built (nnz stays
// 0) when J has more than max_density * N^2 nonzeros.  free_csr drops it.
int build_csr(sb_problem *p, double max_density);
void free_csr(sb_problem *p);
sbk::SparseArgs sparse_args(const sb_problem *prob, int R, int ldr);
sb_step_coeffs prologue_coeffs(const sb_params &p);

// Pick the kernel family for a precision/variant/problem (before tuning).
// Returns the list of families whose results are identical for this case;
// the autotuner may choose any of them.
std::vector<int> candidate_families(int precision, int variant, const sb_problem *prob, int R);

const void *family_J(int family, const sb_problem *prob);

// Autotune (family, shape) for one problem size; cached per context.
int choose_kernel(sb_ctx *ctx, const sb_problem *prob, int precision, int variant, int R,
                  int *family, int *shape);

int ensure_workspace(sb_ctx *ctx, size_t state_elems, int R, int N);

// Runs the full solve on the context's workspace; shared by sb_solve and the
// server's single-request fallback.
int solve_on_workspace(sb_ctx *ctx, sb_problem *prob, const sb_params *p, double c0,
                       sb_result *res);

// Reads energies/spins of R replicas from device state X ([R][ld]).
int read_results(sb_ctx *ctx, const sb_problem *prob, const float *X, const float *Y, int R,
                 sb_result *res, hipStream_t s);

}  // namespace sbh
