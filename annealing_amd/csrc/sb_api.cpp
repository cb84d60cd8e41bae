// sb_api.cpp -- C ABI of the solver: host-only helpers, context, problems and
// the single-instance solve loop (the reference's SB/CUDA_GPU/m2_opt.cu
// main(), lines 100-220, re-designed around one fused launch per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sb_engine.h"
#include "sb_host.h"

using namespace sbk;

// ---------------------------------------------------------------------------
// Errors.

namespace sbh {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SB_E_NOMEM : SB_E_HIP;
}

}  // namespace sbh

#define HIP_TRY(call)                                   \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return sbh::hip_fail(_e, #call); \
    } while (0)

extern "C" SB_API const char *sb_last_error_message(void) { return sbh::g_last_error.c_str(); }

extern "C" SB_API const char *sb_version(void) { return "annealing_amd-sb 0.1.0 (gfx950)"; }

extern "C" SB_API const char *sb_status_string(int status)
{
    switch (status) {
    case SB_OK: return "ok";
    case SB_E_INVALID: return "invalid argument";
    case SB_E_IO: return "i/o error";
    case SB_E_PARSE: return "malformed edge list";
    case SB_E_RANGE: return "spin index out of range";
    case SB_E_INEXACT: return "J not exactly representable in the requested format";
    case SB_E_NOMEM: return "out of memory";
    case SB_E_HIP: return "HIP runtime error";
    case SB_E_BUDGET: return "request exceeds the HBM budget";
    case SB_E_NODEVICE: return "no usable GPU";
    case SB_E_TIMEOUT: return "timed out";
    case SB_E_CANCELLED: return "cancelled";
    case SB_E_NOTFOUND: return "unknown ticket";
    case SB_E_SHUTDOWN: return "server shutting down";
    default: return "unknown status";
    }
}

extern "C" SB_API int sb_default_params(sb_params *p)
{
    if (!p) return SB_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    // SB/CUDA_GPU/m2_opt.cu:10-22
    p->variant = SB_ASB;
    p->precision = SB_PREC_FP32;
    p->init = SB_INIT_REFERENCE;
    p->schedule = SB_SCHED_LINEAR;
    p->steps = 100;   // STEP
    p->replicas = 1;  // TIMES
    p->M = 2;         // M
    p->dt = 1.0;      // DeltaT (local sub-step deltaT = DeltaT / M = 0.5)
    p->a0 = 1.0;      // detune
    p->c0 = 0.0;      // xi: auto, 0.7 a0 / (sigma sqrt(N)) = 0.01565 for K2000 +-1
    p->K = 1.0;       // K
    p->seed = 0;
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Host-only instance helpers.

namespace {

struct Cursor {
    const char *p, *end;
};

void skip_ws(Cursor &c)
{
    while (c.p < c.end && (*c.p == ' ' || *c.p == '\t' || *c.p == '\n' || *c.p == '\r' || *c.p == '\v' ||
                           *c.p == '\f'))
        ++c.p;
}

// 0 = end of input, 1 = token read, -1 = malformed token.
int next_token(Cursor &c, char *buf, size_t cap)
{
    skip_ws(c);
    if (c.p >= c.end) return 0;
    size_t n = 0;
    while (c.p < c.end && !(*c.p == ' ' || *c.p == '\t' || *c.p == '\n' || *c.p == '\r' || *c.p == '\v' ||
                            *c.p == '\f')) {
        if (n + 1 >= cap) return -1;
        buf[n++] = *c.p++;
    }
    buf[n] = 0;
    return 1;
}

int read_index(Cursor &c, long long *v)
{
    char buf[40];
    int t = next_token(c, buf, sizeof(buf));
    if (t <= 0) return t;
    char *e = nullptr;
    long long x = std::strtoll(buf, &e, 10);
    if (e == buf || *e) return -1;
    *v = x;
    return 1;
}

int read_weight(Cursor &c, double *v)
{
    char buf[80];
    int t = next_token(c, buf, sizeof(buf));
    if (t <= 0) return t;
    char *e = nullptr;
    double x = std::strtod(buf, &e);
    if (e == buf || *e || !std::isfinite(x)) return -1;
    *v = x;
    return 1;
}

}  // namespace

// SB/CUDA_GPU/m2_opt.cu:111-123: "%d" for N, then "%d%d%d" triples to EOF,
// couplings[a*N+b] = couplings[b*N+a] = w; assert(a != b).
extern "C" SB_API int sb_edgelist_parse(const char *text, size_t len, int one_based, int32_t *N_out,
                                        float *J, int32_t ldj)
{
    return sb_edgelist_parse_field(text, len, one_based, N_out, J, ldj, nullptr, nullptr);
}

extern "C" SB_API int sb_edgelist_parse_field(const char *text, size_t len, int one_based, int32_t *N_out,
                                              float *J, int32_t ldj, float *h, int32_t *has_field)
{
    if (has_field) *has_field = 0;
    if (!text || !N_out) return SB_E_INVALID;
    Cursor c{text, text + len};
    long long n = 0;
    if (read_index(c, &n) != 1 || n < 1 || n > (1ll << 20)) {
        sbh::set_error("edge list: missing or invalid spin count");
        return SB_E_PARSE;
    }
    *N_out = (int32_t)n;
    if (!J) return SB_OK;
    if (ldj < n) return SB_E_INVALID;
    for (long long i = 0; i < n; ++i) std::memset(J + (size_t)i * ldj, 0, sizeof(float) * (size_t)n);
    if (h) std::memset(h, 0, sizeof(float) * (size_t)n);
    long long line = 1;
    for (;;) {
        long long a, b;
        double w;
        int ra = read_index(c, &a);
        if (ra == 0) break;
        ++line;
        if (ra < 0 || read_index(c, &b) != 1 || read_weight(c, &w) != 1) {
            sbh::set_error("edge list: malformed triple #" + std::to_string(line - 1));
            return SB_E_PARSE;
        }
        if (one_based) {
            a -= 1;
            b -= 1;
        }
        if (a < 0 || b < 0 || a >= n || b >= n) {
            sbh::set_error("edge list: index out of range in triple #" + std::to_string(line - 1));
            return SB_E_RANGE;
        }
        if (a == b) {
            // i i w: external field h_i (the king-graph generators' convention);
            // the reference SB rejects it (m2_opt.cu:119).
            if (!h) {
                sbh::set_error("edge list: self-coupling i == j in triple #" + std::to_string(line - 1));
                return SB_E_INVALID;
            }
            h[a] = (float)w;
            if (has_field) *has_field = 1;
            continue;
        }
        J[(size_t)a * ldj + b] = (float)w;
        J[(size_t)b * ldj + a] = (float)w;
    }
    return SB_OK;
}

extern "C" SB_API int sb_instance_stats(int32_t N, const float *J, int32_t ldj, double *sigma,
                                        int32_t *is_integer, double *max_abs)
{
    if (N < 1 || !J || ldj < N) return SB_E_INVALID;
    double s2 = 0.0, mx = 0.0;
    int integer = 1;
    for (int i = 0; i < N; ++i) {
        const float *row = J + (size_t)i * ldj;
        for (int j = 0; j < N; ++j) {
            double v = row[j];
            if (i == j) {
                if (v != 0.0) return SB_E_INVALID;
                continue;
            }
            if (!std::isfinite(v)) return SB_E_INVALID;
            s2 += v * v;
            mx = std::max(mx, std::fabs(v));
            if (v != std::floor(v)) integer = 0;
        }
    }
    if (sigma) *sigma = N > 1 ? std::sqrt(s2 / ((double)N * (double)(N - 1))) : 0.0;
    if (is_integer) *is_integer = integer;
    if (max_abs) *max_abs = mx;
    return SB_OK;
}

extern "C" SB_API double sb_c0_auto(int32_t variant, int32_t N, double a0, double sigma)
{
    if (N < 1 || !(sigma > 0.0)) return 0.0;
    if (variant == SB_ASB || variant == SB_ASB_M1) return 0.7 * a0 / (sigma * std::sqrt((double)N));
    return 0.5 / (std::sqrt((double)N) * sigma);
}

// Keyed per-pair generator (sbk::gen_coupling): J_ij depends on (seed, i, j)
// only, identical to the device generator of sb_problem_generate.
extern "C" SB_API int sb_generate(int32_t kind, int32_t N, double p, uint64_t seed, float *J)
{
    if (N < 1 || !J || kind < 0 || kind > 3) return SB_E_INVALID;
    if ((kind == 0 || kind == 3) && !(p >= 0.0 && p <= 1.0)) return SB_E_INVALID;
    for (int i = 0; i < N; ++i) {
        J[(size_t)i * N + i] = 0.0f;
        for (int j = i + 1; j < N; ++j) {
            float v = gen_coupling(kind, N, p, seed, (uint32_t)i, (uint32_t)j);
            J[(size_t)i * N + j] = v;
            J[(size_t)j * N + i] = v;
        }
    }
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Coefficients.

extern "C" SB_API int sb_step_coeffs_for(const sb_params *p, double c0, int32_t k, sb_step_coeffs *out)
{
    if (!p || !out || p->steps < 1 || k < 0 || k >= p->steps) return SB_E_INVALID;
    *out = sbk::step_coeffs(p->variant, k, p->steps, p->dt, p->a0, c0, p->K, p->M);
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Context.

extern "C" SB_API int sb_ctx_create(int32_t device, sb_ctx **out)
{
    if (!out) return SB_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        sbh::set_error("no HIP device visible");
        return SB_E_NODEVICE;
    }
    if (device < 0 || device >= n) return SB_E_INVALID;
    HIP_TRY(hipSetDevice(device));
    sb_ctx *c = new sb_ctx();
    c->device = device;
    // Default stream priority only.
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->stream2) (void)hipStreamDestroy(c->stream2);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        delete c;
        return sbh::hip_fail(e, "sb_ctx_create");
    }
    // The device is current: look up its static scratch (sb_energy_noalloc).
    // A failure leaves it null; that entry point then reports it.
    if (noalloc_scratch_address(device, &c->scratch) != hipSuccess) c->scratch = nullptr;
    const char *at = std::getenv("SB_AUTOTUNE");
    if (at && at[0] == '0') c->autotune = false;
    // Debug override SB_SOLO_GRAPH=0: launch every step from the host.
    if (const char *sg = std::getenv("SB_SOLO_GRAPH")) c->solo_graphs = sg[0] != '0';
    const char *vb = std::getenv("SB_VERBOSE");
    if (vb && vb[0] != '0') c->verbose = true;
    // Debug override SB_HALVES_MIN_R=<R> for both families (a large value turns halves off).
    if (const char *hm = std::getenv("SB_HALVES_MIN_R"))
        c->halves_min_r_fp8 = c->halves_min_r_fp32 = std::max(2, std::atoi(hm));
    *out = c;
    return SB_OK;
}

static void free_ws(sb_ctx *c)
{
    (void)hipFree(c->X0);
    (void)hipFree(c->X1);
    (void)hipFree(c->Y);
    (void)hipFree(c->H);
    (void)hipFree(c->part);
    (void)hipFree(c->S);
    (void)hipFree(c->tmp);
    (void)hipFree(c->P0);
    (void)hipFree(c->P1);
    (void)hipFree(c->XT0);
    (void)hipFree(c->XT1);
    (void)hipFree(c->YT);
    c->XT0 = c->XT1 = c->YT = nullptr;
    c->t_cap = 0;
    c->P0 = c->P1 = nullptr;
    c->plane_cap = 0;
    c->X0 = c->X1 = c->Y = nullptr;
    c->H = c->part = nullptr;
    c->S = nullptr;
    c->tmp = nullptr;
    c->state_cap = c->H_cap = c->part_cap = c->S_cap = c->tmp_cap = 0;
}

extern "C" SB_API int sb_ctx_destroy(sb_ctx *ctx)
{
    if (!ctx) return SB_E_INVALID;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        if (ctx->live_problems > 0) {
            sbh::set_error("sb_ctx_destroy: destroy the context's problems first");
            return SB_E_INVALID;
        }
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream2);
    for (auto &e : ctx->solo_graph_cache) (void)hipGraphExecDestroy(e.second);
    ctx->solo_graph_cache.clear();
    free_ws(ctx);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->ev_fork);
    (void)hipEventDestroy(ctx->ev_join);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return SB_OK;
}

namespace sbh {

template <typename T>
static int grow(T **p, size_t *cap, size_t need)
{
    if (*cap >= need) return SB_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc((void **)p, need * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc(workspace)");
    }
    *cap = need;
    return SB_OK;
}

int ensure_workspace(sb_ctx *ctx, size_t state_elems, int R, int N)
{
    int rc;
    // Solves run replica halves on stream2 as well; both streams must be idle
    // before a buffer is freed and reallocated.
    if (ctx->state_cap < state_elems || ctx->plane_cap < plane_bytes(BK_SIGN_BF16) * state_elems)
        (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->state_cap < state_elems) {
        (void)hipStreamSynchronize(ctx->stream);
        size_t c0 = ctx->state_cap, c1 = ctx->state_cap, c2 = ctx->state_cap;
        if ((rc = grow(&ctx->X0, &c0, state_elems))) return rc;
        if ((rc = grow(&ctx->X1, &c1, state_elems))) return rc;
        if ((rc = grow(&ctx->Y, &c2, state_elems))) return rc;
        ctx->state_cap = state_elems;
    }
    const size_t plane_elems = plane_bytes(BK_SIGN_BF16) * state_elems;  // the widest plane kind
    if (ctx->plane_cap < plane_elems) {
        (void)hipStreamSynchronize(ctx->stream);
        size_t c0 = ctx->plane_cap, c1 = ctx->plane_cap;
        if ((rc = grow(&ctx->P0, &c0, plane_elems))) return rc;
        if ((rc = grow(&ctx->P1, &c1, plane_elems))) return rc;
        ctx->plane_cap = plane_elems;
    }
    if ((rc = grow(&ctx->H, &ctx->H_cap, 3 * (size_t)std::max(R, 1)))) return rc;
    if ((rc = grow(&ctx->part, &ctx->part_cap, std::max<size_t>(energy_partial_elems(N, R), 1)))) return rc;
    if ((rc = grow(&ctx->S, &ctx->S_cap, std::max<size_t>((size_t)R * N, 1)))) return rc;
    if ((rc = grow(&ctx->tmp, &ctx->tmp_cap, std::max<size_t>((size_t)R * N, 1)))) return rc;
    return SB_OK;
}

}  // namespace sbh

// ---------------------------------------------------------------------------
// Problems.

extern "C" SB_API int sb_problem_create(sb_ctx *ctx, int32_t N, const float *J, int32_t ldj, sb_problem **out)
{
    if (!ctx || !J || !out || N < 1 || ldj < N) return SB_E_INVALID;
    *out = nullptr;
    double sigma = 0, mx = 0;
    int32_t integer = 0;
    for (int i = 0; i < N; ++i) {
        const float *row = J + (size_t)i * ldj;
        for (int j = i + 1; j < N; ++j)
            if (row[j] != J[(size_t)j * ldj + i]) {
                sbh::set_error("J is not symmetric");
                return SB_E_INVALID;
            }
    }
    int rc = sb_instance_stats(N, J, ldj, &sigma, &integer, &mx);
    if (rc) {
        sbh::set_error("J has a nonzero diagonal or a non-finite entry");
        return rc;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    sb_problem *p = new sb_problem();
    p->ctx = ctx;
    p->N = N;
    p->ld = round_up(N, kSpinPad);
    p->integer = integer != 0;
    p->sigma = sigma;
    p->max_abs = mx;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) p->sum_upper += (double)J[(size_t)i * ldj + j];
    const size_t elems = (size_t)N * p->ld;
    std::vector<float> host(elems, 0.0f);
    for (int i = 0; i < N; ++i) std::memcpy(&host[(size_t)i * p->ld], J + (size_t)i * ldj, sizeof(float) * N);
    hipError_t e = hipMalloc((void **)&p->J32, elems * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(p->J32, host.data(), elems * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMalloc((void **)&p->Jbf16, elems * sizeof(uint16_t));
    if (e == hipSuccess) e = launch_pack_bf16(p->J32, elems, p->Jbf16, ctx->stream);
    if (e == hipSuccess && p->integer && mx <= 16.0) {
        e = hipMalloc((void **)&p->Jfp8, elems);
        if (e == hipSuccess) e = launch_pack_fp8(p->J32, elems, p->Jfp8, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(p->J32);
        (void)hipFree(p->Jbf16);
        (void)hipFree(p->Jfp8);
        delete p;
        return sbh::hip_fail(e, "sb_problem_create");
    }
    ctx->live_problems++;
    *out = p;
    return SB_OK;
}

extern "C" SB_API int sb_problem_from_edgelist_file(sb_ctx *ctx, const char *path, int one_based,
                                                    sb_problem **out)
{
    if (!ctx || !path || !out) return SB_E_INVALID;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        sbh::set_error(std::string("cannot open ") + path);
        return SB_E_IO;
    }
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return SB_E_IO;
    int32_t N = 0, has_field = 0;
    int rc = sb_edgelist_parse(text.data(), text.size(), one_based, &N, nullptr, 0);
    if (rc) return rc;
    std::vector<float> J((size_t)N * N, 0.0f), h((size_t)N, 0.0f);
    rc = sb_edgelist_parse_field(text.data(), text.size(), one_based, &N, J.data(), N, h.data(), &has_field);
    if (rc) return rc;
    rc = sb_problem_create(ctx, N, J.data(), N, out);
    if (rc || !has_field) return rc;
    rc = sb_problem_set_field(*out, h.data());
    if (rc) {
        sb_problem_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

extern "C" SB_API int sb_problem_destroy(sb_problem *p)
{
    if (!p) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    (void)hipStreamSynchronize(p->ctx->stream2);
    (void)hipFree(p->J32);
    (void)hipFree(p->Jbf16);
    (void)hipFree(p->Jfp8);
    (void)hipFree(p->h);
    (void)hipFree(p->csr_ptr);
    (void)hipFree(p->csr_col);
    (void)hipFree(p->csr_w);
    p->ctx->live_problems--;
    delete p;
    return SB_OK;
}

extern "C" SB_API int sb_problem_set_field(sb_problem *p, const float *h)
{
    if (!p) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (!h) {
        (void)hipFree(p->h);
        p->h = nullptr;
        p->h_integer = true;
        return SB_OK;
    }
    bool integer = true;
    for (int i = 0; i < p->N; ++i) {
        if (!std::isfinite(h[i])) {
            sbh::set_error("external field: non-finite h_i");
            return SB_E_INVALID;
        }
        if (h[i] != std::floor(h[i])) integer = false;
    }
    if (!p->h) {
        // Padded to ld with zeros like the state.
        hipError_t e = hipMalloc(&p->h, sizeof(float) * (size_t)p->ld);
        if (e != hipSuccess) {
            p->h = nullptr;
            return sbh::hip_fail(e, "hipMalloc(h)");
        }
    }
    HIP_TRY(hipMemset(p->h, 0, sizeof(float) * (size_t)p->ld));
    HIP_TRY(hipMemcpy(p->h, h, sizeof(float) * (size_t)p->N, hipMemcpyHostToDevice));
    p->h_integer = integer;
    return SB_OK;
}

namespace sbh {

void free_csr(sb_problem *p)
{
    (void)hipFree(p->csr_ptr);
    (void)hipFree(p->csr_col);
    (void)hipFree(p->csr_w);
    p->csr_ptr = p->csr_col = nullptr;
    p->csr_w = nullptr;
    p->nnz = 0;
}

// CSR form of the problem's FP32 J, built on the device (rows in ascending
// column order).  Counts the nonzeros first; when they exceed
// max_density * N^2 nothing is built and p->nnz stays 0.
int build_csr(sb_problem *p, double max_density)
{
    sb_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const int N = p->N;
    free_csr(p);
    int *cnt = nullptr;
    HIP_TRY(hipMalloc((void **)&cnt, sizeof(int) * (size_t)N));
    std::vector<int> c(N);
    hipError_t e = launch_row_nnz(p->J32, N, p->ld, cnt, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c.data(), cnt, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(cnt);
    if (e != hipSuccess) return hip_fail(e, "sparse J: row counts");
    std::vector<int> ptr((size_t)N + 1, 0);
    long long nnz = 0;
    for (int i = 0; i < N; ++i) {
        nnz += c[i];
        if (nnz > INT32_MAX) {
            set_error("sparse J: more than 2^31 - 1 nonzeros");
            return SB_E_INVALID;
        }
        ptr[(size_t)i + 1] = (int)nnz;
    }
    if ((double)nnz > max_density * (double)N * (double)N) return SB_OK;
    const size_t m = (size_t)std::max(nnz, 1ll);
    e = hipMalloc((void **)&p->csr_ptr, sizeof(int) * ptr.size());
    if (e == hipSuccess) e = hipMalloc((void **)&p->csr_col, sizeof(int) * m);
    if (e == hipSuccess) e = hipMalloc((void **)&p->csr_w, sizeof(float) * m);
    if (e == hipSuccess) e = hipMemcpyAsync(p->csr_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_csr_fill(p->J32, N, p->ld, p->csr_ptr, p->csr_col, p->csr_w, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free_csr(p);
        return hip_fail(e, "sparse J upload");
    }
    p->nnz = nnz;
    return SB_OK;
}

}  // namespace sbh

extern "C" SB_API int sb_problem_set_sparse(sb_problem *p, int32_t enable)
{
    if (!p) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream2));
    sbh::free_csr(p);
    if (!enable) return SB_OK;
    return sbh::build_csr(p, 2.0);
}

extern "C" SB_API int sb_problem_sparse_nnz(const sb_problem *p, int64_t *nnz)
{
    if (!p || !nnz) return SB_E_INVALID;
    *nnz = p->nnz;
    return SB_OK;
}

extern "C" SB_API int sb_problem_info(const sb_problem *p, int32_t *N, int32_t *is_integer, double *sigma,
                                      int32_t *split_jfmt)
{
    if (!p) return SB_E_INVALID;
    if (N) *N = p->N;
    if (is_integer) *is_integer = p->integer ? 1 : 0;
    if (sigma) *sigma = p->sigma;
    if (split_jfmt) *split_jfmt = p->Jfp8 ? SB_J_FP8 : SB_J_BF16;
    return SB_OK;
}

extern "C" SB_API int sb_problem_generate(sb_ctx *ctx, int32_t kind, int32_t N, double p, uint64_t seed,
                                          sb_problem **out)
{
    if (!ctx || !out || N < 1 || kind < 0 || kind > 3) return SB_E_INVALID;
    if ((kind == 0 || kind == 3) && !(p >= 0.0 && p <= 1.0)) return SB_E_INVALID;
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    sb_problem *pr = new sb_problem();
    pr->ctx = ctx;
    pr->N = N;
    pr->ld = round_up(N, kSpinPad);
    const size_t elems = (size_t)N * pr->ld;
    double *row_ss = nullptr;
    float *row_max = nullptr;
    int *row_int = nullptr;
    std::vector<double> ss(2 * (size_t)N);
    std::vector<float> mx(N);
    std::vector<int> in(N);
    hipStream_t s = ctx->stream;
    hipError_t e = hipMalloc((void **)&pr->J32, elems * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&row_ss, 2 * (size_t)N * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&row_max, N * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&row_int, N * sizeof(int));
    if (e == hipSuccess) e = launch_generate(kind, N, p, seed, pr->J32, pr->ld, s);
    if (e == hipSuccess) e = launch_row_stats(pr->J32, N, pr->ld, row_ss, row_max, row_int, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ss.data(), row_ss, 2 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(mx.data(), row_max, N * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(in.data(), row_int, N * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    double tot = 0.0, m = 0.0;
    bool integer = true;
    for (int i = 0; i < N; ++i) {
        tot += ss[i];
        pr->sum_upper += ss[(size_t)N + i];
        m = std::max(m, (double)mx[i]);
        integer = integer && in[i] != 0;
    }
    pr->sigma = N > 1 ? std::sqrt(tot / ((double)N * (double)(N - 1))) : 0.0;
    pr->max_abs = m;
    pr->integer = integer;
    if (e == hipSuccess) e = hipMalloc((void **)&pr->Jbf16, elems * sizeof(uint16_t));
    if (e == hipSuccess) e = launch_pack_bf16(pr->J32, elems, pr->Jbf16, s);
    if (e == hipSuccess && pr->integer && m <= 16.0) {
        e = hipMalloc((void **)&pr->Jfp8, elems);
        if (e == hipSuccess) e = launch_pack_fp8(pr->J32, elems, pr->Jfp8, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(row_ss);
    (void)hipFree(row_max);
    (void)hipFree(row_int);
    if (e != hipSuccess) {
        (void)hipFree(pr->J32);
        (void)hipFree(pr->Jbf16);
        (void)hipFree(pr->Jfp8);
        delete pr;
        return sbh::hip_fail(e, "sb_problem_generate");
    }
    ctx->live_problems++;
    *out = pr;
    return SB_OK;
}

extern "C" SB_API int sb_problem_read_j(const sb_problem *p, float *J, int32_t ldj)
{
    if (!p || !J || ldj < p->N) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    HIP_TRY(hipMemcpy2DAsync(J, (size_t)ldj * sizeof(float), p->J32, (size_t)p->ld * sizeof(float),
                             (size_t)p->N * sizeof(float), (size_t)p->N, hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Kernel choice and autotuning.

namespace sbh {

bool has_prologue(int variant) { return variant == SB_ASB || variant == SB_ASB_M1; }

sb_step_coeffs prologue_coeffs(const sb_params &p)
{
    sb_step_coeffs c{};
    c.variant = p.variant;
    c.a0 = p.a0;
    c.K = p.K;
    c.M = p.M;
    c.dts = p.variant == SB_ASB_M1 ? p.dt : p.dt / (double)p.M;
    c.p_next = 0.0f;
    return c;
}

int check_params(const sb_params *p, int N, double sigma, double *c0_out)
{
    if (!p) return SB_E_INVALID;
    if (p->variant < SB_ASB || p->variant > SB_ASB_M1) return SB_E_INVALID;
    if (p->precision < SB_PREC_FP32 || p->precision > SB_PREC_ORDERED) return SB_E_INVALID;
    if (p->init != SB_INIT_UNIFORM && p->init != SB_INIT_REFERENCE) return SB_E_INVALID;
    if (p->schedule != SB_SCHED_LINEAR) return SB_E_INVALID;
    if (p->steps < 1 || p->replicas < 1) return SB_E_INVALID;
    if (p->variant == SB_ASB && p->M < 1) return SB_E_INVALID;
    if (!std::isfinite(p->dt) || !std::isfinite(p->a0) || !std::isfinite(p->c0) || !std::isfinite(p->K))
        return SB_E_INVALID;
    double c0 = p->c0 > 0.0 ? p->c0 : sb_c0_auto(p->variant, N, p->a0, sigma);
    if (c0_out) *c0_out = c0;
    return SB_OK;
}

std::vector<int> candidate_families(int precision, int variant, const sb_problem *prob, int R)
{
    // Few replicas: the matrix-vector families (their field does not depend on
    // which of them runs, see the thin kernels).
    if (R <= kThinMaxR) {
        if (precision == SB_PREC_FP32) return {FAM_THIN_F32};
        if (precision == SB_PREC_SPLIT) {
            if (prob->Jfp8) return {FAM_THIN_FP8, FAM_THIN_BF16};
            return {FAM_THIN_BF16};
        }
        return {};
    }
    if (precision == SB_PREC_FP32) return {FAM_FP32};
    if (precision == SB_PREC_SPLIT) {
        if (prob->Jfp8) {
            if (variant == SB_DSB) return {FAM_DSB_FP8, FAM_SPLIT_FP8, FAM_SPLIT_BF16};
            return {FAM_SPLIT_BF16, FAM_SPLIT_FP8};
        }
        return {FAM_SPLIT_BF16};
    }
    return {};
}

const void *family_J(int family, const sb_problem *prob)
{
    switch (family) {
    case FAM_FP32:
    case FAM_THIN_F32: return prob->J32;
    case FAM_SPLIT_BF16:
    case FAM_THIN_BF16: return prob->Jbf16;
    case FAM_SPLIT_FP8:
    case FAM_DSB_FP8:
    case FAM_THIN_FP8: return prob->Jfp8;
    default: return nullptr;
    }
}

static int bucket(int v)
{
    int b = 1;
    while (b < v) b <<= 1;
    return b;
}

// A step whose launch costs about as much as its work: the matrix-vector
// families, or an MFMA step of at most 512 MFLOP (a few microseconds).
static bool launch_bound(int family, int N, int R)
{
    return is_thin(family) || 2.0 * N * (double)N * R <= 512e6;
}

// Heuristic shape when not tuning: the largest tile that still gives at least
// one workgroup per CU, else the smallest tile.
static int heuristic_shape(int family, int N, int R)
{
    int best = -1;
    long best_area = -1;
    for (int s = 0; s < shape_count(family); ++s) {
        TileShape sh = shape_at(family, s);
        long tiles = (long)((N + sh.mt() - 1) / sh.mt()) * ((R + sh.nt() - 1) / sh.nt());
        long area = (long)sh.mt() * sh.nt();
        if (tiles >= 256 && area > best_area) {
            best = s;
            best_area = area;
        }
    }
    if (best >= 0) return best;
    long smallest = 1L << 40;
    for (int s = 0; s < shape_count(family); ++s) {
        TileShape sh = shape_at(family, s);
        long area = (long)sh.mt() * sh.nt();
        if (area < smallest) {
            smallest = area;
            best = s;
        }
    }
    return best;
}

// Time (family, shape) candidates for a few launches each on the context
// workspace (its contents are overwritten by the init that follows) and return
// the fastest.  The state is random, not zero: MFMA timing depends on the
// operand values (power).
static int tune_candidates(sb_ctx *ctx, const sb_problem *prob, int variant, int R,
                           const std::vector<std::pair<int, int>> &cands, std::pair<int, int> *best)
{
    const int N = prob->N;
    int rc = ensure_workspace(ctx, (size_t)R * prob->ld, R, N);
    if (rc) return rc;
    HIP_TRY(launch_init(SB_INIT_UNIFORM, 0x5eed, N, R, prob->ld, ctx->X0, ctx->Y, ctx->X1, ctx->stream));
    StepArgs a{};
    a.X = ctx->X0;
    a.Xn = ctx->X1;
    a.Y = ctx->Y;
    a.ldj = prob->ld;
    a.ldx = prob->ld;
    a.N = N;
    a.R = R;
    a.c.variant = variant;
    a.c.M = 1;
    float best_ms = 1e30f;
    *best = cands.front();
    const int reps = 3;
    int planes_of = -1;
    for (const auto &fs : cands) {
        const int f = fs.first, s = fs.second;
        a.J = family_J(f, prob);
        a.bkind = plane_kind(f, variant);
        a.B = a.bkind != BK_NONE ? (const void *)ctx->P0 : nullptr;
        a.Bn = a.bkind != BK_NONE ? (void *)ctx->P1 : nullptr;
        if (a.bkind != BK_NONE && planes_of != a.bkind) {
            HIP_TRY(launch_make_planes(a.bkind, ctx->X0, R, prob->ld, ctx->P0, ctx->stream));
            HIP_TRY(launch_make_planes(a.bkind, ctx->X0, R, prob->ld, ctx->P1, ctx->stream));
            planes_of = a.bkind;
        }
        HIP_TRY(launch_step(f, s, a, ctx->stream));  // warm-up
        float ms = 0;
        int n = reps;
        if (launch_bound(f, N, R) && ctx->solo_graphs) {
            // Timed as the solve runs them: back to back in a replayed graph
            // (host launch cost excluded).
            n = 16;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int r = 0; r < n && e == hipSuccess; ++r) e = launch_step(f, s, a, ctx->stream);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(e != hipSuccess ? e : e2, "autotune capture");
            }
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(e, "autotune instantiate");
            e = hipGraphLaunch(x, ctx->stream);  // warm-up
            if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
            if (e == hipSuccess) e = hipGraphLaunch(x, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
            (void)hipGraphExecDestroy(x);
            if (e != hipSuccess) return hip_fail(e, "autotune graph");
        } else {
            HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
            for (int r = 0; r < reps; ++r) HIP_TRY(launch_step(f, s, a, ctx->stream));
            HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
            HIP_TRY(hipEventSynchronize(ctx->ev1));
        }
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ms /= (float)n / (float)reps;
        if (ctx->verbose) {
            TileShape sh = shape_at(f, s);
            std::fprintf(stderr, "[sb autotune] N=%d R=%d family=%d shape=%d (%dx%d, %d thr) %.2f us\n", N, R, f, s,
                         sh.mt(), sh.nt(), sh.threads(), ms * 1e3f / reps);
        }
        if (ms < best_ms) {
            best_ms = ms;
            *best = fs;
        }
    }
    return SB_OK;
}

int choose_kernel(sb_ctx *ctx, const sb_problem *prob, int precision, int variant, int R, int *family,
                  int *shape)
{
    std::vector<int> fams = candidate_families(precision, variant, prob, R);
    if (fams.empty()) return SB_E_INVALID;
    const int N = prob->N;
    // Debug override SB_SHAPE=<family>:<shape> (used by the shape-sweep tests).
    if (const char *force = std::getenv("SB_SHAPE")) {
        int f = -1, s = -1;
        if (std::sscanf(force, "%d:%d", &f, &s) == 2 && std::find(fams.begin(), fams.end(), f) != fams.end() &&
            s >= 0 && s < shape_count(f)) {
            *family = f;
            *shape = s;
            return SB_OK;
        }
    }
    // The key holds the whole candidate set: a problem without an FP8 copy of
    // J must never reuse a choice made among FP8 and BF16 families.
    int fam_set = 0;
    for (int f : fams) fam_set |= 1 << f;
    auto key = std::make_tuple(fam_set * 8 + precision, bucket(N), bucket(R), variant == SB_DSB ? 1 : 0);
    auto it = ctx->tune.find(key);
    if (it != ctx->tune.end()) {
        *family = it->second.first;
        *shape = it->second.second;
        return SB_OK;
    }
    std::pair<int, int> best{fams[0], heuristic_shape(fams[0], N, R)};
    if (ctx->autotune) {
        std::vector<std::pair<int, int>> cands;
        for (int f : fams)
            for (int s = 0; s < shape_count(f); ++s) cands.push_back({f, s});
        int rc = tune_candidates(ctx, prob, variant, R, cands, &best);
        if (rc) return rc;
    }
    ctx->tune[key] = best;
    *family = best.first;
    *shape = best.second;
    return SB_OK;
}

int choose_batch_shape(sb_ctx *ctx, const sb_problem *prob, int family, int variant, int R, int *bshape)
{
    const int N = prob->N;
    // Cache key: -(family + 1) marks the batched-shape table (the solo keys
    // are non-negative).
    auto key = std::make_tuple(-(family + 1), bucket(N), bucket(R), variant == SB_DSB ? 1 : 0);
    auto it = ctx->tune.find(key);
    if (it != ctx->tune.end()) {
        *bshape = it->second.second;
        return SB_OK;
    }
    const int nb = batch_shape_count(family);
    if (nb <= 0) return SB_E_INVALID;
    // Debug override SB_BATCH_SHAPE=<index> (used by the shape-sweep tests).
    if (const char *force = std::getenv("SB_BATCH_SHAPE")) {
        int b = std::atoi(force);
        if (b >= 0 && b < nb) {
            *bshape = b;
            return SB_OK;
        }
    }
    // Candidates: the family's batched shapes, without those wider than R
    // rounded up to 64 replicas (their padding columns would be multiplied
    // for nothing).  A batched step streams every instance's J and there is
    // parallelism across instances, so the shapes are timed on a trial batch
    // of copies of the instance, not as a solo launch (tune_batch_shape).
    // Without the autotuner (SB_AUTOTUNE=0) the pick is the shape with the
    // least modelled traffic (J read once per replica tile, the B operand
    // once per row tile; ties to the larger tile).  Results do not depend on
    // the shape.
    const double bj = family == FAM_FP32 || family == FAM_THIN_F32 ? 4.0
                      : family == FAM_SPLIT_BF16 || family == FAM_THIN_BF16 ? 2.0 : 1.0;
    const int bk = plane_kind(family, variant);
    const double bb = bk == BK_NONE ? 4.0 : (double)plane_bytes(bk);
    int pick = -1;
    double best_cost = 0.0;
    long best_area = 0;
    std::vector<int> cands;
    for (int b = 0; b < nb; ++b) {
        const TileShape sh = batch_shape_at(family, b);
        if (b > 0 && sh.nt() > (R + 63) / 64 * 64) continue;
        cands.push_back(b);
        const double tm = (N + sh.mt() - 1) / sh.mt(), tn = (R + sh.nt() - 1) / sh.nt();
        const double cost = tn * (double)N * N * bj + tm * (double)N * R * bb;
        const long area = (long)sh.mt() * sh.nt();
        if (pick < 0 || cost < best_cost || (cost == best_cost && area > best_area)) {
            pick = b;
            best_cost = cost;
            best_area = area;
        }
    }
    if (ctx->autotune && cands.size() > 1) {
        // A trial that cannot run (device memory short of the trial batch)
        // leaves the modelled pick; the failed call's error is cleared so the
        // next launch check does not report it.
        int tuned = pick;
        if (tune_batch_shape(ctx, const_cast<sb_problem *>(prob), family, variant, R, cands, pick, &tuned) ==
            SB_OK)
            pick = tuned;
        else
            (void)hipGetLastError();
    }
    ctx->tune[key] = {family, pick};
    *bshape = pick;
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Solve.

// Energies are exact integers (int64) when J and the external field are.
static bool exact_energy(const sb_problem *prob) { return prob->integer && (!prob->h || prob->h_integer); }

// Integer-valued J with |J_ij| <= 127: the coupling energy runs as the exact
// integer reduction (int8 MFMA); other J in FP64.
static bool integer_energy(const sb_problem *prob) { return prob->integer && prob->max_abs <= 127.0; }

int read_results(sb_ctx *ctx, const sb_problem *prob, const float *X, const float *Y, int R, sb_result *res,
                 hipStream_t s)
{
    const int N = prob->N, ld = prob->ld;
    // ctx->H: [0, R) coupling part (FP64), [R, 2R) field part, [2R, 3R) the
    // coupling part as int64 (integer form only).
    double *Hc = ctx->H, *Hf = ctx->H + R;
    long long *Hi = (long long *)(ctx->H + 2 * (size_t)R);
    const bool ie = integer_energy(prob);
    if (ie) HIP_TRY(launch_energy_i8(prob->J32, ld, X, ld, N, R, (long long *)ctx->part, Hi, Hc, s));
    else HIP_TRY(launch_energy(prob->J32, ld, X, ld, N, R, ctx->part, Hc, s));
    if (prob->h) HIP_TRY(launch_energy_field(prob->h, X, ld, N, R, Hf, s));
    std::vector<double> hc(R), hf(R, 0.0);
    std::vector<long long> hi(ie ? R : 0);
    HIP_TRY(hipMemcpyAsync(hc.data(), Hc, sizeof(double) * R, hipMemcpyDeviceToHost, s));
    if (prob->h) HIP_TRY(hipMemcpyAsync(hf.data(), Hf, sizeof(double) * R, hipMemcpyDeviceToHost, s));
    if (ie) HIP_TRY(hipMemcpyAsync(hi.data(), Hi, sizeof(long long) * R, hipMemcpyDeviceToHost, s));
    if (res->spins) {
        HIP_TRY(launch_spins(X, ld, N, R, ctx->S, s));
        HIP_TRY(hipMemcpyAsync(res->spins, ctx->S, (size_t)R * N, hipMemcpyDeviceToHost, s));
    }
    if (res->x_final) {
        HIP_TRY(launch_compact(X, ld, N, R, ctx->tmp, s));
        HIP_TRY(hipMemcpyAsync(res->x_final, ctx->tmp, sizeof(float) * (size_t)R * N, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (res->y_final) {
        HIP_TRY(launch_compact(Y, ld, N, R, ctx->tmp, s));
        HIP_TRY(hipMemcpyAsync(res->y_final, ctx->tmp, sizeof(float) * (size_t)R * N, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    const bool exact = exact_energy(prob);
    int best = 0;
    std::vector<double> H(R);
    for (int r = 0; r < R; ++r) {
        H[r] = prob->h ? hc[r] + hf[r] : hc[r];
        if (res->energy_f64) res->energy_f64[r] = H[r];
        if (res->energy_i64)
            res->energy_i64[r] = !exact ? 0
                                 : ie ? (int64_t)hi[r] + (prob->h ? (int64_t)std::llround(hf[r]) : 0)
                                      : (int64_t)std::llround(H[r]);
        // Max-cut mapping W = -J: cut = (sum_{i<j} W_ij - sum_{i<j} W_ij s_i s_j) / 2
        //                             = (-sum_{i<j} J_ij - Hc) / 2.
        if (res->cut) res->cut[r] = (-prob->sum_upper - hc[r]) * 0.5;
        if (H[r] < H[best]) best = r;
    }
    res->best_replica = best;
    res->energy_exact = exact ? 1 : 0;
    return SB_OK;
}

int ensure_sparse_ws(sb_ctx *ctx, size_t elems)
{
    if (ctx->t_cap >= elems) return SB_OK;
    (void)hipStreamSynchronize(ctx->stream);
    size_t c0 = ctx->t_cap, c1 = ctx->t_cap, c2 = ctx->t_cap;
    int rc;
    if ((rc = grow(&ctx->XT0, &c0, elems))) return rc;
    if ((rc = grow(&ctx->XT1, &c1, elems))) return rc;
    if ((rc = grow(&ctx->YT, &c2, elems))) return rc;
    ctx->t_cap = elems;
    return SB_OK;
}

SparseArgs sparse_args(const sb_problem *prob, int R, int ldr)
{
    SparseArgs a{};
    a.ptr = prob->csr_ptr;
    a.col = prob->csr_col;
    a.w = prob->csr_w;
    a.h = prob->h;
    a.N = prob->N;
    a.R = R;
    a.ldr = ldr;
    return a;
}

// Solve on the sparse family: init and readout in the replica-major workspace,
// the step loop on the spin-major copy (padding replicas zero).
static int solve_sparse(sb_ctx *ctx, sb_problem *prob, const sb_params *p, double c0, sb_result *res)
{
    const int N = prob->N, ld = prob->ld, R = p->replicas, ldr = round_up(R, 64);
    hipStream_t s = ctx->stream;
    int rc = ensure_workspace(ctx, (size_t)R * ld, R, N);
    if (rc) return rc;
    if ((rc = ensure_sparse_ws(ctx, (size_t)N * ldr))) return rc;
    float *X = ctx->X0, *Y = ctx->Y;
    HIP_TRY(launch_init(p->init, p->seed, N, R, ld, X, Y, ctx->X1, s));
    if (has_prologue(p->variant)) HIP_TRY(launch_asb_local(prologue_coeffs(*p), N, R, ld, X, Y, s));
    const size_t tb = sizeof(float) * (size_t)N * ldr;
    HIP_TRY(hipMemsetAsync(ctx->XT0, 0, tb, s));
    HIP_TRY(hipMemsetAsync(ctx->XT1, 0, tb, s));
    HIP_TRY(hipMemsetAsync(ctx->YT, 0, tb, s));
    HIP_TRY(launch_to_spin_major(X, ld, N, R, ctx->XT0, ldr, s));
    HIP_TRY(launch_to_spin_major(Y, ld, N, R, ctx->YT, ldr, s));
    SparseArgs a = sparse_args(prob, R, ldr);
    float *XT = ctx->XT0, *XTn = ctx->XT1;
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    for (int k = 0; k < p->steps; ++k) {
        sb_step_coeffs_for(p, c0, k, &a.c);
        a.XT = XT;
        a.XTn = XTn;
        a.YT = ctx->YT;
        hipError_t e = launch_step_sparse(a, s);
        if (e != hipSuccess) return hip_fail(e, "launch_step_sparse");
        std::swap(XT, XTn);
    }
    HIP_TRY(hipEventRecord(ctx->ev1, s));
    HIP_TRY(launch_to_replica_major(XT, ldr, N, R, X, ld, s));
    HIP_TRY(launch_to_replica_major(ctx->YT, ldr, N, R, Y, ld, s));
    rc = read_results(ctx, prob, X, Y, R, res, s);
    if (rc) return rc;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->step_ms = ms;
    res->c0_used = c0;
    return SB_OK;
}

// Replica halves.  Every launch of a step runs the K loop and then the
// epilogue on every CU at once.  From a family's threshold up, the solve runs
// as two halves on two streams: the two kernels fill the device where one
// mid-sized grid leaves CUs idle, and they drift out of phase so one's
// epilogue overlaps the other's K loop.  Measured per step:
//   dSB, FP8 family (state-bound), K2000: R=1536 15.9 against 17.5 us,
//     R=2048 18.8 against about 25 us, R=4096 34.4 against 35.5 us; at
//     R <= 1024 (N = 1000 and 2000) one launch is faster;
//   FP32 family, SK N=2000: R=768 59.2 against 65.4 us, R=1024 63.6 against
//     66.5-69.2 us, R=2048 121.5 against 130 us, R=4096 245.7 against
//     249.8 us; at R <= 512 one launch is faster (39.7 against 47.9 us);
//   split bSB (MFMA-bound): no gain up to R=3000, slower at R=4096.
// A replica's trajectory does not depend on R, so the results are the same.
static bool two_stream_solve(const sb_ctx *ctx, const sb_params *p, int family, int R)
{
    if (p->variant == SB_DSB && family == FAM_DSB_FP8) return R >= ctx->halves_min_r_fp8;
    if (family == FAM_FP32) return R >= ctx->halves_min_r_fp32;
    return false;
}

// The instantiated graph of a whole step loop for `key`, captured on first use
// (the launches of solve_on_workspace's loop, with the state buffers swapped
// every step) and kept in the context's cache (at most kSoloGraphs entries,
// the oldest evicted).
static int solo_graph(sb_ctx *ctx, const SoloGraphKey &key, StepArgs a, hipGraphExec_t *out)
{
    for (auto &e : ctx->solo_graph_cache)
        if (std::memcmp(&e.first, &key, sizeof(key)) == 0) {
            *out = e.second;
            return SB_OK;
        }
    hipStream_t s = ctx->stream;
    float *X = (float *)key.X0, *Xn = (float *)key.X1;
    unsigned char *P = (unsigned char *)key.P0, *Pn = (unsigned char *)key.P1;
    sb_params p{};
    p.variant = key.variant;
    p.steps = key.steps;
    p.M = key.M;
    p.dt = key.dt;
    p.a0 = key.a0;
    p.K = key.K;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < key.steps && e == hipSuccess; ++k) {
        sb_step_coeffs_for(&p, key.c0, k, &a.c);
        a.X = X;
        a.Xn = Xn;
        a.Y = (float *)key.Y;
        if (a.bkind != BK_NONE) {
            a.B = P;
            a.Bn = Pn;
        }
        e = launch_step(key.family, key.shape, a, s);
        std::swap(X, Xn);
        std::swap(P, Pn);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e != hipSuccess ? e : e2, "solo step graph capture");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    if (ctx->solo_graph_cache.size() >= kSoloGraphs) {
        (void)hipGraphExecDestroy(ctx->solo_graph_cache.front().second);
        ctx->solo_graph_cache.erase(ctx->solo_graph_cache.begin());
    }
    ctx->solo_graph_cache.push_back({key, x});
    *out = x;
    return SB_OK;
}

int solve_on_workspace(sb_ctx *ctx, sb_problem *prob, const sb_params *p, double c0, sb_result *res)
{
    if (prob->nnz > 0 && p->precision != SB_PREC_ORDERED) return solve_sparse(ctx, prob, p, c0, res);
    const int N = prob->N, ld = prob->ld, R = p->replicas;
    hipStream_t s = ctx->stream;
    int family = -1, shape = -1;
    int rc;
    if (p->precision != SB_PREC_ORDERED) {
        rc = choose_kernel(ctx, prob, p->precision, p->variant, R, &family, &shape);
        if (rc) return rc;
    }
    // Replica halves [0, R1) on `stream` and [R1, R) on `stream2` (R1 a
    // multiple of 128, the widest replica tile), each with its own kernel.
    bool halves = p->precision != SB_PREC_ORDERED && two_stream_solve(ctx, p, family, R);
    const int R1 = halves ? (R / 2) / 128 * 128 : R, R2 = R - R1;
    int family1 = family, shape1 = shape, family2 = family, shape2 = shape;
    if (halves) {
        if ((rc = choose_kernel(ctx, prob, p->precision, p->variant, R1, &family1, &shape1))) return rc;
        if ((rc = choose_kernel(ctx, prob, p->precision, p->variant, R2, &family2, &shape2))) return rc;
        // The halves share the planes of one family; otherwise run whole.
        halves = family1 == family && family2 == family;
    }
    rc = ensure_workspace(ctx, (size_t)R * ld, R, N);
    if (rc) return rc;
    float *X = ctx->X0, *Xn = ctx->X1, *Y = ctx->Y;
    HIP_TRY(launch_init(p->init, p->seed, N, R, ld, X, Y, Xn, s));
    if (has_prologue(p->variant)) HIP_TRY(launch_asb_local(prologue_coeffs(*p), N, R, ld, X, Y, s));
    StepArgs a{};
    a.J = p->precision == SB_PREC_ORDERED ? (const void *)prob->J32 : family_J(family, prob);
    a.h = prob->h;
    a.ldj = ld;
    a.ldx = ld;
    a.N = N;
    a.R = R;
    a.bkind = p->precision == SB_PREC_ORDERED ? BK_NONE : plane_kind(family, p->variant);
    unsigned char *P = ctx->P0, *Pn = ctx->P1;
    // Both plane buffers start complete: the epilogue rewrites only the
    // 4-spin groups holding real spins, and the K loop reads the padding
    // columns too (J is zero there, but 0 * NaN from stale bytes is NaN).
    if (a.bkind != BK_NONE) {
        HIP_TRY(launch_make_planes(a.bkind, X, R, ld, P, s));
        HIP_TRY(launch_make_planes(a.bkind, X, R, ld, Pn, s));
    }
    // Launch-bound solves (a step of a few microseconds) replay the step loop
    // from a captured graph; the instantiated graph is cached per context,
    // keyed by every argument of its launches.
    if (!halves && p->precision != SB_PREC_ORDERED && ctx->solo_graphs && launch_bound(family, N, R)) {
        SoloGraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.family = family;
        key.shape = shape;
        key.J = a.J;
        key.h = a.h;
        key.X0 = X;
        key.X1 = Xn;
        key.Y = Y;
        key.P0 = P;
        key.P1 = Pn;
        key.bkind = a.bkind;
        key.N = N;
        key.R = R;
        key.ld = ld;
        key.steps = p->steps;
        key.variant = p->variant;
        key.M = p->M;
        key.dt = p->dt;
        key.a0 = p->a0;
        key.c0 = c0;
        key.K = p->K;
        hipGraphExec_t exec = nullptr;
        rc = solo_graph(ctx, key, a, &exec);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ctx->ev0, s));
        HIP_TRY(hipGraphLaunch(exec, s));
        HIP_TRY(hipEventRecord(ctx->ev1, s));
        if (p->steps & 1) std::swap(X, Xn);
        rc = read_results(ctx, prob, X, Y, R, res, s);
        if (rc) return rc;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        res->step_ms = ms;
        res->c0_used = c0;
        return SB_OK;
    }
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    if (halves) {
        HIP_TRY(hipEventRecord(ctx->ev_fork, s));
        HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
    }
    const size_t off2 = halves ? (size_t)R1 * ld : 0;  // first state element of the second half
    const size_t poff2 = off2 * plane_bytes(a.bkind);
    for (int k = 0; k < p->steps; ++k) {
        sb_step_coeffs_for(p, c0, k, &a.c);
        a.X = X;
        a.Xn = Xn;
        a.Y = Y;
        if (a.bkind != BK_NONE) {
            a.B = P;
            a.Bn = Pn;
        }
        hipError_t e;
        if (p->precision == SB_PREC_ORDERED) {
            e = launch_step_ordered(a, s);
        } else if (!halves) {
            e = launch_step(family, shape, a, s);
        } else {
            StepArgs a2 = a;
            a.R = R1;
            e = launch_step(family1, shape1, a, s);
            a.R = R;
            a2.R = R2;
            a2.X = X + off2;
            a2.Xn = Xn + off2;
            a2.Y = Y + off2;
            if (a2.bkind != BK_NONE) {
                a2.B = P + poff2;
                a2.Bn = Pn + poff2;
            }
            if (e == hipSuccess) e = launch_step(family2, shape2, a2, ctx->stream2);
        }
        if (e != hipSuccess) {
            // The other half may still have kernels in flight on the shared
            // workspace: drain stream2 before the workspace can be reused.
            if (halves) (void)hipStreamSynchronize(ctx->stream2);
            return hip_fail(e, "launch_step");
        }
        std::swap(X, Xn);
        std::swap(P, Pn);
    }
    if (halves) {
        HIP_TRY(hipEventRecord(ctx->ev_join, ctx->stream2));
        HIP_TRY(hipStreamWaitEvent(s, ctx->ev_join, 0));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, s));
    rc = read_results(ctx, prob, X, Y, R, res, s);
    if (rc) return rc;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->step_ms = ms;
    res->c0_used = c0;
    return SB_OK;
}

}  // namespace sbh

extern "C" SB_API int sb_solve(sb_ctx *ctx, sb_problem *prob, const sb_params *params, sb_result *res)
{
    if (!ctx || !prob || !params || !res || prob->ctx != ctx) return SB_E_INVALID;
    auto t0 = std::chrono::steady_clock::now();
    double c0 = 0;
    int rc = sbh::check_params(params, prob->N, prob->sigma, &c0);
    if (rc) return rc;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    rc = sbh::solve_on_workspace(ctx, prob, params, c0, res);
    if (rc) return rc;
    res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SB_OK;
}

// ---------------------------------------------------------------------------
// Low-level operators on caller buffers.

extern "C" SB_API int sb_step(sb_ctx *ctx, int32_t precision, const void *dJ, int32_t jfmt, int32_t N,
                              int32_t ldj, const float *dX, float *dXn, float *dY, int32_t R, int32_t ldx,
                              const sb_step_coeffs *c, void *stream)
{
    if (!ctx || !dJ || !dX || !dXn || !dY || !c || N < 1 || R < 1) return SB_E_INVALID;
    if (ldj < N || ldx < N || ldj % kSpinPad || ldx % kSpinPad) return SB_E_INVALID;
    if (c->variant < SB_ASB || c->variant > SB_ASB_M1) return SB_E_INVALID;
    int family;
    const bool thin = R <= kThinMaxR;
    if (precision == SB_PREC_FP32 || precision == SB_PREC_ORDERED) {
        if (jfmt != SB_J_FP32) return SB_E_INVALID;
        family = thin ? FAM_THIN_F32 : FAM_FP32;
    } else if (precision == SB_PREC_SPLIT) {
        if (jfmt == SB_J_BF16) family = thin ? FAM_THIN_BF16 : FAM_SPLIT_BF16;
        else if (jfmt == SB_J_FP8) family = thin ? FAM_THIN_FP8 : c->variant == SB_DSB ? FAM_DSB_FP8 : FAM_SPLIT_FP8;
        else return SB_E_INVALID;
    } else {
        return SB_E_INVALID;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    StepArgs a{};
    a.J = dJ;
    a.X = dX;
    a.Xn = dXn;
    a.Y = dY;
    a.ldj = ldj;
    a.ldx = ldx;
    a.N = N;
    a.R = R;
    a.c = *c;
    hipError_t e;
    if (precision == SB_PREC_ORDERED) {
        e = launch_step_ordered(a, s);
    } else {
        int shape = sbh::heuristic_shape(family, N, R);
        e = launch_step(family, shape, a, s);
    }
    if (e != hipSuccess) return sbh::hip_fail(e, "sb_step");
    return SB_OK;
}

extern "C" SB_API int sb_field(sb_ctx *ctx, sb_problem *prob, int32_t precision, int32_t variant, const float *X,
                               int32_t R, float *F)
{
    if (!ctx || !prob || !X || !F || R < 1 || prob->ctx != ctx) return SB_E_INVALID;
    if (variant < SB_ASB || variant > SB_ASB_M1) return SB_E_INVALID;
    if (precision < SB_PREC_FP32 || precision > SB_PREC_ORDERED) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const int N = prob->N, ld = prob->ld;
    int rc = sbh::ensure_workspace(ctx, (size_t)R * ld, R, N);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    // X into the padded workspace, Y = 0, then one step that returns the field
    // in Y.  bSB, aSB and aSB-m1 run aSB's kick with beta = 1 and no local
    // sub-steps: y' = fmaf(1, F, 0) = F, with no wall whatever |x| is.  dSB
    // runs its own step (the sign of x feeds the product) with
    // alpha = gamma = 0: y' = fmaf(1, F, fmaf(-0, x, 0)) = F, x' = x; its
    // amplitudes are clamped to [-1, 1] first (same signs, so the same field)
    // so the wall never resets y'.
    std::vector<float> Xc;
    const float *Xsrc = X;
    if (variant == SB_DSB) {
        Xc.assign(X, X + (size_t)R * N);
        for (float &v : Xc) v = std::max(-1.0f, std::min(1.0f, v));
        Xsrc = Xc.data();
    }
    HIP_TRY(hipMemsetAsync(ctx->X0, 0, (size_t)R * ld * sizeof(float), s));
    HIP_TRY(hipMemsetAsync(ctx->X1, 0, (size_t)R * ld * sizeof(float), s));
    HIP_TRY(hipMemsetAsync(ctx->Y, 0, (size_t)R * ld * sizeof(float), s));
    HIP_TRY(hipMemcpy2DAsync(ctx->X0, (size_t)ld * sizeof(float), Xsrc, (size_t)N * sizeof(float),
                             (size_t)N * sizeof(float), (size_t)R, hipMemcpyHostToDevice, s));
    StepArgs a{};
    a.X = ctx->X0;
    a.Xn = ctx->X1;
    a.Y = ctx->Y;
    a.ldj = ld;
    a.ldx = ld;
    a.N = N;
    a.R = R;
    a.h = prob->h;
    // J.x for bSB, aSB, aSB-m1 through aSB's kick; J.sign(x) through dSB's step.
    a.c.variant = variant == SB_DSB ? SB_DSB : SB_ASB;
    a.c.beta = 1.0f;
    a.c.do_local = 0;
    hipError_t e;
    if (precision == SB_PREC_ORDERED) {
        a.J = prob->J32;
        e = launch_step_ordered(a, s);
    } else if (prob->nnz > 0) {
        // Sparse family: one step on the spin-major copy of X (Y = 0).
        const int ldr = round_up(R, 64);
        if ((rc = sbh::ensure_sparse_ws(ctx, (size_t)N * ldr))) return rc;
        const size_t tb = sizeof(float) * (size_t)N * ldr;
        HIP_TRY(hipMemsetAsync(ctx->XT0, 0, tb, s));
        HIP_TRY(hipMemsetAsync(ctx->YT, 0, tb, s));
        HIP_TRY(launch_to_spin_major(ctx->X0, ld, N, R, ctx->XT0, ldr, s));
        SparseArgs sa = sbh::sparse_args(prob, R, ldr);
        sa.c = a.c;
        sa.XT = ctx->XT0;
        sa.XTn = ctx->XT1;
        sa.YT = ctx->YT;
        e = launch_step_sparse(sa, s);
        if (e == hipSuccess) e = launch_to_replica_major(ctx->YT, ldr, N, R, ctx->Y, ld, s);
    } else {
        int family = 0, shape = 0;
        rc = sbh::choose_kernel(ctx, prob, precision, a.c.variant, R, &family, &shape);
        if (rc) return rc;
        // choose_kernel may have used the workspace: restore X and Y.
        HIP_TRY(hipMemsetAsync(ctx->X0, 0, (size_t)R * ld * sizeof(float), s));
        HIP_TRY(hipMemsetAsync(ctx->Y, 0, (size_t)R * ld * sizeof(float), s));
        HIP_TRY(hipMemcpy2DAsync(ctx->X0, (size_t)ld * sizeof(float), Xsrc, (size_t)N * sizeof(float),
                                 (size_t)N * sizeof(float), (size_t)R, hipMemcpyHostToDevice, s));
        a.J = sbh::family_J(family, prob);
        e = launch_step(family, shape, a, s);  // B staged from X (no planes)
    }
    if (e != hipSuccess) return sbh::hip_fail(e, "sb_field");
    HIP_TRY(hipMemcpy2DAsync(F, (size_t)N * sizeof(float), ctx->Y, (size_t)ld * sizeof(float),
                             (size_t)N * sizeof(float), (size_t)R, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SB_OK;
}

extern "C" SB_API int sb_init_state(sb_ctx *ctx, int32_t init, uint64_t seed, int32_t N, int32_t R, int32_t ldx,
                                    float *dX, float *dY, void *stream)
{
    if (!ctx || !dX || !dY || N < 1 || R < 1 || ldx < N) return SB_E_INVALID;
    if (init != SB_INIT_UNIFORM && init != SB_INIT_REFERENCE) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(launch_init(init, seed, N, R, ldx, dX, dY, nullptr, s));
    return SB_OK;
}

extern "C" SB_API int sb_asb_prologue(sb_ctx *ctx, const sb_params *p, int32_t N, int32_t R, int32_t ldx,
                                      float *dX, float *dY, void *stream)
{
    if (!ctx || !p || !dX || !dY || N < 1 || R < 1 || ldx < N || p->M < 1) return SB_E_INVALID;
    if (!sbh::has_prologue(p->variant)) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIP_TRY(launch_asb_local(sbh::prologue_coeffs(*p), N, R, ldx, dX, dY, s));
    return SB_OK;
}

extern "C" SB_API int sb_energy(sb_ctx *ctx, const float *dJ, int32_t N, int32_t ldj, const float *dX, int32_t R,
                                int32_t ldx, double *dH, void *stream)
{
    if (!ctx || !dJ || !dX || !dH || N < 1 || R < 1 || ldj < N || ldx < N) return SB_E_INVALID;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    size_t need = energy_partial_elems(N, R);
    int rc = sbh::grow(&ctx->part, &ctx->part_cap, need);
    if (rc) return rc;
    HIP_TRY(launch_energy(dJ, ldj, dX, ldx, N, R, ctx->part, dH, s));
    return SB_OK;
}

// Variant of sb_energy that allocates nothing: the partial sums go to the
// device's static scratch (looked up by sb_ctx_create) instead of the
// context's workspace, which sb_energy grows with hipMalloc.
extern "C" SB_API int sb_energy_noalloc(sb_ctx *ctx, const float *dJ, int32_t N, int32_t ldj, const float *dX,
                                        int32_t R, int32_t ldx, double *dH, void *stream)
{
    if (N >= 1 && R >= 1 && energy_partial_elems(N, R) > noalloc_scratch_elems()) {
        sbh::set_error("sb_energy_noalloc: ceil(N / 64) * R = " + std::to_string(energy_partial_elems(N, R)) +
                       " partial sums exceed the static scratch of " + std::to_string(noalloc_scratch_elems()));
        return SB_E_INVALID;
    }
    if (!ctx || !dJ || !dX || !dH || N < 1 || R < 1 || ldj < N || ldx < N) return SB_E_INVALID;
    if (!ctx->scratch) {
        sbh::set_error("sb_energy_noalloc: no static scratch on this device");
        return SB_E_HIP;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    // Every partial sum the final pass reads is written earlier in the same
    // call, so the scratch needs no clearing.
    HIP_TRY(launch_energy(dJ, ldj, dX, ldx, N, R, ctx->scratch, dH, s));
    return SB_OK;
}
