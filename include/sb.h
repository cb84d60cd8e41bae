/*
 * sb.h -- C ABI of the MI355X (gfx950) simulated-bifurcation solver.
 *
 * Drop-in for the reference's SB/CUDA_GPU solver (strongshih/Annealing).
 * Paths below are relative to the reference repository root.
 *
 * The reference has two boundaries on this path:
 *   - the process boundary `./opt <instance>` (SB/CUDA_GPU/m2_opt.cu:100-103),
 *     which reads an edge list (:111-123) and writes one energy per trial to
 *     output.txt (:206-210).  The CLI annealing_amd/bin/sb_solve keeps it and
 *     is built on sb_problem_from_edgelist_file + sb_solve below.
 *   - the in-process operator pair per step: UpdateTwice<<<N/64,64>>>(x, y, p)
 *     (m2_opt.cu:168) followed by cublasSgemv(handle, OP_N, N, N, &coeff1,
 *     J, N, x, 1, &coeff2, y, 1) (m2_opt.cu:169).  sb_step replaces the pair
 *     with one fused launch over R replicas.
 *
 * Conventions: every entry point returns an sb_status (0 = success) and never
 * exits the process; the library owns device buffers it allocates, the caller
 * owns host buffers.  No torch or HIP types appear in signatures; streams are
 * passed as void* (a hipStream_t, or NULL for the context's own stream).
 * One sb_ctx may be used from several host threads (calls are serialised per
 * context).
 */
#ifndef ANNEALING_AMD_SB_H
#define ANNEALING_AMD_SB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define SB_API __attribute__((visibility("default")))
#else
#define SB_API
#endif

typedef enum sb_status {
    SB_OK = 0,
    SB_E_INVALID = 1,     /* bad argument, i == j edge, unsupported combination */
    SB_E_IO = 2,          /* file cannot be opened / read                       */
    SB_E_PARSE = 3,       /* malformed edge list                                */
    SB_E_RANGE = 4,       /* spin index outside [0, N)                          */
    SB_E_INEXACT = 5,     /* J cannot be stored exactly in the requested format */
    SB_E_NOMEM = 6,       /* device or host allocation failed                   */
    SB_E_HIP = 7,         /* a HIP runtime call failed                          */
    SB_E_BUDGET = 8,      /* request exceeds the server's HBM budget            */
    SB_E_NODEVICE = 9,    /* no usable GPU                                      */
    SB_E_TIMEOUT = 10,    /* sb_server_wait timed out (request still running)   */
    SB_E_CANCELLED = 11,  /* request was cancelled                              */
    SB_E_NOTFOUND = 12,   /* unknown request ticket                             */
    SB_E_SHUTDOWN = 13    /* server is shutting down                            */
} sb_status;

/* Update rule.  SB_ASB is the reference's adiabatic SB with Kerr term and M
 * local sub-steps per coupling kick (m2_opt.cu:58-75,168-169).  SB_ASB_M1 is
 * the reference's one-x-step, one-y-step variant with the coupling inside the
 * y-step (Bifurcation_m1.cu.c:55-92,166-169): per step x += a0 y dt, then
 * y += dt (-K x^3 + p x - a0 x) + xi dt (J x) with the new x, in the
 * reference's mixed FP32/FP64 arithmetic.  SB_BSB and SB_DSB are ballistic and
 * discrete SB (Goto et al., Sci. Adv. 7:eabe7953, 2021), which the reference
 * does not implement. */
typedef enum sb_variant { SB_ASB = 0, SB_BSB = 1, SB_DSB = 2, SB_ASB_M1 = 3 } sb_variant;

/* Arithmetic of the field F = J.X.
 *   SB_PREC_FP32    FP32 J and X on FP32 MFMA (the plain path).
 *   SB_PREC_SPLIT   J in FP8 E4M3 when integer-valued with |J| <= 16, else
 *                   BF16; X split into BF16 hi + lo; FP32 accumulation.
 *                   dSB with FP8 J runs one exact FP8 x FP8 pass.
 *   SB_PREC_ORDERED FP32, each F_i summed sequentially over j with fmaf, the
 *                   order of the reference's naive kernel
 *                   (Bifurcation_m2.cu.c:93-97).  Slow; for exact parity. */
typedef enum sb_precision {
    SB_PREC_FP32 = 0,
    SB_PREC_SPLIT = 1,
    SB_PREC_ORDERED = 2
} sb_precision;

/* Initial (x, y).
 *   SB_INIT_UNIFORM   x, y ~ U[-0.1, 0.1) from a keyed hash of
 *                     (seed, replica, spin, component); independent of batching.
 *   SB_INIT_REFERENCE the reference's prepare_points (m2_opt.cu:48-56): one
 *                     xorshift32 stream per spin seeded i + seed*N, replica r
 *                     taking draws 2r+1, 2r+2; values in [0, 0.1). */
typedef enum sb_init_kind { SB_INIT_UNIFORM = 0, SB_INIT_REFERENCE = 1 } sb_init_kind;

/* Pump schedule: p_k = k / steps (m2_opt.cu:158); bSB/dSB use a(t) = a0 p_k. */
typedef enum sb_schedule { SB_SCHED_LINEAR = 0 } sb_schedule;

/* Storage format of J on the device (sb_step / sb_energy). */
typedef enum sb_jfmt { SB_J_FP32 = 0, SB_J_BF16 = 1, SB_J_FP8 = 2 } sb_jfmt;

/* Solver knobs.  The reference has them as compile-time defines
 * (m2_opt.cu:10-22): steps <- STEP, dt <- DeltaT (coupling kick; the aSB local
 * sub-step is dt/M, deltaT = 0.5 for DeltaT = 1, M = 2), a0 <- detune,
 * c0 <- xi, K <- K, M <- M, replicas <- TIMES. */
typedef struct sb_params {
    int32_t variant;    /* sb_variant                                     */
    int32_t precision;  /* sb_precision                                   */
    int32_t init;       /* sb_init_kind                                   */
    int32_t schedule;   /* sb_schedule                                    */
    int32_t steps;      /* >= 1                                           */
    int32_t replicas;   /* >= 1                                           */
    int32_t M;          /* aSB local sub-steps per kick (>= 1)            */
    int32_t reserved;
    double dt;          /* time step                                      */
    double a0;          /* detune / a0                                    */
    double c0;          /* coupling; <= 0 selects the default (sb_c0_auto) */
    double K;           /* aSB Kerr coefficient                           */
    uint64_t seed;
} sb_params;

/* Outputs of sb_solve.  Pointer fields are caller-owned host buffers and
 * may be NULL; the remaining fields are written by the library. */
typedef struct sb_result {
    int64_t *energy_i64;  /* [R] exact H when J is integer-valued             */
    double *energy_f64;   /* [R] H in FP64 (exact for integer-valued J)       */
    /* Integer-valued J with |J_ij| <= 127: the coupling sum is an exact integer
     * reduction on the device (int8 J x int8 spins on the i8 MFMA, int32
     * accumulators, int64 sums); other J: FP64. */
    int8_t *spins;        /* [R*N] s = x > 0 ? +1 : -1, replica-major          */
    float *x_final;       /* [R*N] final x                                    */
    float *y_final;       /* [R*N] final y                                    */
    int32_t best_replica; /* argmin of H (lowest index on ties)              */
    int32_t energy_exact; /* 1 when energies are exact integers              */
    double c0_used;
    double step_ms;       /* device time of the step loop (HIP events)       */
    double total_ms;      /* host wall time of the whole call                */
    /* [R] cut value of the max-cut mapping W = -J: the sum of W_ij over the
     * pairs i < j with s_i != s_j, i.e. (sum_{i<j} W_ij - sum_{i<j} W_ij s_i s_j) / 2
     * (the field h does not enter).  Exact for integer-valued J.  May be NULL. */
    double *cut;
} sb_result;

typedef struct sb_ctx sb_ctx;
typedef struct sb_problem sb_problem;

/* ---- library ----------------------------------------------------------- */
SB_API const char *sb_version(void);
SB_API const char *sb_status_string(int status);
/* Detail of the last failure on the calling thread ("" if none). */
SB_API const char *sb_last_error_message(void);
/* Reference defaults: aSB, steps 100, dt 1.0, a0 1, c0 auto, K 1, M 2,
 * 1 replica, FP32, SB_INIT_REFERENCE, seed 0 (m2_opt.cu:10-22). */
SB_API int sb_default_params(sb_params *p);

/* ---- host-only instance helpers (no GPU needed) ------------------------ */
/* Parse an edge list (m2_opt.cu:111-123): first integer N, then "i j w"
 * triples; J[i][j] = J[j][i] = w (later lines overwrite).  Pass J = NULL to
 * read N only.  J is written densely as [N][ldj].  Weights may be real. */
SB_API int sb_edgelist_parse(const char *text, size_t len, int one_based, int32_t *N,
                             float *J, int32_t ldj);
/* sb_edgelist_parse that also accepts "i i w" lines as an external field
 * h_i (the king-graph generators' convention; the reference SB rejects them,
 * m2_opt.cu:119).  h holds N floats (zeroed first); *has_field is set when
 * any such line was read. */
SB_API int sb_edgelist_parse_field(const char *text, size_t len, int one_based, int32_t *N,
                                   float *J, int32_t ldj, float *h, int32_t *has_field);
/* sigma_J, integer flag and max |J_ij| of a dense symmetric J. */
SB_API int sb_instance_stats(int32_t N, const float *J, int32_t ldj, double *sigma,
                             int32_t *is_integer, double *max_abs);
/* Default c0: bSB/dSB 0.5/(sqrt(N) sigma_J) (Goto 2021); aSB
 * 0.7 a0/(sigma_J sqrt(N)) (m2_opt.cu:18 comment). */
SB_API double sb_c0_auto(int32_t variant, int32_t N, double a0, double sigma);
/* Seeded synthetic instances, dense [N][N] symmetric, zero diagonal:
 *   kind 0: Erdos-Renyi G(N, p), weights +-1 equiprobable, J = -W (max-cut)
 *   kind 1: Sherrington-Kirkpatrick, J_ij ~ N(0,1)/sqrt(N)
 *   kind 2: complete graph, J_ij = +-1 equiprobable
 *   kind 3: SK (N(0,1)/sqrt(N)) kept with probability p, else 0 */
SB_API int sb_generate(int32_t kind, int32_t N, double p, uint64_t seed, float *J);

/* ---- context and problems ---------------------------------------------- */
SB_API int sb_ctx_create(int32_t device, sb_ctx **out);
/* Fails with SB_E_INVALID while problems created on ctx are still alive. */
SB_API int sb_ctx_destroy(sb_ctx *ctx);
/* Upload a dense host J (FP32, [N][ldj]); stores FP32 and BF16 forms on the
 * device, and an FP8 E4M3 form when exact (integer |J| <= 16). */
SB_API int sb_problem_create(sb_ctx *ctx, int32_t N, const float *J, int32_t ldj,
                             sb_problem **out);
/* Reads the file with sb_edgelist_parse_field; "i i w" lines set the field. */
SB_API int sb_problem_from_edgelist_file(sb_ctx *ctx, const char *path, int one_based,
                                         sb_problem **out);
/* External field h (N host floats; NULL removes it).  The force on spin i
 * becomes (J.x)_i + h_i (J.sign(x) + h for dSB), added after the coupling
 * sum, and energies are H = -sum_{i<j} J_ij s_i s_j - sum_i h_i s_i (exact
 * int64 when J and h are integers).  sb_step and sb_energy on caller buffers
 * take no field. */
SB_API int sb_problem_set_field(sb_problem *prob, const float *h);
/* Sparse form (enable != 0): builds a CSR copy of J (rows in ascending column
 * order); sb_solve and sb_field with SB_PREC_FP32 or SB_PREC_SPLIT then use
 * the sparse family (a gather per nonzero on a spin-major state, FP32 fmaf
 * over the nonzeros in ascending column order: the same sums as
 * SB_PREC_ORDERED).  Pays off below a few percent density.  Batched solves and
 * the server take dense problems only (SB_E_INVALID).  enable = 0 frees it. */
SB_API int sb_problem_set_sparse(sb_problem *prob, int32_t enable);
/* Nonzeros of the sparse form (0 when the problem is dense only). */
SB_API int sb_problem_sparse_nnz(const sb_problem *prob, int64_t *nnz);
/* Generate a synthetic instance on the device (kinds as sb_generate; the
 * same J as sb_generate gives on the host, up to rare last-bit differences of
 * libm in the SK kinds). */
SB_API int sb_problem_generate(sb_ctx *ctx, int32_t kind, int32_t N, double p, uint64_t seed,
                               sb_problem **out);
/* Copy the problem's FP32 J to a host buffer [N][ldj]. */
SB_API int sb_problem_read_j(const sb_problem *prob, float *J, int32_t ldj);
SB_API int sb_problem_destroy(sb_problem *prob);
SB_API int sb_problem_info(const sb_problem *prob, int32_t *N, int32_t *is_integer,
                           double *sigma, int32_t *split_jfmt);

/* ---- solve (replaces the TIMES x STEP loop, m2_opt.cu:152-202) --------- */
SB_API int sb_solve(sb_ctx *ctx, sb_problem *prob, const sb_params *params,
                    sb_result *result);

/* ---- low-level operator (replaces UpdateTwice + cublasSgemv) ----------- */
/* Coefficients of one fused step (see sb_step_coeffs_for). */
typedef struct sb_step_coeffs {
    int32_t variant;
    int32_t do_local; /* aSB, aSB-m1: apply the next step's local (x) update after the kick */
    int32_t M;
    int32_t reserved;
    float alpha, beta, gamma;
    float p_next;     /* aSB: pump of the next step's sub-steps; aSB-m1: pump of this y-step */
    double a0, K, dts;
    double c0;        /* aSB-m1: the coupling xi, applied in FP64 */
} sb_step_coeffs;

/* Fill the coefficients of fused step k of an S-step schedule. */
SB_API int sb_step_coeffs_for(const sb_params *p, double c0, int32_t k,
                              sb_step_coeffs *out);

/* One fused step on device buffers: F = J.X (or J.sign(X) for dSB), then the
 * y/x update, wall and (aSB) local sub-steps; reads X, writes Xn, updates Y
 * in place.  X, Xn, Y are FP32 [R][ldx]; J is [N][ldj] in jfmt; ldj, ldx are
 * multiples of 64 and >= N; spins N..ldx-1 of X must be 0. */
SB_API int sb_step(sb_ctx *ctx, int32_t precision, const void *dJ, int32_t jfmt,
                   int32_t N, int32_t ldj, const float *dX, float *dXn, float *dY,
                   int32_t R, int32_t ldx, const sb_step_coeffs *c, void *stream);

/* Local field of given amplitudes through the product kernels: F = J.X
 * (variant SB_BSB/SB_ASB/SB_ASB_M1) or J.sign(X) (SB_DSB), plus h when the
 * problem has an external field; host X and F [R][N], with the arithmetic of
 * `precision` (FP32 MFMA, BF16 split, ordered; R <= 4 uses the thin kernels).
 * The reference's field alone (cublasSgemv, m2_opt.cu:169); used to report
 * field errors against an FP64 trajectory. */
SB_API int sb_field(sb_ctx *ctx, sb_problem *prob, int32_t precision, int32_t variant, const float *X,
                    int32_t R, float *F);

/* Initial state on device buffers ([R][ldx], padding zeroed). */
SB_API int sb_init_state(sb_ctx *ctx, int32_t init, uint64_t seed, int32_t N, int32_t R,
                         int32_t ldx, float *dX, float *dY, void *stream);
/* aSB only: the local sub-steps of step 0, applied before the first sb_step. */
SB_API int sb_asb_prologue(sb_ctx *ctx, const sb_params *p, int32_t N, int32_t R,
                           int32_t ldx, float *dX, float *dY, void *stream);

/* H_r = -sum_{i<j} J_ij s_i s_j for s = sign(X) on device buffers; dH is a
 * device FP64 array [R].  Exact for integer-valued J (|sum_j J_ij| < 2^52). */
SB_API int sb_energy(sb_ctx *ctx, const float *dJ, int32_t N, int32_t ldj,
                     const float *dX, int32_t R, int32_t ldx, double *dH, void *stream);

/* sb_energy that allocates nothing, for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  No counterpart in the reference.  The partial sums go to a
 * static device array of 2^20 doubles (8 MiB) that each GPU holds for the
 * life of the process; calls with ceil(N / 64) * R above 2^20 are rejected
 * (SB_E_INVALID).  Every call on one GPU shares that array: calls that may
 * overlap on the device (different streams, or contexts on the same GPU)
 * race and must be ordered by the caller. */
SB_API int sb_energy_noalloc(sb_ctx *ctx, const float *dJ, int32_t N, int32_t ldj, const float *dX,
                             int32_t R, int32_t ldx, double *dH, void *stream);

/* ---- batched solve of many instances (one launch per step) ------------- */
typedef struct sb_batch_stats {
    int32_t work_items;   /* tiles per step across all instances */
    int32_t workgroups;   /* persistent grid size                */
    int32_t graph_used;   /* 1 if the step loop was a replayed hipGraph */
    int32_t reserved;
    double step_ms;       /* device time of the step loop        */
    double total_ms;
} sb_batch_stats;

/* Solve n independent problems; params[i] / results[i] per problem.  All
 * problems must share precision; steps, N, R, variant may differ.  Each
 * problem's result is identical to sb_solve on it alone. */
SB_API int sb_batch_solve(sb_ctx *ctx, int32_t n, sb_problem *const *probs,
                          const sb_params *params, sb_result *results,
                          sb_batch_stats *stats);

/* ---- serving (continuous batching) -------------------------------------- */
typedef struct sb_server sb_server;

/* Queue order of the server.  Resident requests are admitted while their J
 * and state fit the HBM budget and the resident set stays within a tile
 * budget of about two work items per workgroup of the persistent grid (the
 * device is then busy; more resident requests would only share it). */
typedef enum sb_serve_policy {
    SB_SERVE_SHORTEST = 0,  /* least total work (N^2 R steps) first; ties by arrival */
    SB_SERVE_ARRIVAL = 1    /* arrival order (smaller requests backfill the HBM budget) */
} sb_serve_policy;

typedef struct sb_server_opts {
    size_t hbm_budget_bytes;  /* device bytes for J + state of admitted requests */
    int32_t chunk_steps;      /* steps run between admission points (>= 1)     */
    int32_t max_active;       /* cap on concurrently resident requests (0: none) */
    int32_t precision;        /* sb_precision for all requests                   */
    int32_t policy;           /* sb_serve_policy                                 */
    double sparse_density;    /* requests with at most this fraction of nonzero
                                 couplings run the sparse (CSR) family; 0: never */
} sb_server_opts;

typedef struct sb_server_stats {
    int64_t submitted, completed, rejected, cancelled;
    int64_t active, queued;
    size_t bytes_in_use;
    double busy_ms;           /* device time spent in step launches */
    int64_t chunks;           /* chunks (graph replays) run            */
    int64_t max_active_seen;  /* most requests resident at once        */
} sb_server_stats;

SB_API int sb_server_create(sb_ctx *ctx, const sb_server_opts *opts, sb_server **out);
/* Copies J (dense FP32 host [N][ldj]) and params; returns a ticket.  Rejects
 * with SB_E_BUDGET when the request alone exceeds the budget. */
SB_API int sb_server_submit(sb_server *srv, int32_t N, const float *J, int32_t ldj,
                            const sb_params *params, uint64_t *ticket);
/* Submit a synthetic instance (sb_generate kinds) that the server generates
 * on the device at admission; no host copy of J is kept. */
SB_API int sb_server_submit_generated(sb_server *srv, int32_t kind, int32_t N, double p, uint64_t seed,
                                      const sb_params *params, uint64_t *ticket);
/* Wait for a ticket (timeout_ms < 0: forever).  Fills energy_f64 / energy_i64
 * / spins of result (x_final, y_final are not kept by the server) and the
 * best replica, then frees the ticket.  result->total_ms is the request's
 * latency from submission to completion. */
SB_API int sb_server_wait(sb_server *srv, uint64_t ticket, sb_result *result,
                          double timeout_ms);
SB_API int sb_server_cancel(sb_server *srv, uint64_t ticket);
SB_API int sb_server_stats_get(sb_server *srv, sb_server_stats *out);
SB_API int sb_server_destroy(sb_server *srv);

#ifdef __cplusplus
}
#endif
#endif
