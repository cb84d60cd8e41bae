// sb_internal.h -- types shared by the gfx950 kernels (sb_kernels.hip) and the
// host library (sb_api.cpp, sb_batch.cpp, sb_server.cpp).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <cmath>

#include "../../include/sb.h"

namespace sbk {

// Spin padding of every device row (J rows and X/Y replica rows): all K loops
// read up to round_up(N, kSpinPad) and rely on zeros there.
constexpr int kSpinPad = 64;

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Arguments of one fused step over one instance.
// B-operand planes: the sign of X in the MFMA operand format, written once per
// step by the previous step's epilogue (or k_make_planes after init), so
// staging it is a plain copy of 1-2 bytes per element instead of 4.  Layout
// [R][ldx].  The split bSB/aSB operand (BF16 hi and lo of x) is not kept as
// planes: the K loop reads FP32 X (the same 4 bytes per element as hi + lo)
// and splits it while staging, so the epilogue writes 16 bytes per element
// instead of 20.
enum BKind {
    BK_NONE = 0,      // B staged from X
    BK_SIGN_FP8 = 2,  // E4M3 +-1 bytes
    BK_SIGN_BF16 = 3  // BF16 +-1
};

struct StepArgs {
    const void *J;   // [N][ldj] in the kernel's J format
    const float *X;  // [R][ldx] state read this step
    float *Xn;       // [R][ldx] state written this step
    float *Y;        // [R][ldx] updated in place
    const float *h;  // external field [N] added to J.x (null: none)
    const void *B;   // B-operand planes of X (bkind != BK_NONE)
    void *Bn;        // planes of Xn written by the epilogue (may be null)
    int bkind;
    int ldj, ldx;
    int N, R;
    int tiles_m, tiles_n;  // row tiles x replica tiles of this launch
    int rpg;               // row tiles per XCD group of the tile order (0: all)
    sb_step_coeffs c;
};

// Row tiles per group of the XCD-blocked tile order: each of the 8 XCDs runs
// ~T/8 consecutive logical tiles; a block of g rows x (T/8)/g columns needs
// g*MT*bytesJ + (T/8/g)*NT*bytesB bytes per K, minimal at
// g = sqrt(T/8 * NT*bytesB / (MT*bytesJ)).
int rows_per_group(int tiles_m, int tiles_n, int mt, int nt, int bytes_j, int bytes_b);

// Plane kind a family reads for a variant (BK_NONE: stages from X).
inline int plane_kind(int family, int variant)
{
    if (family == 1 || family == 2) return variant == SB_DSB ? BK_SIGN_BF16 : BK_NONE;
    if (family == 3) return BK_SIGN_FP8;
    return BK_NONE;
}
// Bytes per state element of a plane set of this kind.
inline size_t plane_bytes(int kind)
{
    return kind == BK_SIGN_BF16 ? 2 : kind == BK_SIGN_FP8 ? 1 : 0;
}

// Tile shapes.  A workgroup has WM x WN waves; each wave owns BM x BN blocks
// of 32 x 32 (spins x replicas).  Every shape accumulates each F_i in the same
// k order, so the choice never changes results.
// Replica counts up to this run on the thin (matrix-vector) families.
constexpr int kThinMaxR = 4;

// MFMA tile: wm x wn waves, each bm x bn blocks of 32x32.  Thin tile (rows >
// 0): wm waves, each owning `rows` spins for all R <= kThinMaxR replicas.
struct TileShape {
    int wm, wn, bm, bn;
    int rows;
    int mt() const { return rows ? wm * rows : 32 * bm * wm; }
    int nt() const { return rows ? kThinMaxR : 32 * bn * wn; }
    int threads() const { return 64 * wm * wn; }
};

// Kernel families.
enum Family {
    FAM_FP32 = 0,       // FP32 J x FP32 X on v_mfma_f32_32x32x2_f32
    FAM_SPLIT_BF16 = 1, // BF16 J x BF16 hi/lo X (or BF16 sign X for dSB)
    FAM_SPLIT_FP8 = 2,  // FP8 J widened to BF16 x BF16 hi/lo X
    FAM_DSB_FP8 = 3,    // FP8 J x FP8 sign(X) on the f8f6f4 MFMA (dSB only)
    FAM_THIN_F32 = 4,   // R <= kThinMaxR: FP32 J streamed, FP32 FMA on the VALU
    FAM_THIN_BF16 = 5,  // same with BF16 J
    FAM_THIN_FP8 = 6,   // same with FP8 J
    FAM_SPARSE = 7,     // CSR J gathered over a spin-major state (batched launches)
    FAM_COUNT = 8
};
inline bool is_thin(int family) { return family >= FAM_THIN_F32 && family <= FAM_THIN_FP8; }

// Number of compiled tile shapes per family and their table.
int shape_count(int family);
TileShape shape_at(int family, int idx);

// Launchers (sb_kernels.hip).  Return a hipError_t.
hipError_t launch_step(int family, int shape_idx, const StepArgs &a, hipStream_t s);
hipError_t launch_step_ordered(const StepArgs &a, hipStream_t s);

// Sparse family: CSR J (rows in ascending column order) and a spin-major state
// XT, YT [N][ldr] (ldr = round_up(R, 64); replicas of one spin contiguous).
struct SparseArgs {
    const int *ptr;    // [N + 1] row starts
    const int *col;    // [nnz] column indices
    const float *w;    // [nnz] values
    const float *h;    // external field [N] or null
    const float *XT;   // state read this step
    float *XTn;        // state written this step
    float *YT;         // updated in place
    int N, R, ldr;
    sb_step_coeffs c;
};
hipError_t launch_step_sparse(const SparseArgs &a, hipStream_t s);
// CSR of a dense FP32 J [N][ld] on the device: nonzeros per row into cnt [N];
// then, with ptr [N + 1] their prefix sums, columns and values per row in
// ascending column order.
hipError_t launch_row_nnz(const float *J, int N, int ld, int *cnt, hipStream_t s);
hipError_t launch_csr_fill(const float *J, int N, int ld, const int *ptr, int *col, float *w, hipStream_t s);
// [R][ldx] replica-major <-> [N][ldr] spin-major (rows/columns past R or N untouched).
hipError_t launch_to_spin_major(const float *X, int ldx, int N, int R, float *XT, int ldr, hipStream_t s);
hipError_t launch_to_replica_major(const float *XT, int ldr, int N, int R, float *X, int ldx, hipStream_t s);
// Planes of kind `kind` for state X ([R][ldx]) into B.
hipError_t launch_make_planes(int kind, const float *X, int R, int ldx, void *B, hipStream_t s);
hipError_t launch_init(int kind, uint64_t seed, int N, int R, int ldx, float *X, float *Y,
                       float *X2, hipStream_t s);
hipError_t launch_asb_local(const sb_step_coeffs &c, int N, int R, int ldx, float *X, float *Y,
                            hipStream_t s);
// Coupling energy Hc_r = -sum_{i<j} J_ij s_i s_j of s = sign(X) in FP64
// (exact for integer J below 2^53); `partial` holds energy_partial_elems
// doubles.
hipError_t launch_energy(const float *J, int ldj, const float *X, int ldx, int N, int R, double *partial,
                         double *H, hipStream_t s);
// The same for integer-valued J with |J_ij| <= 127 as an exact integer
// reduction (int8 MFMA, int32 accumulators, int64 sums): Hi (int64) and H
// (its FP64 value); `partial` holds energy_partial_elems int64.  ldj and ldx
// are multiples of 64.
hipError_t launch_energy_i8(const float *J, int ldj, const float *X, int ldx, int N, int R, long long *partial,
                            long long *Hi, double *H, hipStream_t s);
// Field part Hf_r = -sum_i h_i s_i.
hipError_t launch_energy_field(const float *h, const float *X, int ldx, int N, int R, double *Hf, hipStream_t s);
size_t energy_partial_elems(int N, int R);
// The static device scratch of sb_energy_noalloc: its capacity in doubles, and
// its address on `device` (which must be current; looked up once per device).
size_t noalloc_scratch_elems();
hipError_t noalloc_scratch_address(int device, double **ptr);
hipError_t launch_spins(const float *X, int ldx, int N, int R, int8_t *S, hipStream_t s);
hipError_t launch_compact(const float *X, int ldx, int N, int R, float *out, hipStream_t s);
hipError_t launch_pack_bf16(const float *J, size_t n_elems, uint16_t *out, hipStream_t s);
hipError_t launch_pack_fp8(const float *J, size_t n_elems, uint8_t *out, hipStream_t s);

// ---- batched launch ------------------------------------------------------
// Many instances advance one step per launch.  A work item is one tile
// (row tile x replica tile) of one instance; each workgroup of the persistent
// grid walks its own item list.  An instance's step index in a launch is
// step_base + chunk_step, so one captured chunk of launches can be replayed
// while the host moves step_base (and the instance set) between replays.
struct BatchInstance {
    const void *J;      // J in the family's format, [N][ldx]
    const float *h;     // external field [N] or null
    float *X[2];        // ping-pong state, read X[k & 1], write X[(k + 1) & 1]
    float *Y;
    void *P[2];         // B-operand sign planes (dSB), same ping-pong
    int bkind;
    int ldx;
    int N, R;
    int family, bshape;  // kernel family and batch tile shape
    int tiles_m, tiles_n;
    int steps;           // schedule length S
    int step_base;       // absolute step of chunk_step 0
    int active;          // 0: skip (slot free or finished)
    int variant, M;
    double dt, a0, c0, K;
    // Sparse family: CSR J and the spin-major state ([N][ldr], ldr = round_up(R, 64)).
    const int *csr_ptr, *csr_col;
    const float *csr_w;
    float *XT[2];
    float *YT;
    int ldr;
};

struct BatchItem {
    int inst;  // index into the instance table
    int tile;  // tile index within the instance (tm + tiles_m * tn)
};

// Tile shapes the batched kernel compiles per family (4 waves, <= 74 KB LDS).
int batch_shape_count(int family);
TileShape batch_shape_at(int family, int idx);
size_t batch_lds_bytes();  // dynamic LDS of every batched launch
// Index of a batch shape in the single-instance table of the same family.
int batch_shape_solo_index(int family, int idx);

// One persistent launch of `grid` workgroups over the item lists.
hipError_t launch_batch_step(int grid, const BatchInstance *insts, const BatchItem *items, const int *wg_begin,
                             int chunk_step, hipStream_t s);
// Workgroups of the batched kernel that fit on the device at once.
int batch_grid_size(int device);

// ---- synthetic instances ----------------------------------------------------
// J_ij (i < j) is a function of (seed, i, j) only: a splitmix64 finaliser of
// the pair key, so host and device generate the same instance in any order.
// (SK values go through log/cos/sqrt; host and device libm may round those
// differently in rare last bits, so read J back when the host needs it.)
#ifdef __HIPCC__
#define SB_HD __host__ __device__
#else
#define SB_HD
#endif

SB_HD inline uint64_t gen_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

SB_HD inline uint64_t gen_hash(uint64_t seed, uint32_t i, uint32_t j, uint32_t c)
{
    uint64_t key = ((uint64_t)i << 32) | (uint64_t)j;
    return gen_mix(seed + 0x9E3779B97F4A7C15ull * (4ull * key + (uint64_t)c + 1ull));
}

SB_HD inline double gen_u01(uint64_t h) { return (double)(h >> 11) * 0x1p-53; }

// kind: 0 ER(p) +-1 with J = -W, 1 SK N(0,1)/sqrt(N), 2 K_N +-1, 3 SK kept with prob. p
SB_HD inline float gen_coupling(int kind, int N, double p, uint64_t seed, uint32_t i, uint32_t j)
{
    switch (kind) {
    case 0:
        if (gen_u01(gen_hash(seed, i, j, 0)) < p) return (gen_hash(seed, i, j, 1) & 1) ? 1.0f : -1.0f;
        return 0.0f;
    case 2: return (gen_hash(seed, i, j, 0) & 1) ? 1.0f : -1.0f;
    case 1:
    case 3: {
        if (kind == 3 && !(gen_u01(gen_hash(seed, i, j, 2)) < p)) return 0.0f;
        double u1 = gen_u01(gen_hash(seed, i, j, 0)), u2 = gen_u01(gen_hash(seed, i, j, 1));
        if (u1 < 0x1p-60) u1 = 0x1p-60;
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        return (float)(z / sqrt((double)N));
    }
    default: return 0.0f;
    }
}

// Device generation of a dense J [N][ld] (zero diagonal and padding), and
// per-row statistics: sum of squares and sum right of the diagonal (FP64,
// row_ss holds 2 N doubles), max |J|, integer flag.
hipError_t launch_generate(int kind, int N, double p, uint64_t seed, float *J, int ld, hipStream_t s);
hipError_t launch_row_stats(const float *J, int N, int ld, double *row_ss, float *row_max, int *row_int,
                            hipStream_t s);

// Step coefficients (identical arithmetic on host and device).
#ifdef __HIPCC__
__host__ __device__
#endif
inline sb_step_coeffs step_coeffs(int variant, int k, int steps, double dt, double a0, double c0, double K, int M)
{
    sb_step_coeffs c{};
    c.variant = variant;
    c.M = M;
    c.a0 = a0;
    c.K = K;
    c.dts = M > 0 ? dt / (double)M : 0.0;
    if (variant == SB_ASB_M1) {
        // Bifurcation_m1.cu.c:166-169: step k is UpdateX, then UpdateY with
        // amplitude k/(float)STEP; the fused kernel applies the y-step of step
        // k and then the x-step of step k+1 (a prologue does step 0's).
        c.dts = dt;
        c.c0 = c0;
        c.p_next = (float)k / (float)steps;
        c.do_local = (k + 1 < steps) ? 1 : 0;
    } else if (variant == SB_ASB) {
        c.beta = (float)(dt * c0);  // m2_opt.cu:148 coeff1 = DeltaT * xi
        c.do_local = (k + 1 < steps) ? 1 : 0;
        c.p_next = (k + 1 < steps) ? (float)(k + 1) / (float)steps : 0.0f;  // m2_opt.cu:158
    } else {
        double pk = (double)k / (double)steps;
        double a = a0 * pk;
        c.alpha = (float)(dt * (a0 - a));
        c.beta = (float)(dt * c0);
        c.gamma = (float)(dt * a0);
        c.p_next = (float)pk;
    }
    return c;
}

}  // namespace sbk
