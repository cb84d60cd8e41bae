// sb_kernels.hip -- kernel dispatch (shape tables, launch_step) and the
// non-MFMA kernels of the simulated-bifurcation solver: ordered and sparse
// step kernels, state transposes, instance generation, init, planes,
// energies, readout and J packing.  The MFMA step kernels live in
// sb_step_*.hip, the batched kernel in sb_batch_kernel.hip; all share the
// device building blocks of sb_tiles.h.
#include "sb_tiles.h"

#include <map>
#include <mutex>

namespace sbk {

// Kernels indexed by replica in blockIdx.y stride over replicas, so R is not
// bounded by the 65535 limit of a grid's y dimension.
constexpr int kMaxGridY = 65535;
static inline unsigned grid_y(int R) { return (unsigned)(R < kMaxGridY ? R : kMaxGridY); }

// Ordered-sum debug kernel: one thread per (spin, replica), F_i summed over
// j = 0..N-1 with fmaf in order (Bifurcation_m2.cu.c:93-97).
__global__ __launch_bounds__(256) void k_step_ordered(StepArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const float *Jr = (const float *)a.J + (size_t)i * a.ldj;
    const bool sgn = a.c.variant == SB_DSB;
    for (int r = blockIdx.y; r < a.R; r += gridDim.y) {
        const float *xr = a.X + (size_t)r * a.ldx;
        float F = 0.0f;
        for (int j = 0; j < a.N; ++j) {
            float v = xr[j];
            if (sgn) v = sgn1(v);
            F = __builtin_fmaf(Jr[j], v, F);
        }
        if (a.h) F = F + a.h[i];  // external field, after the sum
        float x = xr[i], y = a.Y[(size_t)r * a.ldx + i];
        sb_update(a.c, F, x, y);
        a.Xn[(size_t)r * a.ldx + i] = x;
        a.Y[(size_t)r * a.ldx + i] = y;
    }
}

#define SB_SHAPE_ENTRY(wm, wn, bm, bn) TileShape{wm, wn, bm, bn, 0},
static const TileShape kFp32Shapes[] = {SB_FP32_SHAPES(SB_SHAPE_ENTRY)};
static const TileShape kSplitShapes[] = {SB_SPLIT_SHAPES(SB_SHAPE_ENTRY)};
static const TileShape kFp8Shapes[] = {SB_FP8_SHAPES(SB_SHAPE_ENTRY)};
static const TileShape kThinShapes[] = {SB_THIN_SHAPES(SB_THIN_ENTRY)};
#undef SB_SHAPE_ENTRY

int shape_count(int family)
{
    switch (family) {
    case FAM_FP32: return (int)(sizeof(kFp32Shapes) / sizeof(kFp32Shapes[0]));
    case FAM_SPLIT_BF16:
    case FAM_SPLIT_FP8: return (int)(sizeof(kSplitShapes) / sizeof(kSplitShapes[0]));
    case FAM_DSB_FP8: return (int)(sizeof(kFp8Shapes) / sizeof(kFp8Shapes[0]));
    case FAM_THIN_F32:
    case FAM_THIN_BF16:
    case FAM_THIN_FP8: return (int)(sizeof(kThinShapes) / sizeof(kThinShapes[0]));
    default: return 0;
    }
}

TileShape shape_at(int family, int idx)
{
    switch (family) {
    case FAM_FP32: return kFp32Shapes[idx];
    case FAM_SPLIT_BF16:
    case FAM_SPLIT_FP8: return kSplitShapes[idx];
    case FAM_THIN_F32:
    case FAM_THIN_BF16:
    case FAM_THIN_FP8: return kThinShapes[idx];
    default: return kFp8Shapes[idx];
    }
}

int rows_per_group(int tiles_m, int tiles_n, int mt, int nt, int bytes_j, int bytes_b)
{
    double per_xcd = (double)tiles_m * tiles_n / 8.0;
    double g = std::sqrt(per_xcd * nt * bytes_b / ((double)mt * bytes_j));
    int r = (int)(g + 0.5);
    if (r < 1) r = 1;
    if (r > tiles_m) r = tiles_m;
    return r;
}

hipError_t launch_step(int family, int shape_idx, const StepArgs &a_in, hipStream_t s)
{
    if (shape_idx < 0 || shape_idx >= shape_count(family)) return hipErrorInvalidValue;
    // A family whose form of J the problem does not hold is never launched.
    if (!a_in.J || !a_in.X || !a_in.Xn || !a_in.Y) return hipErrorInvalidValue;
    TileShape sh = shape_at(family, shape_idx);
    StepArgs a = a_in;
    if (is_thin(family)) {
        if (a.R > kThinMaxR) return hipErrorInvalidValue;
        a.tiles_m = (a.N + sh.mt() - 1) / sh.mt();
        a.tiles_n = 1;
        a.rpg = 1;
        StepKernel k = kThinKernels[family - FAM_THIN_F32][a.c.variant == SB_DSB][thin_rm_index(a.R)][shape_idx];
        hipLaunchKernelGGL(k, dim3(a.tiles_m), dim3(sh.threads()), 0, s, a);
        return hipGetLastError();
    }
    a.tiles_m = (a.N + sh.mt() - 1) / sh.mt();
    a.tiles_n = (a.R + sh.nt() - 1) / sh.nt();
    int grid = a.tiles_m * a.tiles_n;
    if (grid <= 0) return hipSuccess;
    // Planes are read only when the caller supplies ones of the family's kind.
    const bool pl = a.B != nullptr && a.bkind == plane_kind(family, a.c.variant) && a.bkind != BK_NONE;
    if (a.Bn && a.bkind != plane_kind(family, a.c.variant)) return hipErrorInvalidValue;
    const int bytes_j = family == FAM_FP32 ? 4 : family == FAM_SPLIT_BF16 ? 2 : 1;
    const int bytes_b = pl ? (int)plane_bytes(a.bkind) : 4;
    a.rpg = rows_per_group(a.tiles_m, a.tiles_n, sh.mt(), sh.nt(), bytes_j, bytes_b);
    StepKernel k = nullptr;
    switch (family) {
    case FAM_FP32: k = kFp32Kernels[shape_idx]; break;
    case FAM_SPLIT_BF16:
    case FAM_SPLIT_FP8:
        k = (family == FAM_SPLIT_FP8 ? kSplitFp8Kernels
                                     : kSplitBf16Kernels)[a.c.variant != SB_DSB ? 0 : pl ? 2 : 1][shape_idx];
        break;
    case FAM_DSB_FP8: k = kDsbFp8Kernels[pl ? 1 : 0][shape_idx]; break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(sh.threads()), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sparse family: one workgroup per 4 spins x 64 replicas (sparse_tile).

__global__ __launch_bounds__(256) void k_step_sparse(SparseArgs a) { sparse_tile(a, blockIdx.x, blockIdx.y); }

hipError_t launch_step_sparse(const SparseArgs &a, hipStream_t s)
{
    if (a.N <= 0 || a.R <= 0) return hipSuccess;
    if (a.ldr % 64 || a.ldr < a.R) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_step_sparse, dim3((a.N + 3) / 4, a.ldr / 64), dim3(256), 0, s, a);
    return hipGetLastError();
}

// 32 x 32 tiles through LDS (one side of each copy coalesced on each pass).
__global__ __launch_bounds__(256) void k_transpose(const float *src, int lds_, int rows, int cols, float *dst,
                                                   int ldd)
{
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) t[k][tx] = src[(size_t)r * lds_ + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;  // dst row c, column r
        if (r < rows && c < cols) dst[(size_t)c * ldd + r] = t[tx][k];
    }
}

hipError_t launch_to_spin_major(const float *X, int ldx, int N, int R, float *XT, int ldr, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_transpose, dim3((N + 31) / 32, (R + 31) / 32), dim3(256), 0, s, X, ldx, R, N, XT, ldr);
    return hipGetLastError();
}

hipError_t launch_to_replica_major(const float *XT, int ldr, int N, int R, float *X, int ldx, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_transpose, dim3((R + 31) / 32, (N + 31) / 32), dim3(256), 0, s, XT, ldr, N, R, X, ldx);
    return hipGetLastError();
}

hipError_t launch_step_ordered(const StepArgs &a, hipStream_t s)
{
    if (a.N <= 0 || a.R <= 0) return hipSuccess;
    dim3 grid((a.N + 255) / 256, grid_y(a.R));
    hipLaunchKernelGGL(k_step_ordered, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Synthetic instances on the device (gen_coupling in sb_internal.h).

__global__ __launch_bounds__(256) void k_generate(int kind, int N, double p, uint64_t seed, float *J, int ld)
{
    for (int i = blockIdx.y; i < N; i += gridDim.y)
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ld; j += gridDim.x * blockDim.x) {
            float v = 0.0f;
            if (j < N && j != i)
                v = i < j ? gen_coupling(kind, N, p, seed, (uint32_t)i, (uint32_t)j)
                          : gen_coupling(kind, N, p, seed, (uint32_t)j, (uint32_t)i);
            J[(size_t)i * ld + j] = v;
        }
}

hipError_t launch_generate(int kind, int N, double p, uint64_t seed, float *J, int ld, hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    int gx = (ld + 255) / 256;
    hipLaunchKernelGGL(k_generate, dim3(gx < 16 ? gx : 16, grid_y(N)), dim3(256), 0, s, kind, N, p, seed, J, ld);
    return hipGetLastError();
}

// One workgroup per row: sum of squares and sum of the entries right of the
// diagonal (FP64, fixed tree), max |J|, and whether every entry is an
// integer.  row_ss holds 2 N doubles: the sums of squares, then the upper sums.
__global__ __launch_bounds__(256) void k_row_stats(const float *J, int N, int ld, double *row_ss, float *row_max,
                                                   int *row_int)
{
    __shared__ double ss[256], su[256];
    __shared__ float mx[256];
    __shared__ int in[256];
    const int i = blockIdx.x, t = threadIdx.x;
    double s = 0.0, u = 0.0;
    float m = 0.0f;
    int integ = 1;
    for (int j = t; j < N; j += 256) {
        float v = J[(size_t)i * ld + j];
        s += (double)v * (double)v;
        if (j > i) u += (double)v;
        m = fmaxf(m, fabsf(v));
        if (v != floorf(v)) integ = 0;
    }
    ss[t] = s;
    su[t] = u;
    mx[t] = m;
    in[t] = integ;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            su[t] += su[t + w];
            mx[t] = fmaxf(mx[t], mx[t + w]);
            in[t] &= in[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        row_ss[i] = ss[0];
        row_ss[N + i] = su[0];
        row_max[i] = mx[0];
        row_int[i] = in[0];
    }
}

hipError_t launch_row_stats(const float *J, int N, int ld, double *row_ss, float *row_max, int *row_int,
                            hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_stats, dim3(N), dim3(256), 0, s, J, N, ld, row_ss, row_max, row_int);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_make_planes(int kind, const float *X, size_t n4, void *B)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (size_t)gridDim.x * blockDim.x) {
        v4f xv = *(const v4f *)(X + 4 * q);
        write_planes(kind, xv, B, 4 * q);
    }
}

hipError_t launch_make_planes(int kind, const float *X, int R, int ldx, void *B, hipStream_t s)
{
    if (kind == BK_NONE || R <= 0) return hipSuccess;
    size_t n = (size_t)R * ldx, n4 = n / 4;
    size_t g = (n4 + 255) / 256;
    hipLaunchKernelGGL(k_make_planes, dim3((unsigned)(g < 65536 ? g : 65536)), dim3(256), 0, s, kind, X, n4, B);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Initial state.

__device__ __forceinline__ uint32_t xorshift32(uint32_t &st)
{
    uint32_t v = st;
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    st = v;
    return v;
}

// prepare_points value (m2_opt.cu:54-55): (float)v / 2^32 exactly, / 10.0 in FP64.
__device__ __forceinline__ float ref_draw_value(uint32_t v)
{
    float q = (float)v * 0x1p-32f;
    return (float)((double)q / 10.0);
}

__device__ __forceinline__ uint32_t keyed_u24(uint64_t seed, uint32_t replica, uint32_t spin,
                                              uint32_t comp)
{
    uint64_t key = ((uint64_t)replica << 32) | (uint64_t)spin;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (2ull * key + (uint64_t)comp + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (uint32_t)(z >> 40);
}

__device__ __forceinline__ float keyed_value(uint32_t u24)
{
    float v = (float)u24 * 0x1p-23f;
    v = v - 1.0f;
    return v * 0.1f;
}

__global__ __launch_bounds__(256) void k_init_uniform(uint64_t seed, int N, int R, int ldx, float *X,
                                                      float *Y, float *X2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ldx) return;
    for (int r = blockIdx.y; r < R; r += gridDim.y) {
        size_t o = (size_t)r * ldx + i;
        float x = 0.0f, y = 0.0f;
        if (i < N) {
            x = keyed_value(keyed_u24(seed, (uint32_t)r, (uint32_t)i, 0u));
            y = keyed_value(keyed_u24(seed, (uint32_t)r, (uint32_t)i, 1u));
        }
        X[o] = x;
        Y[o] = y;
        if (X2) X2[o] = 0.0f;
    }
}

// One thread per spin walks its xorshift32 stream over the replicas, as the
// reference's trial loop does (m2_opt.cu:136-137,152-154).
__global__ __launch_bounds__(256) void k_init_reference(uint32_t seed, int N, int R, int ldx, float *X,
                                                        float *Y, float *X2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ldx) return;
    uint32_t st = (uint32_t)i + seed * (uint32_t)N;
    for (int r = 0; r < R; ++r) {
        size_t o = (size_t)r * ldx + i;
        float x = 0.0f, y = 0.0f;
        if (i < N) {
            x = ref_draw_value(xorshift32(st));
            y = ref_draw_value(xorshift32(st));
        }
        X[o] = x;
        Y[o] = y;
        if (X2) X2[o] = 0.0f;
    }
}

hipError_t launch_init(int kind, uint64_t seed, int N, int R, int ldx, float *X, float *Y, float *X2,
                       hipStream_t s)
{
    if (R <= 0 || ldx <= 0) return hipSuccess;
    if (kind == SB_INIT_REFERENCE) {
        hipLaunchKernelGGL(k_init_reference, dim3((ldx + 255) / 256), dim3(256), 0, s, (uint32_t)seed, N, R,
                           ldx, X, Y, X2);
    } else {
        hipLaunchKernelGGL(k_init_uniform, dim3((ldx + 255) / 256, grid_y(R)), dim3(256), 0, s, seed, N, R, ldx,
                           X, Y, X2);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_asb_local(sb_step_coeffs c, int N, int R, int ldx, float *X,
                                                   float *Y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int r = blockIdx.y; r < R; r += gridDim.y) {
        size_t o = (size_t)r * ldx + i;
        float x = X[o], y = Y[o];
        if (c.variant == SB_ASB_M1)
            x = m1_x(x, y, c.a0, c.dts);  // step 0's x-step
        else
            asb_local(x, y, c.p_next, c.a0, c.K, c.dts, c.M);
        X[o] = x;
        Y[o] = y;
    }
}

hipError_t launch_asb_local(const sb_step_coeffs &c, int N, int R, int ldx, float *X, float *Y,
                            hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_asb_local, dim3((N + 255) / 256, grid_y(R)), dim3(256), 0, s, c, N, R, ldx, X, Y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Energies.  The coupling part Hc_r = -sum_{i<j} J_ij s_i s_j = -1/2 sum_i s_i G_i
// with G = J s, s = x > 0 ? +1 : -1, and the field part -sum_i h_i s_i are
// computed separately (the cut value of the max-cut mapping needs Hc alone).
//
// FP64 form (any J): tile 64 spins x 64 replicas per workgroup; every
// product is +-J_ij (exact), so for integer J all partial sums are exact
// integers below 2^53.  Deterministic: fixed per-thread order and a fixed
// reduction tree.

constexpr int EN_T = 64, EN_KC = 32;

__global__ __launch_bounds__(256) void k_energy_partial(const float *J, int ldj, const float *X, int ldx,
                                                        int N, int R, double *part)
{
    __shared__ float Js[EN_KC][EN_T + 1];
    __shared__ float Ss[EN_KC][EN_T + 1];
    __shared__ double red[16][EN_T + 1];
    const int tiles_m = (N + EN_T - 1) / EN_T;
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int row0 = tm * EN_T, col0 = tn * EN_T;
    const int tid = threadIdx.x, ti = tid & 15, tr = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int j0 = 0; j0 < N; j0 += EN_KC) {
        for (int q = tid; q < EN_T * EN_KC; q += 256) {
            int rr = q / EN_KC, c = q % EN_KC;
            int gi = row0 + rr, gj = j0 + c;
            Js[c][rr] = (gi < N && gj < N) ? J[(size_t)gi * ldj + gj] : 0.0f;
            int gr = col0 + rr;
            Ss[c][rr] = (gr < R && gj < N) ? sgn1(X[(size_t)gr * ldx + gj]) : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int c = 0; c < EN_KC; ++c) {
            float av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = Js[c][ti * 4 + u];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Ss[c][tr * 4 + v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += (double)av[u] * (double)bv[v];
        }
        __syncthreads();
    }
    // s_i * G_i summed over this thread's 4 rows, per replica.
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        int gr = col0 + tr * 4 + v;
        double sum = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int gi = row0 + ti * 4 + u;
            if (gi < N && gr < R) sum += sgn1(X[(size_t)gr * ldx + gi]) > 0.0f ? acc[u][v] : -acc[u][v];
        }
        red[ti][tr * 4 + v] = sum;
    }
    __syncthreads();
    if (tid < EN_T) {
        double s = 0.0;
        for (int t = 0; t < 16; ++t) s += red[t][tid];
        int gr = col0 + tid;
        if (gr < R) part[(size_t)tm * R + gr] = s;
    }
}

__global__ __launch_bounds__(256) void k_energy_final(const double *part, int tiles_m, int R, double *H)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double s = 0.0;
    for (int t = 0; t < tiles_m; ++t) s += part[(size_t)t * R + r];
    H[r] = -0.5 * s;
}

// Integer form, for integer-valued J with |J_ij| <= 127: G = J s on
// v_mfma_i32_32x32x32_i8 (int8 J x int8 signs, int32 accumulators, exact
// for |G_i| <= 127 N), then s_i G_i summed in int64 over the tile's spins
// and, in k_energy_final_i64, over the row tiles.  Tile 64 spins x 64
// replicas, four waves of one 32 x 32 block; the signs are the MFMA's first
// operand, so accumulator v of lane l is replica 8 (v / 4) + 4 (l / 32) + v % 4
// and spin l % 32 of the wave's block.  A and B take the same k positions,
// so the contraction is exact whatever the hardware's k numbering.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr int EI_T = 64, EI_KC = 64, EI_LD = EI_KC + 16;

__device__ __forceinline__ unsigned int pack_i8x4(int a, int b, int c, int d)
{
    return (unsigned)(a & 0xff) | ((unsigned)(b & 0xff) << 8) | ((unsigned)(c & 0xff) << 16) |
           ((unsigned)(d & 0xff) << 24);
}

__global__ __launch_bounds__(256) void k_energy_i8(const float *J, int ldj, const float *X, int ldx, int N,
                                                   int R, long long *part)
{
    __shared__ __attribute__((aligned(16))) unsigned char Js[EI_T][EI_LD];
    __shared__ __attribute__((aligned(16))) unsigned char Ss[EI_T][EI_LD];
    __shared__ long long red[2][EI_T];
    const int tiles_m = (N + EI_T - 1) / EI_T;
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int row0 = tm * EI_T, col0 = tn * EI_T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    // Staging: thread t converts 16 consecutive k (at 16 (t % 4)) of J row
    // t / 4 and of replica t / 4; rows past N and replicas past R are zeros.
    const int sr = tid >> 2, sc = 16 * (tid & 3);
    const bool jrow = row0 + sr < N, srep = col0 + sr < R;
    const float *Jp = J + (size_t)(jrow ? row0 + sr : 0) * ldj + sc;
    const float *Xp = X + (size_t)(srep ? col0 + sr : 0) * ldx + sc;
    v16i acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0;
    // k < round_up(N, 64) <= ldj, ldx; J is zero in the padding columns.
    const int Kp = round_up_dev(N, EI_KC);
    for (int k0 = 0; k0 < Kp; k0 += EI_KC) {
        u32x4 jw = {0u, 0u, 0u, 0u}, sw = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (jrow) {
                const v4f v = *(const v4f *)(Jp + k0 + 4 * g);
                jw[g] = pack_i8x4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
            }
            if (srep) {
                const v4f v = *(const v4f *)(Xp + k0 + 4 * g);
                sw[g] = pack_i8x4(v[0] > 0.0f ? 1 : -1, v[1] > 0.0f ? 1 : -1, v[2] > 0.0f ? 1 : -1,
                                  v[3] > 0.0f ? 1 : -1);
            }
        }
        *(u32x4 *)&Js[sr][sc] = jw;
        *(u32x4 *)&Ss[sr][sc] = sw;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < EI_KC / 32; ++kk) {
            const v4i a = *(const v4i *)&Ss[wn * 32 + (lane & 31)][32 * kk + 16 * (lane >> 5)];
            const v4i b = *(const v4i *)&Js[wm * 32 + (lane & 31)][32 * kk + 16 * (lane >> 5)];
            acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int i = row0 + wm * 32 + (lane & 31);
    long long c[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int r = col0 + wn * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int si = (i < N && r < R) ? (X[(size_t)r * ldx + i] > 0.0f ? 1 : -1) : 0;
        c[v] = (long long)si * (long long)acc[v];
    }
    // Sum over the 32 spins of the block (lanes of one half-wave).
#pragma unroll
    for (int v = 0; v < 16; ++v)
#pragma unroll
        for (int m = 16; m >= 1; m >>= 1) c[v] += __shfl_xor(c[v], m);
    if ((lane & 31) == 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) red[wm][wn * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)] = c[v];
    }
    __syncthreads();
    if (tid < EI_T && col0 + tid < R) part[(size_t)tm * R + col0 + tid] = red[0][tid] + red[1][tid];
}

// Hc_r = -E_r / 2 with E_r = sum_i s_i G_i (even: J is symmetric with a zero
// diagonal), in int64, and its FP64 value (exact below 2^53).
__global__ __launch_bounds__(256) void k_energy_final_i64(const long long *part, int tiles_m, int R,
                                                          long long *Hi, double *H)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    long long s = 0;
    for (int t = 0; t < tiles_m; ++t) s += part[(size_t)t * R + r];
    const long long hc = -(s / 2);
    Hi[r] = hc;
    H[r] = (double)hc;
}

// The field term Hf_r = -sum_i h_i s_i: one workgroup per replica, a fixed
// per-thread order and a fixed tree (deterministic; exact for integer h).
__global__ __launch_bounds__(256) void k_energy_field(const float *h, const float *X, int ldx, int N, double *Hf)
{
    __shared__ double red[256];
    const int r = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < N; i += 256) s += X[(size_t)r * ldx + i] > 0.0f ? (double)h[i] : -(double)h[i];
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) Hf[r] = -red[0];
}

size_t energy_partial_elems(int N, int R)
{
    static_assert(EN_T == EI_T, "both energy forms use one partial layout");
    return (size_t)((N + EN_T - 1) / EN_T) * (size_t)R;
}

hipError_t launch_energy(const float *J, int ldj, const float *X, int ldx, int N, int R, double *partial,
                         double *H, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    int tiles_m = (N + EN_T - 1) / EN_T, tiles_n = (R + EN_T - 1) / EN_T;
    hipLaunchKernelGGL(k_energy_partial, dim3(tiles_m * tiles_n), dim3(256), 0, s, J, ldj, X, ldx, N, R,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_final, dim3((R + 255) / 256), dim3(256), 0, s, partial, tiles_m, R, H);
    return hipGetLastError();
}

hipError_t launch_energy_i8(const float *J, int ldj, const float *X, int ldx, int N, int R, long long *partial,
                            long long *Hi, double *H, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    if (ldj % EI_KC || ldx % EI_KC || ldj < N || ldx < N) return hipErrorInvalidValue;
    int tiles_m = (N + EI_T - 1) / EI_T, tiles_n = (R + EI_T - 1) / EI_T;
    hipLaunchKernelGGL(k_energy_i8, dim3(tiles_m * tiles_n), dim3(256), 0, s, J, ldj, X, ldx, N, R, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_final_i64, dim3((R + 255) / 256), dim3(256), 0, s, partial, tiles_m, R, Hi, H);
    return hipGetLastError();
}

hipError_t launch_energy_field(const float *h, const float *X, int ldx, int N, int R, double *Hf, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy_field, dim3(R), dim3(256), 0, s, h, X, ldx, N, Hf);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CSR form of a dense J on the device: nonzeros per row, then (after a host
// prefix sum into ptr) the column indices and values of each row in
// ascending column order.

__global__ __launch_bounds__(256) void k_row_nnz(const float *J, int N, int ld, int *cnt)
{
    __shared__ int red[4];
    const int i = blockIdx.x, t = threadIdx.x;
    int c = 0;
    for (int j = t; j < N; j += 256) c += J[(size_t)i * ld + j] != 0.0f;
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
    if ((t & 63) == 0) red[t >> 6] = c;
    __syncthreads();
    if (t == 0) cnt[i] = red[0] + red[1] + red[2] + red[3];
}

// One workgroup per row; each pass over 256 columns places its nonzeros by a
// workgroup prefix count (wave ballots), so the order stays ascending.
__global__ __launch_bounds__(256) void k_csr_fill(const float *J, int N, int ld, const int *ptr, int *col,
                                                  float *w)
{
    __shared__ int wave_cnt[4];
    const int i = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int base = ptr[i];
    for (int j0 = 0; j0 < N; j0 += 256) {
        const int j = j0 + t;
        const float v = j < N ? J[(size_t)i * ld + j] : 0.0f;
        const bool nz = v != 0.0f;
        const unsigned long long m = __ballot(nz);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int q = 0; q < wave; ++q) off += wave_cnt[q];
        if (nz) {
            col[off + before] = j;
            w[off + before] = v;
        }
        base += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        __syncthreads();
    }
}

hipError_t launch_row_nnz(const float *J, int N, int ld, int *cnt, hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_nnz, dim3(N), dim3(256), 0, s, J, N, ld, cnt);
    return hipGetLastError();
}

hipError_t launch_csr_fill(const float *J, int N, int ld, const int *ptr, int *col, float *w, hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_csr_fill, dim3(N), dim3(256), 0, s, J, N, ld, ptr, col, w);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Static scratch of sb_energy_noalloc: the per-row-tile FP64 partial sums of
// one call (energy_partial_elems(N, R) doubles), in a file-scope device array
// instead of allocated memory.  Each GPU the library runs on holds its own
// copy, at its own address, until the process ends.  In host code the name
// `scratch` is a host-side stand-in, not a device address: host code passes
// the address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchDoubles = size_t(1) << 20;  // 8 MiB per device
static __device__ __align__(256) double scratch[kScratchDoubles];

size_t noalloc_scratch_elems() { return kScratchDoubles; }

hipError_t noalloc_scratch_address(int device, double **ptr)
{
    // Looked up once per device, with that device current (the caller has
    // made it current), and kept.
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    double *p = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
    *ptr = p;
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Readout and packing.

__global__ __launch_bounds__(256) void k_spins(const float *X, int ldx, int N, int R, int8_t *S)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int r = blockIdx.y; r < R; r += gridDim.y)
        S[(size_t)r * N + i] = X[(size_t)r * ldx + i] > 0.0f ? (int8_t)1 : (int8_t)-1;
}

hipError_t launch_spins(const float *X, int ldx, int N, int R, int8_t *S, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spins, dim3((N + 255) / 256, grid_y(R)), dim3(256), 0, s, X, ldx, N, R, S);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_compact(const float *X, int ldx, int N, int R, float *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int r = blockIdx.y; r < R; r += gridDim.y) out[(size_t)r * N + i] = X[(size_t)r * ldx + i];
}

hipError_t launch_compact(const float *X, int ldx, int N, int R, float *out, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact, dim3((N + 255) / 256, grid_y(R)), dim3(256), 0, s, X, ldx, N, R, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_pack_bf16(const float *J, size_t n, uint16_t *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f2bf(J[i]);
}

// Integer |v| <= 16 -> E4M3 bits (v = m * 2^e with a 3-bit mantissa).
__device__ __forceinline__ uint8_t int_to_e4m3(float v)
{
    if (v == 0.0f) return 0;
    uint32_t b = __builtin_bit_cast(uint32_t, v);
    uint32_t sign = (b >> 31) & 1u;
    int e = (int)((b >> 23) & 0xffu) - 127;  // 0..4 for 1..16
    uint32_t m = (b >> 20) & 0x7u;           // top 3 mantissa bits
    return (uint8_t)((sign << 7) | ((uint32_t)(e + 7) << 3) | m);
}

__global__ __launch_bounds__(256) void k_pack_fp8(const float *J, size_t n, uint8_t *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = int_to_e4m3(J[i]);
}

static unsigned pack_grid(size_t n)
{
    size_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? g : 65536);
}

hipError_t launch_pack_bf16(const float *J, size_t n, uint16_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_bf16, dim3(pack_grid(n)), dim3(256), 0, s, J, n, out);
    return hipGetLastError();
}

hipError_t launch_pack_fp8(const float *J, size_t n, uint8_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_fp8, dim3(pack_grid(n)), dim3(256), 0, s, J, n, out);
    return hipGetLastError();
}

}  // namespace sbk

