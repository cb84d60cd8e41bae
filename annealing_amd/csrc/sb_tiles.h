// sb_tiles.h -- device building blocks of the gfx950 step kernels (included by
// the sb_step_*.hip and sb_kernels.hip translation units).
//
// One fused launch per SB step computes the field F = J.X for a tile of
// (spins x replicas) on MFMA, with J and X staged through LDS, and applies the
// update of that tile in the same kernel:
//   bSB/dSB (Goto et al. 2021): y += dt(-(a0-a)x + c0 F); x += dt a0 y; wall
//   aSB (reference SB/CUDA_GPU/m2_opt.cu:58-75,168-169): kick y += dt c0 F,
//        then the next step's M local sub-steps (FP64-evaluated, FP32-stored)
// Data layout (HBM): J [N][ldj] row-major (FP32 / BF16 / FP8-E4M3), state
// X, Y [R][ldx] FP32 replica-major with spins contiguous; ldj, ldx multiples
// of 64 with zero padding.  X is double-buffered across steps, Y in place.
//
// Build flags: -O3 --offload-arch=gfx950 -ffp-contract=off (only the explicit
// fmaf / fma calls fuse, matching the CPU restatement op for op).
#pragma once
#include "sb_internal.h"

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace sbk {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Workgroup -> tile mapping.  Workgroup b is dispatched to XCD b % 8; xcd_remap
// gives each XCD a contiguous range of logical tiles, and tile_coords lays the
// logical tiles out in groups of `rpg` row tiles (row fastest inside a group),
// so one XCD's range is a near-square block of rows x replica columns whose J
// rows and X columns stay in that XCD's L2.
__device__ __forceinline__ int round_up_dev(int v, int m) { return (v + m - 1) / m * m; }

__device__ __forceinline__ int xcd_remap(int bid, int nblk)
{
    int xcd = bid & 7, local = bid >> 3;
    int per = nblk >> 3, rem = nblk & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

__device__ __forceinline__ void tile_coords(const StepArgs &a, int L, int &tm, int &tn)
{
    const int rpg = a.rpg > 0 ? a.rpg : a.tiles_m;
    const int full = rpg * a.tiles_n;
    const int g = L / full;
    const int rg = min(rpg, a.tiles_m - g * rpg);
    const int local = L - g * full;
    tn = local / rg;
    tm = g * rpg + local % rg;
}

// ---------------------------------------------------------------------------
// Update rules (single definition used by every kernel family).

__device__ __forceinline__ void asb_local(float &x, float &y, float p, double a0, double K,
                                          double dts, int M)
{
    // m2_opt.cu:66-71 with double constants: each statement in FP64, stored FP32.
    for (int m = 0; m < M; ++m) {
        double t = a0 * (double)y;
        t = t * dts;
        x = (float)((double)x + t);
        double xd = (double)x;
        double cube = K * xd;
        cube = cube * xd;
        cube = cube * xd;
        double lin = a0 - (double)p;
        lin = lin * xd;
        double u = cube + lin;
        u = dts * u;
        y = (float)((double)y - u);
    }
}

// Bifurcation_m1.cu.c:55-69 (UpdateX): detune * y in FP32, then the product
// with DeltaT and the sum in FP64, stored as FP32.
__device__ __forceinline__ float m1_x(float x, float y, double a0, double dt)
{
    const float t = (float)a0 * y;
    return (float)((double)t * dt + (double)x);
}

// Bifurcation_m1.cu.c:71-92 (UpdateY): x^3, p x and detune x in FP32, each
// times DeltaT and summed in FP64 ((A + B) - C), stored as FP32; the coupling
// sum times xi and DeltaT in FP64, stored as FP32; then two FP32 additions.
__device__ __forceinline__ float m1_y(float F, float x, float y, float p, double a0, double K, double dt,
                                      double xi)
{
    const float x3 = x * x * x;
    const double A = (double)(-(float)K * x3) * dt;
    const double B = (double)(p * x) * dt;
    const double C = (double)((float)a0 * x) * dt;
    float ny = (float)((A + B) - C);
    const float s = (float)(((double)F * xi) * dt);
    ny = ny + s;
    ny = ny + y;
    return ny;
}

// The update of one spin-replica for a compile-time variant (bSB and dSB share
// it: they differ only in the field).
template <int V>
__device__ __forceinline__ void sb_update_v(const sb_step_coeffs &c, float F, float &x, float &y)
{
    if constexpr (V == SB_ASB_M1) {
        y = m1_y(F, x, y, c.p_next, c.a0, c.K, c.dts, c.c0);
        if (c.do_local) x = m1_x(x, y, c.a0, c.dts);
    } else if constexpr (V == SB_ASB) {
        y = __builtin_fmaf(c.beta, F, y);  // cublasSgemv(alpha = dt xi, beta = 1)
        if (c.do_local) asb_local(x, y, c.p_next, c.a0, c.K, c.dts, c.M);
    } else {
        float yn = __builtin_fmaf(c.beta, F, __builtin_fmaf(-c.alpha, x, y));
        float xn = __builtin_fmaf(c.gamma, yn, x);
        if (__builtin_fabsf(xn) > 1.0f) {
            xn = __builtin_copysignf(1.0f, xn);
            yn = 0.0f;
        }
        x = xn;
        y = yn;
    }
}

__device__ __forceinline__ void sb_update(const sb_step_coeffs &c, float F, float &x, float &y)
{
    if (c.variant == SB_ASB_M1)
        sb_update_v<SB_ASB_M1>(c, F, x, y);
    else if (c.variant == SB_ASB)
        sb_update_v<SB_ASB>(c, F, x, y);
    else
        sb_update_v<SB_BSB>(c, F, x, y);
}

__device__ __forceinline__ float sgn1(float v) { return v > 0.0f ? 1.0f : -1.0f; }

__device__ __forceinline__ unsigned short f2bf(float f)
{
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float bf2f(unsigned short h)
{
    return __builtin_bit_cast(float, (unsigned int)h << 16);
}

// B-operand planes of four consecutive spins (one float4 of X) at element o:
//   BK_SIGN_FP8  E4M3 +1.0 = 0x38, -1.0 = 0xB8 (x > 0 ? +1 : -1)
//   BK_SIGN_BF16 BF16 +1.0 = 0x3F80, -1.0 = 0xBF80
__device__ __forceinline__ void write_planes(int kind, const v4f &xv, void *B, size_t o)
{
    if (kind == BK_SIGN_FP8) {
        unsigned int s4 = 0u;
#pragma unroll
        for (int e = 0; e < 4; ++e) s4 |= (xv[e] > 0.0f ? 0x38u : 0xB8u) << (8 * e);
        *(unsigned int *)((unsigned char *)B + o) = s4;
    } else if (kind == BK_SIGN_BF16) {
        u16x4 h;
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = xv[e] > 0.0f ? (unsigned short)0x3f80 : (unsigned short)0xbf80;
        *(u16x4 *)((unsigned short *)B + o) = h;
    }
}

// Plane entry of one spin (same encodings as write_planes).
__device__ __forceinline__ void write_plane1(int kind, float x, void *B, size_t o)
{
    if (kind == BK_SIGN_FP8) {
        ((unsigned char *)B)[o] = x > 0.0f ? (unsigned char)0x38 : (unsigned char)0xB8;
    } else if (kind == BK_SIGN_BF16) {
        ((unsigned short *)B)[o] = x > 0.0f ? (unsigned short)0x3f80 : (unsigned short)0xbf80;
    }
}

// Epilogue of one 32x32 accumulator block.  The tiles multiply with the state
// fragment in the MFMA's A operand and the coupling fragment in its B operand,
// so the block is F^T: lane l, register v holds replica
// rep_base + 8*(v/4) + 4*(l/32) + v%4 and spin spin_base + l%32.  Every load
// and store of a half-wave then covers 32 consecutive spins (128 bytes of X or
// Y) of one replica: whole cache lines, not 16-byte pieces of 32 lines.
template <int V>
__device__ __forceinline__ void epilogue_block_v(const StepArgs &a, const v16f &acc, int spin_base, int rep_base,
                                                 int lane)
{
    const int i = spin_base + (lane & 31);
    if (i >= a.N) return;
    // External field: F_i = (J.x)_i + h_i, added after the sum.
    const float hi = a.h ? a.h[i] : 0.0f;
    auto field = [&](float f) { return a.h ? f + hi : f; };
    const int r0 = rep_base + 4 * (lane >> 5);
    const size_t o0 = (size_t)r0 * a.ldx + i;
    auto off = [&](int v) { return o0 + (size_t)(8 * (v >> 2) + (v & 3)) * a.ldx; };
    if (rep_base + 32 <= a.R) {
        // Whole block: all loads issue before the first use.
        float xv[16], yv[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            xv[v] = a.X[off(v)];
            yv[v] = a.Y[off(v)];
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            sb_update_v<V>(a.c, field(acc[v]), xv[v], yv[v]);
            a.Xn[off(v)] = xv[v];
            a.Y[off(v)] = yv[v];
        }
        if (a.Bn) {
#pragma unroll
            for (int v = 0; v < 16; ++v) write_plane1(a.bkind, xv[v], a.Bn, off(v));
        }
        return;
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        if (r0 + 8 * (v >> 2) + (v & 3) >= a.R) continue;
        float x = a.X[off(v)], y = a.Y[off(v)];
        sb_update_v<V>(a.c, field(acc[v]), x, y);
        a.Xn[off(v)] = x;
        a.Y[off(v)] = y;
        if (a.Bn) write_plane1(a.bkind, x, a.Bn, off(v));
    }
}

// Epilogue of a wave's BM x BN accumulator blocks.  The variant is dispatched
// once, outside the blocks, and the blocks are expanded at compile time (a
// fold over an index sequence, not a loop left to the unroller), so every
// accumulator is indexed statically and stays in registers at any BM x BN.
template <int BM, int BN, int V, int... I>
__device__ __forceinline__ void epilogue_blocks_seq(const StepArgs &a, const v16f (&acc)[BM][BN], int spin_base,
                                                    int rep_base, int lane, std::integer_sequence<int, I...>)
{
    (epilogue_block_v<V>(a, acc[I / BN][I % BN], spin_base + 32 * (I / BN), rep_base + 32 * (I % BN), lane), ...);
}

template <int BM, int BN, int V>
__device__ __forceinline__ void epilogue_blocks_v(const StepArgs &a, const v16f (&acc)[BM][BN], int spin_base,
                                                  int rep_base, int lane)
{
    epilogue_blocks_seq<BM, BN, V>(a, acc, spin_base, rep_base, lane, std::make_integer_sequence<int, BM * BN>{});
}

template <int BM, int BN>
__device__ __forceinline__ void epilogue_tile(const StepArgs &a, const v16f (&acc)[BM][BN], int spin_base,
                                              int rep_base, int lane)
{
    if (a.c.variant == SB_ASB_M1)
        epilogue_blocks_v<BM, BN, SB_ASB_M1>(a, acc, spin_base, rep_base, lane);
    else if (a.c.variant == SB_ASB)
        epilogue_blocks_v<BM, BN, SB_ASB>(a, acc, spin_base, rep_base, lane);
    else
        epilogue_blocks_v<BM, BN, SB_BSB>(a, acc, spin_base, rep_base, lane);
}

// ---------------------------------------------------------------------------
// K pipeline shared by the families.  Two register stages hold chunks in
// flight from global memory and LDS is double-buffered: chunk c+2's loads are
// issued before chunk c is multiplied and stored to LDS after chunk c+1 is, so
// every load has two chunks of MFMA work to arrive.  Stage selection is a
// compile-time constant (the stages stay in registers).

template <int S, typename T>
__device__ __forceinline__ T &pick(T &s0, T &s1)
{
    if constexpr (S == 0) return s0;
    else return s1;
}

// Range-checked tile loads.  A buffer resource over [base, base + bytes)
// returns zeros for any load that starts past the end (the hardware's range
// check), so rows past N and replicas past R need no branch: with the loads
// unconditional the compiler can wait for the older of two chunks in flight
// instead of draining every load before each LDS store.  Word 3 is the raw
// 32-bit data format of the gfx9 family.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void *base, long long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ u32x4 rload16(__amdgpu_buffer_rsrc_t r, unsigned int off)
{
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}
// An offset past any tile range (tiles stay below 2^31 bytes).
constexpr unsigned int kOutOfRange = 0x80000000u;

template <typename Load, typename Store, typename Compute>
__device__ __forceinline__ void kpipeline(int nch, Load gload, Store sstore, Compute compute)
{
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    // Loads and LDS stores are unconditional, also for the (at most two)
    // chunks past the last: the tile loads are range-checked (in-range bytes
    // or zeros, never a fault) and such a stage lands in an LDS buffer that no
    // compute reads.  Every path then has the same loads in flight, so the
    // wait before a store covers only that stage's loads.
    gload(I0{}, 0);
    sstore(I0{}, 0);
    gload(I1{}, 1);
    __syncthreads();
    int c = 0;
    for (; c + 1 < nch; c += 2) {
        gload(I0{}, c + 2);
        compute(0);
        sstore(I1{}, 1);
        __syncthreads();
        gload(I1{}, c + 3);
        compute(1);
        sstore(I0{}, 0);
        __syncthreads();
    }
    if (c < nch) compute(0);  // odd chunk count: the last chunk is in buffer 0
}

// ---------------------------------------------------------------------------
// FP32 family: v_mfma_f32_32x32x2_f32.  K chunk of 32 staged through LDS with
// double buffering.  Each lane reads a float4 of 4 consecutive k for its row
// (A) and its replica (B); the four components feed four MFMAs, lane half
// h = l/32 covering k = 8kk + 4h + c.  A and B use the same k assignment, so
// the contraction is exact regardless of the hardware's k numbering.  In every
// family the staged tiles are A = J rows and B = replica states, and the MFMA
// takes the B fragment as its first operand: each accumulator block is F^T,
// laid out for coalesced state I/O in epilogue_block.

template <int WM, int WN, int BM, int BN>
struct Fp32Tile {
    static constexpr int MT = 32 * BM * WM, NT = 32 * BN * WN, NTHR = 64 * WM * WN;
    static constexpr int KC = 32, LDS_LD = KC + 4;
    static constexpr int LA = (MT * KC / 4 + NTHR - 1) / NTHR;
    static constexpr int LB = (NT * KC / 4 + NTHR - 1) / NTHR;
};

template <int WM, int WN, int BM, int BN>
__device__ __forceinline__ void fp32_tile(const StepArgs &a, int tile, float *smem)
{
    using T = Fp32Tile<WM, WN, BM, BN>;
    constexpr int MT = T::MT, NT = T::NT, NTHR = T::NTHR, KC = T::KC, LD = T::LDS_LD;
    float *As = smem;                  // [2][MT][LD]
    float *Bs = smem + 2 * MT * LD;    // [2][NT][LD]
    int tm, tn;
    tile_coords(a, tile, tm, tn);
    const int row0 = tm * MT, col0 = tn * NT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const float *J = (const float *)a.J;
    const bool sgn = a.c.variant == SB_DSB;
    const int Kp = round_up_dev(a.N, KC);

    // Rows row0.. of J and replicas col0.. of X; loads past N or R read zeros.
    const __amdgpu_buffer_rsrc_t rJ =
        tile_rsrc(J + (size_t)row0 * a.ldj, (long long)min(MT, a.N - row0) * a.ldj * 4);
    const __amdgpu_buffer_rsrc_t rX =
        tile_rsrc(a.X + (size_t)col0 * a.ldx, (long long)min(NT, a.R - col0) * a.ldx * 4);

    struct Stage {
        u32x4 a[T::LA], b[T::LB];
    };
    Stage st0, st1;
    auto gload = [&](auto S, int ch) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
        const int k0 = ch * KC;
#pragma unroll
        for (int u = 0; u < T::LA; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / 4), c4 = q % (KC / 4);
            R.a[u] = rload16(rJ, (unsigned)(rr * a.ldj + k0 + 4 * c4) * 4u);
        }
#pragma unroll
        for (int u = 0; u < T::LB; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / 4), c4 = q % (KC / 4);
            R.b[u] = rload16(rX, (unsigned)(rr * a.ldx + k0 + 4 * c4) * 4u);
        }
    };
    auto sstore = [&](auto S, int buf) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
#pragma unroll
        for (int u = 0; u < T::LA; ++u) {
            int q = tid + u * NTHR;
            if ((MT * KC / 4) % NTHR == 0 || q < MT * KC / 4) {
                int rr = q / (KC / 4), c4 = q % (KC / 4);
                *(u32x4 *)(As + (buf * MT + rr) * LD + 4 * c4) = R.a[u];
            }
        }
#pragma unroll
        for (int u = 0; u < T::LB; ++u) {
            int q = tid + u * NTHR;
            if ((NT * KC / 4) % NTHR == 0 || q < NT * KC / 4) {
                int rr = q / (KC / 4), c4 = q % (KC / 4);
                v4f v = __builtin_bit_cast(v4f, R.b[u]);
                if (sgn) v = (v4f){sgn1(v[0]), sgn1(v[1]), sgn1(v[2]), sgn1(v[3])};
                *(v4f *)(Bs + (buf * NT + rr) * LD + 4 * c4) = v;
            }
        }
    };

    v16f acc[BM][BN];
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
#pragma unroll
        for (int bn = 0; bn < BN; ++bn)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[bm][bn][v] = 0.0f;

    const int arow = wm * 32 * BM + (lane & 31), bcol = wn * 32 * BN + (lane & 31);
    const int koff = 4 * (lane >> 5);
    auto compute = [&](int buf) {
        const float *Ab = As + buf * MT * LD;
        const float *Bb = Bs + buf * NT * LD;
#pragma unroll
        for (int kk = 0; kk < KC / 8; ++kk) {
            v4f fa[BM], fb[BN];
#pragma unroll
            for (int bm = 0; bm < BM; ++bm) fa[bm] = *(const v4f *)(Ab + (arow + 32 * bm) * LD + 8 * kk + koff);
#pragma unroll
            for (int bn = 0; bn < BN; ++bn) fb[bn] = *(const v4f *)(Bb + (bcol + 32 * bn) * LD + 8 * kk + koff);
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int bm = 0; bm < BM; ++bm)
#pragma unroll
                    for (int bn = 0; bn < BN; ++bn)
                        acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x2f32(fb[bn][c], fa[bm][c], acc[bm][bn], 0, 0, 0);
        }
    };
    kpipeline((Kp + KC - 1) / KC, gload, sstore, compute);
    epilogue_tile<BM, BN>(a, acc, row0 + wm * 32 * BM, col0 + wn * 32 * BN, lane);
}

// ---------------------------------------------------------------------------
// Split family: v_mfma_f32_32x32x16_bf16.  K chunk of 64.  A = J in BF16 (read
// as BF16, or as FP8 E4M3 widened to BF16 by v_cvt_scalef32_pk_bf16_fp8 while
// staging).  B = X as BF16 hi = bf16(x) and lo = bf16(x - hi), two MFMA passes
// into one FP32 accumulator per k16 group (hi then lo); dSB uses sign(x) once.
// B is split from FP32 X while staging; for dSB (SGN), PLANES instead copies
// the BF16 sign plane written by the previous step's epilogue (2 bytes per
// element instead of 4; same values).  Lane half h reads 8 consecutive k
// (16 bytes) at 16kk + 8h for A and B.

// 4 FP8 bytes -> 4 BF16 (exact for the integers |J| <= 16 the packer stores).
__device__ __forceinline__ void fp8x4_to_bf16x4(unsigned int b4, unsigned int &lo2, unsigned int &hi2)
{
    lo2 = __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(b4, 1.0f, false));
    hi2 = __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(b4, 1.0f, true));
}

template <int WM, int WN, int BM, int BN>
struct SplitTile {
    static constexpr int MT = 32 * BM * WM, NT = 32 * BN * WN, NTHR = 64 * WM * WN;
    static constexpr int KC = 64, LDS_LD = KC + 8;  // in bf16 elements
};

template <int WM, int WN, int BM, int BN, bool FP8J, bool PLANES, bool SGN>
__device__ __forceinline__ void split_tile(const StepArgs &a, int tile, unsigned short *smem)
{
    static_assert(!PLANES || SGN, "planes hold signs only");
    using T = SplitTile<WM, WN, BM, BN>;
    constexpr int MT = T::MT, NT = T::NT, NTHR = T::NTHR, KC = T::KC, LD = T::LDS_LD;
    // A: 16-byte global loads of 8 BF16 or 16 FP8.
    constexpr int A_PER = FP8J ? 16 : 8;               // elements per global load
    constexpr int A_LOADS = MT * KC / A_PER;
    constexpr int LA = (A_LOADS + NTHR - 1) / NTHR;
    // B: float4 loads of X, or 16-byte loads of 8 BF16 per plane.
    constexpr int B_PER = PLANES ? 8 : 4;
    constexpr int B_LOADS = NT * KC / B_PER;
    constexpr int LB = (B_LOADS + NTHR - 1) / NTHR;
    constexpr int KK = KC / 16;
    unsigned short *As = smem;                          // [2][MT][LD]
    unsigned short *Bh = smem + 2 * MT * LD;            // [2][NT][LD]
    unsigned short *Bl = Bh + 2 * NT * LD;              // [2][NT][LD]
    int tm, tn;
    tile_coords(a, tile, tm, tn);
    const int row0 = tm * MT, col0 = tn * NT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int Kp = round_up_dev(a.N, KC);
    const unsigned short *Ph = (const unsigned short *)a.B;
    // Rows row0.. of J and replicas col0.. of the B source; loads past N or R
    // read zeros.
    constexpr int JBYTES = FP8J ? 1 : 2;
    const int rows = min(MT, a.N - row0), cols = min(NT, a.R - col0);
    const __amdgpu_buffer_rsrc_t rJ = tile_rsrc((const unsigned char *)a.J + (size_t)row0 * a.ldj * JBYTES,
                                                (long long)rows * a.ldj * JBYTES);
    const __amdgpu_buffer_rsrc_t rB =
        PLANES ? tile_rsrc(Ph + (size_t)col0 * a.ldx, (long long)cols * a.ldx * 2)
               : tile_rsrc(a.X + (size_t)col0 * a.ldx, (long long)cols * a.ldx * 4);

    struct Stage {
        u32x4 ra[LA];
        u32x4 rb[LB];
    };
    Stage st0, st1;
    auto gload = [&](auto S, int ch) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
        const int k0 = ch * KC;
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / A_PER), c = q % (KC / A_PER);
            R.ra[u] = rload16(rJ, (unsigned)(rr * a.ldj + k0 + A_PER * c) * JBYTES);
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / B_PER), c = q % (KC / B_PER);
            if (PLANES)
                R.rb[u] = rload16(rB, (unsigned)(rr * a.ldx + k0 + 8 * c) * 2u);
            else
                R.rb[u] = rload16(rB, (unsigned)(rr * a.ldx + k0 + 4 * c) * 4u);
        }
    };
    auto sstore = [&](auto S, int buf) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            int q = tid + u * NTHR;
            if (A_LOADS % NTHR == 0 || q < A_LOADS) {
                int rr = q / (KC / A_PER), c = q % (KC / A_PER);
                unsigned short *dst = As + (buf * MT + rr) * LD + A_PER * c;
                if (FP8J) {
                    unsigned int w[8];
#pragma unroll
                    for (int e = 0; e < 4; ++e) fp8x4_to_bf16x4(R.ra[u][e], w[2 * e], w[2 * e + 1]);
                    *(u32x4 *)dst = (u32x4){w[0], w[1], w[2], w[3]};
                    *(u32x4 *)(dst + 8) = (u32x4){w[4], w[5], w[6], w[7]};
                } else {
                    *(u32x4 *)dst = R.ra[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            int q = tid + u * NTHR;
            if (B_LOADS % NTHR == 0 || q < B_LOADS) {
                int rr = q / (KC / B_PER), c = q % (KC / B_PER);
                if (PLANES) {
                    *(u32x4 *)(Bh + (buf * NT + rr) * LD + 8 * c) = R.rb[u];
                } else {
                    u16x4 h, l;
                    const v4f xv = __builtin_bit_cast(v4f, R.rb[u]);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        float v = xv[e];
                        if (SGN) {
                            h[e] = v > 0.0f ? (unsigned short)0x3f80 : (unsigned short)0xbf80;
                            l[e] = 0;
                        } else {
                            unsigned short hb = f2bf(v);
                            h[e] = hb;
                            l[e] = f2bf(v - bf2f(hb));
                        }
                    }
                    *(u16x4 *)(Bh + (buf * NT + rr) * LD + 4 * c) = h;
                    if (!SGN) *(u16x4 *)(Bl + (buf * NT + rr) * LD + 4 * c) = l;
                }
            }
        }
    };

    v16f acc[BM][BN];
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
#pragma unroll
        for (int bn = 0; bn < BN; ++bn)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[bm][bn][v] = 0.0f;

    const int arow = wm * 32 * BM + (lane & 31), bcol = wn * 32 * BN + (lane & 31);
    const int koff = 8 * (lane >> 5);
    struct Frag {
        bf16x8 a[BM], h[BN], l[BN];
    };
    auto compute = [&](int buf) {
        const unsigned short *Ab = As + buf * MT * LD;
        const unsigned short *Bhb = Bh + buf * NT * LD;
        const unsigned short *Blb = Bl + buf * NT * LD;
        auto ldfrag = [&](Frag &F, int kk) {
#pragma unroll
            for (int bm = 0; bm < BM; ++bm)
                F.a[bm] = __builtin_bit_cast(bf16x8, *(const u16x8 *)(Ab + (arow + 32 * bm) * LD + 16 * kk + koff));
#pragma unroll
            for (int bn = 0; bn < BN; ++bn) {
                F.h[bn] = __builtin_bit_cast(bf16x8, *(const u16x8 *)(Bhb + (bcol + 32 * bn) * LD + 16 * kk + koff));
                if (!SGN)
                    F.l[bn] = __builtin_bit_cast(bf16x8, *(const u16x8 *)(Blb + (bcol + 32 * bn) * LD + 16 * kk + koff));
            }
        };
        auto mma = [&](const Frag &F) {
#pragma unroll
            for (int bm = 0; bm < BM; ++bm)
#pragma unroll
                for (int bn = 0; bn < BN; ++bn)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F.h[bn], F.a[bm], acc[bm][bn], 0, 0, 0);
            if (!SGN) {
#pragma unroll
                for (int bm = 0; bm < BM; ++bm)
#pragma unroll
                    for (int bn = 0; bn < BN; ++bn)
                        acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F.l[bn], F.a[bm], acc[bm][bn], 0, 0, 0);
            }
        };
#pragma unroll
        for (int kk = 0; kk < KK; ++kk) {
            Frag f;
            ldfrag(f, kk);
            mma(f);
        }
    };
    kpipeline((Kp + KC - 1) / KC, gload, sstore, compute);
    epilogue_tile<BM, BN>(a, acc, row0 + wm * 32 * BM, col0 + wn * 32 * BN, lane);
}

// ---------------------------------------------------------------------------
// dSB FP8 family: FP8 J x FP8 sign(x) on v_mfma_scale_f32_32x32x64_f8f6f4
// with unit scales.  Products are +-J_ij exactly and the FP32 sums of
// integers below 2^24 are exact, so F equals the FP32 family's F bit for bit.
// K chunk of 128 bytes; lane half h reads 32 consecutive k at 64kk + 32h.

template <int WM, int WN, int BM, int BN>
struct Fp8Tile {
    static constexpr int MT = 32 * BM * WM, NT = 32 * BN * WN, NTHR = 64 * WM * WN;
    static constexpr int KC = 128, LDS_LD = KC + 16;  // bytes
};

template <int WM, int WN, int BM, int BN, bool PLANES>
__device__ __forceinline__ void dsb_fp8_tile(const StepArgs &a, int tile, unsigned char *smem)
{
    using T = Fp8Tile<WM, WN, BM, BN>;
    constexpr int MT = T::MT, NT = T::NT, NTHR = T::NTHR, KC = T::KC, LD = T::LDS_LD;
    constexpr int A_LOADS = MT * KC / 16;
    constexpr int LA = (A_LOADS + NTHR - 1) / NTHR;
    // B: float4 loads of X -> 4 sign bytes, or 16-byte loads of the sign plane.
    constexpr int B_PER = PLANES ? 16 : 4;
    constexpr int B_LOADS = NT * KC / B_PER;
    constexpr int LB = (B_LOADS + NTHR - 1) / NTHR;
    unsigned char *As = smem;               // [2][MT][LD]
    unsigned char *Bs = smem + 2 * MT * LD; // [2][NT][LD]
    int tm, tn;
    tile_coords(a, tile, tm, tn);
    const int row0 = tm * MT, col0 = tn * NT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    // K loop bound: the last 128-byte chunk is clipped at round_up(N, 64) <= ld
    // (zero bytes beyond it; J is zero there).
    const int Kp = round_up_dev(a.N, 64);
    const unsigned char *P = (const unsigned char *)a.B;

    // Rows row0.. of J and replicas col0.. of the B source; loads past N or R,
    // and K columns at or past Kp (the last chunk can run past ld), read zeros.
    const int rows = min(MT, a.N - row0), cols = min(NT, a.R - col0);
    const __amdgpu_buffer_rsrc_t rJ =
        tile_rsrc((const unsigned char *)a.J + (size_t)row0 * a.ldj, (long long)rows * a.ldj);
    const __amdgpu_buffer_rsrc_t rB = PLANES ? tile_rsrc(P + (size_t)col0 * a.ldx, (long long)cols * a.ldx)
                                             : tile_rsrc(a.X + (size_t)col0 * a.ldx, (long long)cols * a.ldx * 4);

    struct Stage {
        u32x4 ra[LA];
        unsigned int rb[PLANES ? 1 : LB];
        u32x4 rbp[PLANES ? LB : 1];
    };
    Stage st0, st1;
    auto gload = [&](auto S, int ch) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
        const int k0 = ch * KC;
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / 16), c = q % (KC / 16), gk = k0 + 16 * c;
            R.ra[u] = rload16(rJ, gk < Kp ? (unsigned)(rr * a.ldj + gk) : kOutOfRange);
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int q = tid + u * NTHR, rr = q / (KC / B_PER), c = q % (KC / B_PER), gk = k0 + B_PER * c;
            if (PLANES) {
                R.rbp[u] = rload16(rB, gk < Kp ? (unsigned)(rr * a.ldx + gk) : kOutOfRange);
            } else {
                // Without planes (operator calls only) the sign bytes are packed
                // at once, which keeps the stage small.  E4M3 +1.0 = 0x38,
                // -1.0 = 0xB8; zeros read past the range become -1 (J is zero
                // there).
                const v4f v = __builtin_bit_cast(v4f, rload16(rB, gk < Kp ? (unsigned)(rr * a.ldx + gk) * 4u
                                                                         : kOutOfRange));
                unsigned int s4 = 0u;
#pragma unroll
                for (int e = 0; e < 4; ++e) s4 |= (v[e] > 0.0f ? 0x38u : 0xB8u) << (8 * e);
                R.rb[u] = s4;
            }
        }
    };
    auto sstore = [&](auto S, int buf) {
        Stage &R = pick<decltype(S)::value>(st0, st1);
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            int q = tid + u * NTHR;
            if (A_LOADS % NTHR == 0 || q < A_LOADS) {
                int rr = q / (KC / 16), c = q % (KC / 16);
                *(u32x4 *)(As + (buf * MT + rr) * LD + 16 * c) = R.ra[u];
            }
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            int q = tid + u * NTHR;
            if (B_LOADS % NTHR == 0 || q < B_LOADS) {
                int rr = q / (KC / B_PER), c = q % (KC / B_PER);
                if (PLANES)
                    *(u32x4 *)(Bs + (buf * NT + rr) * LD + 16 * c) = R.rbp[u];
                else
                    *(unsigned int *)(Bs + (buf * NT + rr) * LD + 4 * c) = R.rb[u];
            }
        }
    };

    v16f acc[BM][BN];
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
#pragma unroll
        for (int bn = 0; bn < BN; ++bn)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[bm][bn][v] = 0.0f;

    const int arow = wm * 32 * BM + (lane & 31), bcol = wn * 32 * BN + (lane & 31);
    const int koff = 32 * (lane >> 5);
    auto compute = [&](int buf) {
        const unsigned char *Ab = As + buf * MT * LD;
        const unsigned char *Bb = Bs + buf * NT * LD;
#pragma unroll
        for (int kk = 0; kk < KC / 64; ++kk) {
            i32x8 fa[BM], fb[BN];
#pragma unroll
            for (int bm = 0; bm < BM; ++bm) {
                const unsigned char *p = Ab + (arow + 32 * bm) * LD + 64 * kk + koff;
                u32x4 x0 = *(const u32x4 *)p, x1 = *(const u32x4 *)(p + 16);
                fa[bm] = (i32x8){(int)x0[0], (int)x0[1], (int)x0[2], (int)x0[3],
                                 (int)x1[0], (int)x1[1], (int)x1[2], (int)x1[3]};
            }
#pragma unroll
            for (int bn = 0; bn < BN; ++bn) {
                const unsigned char *p = Bb + (bcol + 32 * bn) * LD + 64 * kk + koff;
                u32x4 x0 = *(const u32x4 *)p, x1 = *(const u32x4 *)(p + 16);
                fb[bn] = (i32x8){(int)x0[0], (int)x0[1], (int)x0[2], (int)x0[3],
                                 (int)x1[0], (int)x1[1], (int)x1[2], (int)x1[3]};
            }
#pragma unroll
            for (int bm = 0; bm < BM; ++bm)
#pragma unroll
                for (int bn = 0; bn < BN; ++bn)
                    acc[bm][bn] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        fb[bn], fa[bm], acc[bm][bn], 0, 0, 0, 127, 0, 127);
        }
    };
    kpipeline((Kp + KC - 1) / KC, gload, sstore, compute);
    epilogue_tile<BM, BN>(a, acc, row0 + wm * 32 * BM, col0 + wn * 32 * BN, lane);
}

// ---------------------------------------------------------------------------
// Thin families (R <= kThinMaxR).  With a handful of replicas the field is a
// matrix-vector product that streams J once from HBM; MFMA tiles would be at
// least 7/8 padding, so it runs as FP32 fmaf on the VALU.  Wave w of tile t
// owns spins [(4t + w) * ROWS, +ROWS).  Lane l accumulates, per row and
// replica, the 16 consecutive k at 16 * (l + 64 * it) in ascending k with
// fmaf, over it = 0, 1, ...; the 64 lane sums are then combined by an xor
// butterfly, which leaves the same sum in every lane (each level adds two
// values that are already equal across the pair).  The order depends neither
// on ROWS nor on how J is stored, so every thin shape, and the FP32, BF16 and
// FP8 forms of the same J values, give the same field.

// The xor butterfly over the 64 lanes, levels 32, 16, 8, 4, 2, 1: at level m
// lane l adds lane l ^ m's value to its own.  Level 32 goes through
// ds_bpermute, levels 16..4 through ds_swizzle (xor mask within 32 lanes) and
// levels 2, 1 through DPP quad permutes; each level is the same addition as
// v + __shfl_xor(v, m).
template <int M>
__device__ __forceinline__ float xor_lane(float v)
{
    const int b = __builtin_bit_cast(int, v);
    int p;
    if constexpr (M == 1)
        p = __builtin_amdgcn_update_dpp(b, b, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    else if constexpr (M == 2)
        p = __builtin_amdgcn_update_dpp(b, b, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
    else if constexpr (M < 32)
        p = __builtin_amdgcn_ds_swizzle(b, 0x1F | (M << 10));  // and 0x1F, or 0, xor M
    else
        return __shfl_xor(v, M);
    return __builtin_bit_cast(float, p);
}
__device__ __forceinline__ float wave_sum_xor(float v)
{
    v += xor_lane<32>(v);
    v += xor_lane<16>(v);
    v += xor_lane<8>(v);
    v += xor_lane<4>(v);
    v += xor_lane<2>(v);
    v += xor_lane<1>(v);
    return v;
}

// One lane's 16 consecutive J values of a row, as stored (FP32: 4 x 16 bytes,
// BF16: 2, FP8: 1), and their widening to FP32 (exact for BF16 and FP8 E4M3).
template <int JB>
struct ThinJRaw {
    u32x4 w[JB];
};
template <int JB>
__device__ __forceinline__ void thin_widen_j(const ThinJRaw<JB> &raw, float (&v)[16])
{
    if constexpr (JB == 4) {
        // Whole-vector bit casts: hipcc (ROCm 7.2) folds a bit cast of a
        // single vector element, bit_cast<float>(w[e]), to element 0.
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const v4f t = __builtin_bit_cast(v4f, raw.w[g]);
#pragma unroll
            for (int e = 0; e < 4; ++e) v[4 * g + e] = t[e];
        }
    } else if constexpr (JB == 2) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const u32x4 t = raw.w[g];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[8 * g + 2 * e] = __builtin_bit_cast(float, t[e] << 16);
                v[8 * g + 2 * e + 1] = __builtin_bit_cast(float, t[e] & 0xffff0000u);
            }
        }
    } else {
        const u32x4 t = raw.w[0];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            unsigned int lo2, hi2;  // BF16 pairs of bytes 0,1 and 2,3
            fp8x4_to_bf16x4(t[e], lo2, hi2);
            v[4 * e + 0] = __builtin_bit_cast(float, lo2 << 16);
            v[4 * e + 1] = __builtin_bit_cast(float, lo2 & 0xffff0000u);
            v[4 * e + 2] = __builtin_bit_cast(float, hi2 << 16);
            v[4 * e + 3] = __builtin_bit_cast(float, hi2 & 0xffff0000u);
        }
    }
}

// One group (16 consecutive k at k0) of every row and replica of a wave, as
// loaded.  Replica slots r >= R (RM = 4 with R = 3) load replica 0 again and
// their sums are never used.
template <int ROWS, int JB, int RM>
struct ThinGroup {
    ThinJRaw<JB> j[ROWS];
    u32x4 x[RM][4];
};

// RM: replicas per kernel (1, 2 or 4; R <= RM).  Every load is a range-checked
// buffer load (a row or replica holds round_up(N, 16) in-range elements, the
// rest read as zeros), so the loads carry no branch: with RM a compile-time
// constant a group's loads issue together, and when two groups fit the
// registers (PIPE) the loads of group it + 1 issue before group it is summed,
// so a row of up to 2048 spins takes one memory round trip and a longer row
// streams with a group in flight.  A lane sums a group only if it starts below
// N, exactly as above.
template <int ROWS, int JB, bool SGN, int RM, int VBUDGET = 192>
__device__ __forceinline__ void thin_tile(const StepArgs &a, int tile)
{
    // wave (hence row0 and the buffer resources) is uniform across the wave.
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row0 = (tile * 4 + wave) * ROWS;
    if (row0 >= a.N) return;
    // Lane q * kThinMaxR + r updates spin row0 + q of replica r; its x, y (and
    // h) are loaded before the sum so their latency overlaps the J stream.
    const int oq = lane / kThinMaxR, orr = lane % kThinMaxR, oi = row0 + oq;
    const bool owner = lane < ROWS * kThinMaxR && orr < a.R && oi < a.N;
    const size_t o = (size_t)orr * a.ldx + oi;
    float x = 0.0f, y = 0.0f, hv = 0.0f;
    if (owner) {
        x = a.X[o];
        y = a.Y[o];
        if (a.h) hv = a.h[oi];
    }
    float acc[ROWS][RM];
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
#pragma unroll
        for (int r = 0; r < RM; ++r) acc[q][r] = 0.0f;
    // round_up(N, 16) <= ld elements per row and replica are in range (the
    // padding past N is zero); rows past N have no range and read zeros.
    const int n16 = (a.N + 15) & ~15;
    __amdgpu_buffer_rsrc_t rj[ROWS], rx[RM];
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
        rj[q] = tile_rsrc((const unsigned char *)a.J + (size_t)(row0 + q) * a.ldj * JB,
                          row0 + q < a.N ? (long long)n16 * JB : 0);
#pragma unroll
    for (int r = 0; r < RM; ++r)
        rx[r] = tile_rsrc(a.X + (size_t)(RM == 1 || r < a.R ? r : 0) * a.ldx, (long long)n16 * 4);
    using G = ThinGroup<ROWS, JB, RM>;
    auto load = [&](int k0, G &g) {
#pragma unroll
        for (int q = 0; q < ROWS; ++q)
#pragma unroll
            for (int e = 0; e < JB; ++e) g.j[q].w[e] = rload16(rj[q], (unsigned)(k0 * JB + 16 * e));
#pragma unroll
        for (int r = 0; r < RM; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) g.x[r][e] = rload16(rx[r], (unsigned)(4 * k0 + 16 * e));
    };
    auto sum = [&](const G &g) {
        float xv[RM][16];
#pragma unroll
        for (int r = 0; r < RM; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const v4f t = __builtin_bit_cast(v4f, g.x[r][e]);
#pragma unroll
                for (int c = 0; c < 4; ++c) xv[r][4 * e + c] = SGN ? sgn1(t[c]) : t[c];
            }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            float jv[16];
            thin_widen_j<JB>(g.j[q], jv);
#pragma unroll
            for (int r = 0; r < RM; ++r)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[q][r] = __builtin_fmaf(jv[e], xv[r][e], acc[q][r]);
        }
    };
    // Two groups in flight when their load registers fit VBUDGET (the
    // single-instance kernels have the whole register file; the batched
    // kernel, at two waves per SIMD, passes 0).
    constexpr bool PIPE = 2 * (ROWS * 4 * JB + RM * 16) <= VBUDGET;
    const int groups = (a.N + 1023) >> 10;  // wave-uniform
    int k0 = 16 * lane;
    if constexpr (PIPE) {
        G ga, gb;
        load(k0, ga);
        for (int it = 0; it < groups; it += 2) {
            load(k0 + 1024, gb);
            if (k0 < a.N) sum(ga);
            if (it + 1 >= groups) break;
            load(k0 + 2048, ga);
            if (k0 + 1024 < a.N) sum(gb);
            k0 += 2048;
        }
    } else {
        for (int it = 0; it < groups; ++it, k0 += 1024) {
            G g;
            load(k0, g);
            if (k0 < a.N) sum(g);
        }
    }
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
#pragma unroll
        for (int r = 0; r < RM; ++r) acc[q][r] = wave_sum_xor(acc[q][r]);
    if (!owner) return;
    float F = 0.0f;
#pragma unroll
    for (int qq = 0; qq < ROWS; ++qq)
#pragma unroll
        for (int rr = 0; rr < RM; ++rr)
            if (qq == oq && rr == orr) F = acc[qq][rr];
    if (a.h) F = F + hv;  // external field, after the sum
    sb_update(a.c, F, x, y);
    a.Xn[o] = x;
    a.Y[o] = y;
    if (a.Bn) write_plane1(a.bkind, x, a.Bn, o);
}

// ---------------------------------------------------------------------------
// Sparse family (CSR J, spin-major state).  For low-density J the field is a
// gather: wave w of tile (tm, tn) owns spin i = 4 tm + w and the 64 replicas
// 64 tn + lane (lane = replica).  For each nonzero (j, w_ij) of row i, in
// ascending j, the wave reads the 64 consecutive replicas of X^T[j] (one
// 256-byte load) and accumulates fmaf(w_ij, x_j, acc): the sequential
// j = 0..N-1 chain of Bifurcation_m2.cu.c:93-97 without its zero terms (a zero
// term changes at most the sign of a zero sum), so it equals the ordered path.
// The row's indices and values are wave-uniform (scalar loads).  No barriers.
__device__ __forceinline__ void sparse_tile(const SparseArgs &a, int tm, int tn)
{
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(tm * 4 + (threadIdx.x >> 6));
    if (i >= a.N) return;
    const int r = tn * 64 + lane;  // < ldr: padding replicas hold 0
    const bool sgn = a.c.variant == SB_DSB;
    const int k0 = a.ptr[i], k1 = a.ptr[i + 1];
    const float *xr = a.XT + r;
    float acc = 0.0f;
#pragma unroll 8
    for (int k = k0; k < k1; ++k) {
        const int j = a.col[k];
        float v = xr[(size_t)j * a.ldr];
        if (sgn) v = sgn1(v);
        acc = __builtin_fmaf(a.w[k], v, acc);
    }
    if (r >= a.R) return;
    if (a.h) acc = acc + a.h[i];  // external field, after the sum
    const size_t o = (size_t)i * a.ldr + r;
    float x = a.XT[o], y = a.YT[o];
    sb_update(a.c, acc, x, y);
    a.XTn[o] = x;
    a.YT[o] = y;
}

// ---------------------------------------------------------------------------
// Shape tables (4-wave and 8-wave workgroups).

#define SB_FP32_SHAPES(X) X(2, 2, 1, 2) X(2, 2, 2, 2) X(2, 2, 2, 1) X(2, 2, 1, 1) X(1, 4, 1, 1) X(4, 1, 1, 1) \
    X(2, 4, 1, 1) X(4, 2, 1, 1) X(4, 2, 1, 2) X(2, 4, 2, 1) X(4, 2, 2, 1) X(2, 2, 4, 1) X(2, 2, 2, 4)
#define SB_SPLIT_SHAPES(X) X(2, 2, 2, 2) X(2, 2, 2, 1) X(2, 2, 1, 1) X(4, 1, 1, 1) \
    X(4, 2, 2, 2) X(4, 2, 2, 1) X(4, 2, 1, 1) X(2, 4, 1, 1) X(2, 2, 4, 2) X(2, 2, 4, 1) X(4, 4, 2, 1)
#define SB_FP8_SHAPES(X) X(2, 2, 2, 2) X(2, 2, 2, 1) X(2, 2, 1, 1) X(4, 1, 1, 1) \
    X(4, 2, 2, 2) X(2, 4, 2, 2) X(4, 2, 2, 1) X(4, 2, 1, 1) X(2, 4, 1, 1) X(2, 2, 4, 2) X(2, 2, 4, 1) \
    X(4, 4, 2, 1)

// Thin shapes: spins per wave (4 waves per workgroup).
#define SB_THIN_SHAPES(X) X(1) X(2) X(4) X(8)
#define SB_THIN_ENTRY(rows) TileShape{4, 1, 1, 1, rows},

// Kernel tables (each defined in the translation unit that instantiates it).
typedef void (*StepKernel)(StepArgs);
#define SB_COUNT_ONE(wm, wn, bm, bn) +1
#define SB_COUNT_ONE1(rows) +1
constexpr int kNumFp32Shapes = 0 SB_FP32_SHAPES(SB_COUNT_ONE);
constexpr int kNumSplitShapes = 0 SB_SPLIT_SHAPES(SB_COUNT_ONE);
constexpr int kNumFp8Shapes = 0 SB_FP8_SHAPES(SB_COUNT_ONE);
constexpr int kNumThinShapes = 0 SB_THIN_SHAPES(SB_COUNT_ONE1);
extern const StepKernel kFp32Kernels[kNumFp32Shapes];
// [B source: hi/lo from X, sign from X, sign plane][shape], BF16 J and FP8 J.
extern const StepKernel kSplitBf16Kernels[3][kNumSplitShapes];
extern const StepKernel kSplitFp8Kernels[3][kNumSplitShapes];
// [sign source: X, plane][shape]
extern const StepKernel kDsbFp8Kernels[2][kNumFp8Shapes];
// [J format: FP32, BF16, FP8][sign x][replicas per kernel: 1, 2, 4][shape]
constexpr int kNumThinRm = 3;
__host__ __device__ constexpr int thin_rm_index(int R) { return R <= 1 ? 0 : R == 2 ? 1 : 2; }
extern const StepKernel kThinKernels[3][2][kNumThinRm][kNumThinShapes];

}  // namespace sbk
