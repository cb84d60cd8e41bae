// sb_kernels.hip -- gfx950 kernels of the simulated-bifurcation solver.
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

template <int JB>
__device__ __forceinline__ void thin_load_j(const void *J, size_t o, float (&v)[16])
{
    if constexpr (JB == 4) {
        const v4f *p = (const v4f *)((const float *)J + o);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const v4f t = p[g];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[4 * g + e] = t[e];
        }
    } else if constexpr (JB == 2) {
        const u32x4 *p = (const u32x4 *)((const unsigned short *)J + o);
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const u32x4 t = p[g];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[8 * g + 2 * e] = __builtin_bit_cast(float, t[e] << 16);
                v[8 * g + 2 * e + 1] = __builtin_bit_cast(float, t[e] & 0xffff0000u);
            }
        }
    } else {
        const u32x4 t = *(const u32x4 *)((const unsigned char *)J + o);
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

template <int ROWS, int JB, bool SGN>
__device__ __forceinline__ void thin_tile(const StepArgs &a, int tile)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (tile * 4 + wave) * ROWS;
    if (row0 >= a.N) return;
    float acc[ROWS][kThinMaxR];
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
#pragma unroll
        for (int r = 0; r < kThinMaxR; ++r) acc[q][r] = 0.0f;
    for (int k0 = 16 * lane; k0 < a.N; k0 += 1024) {
        // k0 + 16 <= round_up(N, 16) <= ld: the group is in bounds.
        float jv[ROWS][16];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            if (row0 + q < a.N) {
                thin_load_j<JB>(a.J, (size_t)(row0 + q) * a.ldj + k0, jv[q]);
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) jv[q][e] = 0.0f;
            }
        }
#pragma unroll
        for (int r = 0; r < kThinMaxR; ++r) {
            if (r >= a.R) break;
            const v4f *xp = (const v4f *)(a.X + (size_t)r * a.ldx + k0);
            float xv[16];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const v4f t = xp[g];
#pragma unroll
                for (int e = 0; e < 4; ++e) xv[4 * g + e] = SGN ? sgn1(t[e]) : t[e];
            }
#pragma unroll
            for (int q = 0; q < ROWS; ++q)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[q][r] = __builtin_fmaf(jv[q][e], xv[e], acc[q][r]);
        }
    }
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
#pragma unroll
        for (int r = 0; r < kThinMaxR; ++r)
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) acc[q][r] += __shfl_xor(acc[q][r], m);
    // Lane q * kThinMaxR + r updates spin row0 + q of replica r.
    if (lane >= ROWS * kThinMaxR) return;
    const int q = lane / kThinMaxR, r = lane % kThinMaxR, i = row0 + q;
    if (r >= a.R || i >= a.N) return;
    float F = 0.0f;
#pragma unroll
    for (int qq = 0; qq < ROWS; ++qq)
#pragma unroll
        for (int rr = 0; rr < kThinMaxR; ++rr)
            if (qq == q && rr == r) F = acc[qq][rr];
    if (a.h) F = F + a.h[i];  // external field, after the sum
    const size_t o = (size_t)r * a.ldx + i;
    float x = a.X[o], y = a.Y[o];
    sb_update(a.c, F, x, y);
    a.Xn[o] = x;
    a.Y[o] = y;
    if (a.Bn) write_plane1(a.bkind, x, a.Bn, o);
}

// ---------------------------------------------------------------------------
// Kernel entry points per family and shape.

template <int WM, int WN, int BM, int BN>
__global__ __launch_bounds__(64 * WM * WN) void k_step_fp32(StepArgs a)
{
    using T = Fp32Tile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) float smem[2 * (T::MT + T::NT) * T::LDS_LD];
    fp32_tile<WM, WN, BM, BN>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

template <int ROWS, int JB, bool SGN>
__global__ __launch_bounds__(256) void k_step_thin(StepArgs a)
{
    thin_tile<ROWS, JB, SGN>(a, blockIdx.x);
}

template <int WM, int WN, int BM, int BN, bool FP8J, bool PLANES, bool SGN>
__global__ __launch_bounds__(64 * WM * WN) void k_step_split(StepArgs a)
{
    using T = SplitTile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) unsigned short smem[2 * (T::MT + 2 * T::NT) * T::LDS_LD];
    split_tile<WM, WN, BM, BN, FP8J, PLANES, SGN>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

template <int WM, int WN, int BM, int BN, bool PLANES>
__global__ __launch_bounds__(64 * WM * WN) void k_step_dsb_fp8(StepArgs a)
{
    using T = Fp8Tile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * (T::MT + T::NT) * T::LDS_LD];
    dsb_fp8_tile<WM, WN, BM, BN, PLANES>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

// Ordered-sum debug kernel: one thread per (spin, replica), F_i summed over
// j = 0..N-1 with fmaf in order (Bifurcation_m2.cu.c:93-97).
__global__ __launch_bounds__(256) void k_step_ordered(StepArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (i >= a.N || r >= a.R) return;
    const float *Jr = (const float *)a.J + (size_t)i * a.ldj;
    const float *xr = a.X + (size_t)r * a.ldx;
    const bool sgn = a.c.variant == SB_DSB;
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

#define SB_SHAPE_ENTRY(wm, wn, bm, bn) TileShape{wm, wn, bm, bn, 0},
#define SB_THIN_ENTRY(rows) TileShape{4, 1, 1, 1, rows},
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

typedef void (*StepKernel)(StepArgs);

#define SB_FP32_PTR(wm, wn, bm, bn) &k_step_fp32<wm, wn, bm, bn>,
#define SB_DSB_FP8_PTR(wm, wn, bm, bn) &k_step_dsb_fp8<wm, wn, bm, bn, false>,
#define SB_DSB_FP8_PL_PTR(wm, wn, bm, bn) &k_step_dsb_fp8<wm, wn, bm, bn, true>,
static const StepKernel kFp32Kernels[] = {SB_FP32_SHAPES(SB_FP32_PTR)};
static const StepKernel kDsbFp8Kernels[] = {SB_FP8_SHAPES(SB_DSB_FP8_PTR)};
static const StepKernel kDsbFp8PlKernels[] = {SB_FP8_SHAPES(SB_DSB_FP8_PL_PTR)};
// Split kernels indexed [FP8 J][B source: hi/lo from X, sign from X, sign plane][shape].
#define SB_SPLIT_PTR(F, P, G) \
    {SB_SPLIT_SHAPES(SB_SPLIT_PTR_1##F##P##G)}
#define SB_SPLIT_PTR_1000(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, false, false, false>,
#define SB_SPLIT_PTR_1001(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, false, false, true>,
#define SB_SPLIT_PTR_1011(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, false, true, true>,
#define SB_SPLIT_PTR_1100(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, false, false>,
#define SB_SPLIT_PTR_1101(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, false, true>,
#define SB_SPLIT_PTR_1111(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, true, true>,
#define SB_COUNT_ONE(wm, wn, bm, bn) +1
#define SB_COUNT_ONE1(rows) +1
static const StepKernel kSplitKernels[2][3][0 SB_SPLIT_SHAPES(SB_COUNT_ONE)] = {
    {SB_SPLIT_PTR(0, 0, 0), SB_SPLIT_PTR(0, 0, 1), SB_SPLIT_PTR(0, 1, 1)},
    {SB_SPLIT_PTR(1, 0, 0), SB_SPLIT_PTR(1, 0, 1), SB_SPLIT_PTR(1, 1, 1)}};

// Thin kernels indexed [J format: FP32, BF16, FP8][sign x][shape].
#define SB_THIN_PTR_F32(rows) &k_step_thin<rows, 4, false>,
#define SB_THIN_PTR_F32S(rows) &k_step_thin<rows, 4, true>,
#define SB_THIN_PTR_BF16(rows) &k_step_thin<rows, 2, false>,
#define SB_THIN_PTR_BF16S(rows) &k_step_thin<rows, 2, true>,
#define SB_THIN_PTR_FP8(rows) &k_step_thin<rows, 1, false>,
#define SB_THIN_PTR_FP8S(rows) &k_step_thin<rows, 1, true>,
static const StepKernel kThinKernels[3][2][0 SB_THIN_SHAPES(SB_COUNT_ONE1)] = {
    {{SB_THIN_SHAPES(SB_THIN_PTR_F32)}, {SB_THIN_SHAPES(SB_THIN_PTR_F32S)}},
    {{SB_THIN_SHAPES(SB_THIN_PTR_BF16)}, {SB_THIN_SHAPES(SB_THIN_PTR_BF16S)}},
    {{SB_THIN_SHAPES(SB_THIN_PTR_FP8)}, {SB_THIN_SHAPES(SB_THIN_PTR_FP8S)}}};

hipError_t launch_step(int family, int shape_idx, const StepArgs &a_in, hipStream_t s)
{
    if (shape_idx < 0 || shape_idx >= shape_count(family)) return hipErrorInvalidValue;
    TileShape sh = shape_at(family, shape_idx);
    StepArgs a = a_in;
    if (is_thin(family)) {
        if (a.R > kThinMaxR) return hipErrorInvalidValue;
        a.tiles_m = (a.N + sh.mt() - 1) / sh.mt();
        a.tiles_n = 1;
        a.rpg = 1;
        StepKernel k = kThinKernels[family - FAM_THIN_F32][a.c.variant == SB_DSB][shape_idx];
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
        k = kSplitKernels[family == FAM_SPLIT_FP8][a.c.variant != SB_DSB ? 0 : pl ? 2 : 1][shape_idx];
        break;
    case FAM_DSB_FP8: k = pl ? kDsbFp8PlKernels[shape_idx] : kDsbFp8Kernels[shape_idx]; break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(sh.threads()), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sparse family (CSR J, spin-major state).  For low-density J the field is a
// gather: wave w of workgroup t owns spin i = 4t + w and the 64 replicas
// 64 blockIdx.y + lane (lane = replica).  For each nonzero (j, w_ij) of row i,
// in ascending j, the wave reads the 64 consecutive replicas of X^T[j] (one
// 256-byte load) and accumulates fmaf(w_ij, x_j, acc): the sequential
// j = 0..N-1 chain of Bifurcation_m2.cu.c:93-97 without its zero terms (a zero
// term changes at most the sign of a zero sum), so it equals the ordered path.
// The row's indices and values are wave-uniform (scalar loads).

__global__ __launch_bounds__(256) void k_step_sparse(SparseArgs a)
{
    const int lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= a.N) return;
    const int r = blockIdx.y * 64 + lane;  // < ldr: padding replicas hold 0
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
    dim3 grid((a.N + 255) / 256, a.R);
    hipLaunchKernelGGL(k_step_ordered, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched persistent step: many instances, one launch per step.  Each
// workgroup walks its item list; an item runs exactly the tile function the
// single-instance kernel of the same family and shape runs, so an instance's
// result does not depend on what it is batched with.

#define SB_BATCH_FP32(X) X(2, 2, 1, 1) X(2, 2, 2, 1) X(4, 1, 1, 1) X(1, 4, 1, 1) X(2, 2, 1, 2)
#define SB_BATCH_SPLIT(X) X(2, 2, 1, 1) X(2, 2, 2, 1) X(4, 1, 1, 1)
#define SB_BATCH_FP8(X) X(2, 2, 1, 1) X(2, 2, 2, 1) X(4, 1, 1, 1)

#define SB_SHAPE_ENTRY(wm, wn, bm, bn) TileShape{wm, wn, bm, bn},
static const TileShape kBatchFp32[] = {SB_BATCH_FP32(SB_SHAPE_ENTRY)};
static const TileShape kBatchSplit[] = {SB_BATCH_SPLIT(SB_SHAPE_ENTRY)};
static const TileShape kBatchFp8[] = {SB_BATCH_FP8(SB_SHAPE_ENTRY)};
static const TileShape kBatchThin[] = {SB_THIN_ENTRY(1) SB_THIN_ENTRY(2) SB_THIN_ENTRY(4)};
#undef SB_SHAPE_ENTRY

constexpr size_t kBatchLds = 73728;  // max over the shapes above (checked below)

// LDS bytes of one tile of each family (the layouts of fp32_tile, split_tile
// and dsb_fp8_tile): every batched shape must fit the dynamic allocation.
template <int WM, int WN, int BM, int BN>
constexpr size_t fp32_lds()
{
    using T = Fp32Tile<WM, WN, BM, BN>;
    return 2 * (size_t)(T::MT + T::NT) * T::LDS_LD * sizeof(float);
}
template <int WM, int WN, int BM, int BN>
constexpr size_t split_lds()
{
    using T = SplitTile<WM, WN, BM, BN>;
    return 2 * (size_t)(T::MT + 2 * T::NT) * T::LDS_LD * sizeof(unsigned short);
}
template <int WM, int WN, int BM, int BN>
constexpr size_t fp8_lds()
{
    using T = Fp8Tile<WM, WN, BM, BN>;
    return 2 * (size_t)(T::MT + T::NT) * T::LDS_LD;
}
#define SB_CHECK_FP32(wm, wn, bm, bn) static_assert(fp32_lds<wm, wn, bm, bn>() <= kBatchLds, "batched FP32 LDS");
#define SB_CHECK_SPLIT(wm, wn, bm, bn) static_assert(split_lds<wm, wn, bm, bn>() <= kBatchLds, "batched split LDS");
#define SB_CHECK_FP8(wm, wn, bm, bn) static_assert(fp8_lds<wm, wn, bm, bn>() <= kBatchLds, "batched FP8 LDS");
SB_BATCH_FP32(SB_CHECK_FP32)
SB_BATCH_SPLIT(SB_CHECK_SPLIT)
SB_BATCH_FP8(SB_CHECK_FP8)
#undef SB_CHECK_FP32
#undef SB_CHECK_SPLIT
#undef SB_CHECK_FP8

int batch_shape_count(int family)
{
    switch (family) {
    case FAM_FP32: return (int)(sizeof(kBatchFp32) / sizeof(kBatchFp32[0]));
    case FAM_SPLIT_BF16:
    case FAM_SPLIT_FP8: return (int)(sizeof(kBatchSplit) / sizeof(kBatchSplit[0]));
    case FAM_DSB_FP8: return (int)(sizeof(kBatchFp8) / sizeof(kBatchFp8[0]));
    case FAM_THIN_F32:
    case FAM_THIN_BF16:
    case FAM_THIN_FP8: return (int)(sizeof(kBatchThin) / sizeof(kBatchThin[0]));
    default: return 0;
    }
}

TileShape batch_shape_at(int family, int idx)
{
    switch (family) {
    case FAM_FP32: return kBatchFp32[idx];
    case FAM_SPLIT_BF16:
    case FAM_SPLIT_FP8: return kBatchSplit[idx];
    case FAM_THIN_F32:
    case FAM_THIN_BF16:
    case FAM_THIN_FP8: return kBatchThin[idx];
    default: return kBatchFp8[idx];
    }
}

int batch_shape_solo_index(int family, int idx)
{
    TileShape b = batch_shape_at(family, idx);
    for (int s = 0; s < shape_count(family); ++s) {
        TileShape t = shape_at(family, s);
        if (t.wm == b.wm && t.wn == b.wn && t.bm == b.bm && t.bn == b.bn && t.rows == b.rows) return s;
    }
    return -1;
}

size_t batch_lds_bytes() { return kBatchLds; }

__global__ __launch_bounds__(256, 2) void k_batch_step(const BatchInstance *__restrict__ insts,
                                                    const BatchItem *__restrict__ items,
                                                    const int *__restrict__ wg_begin, int chunk_step)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_dyn[];
    const int w = blockIdx.x;
    const int b = wg_begin[w], e = wg_begin[w + 1];
    for (int it = b; it < e; ++it) {
        const BatchItem item = items[it];
        const BatchInstance &I = insts[item.inst];
        const int k = I.step_base + chunk_step;
        if (!I.active || k < 0 || k >= I.steps) continue;
        StepArgs a;
        a.J = I.J;
        a.h = I.h;
        a.X = I.X[k & 1];
        a.Xn = I.X[(k + 1) & 1];
        a.Y = I.Y;
        a.B = I.P[k & 1];
        a.Bn = I.P[(k + 1) & 1];
        a.bkind = I.bkind;
        a.ldj = I.ldx;
        a.ldx = I.ldx;
        a.N = I.N;
        a.R = I.R;
        a.tiles_m = I.tiles_m;
        a.tiles_n = I.tiles_n;
        a.rpg = 0;
        a.c = step_coeffs(I.variant, k, I.steps, I.dt, I.a0, I.c0, I.K, I.M);
        const int sel = (I.family * 2 + (I.variant == SB_DSB)) * 8 + I.bshape;
        unsigned short *sm16 = (unsigned short *)smem_dyn;
        // One case per (family, sign B, batch shape); must follow the SB_BATCH_* lists.
        switch (sel) {
        case (FAM_FP32 * 2 + 0) * 8 + 0:
        case (FAM_FP32 * 2 + 1) * 8 + 0: fp32_tile<2, 2, 1, 1>(a, item.tile, (float *)smem_dyn); break;
        case (FAM_FP32 * 2 + 0) * 8 + 1:
        case (FAM_FP32 * 2 + 1) * 8 + 1: fp32_tile<2, 2, 2, 1>(a, item.tile, (float *)smem_dyn); break;
        case (FAM_FP32 * 2 + 0) * 8 + 2:
        case (FAM_FP32 * 2 + 1) * 8 + 2: fp32_tile<4, 1, 1, 1>(a, item.tile, (float *)smem_dyn); break;
        case (FAM_FP32 * 2 + 0) * 8 + 3:
        case (FAM_FP32 * 2 + 1) * 8 + 3: fp32_tile<1, 4, 1, 1>(a, item.tile, (float *)smem_dyn); break;
        case (FAM_FP32 * 2 + 0) * 8 + 4:
        case (FAM_FP32 * 2 + 1) * 8 + 4: fp32_tile<2, 2, 1, 2>(a, item.tile, (float *)smem_dyn); break;
        case (FAM_SPLIT_BF16 * 2 + 0) * 8 + 0: split_tile<2, 2, 1, 1, false, false, false>(a, item.tile, sm16); break;
        case (FAM_SPLIT_BF16 * 2 + 0) * 8 + 1: split_tile<2, 2, 2, 1, false, false, false>(a, item.tile, sm16); break;
        case (FAM_SPLIT_BF16 * 2 + 0) * 8 + 2: split_tile<4, 1, 1, 1, false, false, false>(a, item.tile, sm16); break;
        case (FAM_SPLIT_BF16 * 2 + 1) * 8 + 0: split_tile<2, 2, 1, 1, false, true, true>(a, item.tile, sm16); break;
        case (FAM_SPLIT_BF16 * 2 + 1) * 8 + 1: split_tile<2, 2, 2, 1, false, true, true>(a, item.tile, sm16); break;
        case (FAM_SPLIT_BF16 * 2 + 1) * 8 + 2: split_tile<4, 1, 1, 1, false, true, true>(a, item.tile, sm16); break;
        case (FAM_SPLIT_FP8 * 2 + 0) * 8 + 0: split_tile<2, 2, 1, 1, true, false, false>(a, item.tile, sm16); break;
        case (FAM_SPLIT_FP8 * 2 + 0) * 8 + 1: split_tile<2, 2, 2, 1, true, false, false>(a, item.tile, sm16); break;
        case (FAM_SPLIT_FP8 * 2 + 0) * 8 + 2: split_tile<4, 1, 1, 1, true, false, false>(a, item.tile, sm16); break;
        case (FAM_DSB_FP8 * 2 + 1) * 8 + 0: dsb_fp8_tile<2, 2, 1, 1, true>(a, item.tile, smem_dyn); break;
        case (FAM_DSB_FP8 * 2 + 1) * 8 + 1: dsb_fp8_tile<2, 2, 2, 1, true>(a, item.tile, smem_dyn); break;
        case (FAM_DSB_FP8 * 2 + 1) * 8 + 2: dsb_fp8_tile<4, 1, 1, 1, true>(a, item.tile, smem_dyn); break;
#define SB_THIN_CASES(fam, jb)                                                                  \
        case (fam * 2 + 0) * 8 + 0: thin_tile<1, jb, false>(a, item.tile); break;              \
        case (fam * 2 + 0) * 8 + 1: thin_tile<2, jb, false>(a, item.tile); break;              \
        case (fam * 2 + 0) * 8 + 2: thin_tile<4, jb, false>(a, item.tile); break;              \
        case (fam * 2 + 1) * 8 + 0: thin_tile<1, jb, true>(a, item.tile); break;               \
        case (fam * 2 + 1) * 8 + 1: thin_tile<2, jb, true>(a, item.tile); break;               \
        case (fam * 2 + 1) * 8 + 2: thin_tile<4, jb, true>(a, item.tile); break;
        SB_THIN_CASES(FAM_THIN_F32, 4)
        SB_THIN_CASES(FAM_THIN_BF16, 2)
        SB_THIN_CASES(FAM_THIN_FP8, 1)
#undef SB_THIN_CASES
        default: break;
        }
        __syncthreads();  // LDS is reused by the next item
    }
}

hipError_t launch_batch_step(int grid, const BatchInstance *insts, const BatchItem *items, const int *wg_begin,
                             int chunk_step, hipStream_t s)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_step, dim3(grid), dim3(256), kBatchLds, s, insts, items, wg_begin, chunk_step);
    return hipGetLastError();
}

int batch_grid_size(int device)
{
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_batch_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kBatchLds);
        attr_set = true;
    }
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_step, 256, kBatchLds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return cus * per_cu;
}

// ---------------------------------------------------------------------------
// Synthetic instances on the device (gen_coupling in sb_internal.h).

__global__ __launch_bounds__(256) void k_generate(int kind, int N, double p, uint64_t seed, float *J, int ld)
{
    const int i = blockIdx.y;
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
    hipLaunchKernelGGL(k_generate, dim3(gx < 16 ? gx : 16, N), dim3(256), 0, s, kind, N, p, seed, J, ld);
    return hipGetLastError();
}

// One workgroup per row: sum of squares (FP64, fixed tree), max |J|, and
// whether every entry is an integer.
__global__ __launch_bounds__(256) void k_row_stats(const float *J, int N, int ld, double *row_ss, float *row_max,
                                                   int *row_int)
{
    __shared__ double ss[256];
    __shared__ float mx[256];
    __shared__ int in[256];
    const int i = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    float m = 0.0f;
    int integ = 1;
    for (int j = t; j < N; j += 256) {
        float v = J[(size_t)i * ld + j];
        s += (double)v * (double)v;
        m = fmaxf(m, fabsf(v));
        if (v != floorf(v)) integ = 0;
    }
    ss[t] = s;
    mx[t] = m;
    in[t] = integ;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            mx[t] = fmaxf(mx[t], mx[t + w]);
            in[t] &= in[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        row_ss[i] = ss[0];
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
    int r = blockIdx.y;
    if (i >= ldx || r >= R) return;
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
        hipLaunchKernelGGL(k_init_uniform, dim3((ldx + 255) / 256, R), dim3(256), 0, s, seed, N, R, ldx, X,
                           Y, X2);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_asb_local(sb_step_coeffs c, int N, int R, int ldx, float *X,
                                                   float *Y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (i >= N || r >= R) return;
    size_t o = (size_t)r * ldx + i;
    float x = X[o], y = Y[o];
    if (c.variant == SB_ASB_M1)
        x = m1_x(x, y, c.a0, c.dts);  // step 0's x-step
    else
        asb_local(x, y, c.p_next, c.a0, c.K, c.dts, c.M);
    X[o] = x;
    Y[o] = y;
}

hipError_t launch_asb_local(const sb_step_coeffs &c, int N, int R, int ldx, float *X, float *Y,
                            hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_asb_local, dim3((N + 255) / 256, R), dim3(256), 0, s, c, N, R, ldx, X, Y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Energy: H_r = -1/2 sum_i s_i G_i with G = J s, in FP64.  Tile 64 spins x 64
// replicas per workgroup; every product is +-J_ij (exact), so for integer J
// all partial sums are exact integers below 2^53 and the result is exact.
// Deterministic: fixed per-thread order and a fixed reduction tree.

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

// The field term -sum_i h_i s_i per replica: one workgroup per replica, a
// fixed per-thread order and a fixed tree (deterministic; exact for integer h).
__global__ __launch_bounds__(256) void k_energy_field(const float *h, const float *X, int ldx, int N, double *H)
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
    if (t == 0) H[r] -= red[0];
}

size_t energy_partial_elems(int N, int R)
{
    return (size_t)((N + EN_T - 1) / EN_T) * (size_t)R;
}

hipError_t launch_energy(const float *J, int ldj, const float *h, const float *X, int ldx, int N, int R,
                         double *partial, double *H, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    int tiles_m = (N + EN_T - 1) / EN_T, tiles_n = (R + EN_T - 1) / EN_T;
    hipLaunchKernelGGL(k_energy_partial, dim3(tiles_m * tiles_n), dim3(256), 0, s, J, ldj, X, ldx, N, R,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_final, dim3((R + 255) / 256), dim3(256), 0, s, partial, tiles_m, R, H);
    e = hipGetLastError();
    if (e != hipSuccess || !h) return e;
    hipLaunchKernelGGL(k_energy_field, dim3(R), dim3(256), 0, s, h, X, ldx, N, H);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Readout and packing.

__global__ __launch_bounds__(256) void k_spins(const float *X, int ldx, int N, int R, int8_t *S)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (i >= N || r >= R) return;
    S[(size_t)r * N + i] = X[(size_t)r * ldx + i] > 0.0f ? (int8_t)1 : (int8_t)-1;
}

hipError_t launch_spins(const float *X, int ldx, int N, int R, int8_t *S, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spins, dim3((N + 255) / 256, R), dim3(256), 0, s, X, ldx, N, R, S);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_compact(const float *X, int ldx, int N, int R, float *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (i >= N || r >= R) return;
    out[(size_t)r * N + i] = X[(size_t)r * ldx + i];
}

hipError_t launch_compact(const float *X, int ldx, int N, int R, float *out, hipStream_t s)
{
    if (N <= 0 || R <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact, dim3((N + 255) / 256, R), dim3(256), 0, s, X, ldx, N, R, out);
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
