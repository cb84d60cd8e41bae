// sb_batch_kernel.hip -- the persistent multi-instance step kernel
// (k_batch_step) and its shape tables.
#include "sb_tiles.h"

namespace sbk {

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
// Sparse tile: 4 spins (one per wave) x 64 replicas.
static const TileShape kBatchSparse[] = {TileShape{4, 1, 1, 1, 1}};
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
    case FAM_SPARSE: return 1;
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
    case FAM_SPARSE: return kBatchSparse[idx];
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
        // Thin items load one 16-element group at a time (no second group in
        // flight): at two waves per SIMD the kernel has 256 VGPRs, and the
        // pipelined thin loop at four replicas would spill the whole kernel
        // to scratch.  The sums are the same either way.
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
#define SB_THIN_CASES(fam, jb)                                                                          \
        case (fam * 2 + 0) * 8 + 0: thin_tile<1, jb, false, kThinMaxR, 0>(a, item.tile); break;         \
        case (fam * 2 + 0) * 8 + 1: thin_tile<2, jb, false, kThinMaxR, 0>(a, item.tile); break;         \
        case (fam * 2 + 0) * 8 + 2: thin_tile<4, jb, false, kThinMaxR, 0>(a, item.tile); break;         \
        case (fam * 2 + 1) * 8 + 0: thin_tile<1, jb, true, kThinMaxR, 0>(a, item.tile); break;          \
        case (fam * 2 + 1) * 8 + 1: thin_tile<2, jb, true, kThinMaxR, 0>(a, item.tile); break;          \
        case (fam * 2 + 1) * 8 + 2: thin_tile<4, jb, true, kThinMaxR, 0>(a, item.tile); break;
        SB_THIN_CASES(FAM_THIN_F32, 4)
        SB_THIN_CASES(FAM_THIN_BF16, 2)
        SB_THIN_CASES(FAM_THIN_FP8, 1)
#undef SB_THIN_CASES
        case (FAM_SPARSE * 2 + 0) * 8 + 0:
        case (FAM_SPARSE * 2 + 1) * 8 + 0: {
            SparseArgs sa;
            sa.ptr = I.csr_ptr;
            sa.col = I.csr_col;
            sa.w = I.csr_w;
            sa.h = I.h;
            sa.XT = I.XT[k & 1];
            sa.XTn = I.XT[(k + 1) & 1];
            sa.YT = I.YT;
            sa.N = I.N;
            sa.R = I.R;
            sa.ldr = I.ldr;
            sa.c = a.c;
            sparse_tile(sa, item.tile % I.tiles_m, item.tile / I.tiles_m);
            break;
        }
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

}  // namespace sbk
