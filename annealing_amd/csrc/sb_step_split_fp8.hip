// sb_step_split_fp8.hip -- split family step kernels with FP8 E4M3 J widened to BF16 while staging
// (v_mfma_f32_32x32x16_bf16; X as BF16 hi + lo, or the BF16 sign for dSB).
// Indexed [B source: hi/lo from X, sign from X, sign plane][shape].
#include "sb_tiles.h"

namespace sbk {

template <int WM, int WN, int BM, int BN, bool FP8J, bool PLANES, bool SGN>
__global__ __launch_bounds__(64 * WM * WN) void k_step_split(StepArgs a)
{
    using T = SplitTile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) unsigned short smem[2 * (T::MT + 2 * T::NT) * T::LDS_LD];
    split_tile<WM, WN, BM, BN, FP8J, PLANES, SGN>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

#define SB_SPLIT_PTR_HL(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, false, false>,
#define SB_SPLIT_PTR_SX(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, false, true>,
#define SB_SPLIT_PTR_SP(wm, wn, bm, bn) &k_step_split<wm, wn, bm, bn, true, true, true>,
const StepKernel kSplitFp8Kernels[3][kNumSplitShapes] = {{SB_SPLIT_SHAPES(SB_SPLIT_PTR_HL)},
                                                        {SB_SPLIT_SHAPES(SB_SPLIT_PTR_SX)},
                                                        {SB_SPLIT_SHAPES(SB_SPLIT_PTR_SP)}};

}  // namespace sbk
