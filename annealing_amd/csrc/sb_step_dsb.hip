// sb_step_dsb.hip -- dSB on FP8 J x FP8 sign(X)
// (v_mfma_scale_f32_32x32x64_f8f6f4 with unit scales).  Indexed
// [sign source: from X, from the sign plane][shape].
#include "sb_tiles.h"

namespace sbk {

template <int WM, int WN, int BM, int BN, bool PLANES>
__global__ __launch_bounds__(64 * WM * WN) void k_step_dsb_fp8(StepArgs a)
{
    using T = Fp8Tile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * (T::MT + T::NT) * T::LDS_LD];
    dsb_fp8_tile<WM, WN, BM, BN, PLANES>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

#define SB_DSB_FP8_PTR(wm, wn, bm, bn) &k_step_dsb_fp8<wm, wn, bm, bn, false>,
#define SB_DSB_FP8_PL_PTR(wm, wn, bm, bn) &k_step_dsb_fp8<wm, wn, bm, bn, true>,
const StepKernel kDsbFp8Kernels[2][kNumFp8Shapes] = {{SB_FP8_SHAPES(SB_DSB_FP8_PTR)},
                                                    {SB_FP8_SHAPES(SB_DSB_FP8_PL_PTR)}};

}  // namespace sbk
