// sb_step_fp32.hip -- FP32 family step kernels (v_mfma_f32_32x32x2_f32), one
// instantiation per tile shape of SB_FP32_SHAPES.
#include "sb_tiles.h"

namespace sbk {

template <int WM, int WN, int BM, int BN>
__global__ __launch_bounds__(64 * WM * WN) void k_step_fp32(StepArgs a)
{
    using T = Fp32Tile<WM, WN, BM, BN>;
    __shared__ __attribute__((aligned(16))) float smem[2 * (T::MT + T::NT) * T::LDS_LD];
    fp32_tile<WM, WN, BM, BN>(a, xcd_remap(blockIdx.x, gridDim.x), smem);
}

#define SB_FP32_PTR(wm, wn, bm, bn) &k_step_fp32<wm, wn, bm, bn>,
const StepKernel kFp32Kernels[kNumFp32Shapes] = {SB_FP32_SHAPES(SB_FP32_PTR)};

}  // namespace sbk
