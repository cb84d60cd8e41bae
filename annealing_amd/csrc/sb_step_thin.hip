// sb_step_thin.hip -- matrix-vector step kernels for R <= kThinMaxR (FP32
// fmaf on the VALU, J streamed once per step).  Indexed
// [J format: FP32, BF16, FP8][sign x][shape].
#include "sb_tiles.h"

namespace sbk {

template <int ROWS, int JB, bool SGN>
__global__ __launch_bounds__(256) void k_step_thin(StepArgs a)
{
    thin_tile<ROWS, JB, SGN>(a, blockIdx.x);
}

#define SB_THIN_PTR_F32(rows) &k_step_thin<rows, 4, false>,
#define SB_THIN_PTR_F32S(rows) &k_step_thin<rows, 4, true>,
#define SB_THIN_PTR_BF16(rows) &k_step_thin<rows, 2, false>,
#define SB_THIN_PTR_BF16S(rows) &k_step_thin<rows, 2, true>,
#define SB_THIN_PTR_FP8(rows) &k_step_thin<rows, 1, false>,
#define SB_THIN_PTR_FP8S(rows) &k_step_thin<rows, 1, true>,
const StepKernel kThinKernels[3][2][kNumThinShapes] = {
    {{SB_THIN_SHAPES(SB_THIN_PTR_F32)}, {SB_THIN_SHAPES(SB_THIN_PTR_F32S)}},
    {{SB_THIN_SHAPES(SB_THIN_PTR_BF16)}, {SB_THIN_SHAPES(SB_THIN_PTR_BF16S)}},
    {{SB_THIN_SHAPES(SB_THIN_PTR_FP8)}, {SB_THIN_SHAPES(SB_THIN_PTR_FP8S)}}};

}  // namespace sbk
