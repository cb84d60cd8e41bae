// sb_step_thin.hip -- matrix-vector step kernels for R <= kThinMaxR (FP32
// fmaf on the VALU, J streamed once per step).  Indexed
// [J format: FP32, BF16, FP8][sign x][replicas per kernel: 1, 2, 4][shape].
#include "sb_tiles.h"

namespace sbk {

template <int ROWS, int JB, bool SGN, int RM>
__global__ __launch_bounds__(256) void k_step_thin(StepArgs a)
{
    thin_tile<ROWS, JB, SGN, RM>(a, blockIdx.x);
}

// One row of the table: the shapes of SB_THIN_SHAPES (1, 2, 4, 8 spins per wave).
#define SB_THIN_ROW(jb, sgn, rm)                                                               \
    {&k_step_thin<1, jb, sgn, rm>, &k_step_thin<2, jb, sgn, rm>, &k_step_thin<4, jb, sgn, rm>, \
     &k_step_thin<8, jb, sgn, rm>}
#define SB_THIN_SIGN(jb, sgn) {SB_THIN_ROW(jb, sgn, 1), SB_THIN_ROW(jb, sgn, 2), SB_THIN_ROW(jb, sgn, 4)}
static_assert(kNumThinShapes == 4 && kNumThinRm == 3, "thin kernel table layout");
const StepKernel kThinKernels[3][2][kNumThinRm][kNumThinShapes] = {
    {SB_THIN_SIGN(4, false), SB_THIN_SIGN(4, true)},
    {SB_THIN_SIGN(2, false), SB_THIN_SIGN(2, true)},
    {SB_THIN_SIGN(1, false), SB_THIN_SIGN(1, true)}};

}  // namespace sbk
