// sb_batch.cpp -- batched solve of many instances.
#include "sb_host.h"

extern "C" SB_API int sb_batch_solve(sb_ctx *, int32_t, sb_problem *const *, const sb_params *, sb_result *,
                                     sb_batch_stats *)
{
    sbh::set_error("sb_batch_solve: not built yet");
    return SB_E_INVALID;
}
