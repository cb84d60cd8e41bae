// sb_server.cpp -- continuous-batching request server.
#include "sb_host.h"

extern "C" {
SB_API int sb_server_create(sb_ctx *, const sb_server_opts *, sb_server **) { return SB_E_INVALID; }
SB_API int sb_server_submit(sb_server *, int32_t, const float *, int32_t, const sb_params *, uint64_t *) { return SB_E_INVALID; }
SB_API int sb_server_wait(sb_server *, uint64_t, sb_result *, double) { return SB_E_INVALID; }
SB_API int sb_server_cancel(sb_server *, uint64_t) { return SB_E_INVALID; }
SB_API int sb_server_stats_get(sb_server *, sb_server_stats *) { return SB_E_INVALID; }
SB_API int sb_server_destroy(sb_server *) { return SB_E_INVALID; }
}
