// sb_batch_kernels.hip -- persistent batched step kernel (filled in below).
#include "sb_internal.h"
