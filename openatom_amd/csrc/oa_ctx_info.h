// oa_ctx_info.h -- library-internal facts of a context for host-only code
// (not part of the C ABI; no HIP types).
#pragma once

#include "oa_step.h"

int oa_ctx_nstates(const oa_ctx* c);
void oa_ctx_grid(const oa_ctx* c, int grid[3]);
