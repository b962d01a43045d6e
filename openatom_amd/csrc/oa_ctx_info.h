// oa_ctx_info.h -- library-internal facts of a context for host-only code
// (not part of the C ABI; no HIP types).
#pragma once

#include "oa_step.h"

int oa_ctx_nstates(const oa_ctx* c);
void oa_ctx_grid(const oa_ctx* c, int grid[3]);
// The conjugate-gradient state of src (direction and gamma history) into
// dst, a context of the same layout whose coefficients were just set: its
// next CG step continues src's minimisation instead of restarting it.
// Waits for both contexts' streams.
oa_status oa_ctx_copy_cg(oa_ctx* dst, oa_ctx* src);
