// oa_internal.h -- library-internal access to a context (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "oa_step.h"

// the HIP stream a context's work is enqueued on
hipStream_t oa_ctx_stream(oa_ctx* c);
