// oa_serve_poll.h -- how the request loop reads the event that follows a
// request's step (host code; no device code, so a host compiler can test it).
#pragma once

#include <hip/hip_runtime_api.h>

namespace oa_serve {

// What an event query says about a request's step in flight.
enum class StepPoll { Running, Done, Failed };

// hipSuccess: the step finished; hipErrorNotReady: still running; anything
// else (a launch failure, a sticky device error) ends the request with
// OA_EHIP, so that its context and budget are released instead of being
// polled for ever.
inline StepPoll classify_event(hipError_t q) {
  if (q == hipSuccess) return StepPoll::Done;
  if (q == hipErrorNotReady) return StepPoll::Running;
  return StepPoll::Failed;
}

}  // namespace oa_serve
