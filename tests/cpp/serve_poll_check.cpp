// Host-only check of the request loop's reading of a step's event
// (openatom_amd/csrc/oa_serve_poll.h): completion retires the step, "not
// ready" keeps polling, and every other status fails the request.
#include <cstdio>

#include "oa_serve_poll.h"

using oa_serve::classify_event;
using oa_serve::StepPoll;

int main() {
  int bad = 0;
  auto expect = [&](hipError_t q, StepPoll want, const char* name) {
    if (classify_event(q) != want) {
      std::printf("FAIL %s\n", name);
      bad++;
    }
  };
  expect(hipSuccess, StepPoll::Done, "hipSuccess");
  expect(hipErrorNotReady, StepPoll::Running, "hipErrorNotReady");
  expect(hipErrorLaunchFailure, StepPoll::Failed, "hipErrorLaunchFailure");
  expect(hipErrorIllegalAddress, StepPoll::Failed, "hipErrorIllegalAddress");
  expect(hipErrorInvalidHandle, StepPoll::Failed, "hipErrorInvalidHandle");
  expect(hipErrorInvalidValue, StepPoll::Failed, "hipErrorInvalidValue");
  expect(hipErrorUnknown, StepPoll::Failed, "hipErrorUnknown");
  if (bad == 0) std::printf("serve poll: ok\n");
  return bad == 0 ? 0 : 1;
}
