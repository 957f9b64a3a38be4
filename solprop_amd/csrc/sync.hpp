// Host-side synchronisation helpers: graph capture that always ends, and a bounded wait.
#pragma once

#include <chrono>
#include <string>
#include <thread>

#include "common.hpp"

namespace sp {

// Capture what `enqueue` puts on `stream` into an executable graph. If `enqueue` throws
// (a refused shape, a launch error), the capture is ended and its partial graph
// destroyed before the exception goes on, so the stream leaves capture mode and later
// calls on it work.
template <typename F>
hipGraphExec_t capture_graph(hipStream_t stream, F&& enqueue) {
  SP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue();
  } catch (...) {
    hipGraph_t partial = nullptr;
    if (hipStreamEndCapture(stream, &partial) == hipSuccess && partial)
      (void)hipGraphDestroy(partial);
    (void)hipGetLastError();
    throw;
  }
  hipGraph_t graph = nullptr;
  SP_HIP(hipStreamEndCapture(stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t made = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  SP_HIP(made);
  return exec;
}

// How long a wait for the device may take before it is reported as a failure, seconds.
constexpr double kWaitLimitSeconds = 300.0;

// Wait until `event` completes: spin for the first `spin_us` microseconds (a trial's
// norms are wanted as soon as they land), then yield the core, then sleep in growing
// steps up to a millisecond. A wait past kWaitLimitSeconds throws SP_BACKEND, so a
// device that never finishes cannot hold the host forever.
inline void wait_event(hipEvent_t event, double spin_us = 200.0) {
  using clock = std::chrono::steady_clock;
  const auto start = clock::now();
  auto sleep = std::chrono::microseconds(2);
  for (;;) {
    const hipError_t status = hipEventQuery(event);
    if (status == hipSuccess) return;
    if (status != hipErrorNotReady) SP_HIP(status);
    const double waited =
        std::chrono::duration<double, std::micro>(clock::now() - start).count();
    if (waited < spin_us) continue;
    if (waited > kWaitLimitSeconds * 1e6)
      throw Error(kBackend, "the device did not finish within " +
                                std::to_string(static_cast<int>(kWaitLimitSeconds)) + " s");
    if (waited < 4.0 * spin_us) {
      std::this_thread::yield();
      continue;
    }
    std::this_thread::sleep_for(sleep);
    if (sleep < std::chrono::microseconds(1000)) sleep *= 2;
  }
}

}  // namespace sp
