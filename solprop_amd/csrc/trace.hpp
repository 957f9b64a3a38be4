// Scripted Runge-Kutta decision traces (see trace.cpp).
#pragma once

#include <string>

extern "C" {
#include "../../include/solprop_mi355x.h"
}

namespace sp::march {

using TraceScenario = sp_rk_scenario;
std::string run_trace(const TraceScenario& scenario);

}  // namespace sp::march

namespace sp::adams {
std::string run_adams_trace(const sp_adams_scenario& scenario);
}  // namespace sp::adams
