// Scripted decision traces of the Runge-Kutta march: the march of march.cpp driven by
// an engine that returns prescribed norms and a host that applies the session's step
// caps (src/solprop/backends/numpy_ref/session.py:289-361), as the reference's trace
// harness does (tests/backends/rk_traces.py:145-453). The trace is written as JSON in
// the corpus's layout so that it compares with tests/data/rk_traces.json.gz exactly.
#include "trace.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <string>

#include "march.hpp"

namespace sp::march {
namespace {

struct Exhausted {};

void put_double(std::string& out, double v) {
  if (std::isinf(v)) {
    out += v > 0 ? "Infinity" : "-Infinity";
    return;
  }
  if (std::isnan(v)) {
    out += "NaN";
    return;
  }
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  out += buf;
}

struct Scripted final : Engine, Host {
  const TraceScenario& sc;
  State& st;
  bool quiet = false;
  double step_cap = std::numeric_limits<double>::infinity();
  long accepted = 0, attempts = 0, natural = 0;
  bool has_monitor;
  int order;
  std::map<long, int> quiet_map;
  std::map<long, double> split_map;
  bool pending = false;
  double pend_z = 0, pend_dz = 0;
  std::string events;
  bool first_event = true;
  bool last_open = false;  // the last event is an attempt still missing its "after"

  Scripted(const TraceScenario& s, State& state) : sc(s), st(state) {
    for (int i = 0; i < s.n_quiet; ++i) quiet_map[s.quiet_keys[i]] = s.quiet_values[i];
    for (int i = 0; i < s.n_split; ++i) split_map[s.split_keys[i]] = s.split_values[i];
    quiet = quiet_map.count(0) ? quiet_map[0] != 0 : false;
    if (split_map.count(0)) step_cap = split_map[0];
    has_monitor = s.has_quiet_interval || s.has_ceiling || s.n_split > 0;
    order = s.estimator_order + 1;
  }

  std::string snapshot() const {
    std::string o = "{\"position\": ";
    put_double(o, st.z);
    o += ", \"proposal\": ";
    put_double(o, st.proposal);
    o += ", \"norms\": [";
    bool first = true;
    for (double v : st.norms) {
      if (!first) o += ", ";
      first = false;
      put_double(o, std::isfinite(v) ? v : std::numeric_limits<double>::infinity());
    }
    o += "], \"accepted\": " + std::to_string(accepted) +
         ", \"natural\": " + std::to_string(natural) +
         ", \"quiet\": " + std::string(quiet ? "true" : "false") +
         ", \"cursor\": " + std::to_string(st.cursor) + "}";
    return o;
  }

  void close() {
    if (last_open) {
      events += ", \"after\": " + snapshot() + "}";
      last_open = false;
    }
  }

  void begin_event() {
    if (!first_event) events += ", ";
    first_event = false;
  }

  double trial(double z, double dz) override {
    close();
    if (attempts >= sc.n_norms) throw Exhausted{};
    pending = true;
    pend_z = z;
    pend_dz = dz;
    const double noise = sc.norms[attempts];
    attempts += 1;
    if (!sc.has_h0) return noise;
    return noise * std::pow(dz / sc.h0, order);
  }

  void accept(double) override {}

  double cap(double proposal, double) override {
    if (!has_monitor) return proposal;
    if (quiet) {
      const double interval =
          sc.has_quiet_interval ? sc.quiet_interval : std::numeric_limits<double>::infinity();
      proposal = interval < proposal ? interval : proposal;
    }
    if (sc.has_ceiling) proposal = sc.ceiling < proposal ? sc.ceiling : proposal;
    return step_cap < proposal ? step_cap : proposal;
  }

  void record(double end, double dz, double norm, bool acc) override {
    begin_event();
    events += "{";
    if (pending) {
      events += "\"z\": ";
      put_double(events, pend_z);
      events += ", \"dz\": ";
      put_double(events, pend_dz);
      events += ", ";
    }
    events += "\"end\": ";
    put_double(events, end);
    events += ", \"step\": ";
    put_double(events, dz);
    events += ", \"norm\": ";
    put_double(events, std::isfinite(norm) ? norm : std::numeric_limits<double>::infinity());
    events += std::string(", \"accepted\": ") + (acc ? "true" : "false");
    last_open = true;
    pending = false;
  }

  void stepped() override {
    accepted += 1;
    if (quiet_map.count(accepted)) quiet = quiet_map[accepted] != 0;
    if (has_monitor && split_map.count(accepted)) step_cap = split_map[accepted];
  }

  bool natural_stop(double target) override {
    natural += 1;
    if (sc.step_planes == 0 || natural % sc.step_planes) return false;
    return !reached(st.z, target);
  }

  void target_event(double target) {
    close();
    begin_event();
    std::string snap = snapshot();
    std::string o = "{\"target\": ";
    put_double(o, target);
    o += ", " + snap.substr(1);
    events += o;
  }
};

}  // namespace

std::string run_trace(const TraceScenario& sc) {
  Policy p;
  p.controller.rtol = sc.rtol;
  p.controller.atol = sc.atol;
  p.controller.safety = sc.safety;
  p.controller.max_factor = sc.max_factor;
  p.controller.min_factor = sc.min_factor;
  p.controller.beta1 = sc.beta1;
  p.controller.beta2 = sc.beta2;
  p.controller.lag = sc.lag;
  p.controller.reject_free = sc.reject_free != 0;
  p.controller.estimator_order = sc.estimator_order;
  p.dz_min = sc.dz_min;
  p.dz_max = sc.dz_max;
  p.auto_init = sc.auto_init != 0;
  p.stretch = sc.stretch;
  p.z_end = sc.z_end;
  p.save_planes.assign(sc.save_planes, sc.save_planes + sc.n_planes);
  p.hard_kinks.assign(sc.kinks, sc.kinks + sc.n_kinks);
  std::vector<double> bp(sc.breakpoints, sc.breakpoints + sc.n_breakpoints);
  if (bp.empty()) bp = {sc.z0, sc.z_end};
  if (!sc.interval_cap) bp = {bp.front(), bp.back()};
  p.breakpoints = bp;

  State st;
  st.z = sc.z0;
  st.proposal = sc.proposal;
  st.max_norms = static_cast<std::size_t>(sc.lag + 2);
  Scripted h(sc, st);

  std::vector<double> targets;
  if (sc.n_targets > 0) {
    targets.assign(sc.targets, sc.targets + sc.n_targets);
  } else if (sc.n_planes == 0) {
    targets = {sc.z_end};
  } else {
    targets.assign(sc.save_planes, sc.save_planes + sc.n_planes);
    targets.push_back(sc.z_end);
    std::sort(targets.begin(), targets.end());
    targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
  }
  std::vector<double> steps(sc.schedule, sc.schedule + sc.n_schedule), edges;
  if (sc.n_schedule > 0) {
    edges.push_back(0.0);
    double total = 0.0;
    for (double s : steps) {
      total = total + s;
      edges.push_back(total);
    }
  }
  std::string end = "{\"stop\": \"done\"";
  try {
    for (double target : targets) {
      while (true) {
        if (sc.n_schedule == 0)
          march_to(p, h, h, st, target);
        else
          replay_to(h, h, st, steps, edges, target);
        h.target_event(target);
        if (reached(st.z, target)) break;
      }
    }
  } catch (const Exhausted&) {
    end = "{\"stop\": \"exhausted\"";
  } catch (const ConvergenceError& e) {
    end = std::string("{\"stop\": \"ConvergenceError\", \"message\": \"") + e.what() + "\"";
  } catch (const ReplayError& e) {
    end = std::string("{\"stop\": \"BackendError\", \"message\": \"") + e.what() + "\"";
  }
  h.close();
  end += ", \"in_flight\": {";
  if (h.pending) {
    end += "\"z\": ";
    put_double(end, h.pend_z);
    end += ", \"dz\": ";
    put_double(end, h.pend_dz);
  }
  end += "}, " + h.snapshot().substr(1);
  return "{\"events\": [" + h.events + "], \"end\": " + end + "}";
}

}  // namespace sp::march
