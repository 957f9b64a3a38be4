// Scripted decision traces of the exponential Adams march: the march of
// adams_march.cpp driven by an engine that returns prescribed norms, with the
// session's quiet-collar cap, written as JSON in the layout of the reference's
// tests/data/adams_traces.json.gz (tests/backends/adams_traces.py:131-633).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <string>

#include "adams_march.hpp"
#include "trace.hpp"

namespace sp::adams {
namespace {

struct Exhausted {};

void put(std::string& out, double v) {
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

void put_list(std::string& out, const std::vector<double>& v) {
  out += "[";
  for (std::size_t i = 0; i < v.size(); ++i) {
    if (i) out += ", ";
    put(out, v[i]);
  }
  out += "]";
}

template <typename T>
void put_ints(std::string& out, const std::vector<T>& v) {
  out += "[";
  for (std::size_t i = 0; i < v.size(); ++i) {
    if (i) out += ", ";
    out += std::to_string(v[i]);
  }
  out += "]";
}

struct Scripted final : Engine, Host {
  const sp_adams_scenario& sc;
  March* march = nullptr;
  bool quiet = false;
  long accepted = 0, attempts = 0;
  std::map<long, int> quiet_map;
  std::string events;
  bool first = true;
  bool pending = false;
  std::string pending_body;  // "z": ..., "dz": ..., ... without braces

  explicit Scripted(const sp_adams_scenario& s) : sc(s) {
    for (int i = 0; i < s.n_quiet; ++i) quiet_map[s.quiet_keys[i]] = s.quiet_values[i];
    quiet = quiet_map.count(0) ? quiet_map[0] != 0 : false;
  }

  void emit(const std::string& e) {
    if (!first) events += ", ";
    first = false;
    events += e;
  }

  std::string state() const {
    std::string o = "\"position\": ";
    put(o, march->z);
    o += ", \"proposal\": ";
    put(o, march->proposal);
    o += ", \"stored_order\": " + std::to_string(march->order);
    o += ", \"rejections\": " + std::to_string(march->rejections);
    o += std::string(", \"last_rejected\": ") + (march->last_rejected ? "true" : "false");
    o += std::string(", \"collar_was_quiet\": ") + (march->collar_was_quiet ? "true" : "false");
    o += std::string(", \"quiet\": ") + (quiet ? "true" : "false");
    o += ", \"history\": [";
    for (std::size_t i = 0; i < march->history.size(); ++i) {
      const auto& e = march->history[i];
      if (i) o += ", ";
      o += "[";
      put(o, e.z);
      o += std::string(", \"") + (e.identity == 0 ? "accepted" : "predicted") + "\", " +
           std::to_string(e.row) + "]";
    }
    const Ledger& l = march->ledger;
    o += "], \"ledger\": [" + std::to_string(l.startup) + ", " + std::to_string(l.predictor) +
         ", " + std::to_string(l.rejected) + ", " + std::to_string(l.accepted_steps) + ", " +
         std::to_string(l.rejected_steps) + ", " + std::to_string(l.restarts) + ", " +
         std::to_string(l.failure_restarts) + ", " + std::to_string(l.dense_observations) + "]";
    return o;
  }

  // Engine.
  void start(double z) override {
    std::string e = "{\"start\": ";
    put(e, z);
    e += ", \"stored\": true}";
    emit(e);
  }
  std::vector<double> trial(double z, double dz, const Stencil& s) override {
    if (attempts >= sc.n_attempts) throw Exhausted{};
    std::string b = "\"z\": ";
    put(b, z);
    b += ", \"dz\": ";
    put(b, dz);
    b += ", \"nodes\": ";
    put_list(b, s.nodes);
    b += ", \"slots\": ";
    put_ints(b, s.rows);
    b += ", \"order\": " + std::to_string(s.order) + ", \"orders\": ";
    put_ints(b, s.orders);
    pending_body = b;
    pending = true;
    const double* row = sc.norms + 3 * attempts;
    attempts += 1;
    return std::vector<double>(row, row + s.orders.size());
  }
  void accept(double, int free, const Stencil&, const std::vector<Entry>& kept) override {
    std::vector<int> rows;
    for (const auto& e : kept) rows.push_back(e.row);
    pending_body += ", \"kept\": ";
    put_ints(pending_body, rows);
    pending_body += ", \"free\": " + std::to_string(free);
  }
  // Host.
  bool collar_quiet() override { return quiet; }
  double cap(double proposal, double) override {
    if (!sc.has_quiet_interval) return proposal;
    if (quiet) proposal = sc.quiet_interval < proposal ? sc.quiet_interval : proposal;
    return proposal;
  }
  void record(double end, double dz, double norm, bool acc) override {
    std::string e = "{" + pending_body + ", \"end\": ";
    put(e, end);
    e += ", \"step\": ";
    put(e, dz);
    e += ", \"norm\": ";
    put(e, std::isfinite(norm) ? norm : std::numeric_limits<double>::infinity());
    e += std::string(", \"accepted\": ") + (acc ? "true" : "false") + ", " + state() + "}";
    emit(e);
    pending = false;
    pending_body.clear();
  }
  void stepped() override {
    accepted += 1;
    if (quiet_map.count(accepted)) quiet = quiet_map[accepted] != 0;
  }
};

}  // namespace

std::string run_adams_trace(const sp_adams_scenario& sc) {
  Policy p;
  p.order = sc.order;
  p.variable_order = sc.variable_order != 0;
  p.safety = sc.safety;
  p.max_growth = sc.max_growth;
  p.min_factor = sc.min_factor;
  p.collar_rate = sc.collar_rate;
  p.max_calls = sc.max_calls;
  p.equal_approach = sc.equal_approach != 0;
  p.interval_cap = sc.interval_cap != 0;
  p.dz_min = sc.dz_min;
  p.dz_max = sc.dz_max;
  p.stretch = sc.stretch;
  p.z_end = sc.z_end;
  p.save_planes.assign(sc.save_planes, sc.save_planes + sc.n_planes);
  p.hard_kinks.assign(sc.kinks, sc.kinks + sc.n_kinks);
  p.breakpoints.assign(sc.breakpoints, sc.breakpoints + sc.n_breakpoints);
  if (p.breakpoints.empty()) p.breakpoints = {sc.z0, sc.z_end};
  if (!p.interval_cap) p.breakpoints = {p.breakpoints.front(), p.breakpoints.back()};

  Scripted h(sc);
  March m(p, h, h, sc.z0, sc.proposal);
  h.march = &m;
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
  std::string end = "\"stop\": \"done\"";
  try {
    for (double target : targets) {
      if (sc.n_schedule > 0)
        m.replay_to(steps, edges, target);
      else
        m.march_to(target);
      std::string e = "{\"target\": ";
      put(e, target);
      e += ", \"position\": ";
      put(e, m.z);
      e += "}";
      h.emit(e);
    }
  } catch (const Exhausted&) {
    end = "\"stop\": \"exhausted\"";
  } catch (const BudgetError& e) {
    end = std::string("\"stop\": \"ResponseBudgetError\", \"message\": \"") + e.what() + "\"";
  } catch (const Diverged& e) {
    end = std::string("\"stop\": \"ConvergenceError\", \"message\": \"scripted: ") + e.what() +
          "\"";
  } catch (const march::ConvergenceError& e) {
    end = std::string("\"stop\": \"ConvergenceError\", \"message\": \"") + e.what() + "\"";
  } catch (const march::ReplayError& e) {
    end = std::string("\"stop\": \"BackendError\", \"message\": \"") + e.what() + "\"";
  }
  std::string out = "{\"events\": [" + h.events + "], \"end\": {" + end + ", \"in_flight\": {" +
                    (h.pending ? h.pending_body : std::string()) + "}, " + h.state() +
                    ", \"orders\": ";
  put_ints(out, m.ledger.orders);
  out += "}}";
  return out;
}

}  // namespace sp::adams
