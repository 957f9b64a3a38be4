// The Runge-Kutta march and controller (see march.hpp). Compiled with
// -ffp-contract=off: no fused multiply-add may change a factor's rounding.
#include "march.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>

namespace sp::march {

namespace {
// Python's max(a, b) / min(a, b): the first argument unless the second compares
// strictly greater (smaller), so a NaN first argument is kept.
inline double pymax(double a, double b) { return b > a ? b : a; }
inline double pymin(double a, double b) { return b < a ? b : a; }
}  // namespace

double ulp(double x) {
  x = std::fabs(x);
  if (!std::isfinite(x)) return std::numeric_limits<double>::quiet_NaN();
  return std::nextafter(x, std::numeric_limits<double>::infinity()) - x;
}

double Controller::propose(const std::deque<double>& history) const {
  const long index = static_cast<long>(history.size()) - 1 - lag;
  if (index < 0) return 1.0;
  const double current = history[static_cast<std::size_t>(index)];
  if (!std::isfinite(current)) return min_factor;
  if (current <= 0.0) return max_factor;
  double previous = index >= 1 ? history[static_cast<std::size_t>(index - 1)] : 1.0;
  if (!std::isfinite(previous) || previous <= 0.0) previous = 1.0;
  const double factor =
      safety * std::pow(1.0 / current, beta1) * std::pow(1.0 / previous, beta2);
  return pymin(pymax(factor, min_factor), max_factor);
}

double Controller::start_factor(double norm) const {
  if (!std::isfinite(norm) || norm <= 0.0) return norm == 0.0 ? kStartGrowth : min_factor;
  const double factor = safety * std::pow(1.0 / norm, 1.0 / (estimator_order + 1));
  return pymin(pymax(factor, min_factor), kStartGrowth);
}

bool Controller::accepts(double norm) const {
  if (!std::isfinite(norm)) return false;
  return reject_free || norm <= 1.0;
}

double Controller::rejection_factor(double norm) const {
  if (!std::isfinite(norm)) return min_factor;
  if (norm <= 0.0) return 1.0;
  const double factor = safety * std::pow(1.0 / norm, 1.0 / (estimator_order + 1));
  return pymin(pymax(factor, min_factor), 1.0);
}

double Controller::equivalent_norm(double norm, double dz_taken, double dz_proposed) const {
  if (!std::isfinite(norm)) return norm;
  return norm * std::pow(dz_proposed / dz_taken, estimator_order + 1);
}

double limit_step(double dz, double factor, double dz_min, double dz_max) {
  const double proposed = pymin(dz * factor, dz_max);
  if (proposed < dz_min) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "step control asked for %.3e m, below the floor dz_min = %.3e m; the "
                  "tolerance cannot be met here",
                  proposed, dz_min);
    throw ConvergenceError(buf);
  }
  return proposed;
}

std::pair<double, bool> decide(const Policy& p, double norm, double dz, double proposal) {
  return {p.controller.equivalent_norm(norm, dz, proposal), p.controller.accepts(norm)};
}

std::pair<double, double> start_up(const Policy& p, double proposal, double norm) {
  const double factor = p.controller.start_factor(norm);
  const double following = limit_step(proposal, factor, p.dz_min, p.dz_max);
  const double recorded =
      factor >= kStartGrowth ? 0.0 : p.controller.equivalent_norm(norm, proposal, following);
  return {following, recorded};
}

double propose(const Policy& p, double proposal, const std::deque<double>& norms) {
  return limit_step(proposal, p.controller.propose(norms), p.dz_min, p.dz_max);
}

double reject(const Policy& p, double proposal, double equivalent) {
  return limit_step(proposal, p.controller.rejection_factor(equivalent), p.dz_min, p.dz_max);
}

bool reached(double position, double target) {
  const double tolerance = kLandingUlps * ulp(pymax(std::fabs(position), std::fabs(target)));
  return position >= target - tolerance;
}

Landing landing(double z, double dz, const Targets& t) {
  const auto& edges = *t.edges;
  const double tolerance = kLandingUlps * ulp(pymax(std::fabs(z), std::fabs(t.z_end)));
  long index = static_cast<long>(std::upper_bound(edges.begin(), edges.end(), z) - edges.begin()) - 1;
  index = std::min(std::max(index, 0L), static_cast<long>(edges.size()) - 2);
  const double width = edges[index + 1] - edges[index];
  // Caps: the shortest wins; among equals the higher priority.
  struct Bound { double dz; int kind; double at; };
  auto better = [](const Bound& a, const Bound& b) {
    return a.dz < b.dz || (a.dz == b.dz && a.kind > b.kind);
  };
  Bound cap{dz, kNone, 0.0};
  for (Bound c : {Bound{t.dz_max, kMaxStep, 0.0}, Bound{width, kIntervalWidth, 0.0}})
    if (better(c, cap)) cap = c;
  bool have = false;
  Bound land{0.0, kNone, 0.0};
  auto consider = [&](const std::vector<double>& positions, int kind) {
    auto it = std::lower_bound(positions.begin(), positions.end(), z + tolerance);
    if (it == positions.end()) return;
    Bound b{*it - z, kind, *it};
    if (!have || better(b, land)) {
      land = b;
      have = true;
    }
  };
  consider(*t.kinks, kHardKink);
  consider(*t.planes, kSavePlane);
  const std::vector<double> end{t.z_end};
  consider(end, kEnd);
  if (land.dz <= cap.dz) return {land.dz, static_cast<LandingTarget>(land.kind), true, land.at, false};
  if (land.dz <= pymin(t.stretch * cap.dz, t.dz_max))
    return {land.dz, static_cast<LandingTarget>(land.kind), true, land.at, true};
  return {cap.dz, static_cast<LandingTarget>(cap.kind), false, 0.0, false};
}

bool march_to(const Policy& p, Engine& engine, Host& host, State& st, double target) {
  const std::vector<double> planes = p.save_planes, kinks = p.hard_kinks, edges = p.breakpoints;
  const Targets targets{&planes, &kinks, &edges, p.dz_max, target, p.stretch};
  while (!reached(st.z, target)) {
    if (host.halt()) return false;
    const Landing stop = landing(st.z, host.cap(st.proposal, st.z), targets);
    const double norm = engine.trial(st.z, stop.dz);
    const auto [equivalent, accepted] = decide(p, norm, stop.dz, st.proposal);
    const double end = stop.landed ? stop.z_target : st.z + stop.dz;
    host.record(end, stop.dz, norm, accepted);
    if (accepted) {
      engine.accept(stop.dz);
      st.z = end;
      st.push(equivalent);
      if (p.auto_init && st.norms.size() == 1) {
        const auto [following, recorded] = start_up(p, st.proposal, equivalent);
        st.proposal = following;
        st.norms.back() = recorded;
      }
      st.proposal = propose(p, st.proposal, st.norms);
      host.stepped();
      if (host.natural_stop(target)) return true;
    } else {
      st.proposal = reject(p, st.proposal, equivalent);
    }
  }
  return true;
}

bool replay_to(Engine& engine, Host& host, State& st, const std::vector<double>& steps,
               const std::vector<double>& edges, double target) {
  const long n_steps = static_cast<long>(steps.size());
  const double total = edges.back();
  const double ulps = static_cast<double>(std::max<long>(kLandingUlps, n_steps));
  const double slack = ulps * ulp(pymax(total, std::fabs(target)));
  while (st.cursor < n_steps && st.z < target - slack) {
    if (host.halt()) return false;
    const double dz = steps[static_cast<std::size_t>(st.cursor)];
    const double norm = engine.trial(st.z, dz);
    if (!std::isfinite(norm)) {
      char buf[256];
      std::snprintf(buf, sizeof buf, "replayed step %ld of %ld, %.4g m from z = %.6g m,",
                    st.cursor + 1, n_steps, dz, st.z);
      throw ConvergenceError(buf);
    }
    const double end = edges[static_cast<std::size_t>(st.cursor + 1)];
    engine.accept(dz);
    st.cursor += 1;
    st.z = end;
    host.record(st.z, dz, norm, true);
    st.push(norm);
    host.stepped();
    if (host.natural_stop(target)) return true;
  }
  if (st.z < target - slack) throw ReplayError("the schedule ends before the target");
  return true;
}

}  // namespace sp::march
