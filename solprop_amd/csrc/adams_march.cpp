// The exponential Adams march (see adams_march.hpp). Compiled with -ffp-contract=off.
#include "adams_march.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <set>

namespace sp::adams {

namespace {
inline double pymax(double a, double b) { return b > a ? b : a; }
inline double pymin(double a, double b) { return b < a ? b : a; }
constexpr double kInf = std::numeric_limits<double>::infinity();
}  // namespace

Stencil stencil(const Policy& p, double z, double dz, const std::vector<Entry>& history,
                int stored_order, bool last_rejected) {
  Stencil s;
  for (auto it = history.rbegin(); it != history.rend(); ++it) {
    s.nodes.push_back((it->z - z) / dz);
    s.rows.push_back(it->row);
  }
  const int held = static_cast<int>(history.size());
  s.order = std::min(stored_order, held);
  const bool rising = !last_rejected && held > s.order;
  s.orders.push_back(s.order);
  if (p.variable_order) {
    for (int other : {s.order - 1, s.order + 1}) {
      if (1 <= other && other <= p.order && other <= held && (other < s.order || rising))
        s.orders.push_back(other);
    }
  }
  return s;
}

double factor(const Policy& p, double norm, int order, double growth) {
  if (norm <= 0.0) return growth;
  const double raw = p.safety * std::pow(norm, -1.0 / (order + 1));
  return pymin(growth, pymax(p.min_factor, raw));
}

namespace {
double stable_factor(const Policy& p, int order, double dz, bool quiet) {
  if (p.collar_rate <= 0.0 || quiet) return kInf;
  const double interval = p.collar_stability * kPecRealInterval[order - 1];
  return interval / (p.collar_rate * dz);
}
}  // namespace

double order_weight(const Policy& p, double norm, int order, double dz, bool quiet) {
  const double rate = norm <= 0.0 ? kInf : std::pow(norm, -1.0 / (order + 1));
  const double allowed = p.safety * rate;
  return pymin(allowed, stable_factor(p, order, dz, quiet));
}

March::March(const Policy& p, Engine& e, Host& h, double z0, double proposal_)
    : z(z0), proposal(proposal_), p_(p), engine_(e), host_(h) {
  order = p.variable_order ? 1 : p.order;
  n_slots = p.order + 2;
}

void March::afford() {
  const long cost = (history.empty() ? 1 : 0) + 1;
  const long spent = ledger.total();
  if (p_.max_calls >= 0 && spent + cost > p_.max_calls) {
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "the source-call budget of %ld is spent (%ld charged) at z = %.6g m: the "
                  "next trial needs %ld; the run stops at its last accepted step, "
                  "unfinished",
                  p_.max_calls, spent, z, cost);
    throw BudgetError(buf);
  }
}

void March::start() {
  ledger.startup += 1;
  engine_.start(z);
  history.assign(1, Entry{z, 0, 0});
  order = p_.variable_order ? 1 : p_.order;
}

namespace {
double next_landing(const Policy& p, double z, const march::Targets& t, double target,
                    bool* found) {
  const double tolerance = march::kLandingUlps * march::ulp(pymax(std::fabs(z), std::fabs(target)));
  double nearest = target;
  for (const auto* positions : {t.planes, t.kinks}) {
    auto it = std::upper_bound(positions->begin(), positions->end(), z + tolerance);
    if (it != positions->end()) nearest = pymin(nearest, *it);
  }
  *found = nearest > z + tolerance;
  (void)p;
  return nearest;
}
}  // namespace

bool March::attempt(double dz, bool fixed, double end, bool kink, double* out_norm) {
  const Stencil s = stencil(p_, z, dz, history, order, last_rejected);
  ledger.predictor += 1;
  const std::vector<double> norms = engine_.trial(z, dz, s);
  return decide(s, norms, dz, fixed, end, kink, out_norm);
}

bool March::decide(const Stencil& s, const std::vector<double>& norms, double dz, bool fixed,
                   double end, bool kink, double* out_norm) {
  const double norm = norms[0];
  *out_norm = norm;
  if (!(norm <= 1.0) && !(fixed && std::isfinite(norm))) {
    ledger.rejected += 1;
    reject(dz, norm, s.order);
    return false;
  }
  const double z0 = z;
  std::vector<Entry> keep;
  if (p_.order > 0) {
    const std::size_t n = history.size();
    const std::size_t first = n > static_cast<std::size_t>(p_.order) ? n - p_.order : 0;
    keep.assign(history.begin() + static_cast<long>(first), history.end());
  }
  std::set<int> used;
  for (const auto& e : keep) used.insert(e.row);
  int free = 0;
  while (used.count(free)) ++free;
  engine_.accept(dz, free, s, keep);
  history = keep;
  history.push_back(Entry{z0 + dz, 1, free});
  ledger.accepted_steps += 1;
  rejections = 0;
  ledger.orders.push_back(s.order);
  // The order choice (adams_core.choose_order).
  const bool quiet = host_.collar_quiet();
  int stored = s.order, control = s.order;
  double control_norm = norms[0];
  if (!p_.variable_order) {
    stored = p_.order;
  } else {
    std::size_t best = 0;
    std::vector<double> weights(s.orders.size());
    for (std::size_t i = 0; i < s.orders.size(); ++i)
      weights[i] = order_weight(p_, norms[i], s.orders[i], dz, quiet);
    for (std::size_t i = 1; i < s.orders.size(); ++i)
      if (weights[i] > weights[best]) best = i;
    if (best != 0 && weights[best] > p_.hysteresis * weights[0]) {
      stored = control = s.orders[best];
      control_norm = norms[best];
    }
  }
  order = stored;
  last_rejected = false;
  proposal = march::limit_step(dz, factor(p_, control_norm, control, p_.max_growth), p_.dz_min,
                               p_.dz_max);
  z = end;
  if (kink) {
    history.clear();
    ledger.restarts += 1;
  }
  return true;
}

void March::reject(double dz, double norm, int order_) {
  ledger.rejected_steps += 1;
  last_rejected = true;
  const int count = rejections + 1;
  const bool restart = count >= p_.restart_after && history.size() > 1;
  const double scale = !std::isfinite(norm) ? p_.min_factor
                                            : factor(p_, norm, order_, p_.rejection_growth);
  rejections = restart ? 0 : count;
  if (restart) {
    history.clear();
    ledger.failure_restarts += 1;
  }
  proposal = march::limit_step(dz, scale, p_.dz_min, p_.dz_max);
}

bool March::march_to(double target) {
  stopped_ = false;
  for (;;) {
    Request r;
    switch (plan(target, &r)) {
      case Request::kDone: return true;
      case Request::kHalted: return false;
      case Request::kStart:
        engine_.start(z);
        started();
        break;
      case Request::kTrial:
        conclude(r, engine_.trial(r.z, r.dz, r.stencil));
        break;
    }
  }
}

void March::started() {
  history.assign(1, Entry{z, 0, 0});
  order = p_.variable_order ? 1 : p_.order;
}

bool March::conclude(const Request& r, const std::vector<double>& norms) {
  double norm = 0.0;
  const bool accepted = decide(r.stencil, norms, r.dz, false, r.end, r.kink, &norm);
  host_.attempt_order(r.stencil.order);
  host_.record(r.end, r.dz, norm, accepted);
  if (accepted) {
    host_.stepped();
    if (host_.natural_stop(r.target)) stopped_ = true;
  }
  return accepted;
}

Request::Kind March::plan(double target, Request* out) {
  const std::vector<double> planes = p_.save_planes, kinks = p_.hard_kinks;
  std::vector<double> edges = p_.breakpoints;
  if (!p_.interval_cap && edges.size() > 2) edges = {edges.front(), edges.back()};
  const march::Targets targets{&planes, &kinks, &edges, p_.dz_max, target, p_.stretch};
  Request& r = *out;
  r = Request{};
  if (stopped_) {
    stopped_ = false;
    return r.kind = Request::kDone;
  }
  {
    if (march::reached(z, target)) return r.kind = Request::kDone;
    if (host_.halt()) return r.kind = Request::kHalted;
    afford();
    if (history.empty()) {
      // The start's evaluation (the reference's start): the caller runs it, then
      // `started` opens the history; the next plan goes on to the trial.
      ledger.startup += 1;
      r.z = z;
      return r.kind = Request::kStart;
    }
    // The collar's re-entry cap (adams_core.collar_reentry).
    const bool quiet = host_.collar_quiet();
    double prop = host_.cap(proposal, z);
    if (collar_was_quiet && !quiet && p_.collar_rate > 0.0) {
      const double interval = p_.collar_stability * kPecRealInterval[order - 1];
      prop = pymin(prop, interval / p_.collar_rate);
    }
    collar_was_quiet = quiet;
    const march::Landing stop = march::landing(z, prop, targets);
    // The approach to a forced landing (adams_core.approach).
    double dz = stop.dz, end;
    bool kink = false;
    bool halved = false;
    if (!stop.landed && p_.equal_approach) {
      bool found = false;
      const double ahead = next_landing(p_, z, targets, target, &found);
      if (found && stop.dz < ahead - z && ahead - z < 2.0 * stop.dz) {
        dz = 0.5 * (ahead - z);
        end = z + dz;
        halved = true;
      }
    }
    if (!halved) {
      end = stop.landed ? stop.z_target : z + stop.dz;
      const double tolerance = march::kLandingUlps * march::ulp(pymax(std::fabs(end), 1.0));
      const bool interior = end < p_.z_end - tolerance;
      bool on_kink = false;
      if (interior)
        for (double k : kinks)
          if (std::fabs(k - end) <= tolerance) on_kink = true;
      kink = stop.target == march::kHardKink || on_kink;
    }
    r.stencil = stencil(p_, z, dz, history, order, last_rejected);
    ledger.predictor += 1;
    r.z = z;
    r.dz = dz;
    r.end = end;
    r.kink = kink;
    r.target = target;
    return r.kind = Request::kTrial;
  }
}

bool March::replay_to(const std::vector<double>& steps, const std::vector<double>& edges,
                      double target) {
  const long n = static_cast<long>(steps.size());
  const double slack = static_cast<double>(std::max<long>(march::kLandingUlps, n)) *
                       march::ulp(pymax(edges.back(), std::fabs(target)));
  while (cursor < n && z < target - slack) {
    if (host_.halt()) return false;
    const double dz = steps[static_cast<std::size_t>(cursor)];
    const double end = edges[static_cast<std::size_t>(cursor + 1)];
    afford();
    if (history.empty()) start();
    bool kink = false;
    for (double k : p_.hard_kinks)
      if (std::fabs(k - end) <= slack) kink = true;
    double norm = 0.0;
    attempt(dz, true, end, kink, &norm);
    if (!std::isfinite(norm)) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "a replayed exponential Adams step of %.4g m ending at z = %.6g m", dz,
                    end);
      throw Diverged(buf);
    }
    cursor += 1;
    host_.attempt_order(order);
    host_.record(end, dz, norm, true);
    host_.stepped();
    if (host_.natural_stop(target)) return true;
  }
  if (z < target - slack) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "the replayed schedule ends at %.6g m and cannot reach z = %.6g m",
                  edges.back(), target);
    throw march::ReplayError(buf);
  }
  return true;
}

}  // namespace sp::adams
