// The exponential Adams march and its controller, host C++.
//
// The decisions are the reference's, operation for operation: the controller core of
// src/solprop/backends/adams_core.py:96-531 (stencil, factor, order weight, order
// choice, rejection, landing approach), the march of
// src/solprop/backends/numpy_ref/marches/adams.py:79-700 (natively
// native/src/march/adams_march.cpp) and the policy constants of
// src/solprop/stepper/multistep.py:60-98. Compiled without floating-point contraction,
// and held to the reference's Adams decision corpus (tests/data/adams_traces.json.gz).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "march.hpp"

namespace sp::adams {

constexpr double kPecRealInterval[6] = {1.0, 0.545, 0.3, 0.163, 0.0877, 0.0465};

struct Policy {
  int order = 5;
  bool variable_order = true;
  double safety = 0.8, max_growth = 2.0, min_factor = 0.2;
  double rejection_growth = 0.9, hysteresis = 1.2;
  int restart_after = 3;
  double collar_stability = 1.0, collar_rate = 0.0;
  long max_calls = -1;  // none
  bool equal_approach = true, interval_cap = false;
  double dz_min = 1e-12, dz_max = 1.0, stretch = 1.1, z_end = 1.0;
  std::vector<double> save_planes, hard_kinks, breakpoints;
};

struct Stencil {
  std::vector<double> nodes;  // newest first
  std::vector<int> rows;
  int order = 1;
  std::vector<int> orders;
};

struct Entry {
  double z;
  int identity;  // 0 accepted (a startup sample), 1 predicted
  int row;
};

struct Ledger {
  long startup = 0, predictor = 0, corrected = 0, rejected = 0, accepted_steps = 0,
       rejected_steps = 0, restarts = 0, failure_restarts = 0, dense_observations = 0;
  std::vector<int> orders;
  long total() const { return startup + predictor + corrected; }
};

struct BudgetError : march::ConvergenceError {
  using march::ConvergenceError::ConvergenceError;
};
// A replayed step whose norm is not finite: the field diverged.
struct Diverged : march::ConvergenceError {
  using march::ConvergenceError::ConvergenceError;
};

struct Engine {
  virtual ~Engine() = default;
  virtual void start(double z) = 0;
  // One norm per scored order, the step's own first.
  virtual std::vector<double> trial(double z, double dz, const Stencil& s) = 0;
  virtual void accept(double dz, int free, const Stencil& s, const std::vector<Entry>& kept) = 0;
};

struct Host {
  virtual ~Host() = default;
  virtual bool collar_quiet() { return false; }
  virtual double cap(double proposal, double z) { (void)z; return proposal; }
  virtual void record(double end, double dz, double norm, bool accepted) {
    (void)end; (void)dz; (void)norm; (void)accepted;
  }
  // The predictor order of the attempt about to be recorded.
  virtual void attempt_order(int order) { (void)order; }
  virtual void stepped() {}
  virtual bool natural_stop(double target) { (void)target; return false; }
  virtual bool halt() { return false; }
};

Stencil stencil(const Policy& p, double z, double dz, const std::vector<Entry>& history,
                int stored_order, bool last_rejected);
double factor(const Policy& p, double norm, int order, double growth);
double order_weight(const Policy& p, double norm, int order, double dz, bool quiet);

// One action a march asks of its engine: evaluate the source at the start of a history,
// or take a trial step.
struct Request {
  enum Kind { kDone, kHalted, kStart, kTrial } kind = kDone;
  double z = 0.0, dz = 0.0, end = 0.0;
  double target = 0.0;  // the position the march is headed for
  bool kink = false;
  Stencil stencil;
};

class March {
 public:
  March(const Policy& p, Engine& e, Host& h, double z0, double proposal);
  // Step adaptively to `target`; false when the host halted between attempts.
  bool march_to(double target);
  // The same march driven from outside, one action at a time, for engines that run many
  // marches' trials together: `plan` says what the march needs next towards `target`
  // (kDone: reached, or the host stopped it naturally; kHalted: the host's attempt limit;
  // kStart: evaluate the source at z, then call `started`; kTrial: run the request's
  // trial, then hand its norms to `conclude`). march_to is this loop with the engine's
  // start and trial in between, so both make the same decisions.
  Request::Kind plan(double target, Request* r);
  void started();
  // The attempt's decision from its norms (the step's own order first); true if accepted.
  bool conclude(const Request& r, const std::vector<double>& norms);
  // Replay a fixed schedule towards `target`.
  bool replay_to(const std::vector<double>& steps, const std::vector<double>& edges,
                 double target);

  double z, proposal;
  long cursor = 0;
  std::vector<Entry> history;
  int rejections = 0;
  bool last_rejected = false;
  int order;
  int n_slots;
  Ledger ledger;
  bool collar_was_quiet = false;

 private:
  void afford();
  void start();
  bool attempt(double dz, bool fixed, double end, bool kink, double* norm);
  bool decide(const Stencil& s, const std::vector<double>& norms, double dz, bool fixed,
              double end, bool kink, double* norm);
  bool stopped_ = false;  // the host stopped the march naturally after an accepted step
  void reject(double dz, double norm, int order_);
  const Policy& p_;
  Engine& engine_;
  Host& host_;
};

}  // namespace sp::adams
