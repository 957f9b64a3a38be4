// The Runge-Kutta march and its step-size controller, host C++.
//
// The decisions are the reference's, operation for operation: the PI controller of
// src/solprop/stepper/controller.py:262-676, the controller core of
// src/solprop/backends/rk_core.py:37-166, the landing rule of
// src/solprop/stepper/schedule.py:959-1090 and the march loop of
// src/solprop/backends/numpy_ref/marches/runge_kutta.py:133-260 (natively
// native/src/march/rk_march.cpp:44-171). This file is compiled without floating-point
// contraction, so every factor rounds as the reference's does; the decisions are held
// to the reference's decision corpus (tests/data/rk_traces.json.gz) exactly.
#pragma once

#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

namespace sp::march {

constexpr double kStartGrowth = 100.0;
constexpr int kLandingUlps = 8;

struct ConvergenceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ReplayError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Controller {
  double rtol = 1e-6, atol = 0.0, safety = 0.8, max_factor = 1.1, min_factor = 0.2;
  double beta1 = 3.0 / 25.0, beta2 = -1.0 / 25.0;
  int lag = 1;
  bool reject_free = false;
  int estimator_order = 4;

  double propose(const std::deque<double>& norms) const;
  double start_factor(double norm) const;
  bool accepts(double norm) const;
  double rejection_factor(double norm) const;
  double equivalent_norm(double norm, double dz_taken, double dz_proposed) const;
};

double limit_step(double dz, double factor, double dz_min, double dz_max);

struct Policy {
  Controller controller;
  double dz_min = 0.0, dz_max = 1.0;
  bool auto_init = true;
  double stretch = 1.1, z_end = 1.0;
  std::vector<double> save_planes, hard_kinks, breakpoints;
};

enum LandingTarget : int { kNone = 0, kMaxStep, kIntervalWidth, kHardKink, kSavePlane, kEnd };

struct Landing {
  double dz;
  LandingTarget target;
  bool landed;
  double z_target;
  bool stretched;
};

struct Targets {
  const std::vector<double>* planes;
  const std::vector<double>* kinks;
  const std::vector<double>* edges;
  double dz_max, z_end, stretch;
};

Landing landing(double z, double dz, const Targets& targets);
bool reached(double position, double target);
double ulp(double x);

struct State {
  double z = 0.0, proposal = 0.0;
  long cursor = 0;
  std::deque<double> norms;
  std::size_t max_norms = 3;
  void push(double norm) {
    norms.push_back(norm);
    while (norms.size() > max_norms) norms.pop_front();
  }
};

// What a march asks of its engine and its session.
struct Engine {
  virtual ~Engine() = default;
  virtual double trial(double z, double dz) = 0;
  virtual void accept(double dz) = 0;
};

struct Host {
  virtual ~Host() = default;
  virtual double cap(double proposal, double z) { (void)z; return proposal; }
  virtual void record(double end, double dz, double norm, bool accepted) {
    (void)end; (void)dz; (void)norm; (void)accepted;
  }
  virtual void stepped() {}
  virtual bool natural_stop(double target) { (void)target; return false; }
  virtual bool halt() { return false; }  // the call's bound: stop between attempts
};

std::pair<double, bool> decide(const Policy& p, double norm, double dz, double proposal);
std::pair<double, double> start_up(const Policy& p, double proposal, double norm);
double propose(const Policy& p, double proposal, const std::deque<double>& norms);
double reject(const Policy& p, double proposal, double equivalent);

// Step adaptively to `target`, landing on it exactly. Returns false when the host
// halted the march between attempts (the target not yet reached).
bool march_to(const Policy& p, Engine& engine, Host& host, State& st, double target);
// Replay a fixed schedule (edges precomputed as the running sum of its steps).
bool replay_to(Engine& engine, Host& host, State& st, const std::vector<double>& steps,
               const std::vector<double>& edges, double target);

}  // namespace sp::march
