// One trajectory's march as a session of events, over a workspace of either route.
//
// The reference's native session (native/include/solprop_native/session/session.hpp,
// native/src/session/session.cpp:695; the protocol of docs/dev/kernel-contract.md
// "What a native call takes / returns") advances until the caller has work: the launch
// plane, a save plane, every k-th accepted step, the call's bound (attempts or wall
// time, or a full chunk buffer), the end of the run, a stop or a failure. Each advance
// returns one batch: a sequence number, one event and the chunk of attempt rows made
// since the previous batch (end, step, norm, accepted, order, wall), delivered once.
// An event that observes the state holds a ticket: the committed state stays as it is,
// and can be exported, until the caller acknowledges the batch; advancing across a
// pending batch is a lifecycle error.
//
// Here the session drives the workspace's bounded march calls (Propagator::
// advance_bounded) and takes the plane statistics on the device at each observation.
// Thread safety: advance, acknowledge, export and close serialise on the session's
// mutex; status and request_stop take no lock.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "propagator.hpp"

namespace sp::session {

enum class State : int { ready = 0, pending = 1, finished = 2, closed = 3 };
enum class EventKind : int { observation = 0, progress = 1, snapshot = 2, terminal = 3,
                             warning = 4 };
enum class Observation : int { launch = 0, plane = 1, step = 2, natural = 3 };
enum class Outcome : int { none = 0, completed = 1, stopped = 2, failed = 3 };
// Why a march call returned (the reference's march::Reason numbering).
enum class Reason : int { reached = 0, host_step = 1, bound = 3, floor = 5, error = 6,
                          stopped = 11 };
enum class Stage : int { none = 0, startup = 1, trial = 2, accept = 3 };
enum class Category : int { none = 0, convergence = 1, backend = 2, value = 3 };

struct LifecycleError : std::logic_error {
  using std::logic_error::logic_error;
};

struct Plan {
  std::vector<double> planes;  // save planes, increasing, the last the run's end
  bool launch = true;          // observe the launch state first
  long step_planes = 0;        // observe every k-th accepted step; 0 for none
  std::uint64_t statistics = 0;  // plane statistics taken at each observation
  std::vector<double> windows;   // (lo, hi) pairs for the windowed statistics
  long attempt_capacity = 4096;  // the chunk buffer's attempts
  long call_attempts = 256;      // one march call's attempts at most
  double call_seconds = 0.1;     // and its wall time
};

struct Event {
  EventKind kind = EventKind::progress;
  Observation observation = Observation::launch;
  double z = 0.0;
  long plane_index = -1;
  bool has_stats = false;
  bool ticket = false;
  std::uint64_t generation = 0;
  Outcome outcome = Outcome::none;
  Reason reason = Reason::reached;
  Category category = Category::none;
  Stage stage = Stage::none;
  std::string message;
  double floor_proposal = 0.0;
};

struct Batch {
  std::uint64_t sequence = 0;
  Event event;
  std::vector<double> attempts;  // kWideColumns a row
  std::vector<double> stats;
};

struct Status {
  double z = 0.0;
  long plane_index = 0;
  std::uint64_t attempts = 0, rejected = 0;
  State state = State::ready;
  Outcome outcome = Outcome::none;
  std::uint64_t generation = 0, march_calls = 0, batches = 0, exports = 0, export_bytes = 0;
};

class Session {
 public:
  // Loads `field` at z into the workspace (which the session then leases: nothing else
  // may march it while the session is open).
  Session(Propagator& ws, int n_modes, int n_omega, const Plan& plan, double z,
          const double* field);
  Batch advance(long work_limit, double deadline);
  void acknowledge(std::uint64_t sequence);
  void request_stop() noexcept { stop_.store(true, std::memory_order_release); }
  Status status() const noexcept;
  // The pending batch's state at its ticket's generation, by the export recipe.
  void export_field(std::uint64_t sequence, std::uint64_t generation, const double* weight,
                    bool optics, double* out);
  // A failed terminal batch at the committed state when none is pending (a failure
  // outside the session), carrying whatever attempts the chunk still holds; false when
  // a batch is pending or the session has ended.
  bool fail_outside(const std::string& message, Batch* out);
  void close() noexcept;
  bool closed() const noexcept { return state_.load() == State::closed; }
  std::size_t elements() const { return elements_; }

 private:
  using Clock = std::chrono::steady_clock;
  void require_open() const;
  void observe(Batch& b, Observation kind, long plane);
  void terminal(Batch& b, Outcome outcome);
  Batch issue(Batch b);
  void publish() noexcept;

  Propagator& ws_;
  Plan plan_;
  std::size_t elements_ = 0;
  std::mutex mu_;
  std::vector<double> chunk_;       // attempts since the last batch
  std::size_t next_plane_ = 0;
  bool launch_pending_ = true;
  long since_observed_ = 0;         // accepted steps since the last step observation
  double z_ = 0.0;
  Clock::time_point last_event_;
  std::uint64_t sequence_ = 0, pending_ = 0;
  // Published for status() (no lock).
  std::atomic<State> state_{State::ready};
  std::atomic<Outcome> outcome_{Outcome::none};
  std::atomic<bool> stop_{false};
  std::atomic<double> z_pub_{0.0};
  std::atomic<long> plane_pub_{0};
  std::atomic<std::uint64_t> attempts_{0}, rejected_{0}, generation_{0}, calls_{0},
      batches_{0}, exports_{0}, export_bytes_{0};
};

}  // namespace sp::session
