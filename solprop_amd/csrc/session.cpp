// A march as a session of events (see session.hpp).
#include "session.hpp"

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "march.hpp"

namespace sp::session {

Session::Session(Propagator& ws, int n_modes, int n_omega, const Plan& plan, double z,
                 const double* field)
    : ws_(ws), plan_(plan) {
  SP_CHECK(field != nullptr, "a session starts from a launch field");
  SP_CHECK(!plan_.planes.empty(), "a session needs its planes, the last the run's end");
  for (std::size_t i = 1; i < plan_.planes.size(); ++i)
    SP_CHECK(plan_.planes[i] > plan_.planes[i - 1], "the planes must increase");
  SP_CHECK(plan_.planes.back() > z, "the run's end lies beyond the start");
  SP_CHECK(plan_.step_planes >= 0, "step_planes is a count");
  SP_CHECK(plan_.attempt_capacity >= 1 && plan_.call_attempts >= 1,
           "the chunk and the call hold at least one attempt");
  SP_CHECK(plan_.windows.size() % 2 == 0, "windows are (lo, hi) pairs");
  elements_ = static_cast<std::size_t>(n_modes) * static_cast<std::size_t>(n_omega);
  // The first plane strictly beyond the start (a plane at the start is the launch).
  while (next_plane_ < plan_.planes.size() && plan_.planes[next_plane_] <= z) ++next_plane_;
  launch_pending_ = plan_.launch;
  ws_.load(z, field);
  z_ = z;
  last_event_ = Clock::now();
  chunk_.reserve(static_cast<std::size_t>(plan_.attempt_capacity) * kWideColumns);
  publish();
}

void Session::require_open() const {
  if (state_.load() == State::closed) throw LifecycleError("the session is closed");
}

void Session::publish() noexcept {
  z_pub_.store(z_);
  plane_pub_.store(static_cast<long>(next_plane_));
}

Status Session::status() const noexcept {
  Status s;
  s.z = z_pub_.load();
  s.plane_index = plane_pub_.load();
  s.attempts = attempts_.load();
  s.rejected = rejected_.load();
  s.state = state_.load();
  s.outcome = outcome_.load();
  s.generation = generation_.load();
  s.march_calls = calls_.load();
  s.batches = batches_.load();
  s.exports = exports_.load();
  s.export_bytes = export_bytes_.load();
  return s;
}

void Session::observe(Batch& b, Observation kind, long plane) {
  b.event.kind = EventKind::observation;
  b.event.observation = kind;
  b.event.z = z_;
  b.event.plane_index = plane;
  b.event.ticket = true;
  if (plan_.statistics) {
    b.stats = ws_.statistics(plan_.statistics,
                             plan_.windows.empty() ? nullptr : plan_.windows.data(),
                             static_cast<int>(plan_.windows.size() / 2), StatExtras{});
    b.event.has_stats = true;
  }
}

void Session::terminal(Batch& b, Outcome outcome) {
  b.event.kind = EventKind::terminal;
  b.event.outcome = outcome;
  b.event.z = z_;
  b.event.plane_index = static_cast<long>(next_plane_);
  // The state a terminal event names can be exported (a checkpoint of where it ended).
  b.event.ticket = true;
}

Batch Session::issue(Batch b) {
  b.sequence = ++sequence_;
  b.event.generation = generation_.load();
  b.attempts.swap(chunk_);
  chunk_.clear();
  last_event_ = Clock::now();
  pending_ = b.sequence;
  batches_ += 1;
  publish();
  if (b.event.kind == EventKind::terminal) {
    outcome_.store(b.event.outcome);
    state_.store(State::finished);
  } else {
    state_.store(State::pending);
  }
  return b;
}

Batch Session::advance(long work_limit, double deadline) {
  std::lock_guard<std::mutex> lock(mu_);
  require_open();
  if (state_.load() == State::pending)
    throw LifecycleError("acknowledge the pending batch before advancing the session again");
  if (state_.load() == State::finished)
    throw LifecycleError("the session has ended and cannot advance");
  const auto started = Clock::now();
  const auto elapsed = [&] {
    return std::chrono::duration<double>(Clock::now() - started).count();
  };
  const long budget = std::max(work_limit, 1L);
  long made = 0;
  Batch b;
  if (launch_pending_) {
    launch_pending_ = false;
    observe(b, Observation::launch, -1);
    return issue(std::move(b));
  }
  if (next_plane_ >= plan_.planes.size()) {
    terminal(b, Outcome::completed);
    return issue(std::move(b));
  }
  const double target = plan_.planes[next_plane_];
  const long capacity = plan_.attempt_capacity;
  std::vector<double> rows;
  while (!march::reached(z_, target)) {
    if (stop_.load(std::memory_order_acquire)) {
      b.event.reason = Reason::stopped;
      terminal(b, Outcome::stopped);
      return issue(std::move(b));
    }
    const long held = static_cast<long>(chunk_.size() / kWideColumns);
    if (held >= capacity || (made > 0 && (made >= budget || elapsed() >= deadline))) {
      b.event.kind = EventKind::progress;
      b.event.reason = Reason::bound;
      b.event.z = z_;
      b.event.plane_index = static_cast<long>(next_plane_);
      return issue(std::move(b));
    }
    Bounds bounds;
    bounds.target = target;
    bounds.attempts = std::max(1L, std::min({plan_.call_attempts, budget - made, capacity - held}));
    bounds.accepted = plan_.step_planes > 0 ? plan_.step_planes - since_observed_ : 0;
    bounds.seconds = std::max(0.0, std::min(plan_.call_seconds, deadline - elapsed()));
    bounds.stop = &stop_;
    rows.clear();
    sp_march_result r{};
    const auto call_started = Clock::now();
    Outcome failed = Outcome::none;
    try {
      ws_.advance_bounded(bounds, kWideColumns, rows, &r);
    } catch (const march::ConvergenceError& e) {
      failed = Outcome::failed;
      b.event.reason = Reason::floor;
      b.event.category = Category::convergence;
      b.event.stage = Stage::trial;
      b.event.message = e.what();
      b.event.floor_proposal = r.proposal;
    } catch (const Error& e) {
      failed = Outcome::failed;
      b.event.reason = Reason::error;
      b.event.category = e.code == kValue ? Category::value : Category::backend;
      b.event.stage = Stage::trial;
      b.event.message = e.what();
    } catch (const std::exception& e) {
      failed = Outcome::failed;
      b.event.reason = Reason::error;
      b.event.category = Category::backend;
      b.event.stage = Stage::trial;
      b.event.message = e.what();
    }
    calls_ += 1;
    // Wall times in the chunk count from the previous event.
    const double offset = std::chrono::duration<double>(call_started - last_event_).count();
    const std::size_t n = rows.size() / kWideColumns;
    for (std::size_t i = 0; i < n; ++i) {
      double* row = rows.data() + i * kWideColumns;
      row[5] += offset;
      chunk_.insert(chunk_.end(), row, row + kWideColumns);
      const bool accepted = row[3] != 0.0;
      attempts_ += 1;
      if (accepted) {
        generation_ += 1;
        since_observed_ += 1;
      } else {
        rejected_ += 1;
      }
    }
    made += static_cast<long>(n);
    z_ = r.z;
    publish();
    if (failed != Outcome::none) {
      terminal(b, failed);
      return issue(std::move(b));
    }
    if (plan_.step_planes > 0 && since_observed_ >= plan_.step_planes &&
        !march::reached(z_, target)) {
      since_observed_ = 0;
      observe(b, Observation::step, static_cast<long>(next_plane_));
      return issue(std::move(b));
    }
  }
  const long index = static_cast<long>(next_plane_++);
  if (plan_.step_planes > 0 && since_observed_ >= plan_.step_planes) since_observed_ = 0;
  observe(b, Observation::plane, index);
  return issue(std::move(b));
}

void Session::acknowledge(std::uint64_t sequence) {
  std::lock_guard<std::mutex> lock(mu_);
  require_open();
  if (state_.load() != State::pending || sequence != pending_)
    throw LifecycleError("no pending batch has that sequence number");
  pending_ = 0;
  state_.store(State::ready);
}

void Session::export_field(std::uint64_t sequence, std::uint64_t generation,
                           const double* weight, bool optics, double* out) {
  std::lock_guard<std::mutex> lock(mu_);
  require_open();
  if (pending_ == 0 || sequence != pending_)
    throw LifecycleError("a ticket is valid only while its batch is pending");
  if (generation != generation_.load())
    throw LifecycleError("the ticket names another generation of the state");
  ws_.export_state(weight, optics, out);
  exports_ += 1;
  export_bytes_ += 16 * elements_;
}

bool Session::fail_outside(const std::string& message, Batch* out) {
  std::lock_guard<std::mutex> lock(mu_);
  if (state_.load() != State::ready) return false;
  Batch b;
  b.event.reason = Reason::error;
  b.event.category = Category::backend;
  b.event.message = message;
  terminal(b, Outcome::failed);
  *out = issue(std::move(b));
  return true;
}

void Session::close() noexcept {
  std::lock_guard<std::mutex> lock(mu_);
  state_.store(State::closed);
  chunk_.clear();
  chunk_.shrink_to_fit();
}

}  // namespace sp::session
