// A batched march: many independent instances of the modal field route (a parameter
// scan's points, each with its own grid, modes, operator and adaptive step control)
// advanced together, their trials run as one captured graph of grouped launches.
//
// The reference runs a scan's points one process each (src/solprop/driver/scan.py:30,
// a process pool); here every round of the batch is one graph replay on one stream: the
// records of all instances uploaded in one copy, then each kernel of the trial launched
// once over every instance (an instance index found per workgroup), and the norms of
// all instances landing in mapped host memory. The host then concludes each instance's
// attempt and plans its next one (its own march: decisions, rows and failures are its
// own), spread over a small pool of threads. Each instance's march is the one its
// workspace would make alone, to the bit: the grouped kernels run the plain kernels'
// bodies on the instance's own buffers.
#pragma once

#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "field_ws.hpp"

namespace sp {

// A fixed pool of host threads running one indexed loop at a time.
class WorkerPool {
 public:
  explicit WorkerPool(int threads);
  ~WorkerPool();
  WorkerPool(const WorkerPool&) = delete;
  WorkerPool& operator=(const WorkerPool&) = delete;
  // f(i) for every i < n, over the pool and the calling thread; returns when all are done.
  // The first exception a call throws is thrown again here (the other calls still run).
  void run(int n, const std::function<void(int)>& f);

 private:
  void work();
  void call(int i);
  std::exception_ptr failure_;
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable wake_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0, next_ = 0, running_ = 0;
  long generation_ = 0;
  bool stop_ = false;
};

struct ScanTiming {
  long rounds = 0;
  double device_ms = 0.0;  // the host's wait for the rounds' graphs
  double host_ms = 0.0;    // planning and concluding, over all rounds
};

class Scan {
 public:
  Scan(const sp_problem_desc* descs, int n);
  ~Scan();
  Scan(const Scan&) = delete;
  Scan& operator=(const Scan&) = delete;

  int size() const { return static_cast<int>(inst_.size()); }
  void load(int i, double z, const double* field);
  // Advance every instance towards its target, at most `work_limit` attempts each; each
  // instance's rows (end, step, norm, accepted per attempt), result, status and message.
  void advance(const double* targets, long work_limit, std::vector<std::vector<double>>& rows,
               sp_march_result* out, int* status, std::vector<std::string>& messages);
  void state(int i, double* out);
  FieldWorkspace& instance(int i) { return *inst_[static_cast<std::size_t>(i)]; }
  const ScanTiming& timing() const { return timing_; }
  // The tuned schedule as JSON: per grouped pass, its shape, the choice and each
  // candidate's measured microseconds.
  const std::string& schedule() const { return schedule_; }
  long device_bytes() const;

 private:
  struct Group {  // one grouped launch of the instance kernels
    int offset = 0;  // into first_dev_
    int blocks = 0;
  };
  struct FftGroup {  // one grouped launch of a transform pass shape
    int log2L = 0, log2G = 0;
    FftJob* jobs = nullptr;
    int* first = nullptr;
    int n = 0, blocks = 0;
    std::vector<FftJob> host;    // the jobs as planned (at log2G)
    std::vector<int> sequences;  // each job's sequences a row
  };
  void release();
  void plan_launches();
  // The autotuner: for each grouped transform pass, the sequences a workgroup owns (a
  // launch shape only: each sequence's arithmetic is the same whichever is chosen) timed
  // over every candidate the pass kernels provide, the fastest kept; chosen once per pass
  // shape and process (`SOLPROP_MI355X_AUTOTUNE=0`: the plans' own choice).
  void tune(std::vector<FftGroup>& groups, int sign, const char* direction);
  void set_schedule(FftGroup& g, int log2G);
  std::vector<FftGroup> group_passes(const std::vector<std::vector<FftJob>>& per_instance,
                                     const std::vector<const FftPlan*>& plans);
  void enqueue_round();
  BatchGrid grid(const Group& g) const;

  hipStream_t stream_ = nullptr;
  std::vector<std::unique_ptr<FieldWorkspace>> inst_;
  AdamsRecord* rec_host_ = nullptr;  // (n) pinned
  AdamsRecord* rec_dev_ = nullptr;   // (n)
  BatchInstance* inst_dev_ = nullptr;
  int* first_dev_ = nullptr;
  std::vector<void*> allocations_;
  // prefix_current_: the free-current tile prefix of the instances that do not run the
  // fused node passes (which need none).
  Group commit_, predict_, rate_, prefix_, current_, prefix_current_, response_, finish_,
      correct_, norms_;
  std::vector<FftGroup> inverse_, forward_;
  hipGraphExec_t graph_ = nullptr;
  hipEvent_t done_ = nullptr;
  std::unique_ptr<WorkerPool> pool_;
  ScanTiming timing_;
  std::string schedule_ = "[]";
};

}  // namespace sp
