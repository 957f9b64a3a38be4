// Host launchers of the stage kernels, shared by the workspace and the pass-level
// entry points. Each enqueues one kernel on `s`; none synchronises.
#pragma once

#include "device.hpp"

namespace sp {

constexpr int kAcceptThreads = 256;

inline unsigned accept_blocks(const DeviceProblem& P) {
  return blocks_for(static_cast<std::size_t>(P.C) * P.N, kAcceptThreads);
}

void launch_rk_entry_envelope(const DeviceProblem& P, const TrialRecord* R, const cplx* y,
                              const cplx* k, cplx* fine, cplx* einv, int stage, hipStream_t s);
void launch_env_raman_drive(const DeviceProblem& P, const TrialRecord* R, const cplx* fine,
                            cplx* drive, int stage, hipStream_t s);
// Over `rows` consecutive rows of T packed points, one kernel spectrum for all.
void launch_raman_product(cplx* z, const cplx* H, const cplx* roots2T, int T, int rows,
                          hipStream_t s);
void launch_raman_kernel_fill(const DeviceProblem& P, const TrialRecord* R, int stage, cplx* z,
                              hipStream_t s);
void launch_raman_kernel_post(const cplx* z, cplx* H, const cplx* roots2T, int T, double dt,
                              hipStream_t s);
void launch_env_response(const DeviceProblem& P, const TrialRecord* R, cplx* fine,
                         const cplx* conv, int stage, hipStream_t s);
// Stage `stage - 1`'s finish and stage `stage`'s entry in one launch (stage >= 1).
void launch_rk_finish_entry(const DeviceProblem& P, const TrialRecord* R, const cplx* y,
                            cplx* k, cplx* fine, cplx* einv, int stage, hipStream_t s);
// The step's acceptance; it first finishes the last stage from its transformed response
// (`fine`, `einv`), bin by bin.
void launch_rk_accept(const DeviceProblem& P, const TrialRecord* R, const cplx* y, cplx* k,
                      cplx* adv, cplx* recycled, double* partials, const cplx* fine,
                      const cplx* einv, hipStream_t s);
void launch_commit_masked(const DeviceProblem& P, const cplx* adv, cplx* y, hipStream_t s);
void launch_copy_points(const cplx* src, cplx* dst, long n, hipStream_t s);
// The envelope route's split-step absorber, around the caller's inverse and forward
// transforms of the coarse grid (the unpack is `launch_commit_masked` from the buffer).
void launch_env_split_pack(const DeviceProblem& P, const cplx* y, cplx* buf, hipStream_t s);
void launch_env_split_collar(const DeviceProblem& P, const TrialRecord* R, double dz, cplx* buf,
                             hipStream_t s);

}  // namespace sp
