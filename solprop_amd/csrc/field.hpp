// The modal field-grid route under the exponential Adams march: device record and
// launchers.
//
// One source evaluation is: scatter the lab-frame modal spectrum into the packed
// real-transform input (1/T folded in) -> inverse FFT of T/2 packed points -> the node
// passes (synthesis over the radial nodes, Kerr, the photoionisation rate and the two
// cumulative plasma integrals as wave-tile scans, the collars, and the projection of
// the local polarisation and both held current streams back onto the modes) ->
// forward FFT of the 3C packed streams -> finish (unpack, held-current assembly on the
// active bins, prefactor and spectral collar).
//
// With more than kMaxModes modes (the radial route's Hankel spectrum) the synthesis and
// the projection are matrix-core GEMMs instead: the synthesis writes the node field
// (Q, T), the node passes read it, the response writes each node's un-gated
// polarisation and current, and the projection applies the time collars on the way out.
//
// The scans' tiles are one wavefront each: 64 samples holding 63 trapezoid increments,
// the last sample shared with the next tile, so every increment's two ends sit in one
// wave and no lane reaches outside it. Tile k starts at sample 63 k.
#pragma once

#include "device.hpp"

namespace sp {

constexpr int kMaxModes = 8;       // modes the node passes synthesise and project in registers
constexpr int kMaxAdamsWidth = 8;  // 1 + the highest scored order (MAX_ADAMS_ORDER 6)
constexpr int kMaxScored = 3;      // the step's own order and two neighbours
constexpr int kWave = 64;          // lanes of a wavefront
constexpr int kTileSpan = kWave - 1;  // trapezoid increments per scan tile
constexpr int kNodeWaves = 8;      // waves per workgroup of the rate and current passes
constexpr int kNodesPerWave = 4;   // nodes each wave of those passes runs
constexpr int kLdsTableIntervals = 1024;  // rate-table intervals held in LDS
constexpr int kResponseWaves = 16; // waves per workgroup of the node response
constexpr int kCorrectThreads = 256;  // the corrector's widest workgroup
constexpr int kLsThreads = 256;    // workgroup of the low-storage register passes
constexpr int kLsRun = 4096;       // workgroup sums one workgroup of their reduction reads

// Scan tiles of a fine grid of T samples: ceil((T - 1) / 63).
__host__ __device__ inline int scan_tiles(int T) { return (T - 1 + kTileSpan - 1) / kTileSpan; }

// What one Adams trial uploads (host scalars and the Lagrange tables of its stencil).
struct AdamsRecord {
  double dz;
  double x_w[kMaxTerms];                 // whole-step Φ weights
  double rate_end_w[kMaxTerms];          // D(z + dz) weights
  double pref_w[kMaxTerms];              // prefactor at z + dz
  double density, preionisation;         // at z + dz
  double slot_rate_w[kMaxAdamsWidth][kMaxTerms];  // D(z_j) weights of each store slot
  double nodes[kMaxAdamsWidth];          // θ_j of the rows read, newest first
  double pred_c[kMaxAdamsWidth][kMaxAdamsWidth];  // [j][k], j,k < order
  double corr_c[kMaxAdamsWidth][kMaxAdamsWidth];  // [j][k] over X = (source, rows)
  double rows_c[kMaxScored][kMaxAdamsWidth][kMaxAdamsWidth];
  double comp_pred_w[kMaxScored][kMaxAdamsWidth];      // weights of nodes[:k]
  double comp_corr_w[kMaxScored][kMaxAdamsWidth + 1];  // weights of [1, nodes[:k]]
  int slots[kMaxAdamsWidth];             // store rows read, newest first
  // `action` is read by the grouped kernels of a batched march only: 0 the instance is
  // idle this round, 1 it starts (the source at its state), 2 it takes a trial.
  int order, width, n_orders, n_read, compensated, free_slot, action, pad1;
  int orders[kMaxScored];
  int pad2;
  double raman_rate[kMaxTerms];          // Raman decay rates at the source position
  double accepted_dz;                    // the step the record's commit finishes
};

struct FieldBuffers {
  cplx* y;          // (C, N) committed state
  cplx* store;      // (S, C, N) untransported source rows
  cplx* states;     // (S, C, N) the state each row was evaluated at
  cplx* carried;    // (W-1, C, N) transported compensation of the rows read
  cplx* lab;        // (C, N) predicted lab-frame field
  cplx* source;     // (C, N) source at the prediction
  cplx* advanced;   // (C, N) corrected field
  cplx* modal_t;    // (C, T/2) packed real modal fields
  cplx* streams;    // (3, C, T/2) packed polarisation, gJ, g'J
  double* rate;     // (Q, T) ionisation rate at each node and sample
  double* frac;     // (Q, T) ionised fraction
  double* free_local;  // (Q, T) free current since the start of the sample's scan tile
  double* tile_x;   // (Q, tiles) exposure tile sums, then offsets
  double* tile_f;   // (Q, tiles) free-current tile sums, then offsets
  double* x_local;  // (Q, T) exposure since the start of the sample's scan tile (band rule)
  cplx* raman_drive;     // (Q, T) the drive on the doubled grid, two reals a point; then
                         // its convolution with the kernel (times 2T)
  cplx* raman_H;         // (T + 1) the kernel's spectrum on the doubled grid, dt folded in
  const cplx* roots_2T;  // unit roots of 2T
  cplx* modal_h;         // (C, T/2) packed Hilbert transforms of the modal fields
                         // (the analytic Raman drive only)
  double* node_hilbert;  // (Q, T) Hilbert transform of the node field (dense route,
                         // analytic Raman drive)
  double* node_field;    // (Q, T) field at the nodes (dense route only)
  double* node_streams;  // (Q, 2, T) un-gated polarisation and current (dense route only)
  double* partials; // (1 + kMaxScored, blocks) masked squared sums of each block
  double* norms;    // (1 + kMaxScored) their totals, in the host's mapped memory
  const cplx* roots_T;  // unit roots of T
  const double* rate_c; // (4, n_intervals) rate table
  // The fused current-and-response pass (fused-node route with plasma): each node
  // group's projections of its nodes' polarisation and tile-local current, and each
  // group's projection of its nodes' free-current tile totals.
  double* group_streams;  // (G, 2C, T), G = ceil(Q / 32)
  double* tile_p;         // (G, C, tiles)
  // (Q, tiles): 1 where the band rate pass found a node's whole filter window zero; its
  // rate and within-tile exposure are then zero, and a sparse rate pass does not store
  // them (the fused passes read the flag instead).
  unsigned char* zero_tiles;
};

// Nodes one workgroup of the band passes runs (kNodeWaves waves of kNodesPerWave).
constexpr int kNodesPerGroup = kNodeWaves * kNodesPerWave;
__host__ __device__ inline int node_groups(int Q) {
  return (Q + kNodesPerGroup - 1) / kNodesPerGroup;
}

constexpr int kMaxBandHalf = 8;  // most taps each side of the band rule's filter

struct PlasmaTable {
  double log_field_min, d_log_field, field_min, field_max, rate_max;
  double charge_mass_ratio, ionisation_potential;
  int n_intervals;
  // The running integrals' rule: one tap (1.0) is the trapezoid; more are the band rule,
  // c_0 .. c_K of a symmetric filter applied to each integrand before the trapezoid,
  // the window's ends closed by point reflection and the exposure floored at zero.
  int n_taps;
  double taps[kMaxBandHalf + 1];
};

// ------------------------------------------------------------ batched marches

constexpr int kMaxBatchModes = 6;  // modes an instance of a batched march may have

enum BatchAction : int { kBatchIdle = 0, kBatchStart = 1, kBatchTrial = 2 };

// One instance of a batched march as the grouped kernels read it (device memory): its
// problem, buffers and plasma table, and its corrector's workgroups.
struct BatchInstance {
  DeviceProblem P;
  FieldBuffers B;
  PlasmaTable pt;
  int correct_blocks;
};

// A grouped launch: `n` instances (device arrays `inst` and their records `records`),
// instance i owning workgroups first[i] .. first[i+1] - 1 of a grid of `blocks` (device
// array `first`).
struct BatchGrid {
  const BatchInstance* inst;
  const AdamsRecord* records;
  const int* first;
  int n;
  int blocks;
};

// The grouped kernels of one batched trial, in launch order: each runs its pass for every
// instance whose record asks for one (a start evaluates the source at the state and
// stores it; a trial predicts, evaluates and corrects), idle instances' workgroups
// exiting at once. The commit applies a pending accepted step (free_slot >= 0) whatever
// the action. Grids are sized on the host: commit ceil(C N / 256), predict and correct
// ceil(C N / 64), the two node passes tiles x ceil(Q / 32) (plasma instances only), the
// prefixes Q, the response T / 64, the finish ceil(C (T/4 + 1) / 64), the norms 1.
void launch_batch_commit(const BatchGrid& g, hipStream_t s);
void launch_batch_predict(const BatchGrid& g, hipStream_t s);
void launch_batch_rate(const BatchGrid& g, hipStream_t s);
void launch_batch_prefix(const BatchGrid& g, bool current, hipStream_t s);
void launch_batch_current(const BatchGrid& g, hipStream_t s);
void launch_batch_response(const BatchGrid& g, hipStream_t s);
void launch_batch_finish(const BatchGrid& g, hipStream_t s);
void launch_batch_correct(const BatchGrid& g, hipStream_t s);
void launch_batch_norms(const BatchGrid& g, hipStream_t s);

unsigned correct_threads(const DeviceProblem& P);
unsigned correct_blocks(const DeviceProblem& P);

// The Raman response (real square drive): the damped kernel at the record's rates,
// packed two reals a point over the doubled grid (`z`, T points) for its transform; and
// the drive E² at every node, zero on the padding half, into B.raman_drive.
void launch_field_raman_kernel_fill(const DeviceProblem& P, const AdamsRecord* R, cplx* z,
                                    hipStream_t s);
void launch_field_raman_drive(const DeviceProblem& P, FieldBuffers B, hipStream_t s);

// The in-stage absorber's stream, -alpha(t) a_c(t) / 2 from the modal fields, written as
// stream P.absorber_stream of B.streams.
void launch_field_absorber_stream(const DeviceProblem& P, FieldBuffers B, bool accumulate,
                                  hipStream_t s);
// The split-step absorber on the committed state, in three launches around the caller's
// transforms of B.modal_t: pack y for the real inverse transform; after it, multiply by
// exp(-alpha(t) h / 2) (h: R->accepted_dz, or `dz` when R is null); after the forward
// transform, unpack into y on the active bins (zero elsewhere).
void launch_field_split_pack(const DeviceProblem& P, FieldBuffers B, hipStream_t s);
void launch_field_split_collar(const DeviceProblem& P, const AdamsRecord* R, double dz,
                               FieldBuffers B, hipStream_t s);
void launch_field_split_unpack(const DeviceProblem& P, FieldBuffers B, hipStream_t s);

// The predictor and corrector are compiled per stencil width (2..7); `width` must be
// the record's. With `commit` the predictor first applies the accepted step's commit
// (launch_adams_commit) to each bin it predicts, the record's free slot included.
void launch_adams_predict(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          int width, hipStream_t s, bool commit = false);
void launch_field_scatter(const DeviceProblem& P, const cplx* spectrum, FieldBuffers B,
                          hipStream_t s);
// The modal spectra's Hilbert transforms packed the same way into B.modal_h: -i Y on the
// bins strictly between DC and Nyquist, zero at both (the analytic signal's imaginary
// part, src/solprop/fourier.py:243-300).
void launch_field_scatter_hilbert(const DeviceProblem& P, const cplx* spectrum, FieldBuffers B,
                                  hipStream_t s);
// The dense route (B.node_field set) expects the node field written; it leaves the
// node streams for the projection GEMM. `marks`, when given, holds five events recorded after the rate and exposure tiles,
// their prefix, the free-current tiles, their prefix and the node response (a run
// without plasma records the first four back to back).
// `exposure` runs the band rate pass under either rule (one tap is the trapezoid, the
// same bits), so that the within-tile exposure is stored for the integration
// diagnostics.
void launch_node_passes(const DeviceProblem& P, const AdamsRecord* R, const PlasmaTable& pt,
                        FieldBuffers B, hipStream_t s, hipEvent_t* marks = nullptr,
                        bool exposure = false);
// The plasma column's integration diagnostics from the unfused passes' stored column
// (run with `exposure`): one workgroup a node writes its six parts to `parts` (6 Q) —
// max(-min dS, 0), max dS, |paid - freed|, freed, max |f E| over the first 2K samples
// and anywhere, with S = tile offset + within-tile exposure (before the floor), paid the
// trapezoid of W (1 - f) and freed f(T-1) - f(0) — then their maxima over the nodes go
// to `out` (6).
void launch_plasma_diagnostics(const DeviceProblem& P, const PlasmaTable& pt, FieldBuffers B,
                               double* parts, double* out, hipStream_t s);
// Whether the fused node passes apply: the fused-node route (1..8 modes) with plasma and
// the buffers they write (the band passes' exposure, the group streams and totals).
__host__ __device__ inline bool fused_node_passes(const DeviceProblem& P, const FieldBuffers& B) {
  return P.has_plasma && !B.node_field && B.group_streams && B.tile_p && B.x_local &&
         B.zero_tiles && P.C >= 1 && P.C <= kMaxModes;
}
// The same streams from four launches and no (Q, T) array beyond the rate and the
// within-tile exposure: the rate pass (either rule; the trapezoid is the band rule with
// one tap), the exposure prefix, the free-current pass fused with the node response
// (each workgroup's projection of its nodes' polarisation and tile-local current, and of
// their free-current tile totals), and the combination, which adds every tile's
// projected free-current offset, sums the groups in order and applies the collars and
// the density. The stored fraction and free current are not written. `marks` as above
// (the fourth after the combination, the fifth at once).
void launch_node_passes_fused(const DeviceProblem& P, const AdamsRecord* R,
                              const PlasmaTable& pt, FieldBuffers B, hipStream_t s,
                              hipEvent_t* marks = nullptr);
void launch_field_finish(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         cplx* out, hipStream_t s);
// The corrector, writing its block sums to B.partials; then their reduction into B.norms.
// With `finish` the corrector first finishes its bin's source from the forward transform's
// streams (launch_field_finish into B.source, bin by bin) instead of reading B.source.
void launch_adams_correct(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                          int width, hipStream_t s, bool finish = false);
void launch_adams_norms(const DeviceProblem& P, FieldBuffers B, hipStream_t s);
// The low-storage RK4IP step's register passes (1..4, each after its source evaluation)
// and the three masked sums they leave (|S2 - g|^2, |S3 - S2|^2, |A|^2) into `norms`.
unsigned ls_blocks(const DeviceProblem& P);
// Doubles of the passes' partial sums: three rows of ls_blocks, and above kLsRun
// workgroups three rows of the first level's run sums after them.
std::size_t ls_partials(const DeviceProblem& P);
// With `finish` (the field route) each pass finishes its bin's source from the forward
// transform's streams instead of reading B.source.
void launch_ls_pass(int pass, const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                    cplx* acc, cplx* prev, double* partials, hipStream_t s,
                    bool finish = false);
void launch_ls_norms(const DeviceProblem& P, const double* partials, double* norms,
                     hipStream_t s);
// Commits the held trial into store slot `slot`; reads no record, so the host may go on
// filling the next trial's record while it runs.
void launch_adams_accept(const DeviceProblem& P, int slot, bool compensated, FieldBuffers B,
                         hipStream_t s);
// The commit with its slot read from the record (`free_slot`).
void launch_adams_commit(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                         hipStream_t s);
// The quadrature estimate from two finished response spectra of n bins:
// sqrt(sum |coarse - fine|^2 / sum |fine|^2) into out[0] (NaN when the fine sum is zero),
// one workgroup summing in a fixed order.
void launch_quadrature_score(const cplx* fine, const cplx* coarse, long n, double* out,
                             hipStream_t s);
// The upload of `count` records from pinned host memory (`host` as the device addresses
// it) to `device`, as a kernel of one workgroup that reads them at system scope.
void launch_record_fetch(const AdamsRecord* host, AdamsRecord* device, int count,
                         hipStream_t s);
void launch_adams_start(const DeviceProblem& P, const AdamsRecord* R, FieldBuffers B,
                        hipStream_t s);

}  // namespace sp
