// Device-side views of a loaded problem and the per-trial scalar record.
#pragma once

#include "common.hpp"

namespace sp {

constexpr int kMaxTerms = 4;    // Φ and prefactor terms, Raman rows
constexpr int kMaxStages = 8;   // Runge-Kutta stages (DOPRI5 has 7)
constexpr int kRowDoubles = 24; // doubles per stage row of the scalar record

// One stage's host scalars (the reference's StageScalars, translate/stage.py:170).
struct StageRow {
  double phi_w[kMaxTerms];    // W_t(z_i) - W_t(z_anchor)
  double pref_w[kMaxTerms];   // p_t(z_i)
  double noise_w[kMaxTerms];  // W_t(z_i), for the modified noise phase
  double raman_rate[kMaxTerms];
  double density;
  double preionisation;
  double taper;
  double raman_version;  // changes when the rates change (the memo key)
  double pad[4];
};
static_assert(sizeof(StageRow) == kRowDoubles * sizeof(double), "stage row layout");

// What a trial uploads: the step and its stage rows, plus the whole step's row.
struct TrialRecord {
  double dz;
  double accepted_dz;  // the step the commit at this trial's head finishes (split-step)
  double pad[kRowDoubles - 2];
  StageRow stage[kMaxStages];
  StageRow whole;
};

// The problem's arrays, resident in HBM for the workspace's life.
struct DeviceProblem {
  int kind;       // 0 envelope (complex time), 1 field (real time)
  int C;          // columns = modes (one polarisation component)
  int N;          // coarse bins
  int T;          // fine samples
  int NF;         // fine bins: T (envelope) or T/2+1 (field)
  int n_phi, n_pref;
  int Q;          // quadrature nodes
  int n_raman;    // Raman rows (0: no Raman)
  int has_kerr, kerr_real, has_plasma, noise_modified;
  double kerr_coefficient, kerr_self, kerr_conjugate;
  double raman_fraction, raman_epsilon0;
  // The field routes' Raman drive: 0 the real square E^2 (the harmonic kept), 1 half the
  // analytic signal's squared modulus, (E^2 + H[E]^2) / 2 (a grid without the harmonic).
  int raman_analytic;
  double plasma_cmr, plasma_ip;  // e^2/m_e and the ionisation potential, J
  double dt_fine, inv_T, inv_2T;
  const int* fine_to_coarse;   // (NF) coarse bin of each fine bin, -1 in the gaps
  const int* crop;             // (N) fine bin of each coarse bin
  const unsigned char* active; // (N)
  const unsigned char* normmask;  // (N)
  const double* omega_window;  // (N)
  const double* omega;         // (N)
  const double* t_window;      // (T)
  const double* t_deriv;       // (T)
  const cplx* phi;             // (n_phi, C, N)
  const cplx* pref;            // (n_pref, C, N)
  const cplx* noise;           // (C, N) or null
  const double* synthesis;     // (Q, C)
  const double* projection;    // (C, Q)
  const double* raman_rows;    // (n_raman, 2T)
  // The temporal absorber: the power rate alpha(t), 1/m, on the fine grid (in-stage
  // scheme) or the coarse grid (split-step scheme); null without one. The in-stage term
  // -alpha E / 2 rides the forward transform as stream `absorber_stream` (-1: none); on
  // the envelope route that stream is the fine row after the polarisation's.
  const double* absorber_rate;
  int absorber_stream;
  // The Runge-Kutta tableau.
  int n_stages;
  const double* tab_a;         // (n_stages, n_stages)
  const double* tab_b;         // (n_stages)
  const double* tab_e;         // (n_stages)
};

}  // namespace sp
