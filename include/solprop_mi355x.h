/*
 * C ABI of the MI355X propagation library (libsolprop_mi355x.so).
 *
 * The library replaces the compiled half of solprop's native backend: the nanobind
 * module `_core` (native/src/bindings/module.cpp:786-832; Python stub
 * native/src/solprop_native/_core.pyi). Its Problem record, Workspace, Session and
 * Probe become the plain structs and entry points below: plain pointers and sizes, no
 * Python or torch types. Host arrays are copied in at create/load and copied out at
 * export, as the reference's are (native/src/general/trial.cpp:356-390; backends.md
 * "Outputs: host arrays only"). Device buffers are owned by the workspace for its life.
 *
 * Every entry point returns 0 on success or a status (SP_VALUE, SP_BACKEND,
 * SP_CONVERGENCE, SP_STATE) and writes a message into `msg` (at most `msglen` bytes,
 * NUL-terminated) -- the reference's refusals and errors
 * (native/src/solprop_native/_core.pyi:23-28 BackendRefusal; kernel-contract.md
 * "Native errors carry a stage, a category and a message").
 *
 * Complex arrays are interleaved float64 pairs (numpy complex128 layout).
 */
#ifndef SOLPROP_MI355X_H
#define SOLPROP_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OK 0
#define SP_VALUE 1
#define SP_BACKEND 2
#define SP_CONVERGENCE 3
#define SP_STATE 4

#define SP_MAX_TERMS 4

/* A scalar function of z (the reference's stage tables, translate/stage.py:170 and
 * native/include/solprop_native/march/stage_tables.hpp): kind 0 constant `value`,
 * kind 1 the position z itself, kind 2 piecewise Chebyshev on `n_intervals` intervals
 * with `edges` (n_intervals+1) and `coefficients` (n_intervals, degree+1), evaluated by
 * Clenshaw's recurrence. */
typedef struct sp_scalar_path {
  int kind;
  double value;
  int n_intervals;
  int degree;
  const double* edges;
  const double* coefficients;
} sp_scalar_path;

/* The problem record: the subset of `_core.Problem` (_core.pyi:51-466) that the
 * envelope-GNLSE and modal field-grid routes read. Shapes: C = n_modes, N = n_omega,
 * T = n_to, Q = n_nodes. */
typedef struct sp_problem_desc {
  /* Dims (_core.pyi Problem.Dims) */
  int kind; /* 0 envelope (complex time samples), 1 field (real), 2 the free-space
             * Cartesian envelope on an (nx, ny, nt) grid (grid_nx, grid_ny below) */
  int n_t, n_to, n_modes, n_nodes, n_omega;
  /* Grid (Problem.Grid) */
  const int64_t* crop_index;        /* (N) */
  const uint8_t* active_mask;       /* (N) */
  const uint8_t* norm_mask;         /* (N) */
  const double* omega_window;       /* (N) */
  const double* omega;              /* (N) absolute angular frequency */
  const double* time_window;        /* (T) */
  const double* time_window_derivative; /* (T) 1/s */
  double dt_fine;
  double z_end;
  /* Linear (Problem.Linear): ΔΦ = Σ_t (W_t(z1) - W_t(z0)) Φ_t, prefactor Σ p_t(z) P_t */
  int n_phi;
  const double* phi_terms;          /* complex (n_phi, C, N) */
  sp_scalar_path phi_paths[SP_MAX_TERMS];
  /* dW_t/dz: the operator's rate D(z) = Σ_t W_t'(z) Φ_t (translate/linop.py phi_rate),
   * read by the exponential Adams compensation of a varying operator */
  sp_scalar_path phi_rate_paths[SP_MAX_TERMS];
  int n_prefactor;
  const double* prefactor_terms;    /* complex (n_prefactor, C, N) */
  sp_scalar_path prefactor_paths[SP_MAX_TERMS];
  /* Transform (Problem.Transform) */
  const double* synthesis;          /* (Q, C) */
  const double* projection;         /* (C, Q) */
  sp_scalar_path density;
  sp_scalar_path preionisation;
  /* Kerr (Problem.Kerr) */
  int has_kerr;
  int kerr_real;                    /* 1: real cubic form, 0: analytic/envelope form */
  double kerr_coefficient, kerr_self, kerr_conjugate;
  /* Raman (Problem.Raman) */
  int n_raman;                      /* rows; 0 for none */
  const double* raman_undamped;     /* (n_raman, 2T) */
  double raman_fraction, raman_epsilon0;
  int raman_thg;
  sp_scalar_path raman_rates[SP_MAX_TERMS];
  /* Plasma (Problem.Plasma), one channel */
  int has_plasma;
  double charge_mass_ratio, ionisation_potential;
  double log_field_min, d_log_field, field_min, field_max, rate_max;
  int rate_intervals;
  const double* rate_coefficients;  /* (4, rate_intervals) highest power first */
  int n_taps;                       /* 1: trapezoid; >1: band rule taps c_0..c_K */
  const double* taps;
  /* Stepper (Problem.Stepper) */
  int n_stages;
  const double* tableau_a;          /* (n_stages, n_stages) */
  const double* tableau_c;          /* (n_stages) */
  const double* tableau_b;          /* (n_stages) */
  const double* tableau_e;          /* (n_stages) b - bhat */
  int first_same_as_last;
  /* Noise: 0 none, 1 input (already in the launch field), 2 modified (per stage) */
  int noise_model;
  const double* noise;              /* complex (C, N) or NULL */
  /* Controller (PIController, src/solprop/stepper/controller.py:262) and landings */
  double rtol, atol, safety, max_factor, min_factor, beta1, beta2;
  int lag, reject_free, estimator_order, auto_init;
  double dz_min, dz_max, stretch;
  double dz_init;                   /* the first proposal, m */
  /* Method: 0 the embedded Runge-Kutta march (envelope route), 1 the exponential Adams
   * march (field route), with the settings of Multistep
   * (src/solprop/stepper/multistep.py:136-214), 2 the low-storage RK4IP replay of a fixed
   * schedule (field route; `schedule_edges` below). */
  int method;
  int adams_order, adams_variable_order, adams_equal_approach;
  double adams_safety, adams_max_growth, adams_min_factor;
  /* Temporal absorber (Problem.Absorber: in_stage, split_step, fine, coarse), the collar
   * as a power absorption rate alpha(t) in 1/m: 0 none; 1 in-stage (-alpha E / 2 on the
   * fine time grid, transformed and added to every source on the active band,
   * passes/stepping.py:286-345); 2 split-step (exp(-alpha h / 2) on the coarse time grid
   * once per accepted step, passes/stepping.py:410-456). */
  int absorber_scheme;
  const double* absorber_rate;      /* (n_to) in-stage, (n_t) split-step */
  /* The radial route's transverse collar (Problem.Spatial.absorber_matrix,
   * native/src/solprop_native/_core.pyi:141): F diag(-alpha(r) / 2) B on the mode axis,
   * (n_modes, n_modes) row-major, 1/m (src/solprop/backends/spatial.py:189-225); its
   * product with the modal time field joins the in-stage absorber's term
   * (passes/stepping.py:286-372). NULL without a collar; dense field route only. */
  const double* spatial_absorber;
  /* Method 2, the low-storage RK4IP replay (Problem.Stepper's low-storage march;
   * native/include/solprop_native/march/low_storage.hpp, src/solprop/backends/numpy_ref/
   * engines.py:236-520): the fixed schedule's edges, `n_schedule` steps (n_schedule + 1
   * increasing positions, m), every step accepted. Field and Cartesian routes; 0 / NULL
   * otherwise. */
  int n_schedule;
  const double* schedule_edges;
  /* Kind 2, the free-space Cartesian route (the reference's `prop_free` geometry,
   * docs/dev/benchmarks.md:529-537): the transverse grid, nx rows of ny points, each
   * point a column of n_t = n_to = n_omega time samples, so n_modes = nx ny. The state
   * is the unnormalised 3-D spectrum Y(kx, ky, w), w fastest; one operator table
   * (n_phi = 1, W(z) = z), one prefactor table, the analytic Kerr response, method 2.
   * 0 otherwise. */
  int grid_nx, grid_ny;
  /* The embedded coarse rule's projection (Problem.Transform.projection_coarse_matrix,
   * _core.pyi:222; src/solprop/modes/quadrature.py:265-332): (C, Q), the Gauss weights of
   * the rule embedded in a Gauss-Kronrod radial rule on the nodes the two share, zero on
   * the others. It serves the quadrature estimate (statistics flag 1 << 10) alone. NULL
   * when the rule has no embedded pair (the estimate is then NaN). */
  const double* projection_coarse;
} sp_problem_desc;

typedef struct sp_workspace sp_workspace;

/* `Workspace(problem)` (_core.pyi:595-607; native/include/.../workspace.hpp:139):
 * validates, allocates every buffer and plans every transform once. */
int sp_workspace_create(const sp_problem_desc* desc, sp_workspace** out, char* msg,
                        size_t msglen);
/* `Workspace.close()`; NULL is accepted. A workspace an open session leases is released
 * when that session closes. */
void sp_workspace_destroy(sp_workspace* ws);

/* Load a committed state at z (the reference's upload at start, trial.cpp:356-390).
 * `field` is complex (C, N). Resets the march and any held trial. */
int sp_workspace_load(sp_workspace* ws, double z, const double* field, char* msg,
                      size_t msglen);

/* `Probe.trial(z, dz) -> float` (_core.pyi:36-49; native probe, workspace.hpp:209-214):
 * one embedded Runge-Kutta trial from the committed state, every stage evaluated,
 * committing nothing; returns the scaled error norm. */
int sp_workspace_trial(sp_workspace* ws, double z, double dz, double* norm, char* msg,
                       size_t msglen);
/* `Probe.field()`: the held trial's end state, complex (C, N). */
int sp_workspace_trial_field(sp_workspace* ws, double* out, char* msg, size_t msglen);
/* The committed state, complex (C, N) (the raw state before the export recipe). */
int sp_workspace_state(sp_workspace* ws, double* out, char* msg, size_t msglen);

/* One march call (`Session.advance(work_limit, deadline)`, _core.pyi:468-499;
 * native/src/session/session.cpp:695): step adaptively towards `target`, at most
 * `work_limit` attempts. Writes the attempt rows (z_end, dz, norm, accepted) into
 * `rows` (4 doubles each, up to `max_rows`), sets `n_rows`, the new position `z` and
 * `done` (1 when the target was reached). SP_CONVERGENCE on the step floor. */
typedef struct sp_march_result {
  double z;
  double proposal;
  int done;
  int n_rows;
  long attempts;
  long accepted;
  long evaluations;
} sp_march_result;
int sp_workspace_advance(sp_workspace* ws, double target, long work_limit, double* rows,
                         long max_rows, sp_march_result* out, char* msg, size_t msglen);

/* The batched variant for independent workspaces of any route (the process pools of
 * src/solprop/driver/scan.py:30): `n` workspaces advanced concurrently to their own
 * targets by a fixed pool of host threads, each on its workspace's own stream, so that
 * the instances' launches and host decisions overlap on the one GPU (a workspace listed
 * twice is refused). A scan of field-grid points runs faster as one sp_scan. Each instance's march is the
 * one sp_workspace_advance would make (the same decisions and results, to the bit).
 * `rows` holds n blocks of max_rows x 4 (end, step, norm, accepted), `out` n results,
 * `status` n codes; the call returns SP_OK when every instance did, else the first
 * failing instance's code, with its message in `msg`. */
int sp_batch_advance(sp_workspace* const* ws, int n, const double* targets, long work_limit,
                     double* rows, long max_rows, sp_march_result* out, int* status, char* msg,
                     size_t msglen);

/* ------------------------------------------------------------ batched scans
 *
 * A parameter scan's points as one batched march (the reference runs them on process
 * pools: src/solprop/driver/scan.py:30, docs/dev/architecture.md:774-786). Every instance
 * is a field-grid problem under the exponential Adams march with its own grid (n_to a
 * power of two), modes (1..6), nodes, operator, plasma rule and launch field, and its own
 * adaptive step control; none may carry a Raman response, an absorber or the dense
 * route. Each round of the scan runs every active instance's next trial (or start) as
 * one captured HIP graph of grouped launches on one stream; the host concludes each
 * instance's attempt and plans its next. Each instance's march is the one
 * sp_workspace_advance makes for its problem alone: the same decisions, rows and end
 * state, to the bit. */
typedef struct sp_scan sp_scan;

/* Create the scan's `n` instances (the `Workspace(problem)` of each point,
 * _core.pyi:595-607) and capture its round. SP_VALUE names the first instance the scan
 * refuses. */
int sp_scan_create(const sp_problem_desc* descs, int n, sp_scan** out, char* msg,
                   size_t msglen);
void sp_scan_destroy(sp_scan* scan);
/* Load instance `i`'s state at z (complex (C_i, N_i)), as sp_workspace_load. */
int sp_scan_load(sp_scan* scan, int i, double z, const double* field, char* msg,
                 size_t msglen);
/* Advance every instance towards its own target, at most `work_limit` attempts each (the
 * reference's per-point `Session.advance(work_limit, ...)`, _core.pyi:468-499). `rows`
 * holds n blocks of max_rows x 4 (end, step, norm, accepted), `out` n results, `status`
 * n codes (an instance's convergence failure stops that instance only); returns SP_OK
 * when every instance did, else the first failing instance's code with its message. */
int sp_scan_advance(sp_scan* scan, const double* targets, long work_limit, double* rows,
                    long max_rows, sp_march_result* out, int* status, char* msg,
                    size_t msglen);
/* Instance `i`'s committed state, complex (C_i, N_i). */
int sp_scan_state(sp_scan* scan, int i, double* out, char* msg, size_t msglen);
/* The scan's rounds so far, the host's wait for their graphs and its planning time,
 * milliseconds, and the device bytes its instances hold. */
typedef struct sp_scan_timing {
  long rounds;
  double device_ms;
  double host_ms;
  long device_bytes;
} sp_scan_timing;
int sp_scan_timing_of(sp_scan* scan, sp_scan_timing* out);
/* The scan's autotuned schedule as JSON (the project's own tuner; the reference has
 * none): for every grouped transform pass its direction, points, jobs, the plan's and the
 * chosen sequences a workgroup, and each candidate's measured microseconds. The choice is
 * a launch shape only: results are the same to the bit whichever is chosen. Timed at
 * sp_scan_create, once per pass shape and process; SOLPROP_MI355X_AUTOTUNE=0 keeps the
 * plans' own choice. SP_VALUE when `len` is too small. */
int sp_scan_schedule(sp_scan* scan, char* out, size_t len);

/* ------------------------------------------------------------------ sessions
 *
 * `Workspace.start(run) -> Session` and the session's event protocol (_core.pyi:468-558;
 * native/include/solprop_native/session/session.hpp; native/src/session/session.cpp:695;
 * docs/dev/kernel-contract.md "What a native call takes / returns"). A session leases its
 * workspace: while it is open the workspace's own march, load and trial entries refuse
 * with SP_STATE. Each advance runs bounded march calls until the caller has work and
 * returns one batch: a sequence number, one event and the chunk of attempt rows made
 * since the previous batch, 6 doubles a row (end, step, norm -- infinity when not
 * finite --, accepted, order, wall seconds since the previous event). A batch is pending
 * until acknowledged; advancing across it is SP_STATE. An observation or terminal batch
 * holds a ticket: (sequence, generation) exports the committed state it names while the
 * batch is pending. */
typedef struct sp_session sp_session;

/* The run's plan (the reference's session::Plan): the save planes, increasing, the last
 * one the run's end; whether the launch state is observed first; every `step_planes`-th
 * accepted step observed (0: none); the plane statistics (sp_workspace_statistics's
 * flags and windows) taken at every observation (0: none); the chunk buffer's attempts
 * (a full buffer yields a progress event) and one march call's bound, 256 attempts and
 * 0.1 s in the reference (0 or less: those defaults). */
typedef struct sp_session_plan {
  const double* planes;
  int n_planes;
  int launch;
  long step_planes;
  unsigned long long statistics;
  const double* windows;
  int n_windows;
  long attempt_capacity;
  long call_attempts;
  double call_seconds;
} sp_session_plan;

/* SessionEvent (_core.pyi:517-558): kind 0 observation, 1 progress, 2 snapshot,
 * 3 terminal, 4 warning; observation 0 launch, 1 plane, 2 step, 3 natural; outcome
 * 0 none, 1 completed, 2 stopped, 3 failed; for a failure the march's reason (0 reached,
 * 1 host step, 3 bound, 5 floor, 6 native error, 11 stopped), its stage (1 startup,
 * 2 trial, 3 accept), category (1 convergence, 2 backend, 3 value), message and, at the
 * floor, the proposal below it. */
typedef struct sp_session_event {
  int kind;
  int observation;
  double z;
  long plane_index;
  int has_stats;
  int ticket;
  unsigned long long generation;
  int outcome;
  int reason;
  int category;
  int stage;
  double floor_proposal;
  char message[256];
} sp_session_event;

/* SessionBatch (_core.pyi:501-515): its sequence, event, the number of attempt rows
 * written to the caller's buffer and of plane statistics written to its stats buffer. */
typedef struct sp_session_batch {
  unsigned long long sequence;
  sp_session_event event;
  long n_attempts;
  int n_stats;
} sp_session_batch;

/* SessionStatus (_core.pyi:560-587): state 0 ready, 1 pending, 2 finished, 3 closed. */
typedef struct sp_session_status {
  double z;
  long plane_index;
  unsigned long long attempts, rejected;
  int state, outcome;
  unsigned long long generation, march_calls, batches, exports, export_bytes;
} sp_session_status;

/* `Workspace.start(run)`: load `field` (complex (C, N)) at z and open a session. */
int sp_session_start(sp_workspace* ws, double z, const double* field,
                     const sp_session_plan* plan, sp_session** out, char* msg, size_t msglen);
/* `Session.advance(work_limit, deadline) -> SessionBatch`: at most `work_limit` attempts
 * (at least one) and `deadline` seconds, both read between attempts. `attempts` holds
 * `capacity` rows of 6 (at least the plan's attempt_capacity), `stats` `stats_capacity`
 * doubles. */
int sp_session_advance(sp_session* s, long work_limit, double deadline, sp_session_batch* out,
                       double* attempts, long capacity, double* stats, int stats_capacity,
                       char* msg, size_t msglen);
/* `Session.acknowledge(sequence)`: release the pending batch. */
int sp_session_acknowledge(sp_session* s, unsigned long long sequence, char* msg,
                           size_t msglen);
/* `Session.request_stop()`: a stop at the next attempt boundary; any thread, no lock. */
void sp_session_request_stop(sp_session* s);
/* `Session.status()`: any thread, no lock. */
int sp_session_status_of(sp_session* s, sp_session_status* out);
/* `Session.export(ticket, recipe, out)` (native/src/solprop_native/session.py:960): the
 * pending batch's state by the recipe of sp_workspace_export. */
int sp_session_export(sp_session* s, unsigned long long sequence, unsigned long long generation,
                      const double* power_weight, int optics, double* out, char* msg,
                      size_t msglen);
/* `Session.fail_outside()`: a failed terminal batch at the committed state (with the
 * attempts the chunk still holds) when none is pending; SP_STATE when a batch is pending
 * or the session has ended. */
int sp_session_fail_outside(sp_session* s, const char* reason, sp_session_batch* out,
                            double* attempts, long capacity, char* msg, size_t msglen);
/* `Session.close()`: ends the lease; NULL is accepted. */
void sp_session_close(sp_session* s);

/* `Session.export(ticket, recipe, out)` (native/src/solprop_native/session.py:960;
 * src/solprop/backends/export.py:52-207): the committed state converted by the recipe,
 * to the bit: first the product with sqrt(w) when `power_weight` (complex-state shape
 * (C, N), float64) is given, then, when `optics` is set, conj(.) times (-1)^k on the
 * storage index k. `out` is complex (C, N). */
int sp_workspace_export(sp_workspace* ws, const double* power_weight, int optics,
                        double* out, char* msg, size_t msglen);

/* Plane statistics of the committed state (`GeneralState::statistics`,
 * native/src/general/statistics.cpp:106-360; the definitions of
 * src/solprop/stats/definitions.py:268-520). `flags` are StatisticsFlag's bits
 * (native/include/solprop_native/problem/problem.hpp:44-60); built: omega0 (1 << 0),
 * energy (1 << 1), energy_window (1 << 2), peak_power (1 << 3), peak_power_window
 * (1 << 4), peak_power_total (1 << 5), fwhm (1 << 6), on the modal power weight of one, and on the field routes electron_density
 * (1 << 7) and peak_ionisation_rate (1 << 8): density(z) times the end-of-window
 * fraction of the node of largest fraction, and that node's largest rate
 * (numpy_ref/statistics.py:502-546), from the plasma column at the committed state. `windows` holds `n_windows` (lo, hi)
 * pairs in rad/s, gated lo <= omega < hi. `out` receives, in this order and for the
 * flags set: omega0 (C; NaN for an empty column), energy (C), energy_window
 * (n_windows x C), peak_power (C), peak_power_window (n_windows x C),
 * peak_power_total (1), electron_density (1),
 * peak_ionisation_rate (1), and for fwhm (1 << 6) the total power trace
 * sum_modes |a_m(t)|^2 (n_t; the host measures the width, as from the native backend's
 * trace); `count` the number written.
 * Any other flag is refused with SP_VALUE. */
int sp_workspace_statistics(sp_workspace* ws, unsigned long long flags, const double* windows,
                            int n_windows, double* out, int capacity, int* count, char* msg,
                            size_t msglen);

/* The statistics with every input the native workspace's `statistics` takes for them
 * (GeneralWorkspace::statistics, native/src/general/statistics.cpp:106-360): the flags
 * and windows above, and for peak_intensity (1 << 9) the on-axis synthesis,
 * `on_axis_rows` (1 or 2) rows of n_modes, with the intensity factor: the largest over
 * the coarse samples of sum_rows |sum_m on_axis[row, m] a_m(t)|^2, times the factor
 * (statistics.cpp:517-603; a_m the modes' analytic signals of the peak powers). On the
 * field routes populations (1 << 11) too: every quadrature node's end-of-window ionised
 * fraction and largest ionisation rate, (n_nodes, 2), one channel
 * (statistics.cpp:636-650), and the plasma integration diagnostics (1 << 12,
 * statistics.cpp:184-186,684-697; numpy_ref/passes/plasma.py:488-552): six maxima over
 * the nodes of the column at the committed state — the exposure increments' largest
 * dip max(-min dS, 0) and rise max dS, the largest |paid - freed| (paid the trapezoid of
 * W (1 - f), freed f_end - f_0) and freed, and the largest |f E| over the first 2K
 * samples and anywhere. And quadrature_error (1 << 10; trial_finish.inc:921-1028): the
 * last source evaluation of the march's last accepted step (its field and position; the
 * low-storage replay's last stage)
 * repeated with the projection and with projection_coarse, each response spectrum
 * weighted by the prefactor times the spectral window, sqrt(sum |w (coarse - fine)|^2 /
 * sum |w fine|^2) over modes and bins; NaN without projection_coarse, before the first
 * accepted step, on the envelope route or when the fine spectrum vanishes. Packed in the
 * native order: ..., electron_density, peak_ionisation_rate, peak_intensity (1),
 * quadrature_error (1), the fwhm trace (n_t), populations (2 n_nodes), the integration
 * diagnostics (6). The plasma statistics on a problem without a plasma column are
 * refused with SP_VALUE. */
typedef struct sp_statistics_request {
  unsigned long long flags;
  const double* windows;
  int n_windows;
  const double* on_axis;
  int on_axis_rows;
  double intensity_factor;
} sp_statistics_request;
int sp_workspace_statistics_request(sp_workspace* ws, const sp_statistics_request* request,
                                    double* out, int capacity, int* count, char* msg,
                                    size_t msglen);

/* The boundary monitor (`_bind_monitor`, src/solprop/monitor.py:440-606; the native
 * workspace's monitor reading, native/src/general/statistics.cpp and
 * statistics_monitor.inc). `weights` holds the monitor's four rows on the coarse time
 * axis, (4, n_t): the low guard, the high guard, the collar and the collar's rate; NULL
 * unbinds. `columns` is a spatial run's column quadrature (n_modes) or NULL. Binding
 * empties the split collar's sums, as loading a state does. Under the split-step
 * absorber every accepted step's finish then adds, from the masked state before the
 * split, what it takes, sum P (1 - e^{-alpha h}), and keeps the energy and the guards'
 * and collar's energies of P e^{-alpha h / 2}. */
int sp_workspace_bind_monitor(sp_workspace* ws, const double* weights, const double* columns,
                              char* msg, size_t msglen);
/* The reading (`BoundaryMonitor.energies` and `split_sums`): out[0..4] the committed
 * state's energy sum_t P(t) and its four weighted energies (weights . P), with
 * P(t) = sum_columns w_c |x_c(t)|^2 on the coarse samples; out[5] the split's removal
 * since the last reading (then emptied); out[6..9] the last accepted step's found
 * energies (total, low guard, high guard, collar). */
int sp_workspace_monitor_reading(sp_workspace* ws, double* out, char* msg, size_t msglen);

/* Counters (`Workspace.counters()`): trials, evaluations, accepted steps, graph
 * replays, device bytes held. */
typedef struct sp_counters {
  long trials;
  long evaluations;
  long accepted;
  long graph_launches;
  long device_bytes;
} sp_counters;
int sp_workspace_counters(sp_workspace* ws, sp_counters* out);

/* Device timing of the last trials: the average duration in microseconds of the
 * whole trial and of each kernel class since the last reset, from HIP events on the
 * workspace's stream. Enabled by `sp_workspace_profile(ws, 1)`. */
int sp_workspace_profile(sp_workspace* ws, int enable);
int sp_workspace_timing(sp_workspace* ws, double* trial_us, long* trials);
/* The same events split by kernel: `*count` phases (at most `capacity`), the average
 * microseconds of each in `us`, their kernel names comma-separated in `names`. A route
 * that times only whole trials reports no phases. */
int sp_workspace_phases(sp_workspace* ws, double* us, int capacity, int* count, char* names,
                        size_t names_len);

/* The source (the stage derivative before any propagator) at a lab-frame field
 * `lab` (complex (C, N)) at z, into `out` (complex (C, N)); when given, the
 * time-domain intermediates: `streams` (3, C, T) real, the projected polarisation,
 * gated current gJ and edge current g'J, and `modal` (C, T) real, the modal fields.
 * The reference's engine-level testing entry points
 * (native/src/bindings/testing.cpp:599-700, `Engines.stage`). Field route only. It
 * leaves the march's history alone but overwrites the last trial's buffers. */
int sp_testing_source(sp_workspace* ws, double z, const double* lab, double* out,
                      double* streams, double* modal, char* msg, size_t msglen);

/* A named device buffer of the field route ("y", "lab", "source", "advanced", "flow",
 * "store", "moments", "states", "carried"; "stiffness", the low-storage replay's last
 * step's indicator h rho, numpy_ref/engines.py:268-276) copied into `out` (at most `count`
 * doubles); `*size` receives the buffer's length in doubles. The reference's
 * `Engines.field` / `Engines.stage` readbacks (native/src/bindings/testing.cpp:601-602). */
int sp_testing_buffer(sp_workspace* ws, const char* name, double* out, long count, long* size,
                      char* msg, size_t msglen);

/* -------------------------------------------------------------------- traces */

/* A scripted Runge-Kutta march (tests/backends/rk_traces.py:63-141 Scenario). */
typedef struct sp_rk_scenario {
  double rtol, atol, safety, max_factor, min_factor, beta1, beta2;
  int lag, reject_free, estimator_order;
  double z0, proposal, dz_min, dz_max, z_end, stretch;
  int auto_init, interval_cap, step_planes;
  int has_h0;
  double h0;
  const double* norms; long n_norms;
  const double* save_planes; int n_planes;
  const double* breakpoints; int n_breakpoints;
  const double* kinks; int n_kinks;
  const double* targets; int n_targets;
  const double* schedule; long n_schedule;
  const long* quiet_keys; const int* quiet_values; int n_quiet;
  int has_quiet_interval;
  double quiet_interval;
  int has_ceiling;
  double ceiling;
  const long* split_keys; const double* split_values; int n_split;
} sp_rk_scenario;

/* Run the library's Runge-Kutta march over a scripted engine and write its decision
 * trace as JSON (the layout of the reference's tests/data/rk_traces.json.gz) into
 * `out`. Returns SP_VALUE if `outlen` is too small. */
int sp_rk_trace(const sp_rk_scenario* scenario, char* out, size_t outlen);

/* A scripted exponential Adams march (tests/backends/adams_traces.py:50-118 Scenario):
 * the method's settings (Multistep, src/solprop/stepper/multistep.py:136-214), the
 * stepper settings, landings and caps, and per attempt three norms, of which the march
 * reads one per order it scores. `max_calls` < 0 is no budget. */
typedef struct sp_adams_scenario {
  int order, variable_order, equal_approach, interval_cap;
  double safety, max_growth, min_factor, collar_rate;
  long max_calls;
  double z0, proposal, dz_min, dz_max, z_end, stretch;
  const double* norms; long n_attempts;     /* (n_attempts, 3) */
  const double* save_planes; int n_planes;
  const double* breakpoints; int n_breakpoints;
  const double* kinks; int n_kinks;
  const double* targets; int n_targets;
  const double* schedule; long n_schedule;
  const long* quiet_keys; const int* quiet_values; int n_quiet;
  int has_quiet_interval;
  double quiet_interval;
} sp_adams_scenario;

/* Run the library's Adams march over a scripted engine and write its decision trace
 * as JSON (the layout of the reference's tests/data/adams_traces.json.gz). */
int sp_adams_trace(const sp_adams_scenario* scenario, char* out, size_t outlen);

/* Testing entry (the counterpart of the reference's `_core._testing`,
 * native/src/bindings/testing.cpp): one batched complex FFT of `n` points over
 * `batch` contiguous rows, unnormalised, sign -1 forward or +1 inverse, through the
 * library's own Stockham plan; host arrays in and out (complex, batch x n). */
/* C (M x N) = A (M x K) . B (K x N), row-major FP64 host arrays, on the matrix cores: the
 * radial route's Hankel-transform GEMM, exposed as the reference's `_testing` module
 * exposes its kernels (native/src/bindings/testing.cpp:732). */
int sp_testing_gemm(long M, long N, long K, const double* A, const double* B, double* C,
                    char* msg, size_t msglen);

int sp_testing_fft(const double* in, double* out, long n, long batch, int sign, char* msg,
                   size_t msglen);

/* A batched complex FFT that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It has no counterpart in the reference: it is sp_testing_fft on device
 * buffers. `in` and `out` are device pointers to `batch` contiguous rows of `n` complex
 * points (they may be equal: in place); `stream` is a hipStream_t (NULL: the default
 * stream). The work is enqueued on `stream` and not waited for. The twiddle tables and
 * the intermediate of transforms above 4096 points live in one static device array per
 * GPU (about 17 MiB, held until the process ends); n <= 65536 and n * batch <= 2^20 are
 * accepted, larger sizes are refused with SP_VALUE. Calls that overlap on one GPU race on
 * that array. A call at a length other than the last call's on that GPU rebuilds the
 * tables first (a copy on `stream`, then a wait for it). Results equal sp_testing_fft's
 * to the bit. */
int sp_fft_noalloc(const double* in, double* out, long n, long batch, int sign, void* stream,
                   char* msg, size_t msglen);

/* ---------------------------------------------------------- learned surrogate
 *
 * A 1-D Fourier neural operator that screens a parameter scan before the solver runs:
 * it maps a point's input spectrum (n bins) and its scan parameters to the exit spectrum
 * (the project's own component; the reference has no counterpart). Trained and run on
 * the device: every dense product (the 1x1 convolutions, the truncated DFT onto the
 * first `modes` bins and its inverse, the head) on the matrix cores in BF16 with FP32
 * accumulation; the per-mode complex mixing, loss, reductions and Adam in FP32. The
 * parameters are one float32 vector in the order of oracle/fno.py's layout; the input
 * channels are the spectrum, the parameters and the grid coordinate s/n. */
typedef struct sp_fno sp_fno;
typedef struct sp_fno_shape {
  int n_params, width, modes, layers, hidden, n;
} sp_fno_shape;
int sp_fno_create(const sp_fno_shape* shape, int max_batch, sp_fno** out, char* msg,
                  size_t msglen);
void sp_fno_destroy(sp_fno* f);
long sp_fno_parameter_count(const sp_fno* f);
int sp_fno_set_parameters(sp_fno* f, const float* flat, char* msg, size_t msglen);
int sp_fno_get_parameters(sp_fno* f, float* flat, char* msg, size_t msglen);
/* out (batch, n) from spectra (batch, n) and params (batch, n_params), host arrays;
 * batch at most max_batch. */
int sp_fno_forward(sp_fno* f, const float* spectra, const float* params, int batch, float* out,
                   char* msg, size_t msglen);
/* The mean squared error against target (batch, n) and its gradient (flat, host). */
int sp_fno_gradient(sp_fno* f, const float* spectra, const float* params, const float* target,
                    int batch, double* loss, float* grad, char* msg, size_t msglen);
/* A device-resident training set of n rows; then `steps` Adam steps (beta 0.9, 0.999,
 * eps 1e-8) over minibatches of its rows, `indices` (steps x batch) drawn by the caller,
 * each step's loss into `losses` and the device time into `device_ms`. */
int sp_fno_load_dataset(sp_fno* f, const float* spectra, const float* params,
                        const float* targets, long n, char* msg, size_t msglen);
int sp_fno_train(sp_fno* f, const int* indices, long steps, int batch, float lr, float* losses,
                 double* device_ms, char* msg, size_t msglen);
int sp_fno_reset_optimizer(sp_fno* f);
/* C (M x N) = A (M x K) . B (K x N) on the BF16 matrix cores, host float32 arrays
 * rounded to BF16 (a lane-layout check of the surrogate's product). */
int sp_testing_gemm_bf16(long M, long N, long K, const float* A, const float* B, float* C,
                         char* msg, size_t msglen);

/* Library identification: the gfx target and build string. */
const char* sp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SOLPROP_MI355X_H */
