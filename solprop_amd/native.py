"""Host layer over the C ABI: the reference's native adapter shape.

Mirrors ``native/src/solprop_native`` -- ``_lowering._lower`` (``_lowering.py:334-460``)
fills the problem record, :class:`Workspace` is ``_core.Workspace``
(``_core.pyi:595-607``), :meth:`Workspace.probe` returns a :class:`Probe`
(``_core.pyi:36-49``, ``_probe.py:33-70``) and :meth:`Workspace.start` a
:class:`Session` whose :meth:`Session.advance` runs until the caller has work and
returns one :class:`SessionBatch` with one :class:`SessionEvent`
(``_core.pyi:468-558``); :meth:`Workspace.march` marches to a target directly. Errors
map onto the reference's classes:
a refusal before allocation is :class:`BackendRefusal`, the step floor is
:class:`ConvergenceError`, everything else :class:`BackendError`.
"""

from __future__ import annotations

import ctypes as C
import math
import weakref
from collections.abc import Mapping
from dataclasses import dataclass, field

import numpy as np

from solprop_amd import _lib
from solprop_amd.problem import Problem, ScalarPath


class BackendError(RuntimeError):
    """A device or runtime failure (``src/solprop/exceptions.py:260``)."""


class BackendRefusal(BackendError):
    """The problem is outside what the library runs; raised before allocation."""


class ConvergenceError(RuntimeError):
    """The controller asked for a step below the floor (``controller.py:34-43``)."""


class LifecycleError(BackendError):
    """A call out of order (``_core.pyi`` SessionLifecycleError)."""


def _raise(status: int, msg: bytes) -> None:
    text = msg.value.decode(errors="replace") if hasattr(msg, "value") else str(msg)
    if status == _lib.SP_VALUE:
        raise BackendRefusal(text)
    if status == _lib.SP_CONVERGENCE:
        raise ConvergenceError(text)
    if status == _lib.SP_STATE:
        raise LifecycleError(text)
    raise BackendError(text)


def _check(status: int, msg) -> None:
    if status != _lib.SP_OK:
        _raise(status, msg)


def _ptr(array: np.ndarray, ctype=C.c_double):
    return array.ctypes.data_as(C.POINTER(ctype))


class _Keep:
    """Holds the arrays a lowered record points into for the record's life."""

    def __init__(self):
        self.arrays = []

    def f64(self, a) -> C.POINTER(C.c_double):
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        self.arrays.append(arr)
        return _ptr(arr)

    def c128(self, a) -> C.POINTER(C.c_double):
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        self.arrays.append(arr)
        return arr.ctypes.data_as(C.POINTER(C.c_double))

    def u8(self, a):
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.uint8))
        self.arrays.append(arr)
        return _ptr(arr, C.c_uint8)

    def i64(self, a):
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
        self.arrays.append(arr)
        return _ptr(arr, C.c_int64)

    def path(self, p: ScalarPath) -> _lib.ScalarPathC:
        out = _lib.ScalarPathC()
        out.kind = p.kind
        out.value = float(p.value)
        if p.kind == 2:
            out.n_intervals = len(p.edges) - 1
            out.degree = p.coefficients.shape[1] - 1
            out.edges = self.f64(p.edges)
            out.coefficients = self.f64(p.coefficients)
        return out


@dataclass
class Settings:
    """The stepper and controller settings a run lowers (``PIController`` defaults)."""

    safety: float = 0.8
    max_factor: float = 1.1
    min_factor: float = 0.2
    beta1: float = 3 / 25
    beta2: float = -1 / 25
    lag: int = 1
    reject_free: bool = False
    estimator_order: int = 4
    auto_init: bool = True
    dz_min: float | None = None
    dz_max: float | None = None
    stretch: float = 1.1
    dz_init: float | None = None
    first_same_as_last: bool = True


def lower(problem: Problem, settings: Settings | None = None) -> tuple[_lib.ProblemDescC, _Keep]:
    """Fill the C record from the problem (``_lowering._lower``, ``_lowering.py:334``)."""
    problem.validate()
    s = settings or Settings()
    k = _Keep()
    d = _lib.ProblemDescC()
    d.kind = {"envelope": 0, "field": 1, "cartesian": 2}[problem.kind]
    if problem.transverse is not None:
        d.grid_nx, d.grid_ny = (int(v) for v in problem.transverse)
    d.n_t, d.n_to = problem.n_t, problem.n_to
    d.n_modes, d.n_nodes, d.n_omega = problem.n_modes, problem.n_nodes, problem.n_omega
    d.crop_index = k.i64(problem.crop_index)
    d.active_mask = k.u8(problem.active_mask)
    d.norm_mask = k.u8(problem.norm_mask)
    d.omega_window = k.f64(problem.omega_window)
    d.omega = k.f64(problem.omega)
    d.time_window = k.f64(problem.time_window)
    d.time_window_derivative = k.f64(problem.time_window_derivative)
    d.dt_fine = problem.dt_fine
    d.z_end = problem.z_end
    if len(problem.phi_paths) > _lib.SP_MAX_TERMS or len(problem.prefactor_paths) > 4:
        raise BackendRefusal("at most four operator and prefactor terms")
    d.n_phi = len(problem.phi_paths)
    d.phi_terms = k.c128(problem.phi_terms)
    for i, p in enumerate(problem.phi_paths):
        d.phi_paths[i] = k.path(p)
    for i, p in enumerate(problem.phi_rate_paths):
        d.phi_rate_paths[i] = k.path(p)
    d.n_prefactor = len(problem.prefactor_paths)
    d.prefactor_terms = k.c128(problem.prefactor_terms)
    for i, p in enumerate(problem.prefactor_paths):
        d.prefactor_paths[i] = k.path(p)
    d.synthesis = k.f64(problem.synthesis)
    d.projection = k.f64(problem.projection)
    if problem.projection_coarse is not None:
        d.projection_coarse = k.f64(problem.projection_coarse)
    d.density = k.path(problem.density)
    d.preionisation = k.path(problem.preionisation)
    if problem.kerr is not None:
        d.has_kerr = 1
        d.kerr_real = 1 if problem.kerr.cubic_form == "real" else 0
        d.kerr_coefficient = problem.kerr.coefficient
        d.kerr_self = problem.kerr.self_factor
        d.kerr_conjugate = problem.kerr.conjugate_factor
    if problem.raman is not None:
        r = problem.raman
        if len(r.rates) > _lib.SP_MAX_TERMS:
            raise BackendRefusal("at most four Raman rows")
        d.n_raman = r.undamped.shape[0]
        d.raman_undamped = k.f64(r.undamped)
        d.raman_fraction = r.fraction
        d.raman_epsilon0 = r.epsilon0
        d.raman_thg = int(r.thg)
        for i, p in enumerate(r.rates):
            d.raman_rates[i] = k.path(p)
    if problem.plasma is not None:
        g = problem.plasma
        d.has_plasma = 1
        d.charge_mass_ratio = g.charge_mass_ratio
        d.ionisation_potential = g.ionisation_potential
        d.log_field_min, d.d_log_field = g.log_field_min, g.d_log_field
        d.field_min, d.field_max, d.rate_max = g.field_min, g.field_max, g.rate_max
        d.rate_intervals = g.rate_coefficients.shape[1]
        d.rate_coefficients = k.f64(g.rate_coefficients)
        d.n_taps = len(g.taps)
        d.taps = k.f64(g.taps)
    tab = problem.tableau
    d.n_stages = len(tab.nodes)
    d.tableau_a = k.f64(tab.matrix)
    d.tableau_c = k.f64(tab.nodes)
    d.tableau_b = k.f64(tab.solution_weights)
    d.tableau_e = k.f64(tab.error_weights)
    d.first_same_as_last = int(tab.first_same_as_last and s.first_same_as_last)
    d.noise_model = {"none": 0, "input": 1, "modified": 2}[problem.noise_model]
    if problem.noise_model == "modified" and problem.noise0 is not None:
        d.noise = k.c128(problem.noise0)
    d.rtol, d.atol = problem.rtol, problem.atol
    d.safety, d.max_factor, d.min_factor = s.safety, s.max_factor, s.min_factor
    d.beta1, d.beta2 = s.beta1, s.beta2
    d.lag, d.reject_free, d.estimator_order = s.lag, int(s.reject_free), s.estimator_order
    d.auto_init = int(s.auto_init)
    d.dz_min = 1e-9 * problem.z_end if s.dz_min is None else s.dz_min
    d.dz_max = problem.z_end if s.dz_max is None else s.dz_max
    d.stretch = s.stretch
    d.dz_init = 1e-4 * problem.z_end if s.dz_init is None else s.dz_init
    d.method = {"rk": 0, "adams": 1, "low_storage": 2}[problem.method]
    if problem.schedule is not None:
        edges = np.ascontiguousarray(problem.schedule, dtype=np.float64)
        d.n_schedule = len(edges) - 1
        d.schedule_edges = k.f64(edges)
    d.adams_order = problem.adams_order
    d.adams_variable_order = int(problem.adams_variable_order)
    d.adams_equal_approach = 1
    d.adams_safety, d.adams_max_growth, d.adams_min_factor = 0.8, 2.0, 0.2
    d.absorber_scheme = {"none": 0, "in_stage": 1, "split_step": 2}[problem.absorber_scheme]
    if problem.absorber_rate is not None:
        d.absorber_rate = k.f64(problem.absorber_rate)
    if problem.spatial_absorber is not None:
        d.spatial_absorber = k.f64(problem.spatial_absorber)
    return d, k


class Workspace:
    """``_core.Workspace``: everything allocated and planned at construction."""

    def __init__(self, problem: Problem, settings: Settings | None = None):
        self._lib = _lib.lib()
        self.problem = problem
        desc, keep = lower(problem, settings)
        self._keep = keep
        handle = C.c_void_p()
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_workspace_create(C.byref(desc), C.byref(handle), msg, 1024), msg)
        self._handle = handle
        self._shape = (problem.n_modes, problem.n_omega)
        self._session = None  # a weak reference to the open session

    @property
    def closed(self) -> bool:
        return self._handle is None

    def close(self) -> None:
        """Close the open session first, if any, then release the workspace."""
        session = self._session() if self._session is not None else None
        if session is not None:
            session.close()
        self._session = None
        if self._handle is not None:
            self._lib.sp_workspace_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        if self._handle is None:
            raise LifecycleError("the workspace is closed")
        msg = C.create_string_buffer(1024)
        _check(getattr(self._lib, name)(self._handle, *args, msg, 1024), msg)

    def load(self, z: float, field: np.ndarray) -> None:
        arr = np.ascontiguousarray(np.asarray(field, dtype=np.complex128).reshape(self._shape))
        self._call("sp_workspace_load", float(z), arr.ctypes.data_as(C.POINTER(C.c_double)))

    def probe(self, z: float, field: np.ndarray) -> "Probe":
        self.load(z, field)
        return Probe(self, z)

    def march(self, z: float, field: np.ndarray) -> "March":
        """Load a state and march it to targets directly (``sp_workspace_advance``)."""
        self.load(z, field)
        return March(self, z)

    def start(self, run: Mapping) -> "Session":
        """``_core.Workspace.start(run) -> Session`` (``_core.pyi:595-607``): a session
        of events over this workspace from ``run``'s launch (see :class:`Session`)."""
        if self._handle is None:
            raise LifecycleError("the workspace is closed")
        session = Session(self, run)
        self._session = weakref.ref(session)
        return session

    def trial(self, z: float, dz: float) -> float:
        norm = C.c_double()
        self._call("sp_workspace_trial", float(z), float(dz), C.byref(norm))
        return norm.value

    def trial_field(self) -> np.ndarray:
        out = np.empty(self._shape, dtype=np.complex128)
        self._call("sp_workspace_trial_field", out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def state(self) -> np.ndarray:
        out = np.empty(self._shape, dtype=np.complex128)
        self._call("sp_workspace_state", out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def export(self, weight: np.ndarray | None = None, optics: bool = True) -> np.ndarray:
        out = np.empty(self._shape, dtype=np.complex128)
        w = None
        if weight is not None:
            w = np.ascontiguousarray(np.broadcast_to(weight, self._shape), dtype=np.float64)
        self._call(
            "sp_workspace_export",
            None if w is None else _ptr(w),
            int(optics),
            out.ctypes.data_as(C.POINTER(C.c_double)),
        )
        return out

    # StatisticsFlag's bits (native/include/solprop_native/problem/problem.hpp:44-60) for
    # the statistics this library forms, in their packing order.
    STATISTICS = {"omega0": 1 << 0, "energy": 1 << 1, "energy_window": 1 << 2,
                  "peak_power": 1 << 3, "peak_power_window": 1 << 4,
                  "peak_power_total": 1 << 5,
                  "fwhm": 1 << 6, "electron_density": 1 << 7,
                  "peak_ionisation_rate": 1 << 8, "peak_intensity": 1 << 9,
                  "quadrature_error": 1 << 10, "populations": 1 << 11,
                  "plasma_integration": 1 << 12}

    def statistics(self, names, windows=None, on_axis=None, intensity_factor=1.0) -> dict:
        """Plane statistics of the committed state (``stats/definitions.py``): a dict of
        the requested names. ``windows`` is an (n, 2) array of (lo, hi) in rad/s for
        ``energy_window``, whose value is (n, n_modes). ``peak_intensity`` reads the
        on-axis synthesis ``on_axis`` ((rows, n_modes), one or two rows) and the
        intensity factor; ``populations`` is (n_nodes, 2): each node's end-of-window
        ionised fraction and largest rate; ``plasma_integration`` the column's six
        integration-diagnostic maxima (:func:`plasma_integration_statistics` forms the
        three diagnostics of them); ``quadrature_error`` the last accepted step's source
        evaluation repeated with the embedded coarse projection, their relative
        difference (NaN without an embedded rule or before the first accepted step)."""
        unknown = set(names) - set(self.STATISTICS)
        if unknown:
            raise BackendRefusal(f"statistics not built: {sorted(unknown)}")
        flags = 0
        for name in names:
            flags |= self.STATISTICS[name]
        w = np.zeros((0, 2)) if windows is None else np.ascontiguousarray(windows, np.float64)
        n_modes, n_nodes = self.problem.n_modes, self.problem.n_nodes
        capacity = 3 * n_modes + 2 * len(w) * n_modes + 5 + self.problem.n_t + 2 * n_nodes + 6
        out = np.zeros(capacity)
        count = C.c_int()
        axis = None
        if on_axis is not None:
            axis = np.ascontiguousarray(np.atleast_2d(on_axis), np.float64)
        request = _lib.StatisticsRequestC(
            flags, _ptr(w) if len(w) else None, len(w), None if axis is None else _ptr(axis),
            0 if axis is None else axis.shape[0], float(intensity_factor))
        self._call("sp_workspace_statistics_request", C.byref(request), _ptr(out), capacity,
                   C.byref(count))
        values, at, result = out[: count.value], 0, {}
        sizes = {"omega0": n_modes, "energy": n_modes, "energy_window": len(w) * n_modes,
                 "peak_power": n_modes, "peak_power_window": len(w) * n_modes,
                 "peak_power_total": 1, "electron_density": 1,
                 "peak_ionisation_rate": 1, "peak_intensity": 1, "quadrature_error": 1,
                 "fwhm": self.problem.n_t,
                 "populations": 2 * n_nodes, "plasma_integration": 6}
        # The packing order (native/src/general/statistics.cpp:160-186): the flags'
        # order, except that the power trace (the fwhm flag's) follows the quadrature
        # estimate, and the populations and then the integration diagnostics come last.
        order = ["omega0", "energy", "energy_window", "peak_power", "peak_power_window",
                 "peak_power_total", "electron_density", "peak_ionisation_rate",
                 "peak_intensity", "quadrature_error", "fwhm", "populations",
                 "plasma_integration"]
        scalars = ("peak_power_total", "electron_density", "peak_ionisation_rate",
                   "peak_intensity", "quadrature_error")
        for name in order:
            if name in names:
                chunk = values[at: at + sizes[name]]
                at += sizes[name]
                windowed = name in ("energy_window", "peak_power_window")
                result[name] = (chunk.reshape(len(w), n_modes) if windowed
                                else float(chunk[0]) if name in scalars
                                else chunk.reshape(n_nodes, 2) if name == "populations"
                                else chunk)
        return result

    def bind_monitor(self, weights: np.ndarray | None, columns: np.ndarray | None = None):
        """Bind the boundary monitor's (4, n_t) weights (low guard, high guard, collar,
        rate) on the coarse axis, or unbind with None (``_bind_monitor``)."""
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and w.shape != (4, self.problem.n_t):
            raise ValueError(f"monitor weights: expected (4, {self.problem.n_t})")
        col = None if columns is None else np.ascontiguousarray(columns, dtype=np.float64)
        self._call("sp_workspace_bind_monitor", None if w is None else _ptr(w),
                   None if col is None else _ptr(col))
        self._monitor_keep = (w, col)

    def monitor_reading(self) -> dict:
        """The state's five energies, the split's removal since the last reading and
        the last step's found energies (``BoundaryEnergies``)."""
        out = np.zeros(10)
        self._call("sp_workspace_monitor_reading", _ptr(out))
        return {"total": out[0], "weighted": out[1:5].copy(), "taken": out[5],
                "found": out[6:10].copy()}

    def counters(self) -> dict:
        c = _lib.CountersC()
        self._lib.sp_workspace_counters(self._handle, C.byref(c))
        return {name: getattr(c, name) for name, _ in _lib.CountersC._fields_}

    def profile(self, enable: bool = True) -> None:
        self._lib.sp_workspace_profile(self._handle, int(enable))

    def timing(self) -> tuple[float, int]:
        us = C.c_double()
        n = C.c_long()
        self._lib.sp_workspace_timing(self._handle, C.byref(us), C.byref(n))
        return us.value, n.value

    def source(self, z: float, lab: np.ndarray, intermediates: bool = False):
        """The source at a lab-frame field (the engines' stage evaluation); with
        `intermediates`, also the (3, C, T) time-domain streams and (C, T) modal fields."""
        lab = np.ascontiguousarray(lab, dtype=np.complex128).reshape(self._shape)
        out = np.zeros(self._shape, dtype=np.complex128)
        C_, T = self.problem.n_modes, self.problem.n_to
        streams = np.zeros((3, C_, T)) if intermediates else None
        modal = np.zeros((C_, T)) if intermediates else None
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_testing_source(
            self._handle, float(z), _ptr(lab.view(np.float64)), _ptr(out.view(np.float64)),
            None if streams is None else _ptr(streams),
            None if modal is None else _ptr(modal), msg, 1024), msg)
        return (out, streams, modal) if intermediates else out

    def buffer(self, name: str) -> np.ndarray:
        """A named device buffer of the field route: complex rows of (C, N); real (Q, T)
        for the node arrays (``node_field``, the dense route's synthesised field;
        ``rate``, ``frac``, ``free_local``, ``x_local``); real (Q, tiles) for the scan
        tiles' offsets (``tile_x``, ``tile_f``)."""
        size = C.c_long()
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_testing_buffer(self._handle, name.encode(), None, 0,
                                           C.byref(size), msg, 1024), msg)
        out = np.zeros(size.value)
        _check(self._lib.sp_testing_buffer(self._handle, name.encode(), _ptr(out), size.value,
                                           C.byref(size), msg, 1024), msg)
        if name in ("node_field", "rate", "frac", "free_local", "x_local"):
            return out.reshape(self.problem.n_nodes, self.problem.n_to)
        if name in ("tile_x", "tile_f"):
            return out.reshape(self.problem.n_nodes, -1)
        if name == "stiffness":
            return out
        return out.view(np.complex128).reshape(-1, *self._shape)

    def phases(self) -> dict:
        """Average device microseconds of each kernel of a trial, in launch order."""
        us = np.zeros(32)
        count = C.c_int()
        names = C.create_string_buffer(2048)
        _check(self._lib.sp_workspace_phases(self._handle, _ptr(us), len(us), C.byref(count),
                                             names, 2048), names)
        labels = names.value.decode().split(",") if count.value else []
        return dict(zip(labels, us[: count.value].tolist(), strict=True))


def plasma_integration_statistics(parts) -> dict:
    """The plasma column's three integration diagnostics from the six maxima the
    ``plasma_integration`` statistic packs (``numpy_ref/passes/plasma.py:555-575``,
    ``numpy_ref/statistics.py:414-455``): the undershoot of the exposure increments
    ``max(-dS) / max(dS)``, the work defect ``max |paid - freed| / max freed`` and the
    boundary level of the free-current integrand, each zero where its denominator is."""
    dip, rise, defect, freed, early, peak = (float(part) for part in parts)
    return {"plasma_undershoot": dip / rise if rise > 0.0 else 0.0,
            "plasma_work_defect": defect / freed if freed > 0.0 else 0.0,
            "plasma_boundary_level": early / peak if peak > 0.0 else 0.0}


def advance_batch(sessions, targets, work_limit: int = 256) -> list:
    """Advance independent sessions concurrently (``sp_batch_advance``): a fixed pool of
    host threads, each workspace on its own stream. Each session's march is the one
    :meth:`Session.advance` would make. Returns each session's :class:`Batch`."""
    n = len(sessions)
    targets = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, float), (n,)))
    handles = (C.c_void_p * n)(*[s._ws._handle for s in sessions])
    rows = np.zeros((n, work_limit, 4))
    results = (_lib.MarchResultC * n)()
    status = (C.c_int * n)()
    msg = C.create_string_buffer(1024)
    lib = sessions[0]._ws._lib
    _check(lib.sp_batch_advance(handles, n, _ptr(targets), int(work_limit), _ptr(rows),
                                int(work_limit), C.cast(results, C.c_void_p), status, msg,
                                1024), msg)
    out = []
    for i, session in enumerate(sessions):
        r = results[i]
        session.z = r.z
        out.append(Batch(r.z, bool(r.done), rows[i, : r.n_rows].copy(), r.evaluations))
    return out


class Scan:
    """A parameter scan's points as one batched march (``sp_scan_*``).

    The reference runs a scan's points on a process pool, one ``Session`` each
    (``src/solprop/driver/scan.py:30``); here all points advance together, every round
    of trials one captured graph of grouped launches on the GPU, each point keeping its
    own adaptive march. ``advance`` returns one :class:`Batch` per point, each the one
    that point's own :meth:`Session.advance` would return.
    """

    def __init__(self, problems, settings: Settings | None = None):
        self._lib = _lib.lib()
        self._keeps = []
        descs = (_lib.ProblemDescC * len(problems))()
        for i, p in enumerate(problems):
            p.validate()
            desc, keep = lower(p, settings)
            descs[i] = desc
            self._keeps.append(keep)
        self.problems = list(problems)
        handle = C.c_void_p()
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_scan_create(descs, len(problems), C.byref(handle), msg, 1024), msg)
        self._handle = handle
        self.z = np.zeros(len(problems))

    def __len__(self) -> int:
        return len(self.problems)

    def close(self) -> None:
        if self._handle:
            self._lib.sp_scan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, i: int, z: float, field: np.ndarray) -> None:
        field = np.ascontiguousarray(field, dtype=np.complex128)
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_scan_load(self._handle, int(i), float(z), _ptr(field), msg, 1024),
               msg)
        self.z[i] = z

    def start(self, z: float = 0.0) -> None:
        """Load every point's own launch field at z."""
        for i, p in enumerate(self.problems):
            self.load(i, z, p.field0)

    def advance(self, targets, work_limit: int = 256) -> list:
        n = len(self.problems)
        targets = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, float), (n,)))
        rows = np.zeros((n, work_limit, 4))
        results = (_lib.MarchResultC * n)()
        status = (C.c_int * n)()
        msg = C.create_string_buffer(1024)
        code = self._lib.sp_scan_advance(self._handle, _ptr(targets), int(work_limit), _ptr(rows),
                                         int(work_limit), C.cast(results, C.c_void_p), status,
                                         msg, 1024)
        out = []
        for i in range(n):
            r = results[i]
            self.z[i] = r.z
            out.append(Batch(r.z, bool(r.done), rows[i, : r.n_rows].copy(), r.evaluations))
        self.status = [int(v) for v in status]
        _check(code, msg)
        return out

    def state(self, i: int) -> np.ndarray:
        p = self.problems[i]
        out = np.zeros((p.n_modes, p.n_omega), dtype=np.complex128)
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_scan_state(self._handle, int(i), _ptr(out), msg, 1024), msg)
        return out

    def schedule(self) -> list:
        """The autotuned launch schedule of the grouped transform passes."""
        import json

        buf = C.create_string_buffer(1 << 16)
        _check(self._lib.sp_scan_schedule(self._handle, buf, len(buf)), b"schedule too long")
        return json.loads(buf.value.decode())

    def timing(self) -> dict:
        t = _lib.ScanTimingC()
        _check(self._lib.sp_scan_timing_of(self._handle, C.byref(t)), b"")
        return {"rounds": t.rounds, "device_ms": t.device_ms, "host_ms": t.host_ms,
                "device_bytes": t.device_bytes}


SessionLifecycleError = LifecycleError

# SessionEvent's kinds, observations and outcomes (native/include/solprop_native/session/
# session.hpp: EventKind, Observation, Outcome).
EVENT_KINDS = ("observation", "progress", "snapshot", "terminal", "warning")
OBSERVATIONS = ("launch", "plane", "step", "natural")
OUTCOMES = ("none", "completed", "stopped", "failed")
STATES = ("ready", "pending", "finished", "closed")


@dataclass
class SessionEvent:
    """``_core.SessionEvent`` (``_core.pyi:517-558``)."""

    kind: int
    observation: int
    z: float
    plane_index: int
    has_stats: bool
    ticket: bool
    generation: int
    outcome: int
    reason: int
    category: int
    stage: int
    message: str
    floor_proposal: float

    @property
    def kind_name(self) -> str:
        return EVENT_KINDS[self.kind]

    @property
    def observation_name(self) -> str:
        return OBSERVATIONS[self.observation]

    @property
    def outcome_name(self) -> str:
        return OUTCOMES[self.outcome]


@dataclass
class SessionBatch:
    """``_core.SessionBatch`` (``_core.pyi:501-515``): one event and the chunk of attempt
    rows (end, step, norm, accepted, order, wall) made since the previous batch."""

    sequence: int
    event: SessionEvent
    attempts: np.ndarray
    plane_stats: np.ndarray
    row_z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    row_step: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    row_values: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def ticket(self) -> tuple[int, int]:
        return (self.sequence, self.event.generation)


@dataclass
class SessionStatus:
    """``_core.SessionStatus`` (``_core.pyi:560-587``)."""

    z: float
    plane_index: int
    attempts: int
    rejected: int
    state: int
    outcome: int
    generation: int
    march_calls: int
    rows_taken: int
    batches: int
    exports: int
    export_bytes: int
    buffer_bytes: int


class Session:
    """``_core.Session`` (``_core.pyi:468-499``): one trajectory's march as events.

    ``run`` holds the launch (``z``, ``field``), the plan (``planes``: save planes,
    increasing, the last the run's end; ``launch``: observe the launch first;
    ``step_planes``: observe every k-th accepted step; ``statistics``: plane statistics
    names taken at each observation, with ``windows``) and the buffers' and calls' bounds
    (``attempt_capacity``, ``call_attempts``, ``call_seconds``). The session leases its
    workspace until :meth:`close`.
    """

    def __init__(self, workspace: Workspace, run: Mapping):
        self._ws = workspace
        self._lib = workspace._lib
        problem = workspace.problem
        z = float(run.get("z", 0.0))
        launch = run.get("field", problem.field0)
        launch = np.ascontiguousarray(np.asarray(launch, dtype=np.complex128)
                                      .reshape(workspace._shape))
        planes = np.ascontiguousarray(run.get("planes", [problem.z_end]), dtype=np.float64)
        names = list(run.get("statistics", ()))
        unknown = set(names) - set(Workspace.STATISTICS)
        if unknown:
            raise BackendRefusal(f"statistics not built: {sorted(unknown)}")
        if "peak_intensity" in names:
            raise BackendRefusal("a session's plan carries no on-axis synthesis: take "
                                 "peak_intensity with Workspace.statistics")
        flags = 0
        for name in names:
            flags |= Workspace.STATISTICS[name]
        windows = run.get("windows")
        windows = (np.zeros((0, 2)) if windows is None
                   else np.ascontiguousarray(windows, dtype=np.float64).reshape(-1, 2))
        self._capacity = int(run.get("attempt_capacity", 4096))
        plan = _lib.SessionPlanC()
        plan.planes = _ptr(planes)
        plan.n_planes = len(planes)
        plan.launch = int(bool(run.get("launch", True)))
        plan.step_planes = int(run.get("step_planes", 0))
        plan.statistics = flags
        plan.windows = _ptr(windows)
        plan.n_windows = len(windows)
        plan.attempt_capacity = self._capacity
        plan.call_attempts = int(run.get("call_attempts", 256))
        plan.call_seconds = float(run.get("call_seconds", 0.1))
        self._keep = (planes, windows, launch)
        self._names = names
        self._n_windows = len(windows)
        handle = C.c_void_p()
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_session_start(workspace._handle, z, _ptr(launch), C.byref(plan),
                                          C.byref(handle), msg, 1024), msg)
        self._handle = handle
        self._rows = np.zeros((self._capacity, 6))
        # As many doubles as Workspace.statistics packs at most.
        n_modes = problem.n_modes
        self._stats = np.zeros(3 * n_modes + 2 * len(windows) * n_modes + 4 + problem.n_t
                               + 2 * problem.n_nodes)

    def _call(self, name, *args):
        if self._handle is None:
            raise LifecycleError("the session is closed")
        msg = C.create_string_buffer(1024)
        _check(getattr(self._lib, name)(self._handle, *args, msg, 1024), msg)

    def _batch(self, b) -> SessionBatch:
        e = b.event
        event = SessionEvent(e.kind, e.observation, e.z, e.plane_index, bool(e.has_stats),
                             bool(e.ticket), int(e.generation), e.outcome, e.reason,
                             e.category, e.stage, e.message.decode(errors="replace"),
                             e.floor_proposal)
        return SessionBatch(int(b.sequence), event, self._rows[: b.n_attempts].copy(),
                            self._stats[: b.n_stats].copy())

    def advance(self, work_limit: int = 256, deadline: float = 0.1) -> SessionBatch:
        b = _lib.SessionBatchC()
        self._call("sp_session_advance", int(work_limit), float(deadline), C.byref(b),
                   _ptr(self._rows), self._capacity, _ptr(self._stats), len(self._stats))
        return self._batch(b)

    def acknowledge(self, sequence: int) -> None:
        self._call("sp_session_acknowledge", int(sequence))

    def request_stop(self) -> None:
        if self._handle is not None:
            self._lib.sp_session_request_stop(self._handle)

    def status(self) -> SessionStatus:
        if self._handle is None:
            raise LifecycleError("the session is closed")
        s = _lib.SessionStatusC()
        _check(self._lib.sp_session_status_of(self._handle, C.byref(s)), b"")
        return SessionStatus(s.z, s.plane_index, s.attempts, s.rejected, s.state, s.outcome,
                             s.generation, s.march_calls, 0, s.batches, s.exports,
                             s.export_bytes, self._rows.nbytes)

    def export(self, ticket: tuple[int, int], recipe: Mapping, out: np.ndarray) -> None:
        """The ticket's state into ``out`` (complex (C, N), C-contiguous, writable) by
        ``recipe``: ``power_weight`` (or None) then ``optics`` (``apply_recipe``)."""
        if out.dtype != np.complex128 or out.shape != self._ws._shape or \
                not out.flags.c_contiguous or not out.flags.writeable:
            raise BackendRefusal("export needs a writable C-contiguous complex128 "
                                 f"array of shape {self._ws._shape}")
        weight = recipe.get("power_weight")
        if weight is not None:
            weight = np.ascontiguousarray(np.broadcast_to(weight, self._ws._shape),
                                          dtype=np.float64)
        sequence, generation = ticket
        self._call("sp_session_export", int(sequence), int(generation),
                   None if weight is None else _ptr(weight), int(bool(recipe.get("optics", True))),
                   out.ctypes.data_as(C.POINTER(C.c_double)))

    def fail_outside(self) -> SessionBatch | None:
        if self._handle is None:
            return None
        b = _lib.SessionBatchC()
        msg = C.create_string_buffer(1024)
        status = self._lib.sp_session_fail_outside(self._handle, b"failed outside the session",
                                                   C.byref(b), _ptr(self._rows), self._capacity,
                                                   msg, 1024)
        if status == _lib.SP_STATE:
            return None
        _check(status, msg)
        return self._batch(b)

    def close(self) -> None:
        if self._handle is not None:
            self._lib.sp_session_close(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Probe:
    """One trial from a loaded state, committing nothing (``_probe.py:33-70``)."""

    def __init__(self, workspace: Workspace, z: float):
        self._ws = workspace
        self.z = z
        self.trials = 0

    def trial(self, z: float, dz: float) -> float:
        self.trials += 1
        return self._ws.trial(z, dz)

    def field(self) -> np.ndarray:
        return self._ws.trial_field()


@dataclass
class Batch:
    """What one ``advance`` returns: the attempts rows and where the march stands."""

    z: float
    done: bool
    attempts: np.ndarray  # (n, 4): end, step, norm, accepted
    evaluations: int


class March:
    """A march over the workspace to a target z, attempt rows returned per call (the
    native march call under the reference's session, without its events)."""

    def __init__(self, workspace: Workspace, z: float):
        self._ws = workspace
        self.z = z
        self.finished = False

    def advance(self, target: float, work_limit: int = 256) -> Batch:
        rows = np.zeros((work_limit, 4))
        res = _lib.MarchResultC()
        msg = C.create_string_buffer(1024)
        status = self._ws._lib.sp_workspace_advance(
            self._ws._handle, float(target), int(work_limit), _ptr(rows), int(work_limit),
            C.byref(res), msg, 1024,
        )
        _check(status, msg)
        self.z = res.z
        return Batch(res.z, bool(res.done), rows[: res.n_rows].copy(), res.evaluations)

    def run(self, target: float, work_limit: int = 256) -> np.ndarray:
        """Advance until the target; return every attempt row."""
        out = []
        while True:
            batch = self.advance(target, work_limit)
            out.append(batch.attempts)
            if batch.done:
                break
        self.finished = True
        return np.concatenate(out) if out else np.zeros((0, 4))

    def state(self) -> np.ndarray:
        return self._ws.state()

    def export(self, weight=None, optics=True) -> np.ndarray:
        return self._ws.export(weight, optics)


def testing_gemm(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``a @ b`` in FP64 on the matrix cores (the radial route's Hankel GEMM)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    (m, k), (k2, n) = a.shape, b.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    out = np.zeros((m, n))
    msg = C.create_string_buffer(1024)
    _check(_lib.lib().sp_testing_gemm(m, n, k, _ptr(a), _ptr(b), _ptr(out), msg, 1024), msg)
    return out


def testing_fft(x: np.ndarray, sign: int) -> np.ndarray:
    """The library's batched complex FFT over the last axis (unnormalised)."""
    arr = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    n = arr.shape[-1]
    batch = arr.size // n
    out = np.empty_like(arr)
    msg = C.create_string_buffer(1024)
    status = _lib.lib().sp_testing_fft(
        arr.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)),
        n, batch, int(sign), msg, 1024,
    )
    _check(status, msg)
    return out


# ---------------------------------------------------------------------- traces


def rk_trace(scenario: dict) -> dict:
    """Run a scripted scenario through the library's march; return its trace dict."""
    import json

    lib = _lib.lib()
    keep = _Keep()
    s = _lib.RKScenarioC()
    c = scenario["controller"]
    s.rtol = c.get("rtol", 1e-6)
    s.atol = c.get("atol", 0.0)
    s.safety = c.get("safety", 0.8)
    s.max_factor = c.get("max_factor", 1.1)
    s.min_factor = c.get("min_factor", 0.2)
    s.beta1 = c.get("beta1", 3 / 25)
    s.beta2 = c.get("beta2", -1 / 25)
    s.lag = c.get("lag", 1)
    s.reject_free = int(c.get("reject_free", False))
    s.estimator_order = c.get("estimator_order", 4)
    for name in ("z0", "proposal", "dz_min", "dz_max", "z_end", "stretch"):
        setattr(s, name, float(scenario[name]))
    s.auto_init = int(scenario["auto_init"])
    s.interval_cap = int(scenario["interval_cap"])
    s.step_planes = int(scenario["step_planes"])
    s.has_h0 = int(scenario["h0"] is not None)
    s.h0 = 0.0 if scenario["h0"] is None else float(scenario["h0"])
    s.norms = keep.f64(scenario["norms"])
    s.n_norms = len(scenario["norms"])
    s.save_planes = keep.f64(scenario["save_planes"] or [0.0])
    s.n_planes = len(scenario["save_planes"])
    s.breakpoints = keep.f64(scenario["breakpoints"] or [0.0])
    s.n_breakpoints = len(scenario["breakpoints"])
    s.kinks = keep.f64(scenario["kinks"] or [0.0])
    s.n_kinks = len(scenario["kinks"])
    targets = scenario["targets"]
    s.targets = keep.f64(targets or [0.0])
    s.n_targets = 0 if targets is None else len(targets)
    schedule = scenario["schedule"]
    s.schedule = keep.f64(schedule or [0.0])
    s.n_schedule = 0 if schedule is None else len(schedule)
    quiet = {int(a): bool(b) for a, b in scenario["quiet"].items()}
    qk = np.ascontiguousarray(np.array(list(quiet) or [0], dtype=np.int64))
    qv = np.ascontiguousarray(np.array([int(v) for v in quiet.values()] or [0], dtype=np.int32))
    keep.arrays += [qk, qv]
    s.quiet_keys = qk.ctypes.data_as(C.POINTER(C.c_long))
    s.quiet_values = qv.ctypes.data_as(C.POINTER(C.c_int))
    s.n_quiet = len(quiet)
    s.has_quiet_interval = int(scenario["quiet_interval"] is not None)
    s.quiet_interval = float(scenario["quiet_interval"] or 0.0)
    s.has_ceiling = int(bool(scenario["ceiling"]))
    s.ceiling = float(scenario["ceiling"] or 0.0)
    split = {int(a): float(b) for a, b in scenario["split"].items()}
    sk = np.ascontiguousarray(np.array(list(split) or [0], dtype=np.int64))
    keep.arrays.append(sk)
    s.split_keys = sk.ctypes.data_as(C.POINTER(C.c_long))
    s.split_values = keep.f64(list(split.values()) or [0.0])
    s.n_split = len(split)
    size = 1 << 22
    buf = C.create_string_buffer(size)
    status = lib.sp_rk_trace(C.byref(s), buf, size)
    if status != _lib.SP_OK:
        raise BackendError(buf.value.decode(errors="replace"))
    return json.loads(buf.value.decode())


def adams_trace(scenario: dict) -> dict:
    """Run a scripted scenario through the library's Adams march; return its trace."""
    import json

    lib = _lib.lib()
    keep = _Keep()
    s = _lib.AdamsScenarioC()
    m = scenario["method"]
    s.order = m.get("order", 5)
    s.variable_order = int(m.get("variable_order", True))
    s.equal_approach = int(m.get("equal_approach", True))
    s.interval_cap = int(m.get("interval_cap", False))
    s.safety = m.get("safety", 0.8)
    s.max_growth = m.get("max_growth", 2.0)
    s.min_factor = m.get("min_factor", 0.2)
    s.collar_rate = float(scenario["collar_rate"])
    s.max_calls = -1 if m.get("max_calls") is None else int(m["max_calls"])
    for name in ("z0", "proposal", "dz_min", "dz_max", "z_end", "stretch"):
        setattr(s, name, float(scenario[name]))
    norms = np.zeros((max(len(scenario["norms"]), 1), 3))
    for i, row in enumerate(scenario["norms"]):
        norms[i, : len(row)] = row
    s.norms = keep.f64(norms)
    s.n_attempts = len(scenario["norms"])
    planes = [] if scenario["interpolate"] else scenario["save_planes"]
    s.save_planes = keep.f64(planes or [0.0])
    s.n_planes = len(planes)
    s.breakpoints = keep.f64(scenario["breakpoints"] or [0.0])
    s.n_breakpoints = len(scenario["breakpoints"])
    s.kinks = keep.f64(scenario["kinks"] or [0.0])
    s.n_kinks = len(scenario["kinks"])
    targets = scenario["targets"]
    s.targets = keep.f64(targets or [0.0])
    s.n_targets = 0 if targets is None else len(targets)
    schedule = scenario["schedule"]
    s.schedule = keep.f64(schedule or [0.0])
    s.n_schedule = 0 if schedule is None else len(schedule)
    quiet = {int(a): bool(b) for a, b in scenario["quiet"].items()}
    qk = np.ascontiguousarray(np.array(list(quiet) or [0], dtype=np.int64))
    qv = np.ascontiguousarray(np.array([int(v) for v in quiet.values()] or [0], dtype=np.int32))
    keep.arrays += [qk, qv]
    s.quiet_keys = qk.ctypes.data_as(C.POINTER(C.c_long))
    s.quiet_values = qv.ctypes.data_as(C.POINTER(C.c_int))
    s.n_quiet = len(quiet)
    s.has_quiet_interval = int(scenario["quiet_interval"] is not None)
    s.quiet_interval = float(scenario["quiet_interval"] or 0.0)
    size = 1 << 23
    buf = C.create_string_buffer(size)
    status = lib.sp_adams_trace(C.byref(s), buf, size)
    if status != _lib.SP_OK:
        raise BackendError(buf.value.decode(errors="replace"))
    return json.loads(buf.value.decode())


__all__ = [
    "BackendError", "BackendRefusal", "Batch", "ConvergenceError", "LifecycleError",
    "March", "Probe", "Session", "SessionBatch", "SessionEvent", "SessionLifecycleError",
    "SessionStatus", "Settings", "Workspace", "adams_trace", "lower", "rk_trace",
]
_ = math
