"""The problem record that crosses into the device library, and its builders.

The record mirrors the subset of solprop-native's ``_core.Problem``
(``native/src/solprop_native/_core.pyi:51-466``) that the envelope-GNLSE and the modal
field-grid routes read: dims, grid, stepper tableau, linear terms, transform matrices,
Kerr, Raman and plasma groups. Every array is C-contiguous with a fixed dtype, as the
reference's lowering makes them (``native/src/solprop_native/_lowering.py:334-460``).

The z-dependence of a stage is carried, as in the reference, by host scalars against
arrays uploaded once (``docs/dev/backends.md:144-163``): ``ΔΦ_i = Σ_t w_{i,t} Φ_t`` with
``w_{i,t} = W_t(z_i) - W_t(z_anchor)`` from cumulative scalar paths, and the prefactor,
density, pre-ionisation and Raman rates read from scalar paths at the stage position.
A scalar path is a constant, the position itself, or a piecewise Chebyshev series
evaluated by Clenshaw's recurrence (``src/solprop/translate/stage.py:170``), so the
library evaluates exactly the same floats as the host.

The builders at the bottom make the synthetic inputs of the benchmarked workloads
(solprop's translation layer, which would make them from materials and mode solvers, is
outside this package): an envelope GNLSE (sech soliton, Taylor dispersion, damped
oscillator Raman, shot noise), and a gas-filled capillary on a real field grid with
modal synthesis/projection over a radial quadrature, Kerr and tunnelling ionisation.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
EPS0 = 8.8541878188e-12
HBAR = 1.054571817e-34
E_CHARGE = 1.602176634e-19
M_E = 9.1093837139e-31
KB = 1.380649e-23


# ------------------------------------------------------------------- scalar paths


@dataclass(frozen=True)
class ScalarPath:
    """A scalar function of z: constant, the position, or piecewise Chebyshev.

    ``kind`` is 0 (``value`` everywhere), 1 (``z`` itself, exact) or 2 (piecewise
    Chebyshev: ``edges`` of shape (n+1,), ``coefficients`` of shape (n, degree+1),
    interval ``j`` mapping ``[edges[j], edges[j+1]]`` onto ``[-1, 1]``).
    """

    kind: int = 0
    value: float = 0.0
    edges: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coefficients: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))

    @staticmethod
    def constant(value: float) -> "ScalarPath":
        return ScalarPath(0, float(value))

    @staticmethod
    def position() -> "ScalarPath":
        return ScalarPath(1, 0.0)

    @staticmethod
    def chebyshev(function, z0: float, z1: float, intervals: int = 8, degree: int = 24,
                  edges=None):
        """Interpolate ``function`` on ``intervals`` equal pieces of ``[z0, z1]``
        (or on the given ``edges``)."""
        edges = np.linspace(z0, z1, intervals + 1) if edges is None else np.asarray(edges)
        rows = []
        for a, b in zip(edges[:-1], edges[1:], strict=True):
            rows.append(
                np.polynomial.chebyshev.chebinterpolate(
                    lambda x, a=a, b=b: function(0.5 * (b - a) * x + 0.5 * (b + a)),
                    degree,
                )
            )
        return ScalarPath(2, 0.0, np.ascontiguousarray(edges), np.ascontiguousarray(rows))

    def derivative(self) -> "ScalarPath":
        """The path's derivative in z: the rate of a cumulative operator term."""
        if self.kind == 0:
            return ScalarPath.constant(0.0)
        if self.kind == 1:
            return ScalarPath.constant(1.0)
        rows = []
        for j in range(len(self.edges) - 1):
            a, b = float(self.edges[j]), float(self.edges[j + 1])
            d = np.polynomial.chebyshev.chebder(self.coefficients[j]) * (2.0 / (b - a))
            rows.append(d)
        return ScalarPath(2, 0.0, self.edges.copy(), np.ascontiguousarray(rows))

    def __call__(self, z: float) -> float:
        """Evaluate exactly as the library does (same operations, same order)."""
        if self.kind == 0:
            return self.value
        if self.kind == 1:
            return float(z)
        edges = self.edges
        n = len(edges) - 1
        j = 0
        while j < n - 1 and z >= edges[j + 1]:
            j += 1
        a, b = float(edges[j]), float(edges[j + 1])
        x = (2.0 * z - a - b) / (b - a)
        c = self.coefficients[j]
        b1 = 0.0
        b2 = 0.0
        for k in range(len(c) - 1, 0, -1):
            t = 2.0 * x * b1 - b2 + float(c[k])
            b2 = b1
            b1 = t
        return x * b1 - b2 + float(c[0])


# ------------------------------------------------------------------------ records


@dataclass(frozen=True)
class Tableau:
    """An embedded explicit Runge-Kutta pair (``src/solprop/stepper/tableau.py:104``)."""

    nodes: np.ndarray
    matrix: np.ndarray
    solution_weights: np.ndarray
    error_weights: np.ndarray
    first_same_as_last: bool


def dopri5() -> Tableau:
    """Dormand and Prince's RK5(4)7M pair, error weights as solution less embedded.

    The published coefficients (J. R. Dormand and P. J. Prince, J. Comput. Appl. Math.
    6, 19 (1980)); the reference holds the same table at
    ``src/solprop/stepper/tableau.py:475-506``.
    """
    nodes = np.array([0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
    matrix = np.array(
        [
            [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
            [1 / 5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
            [3 / 40, 9 / 40, 0.0, 0.0, 0.0, 0.0, 0.0],
            [44 / 45, -56 / 15, 32 / 9, 0.0, 0.0, 0.0, 0.0],
            [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729, 0.0, 0.0, 0.0],
            [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656, 0.0, 0.0],
            [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0],
        ]
    )
    solution = np.array([35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0])
    embedded = np.array(
        [5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40]
    )
    return Tableau(nodes, matrix, solution, solution - embedded, True)


@dataclass(frozen=True)
class KerrGroup:
    """Kerr coefficients (``_core.pyi`` ``Problem.Kerr``)."""

    coefficient: float
    cubic_form: str  # "real" or "analytic"
    self_factor: float = 0.75
    conjugate_factor: float = 0.0
    harmonic_factor: float = 0.0
    take_real: bool = False
    analytic_signal: bool = False


@dataclass(frozen=True)
class RamanGroup:
    """Scalar Raman: undamped rows on the doubled grid (``Problem.Raman``)."""

    undamped: np.ndarray  # (n_terms, 2 n_to) float64
    fraction: float
    epsilon0: float
    rates: tuple  # one ScalarPath per row: decay rate at z, 1/s
    thg: bool = False


@dataclass(frozen=True)
class PlasmaGroup:
    """One ionisation channel and its rate table (``Problem.Plasma``)."""

    charge_mass_ratio: float
    ionisation_potential: float
    log_field_min: float
    d_log_field: float
    field_min: float
    field_max: float
    rate_max: float
    rate_coefficients: np.ndarray  # (4, n_intervals) highest power first
    taps: np.ndarray = field(default_factory=lambda: np.ones(1))


@dataclass
class Problem:
    """Everything the library needs to run one propagation (the seam's arrays).

    Shapes use ``C = n_modes`` (one polarisation component), ``N = n_omega`` and
    ``T = n_to``.
    """

    # "envelope" (complex time samples), "field" (real) or "cartesian" (the free-space
    # envelope on an (ny, nx, nt) grid: columns iy nx + ix, `transverse` = (nx, ny))
    kind: str
    n_t: int
    n_to: int
    n_modes: int
    n_nodes: int
    crop_index: np.ndarray  # (N,) int64
    active_mask: np.ndarray  # (N,) bool
    norm_mask: np.ndarray  # (N,) bool
    omega_window: np.ndarray  # (N,) float64
    omega: np.ndarray  # (N,) float64, absolute angular frequency of each bin
    time_window: np.ndarray  # (T,) float64, the temporal collar g
    time_window_derivative: np.ndarray  # (T,) float64, g', 1/s
    dt_fine: float
    z_end: float
    phi_terms: np.ndarray  # (n_phi, C, N) complex128
    phi_paths: tuple  # n_phi cumulative ScalarPaths W_t(z)
    prefactor_terms: np.ndarray  # (n_pref, C, N) complex128
    prefactor_paths: tuple  # n_pref ScalarPaths p_t(z)
    synthesis: np.ndarray  # (Q, M) float64, carries the pad scale
    projection: np.ndarray  # (M, Q) float64, weights folded in
    density: ScalarPath
    preionisation: ScalarPath
    kerr: KerrGroup | None = None
    raman: RamanGroup | None = None
    plasma: PlasmaGroup | None = None
    tableau: Tableau = field(default_factory=dopri5)
    field0: np.ndarray | None = None  # (C, N) complex128 launch spectrum
    noise0: np.ndarray | None = None  # (C, N) complex128, shot-noise spectrum
    noise_model: str = "none"  # "none", "input" or "modified"
    rtol: float = 1e-6
    atol: float = 0.0
    # "rk" (embedded Runge-Kutta), "adams" (exponential Adams) or "low_storage" (the
    # RK4IP replay of the fixed schedule `schedule`, its edges in m; field route)
    method: str = "rk"
    schedule: np.ndarray | None = None
    adams_order: int = 5
    adams_variable_order: bool = True
    # The temporal absorber (``Problem.Absorber``): "none", "in_stage" (the rate on the fine
    # grid, n_to) or "split_step" (on the coarse grid, n_t), a power rate in 1/m.
    absorber_scheme: str = "none"
    absorber_rate: np.ndarray | None = None
    # The radial route's transverse collar (``Problem.Spatial.absorber_matrix``):
    # F diag(-alpha(r) / 2) B on the mode axis, (C, C) float64, 1/m; None without one.
    spatial_absorber: np.ndarray | None = None
    # The Cartesian route's transverse grid (nx, ny); None on the other routes.
    transverse: tuple | None = None
    # The embedded coarse rule's projection (``Problem.Transform.projection_coarse_matrix``,
    # _core.pyi:222): (M, Q) float64 with the Gauss weights on the Kronrod nodes they share
    # and zero elsewhere; None when the rule has no embedded pair.
    projection_coarse: np.ndarray | None = None
    description: dict = field(default_factory=dict)

    @property
    def n_omega(self) -> int:
        return int(self.crop_index.size)

    @property
    def phi_rate_paths(self) -> tuple:
        """``dW_t/dz``: the operator's rate ``D(z) = Σ_t W_t'(z) Φ_t`` (cached)."""
        cached = self.__dict__.get("_rate_paths")
        if cached is None or cached[0] is not self.phi_paths:
            cached = (self.phi_paths, tuple(p.derivative() for p in self.phi_paths))
            self.__dict__["_rate_paths"] = cached
        return cached[1]

    def phi_rate(self, z: float) -> np.ndarray:
        """``D(z)`` in the state's layout (``translate/linop.py`` ``phi_rate``)."""
        w = [p(z) for p in self.phi_rate_paths]
        total = np.zeros(self.phi_terms.shape[1:], dtype=np.complex128)
        for weight, term in zip(w, self.phi_terms, strict=True):
            total = total + weight * term
        return total.reshape(self.shape)

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.n_modes, 1, self.n_omega)

    @property
    def n_fine_bins(self) -> int:
        return self.n_to // 2 + 1 if self.kind == "field" else self.n_to

    # -- host scalars of a stage (the reference's StageScalars, translate/stage.py:170)

    def phi_weights(self, z0: float, z1: float) -> np.ndarray:
        return np.array([p(z1) - p(z0) for p in self.phi_paths])

    def phi_at(self, z: float) -> np.ndarray:
        return np.array([p(z) for p in self.phi_paths])

    def prefactor_weights(self, z: float) -> np.ndarray:
        return np.array([p(z) for p in self.prefactor_paths])

    def raman_rates(self, z: float) -> np.ndarray:
        if self.raman is None:
            return np.zeros(0)
        return np.array([p(z) for p in self.raman.rates])

    def validate(self) -> None:
        """Refuse shapes and dtypes the library cannot run (before any allocation)."""
        n, t, c = self.n_omega, self.n_to, self.n_modes
        if self.kind == "cartesian":
            self._validate_cartesian()
            return
        if self.transverse is not None:
            raise ValueError("a transverse grid belongs to the Cartesian route")
        if self.kind not in ("field", "envelope"):
            raise ValueError(f"kind must be 'field', 'envelope' or 'cartesian', got "
                             f"{self.kind!r}")
        if t & (t - 1) or t < 16:
            raise ValueError(f"n_to must be a power of two of at least 16, got {t}")
        if self.n_t > t or t % self.n_t:
            raise ValueError("n_to must be an integer multiple of n_t")
        expected = self.n_t // 2 + 1 if self.kind == "field" else self.n_t
        if n != expected:
            raise ValueError(f"crop_index has {n} bins, the grid needs {expected}")
        for name, arr, shape, dtype in (
            ("active_mask", self.active_mask, (n,), np.bool_),
            ("norm_mask", self.norm_mask, (n,), np.bool_),
            ("omega_window", self.omega_window, (n,), np.float64),
            ("omega", self.omega, (n,), np.float64),
            ("time_window", self.time_window, (t,), np.float64),
            ("time_window_derivative", self.time_window_derivative, (t,), np.float64),
            ("phi_terms", self.phi_terms, (len(self.phi_paths), c, n), np.complex128),
            (
                "prefactor_terms",
                self.prefactor_terms,
                (len(self.prefactor_paths), c, n),
                np.complex128,
            ),
            ("synthesis", self.synthesis, (self.n_nodes, c), np.float64),
            ("projection", self.projection, (c, self.n_nodes), np.float64),
        ):
            if arr.shape != shape or arr.dtype != dtype:
                raise ValueError(f"{name}: expected {shape} {np.dtype(dtype)}, got "
                                 f"{arr.shape} {arr.dtype}")
        if self.raman is not None and self.raman.undamped.shape[1] != 2 * t:
            raise ValueError("Raman rows must span the doubled grid, 2 n_to samples")
        if self.plasma is not None and self.kind != "field":
            raise ValueError("the field-grid plasma needs a real field grid")
        if self.absorber_scheme not in ("none", "in_stage", "split_step"):
            raise ValueError(f"unknown absorber scheme {self.absorber_scheme!r}")
        if self.absorber_scheme != "none":
            length = t if self.absorber_scheme == "in_stage" else self.n_t
            rate = self.absorber_rate
            if rate is None or rate.shape != (length,) or rate.dtype != np.float64:
                raise ValueError(f"absorber_rate: expected ({length},) float64 for the "
                                 f"{self.absorber_scheme} scheme")
            if np.any(rate < 0.0) or not np.all(np.isfinite(rate)):
                raise ValueError("absorber_rate must be finite and non-negative")
        if self.spatial_absorber is not None:
            m = self.spatial_absorber
            if m.shape != (self.n_modes, self.n_modes) or m.dtype != np.float64:
                raise ValueError("spatial_absorber: expected (n_modes, n_modes) float64")
            if not np.all(np.isfinite(m)):
                raise ValueError("spatial_absorber must be finite")
        if self.projection_coarse is not None:
            m = self.projection_coarse
            if m.shape != (c, self.n_nodes) or m.dtype != np.float64:
                raise ValueError(f"projection_coarse: expected {(c, self.n_nodes)} float64")
            if not np.all(np.isfinite(m)):
                raise ValueError("projection_coarse must be finite")

    def _validate_cartesian(self) -> None:
        """The free-space route's refusals: a power-of-two (ny, nx, nt) grid of at most
        4096 points an axis, one constant operator table with W(z) = z, one prefactor
        table, the analytic Kerr response alone, the low-storage replay."""
        if self.transverse is None or len(self.transverse) != 2:
            raise ValueError("the Cartesian route needs its transverse grid (nx, ny)")
        nx, ny = (int(v) for v in self.transverse)
        t = self.n_to
        for name, v in (("nx", nx), ("ny", ny), ("n_t", t)):
            if v < 4 or v > 4096 or v & (v - 1):
                raise ValueError(f"the Cartesian grid's {name} must be a power of two from "
                                 f"4 to 4096, got {v}")
        if self.n_t != t or self.n_omega != t:
            raise ValueError("a Cartesian grid keeps every bin: n_t = n_to = n_omega")
        if not np.array_equal(self.crop_index, np.arange(t)):
            raise ValueError("a Cartesian grid's crop is the identity")
        if self.n_modes != nx * ny:
            raise ValueError(f"a Cartesian grid has nx ny = {nx * ny} columns, got "
                             f"n_modes = {self.n_modes}")
        c, n = self.n_modes, t
        for name, arr, shape, dtype in (
            ("active_mask", self.active_mask, (n,), np.bool_),
            ("norm_mask", self.norm_mask, (n,), np.bool_),
            ("phi_terms", self.phi_terms, (1, c, n), np.complex128),
            ("prefactor_terms", self.prefactor_terms, (1, c, n), np.complex128),
        ):
            if arr.shape != shape or arr.dtype != dtype:
                raise ValueError(f"{name}: expected {shape} {np.dtype(dtype)}, got "
                                 f"{arr.shape} {arr.dtype}")
        if not self.active_mask.all():
            raise ValueError("a Cartesian grid's bins are all active (the tables carry the "
                             "band)")
        if len(self.phi_paths) != 1 or self.phi_paths[0].kind != 1:
            raise ValueError("the Cartesian operator is one constant table, W(z) = z")
        if len(self.prefactor_paths) != 1 or self.prefactor_paths[0].kind != 0:
            raise ValueError("the Cartesian prefactor is one constant table")
        if self.density.kind != 0:
            raise ValueError("the Cartesian medium is homogeneous (a constant density)")
        if self.kerr is not None and self.kerr.cubic_form != "analytic":
            raise ValueError("the Cartesian envelope takes the analytic Kerr form")
        if self.raman is not None or self.plasma is not None:
            raise ValueError("the Cartesian route is built with the Kerr response alone")
        if self.absorber_scheme != "none" or self.spatial_absorber is not None:
            raise ValueError("the Cartesian route has no absorbers")
        if self.method != "low_storage" or self.schedule is None:
            raise ValueError("the Cartesian route runs the low-storage replay of a schedule")


# ----------------------------------------------------------------------- builders


def _collar(n: int, width: int) -> np.ndarray:
    """A smooth 0..1..0 collar: raised-cosine edges of ``width`` samples."""
    w = np.ones(n)
    if width > 0:
        ramp = 0.5 - 0.5 * np.cos(np.pi * np.arange(width) / width)
        w[:width] = ramp
        w[n - width :] = ramp[::-1]
    return w


def _band(omega: np.ndarray, band) -> tuple[np.ndarray, np.ndarray]:
    """The active mask and spectral collar of a wavelength band (``grid.py:1359-1387``).

    Active: ``ω_min/2 < ω < 1.1 ω_max``; the collar is one inside ``[ω_min, ω_max]`` and
    a raised cosine to zero at the active edges.
    """
    w_lo = 2 * np.pi * C_LIGHT / band[1]
    w_hi = 2 * np.pi * C_LIGHT / band[0]
    active = (omega > 0.5 * w_lo) & (omega < 1.1 * w_hi)
    window = np.zeros(omega.shape)
    window[(omega >= w_lo) & (omega <= w_hi)] = 1.0
    lo = (omega > 0.5 * w_lo) & (omega < w_lo)
    window[lo] = 0.5 - 0.5 * np.cos(np.pi * (omega[lo] - 0.5 * w_lo) / (0.5 * w_lo))
    hi = (omega > w_hi) & (omega < 1.1 * w_hi)
    window[hi] = 0.5 + 0.5 * np.cos(np.pi * (omega[hi] - w_hi) / (0.1 * w_hi))
    return active, window


def _field_band(omega: np.ndarray, band) -> tuple[np.ndarray, np.ndarray]:
    """`_band` on a real field grid, whose zero-frequency and Nyquist bins stay inactive:
    a real inverse transform keeps only their real parts (``stepping.py:419-423``)."""
    active, window = _band(omega, band)
    active[0] = active[-1] = False
    window[0] = window[-1] = 0.0
    return active, window


def gnlse_problem(
    *,
    n_t: int = 8192,
    ffac: int = 2,
    window: float = 10e-12,
    wavelength: float = 1550e-9,
    fwhm: float = 50e-15,
    soliton_order: float = 4.0,
    betas=(-20e-27, 0.1e-39, -0.0001e-51, 0.0, 0.0),
    gamma: float = 1.1e-3,
    raman_fraction: float = 0.18,
    tau1: float = 12.2e-15,
    tau2: float = 32e-15,
    shock: bool = True,
    length: float = 1.0,
    rtol: float = 1e-6,
    noise_seed: int | None = 1,
    noise_model: str = "input",
    time_collar: float = 0.05,
    band=(700e-9, 5e-6),
    absorber: str = "none",
    absorber_length: float | None = None,
) -> Problem:
    """Single-mode envelope GNLSE: sech pulse at soliton order N, Kerr + Raman.

    The physics lives on the active bins of the wavelength `band`: the dispersion, the
    prefactor, the noise and the launch field are zero outside it, as the reference
    evaluates them on ``active_mask`` only (``grid.py:1359``, ``noise.py:195-214``).

    In the internal (unnormalised numpy) convention the envelope bin ``k`` sits at
    ``ω_0 + Δω_k``; ``Φ = -i Σ_{n≥2} β_n Δω^n / n! · z``; the prefactor is
    ``-i γ ω / ω_0`` (``ω/ω_0`` frozen without the shock) times the crop scale; Kerr is
    ``(1 - f_R)|A|^2 A`` and Raman ``f_R A (h * |A|^2)`` with ``∫h = 1`` (the row holds
    ``2 h`` against the ``|A|^2/2`` drive; ``ε_0 = 1`` in these units).
    ``betas`` are β_2..β_6 in s^n/m (β_2 = -20 ps^2/km by default).
    ``absorber`` ("none", "in_stage", "split_step") turns the time collar into a
    temporal absorber of reference length ``absorber_length`` (default ``length / 20``):
    its rate on the fine grid for the in-stage scheme, on the coarse grid (the collar
    at ``n_t`` samples) for the split-step one.
    """
    from .translate import dispersion_operator, taylor_beta

    n_to = ffac * n_t
    dt = window / n_t
    dt_fine = window / n_to
    omega0 = 2 * np.pi * C_LIGHT / wavelength
    dw = 2 * np.pi * np.fft.fftfreq(n_t, dt)
    omega = omega0 + dw
    active, omega_window = _band(omega, band)
    # A loss-free Taylor mode in the comoving envelope frame, which removes beta_0 and
    # beta_1 (so the expansion starts at beta_2).
    beta = taylor_beta((0.0, 0.0, *betas), omega0, omega[active])
    phi = dispersion_operator(omega, active, beta, np.zeros(beta.size), kind="envelope",
                              omega0=omega0, beta0=0.0, beta1=0.0, subtract_beta0=True,
                              length=length).reshape(1, 1, n_t)
    ratio = omega / omega0 if shock else np.ones_like(omega)
    crop_scale = n_t / n_to
    prefactor = np.where(active, -1j * gamma * ratio * crop_scale, 0.0).reshape(1, 1, n_t)
    t_window = _collar(n_to, int(time_collar * n_to))
    t_deriv = np.gradient(t_window, dt_fine)
    # Launch field: sech at soliton order N, P0 = N^2 |β2| / (γ T0^2).
    t0 = fwhm / (2 * np.arccosh(np.sqrt(2.0)))
    p0 = soliton_order**2 * abs(betas[0]) / (gamma * t0**2)
    t = (np.arange(n_t) - n_t // 2) * dt
    a_t = np.sqrt(p0) / np.cosh(t / t0)
    spectrum = (np.fft.fft(a_t) * active).reshape(1, n_t).astype(np.complex128)
    noise = None
    if noise_seed is not None and noise_model != "none":
        rng = np.random.default_rng(noise_seed)
        # One photon per bin: |A_k| = sqrt(hbar ω N / dt) (src/solprop/noise.py:195-246).
        phase = rng.uniform(0.0, 2.0 * np.pi, size=(1, 1, n_t))
        amplitude = np.zeros(n_t)
        amplitude[active] = np.sqrt(HBAR * omega[active] * n_t / dt)
        noise = (amplitude * np.exp(1j * phase)).reshape(1, n_t)
        if noise_model == "input":
            spectrum = spectrum + noise
    raman = None
    kerr_coefficient = 1.0
    if raman_fraction > 0.0:
        kerr_coefficient = 1.0 - raman_fraction
        delays = np.arange(2 * n_to) * dt_fine
        frequency = 1.0 / tau1
        amplitude = (frequency**2 + 1.0 / tau2**2) / frequency
        row = np.where(delays > 0.0, amplitude * np.sin(frequency * delays), 0.0)
        row[n_to:] = 0.0  # the padding half of the doubled grid stays zero
        raman = RamanGroup(
            undamped=np.ascontiguousarray(2.0 * row.reshape(1, -1)),
            fraction=raman_fraction,
            epsilon0=1.0,
            rates=(ScalarPath.constant(1.0 / tau2),),
        )
    kerr = KerrGroup(kerr_coefficient, "analytic", self_factor=1.0)
    return Problem(
        kind="envelope",
        n_t=n_t,
        n_to=n_to,
        n_modes=1,
        n_nodes=1,
        crop_index=np.ascontiguousarray(
            np.concatenate((np.arange(n_t // 2), np.arange(n_to - n_t // 2, n_to)))
        ).astype(np.int64),
        active_mask=active,
        norm_mask=active.copy(),
        omega_window=omega_window,
        omega=omega,
        time_window=t_window,
        time_window_derivative=t_deriv,
        dt_fine=dt_fine,
        z_end=length,
        phi_terms=np.ascontiguousarray(phi),
        phi_paths=(ScalarPath.position(),),
        prefactor_terms=np.ascontiguousarray(prefactor),
        prefactor_paths=(ScalarPath.constant(1.0),),
        synthesis=np.array([[n_to / n_t]]),
        projection=np.array([[1.0]]),
        density=ScalarPath.constant(1.0),
        preionisation=ScalarPath.constant(0.0),
        kerr=kerr,
        raman=raman,
        field0=np.ascontiguousarray(spectrum),
        noise0=None if noise is None else np.ascontiguousarray(noise),
        noise_model=noise_model if noise is not None else "none",
        rtol=rtol,
        absorber_scheme=absorber,
        absorber_rate=None if absorber == "none" else temporal_absorber_rate(
            t_window if absorber == "in_stage" else _collar(n_t, int(time_collar * n_t)),
            absorber_length if absorber_length else length / 20),
        description={
            "route": "envelope GNLSE",
            "n_t": n_t,
            "n_to": n_to,
            "window_s": window,
            "wavelength_m": wavelength,
            "fwhm_s": fwhm,
            "soliton_order": soliton_order,
            "gamma_per_W_m": gamma,
            "raman_fraction": raman_fraction,
            "length_m": length,
            "rtol": rtol,
            "noise_seed": noise_seed,
            "band_m": tuple(band),
        },
    )


# ------------------------------------------------------------ gas-filled capillary

HE_SELLMEIER = ((4977.77e-8, 28.54e-6), (1856.94e-8, 7.76e-3))
"""Helium's refractivity at 273.15 K and 1 bar, ``n - 1 = Σ B λ²/(λ² - C)`` with λ in
µm (A. Börzsönyi et al., Appl. Opt. 47, 4856 (2008))."""
HE_IONISATION_EV = 24.587387
HE_N2_PER_BAR = 3.4e-25  # m^2/W at 1 bar (the Kerr index of helium; Wahlstrand et al.)


def adk_rate(field, ip_ev: float, z: int = 1):
    """Tunnelling rate of an s-state by the ADK formula, 1/s, for a field in V/m.

    M. V. Ammosov, N. B. Delone and V. P. Krainov, Sov. Phys. JETP 64, 1191 (1986),
    with l = m = 0: in atomic units ``W = C² Ip (2 F0/F)^(2n*-1) exp(-2 F0/(3F))``,
    ``F0 = (2 Ip)^(3/2)``, ``n* = Z / sqrt(2 Ip)``, ``C² = 2^(2n*) / (n* Γ(n*+1) Γ(n*))``.
    """
    from math import gamma

    hartree_ev = 27.211386245988
    au_field = 5.14220674763e11
    au_time = 2.4188843265857e-17
    ip = ip_ev / hartree_ev
    n_star = z / math.sqrt(2.0 * ip)
    c2 = 2.0 ** (2 * n_star) / (n_star * gamma(n_star + 1.0) * gamma(n_star))
    f0 = (2.0 * ip) ** 1.5
    f = np.abs(np.asarray(field, dtype=float)) / au_field
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        w = c2 * ip * (2.0 * f0 / f) ** (2 * n_star - 1) * np.exp(-2.0 * f0 / (3.0 * f))
    return np.where(f > 0, w / au_time, 0.0)


# The band running-integral rule's design (src/solprop/translate/plasma_integration.py:52-74).
BAND_DESIGN_TOLERANCE = 1e-5
BAND_MAX_HALF_WIDTH = 8
BAND_ABOVE_BOUND = 1.05
BAND_FIT_NODES = 4097
BAND_CERTIFICATION_SAMPLES = 8193
BAND_CONDITION_LIMIT = 1e10


def trapezoid_multiplier(theta):
    """``(theta/2) cot(theta/2)``, the trapezoid's multiplier relative to the exact
    integral, with its series below 1e-4 (``plasma_integration.py:78-109``)."""
    theta = np.asarray(theta, dtype=np.float64)
    out = np.empty_like(theta)
    small = theta < 1e-4
    out[small] = 1.0 - theta[small] ** 2 / 12.0 - theta[small] ** 4 / 720.0
    half = theta[~small] / 2.0
    out[~small] = half / np.tan(half)
    return out


def band_composite(taps, theta):
    """``H = C rho``: the filtered trapezoid's multiplier (``plasma_integration.py:278-292``)."""
    theta = np.asarray(theta, dtype=np.float64)
    response = np.full_like(theta, taps[0])
    for k in range(1, len(taps)):
        response = response + 2.0 * taps[k] * np.cos(k * theta)
    return response * trapezoid_multiplier(theta)


def _band_fit(half: int, theta_b: float):
    """Least squares for ``C = 1/rho`` on ``[0, theta_b]`` under ``C(0) = 1``, columns
    scaled to unit norm (``plasma_integration.py:255-276``)."""
    if half == 0:
        return np.ones(1), 1.0
    theta = np.linspace(0.0, theta_b, BAND_FIT_NODES)
    k = np.arange(1, half + 1)
    design = 2.0 * (np.cos(np.outer(theta, k)) - 1.0)
    scale = np.linalg.norm(design, axis=0)
    scaled = design / scale
    target = 1.0 / trapezoid_multiplier(theta) - 1.0
    side, *_ = np.linalg.lstsq(scaled, target, rcond=None)
    side = side / scale
    return np.concatenate([[1.0 - 2.0 * side.sum()], side]), float(np.linalg.cond(scaled))


def _certified_max(function, lower: float, upper: float) -> float:
    """``max |function|`` on ``[lower, upper]``: dense samples, then a bounded search about
    every sampled local maximum above half the largest (``plasma_integration.py:295-330``)."""
    from scipy.optimize import minimize_scalar

    theta = np.linspace(lower, upper, BAND_CERTIFICATION_SAMPLES)
    values = np.abs(function(theta))
    best = float(values.max())
    inner = values[1:-1]
    peaks = np.flatnonzero((inner >= values[:-2]) & (inner >= values[2:]) & (inner >= 0.5 * best))
    for index in peaks + 1:
        result = minimize_scalar(lambda x: -abs(float(function(np.array([x]))[0])),
                                 bounds=(theta[index - 1], theta[index + 1]),
                                 method="bounded", options={"xatol": 1e-12})
        best = max(best, -float(result.fun))
    return best


def design_band_taps(dt: float, band_edge: float) -> np.ndarray:
    """The band running-integral rule's taps for one sample spacing and band edge: the
    smallest half-width K whose composite multiplier is within 1e-5 of one on the band
    and at most 1.05 above it (``plasma_integration.py:112-253``)."""
    theta_b = band_edge * dt
    if not 0.0 < theta_b < math.pi:
        raise ValueError(f"the band edge lies at or above Nyquist (theta_b = {theta_b:.4g})")
    for half in range(BAND_MAX_HALF_WIDTH + 1):
        taps, condition = _band_fit(half, theta_b)
        if condition > BAND_CONDITION_LIMIT:
            raise ValueError(f"the band-rule fit with K = {half} is ill-conditioned")
        in_band = _certified_max(lambda t, c=taps: band_composite(c, t) - 1.0, 0.0, theta_b)
        above = _certified_max(lambda t, c=taps: band_composite(c, t), theta_b, math.pi)
        if in_band <= BAND_DESIGN_TOLERANCE and above <= BAND_ABOVE_BOUND:
            return np.ascontiguousarray(taps)
    raise ValueError(f"no band-rule design with K <= {BAND_MAX_HALF_WIDTH} at theta_b = {theta_b:.4g}")


def with_integration(group: PlasmaGroup, integration: str, dt: float, omega, active):
    """The plasma group under a running-integral rule: ``"band"`` (the reference's
    default, ``nonlinear/responses.py:1091``) designs taps for this grid's spacing and the
    active band's upper edge; ``"trapezoid"`` keeps the single identity tap."""
    if integration == "trapezoid":
        return group
    if integration != "band":
        raise ValueError(f"integration must be 'band' or 'trapezoid', got {integration!r}")
    edge = float(np.max(np.asarray(omega)[np.asarray(active)]))
    import dataclasses

    return dataclasses.replace(group, taps=design_band_taps(dt, edge))


def rate_table(rate, field_min: float, field_max: float, n_intervals: int = 512):
    """A cubic spline of ``log W`` in ``log |E|`` on uniform knots, as the reference's
    ``translate/ionisation.py:95`` RateTable holds it (coefficients highest first)."""
    from scipy.interpolate import CubicSpline

    log_min, log_max = math.log(field_min), math.log(field_max)
    knots = np.linspace(log_min, log_max, n_intervals + 1)
    log_rate = np.log(rate(np.exp(knots)))
    spline = CubicSpline(knots, log_rate)
    coefficients = np.ascontiguousarray(spline.c)  # (4, n_intervals), highest first
    return PlasmaGroup(
        charge_mass_ratio=E_CHARGE**2 / M_E,
        ionisation_potential=HE_IONISATION_EV * E_CHARGE,
        log_field_min=log_min,
        d_log_field=(log_max - log_min) / n_intervals,
        field_min=field_min,
        field_max=field_max,
        rate_max=float(rate(np.array([field_max]))[0]),
        rate_coefficients=coefficients,
    )


# The depth of the temporal absorber at the window's edge, as alpha * ell
# (src/solprop/boundaries.py:101).
TEMPORAL_ALPHA_LENGTH = 12.6


def temporal_absorber_rate(profile: np.ndarray, reference_length: float) -> np.ndarray:
    """The temporal collar as a power absorption rate (``boundaries.py:433-470``).

    ``alpha(t) = TEMPORAL_ALPHA_LENGTH (1 - W(t)) / ell``: zero where the collar profile
    ``W`` is one, ``TEMPORAL_ALPHA_LENGTH / ell`` where it reaches zero.
    """
    if not (np.isfinite(reference_length) and reference_length > 0.0):
        raise ValueError("the reference length must be positive and finite")
    w = np.asarray(profile, dtype=np.float64)
    return np.ascontiguousarray(TEMPORAL_ALPHA_LENGTH * (1.0 - w) / reference_length + 0.0)


def oscillator_raman(n_to: int, dt: float, per_molecule: float, tau1: float,
                     rate: "ScalarPath", drive: str = "real") -> RamanGroup:
    """A molecular Raman response on a real field grid: one damped-oscillator row.

    ``h(t) = c (ω² + Γ²)/ω e^{-Γt} sin(ωt)`` with ``ω = 1/tau1`` and the decay ``Γ`` read
    from the path ``rate``; the undamped row (without ``e^{-Γt}``) fills the first half of
    the doubled grid, so ``∫h dt = c`` with ``c = per_molecule`` (the response per
    molecule, in the units of the Kerr coefficient). The drive is the real square ``E²``
    (the field grid keeps the harmonic): ``P = ε0 ρ E (h * E²)``
    (``passes/responses.py:395-499``); ``drive="analytic"`` takes half the analytic
    signal's squared modulus instead, ``|E_a|² / 2`` (a grid without the harmonic,
    ``raman.py:929-990``).
    """
    if drive not in ("real", "analytic"):
        raise ValueError(f"unknown Raman drive {drive!r}")
    delays = np.arange(2 * n_to) * dt
    frequency = 1.0 / tau1
    decay = rate(0.0)
    amplitude = (frequency**2 + decay**2) / frequency
    row = np.where(delays > 0.0, amplitude * np.sin(frequency * delays), 0.0)
    row[n_to:] = 0.0  # the padding half of the doubled grid stays zero
    return RamanGroup(
        undamped=np.ascontiguousarray(row.reshape(1, -1)),
        fraction=per_molecule / EPS0,
        epsilon0=EPS0,
        rates=(rate,),
        thg=drive == "real",
    )


def capillary_problem(
    *,
    n_t: int = 2**14,
    window: float = 400e-15,
    radius: float = 125e-6,
    length: float = 1.5,
    n_modes: int = 4,
    n_nodes: int = 128,
    p0: float = 0.8,
    p1: float = 0.0,
    wavelength: float = 800e-9,
    fwhm: float = 7.5e-15,
    energy: float = 275e-6,
    band=(100e-9, 4e-6),
    temperature: float = 294.0,
    rtol: float = 1e-6,
    plasma: bool = True,
    kerr: bool = True,
    cladding_index: float = 1.45,
    time_collar: float = 0.06,
    raman_share: float = 0.0,
    raman_tau: tuple = (60e-15, 200e-15),
    raman_drive: str = "real",
    absorber: str = "none",
    absorber_length: float | None = None,
    integration: str = "band",
    radial_rule: str = "gauss",
) -> Problem:
    """Gas-filled capillary on a real field grid: HE_1m modes, a pressure gradient.

    Modes of a hollow capillary in Marcatili's model (E. A. J. Marcatili and R. A.
    Schmeltzer, Bell Syst. Tech. J. 43, 1783 (1964)): profile ``J0(u_m r/a)``,
    ``β_m = k n_gas - u_m² / (2 k a²)``, loss ``α_m = (u_m/2π)² λ²/a³ (ν²+1)/√(ν²-1)``.
    The state is the modal spectrum (sqrt(W)); the synthesis samples power-normalised
    profiles at an n_nodes-point Gauss-Legendre rule on [0, a]; the projection carries
    the rule's weights times ``2π r``; the prefactor is ``-i ω/4`` (internal
    convention). Helium's refractivity is linear in the density, so the operator is
    ``Φ(z) = z Φ_0 + (∫ρ dz) Φ_1`` with the steady-flow profile
    ``p(z) = sqrt(p0² + (z/L)(p1² - p0²))``. Kerr: ``P = c ρ E³`` (real form);
    tunnelling ionisation by the ADK rate tabulated as a log-log cubic spline.
    """
    from scipy.special import jn_zeros, j0, j1

    from .translate import dispersion_operator, fill_density, fill_integral, fill_pressure
    from .translate import ideal_density
    from .translate import kerr_group as translated_kerr

    n_to = n_t
    dt = window / n_t
    n = n_t // 2 + 1
    omega = 2 * np.pi * np.arange(n) / (n_t * dt)
    c = C_LIGHT
    active, omega_window = _field_band(omega, band)
    norm_mask = active.copy()
    # Temporal collar and its derivative.
    t = (np.arange(n_to) - n_to // 2) * dt
    width = time_collar * window
    edge = 0.5 * window - width
    g = np.ones(n_to)
    gp = np.zeros(n_to)
    a = np.abs(t) > edge
    s = (np.abs(t[a]) - edge) / width
    g[a] = np.where(s < 1, 0.5 + 0.5 * np.cos(np.pi * s), 0.0)
    gp[a] = np.where(s < 1, -0.5 * np.pi / width * np.sin(np.pi * s) * np.sign(t[a]), 0.0)
    # Modes.
    u = jn_zeros(0, n_modes)
    k = omega / c
    safe = np.where(omega > 0, omega, 1.0)
    lam = 2 * np.pi * c / safe
    nu = cladding_index
    beta_vac = np.zeros((n_modes, n))
    alpha = np.zeros((n_modes, n))
    for m in range(n_modes):
        beta_vac[m] = np.where(omega > 0, k - u[m] ** 2 / (2 * np.where(k > 0, k, 1.0) * radius**2), 0.0)
        alpha[m] = np.where(
            omega > 0, (u[m] / (2 * np.pi)) ** 2 * lam**2 / radius**3 * (nu**2 + 1) / np.sqrt(nu**2 - 1), 0.0
        )
    # Refractivity per unit density (m^3): (n - 1)/ρ_ref at 273.15 K, 1 bar.
    lam_um = lam * 1e6
    refr = np.zeros(n)
    for b_coef, c_coef in HE_SELLMEIER:
        refr += np.where(omega > 0, b_coef * lam_um**2 / (lam_um**2 - c_coef), 0.0)
    rho_ref = 1e5 / (KB * 273.15)
    beta_rho = k * refr / rho_ref  # dβ/dρ
    # The density expansion of the operator: the vacuum part in a frame moving at the
    # speed of light (a field grid's delay through zero frequency, beta_1 = 1/c), with
    # the loss clamped as the reference clamps it, and the part per unit density.
    phi0 = np.zeros((n_modes, n), dtype=np.complex128)
    phi1 = np.zeros((n_modes, n), dtype=np.complex128)
    # On a field grid the frame's delay runs from zero frequency: its reference frequency
    # enters only the check that it lies on the grid, which a grid too coarse for the
    # carrier (a refusal test's) would fail.
    omega_ref = float(np.clip(2 * np.pi * c / wavelength, omega.min(), omega.max()))
    for m in range(n_modes):
        phi0[m] = dispersion_operator(omega, active, beta_vac[m][active], alpha[m][active],
                                      kind="field", omega0=omega_ref, beta0=0.0,
                                      beta1=1.0 / c, subtract_beta0=False, length=length)
        phi1[m] = np.where(active, -1j * beta_rho, 0.0)
    phi_terms = np.ascontiguousarray(np.stack([phi0, phi1]))

    # The fill: the steady-flow pressure, its ideal-gas density and ∫ρ dz.
    def pressure(z):
        return fill_pressure(z, p0, p1, length)

    def density(z):
        return fill_density(z, p0, p1, length, temperature)

    def cumulative(z):
        return ideal_density(fill_integral(z, p0, p1, length), temperature)

    # Geometric refinement towards the end, where sqrt(1 - z/L) is steepest.
    fractions = 1.0 - 0.5 ** np.arange(0, 14)
    edges = np.unique(np.concatenate([np.linspace(0, 0.5, 5), fractions, [1.0]])) * length
    w_path = ScalarPath.chebyshev(lambda z: cumulative(z), 0.0, length, degree=24, edges=edges)
    rho_path = ScalarPath.chebyshev(lambda z: density(z), 0.0, length, degree=24, edges=edges)
    # Synthesis and projection on a Gauss-Legendre rule in r, or on the Kronrod extension
    # of one (rounded up to 2 n + 1 nodes) with the Gauss rule embedded in it.
    if radial_rule not in ("gauss", "gauss_kronrod"):
        raise ValueError(f"unknown radial rule {radial_rule!r}")
    if radial_rule == "gauss":
        x, wq = np.polynomial.legendre.leggauss(n_nodes)
        r = 0.5 * radius * (x + 1.0)
        wr = 0.5 * radius * wq * 2 * np.pi * r
    else:
        from .quadrature import radial_rule as radial_nodes

        xi, w_fine, w_coarse = radial_nodes(n_nodes, True)
        n_nodes = xi.size
        r = radius * xi
        wr = radius * w_fine * 2 * np.pi * r
    synthesis = np.zeros((n_nodes, n_modes))
    for m in range(n_modes):
        norm = math.sqrt(2.0 / (EPS0 * c)) / (math.sqrt(math.pi) * radius * abs(j1(u[m])))
        synthesis[:, m] = norm * j0(u[m] * r / radius)
    projection = np.ascontiguousarray((synthesis * wr[:, None]).T)
    projection_coarse = None
    if radial_rule == "gauss_kronrod":
        wr_coarse = radius * w_coarse * 2 * np.pi * r
        projection_coarse = np.ascontiguousarray((synthesis * wr_coarse[:, None]).T)
    prefactor = np.zeros((1, n_modes, n), dtype=np.complex128)
    for m in range(n_modes):
        prefactor[0, m] = np.where(active, -1j * omega / 4.0, 0.0)
    # Launch: a Gaussian at the carrier in HE11, cycle-averaged power P(t).
    omega0 = 2 * np.pi * c / wavelength
    p_peak = energy / (fwhm * math.sqrt(math.pi / (4 * math.log(2))))
    power = p_peak * np.exp(-4 * math.log(2) * t**2 / fwhm**2)
    a_t = np.sqrt(2 * power) * np.cos(omega0 * t)
    field0 = np.zeros((n_modes, n), dtype=np.complex128)
    field0[0] = np.fft.rfft(a_t) * active
    kerr_group = None
    chi3_per_rho = (4.0 / 3.0) * HE_N2_PER_BAR * EPS0 * c / (1e5 / (KB * 273.15))
    if kerr:
        # The real cubic of a field grid with its harmonic: eps0 chi3 rho (sum E^2) E.
        kerr_group = translated_kerr(chi3_per_rho * (1.0 - raman_share), "field", True)
    raman_group = None
    if raman_share > 0.0:
        # A molecular response carrying `raman_share` of the third-order response, its
        # dephasing rate growing with the local pressure (collisions).
        tau1, tau2 = raman_tau

        def decay(z):
            return (0.5 + 0.5 * pressure(z) / max(p0, 1e-12)) / tau2

        raman_group = oscillator_raman(
            n_to, dt, EPS0 * chi3_per_rho * raman_share, tau1,
            ScalarPath.chebyshev(decay, 0.0, length, degree=24, edges=edges), raman_drive)
    plasma_group = None
    if plasma:
        # Below 3 GV/m helium's rate is under 1e-150 /s: the table is exactly zero there.
        plasma_group = with_integration(
            rate_table(lambda f: adk_rate(f, HE_IONISATION_EV), 3e9, 2e12), integration, dt,
            omega, active)
    return Problem(
        kind="field",
        n_t=n_t,
        n_to=n_to,
        n_modes=n_modes,
        n_nodes=n_nodes,
        crop_index=np.arange(n, dtype=np.int64),
        active_mask=active,
        norm_mask=norm_mask,
        omega_window=omega_window,
        omega=omega,
        time_window=g,
        time_window_derivative=gp,
        dt_fine=dt,
        z_end=length,
        phi_terms=phi_terms,
        phi_paths=(ScalarPath.position(), w_path),
        prefactor_terms=np.ascontiguousarray(prefactor),
        prefactor_paths=(ScalarPath.constant(1.0),),
        synthesis=np.ascontiguousarray(synthesis),
        projection=projection,
        projection_coarse=projection_coarse,
        density=rho_path,
        preionisation=ScalarPath.constant(0.0),
        kerr=kerr_group,
        raman=raman_group,
        plasma=plasma_group,
        field0=np.ascontiguousarray(field0),
        noise0=None,
        noise_model="none",
        rtol=rtol,
        method="adams",
        absorber_scheme=absorber,
        absorber_rate=None if absorber == "none" else temporal_absorber_rate(
            g, absorber_length if absorber_length else length / 20),
        description={
            "route": "capillary UPPE, modal field grid",
            "n_t": n_t, "n_omega": n, "window_s": window, "radius_m": radius,
            "length_m": length, "modes": n_modes, "nodes": n_nodes,
            "pressure_bar": [p0, p1], "gas": "He", "wavelength_m": wavelength,
            "fwhm_s": fwhm, "energy_J": energy, "rtol": rtol,
            "stepper": "exponential Adams, variable order <= 5",
        },
    )


# ------------------------------------------------------------ free-space radial beam

AR_SELLMEIER = ((20332.29e-8, 206.12e-6), (34458.31e-8, 8.066e-3))
"""Argon's refractivity at 273.15 K and 1 bar, ``n - 1 = Σ B λ²/(λ² - C)`` with λ in µm
(A. Börzsönyi et al., Appl. Opt. 47, 4856 (2008))."""
AR_IONISATION_EV = 15.7596119
AR_N2_PER_BAR = 1.0e-23  # m^2/W at 1 bar (argon's Kerr index; Wahlstrand et al.)


def qdht(radius: float, n_r: int):
    """The raw order-zero quasi-discrete Hankel transform (``hankel.py:29-190``).

    M. Guizar-Sicairos and J. C. Gutiérrez-Vega, JOSA A 21, 53 (2004): with ``j_n`` the
    zeros of J0, ``S = j_{N+1}`` and ``J_n = |J1(j_n)|``, the symmetric matrix
    ``T_mn = 2 J0(j_m j_n / S) / (S J_m J_n)`` acts between the diagonal scalings
    ``J``: the forward transform of samples at ``r_n = R j_n / S`` is
    ``H_m = (R² / S) J_m Σ_n T_mn f_n / J_n`` (the quadrature of ``∫ r f J0(k r) dr`` at
    ``k_m = j_m / R``), the inverse the same with ``S / R²``.
    Returns ``(r, k_r, forward, inverse)``, the last two as (N, N) matrices acting on
    column vectors.
    """
    from scipy.special import jn_zeros, jv

    zeros = jn_zeros(0, n_r + 1)
    edge = float(zeros[-1])
    j1 = np.abs(jv(1, zeros[:-1]))
    matrix = 2 * jv(0, np.outer(zeros[:-1], zeros[:-1]) / edge)
    matrix /= edge * np.outer(j1, j1)
    forward_scale = radius * radius / edge
    conjugated = j1[:, None] * matrix / j1[None, :]
    return (radius * zeros[:-1] / edge, zeros[:-1] / radius, forward_scale * conjugated,
            conjugated / forward_scale)


def planck_edge(x: np.ndarray, outer: float, inner: float) -> np.ndarray:
    """One Planck-taper edge strictly between its breakpoints (``grid.py:153-166``):
    ``1 / (1 + exp(w/(x - outer) + w/(x - inner)))`` with ``w = inner - outer``, the
    exponent clipped to +-700 (past which the value is already 0 or 1)."""
    width = inner - outer
    exponent = width / (x - outer) + width / (x - inner)
    return 1.0 / (1.0 + np.exp(np.clip(exponent, -700.0, 700.0)))


def collar_profile(s: np.ndarray) -> np.ndarray:
    """The spatial collar's profile ``1 - W`` on its normalised depth ``s``
    (``spatial_absorber.py:110-142``): 0 at and before ``s = 0``, 1 at and beyond
    ``s = 1``, the complement of a Planck edge between (every derivative continuous)."""
    s = np.asarray(s, dtype=np.float64)
    out = np.where(s >= 1.0, 1.0, 0.0)
    inside = (s > 0.0) & (s < 1.0)
    out[inside] = 1.0 - planck_edge(s[inside], 1.0, 0.0)
    return out


def radial_collar_rate(r: np.ndarray, radius: float, width: float, alpha0: float) -> np.ndarray:
    """A transverse collar's power rate on the radial nodes: ``alpha0 P(s)`` with ``P``
    the collar profile and ``s`` the depth into the outer ``width`` of the aperture
    (the rate a ``SpatialAbsorberDesign`` places on the nodes,
    ``spatial_absorber.py:550-660``; its calibration of ``alpha0`` and the width is the
    caller's here)."""
    if not (0.0 < width < radius and np.isfinite(alpha0) and alpha0 >= 0.0):
        raise ValueError("the collar needs 0 < width < radius and a finite alpha0 >= 0")
    return alpha0 * collar_profile((np.asarray(r) - (radius - width)) / width)


def radial_absorber_matrix(forward: np.ndarray, inverse: np.ndarray, rate: np.ndarray,
                           pad_scale: float = 1.0) -> np.ndarray:
    """The collar composed on the mode axis, ``F diag(-alpha/2) B / pad_scale``
    (``src/solprop/backends/spatial.py:189-225``): the rate is fixed in time and z, so
    it commutes with the temporal transforms and is formed once."""
    return np.ascontiguousarray(forward @ ((-0.5 * rate[:, None]) * (inverse / pad_scale)))


def radial_problem(
    *,
    n_t: int = 2**13,
    window: float = 1.0e-12,
    n_r: int = 512,
    radius: float = 4.0e-3,
    length: float = 0.5,
    pressure: float = 1.0,
    wavelength: float = 800e-9,
    fwhm: float = 100e-15,
    w0: float = 1.0e-3,
    energy: float = 2.0e-3,
    band=(200e-9, 4e-6),
    temperature: float = 294.0,
    rtol: float = 1e-6,
    plasma: bool = True,
    kerr: bool = True,
    time_collar: float = 0.05,
    raman_share: float = 0.0,
    raman_tau: tuple = (60e-15, 200e-15),
    raman_drive: str = "real",
    absorber: str = "none",
    absorber_length: float | None = None,
    integration: str = "band",
    collar: tuple | None = None,
) -> Problem:
    """Free-space radially symmetric beam in argon on a real field grid.

    The state is the Hankel spectrum ``H(k_r, ω)`` at the QDHT's transverse frequencies
    (one "mode" per k_r sample); the nonlinear response acts at the QDHT's radial samples,
    so the synthesis is the inverse transform and the projection the forward one, both
    dense N_r x N_r matrices. Per mode the linear operator is ``Φ = -i (k_z - ω/c) z`` with
    ``k_z = sqrt(k² - k_r²)`` (evanescent components decay), the prefactor
    ``-i ω² / (2 ε0 c² k_z)`` on propagating components. Kerr from argon's n2 and
    tunnelling ionisation (ADK, argon's potential); argon, an atom, has no rotational
    Raman response. Constant density, so one operator term. ``collar = (width, alpha0)``
    adds a transverse collar over the aperture's outer ``width`` (m) with peak power rate
    ``alpha0`` (1/m), as the matrix ``F diag(-alpha(r)/2) B`` on the mode axis.
    """
    n_to = n_t
    dt = window / n_t
    n = n_t // 2 + 1
    omega = 2 * np.pi * np.arange(n) / (n_t * dt)
    c = C_LIGHT
    active, omega_window = _field_band(omega, band)
    norm_mask = active.copy()
    t = (np.arange(n_to) - n_to // 2) * dt
    width = time_collar * window
    edge = 0.5 * window - width
    g = np.ones(n_to)
    gp = np.zeros(n_to)
    a = np.abs(t) > edge
    s = (np.abs(t[a]) - edge) / width
    g[a] = np.where(s < 1, 0.5 + 0.5 * np.cos(np.pi * s), 0.0)
    gp[a] = np.where(s < 1, -0.5 * np.pi / width * np.sin(np.pi * s) * np.sign(t[a]), 0.0)
    r, k_r, forward, inverse = qdht(radius, n_r)
    # Argon's index at this pressure and temperature.
    safe = np.where(omega > 0, omega, 1.0)
    lam_um = 2 * np.pi * c / safe * 1e6
    refr = np.zeros(n)
    for b_coef, c_coef in AR_SELLMEIER:
        refr += np.where(omega > 0, b_coef * lam_um**2 / (lam_um**2 - c_coef), 0.0)
    refr *= pressure * 273.15 / temperature
    k = omega * (1.0 + refr) / c
    kz2 = k[None, :] ** 2 - k_r[:, None] ** 2
    propagating = kz2 > 0
    kz = np.where(propagating, np.sqrt(np.abs(kz2)), -1j * np.sqrt(np.abs(kz2)))
    phi0 = np.where(active[None, :], -1j * (kz - omega[None, :] / c), 0.0)
    phi_terms = np.ascontiguousarray(phi0.reshape(1, n_r, n).astype(np.complex128))
    safe_kz = np.where(propagating, kz.real, 1.0)
    pref = np.where(active[None, :] & propagating,
                    -1j * omega[None, :] ** 2 / (2 * EPS0 * c**2 * safe_kz), 0.0)
    prefactor = np.ascontiguousarray(pref.reshape(1, n_r, n).astype(np.complex128))
    # Launch: a Gaussian beam with a Gaussian envelope, its energy from the real field.
    omega0 = 2 * np.pi * c / wavelength
    tau = fwhm / math.sqrt(2 * math.log(2))
    e0 = math.sqrt(energy / (0.5 * EPS0 * c * (math.pi * w0**2 / 2) * tau * math.sqrt(math.pi / 2)))
    field_rt = (e0 * np.exp(-(r[:, None] / w0) ** 2) * np.exp(-(t[None, :] / tau) ** 2)
                * np.cos(omega0 * t[None, :]))
    spectrum_r = np.fft.rfft(field_rt, axis=-1) * active[None, :]
    field0 = np.ascontiguousarray(forward @ spectrum_r)
    density = 1e5 * pressure / (KB * temperature)
    kerr_group = None
    chi3 = (4.0 / 3.0) * AR_N2_PER_BAR * pressure * EPS0 * c
    if kerr:
        kerr_group = KerrGroup(EPS0 * chi3 / density * (1.0 - raman_share), "real")
    raman_group = None
    if raman_share > 0.0:
        # A molecular response carrying `raman_share` of the third-order response (argon
        # itself, an atom, has none: this is for molecular gases and the tests).
        raman_group = oscillator_raman(n_to, dt, EPS0 * chi3 / density * raman_share,
                                       raman_tau[0], ScalarPath.constant(1.0 / raman_tau[1]),
                                       raman_drive)
    plasma_group = None
    if plasma:
        plasma_group = with_integration(
            rate_table(lambda f: adk_rate(f, AR_IONISATION_EV), 1e9, 2e12), integration, dt,
            omega, active)
    return Problem(
        kind="field",
        n_t=n_t,
        n_to=n_to,
        n_modes=n_r,
        n_nodes=n_r,
        crop_index=np.arange(n, dtype=np.int64),
        active_mask=active,
        norm_mask=norm_mask,
        omega_window=omega_window,
        omega=omega,
        time_window=g,
        time_window_derivative=gp,
        dt_fine=dt,
        z_end=length,
        phi_terms=phi_terms,
        phi_paths=(ScalarPath.position(),),
        prefactor_terms=prefactor,
        prefactor_paths=(ScalarPath.constant(1.0),),
        synthesis=np.ascontiguousarray(inverse),
        projection=np.ascontiguousarray(forward),
        density=ScalarPath.constant(density),
        preionisation=ScalarPath.constant(0.0),
        kerr=kerr_group,
        raman=raman_group,
        plasma=plasma_group,
        field0=field0,
        noise0=None,
        noise_model="none",
        rtol=rtol,
        method="adams",
        absorber_scheme=absorber,
        absorber_rate=None if absorber == "none" else temporal_absorber_rate(
            g, absorber_length if absorber_length else length / 20),
        spatial_absorber=None if collar is None else radial_absorber_matrix(
            forward, inverse, radial_collar_rate(r, radius, collar[0], collar[1])),
        description={
            "route": "free-space radial beam, QDHT field grid",
            "n_t": n_t, "n_omega": n, "window_s": window, "n_r": n_r, "radius_m": radius,
            "length_m": length, "pressure_bar": pressure, "gas": "Ar",
            "wavelength_m": wavelength, "fwhm_s": fwhm, "w0_m": w0, "energy_J": energy,
            "rtol": rtol, "stepper": "exponential Adams, variable order <= 5",
        },
    )


# -------------------------------------------------------------- free-space Cartesian

# Fused silica (I. H. Malitson, JOSA 55, 1205 (1965)): n^2 - 1 = sum B lambda^2 /
# (lambda^2 - C^2), lambda and C in um (the material of the reference's Kerr-beam cases,
# benchmarks/builders.py `kerr`).
SILICA_SELLMEIER = ((0.6961663, 0.0684043), (0.4079426, 0.1162414), (0.8974794, 9.896161))


def silica_wavenumber(omega: np.ndarray) -> np.ndarray:
    """k(omega) = n(omega) omega / c in fused silica, 1/m (positive omega only)."""
    lam_um = 2e6 * np.pi * C_LIGHT / np.asarray(omega, dtype=np.float64)
    n2 = np.ones_like(lam_um)
    for b, c in SILICA_SELLMEIER:
        n2 = n2 + b * lam_um**2 / (lam_um**2 - c**2)
    return np.sqrt(n2) * np.asarray(omega) / C_LIGHT


def transverse_axis(n: int, half_width: float) -> tuple[np.ndarray, np.ndarray]:
    """A periodic axis of n samples over [-a, a): positions (centred) and wavenumbers
    (unshifted FFT order), rad/m."""
    dx = 2.0 * half_width / n
    x = (np.arange(n) - n // 2) * dx
    return x, 2.0 * np.pi * np.fft.fftfreq(n, dx)


def cartesian_tables(k: np.ndarray, kperp_sq: np.ndarray, omega: np.ndarray,
                     active: np.ndarray, beta0: float, beta1: float, omega0: float,
                     rows: int = 4096) -> tuple[np.ndarray, np.ndarray]:
    """The free-space operator and prefactor tables, (C, N) each, in the internal
    convention (bin w at omega0 + dw, fields as exp(+i omega t)):

        D = -i (k_z - beta0 - beta1 dw),   p = -i omega^2 / (2 eps0 c^2 k_z)

    with k_z = sqrt(k^2 - kperp^2) on the propagating bins; an evanescent bin decays at
    |k_z| and has no source; inactive bins have neither (the spatial split of the
    reference's separable operator, src/solprop/translate/free.py:161-249, formed here
    whole, per element). Built a block of `rows` columns at a time."""
    c_count, n = kperp_sq.size, k.size
    phi = np.zeros((c_count, n), np.complex128)
    pref = np.zeros((c_count, n), np.complex128)
    k2 = np.where(active, k, 0.0) ** 2
    dw = omega - omega0
    for r0 in range(0, c_count, rows):
        arg = k2[None, :] - kperp_sq[r0:r0 + rows, None]
        prop = (arg > 0.0) & active[None, :]
        kz = np.sqrt(np.abs(arg))
        phi[r0:r0 + rows] = np.where(prop, -1j * (kz - beta0 - beta1 * dw[None, :]),
                                     np.where(active[None, :], -kz, 0.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            pref[r0:r0 + rows] = np.where(
                prop, -1j * omega[None, :] ** 2 / (2.0 * EPS0 * C_LIGHT**2 * kz), 0.0)
    return phi, pref


def cartesian_problem(
    *,
    nx: int = 64,
    ny: int = 64,
    n_t: int = 256,
    spacing: float = 1.05e-6,
    window: float = 250e-15,
    wavelength: float = 260e-9,
    band=(143e-9, 600e-9),
    fwhm: float = 30e-15,
    energy: float = 1e-7,
    waist: float = 40e-6,
    chi3: float = 2.2e-22,
    length: float = 2e-6,
    steps: int = 4,
    kerr: bool = True,
) -> Problem:
    """A Kerr beam in fused silica on a Cartesian (ny, nx, n_t) envelope grid, marched
    by `steps` fixed RK4IP steps over `length`.

    The reference's `kerr_beam_*` cases (benchmarks/registry.py:348-379,
    benchmarks/builders.py `kerr`): a 260 nm, 30 fs, 0.1 uJ Gaussian of 40 um waist
    in Malitson's silica with chi3 = 2.2e-22 m^2/V^2, a 250 fs window, the band 143-600
    nm, four 0.5 um steps over 2 um, the transverse spacing of the 384-point case (1.05
    um). Power-of-two axes here. Kerr: P = eps0 chi3 (3/4) |E|^2 E (E in V/m), so the
    coefficient is eps0 chi3 with the self factor 3/4. The launch is
    E = E0 exp(-r^2 / w0^2) exp(-2 ln 2 t^2 / tau^2), E0 from the energy at the carrier's
    index (U = n0 eps0 c E0^2 / 2 * pi w0^2 / 2 * tau sqrt(pi / (4 ln 2)))."""
    c_count = nx * ny
    dt = window / n_t
    omega0 = 2.0 * np.pi * C_LIGHT / wavelength
    dw = 2.0 * np.pi * np.fft.fftfreq(n_t, dt)
    omega = omega0 + dw
    lo, hi = 2.0 * np.pi * C_LIGHT / band[1], 2.0 * np.pi * C_LIGHT / band[0]
    active = (omega > lo) & (omega < hi)
    k = np.zeros(n_t)
    k[active] = silica_wavenumber(omega[active])
    beta0 = float(silica_wavenumber(np.array([omega0]))[0])
    h = 1e-4 * omega0
    beta1 = float((silica_wavenumber(np.array([omega0 + h]))[0]
                   - silica_wavenumber(np.array([omega0 - h]))[0]) / (2.0 * h))
    x, kx = transverse_axis(nx, 0.5 * nx * spacing)
    y, ky = transverse_axis(ny, 0.5 * ny * spacing)
    kperp_sq = (ky[:, None] ** 2 + kx[None, :] ** 2).reshape(c_count)
    phi, pref = cartesian_tables(k, kperp_sq, omega, active, beta0, beta1, omega0)
    n0 = beta0 * C_LIGHT / omega0
    tau = fwhm
    area = 0.5 * np.pi * waist**2
    duration = tau * math.sqrt(math.pi / (4.0 * math.log(2.0)))
    e0 = math.sqrt(2.0 * energy / (n0 * EPS0 * C_LIGHT * area * duration))
    t = (np.arange(n_t) - n_t // 2) * dt
    envelope_t = np.exp(-2.0 * math.log(2.0) * t**2 / tau**2)
    profile_x = np.exp(-(x**2) / waist**2)
    profile_y = np.exp(-(y**2) / waist**2)
    # fftn of the separable launch: the product of the axes' transforms.
    spec_t = np.fft.fft(e0 * envelope_t) * active
    spec_x = np.fft.fft(profile_x)
    spec_y = np.fft.fft(profile_y)
    field0 = ((spec_y[:, None] * spec_x[None, :]).reshape(c_count, 1)
              * spec_t[None, :]).astype(np.complex128)
    edges = np.linspace(0.0, length, steps + 1)
    return Problem(
        kind="cartesian",
        n_t=n_t,
        n_to=n_t,
        n_modes=c_count,
        n_nodes=0,
        crop_index=np.arange(n_t, dtype=np.int64),
        active_mask=np.ones(n_t, dtype=bool),
        norm_mask=active.copy(),
        omega_window=active.astype(np.float64),
        omega=omega,
        time_window=np.ones(n_t),
        time_window_derivative=np.zeros(n_t),
        dt_fine=dt,
        z_end=length,
        phi_terms=phi.reshape(1, c_count, n_t),
        phi_paths=(ScalarPath.position(),),
        prefactor_terms=pref.reshape(1, c_count, n_t),
        prefactor_paths=(ScalarPath.constant(1.0),),
        synthesis=np.zeros((0, c_count)),
        projection=np.zeros((c_count, 0)),
        density=ScalarPath.constant(1.0),
        preionisation=ScalarPath.constant(0.0),
        kerr=KerrGroup(EPS0 * chi3, "analytic", self_factor=0.75) if kerr else None,
        field0=field0,
        method="low_storage",
        schedule=edges,
        transverse=(nx, ny),
        description={
            "route": "free-space Cartesian",
            "grid": (ny, nx, n_t),
            "spacing_m": spacing,
            "window_s": window,
            "wavelength_m": wavelength,
            "band_m": tuple(band),
            "fwhm_s": fwhm,
            "energy_J": energy,
            "waist_m": waist,
            "chi3_m2_per_V2": chi3,
            "length_m": length,
            "steps": steps,
        },
    )
