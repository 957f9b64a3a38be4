"""Restated reference passes of one stage -- TEST INFRASTRUCTURE (see ``oracle``).

Each function restates one pass of solprop's NumPy reference backend, in the same
order of operations, for the routes the build runs:

- the transforms (``src/solprop/backends/numpy_ref/passes/transforms.py``) and the
  Fourier conventions (``src/solprop/fourier.py``);
- the step arithmetic (``src/solprop/backends/numpy_ref/passes/stepping.py``);
- the responses (``src/solprop/backends/numpy_ref/passes/responses.py``): Kerr in the
  real and the analytic/envelope cubic forms, scalar Raman on the doubled grid;
- the plasma column (``src/solprop/backends/numpy_ref/passes/plasma.py``): the rate
  table, the two trapezoid scans, the loss current and the held-current spectrum;
- one stage's source evaluation (``src/solprop/backends/numpy_ref/source.py:204-299``
  and ``:599-700``).

Parity pins (what checks this restatement, in ``tests/``): the reference's docstring
known answers for every pass; the DFT definition and Parseval for the transforms; the
Luna plasma witness (``tests/data/luna_plasma.json`` of the reference, copied as a
fixture) for the rate, fraction and polarisation; closed forms (a linear-only stage,
the damped oscillator's integral) for the rest.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.fft as _fft
from scipy import integrate

EPS0 = 8.8541878188e-12
"""Vacuum permittivity, F/m (CODATA 2022, as ``scipy.constants.epsilon_0``)."""


# --------------------------------------------------------------------- conventions


def n_omega_for(n_t: int, kind: str) -> int:
    """Spectral bins of ``n_t`` samples (``src/solprop/fourier.py:94-129``)."""
    return n_t // 2 + 1 if kind == "field" else n_t


def crop_index(n_t: int, n_to: int, kind: str) -> np.ndarray:
    """Coarse bins inside the fine spectrum (``src/solprop/fourier.py:541-587``)."""
    n_omega = n_omega_for(n_t, kind)
    if kind == "field":
        return np.arange(n_omega, dtype=np.intp)
    half = n_t // 2
    return np.concatenate(
        (np.arange(half, dtype=np.intp), np.arange(n_to - half, n_to, dtype=np.intp))
    )


def forward(x: np.ndarray, kind: str) -> np.ndarray:
    """Raw forward DFT, no scale (``src/solprop/fourier.py:156-199``)."""
    if kind == "field":
        return _fft.rfft(x, axis=-1)
    return _fft.fft(x, axis=-1)


def inverse(x: np.ndarray, kind: str) -> np.ndarray:
    """Inverse DFT with NumPy's ``1/N`` (``src/solprop/fourier.py:202-240``)."""
    n = x.shape[-1]
    if kind == "field":
        return _fft.irfft(x, n=2 * (n - 1), axis=-1)
    return _fft.ifft(x, n=n, axis=-1)


def analytic_signal(x: np.ndarray) -> np.ndarray:
    """Analytic signal keeping DC and Nyquist (``src/solprop/fourier.py:243-300``)."""
    n = x.shape[-1]
    spectrum = np.fft.fft(x, axis=-1)
    h = np.zeros(n)
    h[0] = 1.0
    h[n // 2] = 1.0
    h[1 : n // 2] = 2.0
    return np.fft.ifft(spectrum * h, axis=-1)


# ---------------------------------------------------------------------- transforms


def to_time(spectrum: np.ndarray, crop: np.ndarray, n_to: int, kind: str) -> np.ndarray:
    """Pad onto the fine grid and invert (``passes/transforms.py:31-84``)."""
    n_modes, n_pol = spectrum.shape[:2]
    fine = np.zeros((n_modes * n_pol, n_omega_for(n_to, kind)), dtype=np.complex128)
    fine[..., crop] = spectrum.reshape(n_modes * n_pol, -1)
    return inverse(fine, kind).reshape(n_modes, n_pol, n_to)


def synthesise(modal_t: np.ndarray, synthesis: np.ndarray, taper: float, n_pol: int):
    """Field at a chunk of nodes, one GEMM (``passes/transforms.py:87-139``)."""
    scaled = synthesis * taper
    nodes = np.tensordot(scaled, modal_t, axes=([1], [0]))
    n_rows, n_pol_state, n_to = nodes.shape
    return nodes.reshape(n_rows // n_pol, n_pol * n_pol_state, n_to)


def project(polarisation_t: np.ndarray, projection: np.ndarray, taper: float, n_pol: int):
    """Project a chunk back onto the modes (``passes/transforms.py:142-187``)."""
    n_chunk, n_pol_node, n_to = polarisation_t.shape
    rows = polarisation_t.reshape(n_chunk * n_pol, n_pol_node // n_pol, n_to)
    scaled = projection / taper
    return np.tensordot(scaled, rows, axes=([1], [0]))


def to_frequency(polarisation_t: np.ndarray, crop: np.ndarray, kind: str) -> np.ndarray:
    """Forward transform and crop (``passes/transforms.py:190-233``)."""
    n_modes, n_pol = polarisation_t.shape[:2]
    flat = polarisation_t.reshape(n_modes * n_pol, -1)
    spectrum = forward(flat, kind)[..., crop]
    return spectrum.reshape(n_modes, n_pol, -1)


# ------------------------------------------------------------------ step arithmetic


def stage_input(state, stage_sum, dz, propagator, noise):
    """``u = exp(dPhi_i)(y + h sum a_ij k_j) + N`` (``passes/stepping.py:38-85``)."""
    combined = propagator * (state + dz * stage_sum)
    return combined if noise is None else combined + noise


def stage_output(polarisation, prefactor, omega_window, inverse_propagator, absorber=None):
    """``k_i = exp(-dPhi_i)(p W P + A)`` (``passes/stepping.py:88-145``)."""
    normalised = prefactor * omega_window * polarisation
    if absorber is not None:
        normalised = normalised + absorber
    return inverse_propagator * normalised


def accumulate(total, stage, weight):
    """``total + w k_i`` (``passes/stepping.py:148-182``)."""
    return total + weight * stage


def accept_step(state, propagator, dz, accumulated):
    """``y <- exp(dPhi)(y + h sum b_i k_i)`` (``passes/stepping.py:185-226``)."""
    return propagator * (state + dz * accumulated)


def weak_norm_scale(norm_y, norm_y_new, rtol, atol):
    """``rtol max(|y|, |y'|) + atol`` (``src/solprop/stepper/controller.py:60-103``)."""
    return rtol * max(norm_y, norm_y_new) + atol


def error_norm(error, state, state_new, norm_mask, rtol, atol, weights=None):
    """Scaled weak norm over the norm mask (``passes/stepping.py:229-283``)."""
    norm_y = float(np.linalg.norm(state[..., norm_mask]))
    norm_new = float(np.linalg.norm(state_new[..., norm_mask]))
    if weights is None:
        scale = weak_norm_scale(norm_y, norm_new, rtol, atol)
        residual = error[..., norm_mask] / scale
    else:
        scale = rtol * weights * max(norm_y, norm_new) + atol
        residual = error[..., norm_mask] / scale[norm_mask]
    return float(np.linalg.norm(residual))


def absorber_term(modal_t, temporal_rate_fine, crop, active_mask, kind: str, radial=None):
    """The in-stage absorbers ``F[-alpha(t) E(t) / 2 + M E]``, cropped and masked to the
    active band (``passes/stepping.py:286-372``; the temporal term is
    ``boundaries.py:1159`` ``nonlinear_absorber_term``, ``M`` the radial collar's
    matrix on the mode axis, ``spatial.py:189-225``). Either part may be absent."""
    term = np.zeros_like(modal_t)
    if temporal_rate_fine is not None:
        term = -0.5 * temporal_rate_fine * modal_t
    if radial is not None:
        n_modes = modal_t.shape[0]
        term = term + (radial @ modal_t.reshape(n_modes, -1)).reshape(modal_t.shape)
    return to_frequency(term, crop, kind) * active_mask


def apply_temporal_rate(state, rate, dz, kind: str):
    """``y <- F[F^-1[y] exp(-alpha(t) h / 2)]`` on the coarse grid, the split-step
    absorber (``passes/stepping.py:410-456``)."""
    time = inverse(state, kind)
    return forward(time * np.exp(-0.5 * rate * dz), kind)


def apply_hard_mask(state, active_mask):
    """``y * active_mask`` once per accepted step (``passes/stepping.py:372-407``)."""
    return state * active_mask


# ------------------------------------------------------------------------ responses


@dataclass(frozen=True)
class Kerr:
    """The compiled Kerr group (``translate/responses.py:235`` KerrCoefficients)."""

    coefficient: float
    cubic_form: str  # "real" or "analytic"
    self_factor: float = 0.75
    conjugate_factor: float = 0.0
    harmonic_factor: float = 0.0
    analytic_signal: bool = False
    take_real: bool = False


def kerr_polarisation(field_t, kerr: Kerr, density, carrier=None):
    """Instantaneous cubic (``passes/responses.py:65-137``)."""
    values = analytic_signal(field_t) if kerr.analytic_signal else field_t
    factor = kerr.coefficient * density
    if kerr.cubic_form == "real":
        polarisation = factor * np.sum(values**2, axis=-2, keepdims=True) * values
    else:
        squares = np.sum(values**2, axis=-2, keepdims=True)
        moduli = np.sum(np.abs(values) ** 2, axis=-2, keepdims=True)
        polarisation = factor * (
            kerr.self_factor * moduli * values
            + kerr.conjugate_factor * squares * np.conj(values)
        )
        if carrier is not None:
            polarisation = polarisation + (
                factor * kerr.harmonic_factor * carrier * squares * values
            )
    return polarisation.real if kerr.take_real else polarisation


def kernel_spectrum(kernel, dt):
    """``rfft(kernel) dt`` on the doubled grid (``translate/raman.py:317-375``)."""
    return np.fft.rfft(np.asarray(kernel)) * dt


def raman_spectrum(undamped, rates, fraction, dt):
    """Damp, sum, scale, transform (``passes/responses.py:218-281``)."""
    values = np.asarray(rates, dtype=float)
    delays = np.arange(undamped.shape[1]) * dt
    total = np.zeros(undamped.shape[1])
    for row, rate in zip(undamped, values, strict=True):
        total += row * np.exp(-rate * delays)
    return kernel_spectrum(fraction * total, dt)


def raman_drive(column, kind: str, thg: bool):
    """The real quadratic drive (``src/solprop/nonlinear/raman.py:929-990``)."""
    if kind == "envelope":
        return 0.5 * (np.real(column) ** 2 + np.imag(column) ** 2)
    if thg:
        return np.ascontiguousarray(np.real(column) ** 2)
    a = analytic_signal(np.real(column))
    return 0.5 * (np.real(a) ** 2 + np.imag(a) ** 2)


def raman_polarisation(field_t, spectrum, kind: str, thg: bool, density: float,
                       epsilon0: float = EPS0):
    """Scalar delayed response on the doubled grid (``passes/responses.py:395-499``).

    ``epsilon0`` is the problem's ``Raman.epsilon0`` (``native/src/solprop_native/
    _core.pyi`` ``Problem.Raman``): the vacuum permittivity in SI units.
    """
    n_to = field_t.shape[-1]
    drive = np.zeros((field_t.shape[0], 2 * n_to))
    for node in range(field_t.shape[0]):
        drive[node, :n_to] = raman_drive(field_t[node, 0], kind, thg)
    convolved = np.fft.irfft(spectrum * np.fft.rfft(drive, axis=-1), n=2 * n_to, axis=-1)[
        ..., :n_to
    ]
    return (epsilon0 * density) * field_t * convolved[:, np.newaxis, :]


# --------------------------------------------------------------------------- plasma


@dataclass(frozen=True)
class RateTable:
    """Cubic spline of ``log W`` in ``log |E|`` (``translate/ionisation.py:95-313``)."""

    coefficients: np.ndarray  # (4, n_intervals), highest power first
    log_field_min: float
    d_log_field: float
    field_min: float
    field_max: float
    rate_max: float

    def log_evaluate(self, log_field):
        """The interpolant (``translate/ionisation.py:242-272``)."""
        index = np.clip(
            ((log_field - self.log_field_min) / self.d_log_field).astype(np.intp),
            0,
            self.coefficients.shape[1] - 1,
        )
        local = log_field - (self.log_field_min + index * self.d_log_field)
        c = self.coefficients
        return ((c[0, index] * local + c[1, index]) * local + c[2, index]) * local + c[
            3, index
        ]

    def __call__(self, field_amplitude):
        """Zero below ``field_min``, held above ``field_max`` (``:287-313``)."""
        field = np.asarray(np.abs(np.asarray(field_amplitude, dtype=np.float64)))
        flat = np.atleast_1d(field).ravel()
        rate = np.zeros_like(flat)
        above = flat > self.field_max
        inside = (flat >= self.field_min) & ~above
        if bool(np.any(inside)):
            rate[inside] = np.exp(self.log_evaluate(np.log(flat[inside])))
        rate[above] = self.rate_max
        return rate.reshape(field.shape)


@dataclass(frozen=True)
class RunningIntegral:
    """The running-integral rule (``src/solprop/nonlinear/ionisation.py:410-690``).

    ``taps == (1.0,)`` is the trapezoid: cumulative trapezoid, no floor. Longer taps are
    the band rule: a symmetric FIR filter closed by point reflection, then the
    trapezoid, and the exposure floored at zero.
    """

    taps: tuple = (1.0,)

    @property
    def half_width(self) -> int:
        return len(self.taps) - 1

    def filter(self, values):
        """``c_0 y_n + sum_k c_k (y_{n-k} + y_{n+k})`` (``ionisation.py:611-653``)."""
        half = self.half_width
        values = np.asarray(values, dtype=np.float64)
        n_to = values.shape[-1]
        reach = np.arange(half, 0, -1)
        head = 2.0 * values[..., :1] - values[..., reach]
        tail = 2.0 * values[..., -1:] - values[..., n_to - 1 - reach[::-1]]
        extended = np.concatenate([head, values, tail], axis=-1)
        taps = np.asarray(self.taps)
        filtered = taps[0] * values
        for k in range(1, half + 1):
            filtered = filtered + taps[k] * (
                extended[..., half - k : half - k + n_to]
                + extended[..., half + k : half + k + n_to]
            )
        return filtered

    def integrate(self, values, dt):
        """Cumulative trapezoid of the (filtered) integrand (``ionisation.py:567-609``)."""
        if self.half_width == 0:
            return integrate.cumulative_trapezoid(values, dx=dt, initial=0, axis=-1)
        return integrate.cumulative_trapezoid(
            self.filter(values), dx=dt, initial=0, axis=-1
        )

    def floor(self, exposure):
        """``max(X, 0)`` under the band rule only (``ionisation.py:655-676``)."""
        if self.half_width == 0:
            return exposure
        return np.maximum(exposure, 0.0)


@dataclass(frozen=True)
class Plasma:
    """One plasma channel (``translate/responses.py`` PlasmaCoefficients)."""

    rate: RateTable
    charge_mass_ratio: float
    ionisation_potential: float
    integral: RunningIntegral = RunningIntegral()


def magnitude(field_t):
    """``|E|`` or ``hypot(E_x, E_y)`` (``passes/plasma.py:616-625``)."""
    if field_t.shape[-2] == 1:
        return np.abs(field_t[..., 0, :])
    return np.hypot(field_t[..., 0, :], field_t[..., 1, :])


def ionisation_rate(field_t, table: RateTable):
    """Rate on the field's magnitude (``passes/plasma.py:59-101``)."""
    return table(magnitude(field_t))


def plasma_scan(field_t, rate, plasma: Plasma, preionisation, dt):
    """Fraction and free current (``passes/plasma.py:104-191``)."""
    rule = plasma.integral
    exposure = rule.floor(rule.integrate(rate, dt))
    fraction = 1.0 - (1.0 - preionisation) * np.exp(-exposure)
    integrand = fraction[..., np.newaxis, :] * plasma.charge_mass_ratio * field_t
    free_current = rule.integrate(integrand, dt)
    return fraction, free_current


def plasma_total_current(field_t, rate, fraction, free_current, plasma: Plasma):
    """Free plus guarded loss current (``passes/plasma.py:280-323``)."""
    squared = magnitude(field_t) ** 2
    scale = np.zeros_like(squared)
    np.divide(
        plasma.ionisation_potential * rate * (1.0 - fraction),
        squared,
        out=scale,
        where=(rate > 0.0) & (squared > 0.0),
    )
    return free_current + scale[..., np.newaxis, :] * field_t


def plasma_current(field_t, rate, fraction, free_current, plasma: Plasma, density, dt):
    """``P = rho int J dt``, the time-domain oracle form (``passes/plasma.py:202-277``)."""
    current = plasma_total_current(field_t, rate, fraction, free_current, plasma)
    return density * plasma.integral.integrate(current, dt)


def held_plasma_spectrum(gated, edge, omega, crop, active):
    """``DFT[gJ]/(iw) - DFT[g'J]/(iw)^2`` on active bins (``passes/plasma.py:579-613``)."""
    current = np.fft.rfft(gated, axis=-1)[..., crop]
    derivative = np.fft.rfft(edge, axis=-1)[..., crop]
    result = np.zeros_like(current)
    q = 1j * omega[active]
    result[..., active] = current[..., active] / q - derivative[..., active] / q**2
    return result


@dataclass
class NonlinearSources:
    """Collared local polarisation and ungated current (``passes/responses.py:699``)."""

    polarisation: np.ndarray
    current: np.ndarray | None


def nonlinear_sources(
    field_t,
    *,
    kerr: Kerr | None,
    raman: np.ndarray | None,
    raman_thg: bool,
    plasma: Plasma | None,
    raman_epsilon0: float = EPS0,
    t_window_fine,
    kind: str,
    density: float,
    preionisation: float,
    dt: float,
    carrier=None,
):
    """One column's sources in the compiled order (``passes/responses.py:715-836``).

    The order is the reference's for a single plasma channel: the rate, the scan and
    the current, then Kerr, then Raman; the collar multiplies the local sum.
    """
    total = np.zeros_like(field_t)
    current = None
    if plasma is not None:
        rate = ionisation_rate(field_t, plasma.rate)
        fraction, free_current = plasma_scan(field_t, rate, plasma, preionisation, dt)
        current = density * plasma_total_current(
            field_t, rate, fraction, free_current, plasma
        )
    if kerr is not None:
        total = total + kerr_polarisation(field_t, kerr, density, carrier)
    if raman is not None:
        total = total + raman_polarisation(
            field_t, raman, kind=kind, thg=raman_thg, density=density,
            epsilon0=raman_epsilon0,
        )
    return NonlinearSources(total * t_window_fine, current)
