"""The translation step on the host: a mode's data to the arrays the propagation path reads.

The reference turns a waveguide mode (its propagation constant ``beta(omega)``, power
attenuation ``alpha(omega)`` and effective area), a Kerr susceptibility and a gas fill
into the seam's arrays before the march starts: the linear operator ``D(omega)`` on the
coarse frequency axis, the mode-average prefactor, the Kerr group and the density path.
This module restates those rules from supplied arrays, so that a builder given a
tabulated mode (or a Taylor expansion of one) produces the operator the reference would:

- the linear operator, its co-moving frame and the loss clamp
  (``src/solprop/linear.py:165-607``);
- a Taylor-specified mode and the dB-to-attenuation conversion
  (``src/solprop/nonlinear/gnlse.py:442-466``, ``:1453-1509``);
- the mode-average prefactor (``src/solprop/nonlinear/normalisation.py:56-243``);
- the GNLSE's ``gamma -> n2 -> chi3`` mapping (``gnlse.py:468-505``, ``:2290-2420``);
- the Kerr response's cubic forms and the compiled group they become
  (``src/solprop/nonlinear/responses.py:238-362``, ``:1421-1445``, ``:140-170``);
- a steadily flowing fill's pressure and density along a capillary
  (``src/solprop/profiles.py:1289-1340``: ``p^2`` linear in ``z``).

Constants are CODATA 2022 (``scipy.constants``), as the reference's are.
"""

from __future__ import annotations

import math

import numpy as np

from .problem import C_LIGHT, EPS0, KB, KerrGroup

MAX_ATTENUATION_EXPONENT = 708.0
"""The exponent past which ``exp(-x)`` underflows in binary64 (``linear.py:150-162``)."""


# ------------------------------------------------------------------ the mode's data


def taylor_beta(betas, omega_ref: float, omega, derivative: int = 0) -> np.ndarray:
    """``beta(omega)`` of a Taylor-specified mode, or its derivative.

    ``beta = sum_k beta_k (omega - omega_ref)^k / k!`` with ``betas = (beta_0, beta_1,
    ...)`` in s^k/m, summed by Horner's rule from the highest order
    (``gnlse.py:1453-1509``, ``TaylorDispersion``); ``derivative = 1`` gives
    ``d beta / d omega`` (the analytic group delay a frame removes).
    """
    coefficients = [float(b) for b in betas][derivative:]
    if not coefficients:
        return np.zeros_like(np.asarray(omega, dtype=float))
    delta = np.asarray(omega, dtype=float) - omega_ref
    total = np.zeros_like(delta)
    for k in range(len(coefficients) - 1, -1, -1):
        total = total * delta + coefficients[k] / math.factorial(k)
    return total


def attenuation_from_db(loss_db_per_m: float) -> float:
    """Power attenuation in 1/m of a loss in dB/m: ``ln(10) / 10 * loss`` (``gnlse.py:442``)."""
    return math.log(10.0) / 10.0 * loss_db_per_m


def loss_limit(length: float, max_alpha: float | None = None) -> float:
    """The largest attenuation the operator carries (``linear.py:165-240`` LossPolicy).

    A fixed ``max_alpha`` when given; else ``2 * 708 / length``, the loss at which the
    field has underflowed over the whole fibre, so that clamping there changes nothing
    at its end.
    """
    if not (math.isfinite(length) and length > 0.0):
        raise ValueError(f"length must be positive and finite, got {length!r}")
    if max_alpha is not None:
        if not (math.isfinite(max_alpha) and max_alpha > 0.0):
            raise ValueError(f"max_alpha must be positive and finite, got {max_alpha!r}")
        return float(max_alpha)
    return 2.0 * MAX_ATTENUATION_EXPONENT / length


def frame_origin(omega0: float, kind: str, subtract_beta0: bool) -> float:
    """Where the frame's delay term ``beta_1 (omega - origin)`` starts (``linear.py:379``).

    The carrier frequency on an envelope grid whose frame removes ``beta_0``; zero
    otherwise (a field grid, or an envelope grid that keeps the carrier phase, where the
    frame is a pure change of time origin).
    """
    return omega0 if kind == "envelope" and subtract_beta0 else 0.0


def dispersion_operator(omega, active, beta, alpha, *, kind: str, omega0: float,
                        beta0: float, beta1: float, subtract_beta0: bool, length: float,
                        max_alpha: float | None = None, absorber_rate=None) -> np.ndarray:
    """The linear operator ``D(omega)`` on the coarse axis (``linear.py:492-607``).

    ``beta`` and ``alpha`` are the mode's propagation constant (1/m) and power
    attenuation (1/m) at the **active** bins, in the order of ``omega[active]``. On the
    active band ``D = -i (beta - beta_0 - beta_1 (omega - origin)) - min(alpha, limit) /
    2``, with ``beta_0`` removed only when ``subtract_beta0`` and the origin from
    :func:`frame_origin`; outside it ``D`` is zero. An absorber rate (1/m, on every bin)
    adds ``-rate / 2`` over the whole axis.
    """
    omega = np.asarray(omega, dtype=float)
    active = np.asarray(active, dtype=bool)
    if not omega.min() <= omega0 <= omega.max():
        raise ValueError("the frame's reference frequency lies outside the grid")
    beta = np.asarray(beta, dtype=float)
    alpha = np.asarray(alpha, dtype=float)
    if beta.shape != (int(active.sum()),) or alpha.shape != beta.shape:
        raise ValueError("beta and alpha are given at the active bins")
    if not (np.all(np.isfinite(beta)) and np.all(np.isfinite(alpha))):
        raise ValueError("the mode's data must be finite on the active band")
    clamped = np.minimum(alpha, loss_limit(length, max_alpha))
    removed = beta0 if subtract_beta0 else 0.0
    origin = frame_origin(omega0, kind, subtract_beta0)
    operator = np.zeros(omega.size, dtype=complex)
    operator[active] = (-1j * (beta - removed - beta1 * (omega[active] - origin))
                        - 0.5 * clamped)
    if absorber_rate is not None:
        rate = np.asarray(absorber_rate, dtype=float)
        if rate.shape != omega.shape or np.any(rate < 0.0) or not np.all(np.isfinite(rate)):
            raise ValueError("an absorber rate is finite and non-negative on every bin")
        operator = operator - 0.5 * rate
    return operator


# ---------------------------------------------------------------- the prefactor


def sqrtw_to_vm(aeff: float) -> float:
    """Field in V/m per unit amplitude in sqrt(W): ``1 / sqrt(eps0 c A_eff / 2)``
    (``normalisation.py:56-90``)."""
    if not (math.isfinite(aeff) and aeff > 0.0):
        raise ValueError(f"aeff must be positive and finite, got {aeff!r}")
    return 1.0 / math.sqrt(0.5 * EPS0 * C_LIGHT * aeff)


def mode_average_prefactor(omega, active, aeff: float, omega0: float, *, shock: bool = True,
                           index=None) -> np.ndarray:
    """``-i omega / (2 eps0 c n_eff) / sqrtw_to_vm(A_eff)`` on the active band, zero outside
    (``normalisation.py:93-243``).

    ``index`` is the real effective index at the active bins (an array), or a constant
    (the GNLSE passes 1.0); without the shock the numerator is ``omega0`` throughout.
    """
    omega = np.asarray(omega, dtype=float)
    active = np.asarray(active, dtype=bool)
    band = omega[active]
    effective = (np.full(band.shape, float(index)) if np.isscalar(index)
                 else np.asarray(index, dtype=float))
    if effective.shape != band.shape or not np.all(np.isfinite(effective)) or np.any(
            effective == 0.0):
        raise ValueError("the effective index is finite and non-zero on the active band")
    numerator = band if shock else np.full(band.shape, omega0)
    prefactor = np.zeros(omega.size, dtype=complex)
    prefactor[active] = -1j * numerator / (2.0 * EPS0 * C_LIGHT * effective)
    return prefactor / sqrtw_to_vm(aeff)


# ------------------------------------------------------------------ the Kerr term


def gamma_mapping(gamma: float, aeff: float, omega_ref: float,
                  raman_fraction: float = 0.0) -> tuple[float, float]:
    """``(n2, chi3)`` of a GNLSE's ``gamma`` (``gnlse.py:468-505``, ``:2290-2420``).

    ``k0 = omega_ref / c``, ``n2 = gamma A_eff / k0`` and ``chi3 = (4/3) (1 - f_R) n2
    eps0 c``: the 4/3 undoes the 3/4 of the envelope's cubic, the linear index is one.
    """
    if not (0.0 <= raman_fraction <= 1.0):
        raise ValueError("the Raman fraction lies in [0, 1]")
    if not (omega_ref > 0.0 and aeff > 0.0):
        raise ValueError("omega_ref and aeff are positive")
    n2 = gamma * aeff / (omega_ref / C_LIGHT)
    return n2, 4.0 / 3.0 * (1.0 - raman_fraction) * n2 * EPS0 * C_LIGHT


def kerr_group(chi3: float, kind: str, thg: bool) -> KerrGroup:
    """The compiled Kerr group of a susceptibility ``chi3`` per unit density
    (``responses.py:238-362``; the factors ``responses.py:1421-1445`` name).

    A field grid with the harmonic: ``eps0 chi3 (sum_p E_p^2) E`` (the real cubic).
    Without it, or on an envelope grid: ``eps0 chi3 [(1/2)(sum |a|^2) a + (1/4)(sum a^2)
    conj(a)]`` of the analytic signal (taken real on a field grid), and on an envelope
    grid with the harmonic ``+ (1/4) eps0 chi3 carrier (sum a^2) a``.
    """
    factor = EPS0 * chi3
    if kind == "field" and thg:
        return KerrGroup(factor, "real")
    if kind == "field":
        return KerrGroup(factor, "analytic", self_factor=0.5, conjugate_factor=0.25,
                         take_real=True, analytic_signal=True)
    if kind != "envelope":
        raise ValueError(f"unknown grid kind {kind!r}")
    return KerrGroup(factor, "analytic", self_factor=0.5, conjugate_factor=0.25,
                     harmonic_factor=0.25 if thg else 0.0)


def thg_carrier(omega0: float, t) -> np.ndarray:
    """``exp(2 i omega0 t)``: what places the envelope's harmonic near ``3 omega0``
    (``responses.py:140-170``)."""
    return np.exp(2j * omega0 * np.asarray(t, dtype=float))


def analytic_signal(x) -> np.ndarray:
    """The analytic signal of real samples along the last axis: the positive
    frequencies doubled, DC and (for even length) Nyquist kept once, the negative ones
    removed (``src/solprop/fourier.py:243-300``, scipy's convention)."""
    x = np.asarray(x, dtype=float)
    n = x.shape[-1]
    spectrum = np.fft.fft(x, axis=-1)
    weight = np.zeros(n)
    weight[0] = 1.0
    if n % 2 == 0:
        weight[n // 2] = 1.0
        weight[1:n // 2] = 2.0
    else:
        weight[1:(n + 1) // 2] = 2.0
    return np.fft.ifft(spectrum * weight, axis=-1)


def kerr_polarisation(values, group: KerrGroup, density: float, carrier=None) -> np.ndarray:
    """The Kerr polarisation a compiled group describes, columns along axis -2
    (``responses.py:303-362``)."""
    values = np.asarray(values)
    if group.analytic_signal:
        values = analytic_signal(values)
    factor = group.coefficient * density
    squares = np.sum(values**2, axis=-2, keepdims=True)
    if group.cubic_form == "real":
        return factor * squares * values
    moduli = np.sum(np.abs(values) ** 2, axis=-2, keepdims=True)
    out = factor * (group.self_factor * moduli * values
                    + group.conjugate_factor * squares * np.conj(values))
    if carrier is not None and group.harmonic_factor:
        out = out + factor * group.harmonic_factor * np.asarray(carrier) * squares * values
    return np.ascontiguousarray(out.real) if group.take_real else out


# ------------------------------------------------------------------- the gas fill


def fill_pressure(z, p_in: float, p_out: float, length: float) -> np.ndarray:
    """The pressure of a steady laminar flow along a uniform bore, bar: ``p^2`` linear in
    ``z``, evaluated as the convex combination ``sqrt((1 - f) p_in^2 + f p_out^2)`` with
    ``f = z / L`` (which cannot cancel) and held constant outside ``[0, L]``
    (``src/solprop/profiles.py:1289-1340``, ``PressureGradient``)."""
    f = np.clip(np.asarray(z, dtype=float) / length, 0.0, 1.0)
    return np.sqrt((1.0 - f) * p_in**2 + f * p_out**2)


def ideal_density(pressure_bar, temperature: float) -> np.ndarray:
    """Number density of an ideal gas, m^-3: ``p / (k_B T)`` with ``p`` in bar."""
    return np.asarray(pressure_bar, dtype=float) * 1e5 / (KB * temperature)


def fill_density(z, p_in: float, p_out: float, length: float, temperature: float) -> np.ndarray:
    """The ideal-gas density along the flow (the builders' density path)."""
    return ideal_density(fill_pressure(z, p_in, p_out, length), temperature)


def fill_integral(z, p_in: float, p_out: float, length: float) -> np.ndarray:
    """``int_0^z p(z') dz'`` of :func:`fill_pressure`, bar m, in closed form:
    ``(2 L / (3 d)) ((p_in^2 + z d / L)^{3/2} - p_in^3)`` with ``d = p_out^2 - p_in^2``
    (``p_in z`` for a uniform fill)."""
    z = np.asarray(z, dtype=float)
    d = p_out**2 - p_in**2
    if abs(d) < 1e-300:
        return p_in * z
    return (2.0 * length / (3.0 * d)) * (
        np.maximum(p_in**2 + (z / length) * d, 0.0) ** 1.5 - p_in**3)
