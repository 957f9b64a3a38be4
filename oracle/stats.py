"""Restated plane statistics -- TEST INFRASTRUCTURE.

The saved-plane statistics of ``src/solprop/stats/definitions.py`` on a stored spectrum
of shape ``(n_modes, n_omega)`` (one polarisation), unit modal power weight:

- :func:`energy` restates ``energy`` (``:363-397``) through ``fourier.energy_omega``
  (``src/solprop/fourier.py:805-856``): ``(dt / n_t) sum_k w_k |X_k|^2`` with ``w`` one on
  the envelope grid and ``[1, 4, ..., 4, 1]`` on the field grid.
- :func:`omega0` restates ``omega0`` (``:268-361``): the centroid of ``|X|^2`` over every
  bin, NaN for a column at most one eps of the largest column's total.
- :func:`energy_window` restates ``energy_window`` (``:399-441``) and ``_window_masks``
  (``:249-266``): hard half-open gates ``lo <= omega < hi``.
- :func:`peak_power` and :func:`peak_power_total` restate ``peak_power`` (``:443-477``)
  and ``peak_power_total`` (``:479-520``) over ``analytic_field`` (``:148-178``): the
  coarse inverse transform, then on a field grid ``fourier.analytic_signal``.

- :func:`ionisation` restates the plasma pair (``numpy_ref/statistics.py:502-546``);
  :func:`populations` every node's pair of the same scan (``native/src/general/
  statistics.cpp:636-650``); :func:`peak_intensity` the on-axis synthesis's peak
  (``statistics.cpp:517-603``).
- :func:`monitor_energies` and :func:`split_sums` restate the boundary monitor's
  reductions (``src/solprop/monitor.py:493-606``).
- :func:`quadrature_error` restates the quadrature estimate of a source evaluation
  (``native/src/general/trial_finish.inc:921-1028``, ``numpy_ref/source.py:929-995``):
  ``|w (F_coarse - F_fine)| / |w F_fine|`` over every mode and coarse bin, ``F`` the
  response's spectrum projected with the Kronrod and the embedded Gauss weights, ``w``
  the prefactor at the evaluation's position times the spectral window; NaN without an
  embedded rule or where the fine spectrum vanishes.

Pins (``tests/test_oracle.py``): Parseval's identity with the analytic signal's time
sum, the window tiling summing to the total, the centroid of a single bin, a sampled
envelope's peak, and an empty column's NaN; the monitor's total against the reference's
docstring identity, and its removal against the split collar's own factor; the peak
intensity of a one-hot row against the peak power of that mode and of a row of ones
against the coherent sum's peak; the populations' densest row against the plasma pair;
the quadrature estimate of a Kerr response launched in one mode against the closed form
its separable radial integrals give (``tests/test_quadrature.py``).
"""

from __future__ import annotations

import numpy as np

from oracle import passes as P


def _n_t(n_omega: int, kind: str) -> int:
    return 2 * (n_omega - 1) if kind == "field" else n_omega


def energy(field: np.ndarray, dt: float, kind: str) -> np.ndarray:
    power = np.abs(field) ** 2
    n_t = _n_t(field.shape[-1], kind)
    if kind == "envelope":
        return (dt / n_t) * np.sum(power, axis=-1)
    interior = 4.0 * np.sum(power[..., 1:-1], axis=-1)
    return (dt / n_t) * (power[..., 0] + interior + power[..., -1])


def omega0(field: np.ndarray, omega: np.ndarray) -> np.ndarray:
    density = np.abs(field) ** 2
    total = np.sum(density, axis=-1)
    weighted = np.sum(density * omega, axis=-1)
    centroid = np.full(total.shape, np.nan)
    floor = float(np.finfo(float).eps) * float(np.max(total, initial=0.0))
    np.divide(weighted, total, out=centroid, where=total > floor)
    return centroid


def energy_window(field, omega, windows, dt, kind):
    windows = np.asarray(windows, dtype=float).reshape(-1, 2)
    masks = (omega[None, :] >= windows[:, :1]) & (omega[None, :] < windows[:, 1:])
    return energy(field[None, ...] * masks[:, None, :], dt, kind)


def analytic_field(field: np.ndarray, kind: str) -> np.ndarray:
    samples = P.inverse(field, kind)
    return P.analytic_signal(samples) if kind == "field" else samples


def peak_power(field: np.ndarray, kind: str) -> np.ndarray:
    return np.max(np.abs(analytic_field(field, kind)) ** 2, axis=-1)


def power_trace(field: np.ndarray, kind: str) -> np.ndarray:
    """The total power ``sum_modes |a_m(t)|^2``, the trace ``fwhm_t`` measures
    (``definitions.py:865-923``; the native backend's packed trace,
    ``native/src/general/statistics.cpp:80-93``)."""
    return np.sum(np.abs(analytic_field(field, kind)) ** 2, axis=0)


def peak_power_window(field, omega, windows, kind):
    """``peak_power_window`` (``definitions.py:725-760``): the spectrum gated by each
    hard half-open window, then each mode's peak power."""
    windows = np.asarray(windows, dtype=float).reshape(-1, 2)
    masks = (omega[None, :] >= windows[:, :1]) & (omega[None, :] < windows[:, 1:])
    return np.stack([peak_power(field * m, kind) for m in masks])


def peak_power_total(field: np.ndarray, kind: str) -> float:
    return float(np.max(np.sum(np.abs(analytic_field(field, kind)) ** 2, axis=0)))


def monitor_power(field: np.ndarray, kind: str, columns=None) -> np.ndarray:
    """``BoundaryMonitor.power`` (``src/solprop/monitor.py:493-516``): the coarse samples'
    squared modulus, weighted by the columns' quadrature, summed over the columns."""
    density = np.abs(P.inverse(field, kind)) ** 2
    if columns is not None:
        density = density * np.asarray(columns)[:, None]
    return density.sum(axis=0)


def monitor_energies(field, weights, kind, columns=None) -> np.ndarray:
    """``BoundaryMonitor.energies`` (``monitor.py:518-549``): ``sum_t P`` and
    ``weights @ P`` (four rows: low guard, high guard, collar, rate)."""
    power = monitor_power(field, kind, columns)
    return np.concatenate([[power.sum()], weights @ power])


def split_sums(field, weights, step, kind, columns=None) -> np.ndarray:
    """``BoundaryMonitor.split_sums`` (``monitor.py:551-606``): what the split takes,
    ``P (1 - e^{-rate h})``, then the energy and the guards' and collar's energies of
    ``P e^{-rate h / 2}``."""
    power = monitor_power(field, kind, columns)
    rate = weights[3]
    taken = power @ -np.expm1(-rate * step)
    half = power * np.exp(-0.5 * rate * step)
    return np.concatenate([[taken, half.sum()], weights[:3] @ half])


def peak_intensity(field: np.ndarray, kind: str, on_axis, factor: float) -> float:
    """``peak_intensity`` (the native backend's ``statistics_intensity``,
    ``native/src/general/statistics.cpp:517-603``): the on-axis synthesis rows applied to
    the modes' analytic signals, the largest over the samples of the rows' summed squared
    moduli, times the intensity factor."""
    rows = np.atleast_2d(np.asarray(on_axis, dtype=np.float64))
    values = rows @ analytic_field(field, kind)
    return float(factor * np.max(np.sum(np.abs(values) ** 2, axis=0)))


def plasma_column(problem, state, z):
    """The plasma scan of the node fields at a plane: the ionised fraction and the rate,
    each (n_nodes, n_to)."""
    from oracle import rk

    src = rk.StageSource(problem)
    crop = problem.crop_index.astype(np.intp)
    modal = P.to_time(state.reshape(problem.shape), crop, problem.n_to, problem.kind)
    nodes = P.synthesise(modal, problem.synthesis, 1.0, 1)
    rate = P.ionisation_rate(nodes, src.plasma.rate)
    fraction, _ = P.plasma_scan(nodes, rate, src.plasma, problem.preionisation(z),
                                problem.dt_fine)
    return fraction, rate


def populations(problem, state, z) -> np.ndarray:
    """``populations`` (``statistics.cpp:636-650``, one channel): each node's ionised
    fraction at the window's end and its largest rate, (n_nodes, 2)."""
    fraction, rate = plasma_column(problem, state, z)
    return np.stack([fraction[:, -1], rate.max(axis=1)], axis=1)


def ionisation(problem, state, z):
    """The electron density and the peak ionisation rate at a plane
    (``src/solprop/backends/numpy_ref/statistics.py:502-546``): the plasma scan of the
    node fields, the node of largest fraction at the window's end (the first), density
    times that fraction, and that node's largest rate."""
    fraction, rate = plasma_column(problem, state, z)
    densest = int(np.argmax(fraction[:, -1]))
    return float(problem.density(z)) * fraction[densest, -1], float(np.max(rate[densest]))


def plasma_integration_parts(problem, state, z) -> np.ndarray:
    """The plasma column's integration diagnostics at a plane, one channel
    (``src/solprop/backends/numpy_ref/passes/plasma.py:488-552``): at every node the
    exposure ``S`` as the rule forms it (before the floor) and its increments ``dS``, the
    fraction, the work the loss current pays ``int W (1 - f) dt`` (the trapezoid of the
    raw rate), the electrons freed ``f_end - f_0`` and the free-current integrand
    ``|f E|``; the six maxima over the nodes: ``[max(-min dS, 0), max dS,
    max |paid - freed|, max freed, max |f E| over the first 2K samples, max |f E|]``."""
    from oracle import rk

    src = rk.StageSource(problem)
    crop = problem.crop_index.astype(np.intp)
    modal = P.to_time(state.reshape(problem.shape), crop, problem.n_to, problem.kind)
    nodes = P.synthesise(modal, problem.synthesis, 1.0, 1)  # (Q, 1, T)
    return integration_parts(nodes, src.plasma, problem.preionisation(z), problem.dt_fine)


def integration_parts(nodes, plasma, preionisation, dt) -> np.ndarray:
    """The six parts of :func:`plasma_integration_parts` over node fields
    ``(Q, n_pol, T)`` in V/m, one channel."""
    rule = plasma.integral
    reach = 2 * rule.half_width
    magnitude = P.magnitude(nodes)
    rate = plasma.rate(magnitude)
    exposure = rule.integrate(rate, dt)
    steps = np.diff(exposure, axis=-1)
    fraction = 1.0 - (1.0 - preionisation) * np.exp(-rule.floor(exposure))
    paid = np.trapezoid(rate * (1.0 - fraction), dx=dt, axis=-1)
    freed = fraction[..., -1] - fraction[..., 0]
    integrand = np.abs(fraction[..., np.newaxis, :] * nodes)
    early = float(np.max(integrand[..., :reach])) if reach else 0.0
    return np.array([max(-float(np.min(steps)), 0.0), float(np.max(steps)),
                     float(np.max(np.abs(paid - freed))), float(np.max(freed)), early,
                     float(np.max(integrand))])


def plasma_integration_statistics(parts) -> tuple[float, float, float]:
    """The three diagnostics of the six parts (``passes/plasma.py:555-575``): the
    undershoot ``max(-dS) / max(dS)``, the work defect ``max |paid - freed| / max freed``
    and the boundary level, each zero where its denominator is."""
    dip, rise, defect, freed, early, peak = (float(part) for part in parts)
    return (dip / rise if rise > 0.0 else 0.0, defect / freed if freed > 0.0 else 0.0,
            early / peak if peak > 0.0 else 0.0)


__all__ = ["analytic_field", "energy", "energy_window", "monitor_energies", "monitor_power",
           "integration_parts", "ionisation", "omega0", "peak_intensity", "peak_power",
           "peak_power_total", "peak_power_window", "plasma_column", "plasma_integration_parts",
           "plasma_integration_statistics", "populations", "power_trace",
           "quadrature_error", "split_sums"]


def quadrature_error(problem, lab, z) -> float:
    """The quadrature estimate of the source evaluation at the lab-frame field ``lab``
    (``(n_modes, n_omega)``) and position ``z``: the response's spectrum projected with
    the problem's projection and with its embedded coarse one, both weighted by the
    prefactor at z times the spectral window, ``sqrt(sum |w (c - f)|^2 / sum |w f|^2)``;
    NaN without an embedded rule or when the fine spectrum is zero."""
    from oracle import rk

    if problem.projection_coarse is None:
        return float("nan")
    src = rk.StageSource(problem)
    lab = np.asarray(lab).reshape(problem.shape)
    fine, _ = src.projected_spectrum(lab, z)
    coarse, _ = src.projected_spectrum(lab, z, problem.projection_coarse)
    weight = src.prefactor(z) * problem.omega_window
    scale = float(np.sum(np.abs(weight * fine) ** 2))
    if scale == 0.0:
        return float("nan")
    return float(np.sqrt(np.sum(np.abs(weight * (coarse - fine)) ** 2) / scale))
