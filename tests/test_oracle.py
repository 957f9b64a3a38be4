"""Pins of the oracle: known answers, closed forms and the Luna plasma witness.

These check the CPU restatement itself, so that the GPU parity tests compare against a
trusted oracle:

- every pass's docstring example in the reference (``passes/transforms.py``,
  ``passes/stepping.py``, ``passes/responses.py``, ``passes/plasma.py``,
  ``src/solprop/fourier.py``), which a dropped term or a transposed operand fails;
- the DFT definition on small sizes and Parseval;
- the linear-only problem, whose trial must be exactly ``exp(ΔΦ) y0``;
- the fundamental soliton, stationary under the full envelope stage (a wrong sign
  between the dispersion and the Kerr prefactor makes it spread or collapse);
- the damped oscillator's integral (``src/solprop/nonlinear/raman.py:243-310``);
- the plasma column against Luna's stages (``tests/data/luna_plasma.json`` of the
  reference), with the analytic eighth-power rate as a rate table, at the witness's own
  tolerances.
"""

import math

import numpy as np
import pytest

from oracle import passes as P
from oracle import rk
from solprop_amd.problem import E_CHARGE, M_E, gnlse_problem
from tests.conftest import GOLDEN

# ------------------------------------------------------------- docstring answers


def test_transform_docstring_answers():
    spectrum = np.array([[[1.0 + 0.0j, 0.0, 0.0]]])
    out = P.to_time(spectrum, np.arange(3), 8, "field")
    assert np.array_equal(out, np.full((1, 1, 8), 0.125))
    samples = np.ones((1, 1, 8))
    assert np.array_equal(P.to_frequency(samples, np.arange(3), "field").real,
                          [[[8.0, 0.0, 0.0]]])
    assert P.synthesise(np.ones((1, 1, 4)), np.array([[2.0], [3.0]]), 1.0, 1).shape == (2, 1, 4)
    assert P.project(np.ones((2, 1, 4)), np.array([[1.0, 3.0]]), 1.0, 1).shape == (1, 1, 4)
    assert P.crop_index(4, 8, "field").tolist() == [0, 1, 2]
    assert P.crop_index(4, 8, "envelope").tolist() == [0, 1, 6, 7]
    assert (P.n_omega_for(8, "field"), P.n_omega_for(8, "envelope")) == (5, 8)


def test_stepping_docstring_answers():
    state = np.ones((1, 1, 2), dtype=complex)
    zero = np.zeros_like(state)
    assert P.stage_input(state, zero, 0.1, 2.0 * np.ones_like(state), None).ravel().tolist() \
        == [2 + 0j, 2 + 0j]
    pol = np.ones((1, 1, 2), dtype=complex)
    out = P.stage_output(pol, np.array([[[-2.0j, 0.0j]]]), np.ones(2), np.ones_like(pol))
    assert out.ravel().tolist() == [-2j, 0j]
    assert P.accumulate(np.zeros((1, 1, 2), complex), np.ones((1, 1, 2)), 0.5).ravel().tolist() \
        == [0.5, 0.5]
    assert P.accept_step(state, 2.0 * np.ones_like(state), 0.5, state).ravel().tolist() \
        == [3 + 0j, 3 + 0j]
    assert round(P.error_norm(1e-2 * state, state, state, np.ones(2, bool), 1e-2, 0.0), 12) \
        == 1.0
    masked = P.apply_hard_mask(np.ones((1, 1, 3), complex), np.array([True, True, False]))
    assert masked.ravel().tolist() == [1, 1, 0]


def test_response_docstring_answers():
    kerr = P.Kerr(1.0, "analytic", 0.5, 0.25)
    field = np.array([[[2.0 + 0.0j]]])
    assert np.allclose(P.kerr_polarisation(field, kerr, 1.0), 0.75 * 4.0 * 2.0)
    rows = np.zeros((1, 8))
    rows[0, 1] = 1.0
    spectrum = P.raman_spectrum(rows, [0.0], 1.0, 2.0)
    assert spectrum.shape == (5,) and np.isclose(spectrum[0], 2.0)
    assert P.raman_drive(np.array([3.0, -2.0]), "field", True).tolist() == [9.0, 4.0]
    assert P.raman_drive(np.array([2.0 + 0j]), "envelope", True).tolist() == [2.0]
    flat = np.zeros(5, dtype=complex)
    assert np.all(P.raman_polarisation(np.ones((1, 1, 4)), flat, "field", True, 1.0) == 0.0)


def test_plasma_scan_docstring_answer():
    table = P.RateTable(np.zeros((4, 1)), 0.0, 1.0, 1.0, 2.0, 0.0)
    plasma = P.Plasma(table, E_CHARGE**2 / M_E, 2.5e-18)
    fraction, current = P.plasma_scan(np.zeros((1, 1, 4)), np.zeros((1, 4)), plasma, 0.25, 1e-16)
    assert fraction.tolist() == [[0.25, 0.25, 0.25, 0.25]] and np.all(current == 0.0)


# ------------------------------------------------------------------ transforms


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_transforms_are_the_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.allclose(P.forward(x, "envelope"), dft, atol=1e-12)
    assert np.allclose(P.inverse(dft, "envelope"), x, atol=1e-12)
    real = rng.normal(size=n)
    assert np.allclose(P.forward(real, "field"), (np.exp(-2j * np.pi * np.outer(k, k) / n)
                                                  @ real)[: n // 2 + 1], atol=1e-12)


def test_parseval():
    rng = np.random.default_rng(3)
    x = rng.normal(size=4096) + 1j * rng.normal(size=4096)
    X = P.forward(x, "envelope")
    assert math.isclose(np.sum(np.abs(x) ** 2), np.sum(np.abs(X) ** 2) / x.size, rel_tol=1e-13)


# -------------------------------------------------------------- closed forms


def small_gnlse(**kw):
    args = dict(n_t=256, ffac=2, window=2e-12, fwhm=50e-15, soliton_order=1.0,
                betas=(-20e-27,), raman_fraction=0.0, shock=False, noise_seed=None,
                length=0.05)
    args.update(kw)
    return gnlse_problem(**args)


def test_linear_only_trial_is_the_exact_flow():
    problem = small_gnlse()
    problem.kerr = None
    source = rk.StageSource(problem)
    y0 = problem.field0.reshape(problem.shape)
    advanced, norm, _ = rk.rk_trial(problem, source, y0, 0.0, 1e-3)
    assert np.array_equal(advanced, rk.linear_solution(problem, y0, 1e-3))
    assert norm == 0.0


def test_damped_oscillator_integrates_to_one():
    problem = gnlse_problem(n_t=2048, ffac=2, window=4e-12, noise_seed=None)
    row = problem.raman.undamped[0] / 2.0  # the row holds 2 h against the |A|^2/2 drive
    delays = np.arange(row.size) * problem.dt_fine
    h = row * np.exp(-delays / 32e-15)
    # The trapezoid's own error at dt = 1 fs on a 77 fs period is 6e-4.
    assert math.isclose(np.trapezoid(h, dx=problem.dt_fine), 1.0, rel_tol=2e-3)


@pytest.mark.slow
def test_fundamental_soliton_is_stationary():
    problem = small_gnlse(length=0.0636, rtol=1e-7)
    y0 = problem.field0.reshape(problem.shape)
    state, records = rk.march(problem, y0)
    assert any(r[3] for r in records)
    a0 = np.abs(np.fft.ifft(y0[0, 0]))
    a1 = np.abs(np.fft.ifft(state[0, 0]))
    assert np.max(np.abs(a1 - a0)) / np.max(a0) < 1e-6  # measured 3.3e-8
    # The same march with the prefactor's sign flipped is far from stationary (0.35).
    problem.prefactor_terms = -problem.prefactor_terms
    flipped, _ = rk.march(problem, y0)
    a2 = np.abs(np.fft.ifft(flipped[0, 0]))
    assert np.max(np.abs(a2 - a0)) / np.max(a0) > 0.1


def test_held_spectrum_is_the_gated_current_derivative_over_iw_squared():
    """``DFT[gJ]/(iω) - DFT[g'J]/(iω)^2 = DFT[g J']/(iω)^2`` on band-limited signals.

    Multiplying by ``(iω)^2``: ``iω DFT[gJ] - DFT[g'J] = DFT[(gJ)' - g'J] = DFT[g J']``,
    exact for periodic signals whose products stay below Nyquist (the spectral
    derivative is then the derivative). A sign or term error fails this.
    """
    n, dt = 256, 1e-15
    t = np.arange(n) * dt
    w = 2 * np.pi * np.fft.rfftfreq(n, dt)
    rng = np.random.default_rng(3)
    J = np.zeros(n)
    dJ = np.zeros(n)
    for k in range(3, 11):
        a, ph = rng.normal(), rng.uniform(0, 2 * np.pi)
        J += a * np.cos(w[k] * t + ph)
        dJ -= a * w[k] * np.sin(w[k] * t + ph)
    g = 1.0 + 0.3 * np.cos(w[2] * t)
    dg = -0.3 * w[2] * np.sin(w[2] * t)
    active = w > 0
    crop = np.arange(w.size)
    held = P.held_plasma_spectrum(g * J, dg * J, w, crop, active)
    expected = np.zeros_like(held)
    expected[active] = np.fft.rfft(g * dJ)[active] / (1j * w[active]) ** 2
    assert np.all(held[~active] == 0)
    assert np.max(np.abs(held - expected)) < 1e-12 * np.max(np.abs(expected))


# ----------------------------------------------------------- the Luna witness

W = np.load(GOLDEN / "luna_plasma_witness.npz")
CMR = E_CHARGE**2 / M_E
RESCALE = CMR / float(W["luna_CONSTANTS"][0])


def _rescaled(case, stage):
    free = W[f"luna_{case}_FREE_{stage}"]
    return free * RESCALE + (W[f"luna_{case}_{stage}"] - free)


def _deviation(got, expected):
    return float(np.max(np.abs(got - expected)) / np.max(np.abs(expected)))


def analytic_table(field):
    """The witness's eighth-power rate as a log-log table: exactly linear per interval."""
    threshold = float(W["param_rate_threshold"])
    scale = float(W["param_rate_scale"])
    reference = float(W["param_rate_reference"])
    field_max = 4.0 * float(np.max(np.abs(field)))
    log_min = math.log(threshold)
    n = 64
    d = (math.log(field_max) - log_min) / n
    knots = log_min + d * np.arange(n)
    coefficients = np.zeros((4, n))
    coefficients[2] = 8.0
    coefficients[3] = math.log(scale) + 8.0 * (knots - math.log(reference))
    rate_max = scale * (field_max / reference) ** 8
    return P.RateTable(coefficients, log_min, d, threshold, field_max, rate_max)


@pytest.mark.parametrize("case,signal", [("SCALAR", "field"), ("ZEROS", "zeroed")])
def test_plasma_column_matches_luna(case, signal):
    field = W[f"signal_{signal}"]
    dt = float(W["param_dt"])
    ip = float(W["param_ionisation_potential"])
    table = analytic_table(field)
    plasma = P.Plasma(table, CMR, ip)
    f3 = field[None, None, :]
    rate = P.ionisation_rate(f3, table)
    assert _deviation(rate[0], W[f"luna_{case}_RATE"]) < 1e-13
    fraction, free = P.plasma_scan(f3, rate, plasma, 0.0, dt)
    assert _deviation(fraction[0], W[f"luna_{case}_FRACTION"]) < 1e-13
    pol = P.plasma_current(f3, rate, fraction, free, plasma, 1.0, dt)
    assert _deviation(pol[0, 0], _rescaled(case, "P")) < 2e-13


def test_plasma_column_on_a_real_rate_matches_luna():
    field = W["signal_field"]
    dt = float(W["param_dt"])
    table = analytic_table(field)  # unused for the rate: the recorded PPT rate is given
    plasma = P.Plasma(table, CMR, float(W["param_ionisation_potential"]))
    rate = W["luna_PPT_RATE"][None, :]
    f3 = field[None, None, :]
    fraction, free = P.plasma_scan(f3, rate, plasma, 0.0, dt)
    pol = P.plasma_current(f3, rate, fraction, free, plasma, 1.0, dt)
    assert _deviation(pol[0, 0], _rescaled("PPT", "P")) < 2e-13


def test_preionised_fraction_stays_below_one():
    field = W["signal_field"]
    table = analytic_table(field)
    plasma = P.Plasma(table, CMR, float(W["param_ionisation_potential"]))
    f3 = field[None, None, :]
    rate = P.ionisation_rate(f3, table)
    fraction, _ = P.plasma_scan(f3, rate, plasma, 0.2, float(W["param_dt"]))
    assert np.all(np.diff(fraction[0]) >= -1e-15)
    assert fraction.min() >= 0.2 - 1e-15 and fraction.max() <= 1.0


# ------------------------------------------------------------------- temporal absorber


def test_absorber_term_docstring_example():
    """``passes/stepping.py:335-340``: a unit field under rate 2 gives -8 at DC, 0 above."""
    from oracle import passes as P

    field = np.ones((1, 1, 8))
    got = P.absorber_term(field, np.full(8, 2.0), np.arange(2, dtype=np.intp),
                          np.ones(2, bool), kind="field")
    assert np.array_equal(got.real, [[[-8.0, 0.0]]])
    assert not np.any(got.imag)


def test_split_absorber_is_an_exact_rate():
    """``passes/stepping.py:410-456``: zero rate leaves the state; a constant rate is the
    closed form y e^{-alpha h / 2} (a factor without the half fails it); a varying rate
    telescopes, two steps equal to one of their sum."""
    from oracle import passes as P

    rng = np.random.default_rng(4)
    n_t = 64
    state = rng.normal(size=(2, 1, n_t // 2 + 1)) + 1j * rng.normal(size=(2, 1, n_t // 2 + 1))
    state[..., 0] = state[..., 0].real  # a real signal's DC and Nyquist bins are real
    state[..., -1] = state[..., -1].real
    assert np.allclose(P.apply_temporal_rate(state, np.zeros(n_t), 0.3, "field"), state,
                       rtol=0, atol=1e-14)
    alpha, h = 3.0, 0.2
    got = P.apply_temporal_rate(state, np.full(n_t, alpha), h, "field")
    scale = np.abs(state).max()
    assert np.allclose(got, state * np.exp(-alpha * h / 2), rtol=0, atol=1e-14 * scale)
    assert not np.allclose(got, state * np.exp(-alpha * h), rtol=0, atol=1e-3 * scale)
    varying = rng.uniform(0.0, 50.0, n_t)
    two = P.apply_temporal_rate(P.apply_temporal_rate(state, varying, 0.1, "field"),
                                varying, 0.25, "field")
    one = P.apply_temporal_rate(state, varying, 0.35, "field")
    assert np.allclose(two, one, rtol=0, atol=1e-13 * np.abs(state).max())


def test_temporal_absorber_rate():
    """``boundaries.py:433-470``: zero on the flat region, 12.6 / ell at the window's edge."""
    from solprop_amd.problem import temporal_absorber_rate

    rate = temporal_absorber_rate(np.array([1.0, 0.5, 0.0]), 0.1)
    assert rate[0] == 0.0 and not np.signbit(rate[0])
    assert np.isclose(rate[1], 63.0) and np.isclose(rate[2], 126.0)


# ------------------------------------------------------- band running-integral rule


def test_band_rule_integrate_docstring_example():
    """``nonlinear/ionisation.py:598-602``: taps (1.1, -0.05) on 0, 1, 2, 3 at dt 0.5."""
    from oracle import passes as P

    rule = P.RunningIntegral((1.1, -0.05))
    assert rule.integrate(np.arange(4.0), 0.5).tolist() == [0.0, 0.25, 1.0, 2.25]


def test_trapezoid_multiplier_docstring_example():
    """``translate/plasma_integration.py:100-101``."""
    from solprop_amd.problem import trapezoid_multiplier

    assert f"{float(trapezoid_multiplier(np.array(np.pi / 3))):.6f}" == "0.906900"


@pytest.mark.parametrize("theta_b", [0.5, 1.26, 1.5])
def test_band_design_meets_its_certificate(theta_b):
    """The designed taps' composite multiplier is within 1e-5 of one over the band (on a
    sampling independent of the design's), one exactly at zero frequency, at most 1.05
    above the band; the plain trapezoid (one identity tap) misses the first by far."""
    from solprop_amd.problem import band_composite, design_band_taps

    taps = design_band_taps(theta_b, 1.0)  # dt = theta_b, edge = 1 rad/s
    band = np.linspace(0.0, theta_b, 30001)
    assert np.abs(band_composite(taps, band) - 1.0).max() <= 1e-5
    assert abs(band_composite(taps, np.array([0.0]))[0] - 1.0) < 1e-15
    assert np.abs(band_composite(taps, np.linspace(theta_b, np.pi, 30001))).max() <= 1.05
    assert np.abs(band_composite(np.ones(1), band) - 1.0).max() > 1e-3


def test_band_design_refuses_a_band_too_close_to_nyquist():
    """Below an oversampling of about 3.5 no design with K <= 8 certifies."""
    from solprop_amd.problem import design_band_taps

    with pytest.raises(ValueError, match="no band-rule design"):
        design_band_taps(2.0, 1.0)
    with pytest.raises(ValueError, match="Nyquist"):
        design_band_taps(3.5, 1.0)


def test_band_rule_integrates_a_band_limited_signal():
    """Between interior samples a cosine inside the band integrates to the difference of
    sin(ωt)/ω within the design tolerance (the trapezoid's error is ~θ²/12); the ends,
    closed by reflection, only add a constant."""
    from oracle import passes as P
    from solprop_amd.problem import design_band_taps

    dt, theta_b = 1.0, 1.2
    rule = P.RunningIntegral(tuple(design_band_taps(dt, theta_b)))
    t = np.arange(4096) * dt
    omega = 1.0
    got = rule.integrate(np.cos(omega * t), dt)
    exact = np.sin(omega * t) / omega
    inner = slice(50, -50)
    got_span = got[inner] - got[50]
    exact_span = exact[inner] - exact[50]
    assert np.abs(got_span - exact_span).max() < 3e-5
    plain = P.RunningIntegral((1.0,)).integrate(np.cos(omega * t), dt)
    assert np.abs(plain[inner] - plain[50] - exact_span).max() > 1e-2


def test_analytic_raman_drive_is_the_hilbert_pair():
    """Half the analytic signal's squared modulus (``raman_drive`` without the harmonic,
    ``nonlinear/raman.py:929-990``) equals (E^2 + H[E]^2)/2 with the Hilbert transform
    formed independently, as the real inverse transform of -i Y strictly between DC and
    Nyquist; the drive with the harmonic (E^2) does not."""
    from oracle import passes as P

    rng = np.random.default_rng(5)
    n = 256
    spectrum = rng.normal(size=n // 2 + 1) + 1j * rng.normal(size=n // 2 + 1)
    spectrum[0] = spectrum[0].real
    spectrum[-1] = spectrum[-1].real
    field = np.fft.irfft(spectrum, n=n)
    rotated = -1j * spectrum
    rotated[0] = rotated[-1] = 0.0
    hilbert = np.fft.irfft(rotated, n=n)
    expected = 0.5 * (field**2 + hilbert**2)
    got = P.raman_drive(field, "field", False)
    assert np.abs(got - expected).max() < 1e-13 * expected.max()
    assert np.abs(P.raman_drive(field, "field", True) - expected).max() > 0.1 * expected.max()


# ------------------------------------------------------------------ plane statistics


def _random_field_spectrum(rng, n_t, n_modes=3):
    x = rng.normal(size=(n_modes, n_t)) * np.exp(-((np.arange(n_t) - n_t / 2) / (n_t / 8)) ** 2)
    return np.fft.rfft(x, axis=-1)


def test_energy_is_parseval_with_the_analytic_signal():
    """``energy_omega``'s weights make it the analytic signal's time sum exactly
    (``fourier.py:805-856``): dt sum_t |a(t)|^2, with ``a`` formed by numpy's FFT on the
    full grid; on an envelope grid it is dt sum_t |A(t)|^2. Simpson-like weights
    [1, 4, 2, 4, ..] (a plausible mistake) fail it."""
    from oracle import stats as S

    rng = np.random.default_rng(11)
    n_t, dt = 256, 2e-15
    X = _random_field_spectrum(rng, n_t)
    x = np.fft.irfft(X, n=n_t, axis=-1)
    h = np.zeros(n_t)
    h[0] = h[n_t // 2] = 1.0
    h[1:n_t // 2] = 2.0
    a = np.fft.ifft(np.fft.fft(x, axis=-1) * h, axis=-1)
    expected = dt * np.sum(np.abs(a) ** 2, axis=-1)
    assert np.allclose(S.energy(X, dt, "field"), expected, rtol=1e-13, atol=0)
    w = np.array([1.0] + [4.0 if k % 2 else 2.0 for k in range(1, n_t // 2)] + [1.0])
    assert not np.allclose((dt / n_t) * np.sum(w * np.abs(X) ** 2, axis=-1), expected,
                           rtol=1e-3, atol=0)
    A = rng.normal(size=(2, 64)) + 1j * rng.normal(size=(2, 64))
    assert np.allclose(S.energy(np.fft.fft(A, axis=-1), dt, "envelope"),
                       dt * np.sum(np.abs(A) ** 2, axis=-1), rtol=1e-13, atol=0)


def test_energy_windows_tile_the_total_and_the_centroid_of_one_bin():
    """Windows that tile the axis partition the bins, so their energies sum to the total
    (the reference's docstring identity, ``definitions.py:431-439``); a spectrum with
    one occupied bin has that bin's frequency as its centroid; an empty column is NaN."""
    from oracle import stats as S

    rng = np.random.default_rng(12)
    n_t, dt = 128, 1e-15
    X = _random_field_spectrum(rng, n_t, 2)
    omega = 2 * np.pi * np.arange(n_t // 2 + 1) / (n_t * dt)
    cut = omega[20] + 0.5 * (omega[21] - omega[20])
    tiles = np.array([[-np.inf, cut], [cut, np.inf]])
    parts = S.energy_window(X, omega, tiles, dt, "field")
    assert np.allclose(parts.sum(axis=0), S.energy(X, dt, "field"), rtol=1e-14, atol=0)
    one = np.zeros((2, n_t // 2 + 1), complex)
    one[0, 7] = 3.0 - 1.0j
    centroid = S.omega0(one, omega)
    assert centroid[0] == omega[7] and np.isnan(centroid[1])


def test_peak_power_of_a_sampled_envelope_and_of_a_carrier():
    """An envelope grid's peak power is the largest |A(t)|^2 sampled; a field grid's
    carrier of amplitude E0 under a slow envelope has analytic power E0^2 at the peak
    (|a| is the envelope, to the Hilbert transform's accuracy on a band-limited pulse);
    the total's peak is the peak of the summed powers, not the sum of the peaks."""
    from oracle import stats as S

    n_t = 512
    t = np.arange(n_t) - n_t / 2
    env = np.exp(-(t / 40.0) ** 2)
    A = np.stack([2.0 * env, np.roll(env, 100)])
    assert np.allclose(S.peak_power(np.fft.fft(A, axis=-1), "envelope"), [4.0, 1.0],
                       rtol=1e-13)
    assert np.isclose(S.peak_power_total(np.fft.fft(A, axis=-1), "envelope"),
                      np.max(np.sum(np.abs(A) ** 2, axis=0)), rtol=1e-13)
    carrier = 3.0 * env * np.cos(2 * np.pi * 60 * np.arange(n_t) / n_t)
    X = np.fft.rfft(carrier[None, :], axis=-1)
    assert np.isclose(S.peak_power(X, "field")[0], 9.0, rtol=1e-10)


def test_peak_intensity_rows():
    """A one-hot on-axis row picks that mode's peak power; a row of ones the peak of the
    modes' coherent sum (two pulses in phase at one instant add in amplitude); two rows
    add their powers before the maximum; the factor scales."""
    from oracle import stats as S

    n_t = 256
    t = np.arange(n_t) - n_t / 2
    env = np.exp(-(t / 20.0) ** 2)
    A = np.stack([2.0 * env, env, np.roll(env, 64)])
    X = np.fft.fft(A, axis=-1)
    peaks = S.peak_power(X, "envelope")
    for m in range(3):
        row = np.eye(3)[m]
        assert np.isclose(S.peak_intensity(X, "envelope", row, 1.0), peaks[m], rtol=1e-13)
    assert np.isclose(S.peak_intensity(X, "envelope", [1.0, 1.0, 0.0], 1.0), 9.0, rtol=1e-13)
    two = S.peak_intensity(X, "envelope", [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], 2.5)
    assert np.isclose(two, 2.5 * np.max(4.0 * env**2 + np.roll(env, 64) ** 2), rtol=1e-13)


def test_populations_hold_the_plasma_pair():
    """The populations' densest node is the plasma pair's: its end fraction times the
    density is the electron density, its largest rate the peak ionisation rate."""
    from oracle import stats as S
    from solprop_amd.problem import capillary_problem

    problem = capillary_problem(n_t=512, n_nodes=8, length=0.05, integration="trapezoid")
    state = problem.field0 * 3.0
    pops = S.populations(problem, state, 0.0)
    density, rate = S.ionisation(problem, state, 0.0)
    assert pops.shape == (8, 2) and np.all(pops >= 0.0)
    densest = int(np.argmax(pops[:, 0]))
    assert np.isclose(problem.density(0.0) * pops[densest, 0], density, rtol=1e-15)
    assert pops[densest, 1] == rate
    assert np.all(pops[:, 0] <= 1.0)


def test_monitor_total_and_split_removal():
    """The monitor's total is sum |pulse|^2 for an envelope built as fft(pulse) (the
    reference's docstring identity, ``monitor.py:536-546``); the split's removal equals
    the energy the collar's own factor takes (``apply_temporal_rate``, an independent
    path: the total minus the power of the collared state); zero rate takes nothing and
    finds the whole energy."""
    from oracle import passes as P
    from oracle import stats as S

    n = 256
    t = np.arange(n) - n / 2
    pulse = np.exp(-((t - 0.45 * n) / 6.0) ** 2) + 0.3 * np.exp(-(t / 20.0) ** 2)
    field = np.fft.fft(pulse.astype(complex))[None, :]
    edge = (np.abs(t) > 0.4 * n).astype(float)
    rate = 50.0 * np.clip((np.abs(t) - 0.35 * n) / (0.15 * n), 0.0, 1.0)
    weights = np.stack([(t < -0.4 * n) * 1.0, (t > 0.4 * n) * 1.0, edge, rate])
    energies = S.monitor_energies(field, weights, "envelope")
    assert np.isclose(energies[0], np.sum(np.abs(pulse) ** 2), rtol=1e-13)
    h = 0.01
    taken, found = S.split_sums(field, weights, h, "envelope")[:2]
    collared = P.apply_temporal_rate(field, rate, h, "envelope")
    after = S.monitor_power(collared, "envelope").sum()
    assert np.isclose(taken, energies[0] - after, rtol=1e-12)
    assert taken > 0.0
    still = S.split_sums(field, np.stack([weights[0], weights[1], weights[2], 0 * rate]), h,
                         "envelope")
    assert still[0] == 0.0 and np.isclose(still[1], energies[0], rtol=1e-14)
