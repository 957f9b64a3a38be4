"""The radial route's quasi-discrete Hankel transform and problem record (CPU).

``qdht`` restates the transform of ``src/solprop/hankel.py:29-190`` (Guizar-Sicairos and
Gutiérrez-Vega, JOSA A 21, 53 (2004)). It is pinned by closed forms, not by retyping it:
- two Hankel pairs, ∫ r e^{-p r²} J0(k r) dr = e^{-k²/4p} / 2p and
  ∫ r³ e^{-p r²} J0(k r) dr = (1 - k²/4p) e^{-k²/4p} / 2p², sampled on the
  transform's own grids;
- the inverse undoing the forward;
- a mutant without the |J1| scalings failing the first pair.
"""

import numpy as np
import pytest

from solprop_amd.problem import qdht, radial_problem

# A Gaussian of 0.7 mm in a 4 mm aperture is e^-33 at the edge: the transform, which
# takes the function to vanish there, is then exact to rounding.
RADIUS, N_R, A = 4.0e-3, 256, 0.7e-3
P = 1.0 / A**2


def gaussian_pair(r, k):
    return np.exp(-P * r**2), np.exp(-k**2 / (4 * P)) / (2 * P)


def second_pair(r, k):
    return r**2 * np.exp(-P * r**2), (1 - k**2 / (4 * P)) * np.exp(-k**2 / (4 * P)) / (2 * P**2)


@pytest.mark.parametrize("pair", [gaussian_pair, second_pair])
def test_qdht_matches_closed_forms(pair):
    r, k, forward, inverse = qdht(RADIUS, N_R)
    f, exact = pair(r, k)
    got = forward @ f
    assert np.abs(got - exact).max() < 1e-12 * np.abs(exact).max()
    # The inverse maps the closed-form spectrum back onto the samples.
    assert np.abs(inverse @ exact - f).max() < 1e-11 * np.abs(f).max()


def test_qdht_round_trip():
    """The kernel is only nearly orthogonal: |inverse . forward - I| falls with N (about
    1.5e-8 at 32 points, 3e-11 at 256, 4e-12 at 512); a smooth profile comes back to
    rounding."""
    r, k, forward, inverse = qdht(RADIUS, N_R)
    rng = np.random.default_rng(3)
    f = rng.normal(size=(N_R, 5))
    assert np.abs(inverse @ (forward @ f) - f).max() < 1e-10 * np.abs(f).max()
    smooth, _ = gaussian_pair(r, k)
    assert np.abs(inverse @ (forward @ smooth) - smooth).max() < 1e-13
    assert np.abs(inverse @ forward - np.eye(N_R)).max() > 1e-13  # not an identity by chance


def test_qdht_without_j1_scalings_fails_the_closed_form():
    """The symmetric kernel alone (a plausible mistake) does not reproduce the pair."""
    from scipy.special import jn_zeros, jv

    r, k, forward, _ = qdht(RADIUS, N_R)
    zeros = jn_zeros(0, N_R + 1)
    j1 = np.abs(jv(1, zeros[:-1]))
    bare = forward / (j1[:, None] / j1[None, :])
    f, exact = gaussian_pair(r, k)
    assert np.abs(bare @ f - exact).max() > 1e-3 * np.abs(exact).max()


def test_radial_record_uses_the_transform_pair():
    problem = radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02, fwhm=30e-15,
                             integration="trapezoid")
    problem.validate()
    r, k, forward, inverse = qdht(4.0e-3, 32)
    assert np.array_equal(problem.synthesis, inverse)
    assert np.array_equal(problem.projection, forward)
    assert problem.n_modes == problem.n_nodes == 32
    # The launch spectrum vanishes outside the band, and the linear operator and
    # prefactor vanish on inactive bins.
    field = problem.field0.reshape(problem.shape)
    assert not np.any(field[..., ~problem.active_mask])
    assert not np.any(problem.phi_terms[..., ~problem.active_mask])
    assert not np.any(problem.prefactor_terms[..., ~problem.active_mask])
    # Evanescent components decay: Re Φ <= 0 everywhere.
    assert np.all(problem.phi_terms.real <= 0)


# ---------------------------------------------------------------- transverse collar


def test_collar_profile_reference_values_and_symmetry():
    """The reference's docstring values (``spatial_absorber.py:128-131``), the Planck
    edge's point symmetry P(s) + P(1 - s) = 1 about the midpoint, and a profile rising
    monotonically from 0 to 1 (a flipped exponent falls instead)."""
    from solprop_amd.problem import collar_profile

    assert collar_profile(np.array([-1.0, 0.0, 0.5, 1.0, 2.0])).tolist() == [
        0.0, 0.0, 0.5, 1.0, 1.0]
    s = np.linspace(0.001, 0.999, 999)
    p = collar_profile(s)
    assert np.abs(p + collar_profile(1.0 - s) - 1.0).max() < 1e-15
    assert np.all(np.diff(p) >= 0.0)
    middle = collar_profile(np.linspace(0.1, 0.9, 81))
    assert np.all(np.diff(middle) > 0.0)
    # Every derivative continuous: the profile leaves zero flatter than any power.
    assert collar_profile(np.array([0.01]))[0] < 1e-40


def test_collar_matrix_is_the_real_space_product():
    """``F diag(-alpha/2) B`` applied to the Hankel spectrum of a smooth profile equals
    the spectrum of ``-alpha(r) f(r) / 2`` (the rate acts pointwise at the nodes); a
    constant rate gives ``-alpha/2`` times the identity to the transform's
    near-orthogonality; the rate on the wrong axis, or without its half, fails."""
    from solprop_amd.problem import radial_absorber_matrix, radial_collar_rate

    r, k, forward, inverse = qdht(RADIUS, N_R)
    rate = radial_collar_rate(r, RADIUS, 1.5e-3, 800.0)
    # A 1.5 mm Gaussian: e^-7 at the aperture, so the collar has content to act on (the
    # identity rests on B F = I, not on the profile vanishing at the edge).
    f = np.exp(-((r / 1.5e-3) ** 2)) * (1.0 + r / RADIUS)
    spectrum = forward @ f
    m = radial_absorber_matrix(forward, inverse, rate)
    exact = forward @ (-0.5 * rate * f)
    assert np.abs(m @ spectrum - exact).max() < 1e-10 * np.abs(exact).max()
    constant = radial_absorber_matrix(forward, inverse, np.full(N_R, 800.0))
    assert np.abs(constant + 400.0 * np.eye(N_R)).max() < 400.0 * 1e-9
    # Plausible mistakes: the rate applied on the spectral axis, or without the half.
    spectral = (-0.5 * rate)[:, None] * np.eye(N_R)
    assert np.abs(spectral @ spectrum - exact).max() > 1e-3 * np.abs(exact).max()
    no_half = radial_absorber_matrix(forward, inverse, 2.0 * rate)
    assert np.abs(no_half @ spectrum - exact).max() > 0.5 * np.abs(exact).max()


def test_radial_record_carries_the_collar():
    problem = radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02, fwhm=30e-15,
                             integration="trapezoid", collar=(1e-3, 2e3))
    problem.validate()
    assert problem.spatial_absorber.shape == (32, 32)
    # The rate is non-negative, so the matrix only takes energy: its symmetric part,
    # in the transform's weighted inner product, is negative semidefinite; here its
    # diagonal is non-positive.
    assert np.all(np.diag(problem.spatial_absorber) <= 1e-12)
