"""The Cartesian route's oracle and host logic (CPU).

The oracle (``oracle/cartesian.py``) is pinned here by what fixes it independently of
its own formulas:

- a linear problem's replay is the exact flow ``exp(D L) Y0``;
- a separable field's source is the product of the axes' transforms of the cubed
  factors (the column layout ``iy nx + ix`` and the 1/V normalisation);
- a transversely uniform beam is the envelope route's oracle (``oracle/rk.py`` through
  ``oracle/low_storage.py``, itself pinned by P2) scaled by nx ny;
- the plane wave's closed form, a phase rotation at a + b g s |y|^2 / V^2, reached at
  fourth order in the step.

And the problem's refusals (the grid's powers of two, the single tables, the method).
"""

import dataclasses

import numpy as np
import pytest

from oracle import cartesian as cart
from oracle import low_storage as ls
from solprop_amd.problem import KerrGroup, Problem, ScalarPath, cartesian_problem


def small(**kw):
    base = dict(nx=16, ny=8, n_t=32, spacing=6e-6, waist=20e-6, energy=2e-5,
                length=4e-4, steps=6)
    base.update(kw)
    return cartesian_problem(**base)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_builder_shapes_and_validation():
    p = small()
    p.validate()
    assert p.phi_terms.shape == (1, 128, 32) and p.field0.shape == (128, 32)
    assert p.transverse == (16, 8)
    # The launch spectrum is the transform of a centred Gaussian in (y, x, t).
    e = np.fft.ifftn(p.field0.reshape(8, 16, 32))
    peak = np.unravel_index(np.argmax(np.abs(e)), e.shape)
    assert peak == (4, 8, 16)


@pytest.mark.parametrize("change, message", [
    (dict(transverse=(12, 8)), "power of two"),
    (dict(transverse=(16, 16)), "columns"),
    (dict(method="rk"), "low-storage"),
    (dict(schedule=None), "low-storage"),
    (dict(phi_paths=(ScalarPath.constant(1.0),)), "W\\(z\\) = z"),
    (dict(kerr=KerrGroup(1.0, "real")), "analytic"),
    (dict(absorber_scheme="split_step"), "absorbers"),
])
def test_refusals(change, message):
    p = dataclasses.replace(small(), **change)
    with pytest.raises(ValueError, match=message):
        p.validate()


def test_transverse_grid_refused_elsewhere():
    from solprop_amd.problem import gnlse_problem
    p = dataclasses.replace(gnlse_problem(n_t=256), transverse=(4, 4))
    with pytest.raises(ValueError, match="Cartesian"):
        p.validate()


def test_linear_replay_is_the_exact_flow():
    p = small(kerr=False)
    state, rows, _ = cart.replay(p, p.field0)
    exact = np.exp(p.z_end * p.phi_terms[0]) * p.field0
    assert rel(state, exact) < 1e-13
    assert [r[3] for r in rows] == [1.0] * 6
    assert np.allclose([r[0] for r in rows], p.schedule[1:], rtol=0, atol=0)


def test_separable_source_is_the_product_of_axis_transforms():
    """E = g(y) f(x) h(t): |E|^2 E = (|g|^2 g)(|f|^2 f)(|h|^2 h), so its fftn is the outer
    product of three 1-D transforms; a transposed layout or a wrong scale fails."""
    rng = np.random.default_rng(3)
    nx, ny, nt = 8, 4, 16
    p = small(nx=nx, ny=ny, n_t=nt)
    f, g, h = (rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in (nx, ny, nt))
    outer = lambda a, b, c: (a[:, None, None] * b[None, :, None] * c[None, None, :])
    y = np.fft.fft(g)[:, None, None] * np.fft.fft(f)[None, :, None] * np.fft.fft(h)[None, None, :]
    got = cart.source(p, y.reshape(nx * ny, nt))
    k = p.kerr
    cube = lambda a: np.abs(a) ** 2 * a
    expected = (cart.gamma(p) * (k.self_factor + k.conjugate_factor)
                * outer(np.fft.fft(cube(g)), np.fft.fft(cube(f)), np.fft.fft(cube(h))))
    expected = p.prefactor_terms[0] * expected.reshape(nx * ny, nt)
    assert rel(got, expected) < 1e-13


def envelope_twin(p, column0):
    """The envelope-route problem of the Cartesian problem's column 0."""
    nt = p.n_t
    return Problem(
        kind="envelope", n_t=nt, n_to=nt, n_modes=1, n_nodes=1,
        crop_index=np.arange(nt, dtype=np.int64), active_mask=np.ones(nt, bool),
        norm_mask=p.norm_mask.copy(), omega_window=p.omega_window, omega=p.omega,
        time_window=np.ones(nt), time_window_derivative=np.zeros(nt), dt_fine=p.dt_fine,
        z_end=p.z_end, phi_terms=p.phi_terms[:, :1].copy(), phi_paths=p.phi_paths,
        prefactor_terms=p.prefactor_terms[:, :1].copy(),
        prefactor_paths=p.prefactor_paths, synthesis=np.ones((1, 1)),
        projection=np.ones((1, 1)), density=p.density, preionisation=p.preionisation,
        kerr=p.kerr, field0=column0.reshape(1, nt), method="low_storage",
        schedule=p.schedule)


def uniform_beam(p):
    """The launch's time profile, uniform across the grid (column 0 only)."""
    nx, ny = p.transverse
    y0 = np.zeros_like(p.field0)
    y0[0] = p.field0[0] / np.abs(p.field0[0]).max() * 2e9 * nx * ny
    return y0


def test_uniform_beam_is_the_envelope_route():
    p = small()
    nx, ny = p.transverse
    y0 = uniform_beam(p)
    state, _, stiff = cart.replay(p, y0)
    twin = envelope_twin(p, y0[0] / (nx * ny))
    expected, _, stiff1 = ls.replay(twin, twin.field0.reshape(twin.shape), p.schedule)
    assert rel(state[0] / (nx * ny), expected.reshape(-1)) < 1e-12
    assert np.abs(state[1:]).max() <= 1e-12 * np.abs(state[0]).max()
    # The Kerr phase is no rounding: against the linear flow alone the march moved.
    linear = np.exp(p.z_end * p.phi_terms[0, 0]) * y0[0]
    assert rel(state[0], linear) > 1e-3
    assert np.allclose(stiff, stiff1, rtol=1e-9)


def plane_wave(steps, a=3.0, b=2.0, amplitude=1.0):
    """One bin (the DC bin) of a 4 x 4 x 4 grid, D = i a and p = i b everywhere."""
    p = cartesian_problem(nx=4, ny=4, n_t=4, length=1.0, steps=steps)
    v = 64.0
    y0 = np.zeros((16, 4), np.complex128)
    y0[0, 0] = amplitude * v
    kerr = KerrGroup(1.0, "analytic", self_factor=0.75, conjugate_factor=0.25)
    return dataclasses.replace(
        p, phi_terms=np.full((1, 16, 4), 1j * a), prefactor_terms=np.full((1, 16, 4), 1j * b),
        kerr=kerr, field0=y0, norm_mask=np.ones(4, bool)), y0


def test_plane_wave_closed_form_at_fourth_order():
    errors = []
    for steps in (10, 20, 40):
        p, y0 = plane_wave(steps)
        state, _, _ = cart.replay(p, y0)
        # dy/dz = i a y + i b g (s + c) |y|^2 y / V^2, |y| conserved.
        rate = 3.0 + 2.0 * 1.0 * (0.75 + 0.25) * 1.0
        exact = y0[0, 0] * np.exp(1j * rate * 1.0)
        assert np.abs(state).sum() - np.abs(state[0, 0]) == 0.0
        errors.append(abs(state[0, 0] - exact) / abs(exact))
    orders = np.log2(np.array(errors[:-1]) / np.array(errors[1:]))
    assert np.all(orders > 3.8) and np.all(orders < 4.3), orders


def test_oracle_catches_a_dropped_normalisation():
    """Planted on the CPU: a source with one 1/V too many (a forward transform taken
    normalised) departs from the plane wave's closed form at once."""
    p, y0 = plane_wave(40)
    good, _, _ = cart.replay(p, y0)
    orig = cart.source

    def wrong(problem, y):
        return orig(problem, y) / 64.0

    try:
        cart.source = wrong
        bad, _, _ = cart.replay(p, y0)
    finally:
        cart.source = orig
    exact = y0[0, 0] * np.exp(1j * 5.0)
    assert abs(good[0, 0] - exact) < 1e-5 * abs(exact)
    assert abs(bad[0, 0] - exact) > 1e-2 * abs(exact)
