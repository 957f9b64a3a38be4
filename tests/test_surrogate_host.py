"""The surrogate's host side (no GPU): the scan's inputs, units and objective.

- The launch spectra it forms for screening equal the scan problems' own launch fields,
  to the bit (the network sees at screening exactly what it saw in training).
- Parameters map onto [0, 1]; the log-spectral units invert; the UV share of a spectrum
  whose energy sits in one band is 1 or 0; the nearest-neighbour baseline returns a
  training row's own target for that row's parameters.
"""

import numpy as np

from solprop_amd import surrogate as s


def test_launch_spectra_equal_the_problems_own():
    pts = s.draw(5, np.random.default_rng(4))
    fields = s.launch_fields(pts[:, 2], pts[:, 1])
    for i in range(5):
        problem = s.scan_problem(*pts[i])
        assert np.array_equal(problem.field0[0], fields[i])
        peak = float(np.max(np.abs(problem.field0) ** 2))
        assert np.array_equal(s.log_spectrum(problem.field0, peak),
                              s.launch_log_spectra(pts[i:i + 1, 2])[0])


def test_units_and_objective():
    pts = s.draw(1000, np.random.default_rng(5))
    x = s.normalise(pts)
    assert x.min() >= 0.0 and x.max() <= 1.0
    spec = np.random.default_rng(6).uniform(0.0, 1.0, (3, s.BINS)).astype(np.float32)
    assert np.allclose(s.density(spec), np.maximum(10.0 ** (8.0 * spec - 8.0) - s.FLOOR, 0.0))
    f = np.arange(s.BINS) / s.GRID["window"]
    c = 299792458.0
    uv = (f >= c / 300e-9) & (f <= c / 150e-9)
    ir = (f >= c / 1000e-9) & (f <= c / 700e-9)
    only_uv = np.where(uv, 1.0, 0.0)
    only_ir = np.where(ir, 1.0, 0.0)
    assert np.allclose(s.uv_fraction(only_uv), 1.0)
    assert np.allclose(s.uv_fraction(only_ir), 0.0)
    assert np.allclose(s.spectral_error(spec, spec), 0.0)


def test_nearest_neighbour_returns_the_training_row():
    rng = np.random.default_rng(7)
    params = rng.uniform(0, 1, (50, 4)).astype(np.float32)
    targets = rng.uniform(0, 1, (50, 16)).astype(np.float32)
    got = s.nearest_neighbour(params, targets, params[[3, 17, 42]])
    assert np.array_equal(got, targets[[3, 17, 42]])


def test_reached_allows_the_landing_tolerance():
    z = np.array([1.0, 1.0 - 4 * np.spacing(1.0), 0.99, 1.1])
    assert list(s.reached(z, np.ones(4))) == [True, True, False, True]
