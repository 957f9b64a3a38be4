"""GPU parity of the plane statistics against the oracle, through the C ABI.

The committed state after a short march on each route, and the launch field: the
spectral centroid, the energy and the energy inside two windows (sums in a fixed order:
within 1e-13 relative), the peak power of each column and of the total, and the total
power trace the width is measured on (an analytic signal through one FP64 transform of
n_t points: within 1e-12 of the largest value).
"""

import numpy as np
import pytest

from oracle import stats as S
from solprop_amd.native import Workspace
from solprop_amd.problem import capillary_problem, gnlse_problem, radial_problem

pytestmark = pytest.mark.gpu

NAMES = ["omega0", "energy", "energy_window", "peak_power", "peak_power_window",
         "peak_power_total", "fwhm"]

PROBLEMS = {
    "capillary": lambda: capillary_problem(n_t=2048, n_nodes=32, length=0.05,
                                           integration="trapezoid"),
    "radial": lambda: radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02,
                                     fwhm=30e-15, integration="trapezoid"),
    "envelope": lambda: gnlse_problem(n_t=512, ffac=2, window=2e-12, fwhm=50e-15,
                                      soliton_order=2.0, length=0.02),
}


def oracle_statistics(problem, state, windows):
    kind = problem.kind
    dt = problem.dt_fine * problem.n_to / problem.n_t
    field = state.reshape(problem.n_modes, problem.n_omega)
    return {
        "omega0": S.omega0(field, problem.omega),
        "energy": S.energy(field, dt, kind),
        "energy_window": S.energy_window(field, problem.omega, windows, dt, kind),
        "peak_power": S.peak_power(field, kind),
        "peak_power_window": S.peak_power_window(field, problem.omega, windows, kind),
        "peak_power_total": S.peak_power_total(field, kind),
        "fwhm": S.power_trace(field, kind),
    }


@pytest.mark.parametrize("route", sorted(PROBLEMS))
@pytest.mark.parametrize("marched", [False, True])
def test_statistics_match_oracle(route, marched):
    print(f"statistics {route} marched={marched}", flush=True)
    problem = PROBLEMS[route]()
    w = problem.omega[problem.active_mask]
    middle = 0.5 * (w.min() + w.max())
    windows = np.array([[w.min(), middle], [middle, np.inf]])
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    if marched:
        session.advance(problem.z_end, work_limit=6)
    state = session.state()
    got = ws.statistics(NAMES, windows)
    ws.close()
    expected = oracle_statistics(problem, state, windows)
    # A column with no energy has no centroid: NaN on both sides, in the same columns.
    empty = np.isnan(expected["omega0"])
    assert np.array_equal(np.isnan(got["omega0"]), empty)
    got["omega0"], expected["omega0"] = got["omega0"][~empty], expected["omega0"][~empty]
    for name in ("omega0", "energy", "energy_window"):
        err = np.abs(np.asarray(got[name]) - expected[name]).max()
        assert err <= 1e-13 * np.abs(expected[name]).max(), (name, err)
    for name in ("peak_power", "peak_power_window", "peak_power_total", "fwhm"):
        err = np.abs(np.asarray(got[name]) - expected[name]).max()
        assert err <= 1e-12 * np.abs(expected[name]).max(), (name, err)
    # The windows tile the active band: their energies sum to the energy there.
    assert np.allclose(got["energy_window"].sum(axis=0), got["energy"], rtol=1e-12, atol=0)


def monitor_weights(problem):
    """A monitor's rows on the coarse axis: guards on the outer tenth each side, the
    collar's complement, and the split collar's rate (or a ramp without one)."""
    n_t = problem.n_t
    t = np.arange(n_t) - n_t / 2
    low = (t < -0.4 * n_t).astype(float)
    high = (t > 0.4 * n_t).astype(float)
    collar = np.clip((np.abs(t) - 0.3 * n_t) / (0.2 * n_t), 0.0, 1.0)
    rate = (problem.absorber_rate if problem.absorber_scheme == "split_step"
            else 100.0 * collar)
    return np.ascontiguousarray(np.stack([low, high, collar, rate]))


@pytest.mark.parametrize("route", sorted(PROBLEMS))
def test_monitor_reading_matches_oracle(route):
    """The five energies of the committed state, within 1e-12 of the total."""
    print(f"monitor reading {route}", flush=True)
    problem = PROBLEMS[route]()
    weights = monitor_weights(problem)
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.advance(problem.z_end, work_limit=4)
    ws.bind_monitor(weights)
    state = session.state()
    got = ws.monitor_reading()
    ws.close()
    field = state.reshape(problem.n_modes, problem.n_omega)
    expected = S.monitor_energies(field, weights, problem.kind)
    values = np.concatenate([[got["total"]], got["weighted"]])
    assert np.abs(values - expected).max() <= 1e-12 * expected[0]
    assert got["taken"] == 0.0


SPLIT = {
    "capillary": lambda: capillary_problem(n_t=2048, n_nodes=32, length=0.05,
                                           integration="trapezoid", absorber="split_step",
                                           absorber_length=2e-2),
    "envelope": lambda: gnlse_problem(n_t=512, ffac=2, window=2e-12, fwhm=50e-15,
                                      soliton_order=2.0, length=5e-3, absorber="split_step",
                                      absorber_length=1e-3),
}


@pytest.mark.parametrize("route", sorted(SPLIT))
def test_split_sums_follow_the_oracle_march(route):
    """Under the split-step absorber with a bound monitor, each accepted step's finish
    adds what its split takes and keeps what it finds: the reading after a short march
    holds the sum over the oracle march's steps (from each masked state before its
    split) and the last step's found energies, within the marches' own agreement
    (1e-7); a second reading has nothing taken."""
    from oracle import adams, rk

    print(f"split sums {route}", flush=True)
    problem = SPLIT[route]()
    weights = monitor_weights(problem)
    finishes = []
    y0 = problem.field0.reshape(problem.shape)
    if problem.method == "adams":
        _, records = adams.march(problem, y0.copy(), max_attempts=12,
                                 on_finish=lambda s, dz: finishes.append((s.copy(), dz)))
    else:
        _, records = rk.march(problem, y0.copy(), max_attempts=12,
                              on_finish=lambda s, dz: finishes.append((s.copy(), dz)))
    sums = [S.split_sums(s.reshape(problem.n_modes, problem.n_omega), weights, dz,
                         problem.kind) for s, dz in finishes]
    ws = Workspace(problem)
    ws.load(0.0, problem.field0)
    ws.bind_monitor(weights)
    session = ws.march(0.0, problem.field0)
    rows = session.advance(problem.z_end, work_limit=len(records)).attempts
    first = ws.monitor_reading()
    second = ws.monitor_reading()
    ws.close()
    assert len(rows) == len(records) and len(finishes) > 2
    taken = sum(v[0] for v in sums)
    assert taken > 0.0
    assert abs(first["taken"] - taken) <= 1e-7 * taken
    assert np.allclose(first["found"], sums[-1][1:], rtol=1e-7, atol=1e-7 * sums[-1][1])
    assert second["taken"] == 0.0


@pytest.mark.parametrize("route", sorted(PROBLEMS))
def test_peak_intensity_matches_oracle(route):
    """The on-axis synthesis's peak (one row, and two rows) after a short march, with
    every other statistic: within 1e-12 relative, and packed between the plasma pair and
    the power trace."""
    print(f"peak intensity {route}", flush=True)
    problem = PROBLEMS[route]()
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.advance(problem.z_end, work_limit=4)
    state = session.state()
    c = problem.n_modes
    rng = np.random.default_rng(5)
    field = state.reshape(c, problem.n_omega)
    for rows in (rng.uniform(-1, 1, (1, c)), rng.uniform(-1, 1, (2, c))):
        got = ws.statistics(["peak_power_total", "peak_intensity", "fwhm"], on_axis=rows,
                            intensity_factor=1.7)
        expected = S.peak_intensity(field, problem.kind, rows, 1.7)
        assert abs(got["peak_intensity"] - expected) <= 1e-12 * expected
        assert np.allclose(got["fwhm"], S.power_trace(field, problem.kind), rtol=0,
                           atol=1e-12 * np.max(got["fwhm"]))
    with pytest.raises(Exception, match="on-axis"):
        ws.statistics(["peak_intensity"])
    ws.close()


def test_populations_match_oracle():
    """Every node's end fraction and largest rate after a short march, with the plasma
    pair: the rates to 1e-12, the fractions to 1e-9 (as the electron density)."""
    print("populations capillary", flush=True)
    problem = capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.advance(problem.z_end, work_limit=4)
    state, z = session.state(), session.z
    got = ws.statistics(["electron_density", "peak_ionisation_rate", "fwhm", "populations"])
    ws.close()
    expected = S.populations(problem, state, z)
    pops = got["populations"]
    assert pops.shape == (32, 2)
    assert np.allclose(pops[:, 1], expected[:, 1], rtol=1e-12, atol=0)
    assert np.allclose(pops[:, 0], expected[:, 0], rtol=1e-9, atol=1e-300)
    densest = int(np.argmax(pops[:, 0]))
    assert got["peak_ionisation_rate"] == pops[densest, 1]


@pytest.mark.parametrize("route", ["capillary", "radial"])
def test_plasma_statistics_match_oracle(route):
    """The electron density and the peak ionisation rate at the committed state after a
    short march, against the oracle's scan of the same state: the rate to 1e-12, the
    density (a fraction 1 - (1 - f0) e^{-X} that cancels at small exposure) to 1e-9."""
    print(f"plasma statistics {route}", flush=True)
    # The trapezoid rule on one route and the band rule on the other.
    if route == "capillary":
        problem = capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")
    else:
        problem = radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02,
                                 fwhm=30e-15, w0=0.25e-3, integration="band",
                                 band=(400e-9, 4e-6))
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.advance(problem.z_end, work_limit=4)
    state, z = session.state(), session.z
    got = ws.statistics(["electron_density", "peak_ionisation_rate"])
    ws.close()
    density, rate = S.ionisation(problem, state, z)
    assert density > 0.0 and rate > 0.0
    assert abs(got["electron_density"] - density) <= 1e-9 * density
    assert abs(got["peak_ionisation_rate"] - rate) <= 1e-12 * rate


@pytest.mark.parametrize("route", ["capillary_trapezoid", "capillary_band", "radial"])
def test_plasma_integration_diagnostics_match_oracle(route):
    """The column's six integration-diagnostic maxima at the committed state after a short
    march, against the oracle's over the same state: the exposure increments' dip and rise
    within 1e-9 of the rise (their differences of a running sum round at eps times it), the
    electrons freed within 1e-9 and the work defect within 1e-9 of them (a difference of
    two nearly equal numbers), the free-current integrand's levels within 1e-9 of its peak.
    """
    print(f"plasma integration diagnostics {route}", flush=True)
    if route == "capillary_trapezoid":
        problem = capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")
    elif route == "capillary_band":
        problem = capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="band",
                                    band=(300e-9, 4e-6))
    else:
        problem = radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02,
                                 fwhm=30e-15, w0=0.25e-3, integration="band",
                                 band=(400e-9, 4e-6))
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.advance(problem.z_end, work_limit=4)
    state, z = session.state(), session.z
    got = ws.statistics(["plasma_integration", "populations"])
    ws.close()
    parts = np.asarray(got["plasma_integration"])
    expected = S.plasma_integration_parts(problem, state, z)
    print("device", parts.tolist(), "oracle", expected.tolist(), flush=True)
    assert parts.shape == (6,) and expected[1] > 0.0 and expected[3] > 0.0
    rise, freed, peak = expected[1], expected[3], expected[5]
    assert abs(parts[0] - expected[0]) <= 1e-9 * rise
    assert abs(parts[1] - rise) <= 1e-9 * rise
    assert abs(parts[3] - freed) <= 1e-9 * freed
    assert abs(parts[2] - expected[2]) <= 1e-9 * freed
    assert abs(parts[4] - expected[4]) <= 1e-9 * peak
    assert abs(parts[5] - peak) <= 1e-9 * peak


@pytest.mark.parametrize("route", ["kerr", "plasma_band"])
def test_quadrature_estimate_matches_oracle(route):
    """The quadrature estimate after a march lands on a plane: the last accepted trial's
    source evaluation (its field, read back from the device, at the plane) repeated by
    the oracle with the Kronrod and the embedded Gauss projections. The estimate is a
    relative difference near 1e-2 of two spectra that agree with the oracle's to about
    1e-12, so it agrees to 1e-7 of itself."""
    print(f"quadrature estimate {route}", flush=True)
    if route == "kerr":
        problem = capillary_problem(n_t=2048, n_nodes=16, n_modes=6, length=0.05,
                                    plasma=False, integration="trapezoid",
                                    radial_rule="gauss_kronrod")
    else:
        problem = capillary_problem(n_t=2048, n_nodes=16, n_modes=6, length=0.05,
                                    integration="band", band=(300e-9, 4e-6),
                                    radial_rule="gauss_kronrod")
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    assert np.isnan(ws.statistics(["quadrature_error"])["quadrature_error"])  # no step yet
    session.run(0.004, work_limit=64)
    assert session.z == 0.004
    got = ws.statistics(["peak_intensity", "quadrature_error", "energy"],
                        on_axis=problem.synthesis[:1])
    lab = ws.buffer("lab").reshape(problem.n_modes, problem.n_omega)
    ws.close()
    expected = S.quadrature_error(problem, lab, 0.004)
    print("device", got["quadrature_error"], "oracle", expected, flush=True)
    assert 1e-4 < expected < 1.0
    assert abs(got["quadrature_error"] - expected) <= 1e-7 * expected
    assert got["energy"].shape == (problem.n_modes,)


def test_quadrature_estimate_without_an_embedded_rule_is_nan():
    problem = capillary_problem(n_t=2048, n_nodes=16, length=0.05, plasma=False,
                                integration="trapezoid")
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    session.run(0.002, work_limit=64)
    value = ws.statistics(["quadrature_error"])["quadrature_error"]
    ws.close()
    assert np.isnan(value)
