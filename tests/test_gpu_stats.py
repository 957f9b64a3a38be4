"""GPU parity of the plane statistics against the oracle, through the C ABI.

The committed state after a short march on each route, and the launch field: the
spectral centroid, the energy and the energy inside two windows (sums in a fixed order:
within 1e-13 relative), the peak power of each column and of the total (an analytic
signal through one FP64 transform of n_t points: within 1e-12 relative).
"""

import numpy as np
import pytest

from oracle import stats as S
from solprop_amd.native import Workspace
from solprop_amd.problem import capillary_problem, gnlse_problem, radial_problem

pytestmark = pytest.mark.gpu

NAMES = ["omega0", "energy", "energy_window", "peak_power", "peak_power_total"]

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
        "peak_power_total": S.peak_power_total(field, kind),
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
    session = ws.start(0.0, problem.field0)
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
    for name in ("peak_power", "peak_power_total"):
        err = np.abs(np.asarray(got[name]) - expected[name]).max()
        assert err <= 1e-12 * np.abs(expected[name]).max(), (name, err)
    # The windows tile the active band: their energies sum to the energy there.
    assert np.allclose(got["energy_window"].sum(axis=0), got["energy"], rtol=1e-12, atol=0)
