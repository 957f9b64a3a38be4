"""The boundary refuses problems outside what it runs, before touching the device.

Refusals are the reference's ``BackendRefusal`` (``_core.pyi:23-28``,
``backend.py:1165-1199``): raised at workspace construction, before any allocation.
They run here without a GPU because the library checks the problem before its first
HIP call.
"""

import dataclasses

import numpy as np
import pytest

from solprop_amd.native import BackendRefusal, Workspace
from solprop_amd.problem import capillary_problem, gnlse_problem


def refused(problem, match):
    with pytest.raises(BackendRefusal, match=match):
        Workspace(problem)


def test_envelope_under_adams_is_refused():
    refused(dataclasses.replace(gnlse_problem(n_t=64, window=1e-12), method="adams"),
            "embedded Runge-Kutta")


def test_field_route_under_rk_is_refused():
    refused(dataclasses.replace(capillary_problem(n_t=256, n_nodes=4, integration="trapezoid"),
                                method="rk"),
            "exponential Adams")


def test_band_rule_beyond_eight_taps_a_side_is_refused():
    p = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid")
    plasma = dataclasses.replace(p.plasma, taps=np.full(10, 0.1))
    refused(dataclasses.replace(p, plasma=plasma), "9 \\(the band rule\\) taps")


def test_short_field_grid_is_refused():
    refused(capillary_problem(n_t=128, n_nodes=4, integration="trapezoid"),
            "power of two >= 256")


def test_envelope_with_several_modes_is_refused():
    p = gnlse_problem(n_t=64, window=1e-12)
    two = dataclasses.replace(
        p, n_modes=2, phi_terms=np.repeat(p.phi_terms, 2, axis=1),
        prefactor_terms=np.repeat(p.prefactor_terms, 2, axis=1),
        field0=np.repeat(p.field0, 2, axis=0), synthesis=np.ones((1, 2)),
        projection=np.ones((2, 1)))
    with pytest.raises((BackendRefusal, ValueError)):
        Workspace(two)


def test_transverse_collar_off_the_dense_route_is_refused():
    p = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid")
    refused(dataclasses.replace(p, spatial_absorber=np.zeros((p.n_modes, p.n_modes))),
            "dense route")


def test_unknown_statistics_are_refused_by_name():
    """A statistic the native flags do not name is refused before the library is asked;
    the refusal names it. Every flag of native/include/solprop_native/problem/
    problem.hpp:44-60 is built."""
    from solprop_amd.native import Workspace as W

    ws = W.__new__(W)  # no device: the check precedes every library call
    ws.problem = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid")
    with pytest.raises(BackendRefusal, match="walk_off"):
        W.statistics(ws, ["energy", "walk_off"])
    assert set(W.STATISTICS.values()) == {1 << k for k in range(13)}


def test_scan_refuses_a_point_it_cannot_batch():
    """A scan refuses an instance with a Raman response (or more than 6 modes, or an
    absorber) before allocating anything, naming the instance."""
    from solprop_amd.native import Scan

    ok = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid")
    raman = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid", raman_share=0.5)
    with pytest.raises(BackendRefusal, match="scan instance 1: .*Raman"):
        Scan([ok, raman])
    wide = capillary_problem(n_t=256, n_nodes=4, n_modes=7, integration="trapezoid")
    with pytest.raises(BackendRefusal, match="scan instance 0: .*1 to 6 modes"):
        Scan([wide])
    absorbing = capillary_problem(n_t=256, n_nodes=4, integration="trapezoid",
                                  absorber="in_stage")
    with pytest.raises(BackendRefusal, match="absorber"):
        Scan([ok, absorbing])
