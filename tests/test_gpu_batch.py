"""GPU: the batched advance of independent instances (a scan's points) through the C
ABI. Each instance's march under ``sp_batch_advance`` must be the one its own
``sp_workspace_advance`` makes: the same attempts and the same end state, to the bit,
whatever the other instances do on the GPU meanwhile (each has its own stream and
graphs; the decisions are the host's, per instance).
"""

import numpy as np
import pytest

from solprop_amd.native import Workspace, advance_batch
from solprop_amd.problem import capillary_problem, gnlse_problem, radial_problem

pytestmark = pytest.mark.gpu


def instances():
    """Points of a scan over energy and pressure on the capillary route, two envelope
    points and a radial one: different shapes and routes in one batch."""
    out = [capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid",
                             energy=e, p0=p)
           for e, p in [(275e-6, 0.8), (200e-6, 0.8), (275e-6, 0.5), (350e-6, 0.6)]]
    out += [gnlse_problem(n_t=512, ffac=2, window=2e-12, fwhm=50e-15, soliton_order=s,
                          length=5e-3) for s in (2.0, 3.0)]
    # The dense route: its GEMMs run on its own stream beside the others' kernels.
    out.append(radial_problem(n_t=1024, window=0.25e-12, n_r=32, length=0.02, fwhm=30e-15,
                              w0=0.25e-3, integration="trapezoid"))
    return out


def test_batch_matches_instances_run_alone():
    print("batched advance of 7 instances", flush=True)
    problems = instances()
    limit = 40
    alone = []
    for p in problems:
        ws = Workspace(p)
        session = ws.march(0.0, p.field0)
        batch = session.advance(p.z_end, work_limit=limit)
        alone.append((batch.attempts.copy(), session.state()))
        ws.close()
    spaces = [Workspace(p) for p in problems]
    sessions = [ws.march(0.0, p.field0) for ws, p in zip(spaces, problems)]
    batches = advance_batch(sessions, [p.z_end for p in problems], work_limit=limit)
    together = [(b.attempts.copy(), s.state()) for b, s in zip(batches, sessions)]
    for ws in spaces:
        ws.close()
    for (rows_a, state_a), (rows_b, state_b) in zip(alone, together):
        assert len(rows_a) > 0
        assert np.array_equal(rows_a, rows_b)
        assert np.array_equal(state_a, state_b)


@pytest.mark.filterwarnings("ignore:invalid value encountered in divide:RuntimeWarning")
@pytest.mark.parametrize("route", ["capillary", "envelope"])
def test_zero_field_is_refused_as_the_oracle_refuses_it(route):
    """An empty launch field (no light at all): the scaled norm is zero over a zero scale,
    every attempt is rejected until the step would fall below the floor, and the march
    stops with a convergence error; the device stops the same way, after the same
    number of attempts."""
    from oracle import adams, rk
    from oracle.controller import ConvergenceError as OracleConvergence
    from solprop_amd.native import ConvergenceError

    print(f"zero field {route}", flush=True)
    if route == "capillary":
        problem = capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")
    else:
        problem = gnlse_problem(n_t=512, ffac=2, window=2e-12, fwhm=50e-15, length=5e-3,
                                noise_seed=None)
    zero = np.zeros_like(problem.field0)
    y0 = zero.reshape(problem.shape)
    seen = []
    march = adams.march if problem.method == "adams" else rk.march
    with pytest.raises(OracleConvergence):
        march(problem, y0.copy(), max_attempts=12, on_attempt=lambda *a: seen.append(a))
    ws = Workspace(problem)
    session = ws.march(0.0, zero)
    with pytest.raises(ConvergenceError):
        session.advance(problem.z_end, work_limit=12)
    counters = ws.counters()
    ws.close()
    # Zero over a zero scale: the norms are not numbers, every attempt is rejected and
    # the step shrinks to the floor, on both sides after the same attempts.
    # The Adams march's last attempt, taken at the floor's fixed step and rejected, raises
    # inside its decision before its row is recorded (the Runge-Kutta march records it
    # first); the workspace took that trial.
    last = 1 if problem.method == "adams" else 0
    assert len(seen) >= 1 and counters["trials"] == len(seen) + last
