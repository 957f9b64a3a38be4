"""GPU: the free-space Cartesian route against the oracle.

The device marches a Kerr beam on an (ny, nx, nt) envelope grid by the low-storage RK4IP
replay, each source evaluation two rank-3 transforms of three axis passes each around
the pointwise Kerr response. Held to ``oracle/cartesian.py`` (pinned by the exact
linear flow, the separable source, the uniform beam's reduction to the envelope route
and the plane wave's closed form at fourth order):

- the source at the launch within 1e-13 of the oracle's, on grids with nx != ny;
- the replay's end state within 1e-12, its rows identical, its stiffness indicator
  within 1e-10;
- a linear problem's replay is the exact flow; graph replays equal direct launches to
  the bit; a probe step leaves the state;
- at 256 x 256 x 512, a transversely uniform beam equals nx ny times the envelope
  route's oracle, every other column zero to rounding;
- a session over the route observes its planes.
"""

import dataclasses

import numpy as np
import pytest

from oracle import cartesian as cart
from oracle import low_storage as ls
from solprop_amd.native import Workspace
from solprop_amd.problem import cartesian_problem
from test_cartesian import envelope_twin, uniform_beam

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def problem_of(nx, ny, nt, **kw):
    base = dict(nx=nx, ny=ny, n_t=nt, spacing=96e-6 / nx, waist=20e-6, energy=1e-7,
                length=4e-4, steps=6)
    base.update(kw)
    return cartesian_problem(**base)


def device_replay(problem, y0=None):
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0 if y0 is None else y0)
    batch = march.advance(problem.z_end, work_limit=100000)
    state = march.state()
    stiffness = float(ws.buffer("stiffness")[0])
    ws.close()
    return state, batch, stiffness


@pytest.mark.parametrize("grid", [(16, 8, 32), (64, 32, 128)])
def test_source_matches_oracle(grid):
    print("Cartesian source vs oracle", grid, flush=True)
    problem = problem_of(*grid)
    ws = Workspace(problem)
    got = ws.source(0.0, problem.field0)
    ws.close()
    expected = cart.source(problem, problem.field0)
    err = rel(got, expected)
    print("source rel L2", err, flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("grid", [(16, 8, 32), (64, 32, 128), (32, 64, 256)])
def test_replay_matches_oracle(grid):
    print("Cartesian replay vs oracle", grid, flush=True)
    problem = problem_of(*grid)
    expected, rows, stiffness = cart.replay(problem, problem.field0)
    state, batch, h_rho = device_replay(problem)
    err = rel(state, expected)
    print("end-state rel L2", err, "stiffness", h_rho, stiffness[-1], flush=True)
    assert batch.done
    assert np.array_equal(batch.attempts, np.array(rows))
    assert err < 1e-12
    assert abs(h_rho - stiffness[-1]) <= 1e-10 * abs(stiffness[-1])
    linear = np.exp(problem.z_end * problem.phi_terms[0]) * problem.field0
    assert rel(state, linear) > 1e-4  # the Kerr term acted


def test_linear_replay_is_the_exact_flow():
    problem = problem_of(32, 16, 64, kerr=False)
    state, _, _ = device_replay(problem)
    exact = np.exp(problem.z_end * problem.phi_terms[0]) * problem.field0
    assert rel(state, exact) < 1e-13


def test_graphs_and_direct_launches_agree(monkeypatch):
    problem = problem_of(32, 16, 64)
    graphed, _, _ = device_replay(problem)
    monkeypatch.setenv("SOLPROP_MI355X_GRAPH", "0")
    direct, _, _ = device_replay(problem)
    assert np.array_equal(graphed, direct)


def test_probe_step_keeps_the_state():
    problem = problem_of(32, 16, 64)
    expected, _, _ = cart.replay(problem, problem.field0, problem.schedule[:2])
    ws = Workspace(problem)
    probe = ws.probe(0.0, problem.field0)
    norm = probe.trial(0.0, problem.schedule[1])
    field = probe.field()
    state = ws.state()
    ws.close()
    assert norm == 0.0
    assert rel(field, expected) < 1e-12
    assert np.array_equal(state, problem.field0)


def test_uniform_beam_at_size():
    """256 x 256 x 512 (33.5 M points), four steps: the transverse passes carry a
    uniform beam, which is the envelope route's march scaled by nx ny."""
    print("Cartesian uniform beam at 256 x 256 x 512", flush=True)
    problem = cartesian_problem(nx=256, ny=256, n_t=512, spacing=1.05e-6, steps=4,
                                energy=1e-7)
    y0 = uniform_beam(problem)
    state, batch, _ = device_replay(problem, y0)
    twin = envelope_twin(problem, y0[0] / (256 * 256))
    expected, _, _ = ls.replay(twin, twin.field0.reshape(twin.shape), problem.schedule)
    err = rel(state[0] / (256 * 256), expected.reshape(-1))
    rest = np.abs(state[1:]).max() / np.abs(state[0]).max()
    print("column 0 rel L2", err, "other columns", rest, flush=True)
    assert len(batch.attempts) == 4
    assert err < 1e-12
    assert rest < 1e-12


def test_session_over_the_route():
    problem = problem_of(32, 16, 64, steps=10)
    planes = [problem.schedule[3], problem.z_end]
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": planes})
    rows, observed = [], []
    while True:
        batch = session.advance(100000, 60.0)
        rows.append(batch.attempts)
        if batch.event.kind_name == "terminal":
            break
        if batch.event.observation_name == "plane":
            observed.append(batch.event.z)
        session.acknowledge(batch.sequence)
    session.close()
    ws.close()
    rows = np.concatenate(rows)
    assert observed == planes
    assert len(rows) == 10 and np.all(rows[:, 4] == 4)


def test_refusals_before_allocation():
    problem = problem_of(16, 8, 32)
    with pytest.raises(ValueError):
        Workspace(dataclasses.replace(problem, transverse=(16, 16)))
