"""GPU: the low-storage RK4IP replay on the field route against the oracle.

The device replays a fixed schedule with the four register passes of the reference's
low-storage engine (``numpy_ref/engines.py:236-520``): each step one captured graph of
four source evaluations and their passes, every step accepted, its row (end, step, 0.0,
1). Held to ``oracle/low_storage.py`` (pinned by the exact linear flow, the stiffness
indicator's closed form and fourth-order convergence on the analytic P2 rung):

- the replay's end state within 1e-12 of the oracle's on a capillary with Kerr and
  plasma, under either running-integral rule and at full size (measured 1.4e-14 after
  40 steps), its rows identical, its last step's stiffness indicator within 1e-10;
- a linear-only problem's replay is the exact flow;
- graph replays and direct launches agree to the bit; a probe step leaves the state;
- a session over the replay observes its planes and carries order 4 in its rows.
"""

import dataclasses

import numpy as np
import pytest

from field_stages import rel_l2
from oracle import low_storage as ls
from oracle import rk
from solprop_amd.native import Workspace
from solprop_amd.problem import capillary_problem

pytestmark = pytest.mark.gpu


def replay_problem(steps=40, **kw):
    base = dict(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")
    base.update(kw)
    problem = capillary_problem(**base)
    return dataclasses.replace(problem, method="low_storage",
                               schedule=ls.uniform_edges(0.0, problem.z_end, steps))


def device_replay(problem, target=None):
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0)
    batch = march.advance(problem.z_end if target is None else target, work_limit=100000)
    state = march.state()
    stiffness = float(ws.buffer("stiffness")[0])
    ws.close()
    return state, batch, stiffness


@pytest.mark.parametrize("case", ["trapezoid", "band"])
def test_replay_matches_oracle(case):
    print(f"low-storage replay vs oracle, {case}", flush=True)
    problem = replay_problem(integration=case, band=(300e-9, 4e-6))
    y0 = problem.field0.reshape(problem.shape)
    expected, rows, stiffness = ls.replay(problem, y0, problem.schedule)
    state, batch, h_rho = device_replay(problem)
    err = rel_l2(state, expected.reshape(state.shape))
    print("end-state rel L2", err, "stiffness", h_rho, stiffness[-1], flush=True)
    assert batch.done
    assert np.array_equal(batch.attempts, np.array(rows))
    assert err < 1e-12
    assert abs(h_rho - stiffness[-1]) <= 1e-10 * abs(stiffness[-1])


def test_replay_at_full_size():
    """The measured capillary (4 modes, 128 nodes, 2^14 samples, band rule, the pressure
    gradient's two-term operator), ten replayed steps."""
    print("low-storage replay at full size, 10 steps", flush=True)
    base = capillary_problem()
    edges = np.linspace(0.0, 10 * 2e-4, 11)
    problem = dataclasses.replace(base, method="low_storage", schedule=edges)
    y0 = problem.field0.reshape(problem.shape)
    expected, rows, _ = ls.replay(problem, y0, edges)
    state, batch, _ = device_replay(problem)
    err = rel_l2(state, expected.reshape(state.shape))
    print("end-state rel L2", err, flush=True)
    assert len(batch.attempts) == 10
    assert err < 1e-12


def test_linear_replay_is_the_exact_flow():
    problem = replay_problem(steps=9, kerr=False, plasma=False)
    y0 = problem.field0.reshape(problem.shape)
    state, _, _ = device_replay(problem)
    exact = rk.linear_solution(problem, y0, problem.z_end)
    assert rel_l2(state, exact.reshape(state.shape)) < 1e-13


def test_graphs_and_direct_launches_agree(monkeypatch):
    problem = replay_problem(steps=12)
    graphed, _, _ = device_replay(problem)
    monkeypatch.setenv("SOLPROP_MI355X_GRAPH", "0")
    direct, _, _ = device_replay(problem)
    assert np.array_equal(graphed, direct)


def test_probe_step_keeps_the_state():
    problem = replay_problem(steps=8)
    y0 = problem.field0.reshape(problem.shape)
    dz = problem.schedule[1]
    expected, _, _ = ls.replay(problem, y0, problem.schedule[:2])
    ws = Workspace(problem)
    probe = ws.probe(0.0, problem.field0)
    norm = probe.trial(0.0, dz)
    field = probe.field()
    state = ws.state()
    ws.close()
    assert norm == 0.0
    assert rel_l2(field, expected.reshape(field.shape)) < 1e-11
    assert np.array_equal(state, problem.field0.reshape(state.shape))


def test_session_over_the_replay():
    problem = replay_problem(steps=20)
    planes = [problem.schedule[5], problem.schedule[12], problem.z_end]
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
    assert len(rows) == 20 and np.all(rows[:, 4] == 4) and np.all(rows[:, 3] == 1)
    state, _, _ = device_replay(problem)
    expected, _, _ = ls.replay(problem, problem.field0.reshape(problem.shape), problem.schedule)
    assert rel_l2(state, expected.reshape(state.shape)) < 1e-12


def test_quadrature_estimate_of_the_last_stage():
    """After a replay lands on an edge, the quadrature estimate repeats the step's last
    source evaluation (its stage field, read back from the device, at the step's end)
    with both projections, as ``low_storage.cpp:219`` scores the last stage; the oracle's
    estimate of that field within 1e-7 of itself."""
    from oracle import stats as S

    print("low-storage quadrature estimate", flush=True)
    problem = replay_problem(steps=10, n_nodes=16, n_modes=6, radial_rule="gauss_kronrod")
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0)
    target = float(problem.schedule[3])
    march.advance(target, work_limit=100000)
    got = ws.statistics(["quadrature_error"])["quadrature_error"]
    lab = ws.buffer("lab").reshape(problem.n_modes, problem.n_omega)
    ws.close()
    expected = S.quadrature_error(problem, lab, target)
    print("device", got, "oracle", expected, flush=True)
    assert 1e-4 < expected < 1.0
    assert abs(got - expected) <= 1e-7 * expected
