"""GPU: the session's event protocol through the C ABI (``_core.pyi:468-558``;
``native/src/session/session.cpp:695``; ``docs/dev/kernel-contract.md`` "What a native
call takes / returns").

A session runs the same march as a direct march to each plane in turn: its attempt rows
(end, step, norm, accepted) and its states at the planes must equal the direct march's to
the bit, however the advances are bounded. Beside that, the protocol itself: the launch
observed first, planes observed where the march lands on them, every k-th accepted step
observed, bounded advances yielding progress events, a terminal event at the end, the
ticket's export, and the lifecycle refusals (advancing across a pending batch, a stale
acknowledgement or ticket, the leased workspace, a stop, the step floor).
"""

import threading

import numpy as np
import pytest

from solprop_amd.native import LifecycleError, Workspace
from solprop_amd.problem import capillary_problem, gnlse_problem

pytestmark = pytest.mark.gpu


def field_problem():
    return capillary_problem(n_t=2048, n_nodes=32, length=0.05, integration="trapezoid")


def envelope_problem():
    return gnlse_problem(n_t=512, ffac=2, window=2e-12, fwhm=50e-15, soliton_order=2.0,
                         length=5e-3)


def drain(session, work_limit=256, deadline=10.0):
    """Advance to the terminal event, acknowledging each batch; every batch returned."""
    out = []
    for _ in range(10000):
        batch = session.advance(work_limit, deadline)
        out.append(batch)
        if batch.event.kind_name == "terminal":
            return out
        session.acknowledge(batch.sequence)
    raise AssertionError("no terminal event")


def direct(problem, planes, limit=100000):
    """The direct march to each plane in turn: its rows and its state at each plane."""
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0)
    rows, states = [], []
    for z in planes:
        while True:
            batch = march.advance(z, work_limit=limit)
            rows.append(batch.attempts)
            if batch.done:
                break
        states.append(march.state())
    ws.close()
    return np.concatenate(rows), states


@pytest.mark.parametrize("route", ["capillary", "envelope"])
@pytest.mark.parametrize("work_limit", [1, 7, 256])
def test_session_march_equals_the_direct_march(route, work_limit):
    print(f"session vs direct march, {route}, work_limit {work_limit}", flush=True)
    problem = field_problem() if route == "capillary" else envelope_problem()
    planes = [0.2 * problem.z_end, 0.5 * problem.z_end, problem.z_end]
    expected_rows, expected_states = direct(problem, planes)
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": planes})
    batches = drain(session, work_limit=work_limit)
    kinds = [b.event.kind_name for b in batches]
    assert kinds[0] == "observation" and batches[0].event.observation_name == "launch"
    assert kinds[-1] == "terminal" and batches[-1].event.outcome_name == "completed"
    observed = [b for b in batches if b.event.kind_name == "observation"
                and b.event.observation_name == "plane"]
    assert [b.event.plane_index for b in observed] == [0, 1, 2]
    assert [b.event.z for b in observed] == planes
    rows = np.concatenate([b.attempts for b in batches])
    assert rows.shape[1] == 6
    assert np.array_equal(rows[:, :4], expected_rows)
    if work_limit < 256:
        assert all(len(b.attempts) <= work_limit for b in batches)
        assert "progress" in kinds
    # The order column: the Adams march ramps from 1 and never exceeds 5; DOPRI5 is 5.
    orders = rows[:, 4]
    if route == "capillary":
        assert orders[0] == 1 and orders.max() <= 5 and orders.min() >= 1
    else:
        assert np.all(orders == 5)
    # The wall column counts from the previous event: non-negative within each chunk.
    assert np.all(rows[:, 5] >= 0)
    session.close()
    ws.close()
    # The states at the planes, by ticket, against the direct march's.
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": planes})
    got = []
    while True:
        batch = session.advance(100000, 100.0)
        if batch.event.kind_name == "terminal":
            break
        if batch.event.observation_name == "plane":
            out = np.empty((problem.n_modes, problem.n_omega), dtype=np.complex128)
            session.export(batch.ticket, {"power_weight": None, "optics": False}, out)
            got.append(out)
        session.acknowledge(batch.sequence)
    status = session.status()
    session.close()
    ws.close()
    assert len(got) == 3
    for a, b in zip(got, expected_states):
        assert np.array_equal(a, b.reshape(a.shape))
    assert status.exports == 3 and status.outcome == 1 and status.state == 2


def test_step_observations_and_statistics():
    """Every 5th accepted step is observed; each observation carries the plane
    statistics of the state it names, equal to the workspace's own statistics there."""
    print("session step observations with statistics", flush=True)
    problem = field_problem()
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end],
                        "step_planes": 5, "statistics": ["energy", "peak_power"]})
    accepted = 0
    seen = 0
    while True:
        batch = session.advance(100000, 100.0)
        accepted += int(batch.attempts[:, 3].sum()) if len(batch.attempts) else 0
        if batch.event.kind_name == "terminal":
            break
        assert batch.event.has_stats
        assert len(batch.plane_stats) == 2 * problem.n_modes
        if batch.event.observation_name == "step":
            assert accepted % 5 == 0
            seen += 1
        session.acknowledge(batch.sequence)
    session.close()
    ws.close()
    assert seen >= 2
    # The statistics at the first step observation, against a direct march of 5 steps.
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end],
                        "step_planes": 5, "statistics": ["energy", "peak_power"],
                        "launch": False})
    batch = session.advance(100000, 100.0)
    assert batch.event.observation_name == "step"
    session.close()
    ws.close()
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0)
    rows = []
    while sum(int(r[3]) for r in rows) < 5:
        rows.extend(march.advance(problem.z_end, work_limit=1).attempts)
    direct_stats = ws.statistics(["energy", "peak_power"])
    ws.close()
    expected = np.concatenate([direct_stats["energy"], direct_stats["peak_power"]])
    assert np.array_equal(batch.plane_stats, expected)


def test_lifecycle_refusals():
    print("session lifecycle refusals", flush=True)
    problem = field_problem()
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end]})
    with pytest.raises(LifecycleError):
        ws.load(0.0, problem.field0)  # leased
    with pytest.raises(LifecycleError):
        ws.start({"planes": [problem.z_end]})  # one session at a time
    launch = session.advance(10, 1.0)
    assert launch.event.observation_name == "launch" and launch.event.ticket
    with pytest.raises(LifecycleError):
        session.advance(10, 1.0)  # across a pending batch
    with pytest.raises(LifecycleError):
        session.acknowledge(launch.sequence + 1)
    out = np.empty((problem.n_modes, problem.n_omega), dtype=np.complex128)
    with pytest.raises(LifecycleError):
        session.export((launch.sequence, launch.event.generation + 1), {}, out)
    session.export(launch.ticket, {"optics": False}, out)
    assert np.array_equal(out, problem.field0.reshape(out.shape))
    session.acknowledge(launch.sequence)
    with pytest.raises(LifecycleError):
        session.export(launch.ticket, {}, out)  # no longer pending
    progress = session.advance(3, 10.0)
    assert progress.event.kind_name == "progress" and len(progress.attempts) == 3
    session.acknowledge(progress.sequence)
    session.request_stop()
    stopped = session.advance(100, 10.0)
    assert stopped.event.kind_name == "terminal" and stopped.event.outcome_name == "stopped"
    with pytest.raises(LifecycleError):
        session.advance(10, 1.0)  # ended
    assert session.status().state == 2
    session.close()
    with pytest.raises(LifecycleError):
        session.advance(10, 1.0)
    ws.load(0.0, problem.field0)  # the lease ends with the session
    ws.close()


def test_stop_from_another_thread():
    """request_stop takes no lock: from another thread while an advance runs, the advance
    ends at the next attempt boundary with a stopped terminal event."""
    print("session stop from another thread", flush=True)
    # Long enough (some thousands of steps) that the stop lands mid-run.
    problem = capillary_problem(n_t=2048, n_nodes=32, length=20.0, integration="trapezoid")
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end],
                        "launch": False})
    timer = threading.Timer(0.02, session.request_stop)
    timer.start()
    batch = session.advance(10**9, 60.0)
    timer.join()
    assert batch.event.kind_name == "terminal" and batch.event.outcome_name == "stopped"
    assert batch.event.z < problem.z_end and len(batch.attempts) > 0
    assert session.status().outcome == 2
    session.close()
    ws.close()


@pytest.mark.filterwarnings("ignore:invalid value encountered in divide:RuntimeWarning")
def test_floor_is_a_failed_terminal_event():
    """An empty launch field drives the step to the floor: a failed terminal event of the
    convergence category carrying every attempt, and the proposal below the floor."""
    print("session failure at the step floor", flush=True)
    problem = field_problem()
    ws = Workspace(problem)
    zero = np.zeros_like(problem.field0)
    session = ws.start({"z": 0.0, "field": zero, "planes": [problem.z_end], "launch": False})
    batches = drain(session, work_limit=10**6)
    last = batches[-1]
    assert last.event.outcome_name == "failed"
    assert last.event.category == 1 and last.event.reason == 5
    assert last.event.message
    rows = np.concatenate([b.attempts for b in batches])
    assert len(rows) > 0 and not rows[:, 3].any()
    # The same attempts as the direct march's refusal.
    ws2 = Workspace(problem)
    march = ws2.march(0.0, zero)
    with pytest.raises(Exception):
        march.advance(problem.z_end, work_limit=10**6)
    ws2.close()
    session.close()
    ws.close()


def test_workspace_closed_under_an_open_session():
    """Destroying a workspace its session leases defers the release to the session's
    close (the C ABI's rule); Python's Workspace.close closes the session first."""
    print("workspace closed under an open session", flush=True)
    import ctypes as C

    from solprop_amd import _lib

    problem = field_problem()
    ws = Workspace(problem)
    session = ws.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end],
                        "launch": False})
    lib = _lib.lib()
    lib.sp_workspace_destroy(ws._handle)  # deferred: the session still leases it
    ws._handle = None
    batch = session.advance(5, 10.0)
    assert batch.event.kind_name == "progress" and len(batch.attempts) == 5
    session.acknowledge(batch.sequence)
    session.close()  # releases the workspace now
    ws2 = Workspace(problem)
    s2 = ws2.start({"z": 0.0, "field": problem.field0, "planes": [problem.z_end]})
    ws2.close()  # closes s2 first
    with pytest.raises(LifecycleError):
        s2.advance(1, 1.0)
    _ = C
