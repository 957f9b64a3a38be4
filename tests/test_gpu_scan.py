"""GPU: a parameter scan as one batched march (``sp_scan_*``), through the C ABI.

Every instance of a scan is a point of the capillary family with its own grid (2^10 to
2^12 samples here), modes (1 to 4), nodes, gas pressure, energy, plasma rule and length.
Each round of the scan runs all instances' trials as one captured graph of grouped
launches. Each instance's march must be the one its own workspace makes alone: the same
attempts and end state, to the bit (the grouped kernels run the plain kernels' bodies
on each instance's buffers, and the decisions are each instance's own). Two small
instances are also held to the oracle's march, as the plain route is.
"""

import numpy as np
import pytest

from field_stages import rel_l2
from oracle import adams
from solprop_amd.native import Scan, Workspace
from solprop_amd.problem import capillary_problem

pytestmark = pytest.mark.gpu


def points():
    """Uneven instances: grids, modes, nodes, rules and physics all differ."""
    specs = [
        dict(n_t=2048, n_modes=4, n_nodes=32, energy=275e-6, p0=0.8, integration="trapezoid"),
        dict(n_t=1024, n_modes=1, n_nodes=16, energy=150e-6, p0=1.5, integration="trapezoid"),
        dict(n_t=4096, n_modes=2, n_nodes=24, energy=400e-6, p0=0.5, integration="trapezoid"),
        dict(n_t=2048, n_modes=3, n_nodes=32, energy=100e-6, p0=2.0, plasma=False,
             integration="trapezoid"),
        dict(n_t=4096, n_modes=4, n_nodes=40, energy=300e-6, p0=0.8, integration="band",
             band=(300e-9, 4e-6)),
        dict(n_t=1024, n_modes=2, n_nodes=8, energy=50e-6, p0=3.0, kerr=False,
             integration="trapezoid"),
    ]
    return [capillary_problem(length=0.04, **s) for s in specs]


def alone(problem, limit):
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    batch = session.advance(problem.z_end, work_limit=limit)
    out = (batch.attempts.copy(), session.state())
    ws.close()
    return out


def test_scan_matches_each_instance_run_alone():
    print("scan of 6 uneven instances vs each alone", flush=True)
    problems = points()
    limit = 30
    expected = [alone(p, limit) for p in problems]
    scan = Scan(problems)
    scan.start(0.0)
    batches = scan.advance([p.z_end for p in problems], work_limit=limit)
    got = [(b.attempts.copy(), scan.state(i)) for i, b in enumerate(batches)]
    timing = scan.timing()
    scan.close()
    print("rounds", timing["rounds"], flush=True)
    assert timing["rounds"] >= limit
    for i, ((rows_a, state_a), (rows_b, state_b)) in enumerate(zip(expected, got)):
        assert len(rows_a) == limit, i
        assert np.array_equal(rows_a, rows_b), i
        assert np.array_equal(state_a, state_b), i


def test_scan_continues_across_calls_as_one_march():
    """Two advance calls of 15 attempts each make the march one call of 30 makes."""
    print("scan across calls", flush=True)
    problems = points()[:3]
    one = Scan(problems)
    one.start(0.0)
    one.advance([p.z_end for p in problems], work_limit=30)
    ends_one = [one.state(i) for i in range(len(problems))]
    one.close()
    two = Scan(problems)
    two.start(0.0)
    for _ in range(2):
        two.advance([p.z_end for p in problems], work_limit=15)
    ends_two = [two.state(i) for i in range(len(problems))]
    two.close()
    for a, b in zip(ends_one, ends_two):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("index", [0, 4])
def test_scan_instance_matches_oracle(index):
    """A scan instance's march against the oracle's: the same decisions, steps within
    1e-7 and the end state within 1e-8 (the plain route's march bar)."""
    print(f"scan instance {index} vs oracle", flush=True)
    problems = points()
    problem = problems[index]
    expected, records = adams.march(problem, problem.field0.reshape(problem.shape).copy(),
                                    max_attempts=30)
    scan = Scan(problems)
    scan.start(0.0)
    batches = scan.advance([p.z_end for p in problems], work_limit=30)
    rows = batches[index].attempts
    state = scan.state(index)
    scan.close()
    assert len(rows) == len(records)
    assert [bool(r[3]) for r in rows] == [bool(r[3]) for r in records]
    assert np.allclose(rows[:, 1], [r[1] for r in records], rtol=1e-7, atol=0)
    assert rel_l2(state, expected.reshape(state.shape)) < 1e-8


def test_scan_runs_instances_to_their_own_ends():
    """Instances of different lengths: each stops at its own end (done), the others go on;
    the short ones' last commits land."""
    print("scan to the ends", flush=True)
    problems = [capillary_problem(n_t=1024, n_modes=2, n_nodes=8, length=length,
                                  integration="trapezoid", energy=50e-6)
                for length in (2e-4, 5e-4, 1e-3)]
    scan = Scan(problems)
    scan.start(0.0)
    batches = scan.advance([p.z_end for p in problems], work_limit=400)
    ends = [scan.state(i) for i in range(len(problems))]
    scan.close()
    for p, b, end in zip(problems, batches, ends):
        assert b.done
        assert abs(b.z - p.z_end) < 1e-15
        ref = alone(p, 400)
        assert np.array_equal(ref[1], end)


def test_autotuned_schedule_changes_no_result(monkeypatch):
    """The autotuner picks each grouped transform pass's sequences a workgroup by timing;
    a launch shape only, so a scan tuned and one held to the plans' own shapes make the
    same attempts and end states to the bit. The schedule names every pass it timed."""
    print("autotuned scan vs untuned", flush=True)
    problems = [capillary_problem(n_t=16384, n_modes=m, n_nodes=16, energy=e, length=0.01,
                                  integration="trapezoid")
                for m, e in [(1, 200e-6), (2, 275e-6), (3, 150e-6)]]
    problems += [capillary_problem(n_t=8192, n_modes=2, n_nodes=16, energy=250e-6,
                                   length=0.01, integration="trapezoid")]
    results = []
    for tuned in (True, False):
        monkeypatch.setenv("SOLPROP_MI355X_AUTOTUNE", "1" if tuned else "0")
        scan = Scan(problems)
        schedule = scan.schedule()
        scan.start(0.0)
        batches = scan.advance([p.z_end for p in problems], work_limit=12)
        results.append(([b.attempts.copy() for b in batches],
                         [scan.state(i) for i in range(len(problems))], schedule))
        scan.close()
    (rows_a, states_a, schedule), (rows_b, states_b, untuned) = results
    print("schedule", schedule, flush=True)
    assert untuned == []
    assert schedule and all(len(entry["us"]) >= 1 for entry in schedule if not entry["cached"])
    assert any(entry["points"] in (64, 128) for entry in schedule)  # the four-step passes
    for a, b in zip(rows_a, rows_b):
        assert len(a) > 0 and np.array_equal(a, b)
    for a, b in zip(states_a, states_b):
        assert np.array_equal(a, b)
