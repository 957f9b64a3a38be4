"""The Runge-Kutta march's decisions, held to the reference's decision corpus.

``tests/golden/rk_traces.json.gz`` is the reference's ``tests/data/rk_traces.json.gz``:
29 scripted marches (norms, caps, landings, start-up, floors, replays) and the full
decision trace the reference's march made on each. Both the oracle's restatement and
the library's compiled march (through the C ABI, no GPU needed) must reproduce every
trace exactly -- the comparison is ``==`` on the parsed JSON, floats bit for bit (glibc
``pow``, as the corpus was recorded). The one tolerated difference is the prose tail
the reference appends to a diverged replay's message.
"""

import gzip
import json

import pytest

from oracle import controller as ctl
from tests.conftest import GOLDEN

CORPUS = json.load(gzip.open(GOLDEN / "rk_traces.json.gz"))["scenarios"]


def _same(got, expected):
    if got == expected:
        return True
    # Only a replay divergence message may differ, by the reference's explanation.
    g_end, e_end = dict(got["end"]), dict(expected["end"])
    gm, em = g_end.pop("message", ""), e_end.pop("message", "")
    return (
        got["events"] == expected["events"]
        and g_end == e_end
        and em.startswith("scripted: ")
        and em.startswith("scripted: " + gm.split("scripted: ")[-1])
    )


@pytest.mark.parametrize("index", range(len(CORPUS)), ids=[
    s["scenario"]["name"] for s in CORPUS])
def test_oracle_controller_reproduces_the_corpus(index):
    case = CORPUS[index]
    assert _same(ctl.run_scenario(case["scenario"]), case["trace"])


@pytest.mark.parametrize("index", range(len(CORPUS)), ids=[
    s["scenario"]["name"] for s in CORPUS])
def test_library_march_reproduces_the_corpus(index):
    from solprop_amd.native import rk_trace

    case = CORPUS[index]
    assert _same(rk_trace(case["scenario"]), case["trace"])


def test_corpus_covers_every_decision():
    kinds = {s["scenario"]["name"] for s in CORPUS}
    assert {"default", "floor_on_rejection", "floor_at_start_up", "floor_on_proposal",
            "replay", "replay_diverged", "stretch_wide", "kinks_and_intervals"} <= kinds
    stops = {s["trace"]["end"]["stop"] for s in CORPUS}
    assert {"done", "exhausted", "ConvergenceError"} <= stops


def test_controller_docstring_answers():
    # src/solprop/stepper/controller.py docstrings (propose, start_factor,
    # rejection_factor, equivalent_norm, limit_step).
    c = ctl.PIController(rtol=1e-8, atol=1e-30)
    assert (c.accepts(0.5), c.accepts(2.0)) == (True, False)
    assert f"{c.propose([0.5, 0.5]):.4f}" == "0.8694"
    lagged = ctl.PIController(rtol=1e-8, lag=1)
    assert lagged.propose([]) == 1.0 and lagged.propose([0.1]) == 1.0
    assert f"{lagged.propose([0.1, 0.2]):.4f}" == "1.0546"
    assert lagged.propose([1e-6, 0.2]) == 1.1
    c6 = ctl.PIController(rtol=1e-6)
    assert f"{c6.start_factor(1e-5):.2f}" == "8.00"
    assert (c6.start_factor(0.0), c6.start_factor(1e-30)) == (100.0, 100.0)
    assert f"{c.rejection_factor(4.0):.4f}" == "0.6063"
    assert c.equivalent_norm(0.3, 1e-3, 1e-3) == 0.3
    assert f"{c.equivalent_norm(1e-6, 1e-4, 1e-3):.2f}" == "0.10"
    assert ctl.limit_step(1e-3, 1.1, 1e-9, 1e-2) == 0.0011
    assert ctl.limit_step(1e-3, 1000.0, 1e-9, 1e-2) == 0.01
    with pytest.raises(ctl.ConvergenceError):
        ctl.limit_step(1e-3, 1e-9, 1e-6, 1e-2)


def test_landing_docstring_answers():
    # src/solprop/stepper/schedule.py:1001-1022.
    t = ctl.StepTargets((0.25, 0.5, 0.75, 1.0), (0.4,), (0.0, 1.0), 0.3, 1.0)
    stop = ctl.landing(0.0, 1.0, t)
    assert (stop.dz, stop.target, stop.z_target) == (0.25, "save_plane", 0.25)
    assert ctl.landing(0.25, 1.0, t).z_target == 0.4
    assert ctl.landing(0.9, 0.05, t).target == "none"
    stop = ctl.landing(0.0, 0.24, t)
    assert (stop.z_target, stop.stretched) == (0.25, True)
    assert (ctl.reached(0.5, 1.0), ctl.reached(1.0, 1.0)) == (False, True)
