"""GPU parity of the radial route (the dense field route), through the C ABI.

A small free-space argon beam: 1024 samples over 250 fs, a 32-point Hankel transform
over a 4 mm aperture, a 30 fs 800 nm pulse of 2 mJ. On this route the synthesis and the
projection are matrix-core GEMMs, and the projection applies the time collars after
the sum over nodes instead of before it.

Tolerances (class R, fixed before the run):
- node field: the GEMM's error against the oracle's product is held to Higham's
  bound, 4 K eps || |S| |a| ||_F with K the number of modes summed;
- modal fields, polarisation, currents and source as on the capillary route (the
  collars are applied once to the projected sum, an extra rounding of one ulp);
- a first trial and a short march as on the capillary route.
"""

import numpy as np
import pytest

from field_stages import (exp_ulp_sensitivity, log_ulp_sensitivity, oracle_first_trial,
                          oracle_source_stages, rel_l2, scan_sensitivity)
from oracle import adams
from solprop_amd.native import Workspace
from solprop_amd.problem import radial_problem

pytestmark = pytest.mark.gpu

CASES = {
    # A 1 mm beam: Kerr dominates, the plasma current is at the table's floor.
    "wide_beam": {"w0": 1.0e-3},
    # A 0.25 mm beam: a strong plasma current.
    "narrow_beam": {"w0": 0.25e-3},
    "kerr_only": {"w0": 0.25e-3, "plasma": False},
    # A molecular response taking half the third-order response (constant rate: the
    # kernel is built once).
    "raman": {"w0": 0.25e-3, "raman_share": 0.5},
    "raman_analytic": {"w0": 0.25e-3, "raman_share": 0.5, "raman_drive": "analytic"},
    "band_rule": {"w0": 0.25e-3, "integration": "band", "band": (400e-9, 4e-6)},
    # A transverse collar over the aperture's outer millimetre (peak 2000 / m) on the
    # wide beam, which reaches it; alone, and with the in-stage temporal absorber added
    # to the same stream.
    "collar": {"w0": 1.0e-3, "collar": (1.0e-3, 2.0e3)},
    "collar_in_stage": {"w0": 1.0e-3, "collar": (1.0e-3, 2.0e3), "absorber": "in_stage",
                        "absorber_length": 2e-2},
}


def small(**kw):
    # The trapezoid rule unless a case asks for the band rule (see the capillary tests).
    args = dict(n_t=1024, window=0.25e-12, n_r=32, length=0.02, fwhm=30e-15,
                integration="trapezoid")
    args.update(kw)
    return radial_problem(**args)


@pytest.mark.parametrize("case", sorted(CASES))
def test_radial_source_stages_match_oracle(case):
    from oracle import passes as P

    print(f"radial source stages {case}", flush=True)
    problem = small(**CASES[case])
    lab = problem.field0.reshape(problem.shape)
    z = 1e-3
    modal, streams, source = oracle_source_stages(problem, lab, z)
    field_t = P.synthesise(modal[:, None], problem.synthesis, 1.0, 1)[:, 0]
    ws = Workspace(problem)
    ws.load(0.0, problem.field0)
    got, got_streams, got_modal = ws.source(z, lab, intermediates=True)
    node_field = ws.buffer("node_field").reshape(field_t.shape)
    ws.close()
    bound = 4 * problem.n_modes * np.finfo(float).eps * np.linalg.norm(
        np.abs(problem.synthesis) @ np.abs(modal))
    node_error = float(np.linalg.norm(node_field - field_t))
    errors = {"modal": rel_l2(got_modal, modal), "source": rel_l2(got, source.reshape(got.shape))}
    names = ("polarisation", "gated current", "edge current")
    for i, name in enumerate(names):
        if np.any(streams[i]):
            errors[name] = rel_l2(got_streams[i], streams[i])
    sensitivity = np.maximum(log_ulp_sensitivity(problem, lab, z, streams),
                             exp_ulp_sensitivity(problem, lab, z, streams))
    print(errors, "node field", node_error, "bound", bound, "ulp sensitivity",
          sensitivity.tolist(), flush=True)
    assert node_error <= bound
    assert errors["modal"] < 1e-14
    assert errors["source"] < 1e-12
    assert errors["polarisation"] < 1e-13
    for i in (1, 2):
        if names[i] in errors:
            assert errors[names[i]] < max(1e-12, 4 * sensitivity[i]), (names[i], errors)
    if not problem.plasma:
        assert not np.any(got_streams[1]) and not np.any(got_streams[2])


@pytest.mark.parametrize("case", sorted(CASES))
def test_radial_first_trial_matches_oracle(case):
    print(f"radial probe trial {case}", flush=True)
    problem = small(**CASES[case])
    dz = 1e-4
    advanced, norms = oracle_first_trial(problem, dz)
    ws = Workspace(problem)
    probe = ws.probe(0.0, problem.field0)
    got = probe.trial(0.0, dz)
    field = probe.field()
    ws.close()
    assert rel_l2(field, advanced.reshape(field.shape)) < 1e-11
    assert abs(got - norms[0]) / norms[0] < 1e-8


@pytest.mark.parametrize("case", ["wide_beam", "narrow_beam", "raman", "raman_analytic",
                                  "band_rule", "collar", "collar_in_stage"])
def test_radial_march_matches_oracle_decisions(case):
    print(f"radial march {case}", flush=True)
    problem = small(**CASES[case])
    y0 = problem.field0.reshape(problem.shape)
    expected, records = adams.march(problem, y0.copy(), max_attempts=30)
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    batch = session.advance(problem.z_end, work_limit=len(records))
    rows = batch.attempts
    state = session.state()
    ws.close()
    assert len(rows) == len(records)
    assert [bool(r[3]) for r in rows] == [bool(r[3]) for r in records]
    steps = np.array([r[1] for r in records])
    assert np.allclose(rows[:, 1], steps, rtol=1e-7, atol=0)
    assert rel_l2(state, expected.reshape(state.shape)) < 1e-8


def test_collar_changes_the_march():
    """The transverse collar acts: a short march with it ends measurably away from the
    same march without it (the oracle's, which the device follows above)."""
    print("radial collar acts", flush=True)
    target = 2e-3
    ends = []
    for case in ("collar", "wide_beam"):
        problem = small(**CASES[case])
        ws = Workspace(problem)
        session = ws.march(0.0, problem.field0)
        batch = session.advance(target, work_limit=200)
        assert batch.done
        ends.append(session.state())
        ws.close()
    assert rel_l2(ends[0], ends[1]) > 1e-8


def test_radial_repeats_agree_to_the_bit():
    print("radial determinism", flush=True)
    problem = small(**CASES["narrow_beam"])
    outs = []
    for _ in range(2):
        ws = Workspace(problem)
        session = ws.march(0.0, problem.field0)
        batch = session.advance(problem.z_end, work_limit=10)
        outs.append((batch.attempts.copy(), session.state()))
        ws.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_radial_wide_grid_matches_oracle():
    """C·N above 4096·64: the corrector runs four-wave workgroups and its block sums are
    fewer than the bins; every bin of the predicted and corrected fields must still be
    covered. A first trial to 1e-11 and a short march with the oracle's decisions."""
    print("radial wide grid", flush=True)
    problem = small(n_r=128, n_t=4096, w0=0.25e-3, integration="band", band=(400e-9, 4e-6))
    assert problem.n_modes * problem.n_omega > 4096 * 64
    dz = 1e-4
    advanced, norms = oracle_first_trial(problem, dz)
    ws = Workspace(problem)
    probe = ws.probe(0.0, problem.field0)
    got = probe.trial(0.0, dz)
    field = probe.field()
    ws.close()
    assert rel_l2(field, advanced.reshape(field.shape)) < 1e-11
    assert abs(got - norms[0]) / norms[0] < 1e-8
    y0 = problem.field0.reshape(problem.shape)
    expected, records = adams.march(problem, y0.copy(), max_attempts=12)
    ws = Workspace(problem)
    session = ws.march(0.0, problem.field0)
    rows = session.advance(problem.z_end, work_limit=len(records)).attempts
    state = session.state()
    ws.close()
    assert [bool(r[3]) for r in rows] == [bool(r[3]) for r in records]
    assert np.allclose(rows[:, 1], [r[1] for r in records], rtol=1e-7, atol=0)
    assert rel_l2(state, expected.reshape(state.shape)) < 1e-8


def test_radial_first_trial_matches_oracle_at_full_size():
    """The measured workload itself: 512-point transform, 2^13 samples over 1 ps."""
    print("radial probe trial at full size", flush=True)
    problem = radial_problem()
    dz = 1e-4
    advanced, norms = oracle_first_trial(problem, dz)
    ws = Workspace(problem)
    probe = ws.probe(0.0, problem.field0)
    got = probe.trial(0.0, dz)
    field = probe.field()
    ws.close()
    assert rel_l2(field, advanced.reshape(field.shape)) < 1e-11
    assert abs(got - norms[0]) / norms[0] < 1e-8


@pytest.mark.parametrize("case", ["narrow_beam", "band_rule"])
def test_radial_plasma_scan_stages_match_oracle(case):
    """The dense route's stored node arrays after one source evaluation, against the
    oracle's scan of the device's own node field (so that the GEMM's rounding is not in
    the comparison): rate, fraction and free current, to the capillary test's
    tolerances."""
    from oracle import passes as P
    from oracle import rk

    print(f"radial plasma stages {case}", flush=True)
    problem = small(**CASES[case])
    lab = problem.field0.reshape(problem.shape)
    z = 1e-3
    ws = Workspace(problem)
    ws.load(0.0, problem.field0)
    ws.source(z, lab)
    node_field = ws.buffer("node_field")
    got_rate, got_frac = ws.buffer("rate"), ws.buffer("frac")
    free_local, tile_f = ws.buffer("free_local"), ws.buffer("tile_f")
    ws.close()
    src = rk.StageSource(problem)
    field_t = node_field[:, None, :]
    rate = P.ionisation_rate(field_t, src.plasma.rate)
    fraction, free = P.plasma_scan(field_t, rate, src.plasma, problem.preionisation(z),
                                   problem.dt_fine)
    tiles = tile_f.shape[1]
    owner = np.minimum(np.arange(problem.n_to) // 63, tiles - 1)
    got_free = tile_f[:, owner] + free_local
    errors = {"rate": rel_l2(got_rate, rate), "fraction": rel_l2(got_frac, fraction),
              "free current": rel_l2(got_free, free[:, 0])}
    sensitivity = scan_sensitivity(field_t, rate, src.plasma, problem.preionisation(z),
                                   problem.dt_fine, fraction, free)
    print(errors, "exp-ulp sensitivity", sensitivity.tolist(), flush=True)
    assert errors["rate"] < 1e-13
    assert errors["fraction"] < max(1e-12, 4 * sensitivity[0])
    assert errors["free current"] < max(1e-11, 4 * sensitivity[1])
