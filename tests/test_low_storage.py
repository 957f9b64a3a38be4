"""The low-storage RK4IP oracle (``oracle/low_storage.py``), pinned before it judges the
device.

- For u' = lambda u with no linear flow, the stiffness indicator is |h lambda| exactly
  (the reference's definition, ``numpy_ref/engines.py:268-276``), and one step is RK4's
  polynomial 1 + z + z^2/2 + z^3/6 + z^4/24 of z = h lambda.
- A linear-only problem's replay is the exact flow exp(Phi(z) - Phi(0)) y0.
- On the analytic rung P2 (second-order soliton under z-varying dispersion and loss,
  ``tests/analytic.py``) the replay converges at fourth order; a planted mistake (pass
  3's stage without the flow Q) falls to first order.
"""

import math

import numpy as np

from analytic import p2_exact, p2_problem
from oracle import low_storage as ls
from oracle import rk
from solprop_amd.problem import capillary_problem


def test_stiffness_and_rk4_polynomial():
    lam = -0.7 + 0.4j
    h = 0.3
    u = np.array([[1.0 + 0.5j, -0.2 + 0.1j]])
    mask = np.array([True, True])
    out, squared, h_rho = ls.step(u, 0.0, h, lambda v, z: lam * v,
                                  lambda z0, z1: np.ones_like(u), mask)
    z = h * lam
    assert np.allclose(out, (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24) * u, rtol=0, atol=1e-15)
    assert math.isclose(h_rho, abs(z), rel_tol=1e-13)
    assert math.isclose(squared, float(np.vdot(out, out).real), rel_tol=1e-15)


def test_linear_problem_is_the_exact_flow():
    problem = capillary_problem(n_t=512, n_nodes=8, length=0.05, kerr=False, plasma=False,
                                integration="trapezoid")
    y0 = problem.field0.reshape(problem.shape)
    state, rows, _ = ls.replay(problem, y0, ls.uniform_edges(0.0, problem.z_end, 7))
    exact = rk.linear_solution(problem, y0, problem.z_end)
    assert len(rows) == 7 and all(r[3] == 1.0 and r[2] == 0.0 for r in rows)
    assert np.linalg.norm(state - exact) <= 1e-13 * np.linalg.norm(exact)


def _p2_error(steps, step=ls.step):
    problem = p2_problem()
    y0 = problem.field0.reshape(problem.shape)
    saved = ls.step
    ls.step = step
    try:
        state, _, _ = ls.replay(problem, y0, ls.uniform_edges(0.0, problem.z_end, steps))
    finally:
        ls.step = saved
    exact = p2_exact(problem, problem.z_end)
    return float(np.linalg.norm(state - exact) / np.linalg.norm(exact))


def test_p2_converges_at_fourth_order():
    coarse, fine = _p2_error(400), _p2_error(800)
    order = math.log(coarse / fine) / math.log(2.0)
    print("P2 RK4IP errors", coarse, fine, "order", order, flush=True)
    assert fine < 1e-6
    assert order >= 3.8


def test_a_missing_flow_fails():
    def planted(state, z, dz, source, flows, norm_mask):
        mid, end = z + 0.5 * dz, z + dz
        p, q = flows(z, mid), flows(mid, end)
        s1 = source(state, z)
        a, g = p * state, p * s1
        b, c = a + (dz / 6.0) * g, a + (0.5 * dz) * g
        s2 = source(c, mid)
        b, c = b + (dz / 3.0) * s2, a + (0.5 * dz) * s2
        s3 = source(c, mid)
        b, c = b + (dz / 3.0) * s3, a + dz * s3  # Q left out
        s4 = source(c, end)
        a = q * b + (dz / 6.0) * s4
        return a, float(np.vdot(a, a).real), float("nan")

    coarse, fine = _p2_error(400, planted), _p2_error(800, planted)
    order = math.log(coarse / fine) / math.log(2.0)
    print("planted errors", coarse, fine, "order", order, flush=True)
    assert order < 2.5
