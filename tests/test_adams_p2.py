"""The oracle's exponential Adams march on the reference's analytic rung P2.

P2 (``docs/dev/stepper.md:839``): a second-order soliton under z-varying dispersion
d(z) and loss alpha(z), gamma = d e^Gamma, whose exact solution is known
(:mod:`analytic`). Its operator has two terms on different z-paths, so every trial
carries the transported compensation of a varying operator -- the part of the field
route's trial (config [1] runs it: the capillary's pressure gradient gives two Φ
terms) that no decision corpus or per-function check fixes.

The march must converge to the exact field at the rate its order promises: the error
at the end follows rtol down three decades and falls with the step count at an
observed order of at least 4. Two planted mistakes in the oracle fail the same bar:
the compensation with its sign flipped, and the compensation left out.
"""

import math

import numpy as np
import pytest

from analytic import p2_exact, p2_problem, second_order_soliton
from oracle import adams


class FlippedCompensation(adams.ExponentialEngine):
    """The residual rate with its sign flipped: x/h - D(z_j)."""

    @staticmethod
    def residual_rate(rate, reference):
        return reference - rate


class NoCompensation(adams.ExponentialEngine):
    """The residual left out: the varying operator's flow taken as the step's mean."""

    @staticmethod
    def residual_rate(rate, reference):
        return 0.0 * rate


def run(rtol, engine=adams.ExponentialEngine):
    problem = p2_problem(rtol=rtol)
    state, rows = adams.march(problem, problem.field0.reshape(problem.shape).copy(),
                              engine=engine)
    exact = p2_exact(problem, problem.z_end)
    error = float(np.linalg.norm(state - exact) / np.linalg.norm(exact))
    accepted = sum(1 for r in rows if r[3])
    return error, accepted


def observed_order(coarse, fine):
    return math.log(coarse[0] / fine[0]) / math.log(fine[1] / coarse[1])


def test_second_order_soliton_solves_its_equation():
    """The closed form satisfies i u_xi + u_tt/2 + |u|^2 u = 0 (finite differences) and
    starts at 2 sech t: a mistyped formula fails here before any march."""
    t = np.linspace(-6, 6, 1201)
    dt = t[1] - t[0]
    h = 1e-3
    for xi in (0.0, 0.3, 0.785, 1.1):
        u = second_order_soliton(xi, t)
        # Fourth-order central differences in xi and t.
        u_xi = (-second_order_soliton(xi + 2 * h, t) + 8 * second_order_soliton(xi + h, t)
                - 8 * second_order_soliton(xi - h, t) + second_order_soliton(xi - 2 * h, t)) / (12 * h)
        u_tt = (-u[4:] + 16 * u[3:-1] - 30 * u[2:-2] + 16 * u[1:-3] - u[:-4]) / (12 * dt**2)
        inner = u[2:-2]
        residual = 1j * u_xi[2:-2] + 0.5 * u_tt + np.abs(inner) ** 2 * inner
        assert np.max(np.abs(residual)) < 1e-4
    assert np.allclose(second_order_soliton(0.0, t), 2.0 / np.cosh(t), rtol=0, atol=1e-14)


def test_p2_linear_paths_are_consistent():
    """The problem's cumulative paths and the rates the compensation reads agree: the
    rate integrates to the increment (the reduction to the constant equation needs it)."""
    problem = p2_problem()
    z = np.linspace(0.0, problem.z_end, 2001)
    for path, rate in zip(problem.phi_paths, problem.phi_rate_paths, strict=True):
        values = np.array([path(v) for v in z])
        rates = np.array([rate(v) for v in z])
        integral = np.concatenate(([0.0], np.cumsum(0.5 * (rates[1:] + rates[:-1]) * np.diff(z))))
        assert np.max(np.abs(integral - (values - values[0]))) < 1e-6


def test_p2_march_converges_at_its_order():
    """Error within 10 rtol at rtol 1e-6 and 1e-9, observed order >= 4 between them."""
    coarse = run(1e-6)
    fine = run(1e-9)
    print("P2 compensated", coarse, fine, flush=True)
    assert coarse[0] < 10 * 1e-6
    assert fine[0] < 10 * 1e-9
    assert observed_order(coarse, fine) >= 4.0


@pytest.mark.parametrize("engine", [FlippedCompensation, NoCompensation])
def test_p2_fails_with_a_wrong_compensation(engine):
    """Either planted mistake stalls: at rtol 1e-9 the error stays above 1e-6, and the
    observed order between rtol 1e-6 and 1e-9 is below 3 (without the compensation the
    varying operator's flow is taken to second order only; the flipped sign stalls)."""
    coarse = run(1e-6, engine)
    fine = run(1e-9, engine)
    print(engine.__name__, coarse, fine, flush=True)
    assert fine[0] > 1e-6
    assert observed_order(coarse, fine) < 3.0
