"""Analytic propagation problems for order tests (the reference's verification ladder).

``p2_problem`` is the reference's rung P2 (``docs/dev/stepper.md:839``, §11; named in
``docs/dev/testing.md:41``): an N-soliton under an arbitrary z-varying dispersion d(z)
and loss alpha(z), with the Kerr coefficient gamma(z) = d(z) e^{Gamma(z)},
Gamma = int alpha. In the optics convention the envelope obeys

    A_z = i (d/2) A_tt - (alpha/2) A + i gamma |A|^2 A,

and A = e^{-Gamma/2} u with xi = int d reduces it to the constant-coefficient
equation i u_xi + u_tt / 2 + |u|^2 u = 0, whose second-order soliton has the closed
form of Satsuma and Yajima (Prog. Theor. Phys. Suppl. 55, 284 (1974)):

    u(xi, t) = 4 (cosh 3t + 3 e^{4 i xi} cosh t) e^{i xi / 2}
               / (cosh 4t + 4 cosh 2t + 3 cos 4 xi).

The problem is posed on the envelope grid in the library's internal convention (the
conjugate of the optics one, as ``backends/export.py:52-207`` converts), so its state
is ``fft(conj(A))``. The operator has two terms with different z-paths, the dispersion
on xi(z) and the loss on Gamma(z), so its rate D(z) is not proportional to any fixed
operator: the exponential Adams march carries the transported compensation
``(D(z_j) - x/h) u_j`` (``numpy_ref/multistep.py:806-1131``) on every step.
"""

from __future__ import annotations

import math

import numpy as np

from solprop_amd.problem import KerrGroup, Problem, ScalarPath


def second_order_soliton(xi, t):
    """u(xi, t) of i u_xi + u_tt/2 + |u|^2 u = 0 with u(0, t) = 2 sech t."""
    t = np.asarray(t, dtype=np.float64)
    numerator = 4.0 * (np.cosh(3 * t) + 3.0 * np.exp(4j * xi) * np.cosh(t)) * np.exp(0.5j * xi)
    denominator = np.cosh(4 * t) + 4.0 * np.cosh(2 * t) + 3.0 * np.cos(4.0 * xi)
    return numerator / denominator


def p2_paths(length: float, d_swing: float = 0.4, alpha0: float = 0.3, alpha_swing: float = 0.25):
    """xi(z), Gamma(z), and their rates d(z), alpha(z), as closed forms."""
    k = 2 * math.pi / length
    q = 3 * math.pi / length

    def xi(z):
        return z + d_swing * (1.0 - np.cos(k * z)) / k

    def d(z):
        return 1.0 + d_swing * np.sin(k * z)

    def gamma_exp(z):
        return alpha0 * z + alpha_swing * np.sin(q * z) / q

    def alpha(z):
        return alpha0 + alpha_swing * np.cos(q * z)

    return xi, d, gamma_exp, alpha


def p2_problem(*, n_t: int = 1024, window: float = 60.0, length: float = 1.5,
               rtol: float = 1e-6, d_swing: float = 0.4, alpha0: float = 0.3,
               alpha_swing: float = 0.25) -> Problem:
    """The P2 rung as a two-term envelope problem under the exponential Adams march.

    Normalised units (time in soliton widths, z in dispersion lengths). The grid is
    ``n_t`` bins over ``window``, doubled for the products (``n_to = 2 n_t``), every bin
    active and scored; the launch is ``2 sech t`` (xi = 0).
    """
    n_to = 2 * n_t
    dt = window / n_t
    omega = 2 * np.pi * np.fft.fftfreq(n_t, dt)
    xi, d, gamma_exp, _ = p2_paths(length, d_swing, alpha0, alpha_swing)
    # Internal convention: A_spec' = D(z) A_spec + prefactor * F[rho |a|^2 a], a = conj(A).
    # The optics equation's dispersion i(d/2) A_tt becomes +i (omega^2 / 2) d(z) here, its
    # loss -alpha/2, and the Kerr term -i gamma |a|^2 a.
    phi = np.stack([(0.5j * omega**2).reshape(1, n_t),
                    np.full((1, n_t), -0.5, dtype=np.complex128)])
    edges = np.linspace(0.0, length, 9)
    xi_path = ScalarPath.chebyshev(xi, 0.0, length, degree=24, edges=edges)
    loss_path = ScalarPath.chebyshev(gamma_exp, 0.0, length, degree=24, edges=edges)
    rho_path = ScalarPath.chebyshev(lambda z: d(z) * np.exp(gamma_exp(z)), 0.0, length,
                                    degree=24, edges=edges)
    t = (np.arange(n_t) - n_t // 2) * dt
    a0 = np.conj(second_order_soliton(0.0, t))
    crop = np.concatenate((np.arange(n_t // 2), np.arange(n_to - n_t // 2, n_to)))
    ones = np.ones(n_t, dtype=bool)
    return Problem(
        kind="envelope",
        n_t=n_t,
        n_to=n_to,
        n_modes=1,
        n_nodes=1,
        crop_index=crop.astype(np.int64),
        active_mask=ones,
        norm_mask=ones.copy(),
        omega_window=np.ones(n_t),
        omega=omega,
        time_window=np.ones(n_to),
        time_window_derivative=np.zeros(n_to),
        dt_fine=window / n_to,
        z_end=length,
        phi_terms=np.ascontiguousarray(phi.astype(np.complex128)),
        phi_paths=(xi_path, loss_path),
        prefactor_terms=np.full((1, 1, n_t), -1j * n_t / n_to, dtype=np.complex128),
        prefactor_paths=(ScalarPath.constant(1.0),),
        synthesis=np.array([[n_to / n_t]]),
        projection=np.array([[1.0]]),
        density=rho_path,
        preionisation=ScalarPath.constant(0.0),
        kerr=KerrGroup(1.0, "analytic", self_factor=1.0),
        field0=np.fft.fft(a0).reshape(1, n_t).astype(np.complex128),
        rtol=rtol,
        method="adams",
        description={"rung": "P2", "n_t": n_t, "window": window, "length": length},
    )


def p2_exact(problem: Problem, z: float) -> np.ndarray:
    """The exact state at z (the problem's own paths xi and Gamma), shape (1, 1, N)."""
    n_t = problem.n_t
    dt = problem.dt_fine * problem.n_to / n_t
    t = (np.arange(n_t) - n_t // 2) * dt
    xi = problem.phi_paths[0](z)
    gamma_exp = problem.phi_paths[1](z)
    a = np.conj(math.exp(-0.5 * gamma_exp) * second_order_soliton(xi, t))
    return np.fft.fft(a).reshape(problem.shape)
