"""Gauss–Kronrod radial rules with their embedded Gauss rule.

The reference's default radial rule on ``xi = r / a`` is the ``2 n + 1`` point Kronrod
extension of the n-point Gauss–Legendre rule, with the Gauss rule embedded in it
(``src/solprop/modes/quadrature.py:98-114,199-332``): the projection carries the
Kronrod weights and the coarse projection the Gauss weights on the shared nodes, and the
difference of the two projections is the quadrature estimate the statistics report
(``native/src/general/trial_finish.inc:921-1028``).

The Jacobi–Kronrod matrix comes from D. P. Laurie's mixed-moment recursion ("Calculation
of Gauss–Kronrod quadrature rules", Math. Comp. 66, 1133–1145 (1997)), in the vectorised
form of W. Gautschi's OPQ routine ``r_kronrod``: the eastward sweep over the first n - 1
rows of mixed moments, then the southward sweep that completes the matrix's lower half.
Nodes and weights follow from the matrix by Golub and Welsch's eigenvalue method.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.linalg import eigh_tridiagonal
from scipy.special import roots_legendre

__all__ = ["legendre_recurrence", "jacobi_kronrod", "gauss_kronrod", "radial_rule"]


def legendre_recurrence(count: int) -> tuple[np.ndarray, np.ndarray]:
    """The monic Legendre three-term recurrence on [-1, 1]: ``alpha`` (zero: the weight
    is even) and ``beta``, ``beta[0]`` the interval's measure, ``beta[k] = k^2 / (4 k^2
    - 1)``."""
    alpha = np.zeros(count)
    beta = np.empty(count)
    beta[0] = 2.0
    k = np.arange(1, count, dtype=np.float64)
    beta[1:] = 1.0 / (4.0 - 1.0 / (k * k))
    return alpha, beta


def jacobi_kronrod(n: int) -> tuple[np.ndarray, np.ndarray]:
    """The ``2 n + 1`` order Jacobi–Kronrod matrix of the Legendre weight: its diagonal
    ``alpha`` and its squared off-diagonal ``beta`` (``beta[0]`` the measure).

    The first ``floor(3 n / 2) + 1`` diagonal and ``ceil(3 n / 2) + 1`` off-diagonal
    entries are the Gauss–Legendre matrix's; Laurie's recursion fills the rest from two
    rows of mixed moments, ``row`` and ``prev``, which trade places after every step.
    """
    if n < 1:
        raise ValueError(f"a Kronrod extension needs n >= 1, got {n}")
    known_a, known_b = legendre_recurrence(math.ceil(3 * n / 2) + 1)
    alpha = np.zeros(2 * n + 1)
    beta = np.zeros(2 * n + 1)
    alpha[: math.floor(3 * n / 2) + 1] = known_a[: math.floor(3 * n / 2) + 1]
    beta[: math.ceil(3 * n / 2) + 1] = known_b[: math.ceil(3 * n / 2) + 1]
    prev = np.zeros(n // 2 + 2)
    row = np.zeros(n // 2 + 2)
    row[1] = beta[n + 1]
    # Eastward: each new row's entries at k = floor((m+1)/2), ..., 0 accumulated in that
    # order (every entry reads the old rows only).
    for m in range(n - 1):
        k = np.arange((m + 1) // 2, -1, -1)
        other = m - k
        terms = ((alpha[k + n + 1] - alpha[other]) * row[k + 1] + beta[k + n + 1] * prev[k]
                 - beta[other] * prev[k + 1])
        prev[k + 1] = np.cumsum(terms)
        prev, row = row, prev
    # Shift the kept row one place up before the southward sweep.
    prev[1: n // 2 + 2] = prev[0: n // 2 + 1].copy()
    # Southward: each step completes one new diagonal (m even) or off-diagonal (m odd)
    # entry of the lower half.
    for m in range(n - 1, 2 * n - 2):
        k = np.arange(m + 1 - n, (m - 1) // 2 + 1)
        other = m - k
        j = n - 1 - other
        terms = (-(alpha[k + n + 1] - alpha[other]) * row[j + 1] - beta[k + n + 1] * prev[j + 1]
                 + beta[other] * prev[j + 2])
        prev[j + 1] = np.cumsum(terms)
        last = int(j[-1])
        top = (m + 1) // 2
        if m % 2 == 0:
            alpha[top + n + 1] = (alpha[top]
                                  + (prev[last + 1] - beta[top + n + 1] * prev[last + 2])
                                  / row[last + 2])
        else:
            beta[top + n + 1] = prev[last + 1] / prev[last + 2]
        prev, row = row, prev
    alpha[2 * n] = alpha[n - 1] - beta[2 * n] * prev[1] / row[1]
    return alpha, beta


def gauss_kronrod(n: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The ``2 n + 1`` point Gauss–Kronrod rule on [-1, 1]: nodes in increasing order,
    their weights, and the embedded Gauss rule's weights on the same nodes (the Gauss
    nodes are every other Kronrod node from the second; zero elsewhere), as
    ``src/solprop/modes/quadrature.py:265-311`` returns them."""
    alpha, beta = jacobi_kronrod(n)
    nodes, vectors = eigh_tridiagonal(alpha, np.sqrt(beta[1:]))
    weights = beta[0] * vectors[0, :] ** 2
    coarse = np.zeros_like(weights)
    coarse[1::2] = roots_legendre(n)[1]
    return nodes, weights, coarse


def radial_rule(n_r: int, kronrod: bool) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Nodes on ``(0, 1)`` with fine and coarse weights (``quadrature.py:314-332``): the
    Kronrod rule of the smallest n with ``2 n + 1 >= n_r``, or the n_r-point Gauss rule
    whose coarse weights are its own (no embedded rule)."""
    if kronrod:
        nodes, weights, coarse = gauss_kronrod(max(-(-(n_r - 1) // 2), 1))
    else:
        nodes, weights = roots_legendre(n_r)
        coarse = weights
    return (nodes + 1.0) / 2.0, weights / 2.0, coarse / 2.0
