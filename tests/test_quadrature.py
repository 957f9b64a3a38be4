"""The Gauss–Kronrod radial rule and the quadrature estimate's oracle, on the host.

The rule (``solprop_amd/quadrature.py``, restating ``src/solprop/modes/quadrature.py:
199-332``) is pinned by QUADPACK's published 15-point table, by its degree of exactness
(through 3n + 1, 3n + 2 for odd n, and no further), by the embedded rule being the
n-point Gauss rule and by the reference's own docstring example. The oracle's estimate
(``oracle/stats.quadrature_error``) is pinned by a Kerr response launched in one mode,
which is separable in r and t: each mode's projected spectrum is its radial integral
times one common spectrum, so the estimate reduces to the radial integrals' relative
difference over the modes.
"""

import math

import numpy as np
import pytest
from numpy.polynomial import legendre

from oracle import stats as S
from solprop_amd.problem import capillary_problem
from solprop_amd.quadrature import gauss_kronrod, radial_rule

# QUADPACK's qk15 (R. Piessens et al., 1983): the positive Kronrod nodes from the
# largest, their weights, and the 7-point Gauss weights from the outermost node.
XGK15 = [0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
         0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
         0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
         0.207784955007898467600689403773245, 0.0]
WGK15 = [0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
         0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
         0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
         0.204432940075298892414161999234649, 0.209482141084727828012999174891714]
WG7 = [0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
       0.381830050505118944950369775488975, 0.417959183673469387755102040816327]


def mirrored(half):
    """A symmetric rule's 15 values from its 8 for nodes >= 0, increasing nodes."""
    return np.array(half[:-1] + [half[-1]] + half[-2::-1])


def test_fifteen_point_rule_matches_quadpack():
    nodes, weights, coarse = gauss_kronrod(7)
    assert np.max(np.abs(nodes - np.array([-v for v in XGK15[:-1]] + XGK15[::-1]))) < 1e-14
    assert np.max(np.abs(weights - mirrored(WGK15))) < 1e-14
    gauss = np.zeros(15)
    gauss[1::2] = WG7[:-1] + [WG7[-1]] + WG7[-2::-1]
    assert np.max(np.abs(coarse - gauss)) < 1e-15


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 13, 20, 64])
def test_degree_of_exactness(n):
    """Exact for every Legendre polynomial through degree 3n + 1 (3n + 2 for odd n, whose
    odd degree the symmetry integrates), and not for the next even degree."""
    nodes, weights, _ = gauss_kronrod(n)
    top = 3 * n + 1 if n % 2 == 0 else 3 * n + 2
    for k in range(top + 1):
        exact = 2.0 if k == 0 else 0.0
        assert abs(weights @ legendre.legval(nodes, [0] * k + [1]) - exact) < 1e-13, k
    assert abs(weights @ legendre.legval(nodes, [0] * (top + 1) + [1])) > 1e-8


@pytest.mark.parametrize("n", [3, 7, 64])
def test_embedded_rule_is_gauss(n):
    nodes, weights, coarse = gauss_kronrod(n)
    g_nodes, g_weights = legendre.leggauss(n)
    assert np.max(np.abs(nodes[1::2] - g_nodes)) < 1e-14
    assert np.max(np.abs(coarse[1::2] - g_weights)) < 1e-14
    assert np.all(coarse[0::2] == 0.0)
    assert np.all(weights > 0.0) and np.all(np.diff(nodes) > 0.0)


def test_reference_docstring_example():
    """``quadrature.py:294-299``: 15 nodes, 7 coarse, weights summing to 2 within 16 eps."""
    nodes, weights, coarse = gauss_kronrod(7)
    assert nodes.size == 15 and int((coarse > 0.0).sum()) == 7
    assert abs(weights.sum() - 2.0) < 16 * np.spacing(2.0)


def test_radial_rule_rounds_up_to_a_kronrod_count():
    xi, w, wc = radial_rule(32, True)
    assert xi.size == 33 and np.all((xi > 0.0) & (xi < 1.0))
    assert abs(w.sum() - 1.0) < 1e-14 and abs(wc.sum() - 1.0) < 1e-14
    xi, w, wc = radial_rule(32, False)
    assert xi.size == 32 and np.array_equal(w, wc)


def kerr_problem(rule):
    """Six modes on 17 Kronrod nodes: the higher modes' profiles separate the two rules
    (an estimate near 1e-2)."""
    return capillary_problem(n_t=1024, n_nodes=16, n_modes=6, plasma=False,
                             integration="trapezoid", radial_rule=rule)


def test_estimate_of_a_separable_kerr_response():
    """One launched mode, Kerr only: P(r_q, t) = c rho S_q0^3 e(t)^3, so mode m's projected
    spectrum is I_m X with I_m = sum_q W_mq S_q0^3 (fine) or the coarse rule's; every
    mode shares X and the weight, and the estimate is |I_coarse - I| / |I| over modes."""
    p = kerr_problem("gauss_kronrod")
    field = p.field0.copy()
    assert np.count_nonzero(np.abs(field).sum(axis=1)) == 1  # the launch is HE11 alone
    cube = p.synthesis[:, 0] ** 3
    fine = p.projection @ cube
    coarse = p.projection_coarse @ cube
    expected = math.sqrt(np.sum((coarse - fine) ** 2) / np.sum(fine ** 2))
    got = S.quadrature_error(p, field, 0.0)
    assert expected > 1e-3  # the two rules differ here
    assert got == pytest.approx(expected, rel=1e-9)


def test_estimate_without_an_embedded_rule_is_nan():
    p = kerr_problem("gauss")
    assert p.projection_coarse is None
    assert math.isnan(S.quadrature_error(p, p.field0, 0.0))
