"""The matrix-core FP64 GEMM of the radial route's Hankel transforms, through the C ABI.

Against numpy's FP64 product on the same seeded operands. Tolerance (class R): each
entry is a sum of K products, so its rounding is at most about K eps |A| |B| (Higham's
gamma_K); the Frobenius error relative to |A|_F |B|_F is held to 4 K eps.
"""

import numpy as np
import pytest

from solprop_amd.native import testing_gemm as device_gemm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (64, 64, 16), (100, 70, 33), (128, 256, 512),
                                   (512, 8192, 512)])
def test_gemm_matches_numpy(m, n, k):
    print(f"gemm {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a = rng.normal(size=(m, k))
    b = rng.normal(size=(k, n))
    got = device_gemm(a, b)
    expected = a @ b
    scale = np.linalg.norm(a) * np.linalg.norm(b)
    assert np.linalg.norm(got - expected) <= 4 * k * np.finfo(np.float64).eps * scale


def test_gemm_layout_on_structured_operands():
    """Each output entry identifies its row and column: a wrong MFMA lane map moves it."""
    print("gemm layout", flush=True)
    m, n, k = 48, 40, 12
    a = np.zeros((m, k))
    b = np.zeros((k, n))
    # C[i, j] = 1000 i + j + 7, exactly representable.
    a[:, 0] = np.arange(m)
    b[0, :] = 1000.0
    a[:, 1] = 1.0
    b[1, :] = np.arange(n)
    a[:, 2] = 7.0
    b[2, :] = 1.0
    got = device_gemm(a, b)
    i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    assert np.array_equal(got, i * 1000.0 + j + 7.0)
