"""The surrogate's CPU oracle (``oracle/fno.py``), pinned before it judges the device.

- The truncated transforms equal numpy's rfft (the first modes) and irfft (of a spectrum
  zero above them).
- The analytic gradient equals central differences of the loss in FP64, parameter by
  parameter (a sample of each tensor), on a small network; a planted mistake (the
  spectral weights' imaginary part with its sign flipped in the backward pass) fails.
- gelu's derivative equals its central difference.
"""

import numpy as np
import pytest

from oracle import fno


def small():
    return fno.Shape(in_ch=3, width=6, modes=5, layers=3, hidden=7, n=32)


def test_transforms_are_rfft_and_irfft():
    f, g = fno.transforms(64, 9)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((3, 64))
    got = v @ f
    ref = np.fft.rfft(v)[:, :9]
    assert np.allclose(got[:, :9], ref.real, atol=1e-12)
    assert np.allclose(got[:, 9:], ref.imag, atol=1e-12)
    y = rng.standard_normal((3, 9)) + 1j * rng.standard_normal((3, 9))
    full = np.zeros((3, 33), dtype=complex)
    full[:, :9] = y
    assert np.allclose(np.concatenate([y.real, y.imag], axis=1) @ g,
                       np.fft.irfft(full, n=64), atol=1e-13)


def test_gelu_gradient():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (fno.gelu(x + h) - fno.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - fno.gelu_grad(x))) < 1e-8


def _fd_check(backward):
    s = small()
    rng = np.random.default_rng(2)
    flat = fno.init(s, seed=3).astype(np.float64)
    # Larger spectral weights than the initial scale, so their terms matter.
    p = fno.unflatten(s, flat)
    for layer in range(s.layers):
        p[f"r_re{layer}"] *= 20.0
        p[f"r_im{layer}"] *= 20.0
    flat = fno.flatten(s, p)
    x = rng.standard_normal((2, s.in_ch, s.n))
    target = rng.standard_normal((2, s.n))
    out, cache = fno.forward(s, flat, x)
    dout = 2.0 * (out - target) / out.size
    grad = backward(s, flat, cache, dout)
    worst = 0.0
    at = 0
    for name, shape in fno.layout(s):
        size = int(np.prod(shape))
        for j in rng.choice(size, size=min(size, 4), replace=False):
            i = at + j
            h = 1e-6 * max(1.0, abs(flat[i]))
            up, down = flat.copy(), flat.copy()
            up[i] += h
            down[i] -= h
            fd = (np.mean((fno.forward(s, up, x)[0] - target) ** 2)
                  - np.mean((fno.forward(s, down, x)[0] - target) ** 2)) / (2 * h)
            scale = max(abs(fd), 1e-6)
            worst = max(worst, abs(fd - grad[i]) / scale)
        at += size
    return worst


def test_backward_matches_central_differences():
    assert _fd_check(fno.backward) < 1e-5


def test_a_flipped_spectral_gradient_fails():
    def planted(s, flat, cache, dout):
        grad = fno.backward(s, flat, cache, dout)
        out, at = grad.copy(), 0
        for name, shape in fno.layout(s):
            size = int(np.prod(shape))
            if name.startswith("r_im"):
                out[at:at + size] *= -1.0
            at += size
        return out

    assert _fd_check(planted) > 1e-2


def test_layout_counts():
    s = fno.Shape()
    n = fno.count(s)
    assert n == 64 * 6 + 64 + 4 * (2 * 64 * 64 * 64 + 64 * 64 + 64) + 128 * 64 + 128 + 128 + 1
    flat = fno.init(s, seed=0)
    assert flat.dtype == np.float32 and len(flat) == n
    assert np.array_equal(fno.flatten(s, fno.unflatten(s, flat)), flat.astype(np.float64))


@pytest.mark.parametrize("layers", [1, 2])
def test_linear_network_is_linear(layers):
    """With the hidden gelu bypassed (one layer, no nonlinearity before the head), the
    map from the lift's output to the last layer is linear: doubling the input's
    non-bias part doubles the change in the last layer's pre-activation."""
    s = fno.Shape(in_ch=2, width=4, modes=3, layers=layers, hidden=5, n=16)
    flat = fno.init(s, seed=5).astype(np.float64)
    p = fno.unflatten(s, flat)
    p["lift_b"][:] = 0.0
    for layer in range(layers):
        p[f"b{layer}"][:] = 0.0
    flat = fno.flatten(s, p)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1, 2, 16))
    if layers == 1:
        _, c1 = fno.forward(s, flat, x)
        _, c2 = fno.forward(s, flat, 2.0 * x)
        assert np.allclose(c2["z"][0], 2.0 * c1["z"][0], atol=1e-12)
    else:
        _, c1 = fno.forward(s, flat, x)
        assert np.allclose(c1["v"][1], fno.gelu(c1["z"][0]), atol=1e-14)
