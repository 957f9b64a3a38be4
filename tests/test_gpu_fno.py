"""GPU: the learned surrogate's kernels against the FP64 oracle (``oracle/fno.py``).

The device computes in BF16 with FP32 accumulation, the oracle in FP64 on the same FP32
parameters: outputs agree to the rounding of BF16 activations (a few parts in 10^3 per
layer), so the tolerances are a relative L2 error of 2 % with a cosine of at least
0.9995 for every gradient tensor (measured: below 0.7 % and above 0.99997) -- a wrong
sign or index in any backward kernel turns a tensor's cosine negative or near zero. The BF16 product itself is exact on
integer operands small enough for BF16 and FP32 (a lane-layout check of the matrix-core
builtin) and within the rounding bound on random ones.
"""

import numpy as np
import pytest

from oracle import fno as oracle
from solprop_amd import surrogate

pytestmark = pytest.mark.gpu


def test_bf16_gemm_layout_is_exact_on_integers():
    print("bf16 gemm on integer operands", flush=True)
    rng = np.random.default_rng(0)
    for m, n, k in [(16, 16, 16), (64, 64, 32), (70, 130, 45), (1, 2048, 128), (128, 64, 2048)]:
        a = rng.integers(-8, 9, (m, k)).astype(np.float32)
        b = rng.integers(-8, 9, (k, n)).astype(np.float32)
        got = surrogate.testing_gemm_bf16(a, b)
        assert np.array_equal(got, a @ b), (m, n, k)


def test_bf16_gemm_random_within_rounding():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((96, 300)).astype(np.float32)
    b = rng.standard_normal((300, 200)).astype(np.float32)
    got = surrogate.testing_gemm_bf16(a, b)
    exact = surrogate.to_bf16(a).astype(np.float64) @ surrogate.to_bf16(b).astype(np.float64)
    bound = 300 * 2.0**-23 * (np.abs(a) @ np.abs(b)) * 4
    assert np.all(np.abs(got - exact) <= bound)


def case(shape, batch, seed):
    rng = np.random.default_rng(seed)
    o = oracle.Shape(in_ch=shape.in_ch, width=shape.width, modes=shape.modes,
                     layers=shape.layers, hidden=shape.hidden, n=shape.n)
    flat = oracle.init(o, seed=seed)
    # Spectral weights scaled up so the Fourier path carries weight in the output.
    p = oracle.unflatten(o, flat.astype(np.float64))
    for layer in range(o.layers):
        p[f"r_re{layer}"] *= 16.0
        p[f"r_im{layer}"] *= 16.0
    flat = oracle.flatten(o, p).astype(np.float32)
    t = np.linspace(0, 1, shape.n)
    spectra = np.stack([np.sin(2 * np.pi * (2 + b) * t) + 0.3 * rng.standard_normal(shape.n)
                        for b in range(batch)]).astype(np.float32)
    params = rng.uniform(0, 1, (batch, shape.n_params)).astype(np.float32)
    target = rng.standard_normal((batch, shape.n)).astype(np.float32)
    return o, flat, spectra, params, target


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


SHAPES = [surrogate.Shape(n_params=4, width=16, modes=8, layers=2, hidden=32, n=256),
          surrogate.Shape()]


@pytest.mark.parametrize("shape", SHAPES, ids=["small", "full"])
def test_forward_and_gradient_match_oracle(shape):
    batch = 3 if shape.n <= 256 else 2
    print(f"fno forward/gradient vs oracle, {shape}, batch {batch}", flush=True)
    o, flat, spectra, params, target = case(shape, batch, seed=7)
    net = surrogate.Network(shape, max_batch=4)
    net.set_parameters(flat)
    assert np.array_equal(net.parameters(), flat)
    got = net.forward(spectra, params)
    x = surrogate.oracle_inputs(spectra, params)
    loss_ref, grad_ref, out_ref = oracle.loss_and_grad(o, flat.astype(np.float64), x, target)
    err = rel(got, out_ref)
    print("forward rel L2", err, flush=True)
    assert err < 2e-2
    loss, grad = net.gradient(spectra, params, target)
    net.close()
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    at = 0
    worst = {}
    for name, shp in oracle.layout(o):
        size = int(np.prod(shp))
        g, r = grad[at:at + size].astype(np.float64), grad_ref[at:at + size]
        at += size
        cos = float(g @ r / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-300))
        worst[name] = (rel(g, r), cos)
        assert cos > 0.9995, (name, cos)
        assert rel(g, r) < 2e-2, (name, rel(g, r))
    print("gradient rel L2 / cosine per tensor", worst, flush=True)


def test_training_reduces_the_loss():
    """A learnable map (the target is a smoothed, parameter-shifted copy of the input):
    Adam on the device brings the loss down by an order of magnitude."""
    print("fno training on a synthetic map", flush=True)
    shape = surrogate.Shape(n_params=4, width=32, modes=16, layers=2, hidden=64, n=512)
    rng = np.random.default_rng(3)
    rows = 512
    t = np.linspace(0, 1, shape.n)
    params = rng.uniform(0, 1, (rows, 4)).astype(np.float32)
    centre = 0.3 + 0.4 * params[:, :1]
    spectra = np.exp(-((t[None, :] - centre) / 0.08) ** 2).astype(np.float32)
    target = (params[:, 1:2] * np.exp(-((t[None, :] - centre - 0.1 * params[:, 2:3]) / 0.12) ** 2)
              + 0.2 * params[:, 3:4]).astype(np.float32)
    net = surrogate.Network(shape, max_batch=32)
    o = oracle.Shape(in_ch=shape.in_ch, width=shape.width, modes=shape.modes,
                     layers=shape.layers, hidden=shape.hidden, n=shape.n)
    net.set_parameters(oracle.init(o, seed=1))
    net.load_dataset(spectra, params, target)
    losses, ms = net.train(epochs=20, batch=32, lr=2e-3, seed=0)
    print("loss first/last", losses[:3], losses[-3:], "device ms", ms, flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-16:].mean() < 0.1 * losses[:16].mean()
    # The trained network's forward agrees with the last steps' loss.
    pred = net.forward(spectra[:64], params[:64])
    net.close()
    assert np.mean((pred - target[:64]) ** 2) < 0.2 * losses[:16].mean()
