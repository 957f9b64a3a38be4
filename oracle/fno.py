"""CPU FP64 restatement of the learned surrogate: a 1-D Fourier neural operator.

Test infrastructure only (the checker of the device's forward and backward kernels):
nothing under ``solprop_amd`` imports it.

The network (Li et al., "Fourier Neural Operator for Parametric Partial Differential
Equations", ICLR 2021, its 1-D form) maps ``in_ch`` input channels on ``n`` points to one
output channel:

    v0 = P x + p                                   (lift, a 1x1 convolution)
    z_l = W_l v_l + b_l + K_l v_l,  v_{l+1} = gelu(z_l)  (l < L-1; the last layer linear)
    K_l v = irfft_n( R_l . rfft(v)[:modes] )       (per retained mode k: a complex
                                                    width x width matrix R_l[k])
    out = Q2 gelu(Q1 v_L + q1) + q2                (projection through `hidden`)

The truncated transforms are dense matrices (what the device runs as matrix products):
``F`` (n, 2 modes) with ``F[s, k] = cos(2 pi k s / n)`` and ``F[s, modes + k] =
-sin(2 pi k s / n)``, so that ``v F`` is the first ``modes`` bins of ``rfft(v)`` (real
parts, then imaginary parts); and ``G`` (2 modes, n) with ``G[k, s] = c_k cos(2 pi k s /
n) / n`` and ``G[modes + k, s] = -c_k sin(2 pi k s / n) / n``, c_0 = 1, c_k = 2, which is
``irfft`` of a spectrum whose bins at and above ``modes`` are zero (the DC bin's
imaginary part drops out, as ``irfft`` discards it). gelu is the tanh form,
``0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))``.

The parameters travel as one float32 vector in the order of :func:`layout`. The loss is
the mean squared error over the batch and the points. ``backward`` is pinned by central
differences in FP64 (``tests/test_fno.py``); ``rfft``/``irfft`` pin the transforms.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GELU_C = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


@dataclass(frozen=True)
class Shape:
    in_ch: int = 6
    width: int = 64
    modes: int = 64
    layers: int = 4
    hidden: int = 128
    n: int = 2048


def layout(s: Shape) -> list[tuple[str, tuple[int, ...]]]:
    """The parameters' names and shapes in the order of the flat vector."""
    out = [("lift_w", (s.width, s.in_ch)), ("lift_b", (s.width,))]
    for layer in range(s.layers):
        out += [(f"r_re{layer}", (s.modes, s.width, s.width)),
                (f"r_im{layer}", (s.modes, s.width, s.width)),
                (f"w{layer}", (s.width, s.width)), (f"b{layer}", (s.width,))]
    out += [("q1_w", (s.hidden, s.width)), ("q1_b", (s.hidden,)),
            ("q2_w", (1, s.hidden)), ("q2_b", (1,))]
    return out


def count(s: Shape) -> int:
    return sum(int(np.prod(shape)) for _, shape in layout(s))


def unflatten(s: Shape, flat: np.ndarray) -> dict[str, np.ndarray]:
    out, at = {}, 0
    for name, shape in layout(s):
        size = int(np.prod(shape))
        out[name] = np.asarray(flat[at:at + size], dtype=np.float64).reshape(shape)
        at += size
    assert at == len(flat)
    return out


def flatten(s: Shape, params: dict[str, np.ndarray]) -> np.ndarray:
    return np.concatenate([np.asarray(params[name], np.float64).ravel()
                           for name, _ in layout(s)])


def init(s: Shape, seed: int = 0) -> np.ndarray:
    """Float32 parameters: 1x1 convolutions uniform in +-1/sqrt(fan_in), the spectral
    weights uniform in [0, 1/(width^2)) times a random sign (the FNO's own scale)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in layout(s):
        if name.startswith("r_"):
            scale = 1.0 / (s.width * s.width)
            parts.append(scale * rng.uniform(-1.0, 1.0, shape))
        else:
            fan_in = shape[-1] if len(shape) > 1 else {"lift_b": s.in_ch, "q1_b": s.width,
                                                      "q2_b": s.hidden}.get(name, s.width)
            bound = 1.0 / math.sqrt(fan_in)
            parts.append(rng.uniform(-bound, bound, shape))
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def transforms(n: int, modes: int) -> tuple[np.ndarray, np.ndarray]:
    """F (n, 2 modes) and G (2 modes, n) of the module docstring, FP64."""
    s = np.arange(n)[:, None]
    k = np.arange(modes)[None, :]
    angle = 2.0 * np.pi * ((s * k) % n) / n
    f = np.concatenate([np.cos(angle), -np.sin(angle)], axis=1)
    c = np.where(np.arange(modes) == 0, 1.0, 2.0)[:, None]
    g = np.concatenate([c * np.cos(angle.T), -c * np.sin(angle.T)], axis=0) / n
    return f, g


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x**3)))


def gelu_grad(x):
    u = GELU_C * (x + GELU_A * x**3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def forward(s: Shape, flat: np.ndarray, x: np.ndarray):
    """x (batch, in_ch, n) -> out (batch, n) and the cache backward needs."""
    p = unflatten(s, flat)
    f, g = transforms(s.n, s.modes)
    x = np.asarray(x, np.float64)
    v = np.einsum("oi,bin->bon", p["lift_w"], x) + p["lift_b"][None, :, None]
    cache = {"x": x, "v": [], "z": [], "vhat": []}
    for layer in range(s.layers):
        vhat = v @ f  # (b, width, 2 modes)
        vr, vi = vhat[..., : s.modes], vhat[..., s.modes:]
        rr, ri = p[f"r_re{layer}"], p[f"r_im{layer}"]  # (modes, in, out)
        yr = np.einsum("bik,kio->bok", vr, rr) - np.einsum("bik,kio->bok", vi, ri)
        yi = np.einsum("bik,kio->bok", vr, ri) + np.einsum("bik,kio->bok", vi, rr)
        spec = np.concatenate([yr, yi], axis=-1) @ g
        z = np.einsum("oi,bin->bon", p[f"w{layer}"], v) + p[f"b{layer}"][None, :, None] + spec
        cache["v"].append(v)
        cache["z"].append(z)
        cache["vhat"].append(vhat)
        v = gelu(z) if layer < s.layers - 1 else z
    h1 = np.einsum("oi,bin->bon", p["q1_w"], v) + p["q1_b"][None, :, None]
    h = gelu(h1)
    out = np.einsum("oi,bin->bon", p["q2_w"], h)[:, 0, :] + p["q2_b"][0]
    cache.update(vL=v, h1=h1, h=h)
    return out, cache


def backward(s: Shape, flat: np.ndarray, cache, dout: np.ndarray) -> np.ndarray:
    """The gradient of sum(dout * out) with respect to the flat parameters."""
    p = unflatten(s, flat)
    f, g = transforms(s.n, s.modes)
    grads = {}
    dout = np.asarray(dout, np.float64)
    grads["q2_w"] = np.einsum("bn,bhn->h", dout, cache["h"])[None, :]
    grads["q2_b"] = np.array([dout.sum()])
    dh = p["q2_w"][0][None, :, None] * dout[:, None, :]
    dh1 = dh * gelu_grad(cache["h1"])
    grads["q1_w"] = np.einsum("bhn,bin->hi", dh1, cache["vL"])
    grads["q1_b"] = dh1.sum(axis=(0, 2))
    dv = np.einsum("hi,bhn->bin", p["q1_w"], dh1)
    for layer in reversed(range(s.layers)):
        z = cache["z"][layer]
        dz = dv * gelu_grad(z) if layer < s.layers - 1 else dv
        v = cache["v"][layer]
        grads[f"w{layer}"] = np.einsum("bon,bin->oi", dz, v)
        grads[f"b{layer}"] = dz.sum(axis=(0, 2))
        dv = np.einsum("oi,bon->bin", p[f"w{layer}"], dz)
        dy = dz @ g.T  # (b, width, 2 modes)
        dyr, dyi = dy[..., : s.modes], dy[..., s.modes:]
        vhat = cache["vhat"][layer]
        vr, vi = vhat[..., : s.modes], vhat[..., s.modes:]
        rr, ri = p[f"r_re{layer}"], p[f"r_im{layer}"]
        grads[f"r_re{layer}"] = (np.einsum("bik,bok->kio", vr, dyr)
                                 + np.einsum("bik,bok->kio", vi, dyi))
        grads[f"r_im{layer}"] = (np.einsum("bik,bok->kio", vr, dyi)
                                 - np.einsum("bik,bok->kio", vi, dyr))
        dvr = np.einsum("bok,kio->bik", dyr, rr) + np.einsum("bok,kio->bik", dyi, ri)
        dvi = -np.einsum("bok,kio->bik", dyr, ri) + np.einsum("bok,kio->bik", dyi, rr)
        dv = dv + np.concatenate([dvr, dvi], axis=-1) @ f.T
    grads["lift_w"] = np.einsum("bon,bin->oi", dv, cache["x"])
    grads["lift_b"] = dv.sum(axis=(0, 2))
    return flatten(s, grads)


def loss_and_grad(s: Shape, flat: np.ndarray, x: np.ndarray, target: np.ndarray):
    """Mean squared error over (batch, n) and its gradient."""
    out, cache = forward(s, flat, x)
    diff = out - np.asarray(target, np.float64)
    loss = float(np.mean(diff * diff))
    dout = 2.0 * diff / diff.size
    return loss, backward(s, flat, cache, dout), out
