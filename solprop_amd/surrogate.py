"""The learned surrogate of a parameter scan, over the C ABI (``sp_fno_*``).

A 1-D Fourier neural operator (4 Fourier layers, width 64, 64 retained modes by
default) maps a scan point -- its input spectrum on ``n`` bins and its scan parameters --
to the log exit spectrum the solver would produce. It is trained on exit fields the
batched solver (:class:`solprop_amd.native.Scan`) generates, and screens scan points
before the solver runs; the solver stays the fallback and the judge of its error
(:func:`screen`). The network trains and runs on the device, every dense product on the
BF16 matrix cores with FP32 accumulation.

The project's own component: the reference runs every scan point through the solver
(``src/solprop/driver/scan.py:30``).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from solprop_amd import _lib
from solprop_amd.native import BackendError, _check


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass(frozen=True)
class Shape:
    n_params: int = 4
    width: int = 64
    modes: int = 64
    layers: int = 4
    hidden: int = 128
    n: int = 2048

    @property
    def in_ch(self) -> int:
        return self.n_params + 2


class Network:
    """The device network: parameters, forward, gradient, Adam training."""

    def __init__(self, shape: Shape = Shape(), max_batch: int = 64):
        self._lib = _lib.lib()
        self.shape = shape
        self.max_batch = max_batch
        s = _lib.FnoShapeC(shape.n_params, shape.width, shape.modes, shape.layers, shape.hidden,
                           shape.n)
        handle = C.c_void_p()
        msg = C.create_string_buffer(1024)
        _check(self._lib.sp_fno_create(C.byref(s), max_batch, C.byref(handle), msg, 1024), msg)
        self._h = handle
        self.n_parameters = int(self._lib.sp_fno_parameter_count(handle))

    def close(self) -> None:
        if self._h is not None:
            self._lib.sp_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        if self._h is None:
            raise BackendError("the network is closed")
        msg = C.create_string_buffer(1024)
        _check(getattr(self._lib, name)(self._h, *args, msg, 1024), msg)

    def set_parameters(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.shape != (self.n_parameters,):
            raise ValueError(f"{self.n_parameters} parameters expected")
        self._call("sp_fno_set_parameters", _pf(flat))

    def parameters(self) -> np.ndarray:
        out = np.empty(self.n_parameters, dtype=np.float32)
        self._call("sp_fno_get_parameters", _pf(out))
        return out

    def _inputs(self, spectra, params):
        spectra = np.ascontiguousarray(spectra, dtype=np.float32).reshape(-1, self.shape.n)
        params = np.ascontiguousarray(params, dtype=np.float32).reshape(len(spectra),
                                                                        self.shape.n_params)
        return spectra, params

    def forward(self, spectra: np.ndarray, params: np.ndarray) -> np.ndarray:
        """Outputs (batch, n), in chunks of ``max_batch``."""
        spectra, params = self._inputs(spectra, params)
        out = np.empty_like(spectra)
        for i in range(0, len(spectra), self.max_batch):
            j = min(len(spectra), i + self.max_batch)
            chunk = np.empty((j - i, self.shape.n), dtype=np.float32)
            self._call("sp_fno_forward", _pf(spectra[i:j]), _pf(params[i:j]), j - i, _pf(chunk))
            out[i:j] = chunk
        return out

    def gradient(self, spectra, params, target) -> tuple[float, np.ndarray]:
        spectra, params = self._inputs(spectra, params)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(spectra.shape)
        grad = np.empty(self.n_parameters, dtype=np.float32)
        loss = C.c_double()
        self._call("sp_fno_gradient", _pf(spectra), _pf(params), _pf(target), len(spectra),
                   C.byref(loss), _pf(grad))
        return loss.value, grad

    def load_dataset(self, spectra, params, targets) -> None:
        spectra, params = self._inputs(spectra, params)
        targets = np.ascontiguousarray(targets, dtype=np.float32).reshape(spectra.shape)
        self._call("sp_fno_load_dataset", _pf(spectra), _pf(params), _pf(targets), len(spectra))
        self.n_rows = len(spectra)

    def train(self, epochs: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
              rows: int | None = None) -> tuple[np.ndarray, float]:
        """Adam over shuffled minibatches of the loaded dataset; the per-step losses and
        the device time, ms."""
        rng = np.random.default_rng(seed)
        n = rows or self.n_rows
        per_epoch = n // batch
        idx = np.concatenate([rng.permutation(n)[: per_epoch * batch] for _ in range(epochs)])
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        steps = len(idx) // batch
        losses = np.empty(steps, dtype=np.float32)
        ms = C.c_double()
        self._call("sp_fno_train", idx.ctypes.data_as(C.POINTER(C.c_int)), steps, batch,
                   C.c_float(lr), _pf(losses), C.byref(ms))
        return losses, ms.value


def testing_gemm_bf16(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``a @ b`` with both operands rounded to BF16, FP32 accumulation, on the matrix
    cores."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    (m, k), (k2, n) = a.shape, b.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    out = np.empty((m, n), dtype=np.float32)
    msg = C.create_string_buffer(1024)
    _check(_lib.lib().sp_testing_gemm_bf16(m, n, k, _pf(a), _pf(b), _pf(out), msg, 1024), msg)
    return out


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def oracle_inputs(spectra: np.ndarray, params: np.ndarray) -> np.ndarray:
    """The network's input channels as one array (batch, in_ch, n): the spectrum, each
    parameter broadcast, the grid coordinate s / n (what the device's lift forms)."""
    spectra = np.asarray(spectra, np.float64)
    params = np.asarray(params, np.float64)
    b, n = spectra.shape
    grid = np.broadcast_to(np.arange(n) / n, (b, 1, n))
    prm = np.broadcast_to(params[:, :, None], (b, params.shape[1], n))
    return np.concatenate([spectra[:, None, :], prm, grid], axis=1)
