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


# ------------------------------------------------------------- the scan it screens

# The scan's parameters: helium fill pressure (bar, uniform), pulse energy (J,
# log-uniform), FWHM duration (s, uniform) and capillary length (m, uniform).
RANGES = {"pressure": (0.2, 3.0), "energy": (20e-6, 500e-6), "fwhm": (5e-15, 30e-15),
          "length": (0.5, 3.0)}
# The grid every point shares: 2^12 samples over 400 fs (2049 bins, the first 2048 the
# network's), one HE_11 mode on a 32-node quadrature, the band 150 nm - 4 um.
GRID = {"n_t": 4096, "window": 400e-15, "n_modes": 1, "n_nodes": 32, "band": (150e-9, 4e-6)}
BINS = 2048
FLOOR = 1e-8  # spectral density floor, relative to the launch peak


def draw(n: int, rng: np.random.Generator, ranges: dict = RANGES) -> np.ndarray:
    """n scan points (pressure, energy, fwhm, length), physical units."""
    p = rng.uniform(*ranges["pressure"], n)
    e = np.exp(rng.uniform(np.log(ranges["energy"][0]), np.log(ranges["energy"][1]), n))
    f = rng.uniform(*ranges["fwhm"], n)
    ell = rng.uniform(*ranges["length"], n)
    return np.stack([p, e, f, ell], axis=1)


def normalise(points: np.ndarray, ranges: dict = RANGES) -> np.ndarray:
    """The network's parameter inputs: each range mapped onto [0, 1] (energy by its log)."""
    p = np.asarray(points, np.float64)
    out = np.empty_like(p)
    lo, hi = ranges["pressure"]
    out[:, 0] = (p[:, 0] - lo) / (hi - lo)
    lo, hi = ranges["energy"]
    out[:, 1] = (np.log(p[:, 1]) - np.log(lo)) / (np.log(hi) - np.log(lo))
    lo, hi = ranges["fwhm"]
    out[:, 2] = (p[:, 2] - lo) / (hi - lo)
    lo, hi = ranges["length"]
    out[:, 3] = (p[:, 3] - lo) / (hi - lo)
    return out.astype(np.float32)


def scan_problem(pressure: float, energy: float, fwhm: float, length: float):
    """One point's propagation problem: the capillary at a constant fill."""
    from solprop_amd.problem import capillary_problem

    return capillary_problem(n_t=GRID["n_t"], window=GRID["window"], n_modes=GRID["n_modes"],
                             n_nodes=GRID["n_nodes"], band=GRID["band"], p0=float(pressure),
                             p1=float(pressure), energy=float(energy), fwhm=float(fwhm),
                             length=float(length))


def log_spectrum(field: np.ndarray, peak: float) -> np.ndarray:
    """The network's spectral units: log10 of the modes' summed spectral density over
    the launch peak (floored at FLOOR), mapped so that FLOOR .. 1 is 0 .. 1."""
    density = np.sum(np.abs(np.asarray(field).reshape(GRID["n_modes"], -1)) ** 2, axis=0)
    return ((np.log10(density[:BINS] / peak + FLOOR) + 8.0) / 8.0).astype(np.float32)


def density(log_spec: np.ndarray) -> np.ndarray:
    """Spectral density over the launch peak from the network's units."""
    return np.maximum(10.0 ** (8.0 * np.asarray(log_spec, np.float64) - 8.0) - FLOOR, 0.0)


def active_bins() -> np.ndarray:
    """The bins of the network's spectrum inside the grid's band."""
    df = 1.0 / GRID["window"]
    f = np.arange(BINS) * df
    c = 299792458.0
    return (f >= c / GRID["band"][1]) & (f <= c / GRID["band"][0])


def uv_fraction(log_spec: np.ndarray, band=(150e-9, 300e-9)) -> np.ndarray:
    """The screening objective: the share of the exit energy between 150 and 300 nm
    (a resonant dispersive wave's band), from spectra in the network's units (rows)."""
    s = density(np.atleast_2d(log_spec))
    df = 1.0 / GRID["window"]
    f = np.arange(BINS) * df
    c = 299792458.0
    uv = (f >= c / band[1]) & (f <= c / band[0])
    act = active_bins()
    return s[:, uv].sum(axis=1) / np.maximum(s[:, act].sum(axis=1), 1e-300)


def spectral_error(pred: np.ndarray, true: np.ndarray) -> np.ndarray:
    """The spectral error the solver judges, per row: the RMS difference of the log10
    spectral densities (floored at FLOOR of the launch peak) over the band, in decades."""
    act = active_bins()
    d = np.atleast_2d(pred)[:, act].astype(np.float64) - np.atleast_2d(true)[:, act]
    return 8.0 * np.sqrt(np.mean(d * d, axis=1))


def density_error(pred: np.ndarray, true: np.ndarray) -> np.ndarray:
    """Relative L2 error of the linear spectral density over the band, per row."""
    a, b = density(np.atleast_2d(pred)), density(np.atleast_2d(true))
    act = active_bins()
    return (np.linalg.norm(a[:, act] - b[:, act], axis=1)
            / np.maximum(np.linalg.norm(b[:, act], axis=1), 1e-300))


def predict(net: "Network", spectra: np.ndarray, params: np.ndarray) -> np.ndarray:
    """Exit spectra (network units): the launch spectrum plus the network's output, which
    is trained on the change (exit minus launch)."""
    spectra = np.asarray(spectra, np.float32)
    return spectra + net.forward(spectra, params)


def reached(z: np.ndarray, target: np.ndarray) -> np.ndarray:
    """A march has landed on its target: within 8 ulps of it, or past it."""
    z, target = np.asarray(z, np.float64), np.asarray(target, np.float64)
    return z >= target - 8.0 * np.spacing(np.maximum(np.abs(z), np.abs(target)))


def launch_fields(fwhm: np.ndarray, energy: np.ndarray | float = 1.0) -> np.ndarray:
    """The launch spectra the scan's problems carry (the capillary builder's Gaussian at
    800 nm, sqrt(2 P(t)) cos(w0 t), its real transform on the band), one row a point."""
    from solprop_amd.problem import C_LIGHT

    n_t, window = GRID["n_t"], GRID["window"]
    dt = window / n_t
    t = (np.arange(n_t) - n_t // 2) * dt
    fwhm = np.atleast_1d(np.asarray(fwhm, np.float64))[:, None]
    energy = np.broadcast_to(np.asarray(energy, np.float64), (len(fwhm),))[:, None]
    p_peak = energy / (fwhm * np.sqrt(np.pi / (4 * np.log(2))))
    power = p_peak * np.exp(-4 * np.log(2) * t[None, :] ** 2 / fwhm**2)
    a_t = np.sqrt(2 * power) * np.cos(2 * np.pi * C_LIGHT / 800e-9 * t[None, :])
    return np.fft.rfft(a_t, axis=1) * _band_mask()[None, :]


_BAND = []


def _band_mask() -> np.ndarray:
    """The shared grid's active bins, as its problems carry them."""
    if not _BAND:
        _BAND.append(scan_problem(1.0, 1e-4, 1e-14, 1.0).active_mask.astype(bool))
    return _BAND[0]


def launch_log_spectra(fwhm: np.ndarray) -> np.ndarray:
    """The network's launch input for each point (normalised by its own peak)."""
    fields = launch_fields(fwhm)
    d = np.abs(fields) ** 2
    peak = d.max(axis=1, keepdims=True)
    return ((np.log10(d[:, :BINS] / peak + FLOOR) + 8.0) / 8.0).astype(np.float32)


@dataclass
class Solved:
    """Exit spectra the solver produced: per run its launch spectrum, and per sampled
    length the exit spectrum (network units) with a flag for the lengths reached."""

    launch: np.ndarray     # (runs, BINS)
    exits: np.ndarray      # (runs, k, BINS)
    ok: np.ndarray         # (runs, k) bool
    attempts: np.ndarray   # (runs,) attempts each run took
    seconds: float         # wall time of the batched marches


def solve(points: np.ndarray, lengths: np.ndarray | None = None, chunk: int = 2048,
          max_attempts: int = 40000, progress=None) -> Solved:
    """Run scan points through the batched solver (:class:`solprop_amd.native.Scan`).

    ``points`` (runs, 3 or 4): pressure, energy, fwhm (and the length, the run's end).
    ``lengths`` (runs, k) increasing: where each run's exit spectrum is taken, every
    length a target its march lands on (default: the point's own length). A run that
    exceeds ``max_attempts``, or fails, stops there; its later lengths are flagged.
    """
    import time

    from solprop_amd.native import BackendError, Scan

    points = np.asarray(points, np.float64)
    runs = len(points)
    if lengths is None:
        lengths = points[:, 3:4]
    lengths = np.asarray(lengths, np.float64)
    k = lengths.shape[1]
    launch = np.zeros((runs, BINS), np.float32)
    exits = np.zeros((runs, k, BINS), np.float32)
    ok = np.zeros((runs, k), bool)
    attempts = np.zeros(runs, np.int64)
    seconds = 0.0
    for c0 in range(0, runs, chunk):
        idx = np.arange(c0, min(runs, c0 + chunk))
        problems = [scan_problem(points[i, 0], points[i, 1], points[i, 2], lengths[i, -1])
                    for i in idx]
        peaks = []
        for j, i in enumerate(idx):
            d = np.sum(np.abs(problems[j].field0.reshape(GRID["n_modes"], -1)) ** 2, axis=0)
            peaks.append(float(d.max()))
            launch[i] = log_spectrum(problems[j].field0, peaks[-1])
        scan = Scan(problems)
        scan.start(0.0)
        alive = np.ones(len(idx), bool)
        t0 = time.perf_counter()
        for col in range(k):
            target = lengths[idx, col].copy()
            while True:
                pending = alive & ~reached(scan.z, target)
                if not pending.any():
                    break
                # A stopped run's target is where it stands: it takes no more attempts.
                t = np.where(alive, target, scan.z)
                try:
                    batches = scan.advance(t, work_limit=4000)
                except BackendError:
                    batches = None
                status = np.array(scan.status)
                if batches is not None:
                    for j, b in enumerate(batches):
                        attempts[idx[j]] += len(b.attempts)
                alive &= status == 0
                alive &= attempts[idx] < max_attempts
                if progress:
                    progress(c0, col, int(pending.sum()))
            for j in np.nonzero(alive & reached(scan.z, target))[0]:
                exits[idx[j], col] = log_spectrum(scan.state(int(j)), peaks[j])
                ok[idx[j], col] = True
        seconds += time.perf_counter() - t0
        scan.close()
    return Solved(launch, exits, ok, attempts, seconds)


def training_rows(points3: np.ndarray, lengths: np.ndarray, solved: Solved,
                  ranges: dict = RANGES):
    """The (launch spectrum, parameters, exit spectrum) rows of every length reached, and
    each row's run. The network trains on exit minus launch (see :func:`predict`)."""
    r, j = np.nonzero(solved.ok)
    pts = np.concatenate([points3[r], lengths[r, j][:, None]], axis=1)
    return solved.launch[r], normalise(pts, ranges), solved.exits[r, j], r


def nearest_neighbour(train_params: np.ndarray, train_targets: np.ndarray,
                      query: np.ndarray) -> np.ndarray:
    """The method without the network: each query's exit spectrum taken from the
    training point nearest in the normalised parameters."""
    out = np.empty((len(query), train_targets.shape[1]), np.float32)
    for i0 in range(0, len(query), 1024):
        q = query[i0:i0 + 1024]
        d = ((q[:, None, :] - train_params[None, :, :]) ** 2).sum(axis=2)
        out[i0:i0 + 1024] = train_targets[np.argmin(d, axis=1)]
    return out


@dataclass
class Screening:
    objective: np.ndarray  # (n,) the network's UV share per point
    top: np.ndarray        # indices of the top fraction
    random: np.ndarray     # indices of the random sample
    predicted: dict        # index -> predicted spectrum, for the picked points
    seconds: float


def screen(net: Network, points: np.ndarray, top: float = 0.01, random: float = 0.01,
           seed: int = 0, chunk: int = 8192, ranges: dict = RANGES) -> Screening:
    """The network's exit spectra for every point, its objective, the top fraction and a
    random sample of the rest for the solver to verify."""
    import time

    rng = np.random.default_rng(seed)
    n = len(points)
    objective = np.empty(n)
    keep = {}
    preds = []
    t0 = time.perf_counter()
    for i0 in range(0, n, chunk):
        pts = points[i0:i0 + chunk]
        pred = predict(net, launch_log_spectra(pts[:, 2]), normalise(pts, ranges))
        objective[i0:i0 + len(pts)] = uv_fraction(pred)
        preds.append((i0, pred))
    seconds = time.perf_counter() - t0
    n_top = max(1, int(round(top * n)))
    order = np.argsort(-objective, kind="stable")
    picked_top = order[:n_top]
    rest = order[n_top:]
    picked_random = rng.choice(rest, size=max(1, int(round(random * n))), replace=False)
    wanted = set(picked_top.tolist()) | set(picked_random.tolist())
    for i0, pred in preds:
        for j in range(len(pred)):
            if i0 + j in wanted:
                keep[i0 + j] = pred[j]
    return Screening(objective, picked_top, picked_random, keep, seconds)
