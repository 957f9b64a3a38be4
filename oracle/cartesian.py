"""CPU restatement of the free-space Cartesian route (test infrastructure only).

The reference's free-space geometry on a Cartesian grid (``prop_free`` with a
``FreeGrid``; ``src/solprop/translate/free.py:97-249`` freezes the FFT layout and keeps
the constant material's operator and prefactor; the Kerr-beam cases of
``benchmarks/registry.py:348-379`` march it by fixed low-storage RK4IP steps,
``src/solprop/backends/numpy_ref/engines.py:236-520``). The state is the unnormalised
spectrum ``Y = fftn(E)`` of the envelope ``E(y, x, t)``, stored (C, N) with column
``iy nx + ix`` (``free.py:137``: x the faster transverse axis) and the time axis fastest.

- ``source(problem, Y)``: ``p * fftn(g (s |E|^2 E + c (E E) conj(E)))`` with
  ``E = ifftn(Y)`` and ``g`` the Kerr coefficient times the constant density and the
  prefactor's constant weight (the analytic cubic of ``passes/responses.py:65-137``).
- ``flows(problem, z0, z1)``: ``exp((z1 - z0) D)``, the one constant table with
  ``W(z) = z``.
- ``replay(problem, Y0, edges)``: ``oracle/low_storage.py``'s RK4IP step over the schedule.

Pinned by: the exact linear flow; a separable field's source, whose transform is the
product of the axes' transforms (fixes the column layout and the 1/V normalisation);
a uniform beam, which reduces to the envelope route's oracle (``oracle/rk.py``) scaled by
nx ny; the plane wave's closed form (a phase rotation at rate a + b g s |y|^2 / V^2),
reached at fourth order (``tests/test_cartesian.py``).
"""

from __future__ import annotations

import numpy as np

from oracle import low_storage as ls


def grid(problem) -> tuple[int, int, int]:
    """(ny, nx, nt)."""
    nx, ny = problem.transverse
    return int(ny), int(nx), int(problem.n_t)


def gamma(problem) -> float:
    if problem.kerr is None:
        return 0.0
    return (problem.kerr.coefficient * problem.density.value
            * problem.prefactor_paths[0].value)


def response(problem, e: np.ndarray) -> np.ndarray:
    """The analytic Kerr response at every sample of the volume."""
    k = problem.kerr
    if k is None:
        return np.zeros_like(e)
    return gamma(problem) * (k.self_factor * np.abs(e) ** 2 * e
                             + k.conjugate_factor * (e * e) * np.conj(e))


def source(problem, y: np.ndarray) -> np.ndarray:
    """The source at the spectrum ``y`` ((C, N) or the state's shape), (C, N)."""
    ny, nx, nt = grid(problem)
    e = np.fft.ifftn(np.asarray(y, np.complex128).reshape(ny, nx, nt))
    spectrum = np.fft.fftn(response(problem, e)).reshape(ny * nx, nt)
    return problem.prefactor_terms[0] * spectrum


def flows(problem, z0: float, z1: float) -> np.ndarray:
    return np.exp((z1 - z0) * problem.phi_terms[0])


def replay(problem, y0, edges=None):
    """March ``y0`` over the schedule; the state (C, N), the rows and each step's
    stiffness indicator."""
    edges = problem.schedule if edges is None else np.asarray(edges, np.float64)
    norm_mask = problem.norm_mask.astype(bool)
    state = np.asarray(y0, np.complex128).reshape(problem.n_modes, problem.n_t).copy()
    rows, stiffness = [], []
    for i in range(len(edges) - 1):
        z, z1 = float(edges[i]), float(edges[i + 1])
        state, squared, h_rho = ls.step(state, z, z1 - z,
                                        lambda u, _z: source(problem, u),
                                        lambda a, b: flows(problem, a, b), norm_mask)
        if not np.isfinite(squared):
            raise FloatingPointError(f"a replayed step ending at z = {z1:.6g} m diverged")
        rows.append((z1, z1 - z, 0.0, 1.0))
        stiffness.append(h_rho)
    return state, rows, stiffness
