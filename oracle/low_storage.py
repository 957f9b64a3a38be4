"""CPU restatement of the low-storage RK4IP replay (test infrastructure only).

The fourth-order Runge-Kutta method in the interaction picture (J. Hult, J. Lightwave
Technol. 25, 3770 (2007)) in the three-register form the reference's engines take
(``src/solprop/backends/numpy_ref/engines.py:236-520`` ``NumpyLowStorage``; natively
``native/src/general/trial_low_storage.inc``, ``native/src/march/low_storage.cpp``):
the state is register A, updated in place; B accumulates, C is the stage field. Per step
from z over h, with z_m = z + h/2, P = exp(Phi(z_m) - Phi(z)), Q = exp(Phi(z + h) -
Phi(z_m)) and S(u, z) the lab-frame source:

    pass 1: S1 = S(A, z);  A <- P A;  g <- P S1;  B <- A + h/6 g;  C <- A + h/2 g
    pass 2: S2 = S(C, z_m);  B <- B + h/3 S2;  C <- A + h/2 S2
    pass 3: S3 = S(C, z_m);  B <- B + h/3 S3;  C <- Q (A + h S3)
    pass 4: S4 = S(C, z + h);  A <- Q B + h/6 S4

then the finish (the hard mask). The flows here are formed for each step from the
operator's increments (the reference's "recompute" treatment). Beside the step, the
stiffness indicator h rho = 2 ||S3 - S2|| / ||S2 - g|| over the norm mask, which is
exactly |h lambda| for u' = lambda u.

The replay takes a fixed schedule (its edges), every step accepted; each row is (end, step,
0.0 -- "not estimated" --, 1). A step whose state is not finite stops the replay.
"""

from __future__ import annotations

import numpy as np

from oracle import rk


def uniform_edges(z0: float, z1: float, steps: int) -> np.ndarray:
    """A uniform schedule's edges."""
    return np.linspace(z0, z1, steps + 1)


def step(state, z, dz, source, flows, norm_mask):
    """One RK4IP step: the new state, its squared norm over the mask and h rho.

    ``source(u, z)`` is the lab-frame source; ``flows(z0, z1)`` returns exp(Phi(z1) -
    Phi(z0)) in the state's layout.
    """
    mid, end = z + 0.5 * dz, z + dz
    p = flows(z, mid)
    q = flows(mid, end)
    s1 = source(state, z)
    a = p * state
    g = p * s1
    b = a + (dz / 6.0) * g
    c = a + (0.5 * dz) * g
    s2 = source(c, mid)
    lower = np.linalg.norm((s2 - g)[..., norm_mask])
    b = b + (dz / 3.0) * s2
    c = a + (0.5 * dz) * s2
    s3 = source(c, mid)
    upper = np.linalg.norm((s3 - s2)[..., norm_mask])
    b = b + (dz / 3.0) * s3
    c = q * (a + dz * s3)
    s4 = source(c, end)
    a = q * b + (dz / 6.0) * s4
    stiffness = 2.0 * upper / lower if lower > 0.0 else float("nan")
    accepted = a[..., norm_mask]
    return a, float(np.vdot(accepted, accepted).real), stiffness


def replay(problem, state, edges, *, source=None, target=None):
    """March `state` over the schedule `edges` (to `target`, default its end); return the
    state, the rows and each step's stiffness indicator."""
    src = source or rk.StageSource(problem)
    shape = problem.shape
    mask = problem.active_mask.astype(bool)
    norm_mask = problem.norm_mask.astype(bool)

    def lab_source(u, z):
        return src.evaluate(u.reshape(shape), z, 1.0).reshape(shape)

    def flows(z0, z1):
        return np.exp(rk.delta_phi(problem, z0, z1)).reshape(shape)

    state = np.asarray(state, np.complex128).reshape(shape).copy()
    edges = np.asarray(edges, np.float64)
    end = edges[-1] if target is None else target
    rows, stiffness = [], []
    for i in range(len(edges) - 1):
        z, z1 = edges[i], edges[i + 1]
        if z >= end:
            break
        dz = z1 - z
        state, squared, h_rho = step(state, z, dz, lab_source, flows, norm_mask)
        if not np.isfinite(squared):
            raise rk.ctl.ConvergenceError(f"a replayed step of {dz:.4g} m ending at "
                                          f"z = {z1:.6g} m diverged")
        state[..., ~mask] = 0.0
        rows.append((z1, dz, 0.0, 1.0))
        stiffness.append(h_rho)
    return state, rows, stiffness
