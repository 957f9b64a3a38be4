"""Restated source evaluation, Runge-Kutta trial and march -- TEST INFRASTRUCTURE.

- :class:`StageSource` restates ``src/solprop/backends/numpy_ref/source.py:204-299``
  (``evaluate``) and ``:599-700`` (``_over_nodes``), with the held plasma streams of
  ``:762-783`` and the Raman kernel memoised on the decay rates (``:805-856``).
- :func:`rk_trial` restates ``NumpyRungeKutta.trial``
  (``src/solprop/backends/numpy_ref/engines.py:99-178``): every stage evaluated (no
  FSAL), the embedded difference carried to the step's end by the whole-step
  propagator.
- The march (:func:`march`) restates ``RungeKuttaMarch.march_to``
  (``src/solprop/backends/numpy_ref/marches/runge_kutta.py:133-210``) over the
  controller of :mod:`oracle.controller`.

Parity pins: a linear-only problem must equal ``exp(Φ(z)) y0`` (closed form); the
N=1 soliton is stationary; the march's decisions are held to the reference's
``tests/data/rk_traces.json.gz`` through :mod:`oracle.controller`.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import controller as ctl
from oracle import passes as P


def _kerr(problem):
    k = problem.kerr
    if k is None:
        return None
    return P.Kerr(
        k.coefficient,
        k.cubic_form,
        k.self_factor,
        k.conjugate_factor,
        k.harmonic_factor,
        k.analytic_signal,
        k.take_real,
    )


def _plasma(problem):
    g = problem.plasma
    if g is None:
        return None
    table = P.RateTable(
        np.asarray(g.rate_coefficients),
        g.log_field_min,
        g.d_log_field,
        g.field_min,
        g.field_max,
        g.rate_max,
    )
    return P.Plasma(
        table,
        g.charge_mass_ratio,
        g.ionisation_potential,
        P.RunningIntegral(tuple(float(t) for t in np.asarray(g.taps))),
    )


class StageSource:
    """One stage's source evaluation over the problem's arrays."""

    def __init__(self, problem):
        self.p = problem
        self.kerr = _kerr(problem)
        self.plasma = _plasma(problem)
        self._raman_rates = None
        self._raman_kernel = None

    def raman_kernel(self, z):
        p = self.p
        if p.raman is None:
            return None
        rates = tuple(p.raman_rates(z).tolist())
        if self._raman_kernel is None or self._raman_rates != rates:
            self._raman_kernel = P.raman_spectrum(
                p.raman.undamped, rates, p.raman.fraction, p.dt_fine
            )
            self._raman_rates = rates
        return self._raman_kernel

    def prefactor(self, z):
        p = self.p
        w = p.prefactor_weights(z)
        total = np.zeros(p.prefactor_terms.shape[1:], dtype=np.complex128)
        for weight, term in zip(w, p.prefactor_terms, strict=True):
            total = total + weight * term
        return total.reshape(p.shape)

    def evaluate(self, lab, z, inverse_propagator):
        """The stage derivative in the anchor's frame, from a lab-frame field at z."""
        p = self.p
        crop = p.crop_index.astype(np.intp)
        spectrum, modal_t = self.projected_spectrum(lab, z)
        absorber = None
        if p.absorber_scheme == "in_stage" or p.spatial_absorber is not None:
            # ADR-025's in-stage absorber, a term of its own (``source.py:391-401``),
            # with the radial collar's matrix product added to it (``source.py:133``).
            absorber = P.absorber_term(
                modal_t, p.absorber_rate if p.absorber_scheme == "in_stage" else None,
                crop, p.active_mask, kind=p.kind, radial=p.spatial_absorber)
        return P.stage_output(
            spectrum, self.prefactor(z), p.omega_window, inverse_propagator, absorber
        )

    def projected_spectrum(self, lab, z, projection=None):
        """The response's spectrum on the coarse bins before the prefactor: the projected
        polarisation's transform plus the held plasma currents' (with `projection` in
        place of the problem's when given); and the modal time fields."""
        p = self.p
        kind = p.kind
        projection = p.projection if projection is None else projection
        crop = p.crop_index.astype(np.intp)
        modal_t = P.to_time(lab, crop, p.n_to, kind)
        raman = self.raman_kernel(z)
        # Whole-quadrature chunk: round-off aside, any chunking gives the same answer.
        field_t = P.synthesise(modal_t, p.synthesis, 1.0, 1)
        density = p.density(z)
        sources = P.nonlinear_sources(
            field_t,
            kerr=self.kerr,
            raman=raman,
            raman_thg=False if p.raman is None else p.raman.thg,
            plasma=self.plasma,
            raman_epsilon0=P.EPS0 if p.raman is None else p.raman.epsilon0,
            t_window_fine=p.time_window,
            kind=kind,
            density=density,
            preionisation=p.preionisation(z),
            dt=p.dt_fine,
        )
        projected = P.project(sources.polarisation, projection, 1.0, 1)
        spectrum = P.to_frequency(projected, crop, kind)
        if sources.current is not None:
            gated = P.project(p.time_window * sources.current, projection, 1.0, 1)
            edge = P.project(
                p.time_window_derivative * sources.current, projection, 1.0, 1
            )
            spectrum = spectrum + P.held_plasma_spectrum(
                gated, edge, p.omega, crop, p.active_mask
            )
        return spectrum, modal_t


def delta_phi(problem, z0, z1):
    """``Φ(z1) - Φ(z0) = Σ_t (W_t(z1) - W_t(z0)) Φ_t`` in the state's layout."""
    w = problem.phi_weights(z0, z1)
    total = np.zeros(problem.phi_terms.shape[1:], dtype=np.complex128)
    for weight, term in zip(w, problem.phi_terms, strict=True):
        total = total + weight * term
    return total.reshape(problem.shape)


def noise_at(problem, z):
    """``N0 exp(i Im Φ(z))`` for the modified model (``numpy_ref/workspace.py:690-711``)."""
    if problem.noise_model != "modified" or problem.noise0 is None:
        return None
    w = problem.phi_at(z)
    phase = np.zeros(problem.phi_terms.shape[1:])
    total = np.zeros(problem.phi_terms.shape[1:], dtype=np.complex128)
    for weight, term in zip(w, problem.phi_terms, strict=True):
        total = total + weight * term
    phase = np.imag(total)
    return (problem.noise0 * np.exp(1j * phase)).reshape(problem.shape)


def rk_trial(problem, source: StageSource, state, z, dz):
    """One interaction-picture embedded RK trial (``numpy_ref/engines.py:99-178``).

    Returns ``(advanced, norm, stages)``.
    """
    tab = problem.tableau
    stages = []
    solution = np.zeros_like(state)
    error = np.zeros_like(state)
    for index, node in enumerate(tab.nodes):
        fraction = float(node)
        combination = np.zeros_like(state)
        for previous in range(index):
            combination = P.accumulate(
                combination, stages[previous], float(tab.matrix[index, previous])
            )
        zi = z + fraction * dz
        exponent = delta_phi(problem, z, zi)
        lab = P.stage_input(state, combination, dz, np.exp(exponent), noise_at(problem, zi))
        back = np.exp(-exponent)
        stage = source.evaluate(lab, zi, back)
        stages.append(stage)
        solution = P.accumulate(solution, stage, float(tab.solution_weights[index]))
        error = P.accumulate(error, stage, float(tab.error_weights[index]))
    propagator = np.exp(delta_phi(problem, z, z + dz))
    advanced = P.accept_step(state, propagator, dz, solution)
    norm = P.error_norm(
        propagator * (dz * error),
        state,
        advanced,
        problem.norm_mask,
        problem.rtol,
        problem.atol,
    )
    return advanced, norm, stages


def march(problem, state, *, z0=0.0, proposal=None, targets=None, policy=None,
          max_attempts=10**7, on_attempt=None, on_finish=None):
    """Adaptive march to each target (``marches/runge_kutta.py:133-210``).

    Returns ``(state, records)`` where records hold ``(z, dz, norm, accepted)`` per
    attempt. The finish of an accepted step is the hard band mask and, under the
    split-step absorber, the collar over the step and the mask again
    (``src/solprop/backends/numpy_ref/workspace.py:1021-1106``); ``on_finish(masked, dz)``
    sees the masked state before the collar.
    """
    source = StageSource(problem)
    if policy is None:
        policy = ctl.RKPolicy(
            controller=ctl.PIController(rtol=problem.rtol, atol=problem.atol),
            dz_min=1e-9 * problem.z_end,
            dz_max=problem.z_end,
            auto_init=True,
            stretch=1.1,
            z_end=problem.z_end,
            save_planes=(),
            hard_kinks=(),
            breakpoints=(0.0, problem.z_end),
        )
    st = ctl.MarchState(z=z0, proposal=proposal or 1e-4 * problem.z_end,
                        lag=policy.controller.lag)
    records = []
    attempts = 0
    for target in targets or [problem.z_end]:
        tg = ctl.step_targets(policy, target)
        while not ctl.reached(st.z, target):
            stop = ctl.landing(st.z, st.proposal, tg)
            advanced, norm, _ = rk_trial(problem, source, state, st.z, stop.dz)
            equivalent, accepted = ctl.decide(policy, norm, stop.dz, st.proposal)
            end = stop.z_target if stop.z_target is not None else st.z + stop.dz
            records.append((st.z, stop.dz, norm, accepted))
            if on_attempt is not None:
                on_attempt(st.z, stop.dz, norm, accepted)
            attempts += 1
            if accepted:
                state = P.apply_hard_mask(advanced, problem.active_mask)
                if on_finish is not None:
                    on_finish(state, stop.dz)
                if problem.absorber_scheme == "split_step":
                    state = P.apply_hard_mask(
                        P.apply_temporal_rate(state, problem.absorber_rate, stop.dz,
                                              problem.kind),
                        problem.active_mask)
                st.z = end
                st.norms.append(equivalent)
                if policy.auto_init and len(st.norms) == 1:
                    st.proposal, st.norms[-1] = ctl.start_up(policy, st.proposal, equivalent)
                st.proposal = ctl.propose(policy, st.proposal, list(st.norms))
            else:
                st.proposal = ctl.reject(policy, st.proposal, equivalent)
            if attempts >= max_attempts:
                return state, records
    return state, records


def linear_solution(problem, state, z):
    """The exact linear flow ``exp(Φ(z) - Φ(0)) y0``, the linear-only closed form."""
    return np.exp(delta_phi(problem, 0.0, z)) * state


__all__ = ["StageSource", "delta_phi", "linear_solution", "march", "noise_at", "rk_trial"]
_ = math
