"""Oracle-side stages of one field-route source evaluation, shared by the GPU parity tests
of the capillary and radial routes (test infrastructure: the oracle is the checker)."""

import numpy as np

from oracle import adams, rk


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_source_stages(problem, lab, z):
    """The oracle's modal fields, its three projected time streams and its source."""
    from oracle import passes as P

    src = rk.StageSource(problem)
    crop = problem.crop_index.astype(np.intp)
    modal = P.to_time(lab, crop, problem.n_to, problem.kind)
    field_t = P.synthesise(modal, problem.synthesis, 1.0, 1)
    raman = problem.raman
    s = P.nonlinear_sources(
        field_t, kerr=src.kerr, raman=src.raman_kernel(z),
        raman_thg=False if raman is None else raman.thg,
        raman_epsilon0=P.EPS0 if raman is None else raman.epsilon0, plasma=src.plasma,
        t_window_fine=problem.time_window, kind=problem.kind, density=problem.density(z),
        preionisation=problem.preionisation(z), dt=problem.dt_fine)
    streams = np.zeros((3, problem.n_modes, problem.n_to))
    streams[0] = P.project(s.polarisation, problem.projection, 1.0, 1)[:, 0]
    if s.current is not None:
        streams[1] = P.project(problem.time_window * s.current, problem.projection, 1.0,
                               1)[:, 0]
        streams[2] = P.project(problem.time_window_derivative * s.current,
                               problem.projection, 1.0, 1)[:, 0]
    return modal[:, 0], streams, src.evaluate(lab, z, 1.0)


def log_ulp_sensitivity(problem, lab, z, streams):
    """How far each current stream moves when log|E| is off by one ulp in one direction.

    The rate is exp of a cubic in log|E| with a log-slope of tens to hundreds, and the
    free current late in the window is a small residual of an oscillating integral: a
    systematic last-ulp difference between two libm ``log`` implementations (the
    device's and the host's) moves the current streams by this much.
    """
    from oracle import passes as P

    base = P.RateTable.log_evaluate
    worst = np.zeros(3)
    try:
        for sign in (1.0, -1.0):
            def shifted(self, lf, sign=sign):
                return base(self, lf + sign * np.spacing(lf))

            P.RateTable.log_evaluate = shifted
            _, moved, _ = oracle_source_stages(problem, lab, z)
            for i in range(3):
                if np.any(streams[i]):
                    worst[i] = max(worst[i], rel_l2(moved[i], streams[i]))
    finally:
        P.RateTable.log_evaluate = base
    return worst


def oracle_first_trial(problem, dz):
    policy = adams.adams_policy_for(problem)
    engine = adams.ExponentialEngine(problem, rk.StageSource(problem), policy.order + 2)
    engine.state = problem.field0.reshape(problem.shape).copy()
    engine.start(0.0)
    st = adams.stencil(policy, 0.0, dz, [0.0], [0], 1, last_rejected=False)
    norms = engine.trial(0.0, dz, st)
    return engine.held[2], norms


# The directions in which the fraction's exponential is moved by one ulp: all samples
# up, all down, and two seeded draws of an independent sign per sample (two libm
# implementations differ sample by sample, and a coherent shift of a zero-mean drive
# largely cancels in the running integral).
EXP_SHIFTS = ("up", "down", 0, 1)


def scan_with_shifted_exp(field_t, rate, plasma, preionisation, dt, shift):
    """The oracle's plasma scan with the fraction's exponential moved by one ulp in the
    direction ``shift`` names (an entry of :data:`EXP_SHIFTS`)."""
    rule = plasma.integral
    exposure = rule.floor(rule.integrate(rate, dt))
    decay = np.exp(-exposure)
    if shift == "up":
        sign = 1.0
    elif shift == "down":
        sign = -1.0
    else:
        sign = np.random.default_rng(shift).choice([-1.0, 1.0], size=decay.shape)
    # Below an ulp of 1 the exponential rounds to 1 in both implementations (the device
    # matches the oracle there to rounding): those samples are not moved.
    decay = decay + np.where(exposure > np.finfo(float).eps, sign * np.spacing(decay), 0.0)
    fraction = 1.0 - (1.0 - preionisation) * decay
    integrand = fraction[..., np.newaxis, :] * plasma.charge_mass_ratio * field_t
    return fraction, rule.integrate(integrand, dt)


def scan_sensitivity(field_t, rate, plasma, preionisation, dt, fraction, free):
    """How far the fraction and the free current move under a one-ulp exponential."""
    worst = np.zeros(2)
    for shift in EXP_SHIFTS:
        f2, j2 = scan_with_shifted_exp(field_t, rate, plasma, preionisation, dt, shift)
        worst[0] = max(worst[0], rel_l2(f2, fraction))
        worst[1] = max(worst[1], rel_l2(j2, free))
    return worst


def exp_ulp_sensitivity(problem, lab, z, streams):
    """How far each current stream moves when the fraction's exponential is off by one ulp
    (each direction of :data:`EXP_SHIFTS`).

    ``f = 1 - (1 - f0) e^{-X}`` cancels to a small fraction: a one-ulp difference between
    two libm ``exp`` implementations moves f by about eps e^{-X}, a relative change of
    eps / X where the exposure X is small, which the free current carries and its late
    residual amplifies.
    """
    from oracle import passes as P

    base = P.plasma_scan
    worst = np.zeros(3)
    try:
        for shift in EXP_SHIFTS:
            def shifted(field_t, rate, plasma, preionisation, dt, shift=shift):
                return scan_with_shifted_exp(field_t, rate, plasma, preionisation, dt, shift)

            P.plasma_scan = shifted
            _, moved, _ = oracle_source_stages(problem, lab, z)
            for i in range(3):
                if np.any(streams[i]):
                    worst[i] = max(worst[i], rel_l2(moved[i], streams[i]))
    finally:
        P.plasma_scan = base
    return worst
