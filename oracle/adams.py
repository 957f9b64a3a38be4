"""Restated exponential Adams march and arithmetic -- TEST INFRASTRUCTURE ONLY.

- The decisions: ``src/solprop/backends/adams_core.py:96-531`` (stencil, factor,
  order weight, order choice, rejection, landing approach) and the march
  ``HostAdamsMarch`` (``src/solprop/backends/numpy_ref/marches/adams.py:79-700``), with
  the policy of ``src/solprop/stepper/multistep.py:434-512`` (constants ``:60-98``).
- The arithmetic: ``phi_functions`` (``src/solprop/backends/exponential.py:131-226``),
  the Lagrange basis (``src/solprop/stepper/multistep.py:515-670``), the corrector and
  defect tables (``src/solprop/backends/numpy_ref/adams_tables.py:109-173``) and the
  exponential trial with the transported compensation of a varying operator
  (``src/solprop/backends/numpy_ref/multistep.py:806-1131``,
  ``adams_kernels.py:404-715``).

Parity pins: the Adams decision corpus (``tests/data/adams_traces.json.gz`` of the
reference, copied under ``tests/golden/``) for the decisions, exactly, except its
dense-observation and carrier-frame scenarios (features this build does not run);
``phi_functions`` against augmented matrix exponentials (``scipy.linalg.expm``) on
both sides of the join; the weights at ``x = 0`` against the Adams-Moulton and
Adams-Bashforth tables; the corrector's order of accuracy on a problem with a known
solution.
"""

from __future__ import annotations

import bisect
import math
from dataclasses import dataclass, field

import numpy as np

from oracle import controller as ctl
from oracle import passes as P

MAX_ADAMS_ORDER = 6
MAX_PHI_ORDER = MAX_ADAMS_ORDER + 2
PEC_REAL_INTERVAL = (1.0, 0.545, 0.3, 0.163, 0.0877, 0.0465)
RESTART_AFTER = 3
ORDER_HYSTERESIS = 1.2
COLLAR_STABILITY = 1.0
REJECTION_GROWTH = 0.9
PHI_JOIN = 3.0
_SERIES_TERMS = 24
_RF = tuple(1.0 / math.factorial(k) for k in range(MAX_PHI_ORDER + 2))


# ------------------------------------------------------------------ the weights


def phi_functions(x, order: int):
    """``phi_0..phi_order`` (``exponential.py:131-226``)."""
    values = np.asarray(x, dtype=np.complex128)
    flat = values.reshape(-1)
    out = np.empty((order + 1, flat.size), dtype=np.complex128)
    far = np.abs(flat) > PHI_JOIN
    if np.any(far):
        xf = flat[far]
        o = np.empty((order + 1, xf.size), dtype=np.complex128)
        o[0] = np.exp(xf)
        for k in range(1, order + 1):
            o[k] = (o[k - 1] - _RF[k - 1]) / xf
        out[:, far] = o
    if not np.all(far):
        xn = flat[~far]
        magnitude = np.abs(xn)
        squarings = np.where(
            magnitude > 0.5, np.ceil(np.log2(np.maximum(magnitude, 0.5) / 0.5)), 0.0
        ).astype(np.int64)
        y = xn / np.exp2(squarings)
        o = np.empty((order + 1, xn.size), dtype=np.complex128)
        term = np.full(xn.size, _RF[order], dtype=np.complex128)
        total = term.copy()
        for m in range(1, _SERIES_TERMS):
            term = term * y / (m + order)
            total = total + term
        o[order] = total
        for k in range(order - 1, 0, -1):
            o[k] = _RF[k] + y * o[k + 1]
        o[0] = np.exp(y)
        for level in range(int(squarings.max(initial=0))):
            active = squarings > level
            old = o[:, active]
            doubled = np.empty_like(old)
            doubled[0] = old[0] * old[0]
            for k in range(1, order + 1):
                t = old[0] * old[k]
                for j in range(1, k + 1):
                    t = t + old[j] * _RF[k - j]
                doubled[k] = t / 2.0**k
            o[:, active] = doubled
        out[:, ~far] = o
    return out.reshape((order + 1, *values.shape))


def _basis_rows(points, upper: float, integrate: bool):
    """The Lagrange basis, coefficients or integrals (``multistep.py:515-564``)."""
    n = points.shape[0]
    rows = np.zeros((n, n))
    weights = np.empty(n)
    coefficients = np.empty(n)
    product = np.empty(n)
    for index in range(n):
        node = points[index]
        length = 1
        coefficients[0] = 1.0
        denominator = 1.0
        for other in range(n):
            if other == index:
                continue
            root = points[other]
            product[0] = 0.0
            for power in range(length):
                product[power + 1] = coefficients[power]
            for power in range(length):
                product[power] -= root * coefficients[power]
            length += 1
            for power in range(length):
                coefficients[power] = product[power]
            denominator *= node - root
        if integrate:
            integral = 0.0
            for power in range(length - 1, -1, -1):
                integral = (integral + coefficients[power] / (power + 1)) * upper
            weights[index] = integral / denominator
        else:
            for power in range(length):
                rows[index, power] = coefficients[power] / denominator
    return rows, weights


def lagrange_coefficients(nodes):
    return _basis_rows(np.asarray(nodes, dtype=np.float64).reshape(-1), 1.0, False)[0]


def integrated_weights(nodes, upper=1.0):
    return _basis_rows(np.asarray(nodes, dtype=np.float64).reshape(-1), float(upper), True)[1]


def exponential_corrector(nodes, order, width):
    corrector = np.zeros((width, width))
    corrector[: order + 1, : order + 1] = lagrange_coefficients([1.0, *nodes[:order]])
    return corrector


def exponential_rows(nodes, orders, width):
    rows = np.zeros((len(orders), width, width))
    for index, order in enumerate(orders):
        stencil = nodes[:order]
        rows[index, : order + 1, : order + 1] = lagrange_coefficients([1.0, *stencil])
        rows[index, 1 : order + 1, :order] -= lagrange_coefficients(stencil)
    return rows


# ------------------------------------------------------------------ the policy


@dataclass(frozen=True)
class MarchPolicy:
    order: int = 5
    variable_order: bool = True
    safety: float = 0.8
    max_growth: float = 2.0
    min_factor: float = 0.2
    rejection_growth: float = REJECTION_GROWTH
    hysteresis: float = ORDER_HYSTERESIS
    restart_after: int = RESTART_AFTER
    collar_stability: float = COLLAR_STABILITY
    real_intervals: tuple = PEC_REAL_INTERVAL
    collar_rate: float = 0.0
    max_calls: int | None = None
    equal_approach: bool = True
    interval_cap: bool = False
    dz_min: float = 1e-12
    dz_max: float = 1.0
    stretch: float = 1.1
    z_end: float = 1.0
    save_planes: tuple = ()
    hard_kinks: tuple = ()
    breakpoints: tuple = (0.0, 1.0)
    landing_ulps: int = ctl.LANDING_ULPS


@dataclass(frozen=True)
class Stencil:
    nodes: list
    rows: np.ndarray
    order: int
    orders: list


def stencil(policy, z, dz, positions, rows, stored_order, *, last_rejected):
    recent = list(positions)[::-1]
    nodes = [(position - z) / dz for position in recent]
    slots = np.array(list(rows)[::-1], dtype=np.int64)
    held = len(recent)
    order = min(stored_order, held)
    rising = not last_rejected and held > order
    orders = [order] + (
        [
            other
            for other in (order - 1, order + 1)
            if 1 <= other <= policy.order and other <= held and (other < order or rising)
        ]
        if policy.variable_order
        else []
    )
    return Stencil(nodes, slots, order, orders)


def factor(policy, norm, order, *, growth):
    if norm <= 0.0:
        return growth
    raw = policy.safety * norm ** (-1.0 / (order + 1))
    return min(growth, max(policy.min_factor, raw))


def stable_factor(policy, order, dz, *, quiet):
    if policy.collar_rate <= 0.0 or quiet:
        return math.inf
    interval = policy.collar_stability * policy.real_intervals[order - 1]
    return interval / (policy.collar_rate * dz)


def order_weight(policy, norm, order, dz, *, quiet):
    rate = math.inf if norm <= 0.0 else norm ** (-1.0 / (order + 1))
    allowed = policy.safety * rate
    return min(allowed, stable_factor(policy, order, dz, quiet=quiet))


def choose_order(policy, current, scores, dz, *, quiet):
    if not policy.variable_order:
        return policy.order, current, scores[current]
    weights = {o: order_weight(policy, n, o, dz, quiet=quiet) for o, n in scores.items()}
    best = max(scores, key=weights.__getitem__)
    if best != current and weights[best] > policy.hysteresis * weights[current]:
        return best, best, scores[best]
    return current, current, scores[current]


def rejection(policy, norm, order, *, rejections, held):
    count = rejections + 1
    restart = count >= policy.restart_after and held > 1
    scale = (
        policy.min_factor
        if not math.isfinite(norm)
        else factor(policy, norm, order, growth=policy.rejection_growth)
    )
    return scale, 0 if restart else count, restart


def collar_reentry(policy, proposal, order, *, was_quiet, quiet):
    returning = was_quiet and not quiet
    if not returning or policy.collar_rate <= 0.0:
        return proposal
    interval = policy.collar_stability * policy.real_intervals[order - 1]
    return min(proposal, interval / policy.collar_rate)


def next_landing(policy, z, targets, target):
    tolerance = policy.landing_ulps * math.ulp(max(abs(z), abs(target)))
    candidates = [target]
    for positions in (targets.planes, targets.kinks):
        index = bisect.bisect_right(positions, z + tolerance)
        if index < len(positions):
            candidates.append(positions[index])
    nearest = min(candidates)
    return nearest if nearest > z + tolerance else None


def approach(policy, z, stop, targets, target):
    landed = stop.z_target is not None
    if not landed and policy.equal_approach:
        ahead = next_landing(policy, z, targets, target)
        if ahead is not None and stop.dz < ahead - z < 2.0 * stop.dz:
            half = 0.5 * (ahead - z)
            return half, z + half, False
    end = stop.z_target if landed else z + stop.dz
    tolerance = policy.landing_ulps * math.ulp(max(abs(end), 1.0))
    interior = end < policy.z_end - tolerance
    on_kink = interior and any(abs(kink - end) <= tolerance for kink in targets.kinks)
    return stop.dz, end, stop.target == "hard_kink" or on_kink


def replay_kink(policy, end, slack):
    return any(abs(kink - end) <= slack for kink in policy.hard_kinks)


def step_targets(policy, target):
    return ctl.StepTargets(
        tuple(policy.save_planes), tuple(policy.hard_kinks), tuple(policy.breakpoints),
        policy.dz_max, target, policy.stretch,
    )


# -------------------------------------------------------------------- the march


class ResponseBudgetError(ctl.ConvergenceError):
    """The source-call budget is spent."""


@dataclass
class Ledger:
    startup: int = 0
    predictor: int = 0
    corrected: int = 0
    rejected: int = 0
    accepted_steps: int = 0
    rejected_steps: int = 0
    restarts: int = 0
    failure_restarts: int = 0
    dense_observations: int = 0
    orders: list = field(default_factory=list)

    @property
    def total(self):
        return self.startup + self.predictor + self.corrected


@dataclass
class Entry:
    z: float
    identity: str
    row: int


class AdamsMarch:
    """``HostAdamsMarch`` over an engine with ``start``, ``trial`` and ``accept``.

    The host supplies ``cap(proposal, z)``, ``record(end, dz, norm, accepted)``,
    ``stepped()``, ``natural_stop(target)`` and ``collar_quiet``.
    """

    def __init__(self, engine, policy: MarchPolicy, host, *, proposal, z0=0.0,
                 schedule=None):
        self.engine = engine
        self.policy = policy
        self.host = host
        self.schedule = schedule
        self.z = z0
        self.proposal = proposal
        self.cursor = 0
        self.history: list[Entry] = []
        self.rejections = 0
        self.last_rejected = False
        self.order = 1 if policy.variable_order else policy.order
        self.n_slots = policy.order + 2
        self.ledger = Ledger()
        self.collar_was_quiet = False
        self.floor = None

    def march_to(self, target):
        if self.schedule is not None:
            return self._replay_to(target)
        targets = step_targets(self.policy, target)
        while not ctl.reached(self.z, target):
            self._afford()
            if not self.history:
                self._start()
            proposal = self._collar_reentry(self.host.cap(self.proposal, self.z))
            stop = ctl.landing(self.z, proposal, targets)
            dz, end, kink = approach(self.policy, self.z, stop, targets, target)
            accepted, norm = self._attempt(dz, landing=(end, kink))
            self.host.record(end, dz, norm, accepted)
            self._end(accepted)
            if accepted:
                self.host.stepped()
                if self.host.natural_stop(target):
                    return

    def _replay_to(self, target):
        steps = self.schedule
        edges = ctl._schedule_edges(steps)
        n = len(steps)
        slack = max(ctl.LANDING_ULPS, n) * math.ulp(max(edges[-1], abs(target)))
        while self.cursor < n and self.z < target - slack:
            dz = float(steps[self.cursor])
            end = float(edges[self.cursor + 1])
            self._afford()
            if not self.history:
                self._start()
            kink = replay_kink(self.policy, end, slack)
            _, norm = self._attempt(dz, fixed=True, landing=(end, kink))
            if not math.isfinite(norm):
                raise ctl.ConvergenceError(
                    f"scripted: a replayed exponential Adams step of {dz:.4g} m ending "
                    f"at z = {end:.6g} m"
                )
            self.cursor += 1
            self.host.record(end, dz, norm, True)
            self._raise_floor()
            self.host.stepped()
            if self.host.natural_stop(target):
                return
        if self.z < target - slack:
            raise ctl.BackendError(
                f"the replayed schedule ends at {edges[-1]:.6g} m and cannot reach "
                f"z = {target:.6g} m"
            )

    def _end(self, accepted):
        if accepted:
            self._raise_floor()
            return
        error, self.floor = self.floor, None
        if error is not None:
            raise error

    def _afford(self):
        self._raise_floor()
        budget = self.policy.max_calls
        cost = (0 if self.history else 1) + 1
        spent = self.ledger.total
        if budget is not None and spent + cost > budget:
            raise ResponseBudgetError(
                f"the source-call budget of {budget} is spent ({spent} charged) at "
                f"z = {self.z:.6g} m: the next trial needs {cost}; the run stops at its "
                "last accepted step, unfinished"
            )

    def _start(self):
        self.ledger.startup += 1
        self.engine.start(self.z)
        self.history = [Entry(self.z, "accepted", 0)]
        self.order = 1 if self.policy.variable_order else self.policy.order

    def _attempt(self, dz, *, fixed=False, landing=None):
        st = stencil(self.policy, self.z, dz, [e.z for e in self.history],
                     [e.row for e in self.history], self.order,
                     last_rejected=self.last_rejected)
        self.ledger.predictor += 1
        norms = self.engine.trial(self.z, dz, st)
        return self._decide(dz, st, norms, fixed=fixed, landing=landing)

    def _decide(self, dz, st, norms, *, fixed, landing=None):
        policy = self.policy
        order = st.order
        norm = float(norms[0])
        if not norm <= 1.0 and not (fixed and math.isfinite(norm)):
            self.ledger.rejected += 1
            self._reject(dz, norm, order)
            if fixed:
                self._raise_floor()
            return False, norm
        z = self.z
        keep = self.history[-policy.order:] if policy.order else []
        used = {e.row for e in keep}
        free = next(slot for slot in range(self.n_slots) if slot not in used)
        self.engine.accept(dz, free, st, keep)
        self.history = [*keep, Entry(z + dz, "predicted", free)]
        self.ledger.accepted_steps += 1
        self.rejections = 0
        self.ledger.orders.append(order)
        stored, control, cnorm = choose_order(
            policy, order, dict(zip(st.orders, [float(v) for v in norms], strict=True)),
            dz, quiet=self.host.collar_quiet,
        )
        self.order = stored
        self.last_rejected = False
        # The decision corpus was recorded from the march whose floor is raised where
        # the proposal falls below it, before the step's end is landed on and before
        # the attempt is recorded.
        self.proposal = ctl.limit_step(
            dz, factor(policy, cnorm, control, growth=policy.max_growth),
            policy.dz_min, policy.dz_max,
        )
        end, kink = (z + dz, False) if landing is None else landing
        self.z = end
        if kink:
            self.history = []
            self.ledger.restarts += 1
        return True, norm

    def _raise_floor(self):
        error, self.floor = self.floor, None
        if error is not None:
            raise error

    def _collar_reentry(self, proposal):
        quiet = self.host.collar_quiet
        capped = collar_reentry(self.policy, proposal, self.order,
                                was_quiet=self.collar_was_quiet, quiet=quiet)
        self.collar_was_quiet = quiet
        return capped

    def _reject(self, dz, norm, order):
        self.ledger.rejected_steps += 1
        self.last_rejected = True
        scale, count, restart = rejection(self.policy, norm, order,
                                          rejections=self.rejections,
                                          held=len(self.history))
        self.rejections = count
        if restart:
            self.history = []
            self.ledger.failure_restarts += 1
        self.proposal = ctl.limit_step(dz, scale, self.policy.dz_min, self.policy.dz_max)


# ------------------------------------------------------- scripted decision traces

LEDGER = ("startup", "predictor", "rejected", "accepted_steps", "rejected_steps",
          "restarts", "failure_restarts", "dense_observations")


def policy_of(scenario: dict) -> MarchPolicy:
    m = dict(scenario["method"])
    breakpoints = scenario["breakpoints"] or [scenario["z0"], scenario["z_end"]]
    interval_cap = bool(m.get("interval_cap", False))
    return MarchPolicy(
        order=m.get("order", 5),
        variable_order=m.get("variable_order", True),
        safety=m.get("safety", 0.8),
        max_growth=m.get("max_growth", 2.0),
        min_factor=m.get("min_factor", 0.2),
        collar_rate=float(scenario["collar_rate"]),
        max_calls=m.get("max_calls"),
        equal_approach=m.get("equal_approach", True),
        interval_cap=interval_cap,
        dz_min=float(scenario["dz_min"]),
        dz_max=float(scenario["dz_max"]),
        stretch=float(scenario["stretch"]),
        z_end=float(scenario["z_end"]),
        save_planes=() if scenario["interpolate"] else tuple(scenario["save_planes"]),
        hard_kinks=tuple(scenario["kinks"]),
        breakpoints=tuple(breakpoints) if interval_cap else (breakpoints[0], breakpoints[-1]),
    )


def run_scenario(scenario: dict) -> dict:
    """Run one scripted Adams scenario (``tests/backends/adams_traces.py:596-633``)."""
    policy = policy_of(scenario)
    quiet_map = {int(k): v for k, v in scenario["quiet"].items()}
    events = []
    h = {"quiet": quiet_map.get(0, False), "accepted": 0, "attempts": 0, "pending": {}}

    class Host:
        @property
        def collar_quiet(self):
            return h["quiet"]

        def cap(self, proposal, z):
            if scenario["quiet_interval"] is None:
                return proposal
            if h["quiet"]:
                proposal = min(proposal, scenario["quiet_interval"])
            return min(proposal, math.inf)

        def record(self, end, dz, norm, accepted):
            row = dict(h["pending"])
            row.update(end=end, step=dz, norm=norm if math.isfinite(norm) else math.inf,
                       accepted=accepted, **state())
            events.append(row)
            h["pending"] = {}

        def stepped(self):
            h["accepted"] += 1
            h["quiet"] = quiet_map.get(h["accepted"], h["quiet"])

        def natural_stop(self, target):
            return False

    class Engine:
        def start(self, z):
            events.append({"start": z, "stored": True})

        def trial(self, z, dz, st):
            index = h["attempts"]
            if index >= len(scenario["norms"]):
                raise RuntimeError("ran out of norms")
            h["pending"] = {"z": z, "dz": dz, "nodes": [float(n) for n in st.nodes],
                            "slots": [int(s) for s in st.rows], "order": st.order,
                            "orders": list(st.orders)}
            h["attempts"] += 1
            return np.array(scenario["norms"][index][: len(st.orders)], dtype=np.float64)

        def accept(self, dz, free, st, kept):
            h["pending"]["kept"] = [int(e.row) for e in kept]
            h["pending"]["free"] = int(free)

    march = AdamsMarch(Engine(), policy, Host(), proposal=scenario["proposal"],
                       z0=scenario["z0"], schedule=scenario["schedule"])

    def state():
        return {
            "position": march.z,
            "proposal": march.proposal,
            "stored_order": march.order,
            "rejections": march.rejections,
            "last_rejected": march.last_rejected,
            "collar_was_quiet": march.collar_was_quiet,
            "quiet": h["quiet"],
            "history": [[e.z, e.identity, e.row] for e in march.history],
            "ledger": [getattr(march.ledger, name) for name in LEDGER],
        }

    if scenario["targets"] is not None:
        targets = list(scenario["targets"])
    elif scenario["interpolate"] or not scenario["save_planes"]:
        targets = [scenario["z_end"]]
    else:
        targets = sorted({*scenario["save_planes"], scenario["z_end"]})
    end = {"stop": "done"}
    try:
        for target in targets:
            march.march_to(target)
            events.append({"target": target, "position": march.z})
    except ResponseBudgetError as error:
        end = {"stop": "ResponseBudgetError", "message": str(error)}
    except (ctl.BackendError, ctl.ConvergenceError) as error:
        end = {"stop": type(error).__name__, "message": str(error)}
    end["in_flight"] = dict(h["pending"])
    end.update(state())
    end["orders"] = list(march.ledger.orders)
    return {"events": events, "end": end}


# ----------------------------------------------------- the exponential engine


def moments(exponent, count):
    """``M_k = k! phi_{k+1}(x)``, ``k < count`` (``adams_kernels.py:421-447``)."""
    phi = phi_functions(exponent, count)
    factors = np.array([math.factorial(k) for k in range(count)], dtype=np.float64)
    return phi[1:] * factors.reshape((count, *([1] * exponent.ndim)))


def apply(M, coefficients, values, dz):
    """``dz Σ_k M_k (Σ_j c_jk X_j)`` (``adams_kernels.py:521-551``)."""
    total = np.zeros_like(values[0])
    for k in range(coefficients.shape[1]):
        combined = np.zeros_like(values[0])
        for weight, value in zip(coefficients[:, k], values, strict=True):
            if weight != 0.0:
                combined = P.accumulate(combined, value, float(weight))
        total += M[k] * combined
    return dz * total


class ExponentialEngine:
    """``ExponentialHistory`` (``numpy_ref/multistep.py:806-1131``) over a problem.

    Untransported source rows; a varying operator (more than one Φ term with a
    non-constant rate) adds the transported compensation ``(D(z_j) - x/h) u_j``.
    """

    def __init__(self, problem, source, n_slots):
        self.on_finish = None
        from oracle import rk

        self.p = problem
        self.source = source
        self.rk = rk
        self.n_slots = n_slots
        self.state = None
        self.store = None
        self.states = None
        self.rates = None
        self.compensated = len(problem.phi_paths) > 1
        self.held = None

    @staticmethod
    def residual_rate(rate, reference):
        """The operator's rate less the step's constant reference, ``D(z_j) - x/h``: the
        linear part left to the quadrature (``numpy_ref/multistep.py:170-204,1072-1085``)."""
        return rate - reference

    def lab_source(self, lab, z):
        """The source at a lab-frame field, unit frame factor (``workspace.py:819``)."""
        return self.source.evaluate(lab, z, 1.0)

    def start(self, z):
        value = self.lab_source(self.state, z)
        if self.store is None:
            self.store = np.zeros((self.n_slots, *value.shape), dtype=np.complex128)
            if self.compensated:
                self.states = np.zeros_like(self.store)
                self.rates = np.zeros_like(self.store)
        self.store[0] = value
        if self.compensated:
            self.states[0] = self.state
            self.rates[0] = self.p.phi_rate(z)

    def trial(self, z, dz, st: Stencil):
        p = self.p
        state = self.state
        nodes, slots, order, orders = st.nodes, st.rows, st.order, st.orders
        width = 1 + max(orders)
        exponent = self.rk.delta_phi(p, z, z + dz)
        M = moments(exponent, width)
        flow = np.exp(exponent)
        rows_read = slots[: width - 1]
        values = [self.store[s] for s in rows_read[:order]]
        lab = flow * state + apply(M, lagrange_coefficients(nodes[:order]), values, dz)
        carried = None
        if self.compensated:
            reference = exponent / dz
            used = list(nodes[: len(rows_read)])
            carried = np.stack([
                np.exp((1.0 - node) * exponent)
                * (self.residual_rate(self.rates[slot], reference) * self.states[slot])
                for slot, node in zip(rows_read, used, strict=True)
            ])

            def csum(ns, first, count):
                w = integrated_weights(np.asarray(ns))
                total = np.tensordot(w[len(ns) - count:], carried[:count], axes=1)
                if first is not None:
                    total = total + w[0] * first
                return dz * total

            lab = lab + csum(used[:order], None, order)
        source = self.lab_source(lab, z + dz)
        X = [source, *[self.store[s] for s in rows_read]]
        corrector = exponential_corrector(nodes, order, width)
        drows = exponential_rows(nodes, orders, width)
        extra = None
        endpoint_rate = None
        if self.compensated:
            endpoint_rate = p.phi_rate(z + dz)
            endpoint = self.residual_rate(endpoint_rate, exponent / dz) * lab

            def corr(k):
                return csum([1.0, *used[:k]], endpoint, k)

            def pred(k):
                return csum(used[:k], None, k)

            extra = (corr(order), [corr(k) - pred(k) for k in orders])
        advanced = flow * state + apply(M, corrector, X, dz)
        if extra is not None:
            advanced = advanced + extra[0]
        mask = p.norm_mask
        magnitude = float(np.linalg.norm(state[..., mask]))
        divisor = p.rtol * max(magnitude, magnitude) + p.atol
        norms = np.empty(len(orders))
        for index, row in enumerate(drows):
            defect = apply(M, row, X, dz)
            if extra is not None:
                defect = defect + extra[1][index]
            norms[index] = float(np.linalg.norm(defect[..., mask] / divisor))
        self.held = (source, lab, advanced, endpoint_rate)
        return norms

    def accept(self, dz, free, st, kept):
        source, lab, advanced, endpoint_rate = self.held
        self.held = None
        self.store[free] = source
        if self.compensated:
            self.states[free] = lab
            self.rates[free] = endpoint_rate
        self.state = P.apply_hard_mask(advanced, self.p.active_mask)
        if self.on_finish is not None:
            self.on_finish(self.state, dz)
        if self.p.absorber_scheme == "split_step":
            # The workspace's finish: mask, the split collar, mask again
            # (``numpy_ref/workspace.py:1020-1108``).
            self.state = P.apply_hard_mask(
                P.apply_temporal_rate(self.state, self.p.absorber_rate, dz, self.p.kind),
                self.p.active_mask)


class _PlainHost:
    collar_quiet = False

    def __init__(self):
        self.rows = []

    def cap(self, proposal, z):
        return proposal

    def record(self, end, dz, norm, accepted):
        self.rows.append((end, dz, norm, accepted))

    def stepped(self):
        pass

    def natural_stop(self, target):
        return False


def adams_policy_for(problem, **overrides) -> MarchPolicy:
    values = dict(
        order=problem.adams_order,
        variable_order=problem.adams_variable_order,
        dz_min=1e-9 * problem.z_end,
        dz_max=problem.z_end,
        z_end=problem.z_end,
        breakpoints=(0.0, problem.z_end),
    )
    values.update(overrides)
    return MarchPolicy(**values)


def march(problem, state, *, target=None, proposal=None, max_attempts=None, policy=None,
          on_attempt=None, on_finish=None, engine=ExponentialEngine, marches=None):
    """Adaptive exponential Adams march of the problem (``HostAdamsMarch.march_to``).

    Returns ``(state, rows)`` with rows ``(end, dz, norm, accepted)`` per attempt.
    ``on_finish(masked, dz)`` sees each accepted step's masked state before its split
    collar (what a bound monitor's split sums read, ``numpy_ref/workspace.py:1021-1067``).
    ``marches``, when given, receives the :class:`AdamsMarch` (its ledger holds the
    orders of the accepted steps).
    """
    from oracle import rk

    policy = policy or adams_policy_for(problem)
    engine = engine(problem, rk.StageSource(problem), policy.order + 2)
    engine.state = state
    engine.on_finish = on_finish
    host = _PlainHost()
    if on_attempt is not None or max_attempts is not None:
        inner = host.record

        def record(end, dz, norm, accepted):
            inner(end, dz, norm, accepted)
            if on_attempt is not None:
                on_attempt(end, dz, norm, accepted)
            if max_attempts is not None and len(host.rows) >= max_attempts:
                raise StopIteration

        host.record = record
    am = AdamsMarch(engine, policy, host, proposal=proposal or 1e-4 * problem.z_end)
    if marches is not None:
        marches.append(am)
    try:
        am.march_to(problem.z_end if target is None else target)
    except StopIteration:
        pass
    return engine.state, host.rows


__all__ = [
    "AdamsMarch", "ExponentialEngine", "MarchPolicy", "exponential_corrector",
    "exponential_rows", "integrated_weights", "lagrange_coefficients", "march",
    "phi_functions", "run_scenario",
]
