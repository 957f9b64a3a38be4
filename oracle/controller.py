"""Restated step-size controller, landings and Runge-Kutta trace march -- TEST ONLY.

- :class:`PIController` restates ``src/solprop/stepper/controller.py:262-676``
  (``propose`` :473, ``start_factor`` :527, ``accepts`` :563, ``rejection_factor`` :590,
  ``equivalent_norm`` :623) and :func:`limit_step` ``:162-201``.
- :func:`decide`, :func:`start_up`, :func:`propose`, :func:`reject` restate
  ``src/solprop/backends/rk_core.py:37-166``.
- :func:`landing` and :func:`reached` restate ``src/solprop/stepper/schedule.py:959-1090``.
- :func:`run_scenario` restates the scripted decision-trace harness of
  ``tests/backends/rk_traces.py:145-453`` around ``RungeKuttaMarch.march_to`` /
  ``replay_to`` (``src/solprop/backends/numpy_ref/marches/runge_kutta.py:133-260``)
  with the session's caps (``src/solprop/backends/numpy_ref/session.py:289-361``).

Parity pin: ``tests/data/rk_traces.json.gz`` of the reference (29 scenarios), compared
exactly (glibc ``pow``), copied under ``tests/golden/``.
"""

from __future__ import annotations

import bisect
import math
from collections import deque
from dataclasses import dataclass, field

START_GROWTH = 100.0
LANDING_ULPS = 8


class ConvergenceError(Exception):
    """The controller asked for a step below the floor."""


class BackendError(Exception):
    """A replay could not reach its target."""


@dataclass(frozen=True)
class PIController:
    rtol: float = 1e-6
    atol: float = 0.0
    safety: float = 0.8
    max_factor: float = 1.1
    min_factor: float = 0.2
    beta1: float = 3 / 25
    beta2: float = -1 / 25
    lag: int = 1
    reject_free: bool = False
    estimator_order: int = 4

    def propose(self, norm_history) -> float:
        index = len(norm_history) - 1 - self.lag
        if index < 0:
            return 1.0
        current = norm_history[index]
        if not math.isfinite(current):
            return self.min_factor
        if current <= 0.0:
            return self.max_factor
        previous = norm_history[index - 1] if index >= 1 else 1.0
        if not math.isfinite(previous) or previous <= 0.0:
            previous = 1.0
        factor = self.safety * (1.0 / current) ** self.beta1 * (1.0 / previous) ** self.beta2
        return min(max(factor, self.min_factor), self.max_factor)

    def start_factor(self, norm: float) -> float:
        if not math.isfinite(norm) or norm <= 0.0:
            return START_GROWTH if norm == 0.0 else self.min_factor
        factor = self.safety * (1.0 / norm) ** (1.0 / (self.estimator_order + 1))
        return min(max(factor, self.min_factor), START_GROWTH)

    def accepts(self, norm: float) -> bool:
        if not math.isfinite(norm):
            return False
        return self.reject_free or norm <= 1.0

    def rejection_factor(self, norm: float) -> float:
        if not math.isfinite(norm):
            return self.min_factor
        if norm <= 0.0:
            return 1.0
        factor = self.safety * (1.0 / norm) ** (1.0 / (self.estimator_order + 1))
        return min(max(factor, self.min_factor), 1.0)

    def equivalent_norm(self, norm: float, dz_taken: float, dz_proposed: float) -> float:
        if not math.isfinite(norm):
            return norm
        return norm * (dz_proposed / dz_taken) ** (self.estimator_order + 1)


def limit_step(dz: float, factor: float, dz_min: float, dz_max: float) -> float:
    proposed = min(dz * factor, dz_max)
    if proposed < dz_min:
        raise ConvergenceError(
            f"step control asked for {proposed:.3e} m, below the floor dz_min ="
            f" {dz_min:.3e} m; the tolerance cannot be met here"
        )
    return proposed


@dataclass(frozen=True)
class RKPolicy:
    controller: PIController
    dz_min: float
    dz_max: float
    auto_init: bool
    stretch: float
    z_end: float
    save_planes: tuple
    hard_kinks: tuple
    breakpoints: tuple


def decide(policy: RKPolicy, norm, dz, proposal):
    c = policy.controller
    return c.equivalent_norm(norm, dz_taken=dz, dz_proposed=proposal), c.accepts(norm)


def start_up(policy: RKPolicy, proposal, norm):
    c = policy.controller
    factor = c.start_factor(norm)
    following = limit_step(proposal, factor, policy.dz_min, policy.dz_max)
    recorded = (
        0.0
        if factor >= START_GROWTH
        else c.equivalent_norm(norm, dz_taken=proposal, dz_proposed=following)
    )
    return following, recorded


def propose(policy: RKPolicy, proposal, norms):
    return limit_step(proposal, policy.controller.propose(norms), policy.dz_min, policy.dz_max)


def reject(policy: RKPolicy, proposal, equivalent):
    return limit_step(
        proposal, policy.controller.rejection_factor(equivalent), policy.dz_min, policy.dz_max
    )


# ---------------------------------------------------------------------- landings

_PRIORITY = {"none": 0, "max_step": 1, "interval_width": 2, "hard_kink": 3,
             "save_plane": 4, "end": 5}


@dataclass(frozen=True)
class StepTargets:
    planes: tuple
    kinks: tuple
    edges: tuple
    dz_max: float
    z_end: float
    stretch: float = 1.1


@dataclass(frozen=True)
class Landing:
    dz: float
    target: str
    z_target: float | None
    stretched: bool = False


def step_targets(policy: RKPolicy, target: float) -> StepTargets:
    return StepTargets(
        tuple(policy.save_planes), tuple(policy.hard_kinks), tuple(policy.breakpoints),
        policy.dz_max, target, policy.stretch,
    )


def _order(bound):
    return bound[0], -_PRIORITY[bound[1]]


def landing(z: float, dz: float, targets: StepTargets) -> Landing:
    edges = targets.edges
    tolerance = LANDING_ULPS * math.ulp(max(abs(z), abs(targets.z_end)))
    index = min(max(bisect.bisect_right(edges, z) - 1, 0), len(edges) - 2)
    width = edges[index + 1] - edges[index]
    caps = [(dz, "none"), (targets.dz_max, "max_step"), (width, "interval_width")]
    landings = []
    for kind, positions in (
        ("hard_kink", targets.kinks),
        ("save_plane", targets.planes),
        ("end", (float(targets.z_end),)),
    ):
        found = bisect.bisect_left(positions, z + tolerance)
        if found < len(positions):
            landings.append((positions[found] - z, kind, positions[found]))
    cap_dz, cap_kind = min(caps, key=_order)
    gap, kind, ahead = min(landings, key=_order)
    if gap <= cap_dz:
        return Landing(gap, kind, ahead)
    if gap <= min(targets.stretch * cap_dz, targets.dz_max):
        return Landing(gap, kind, ahead, True)
    return Landing(cap_dz, cap_kind, None)


def reached(position: float, target: float) -> bool:
    tolerance = LANDING_ULPS * math.ulp(max(abs(position), abs(target)))
    return position >= target - tolerance


# ----------------------------------------------------------------- march state


@dataclass
class MarchState:
    z: float
    proposal: float
    lag: int = 1
    cursor: int = 0
    norms: deque = field(default_factory=deque)

    def __post_init__(self):
        self.norms = deque(self.norms, maxlen=self.lag + 2)


# ------------------------------------------------------- scripted decision traces


class _Exhausted(Exception):
    pass


def _schedule_edges(steps):
    """``FixedSchedule.edges``: the running sum of the steps from zero."""
    edges = [0.0]
    total = 0.0
    for s in steps:
        total = total + s
        edges.append(total)
    return edges


def run_scenario(sc: dict) -> dict:
    """Run one scripted scenario; return its ``events`` and ``end`` like the corpus."""
    c = sc["controller"]
    controller = PIController(**{k: v for k, v in c.items()})
    breakpoints = sc["breakpoints"] or [sc["z0"], sc["z_end"]]
    if not sc["interval_cap"]:
        breakpoints = [breakpoints[0], breakpoints[-1]]
    policy = RKPolicy(
        controller=controller,
        dz_min=float(sc["dz_min"]),
        dz_max=float(sc["dz_max"]),
        auto_init=bool(sc["auto_init"]),
        stretch=float(sc["stretch"]),
        z_end=float(sc["z_end"]),
        save_planes=tuple(float(z) for z in sc["save_planes"]),
        hard_kinks=tuple(float(z) for z in sc["kinks"]),
        breakpoints=tuple(float(z) for z in breakpoints),
    )
    quiet_map = {int(k): v for k, v in sc["quiet"].items()}
    split_map = {int(k): v for k, v in sc["split"].items()}
    order = controller.estimator_order + 1
    has_monitor = sc["quiet_interval"] is not None or sc["ceiling"] or sc["split"]
    interval = math.inf if sc["quiet_interval"] is None else sc["quiet_interval"]
    ceiling = sc["ceiling"]
    h = {
        "quiet": quiet_map.get(0, False),
        "step_cap": split_map.get(0, math.inf),
        "accepted": 0,
        "attempts": 0,
        "natural": 0,
        "pending": {},
    }
    events: list = []
    st = MarchState(z=sc["z0"], proposal=sc["proposal"], lag=controller.lag)

    def snapshot():
        return {
            "position": st.z,
            "proposal": st.proposal,
            "norms": [v if math.isfinite(v) else math.inf for v in st.norms],
            "accepted": h["accepted"],
            "natural": h["natural"],
            "quiet": h["quiet"],
            "cursor": st.cursor,
        }

    def close():
        if events and "after" not in events[-1] and "end" in events[-1]:
            events[-1]["after"] = snapshot()

    def trial(z, dz):
        close()
        index = h["attempts"]
        if index >= len(sc["norms"]):
            raise _Exhausted
        h["pending"] = {"z": z, "dz": dz}
        h["attempts"] += 1
        noise = float(sc["norms"][index])
        if sc["h0"] is None:
            return noise
        return noise * (dz / sc["h0"]) ** order

    def record(end, dz, norm, accepted):
        row = dict(h["pending"])
        row.update(end=end, step=dz, norm=norm if math.isfinite(norm) else math.inf,
                   accepted=accepted)
        events.append(row)
        h["pending"] = {}

    def cap(proposal, z):
        if not has_monitor:
            return proposal
        if h["quiet"]:
            proposal = min(proposal, interval)
        if ceiling is not None:
            proposal = min(proposal, ceiling)
        return min(proposal, h["step_cap"])

    def stepped():
        h["accepted"] += 1
        h["quiet"] = quiet_map.get(h["accepted"], h["quiet"])
        if has_monitor:
            h["step_cap"] = split_map.get(h["accepted"], h["step_cap"])

    def natural_stop(target):
        h["natural"] += 1
        every = sc["step_planes"]
        if every == 0 or h["natural"] % every:
            return False
        return not reached(st.z, target)

    def march_to(target):
        tg = StepTargets(
            policy.save_planes, policy.hard_kinks, policy.breakpoints,
            policy.dz_max, target, policy.stretch,
        )
        while not reached(st.z, target):
            stop = landing(st.z, cap(st.proposal, st.z), tg)
            norm = trial(st.z, stop.dz)
            equivalent, accepted = decide(policy, norm, stop.dz, st.proposal)
            end = stop.z_target if stop.z_target is not None else st.z + stop.dz
            record(end, stop.dz, norm, accepted)
            if accepted:
                st.z = end
                st.norms.append(equivalent)
                if policy.auto_init and len(st.norms) == 1:
                    st.proposal, st.norms[-1] = start_up(policy, st.proposal, equivalent)
                st.proposal = propose(policy, st.proposal, list(st.norms))
                stepped()
                if natural_stop(target):
                    return
            else:
                st.proposal = reject(policy, st.proposal, equivalent)

    steps = sc["schedule"]
    edges = None if steps is None else _schedule_edges(steps)

    def replay_to(target):
        n_steps = len(steps)
        total = edges[-1]
        ulps = max(LANDING_ULPS, n_steps)
        slack = ulps * math.ulp(max(total, abs(target)))
        while st.cursor < n_steps and st.z < target - slack:
            dz = float(steps[st.cursor])
            norm = trial(st.z, dz)
            if not math.isfinite(norm):
                raise ConvergenceError(
                    f"scripted: replayed step {st.cursor + 1} of {n_steps}, {dz:.4g} m "
                    f"from z = {st.z:.6g} m,"
                )
            end = float(edges[st.cursor + 1])
            st.cursor += 1
            st.z = end
            record(st.z, dz, norm, True)
            st.norms.append(norm)
            stepped()
            if natural_stop(target):
                return
        if st.z < target - slack:
            raise BackendError("replay shortfall")

    if sc["targets"] is not None:
        targets = list(sc["targets"])
    elif not sc["save_planes"]:
        targets = [sc["z_end"]]
    else:
        targets = sorted({*sc["save_planes"], sc["z_end"]})
    end = {"stop": "done"}
    try:
        for target in targets:
            while True:
                if steps is None:
                    march_to(target)
                else:
                    replay_to(target)
                close()
                events.append({"target": target, **snapshot()})
                if reached(st.z, target):
                    break
    except _Exhausted:
        end = {"stop": "exhausted"}
    except (BackendError, ConvergenceError) as error:
        end = {"stop": type(error).__name__, "message": str(error)}
    close()
    end["in_flight"] = dict(h["pending"])
    end.update(snapshot())
    return {"events": events, "end": end}
