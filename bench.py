"""Propagation throughput on one MI355X: accepted steps per second of the native march.

``python bench.py --gpus 1 --steps K --warmup W`` builds the workload's problem from a
fixed seed, loads it into the device workspace, takes W accepted steps untimed, then
times exactly K accepted steps of the adaptive march (every trial, rejected ones
included, is inside the timed region) with ``torch.cuda.synchronize`` on both sides,
and prints one JSON line. The inputs are resident in HBM before the timed region.

Beside it: the roofline figure of the dominant unit (algorithmic bytes per trial over
its HIP-event duration on the workspace's stream) and the CPU oracle timed on a bounded
sample of the same workload on this host's cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def workload_envelope():
    from solprop_amd.problem import gnlse_problem

    problem = gnlse_problem()  # n_t = 2^13 over 10 ps, n_to = 2^14, sech N=4, Raman
    n, t = problem.n_omega, problem.n_to
    # Pass-level algorithmic bytes of one source evaluation and of one accepted
    # DOPRI5 step (6 evaluations with FSAL + the accept pass): 232 N + 296 T per
    # evaluation, 96 N for the accept.
    per_eval = 232 * n + 296 * t
    per_step = 6 * per_eval + 96 * n
    config = {"workload": "envelope GNLSE, N=4 sech 50 fs 1550 nm, 2^13 bins over 10 ps, "
                          "n_to=2^14, Kerr + damped-oscillator Raman (f_R=0.18), shot noise "
                          "seed 1, DOPRI5 rtol 1e-6, 1 m",
              "n_t": problem.n_t, "n_to": t, "n_modes": 1, "stepper": "DOPRI5 (FSAL)",
              "parallelism": "single GPU"}
    return problem, per_step, per_eval, config


def run_gpu(args):
    import torch

    from solprop_amd.native import Workspace

    problem, bytes_per_step, bytes_per_eval, config = workload_envelope()
    ws = Workspace(problem)
    session = ws.start(0.0, problem.field0)
    accepted = 0
    while accepted < args.warmup:
        batch = session.advance(problem.z_end, work_limit=1)
        accepted += int(batch.attempts[:, 3].sum())
    torch.cuda.synchronize()
    ws.profile(True)
    c0 = ws.counters()
    t0 = time.perf_counter()
    accepted = 0
    attempts = 0
    while accepted < args.steps:
        batch = session.advance(problem.z_end, work_limit=1)
        accepted += int(batch.attempts[:, 3].sum())
        attempts += len(batch.attempts)
        if batch.done:
            break
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    c1 = ws.counters()
    trial_us, trials = ws.timing()
    evals = c1["evaluations"] - c0["evaluations"]
    ws.close()
    steps = accepted
    value = steps / elapsed
    # The trial graph is the unit timed by events: its bytes are the evaluations it
    # made plus the accept pass.
    bytes_per_trial = (evals / max(trials, 1)) * bytes_per_eval + 96 * problem.n_omega
    achieved = bytes_per_trial / (trial_us * 1e-6) / 1e9 if trial_us > 0 else None
    roofline = {
        "unit": "GB/s",
        "bound": "hbm",
        "kernel": "trial graph (all stage kernels of one DOPRI5 trial)",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
        "traffic": None,
        "bytes_per_launch": bytes_per_trial,
        "avg_launch_us": trial_us,
    }
    return value, elapsed * 1e3 / max(steps, 1), steps, attempts, config, roofline


def cpu_baseline(seconds: float = 12.0):
    """The oracle's march on this host, on the first steps of the same workload."""
    from oracle import rk
    from solprop_amd.problem import gnlse_problem

    problem = gnlse_problem()
    y0 = problem.field0.reshape(problem.shape)
    steps = [0]
    t0 = time.perf_counter()

    def on_attempt(z, dz, norm, accepted):
        if accepted:
            steps[0] += 1
        if time.perf_counter() - t0 > seconds:
            raise StopIteration

    try:
        rk.march(problem, y0, on_attempt=on_attempt)
    except StopIteration:
        pass
    elapsed = time.perf_counter() - t0
    return {
        "value": steps[0] / elapsed,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"the oracle's NumPy/SciPy march (every stage, no FSAL) from z=0 for "
                  f"{elapsed:.1f} s: {steps[0]} accepted steps",
    }


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=400)
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--no-cpu", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py runs one GPU: one propagation problem runs on one device, and this "
              "build measures the single-GPU path only (--gpus 1).", file=sys.stderr)
        sys.exit(2)
    value, ms, steps, attempts, config, roofline = run_gpu(args)
    baseline = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "propagation steps/s per GPU",
        "value": value,
        "unit": "accepted steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded builders; no material database)",
        "config": config,
        "attempts": attempts,
        "roofline": roofline,
        "cpu_baseline": baseline,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    os.environ.setdefault("PYTHONUNBUFFERED", "1")
    main()
