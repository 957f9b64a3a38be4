"""Propagation throughput on one MI355X: accepted steps per second of the native march.

``python bench.py --gpus 1 --steps K --warmup W`` builds the workload's problem from
fixed parameters, loads it into the device workspace, takes W accepted steps untimed,
then times exactly K accepted steps of the adaptive march (every trial, rejected ones
included, is inside the timed region) with ``torch.cuda.synchronize`` on both sides,
and prints one JSON line. The inputs are resident in HBM before the timed region.

The default workload is the gas-filled capillary: 4 coupled HE_1m modes of a 125 um
bore, 1.5 m of helium at 0.8 -> 0 bar, a real field of 2^14 samples over 400 fs (8193
bins), a 7.5 fs 800 nm 275 uJ Gaussian, a 4-mode x 128-node radial quadrature, Kerr and
tunnelling ionisation, the exponential Adams march at rtol 1e-6. ``--workload radial``
runs the free-space argon beam on a 512-point Hankel transform (its synthesis and
projection as FP64 matrix-core GEMMs), ``--workload envelope`` the single-mode GNLSE
fibre case (2^13 bins, DOPRI5).

Beside the rate: the roofline figure of the trial's dominant kernel (its algorithmic
bytes, or for a GEMM its flops, over its average duration), every kernel's duration, and the CPU oracle timed on
a bounded sample of the same workload on this host's cores. The timed region replays
the captured trial graphs; the kernel durations come from HIP events recorded between
consecutive kernels on the workspace's stream while the same march continues for
``--profile-steps`` more accepted steps (graph-recorded events cannot be timed).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6


def trial_bytes(problem, order):
    """Algorithmic HBM bytes of each kernel of one Adams trial at order `order`.

    Each array the kernel must read or write once, at its stored width: complex (C, N)
    rows of 16 bytes per element, real (C, T) time rows and (Q, T) node rows of 8, the
    scan tiles' sums, the (Q, C) synthesis and (C, Q) projection tables; the rate-table
    lookups come from cache. W = order + 1 moments; the corrector reads W - 1 stored
    source rows and the predictor `order` of them; a varying operator adds the
    transported states of the rows read. With more than 8 modes (the dense route) the
    node passes read the node field the synthesis GEMM wrote, and the response writes
    two node streams for the projection GEMM.
    """
    C, N, T, Q = problem.n_modes, problem.n_omega, problem.n_to, problem.n_nodes
    n_phi = len(problem.phi_paths)
    W = order + 1
    cn = 16 * C * N
    ct = 8 * C * T  # one real time row per mode = C * T/2 complex
    qt = 8 * Q * T
    tiles = (T - 1 + 62) // 63
    qk = 8 * Q * tiles
    groups = (Q + 31) // 32  # node groups of the fused passes
    table = 8 * Q * C
    comp = n_phi > 1
    dense = C > 8
    plasma = problem.plasma is not None
    # The band rule's rate pass also stores the within-tile exposure, and so does the fused
    # route's under either rule (with one zero-tile flag a node and tile).
    band = plasma and (len(problem.plasma.taps) > 1 or not dense)
    n_streams = 3 if plasma else 1
    field = qt if dense else ct + table  # what a node pass reads to know E at its nodes
    raman = problem.raman is not None
    in_stage = problem.absorber_scheme == "in_stage"
    spatial = problem.spatial_absorber is not None
    # The doubled-grid drive (Q rows of 2T reals): written, transformed both ways (read and
    # written each time), multiplied by the kernel's spectrum in place.
    qt2 = 16 * Q * T
    return {
        # The predictor of a trial after an accepted step applies that step's commit: it
        # reads the source and advanced rows (and the state) in place of y and of the
        # newest stored row and state (which it reads back from cache right after storing
        # them), and writes the store's free row (and state row) and y beside its output.
        "adams_predict": cn * (n_phi + 1 + order + (order if comp else 0))
        + cn * (3 + (order + 1 if comp else 0)),
        "field_scatter": cn + 8 * T + ct,
        "inverse_fft": 2 * ct,
        "synthesis_gemm": table + ct + qt if dense else 0,
        "raman": field + qt2 * 7 if raman else 0,
        "plasma_rate": field + qt + qk + (qt if band else 0) + (Q * tiles if plasma and not dense else 0),
        "plasma_tile_prefix": 2 * qk,
        "plasma_current": field + qt + 2 * qk + 2 * qt,
        "plasma_tile_prefix_current": 2 * qk,
        # The fused-node route (at most 8 modes) with plasma: the free-current pass fused
        # with the response reads the exposure and rate (and the offsets and zero-tile
        # flags) and writes each node group's 2C projected rows and C projected tile
        # totals; the combination reads those and writes the 3C streams.
        "plasma_current_response": field + 2 * qt + qk + Q * tiles + groups * (2 * ct + 8 * C * tiles)
        + (qt if raman else 0),
        "plasma_combine": groups * (2 * ct + 8 * C * tiles) + 2 * 8 * T + 3 * ct,
        "node_response": (qt + (3 * qt + qk if plasma else 0) + (2 if plasma else 1) * qt)
        + (qt if raman else 0)
        if dense else ct + 2 * table + 3 * qt + qk + 2 * 8 * T + 3 * ct + (qt if raman else 0),
        "projection_gemm": table + (2 if plasma else 1) * qt + 2 * 8 * T + n_streams * ct
        if dense else 0,
        # The temporal term reads the modal fields and writes the stream (and reads it
        # back after the collar's GEMM, which reads its matrix and the modal fields).
        "absorber": (2 * ct + 8 * T + (ct if spatial else 0) if in_stage else 0)
        + (8 * C * C + 2 * ct if spatial else 0),
        "forward_fft": 2 * (n_streams + (in_stage or spatial)) * ct,
        "field_finish": n_streams * ct + cn * len(problem.prefactor_paths) + cn,
        # A trial's corrector finishes each bin's source itself: it reads the forward
        # transform's streams and the prefactor rows in place of the source, and writes
        # the source beside the advanced field.
        "adams_correct": cn * (n_phi + (W - 1) + 1 + (order + 1 if comp else 0))
        + n_streams * ct + cn * len(problem.prefactor_paths) + 2 * cn,
        "adams_norms": 8 * (1 + 3) * ((C * N + (64 if C * N <= 4096 * 64 else 256) - 1)
                                      // (64 if C * N <= 4096 * 64 else 256)),
    }


def trial_flops(problem):
    """Multiply-add flops of the dense route's GEMMs (2 per product term): the
    synthesis, the projection and, with a transverse collar, its matrix product."""
    C, T, Q = problem.n_modes, problem.n_to, problem.n_nodes
    if C <= 8:
        return {}
    width = 2 * T if problem.plasma is not None else T
    flops = {"synthesis_gemm": 2.0 * Q * C * T, "projection_gemm": 2.0 * C * Q * width}
    if problem.spatial_absorber is not None:
        flops["absorber"] = 2.0 * C * C * T
    return flops


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter runs (FETCH_SIZE and
    WRITE_SIZE, each scaled by what it reads for a plain copy of known size), or None when
    no run covers this workload's dominant kernel."""
    path = Path(__file__).resolve().parent / "profiles" / "traffic.json"
    try:
        row = json.loads(path.read_text()).get(workload, {}).get(kernel)
    except (OSError, ValueError, AttributeError):
        return None
    return None if not row else row["bytes"]


def workload_capillary(raman_share=0.0):
    from solprop_amd.problem import capillary_problem

    problem = capillary_problem(raman_share=raman_share)
    config = {
        "workload": "gas-filled capillary UPPE, 4 HE_1m modes, 125 um bore, 1.5 m helium "
                    "0.8->0 bar, real field 2^14 samples (8193 bins) over 400 fs, 7.5 fs "
                    "800 nm 275 uJ Gaussian, 4x128 radial quadrature, Kerr + tunnelling "
                    "plasma (band-limited running integral, "
                    f"{len(problem.plasma.taps) - 1} taps a side), exponential Adams "
                    "(order <= 5) rtol 1e-6",
        "n_t": problem.n_t, "n_modes": problem.n_modes, "n_nodes": problem.n_nodes,
        "stepper": "exponential Adams PEC, variable order", "parallelism": "single GPU",
    }
    return problem, config


def workload_radial(raman_share=0.0, collar=None):
    from solprop_amd.problem import radial_problem

    problem = radial_problem(raman_share=raman_share, collar=collar)
    config = {
        "workload": "free-space radial beam in argon, 1 bar, 0.5 m; 100 fs 800 nm Gaussian, "
                    "w0 1 mm, 2 mJ; 512-point Hankel transform (dense 512x512) over 4097 "
                    "frequency columns (2^13 samples over 1 ps); Kerr + tunnelling plasma "
                    f"(band-limited running integral, {len(problem.plasma.taps) - 1} taps a "
                    "side); "
                    "FP64 matrix-core GEMMs; exponential Adams (order <= 5) rtol 1e-6",
        "n_t": problem.n_t, "n_modes": problem.n_modes, "n_nodes": problem.n_nodes,
        "stepper": "exponential Adams PEC, variable order", "parallelism": "single GPU",
    }
    return problem, config


def workload_envelope():
    from solprop_amd.problem import gnlse_problem

    problem = gnlse_problem()
    config = {"workload": "envelope GNLSE, N=4 sech 50 fs 1550 nm, 2^13 bins over 10 ps, "
                          "n_to=2^14, Kerr + damped-oscillator Raman (f_R=0.18), shot noise "
                          "seed 1, DOPRI5 rtol 1e-6, 1 m",
              "n_t": problem.n_t, "n_to": problem.n_to, "n_modes": 1,
              "stepper": "DOPRI5 (FSAL)", "parallelism": "single GPU"}
    return problem, config


SCAN_DT = 400e-15 / 2**14  # the capillary grid's sample spacing, kept by every instance


def scan_problems(n, seed=3):
    """A batched parameter scan's instances: pressure uniform in 0.2-3 bar (a constant
    fill), energy log-uniform in 20-500 uJ, FWHM uniform in 5-30 fs, length uniform in
    0.5-3 m, 2^12, 2^13 or 2^14 samples (at the capillary grid's spacing, so 100, 200 or
    400 fs windows) and 1-6 HE_1m modes on a 128-node quadrature, from one seed."""
    from solprop_amd.problem import capillary_problem

    rng = np.random.default_rng(seed)
    pressure = rng.uniform(0.2, 3.0, n)
    energy = np.exp(rng.uniform(np.log(20e-6), np.log(500e-6), n))
    fwhm = rng.uniform(5e-15, 30e-15, n)
    length = rng.uniform(0.5, 3.0, n)
    n_t = rng.choice([2**12, 2**13, 2**14], n)
    modes = rng.integers(1, 7, n)
    return [capillary_problem(n_t=int(t), window=int(t) * SCAN_DT, n_modes=int(m), p0=float(p),
                              p1=float(p), energy=float(e), fwhm=float(f), length=float(ell))
            for t, m, p, e, f, ell in zip(n_t, modes, pressure, energy, fwhm, length)]


def run_scan(args):
    """A parameter scan as one batched march (``sp_scan``): every round runs each active
    instance's next trial as one captured graph of grouped launches; a "step" here is one
    such round. Beside it, the plain path: a sample of the same instances, each marched
    alone by its own workspace for the same number of attempts. Both report accepted
    instance-steps per second; the batched march's every instance equals its plain march
    to the bit, so the field error between them is zero."""
    from solprop_amd.native import Scan, Workspace

    n = args.instances
    problems = scan_problems(n)
    t0 = time.perf_counter()
    scan = Scan(problems)
    setup_s = time.perf_counter() - t0
    scan.start(0.0)
    targets = np.array([p.z_end for p in problems])
    if args.warmup > 0:
        scan.advance(targets, work_limit=args.warmup)
    t0 = time.perf_counter()
    batches = scan.advance(targets, work_limit=args.steps)
    elapsed = time.perf_counter() - t0
    accepted = int(sum(b.attempts[:, 3].sum() for b in batches))
    attempts = int(sum(len(b.attempts) for b in batches))
    timing = scan.timing()
    schedule = scan.schedule()
    finished = None
    if args.to_completion:
        t1 = time.perf_counter()
        rounds = 0
        while not all(scan.z[i] >= targets[i] for i in range(n)) and rounds < 1000:
            scan.advance(targets, work_limit=256)
            rounds += 1
        finished = {"seconds": time.perf_counter() - t1 + elapsed,
                    "done": int(sum(scan.z[i] >= targets[i] for i in range(n)))}
    scan.close()
    # The plain path on a sample: every (n / sample)-th instance alone, its graphs captured
    # by an untimed march first, then the same attempts timed.
    sample = list(range(0, n, max(1, n // args.plain_sample)))[: args.plain_sample]
    plain_accepted = 0
    plain_s = 0.0
    for i in sample:
        problem = problems[i]
        ws = Workspace(problem)
        ws.march(0.0, problem.field0).advance(problem.z_end, work_limit=args.graph_warmup)
        session = ws.march(0.0, problem.field0)
        if args.warmup > 0:
            session.advance(problem.z_end, work_limit=args.warmup)
        t1 = time.perf_counter()
        batch = session.advance(problem.z_end, work_limit=args.steps)
        plain_s += time.perf_counter() - t1
        plain_accepted += int(batch.attempts[:, 3].sum())
        ws.close()
    value = accepted / elapsed
    plain = plain_accepted / plain_s
    grids = {int(t): int(sum(p.n_to == t for p in problems)) for t in (2**12, 2**13, 2**14)}
    config = {
        "workload": f"batched parameter scan of {n} gas-filled capillary instances in one "
                    "captured graph per round: helium 0.2-3 bar, 20-500 uJ (log-uniform), "
                    "5-30 fs, 0.5-3 m, 2^12/2^13/2^14 samples, 1-6 HE_1m modes, 128 nodes, "
                    "Kerr + tunnelling plasma, exponential Adams rtol 1e-6 per instance",
        "instances": n, "grids": grids,
        "modes": {int(m): int(sum(p.n_modes == m for p in problems)) for m in range(1, 7)},
        "parallelism": "single GPU, one batched march",
    }
    extra = {
        "unit": "accepted instance-steps/s",
        "aggregate": "accepted instance-steps/s over all instances (a step: one round)",
        "attempts": attempts, "setup_s": setup_s,
        "scan_timing": timing,
        "autotuned_schedule": schedule,
        "plain": {"value": plain, "unit": "accepted instance-steps/s",
                  "sample": f"{len(sample)} of the instances, each marched alone "
                            f"({args.steps} attempts, graphs captured first)"},
        "speedup_vs_plain": value / plain,
        "field_error_vs_plain": 0.0,
    }
    if finished:
        extra["to_completion"] = finished
        extra["runs_per_hour"] = 3600.0 * finished["done"] / finished["seconds"]
    return value, 1e3 * elapsed / max(args.steps, 1), args.steps, config, None, extra, \
        problems[0]


def run_surrogate(args):
    """The learned scan (the project's own screening method): the batched solver
    generates training exit spectra (each training run marched to the longest length
    and sampled at `--lengths-per-run` random lengths on the way), the network trains on
    the device, screens `--screen-points` points, and the solver re-runs the top 1 %
    plus a 1 % random sample, judging the network's spectral error there. Reported as
    scan points resolved per hour against the plain path's runs per hour (points each
    marched alone)."""
    from oracle import fno as fno_oracle  # initial parameters only (a seeded draw)
    from solprop_amd import surrogate as sg
    from solprop_amd.native import Workspace

    rng = np.random.default_rng(args.seed)
    runs, k = args.train_runs, args.lengths_per_run
    pts = sg.draw(runs, rng)
    lengths = np.sort(rng.uniform(*sg.RANGES["length"], (runs, k)), axis=1)
    last = [time.perf_counter()]

    def progress(first, column, pending):
        # A line every 20 s on stderr, so that a long generation shows it is alive.
        if time.perf_counter() - last[0] > 20.0:
            last[0] = time.perf_counter()
            print(f"solver: chunk from run {first}, length {column}, {pending} marching",
                  file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    solved = sg.solve(pts[:, :3], lengths, chunk=args.chunk, max_attempts=args.max_attempts,
                      progress=progress)
    gen_s = time.perf_counter() - t0
    spectra, params, targets, run_of = sg.training_rows(pts[:, :3], lengths, solved)
    held = run_of % 10 == 0
    shape = sg.Shape()
    net = sg.Network(shape, max_batch=64)
    o = fno_oracle.Shape(in_ch=shape.in_ch, width=shape.width, modes=shape.modes,
                         layers=shape.layers, hidden=shape.hidden, n=shape.n)
    net.set_parameters(fno_oracle.init(o, seed=args.seed))
    net.load_dataset(spectra[~held], params[~held], targets[~held] - spectra[~held])
    print(f"generated {len(spectra)} rows in {gen_s:.0f} s; training", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    losses, train_ms = net.train(epochs=args.epochs, batch=64, lr=args.lr, seed=args.seed)
    train_s = time.perf_counter() - t0
    pred_held = sg.predict(net, spectra[held], params[held])
    err_held = sg.spectral_error(pred_held, targets[held])
    err_nn = sg.spectral_error(sg.nearest_neighbour(params[~held], targets[~held],
                                                    params[held]), targets[held])
    points = sg.draw(args.screen_points, np.random.default_rng(args.seed + 1))
    screening = sg.screen(net, points, seed=args.seed)
    net.close()
    picks = np.concatenate([screening.top, screening.random])
    print(f"trained ({len(losses)} steps), screened {args.screen_points} points; verifying "
          f"{len(picks)}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    verified = sg.solve(points[picks], chunk=args.chunk, max_attempts=args.max_attempts,
                        progress=progress)
    verify_s = time.perf_counter() - t0
    ok = verified.ok[:, 0]
    pred = np.stack([screening.predicted[int(i)] for i in picks])
    err = sg.spectral_error(pred[ok], verified.exits[ok, 0])
    true_uv = sg.uv_fraction(verified.exits[:, 0])
    n_top = len(screening.top)
    top_ok, rand_ok = ok[:n_top], ok[n_top:]
    threshold = np.quantile(true_uv[n_top:][rand_ok], 0.99)
    # The plain path: a sample of the screened points, each marched alone to its end.
    sample = points[rng.choice(len(points), size=args.plain_sample, replace=False)]
    plain_s = 0.0
    plain_done = 0
    for p in sample:
        problem = sg.scan_problem(*p)
        ws = Workspace(problem)
        march = ws.march(0.0, problem.field0)
        t1 = time.perf_counter()
        attempts = 0
        try:
            while attempts < args.max_attempts:
                batch = march.advance(problem.z_end, work_limit=4096)
                attempts += len(batch.attempts)
                if batch.done:
                    plain_done += 1
                    break
        except Exception:
            pass
        plain_s += time.perf_counter() - t1
        ws.close()
    plain_rate = 3600.0 * args.plain_sample / plain_s  # runs per hour, every attempt paid
    total_s = gen_s + train_s + screening.seconds + verify_s
    value = 3600.0 * args.screen_points / total_s
    config = {
        "workload": f"surrogate-screened scan: {args.screen_points} capillary points "
                    "(helium 0.2-3 bar, 20-500 uJ, 5-30 fs, 0.5-3 m; 2^12 samples over 400 fs, "
                    "HE11, 32 nodes, Kerr + tunnelling plasma, exponential Adams rtol 1e-6) "
                    "screened by a 1-D Fourier neural operator (4 layers, width 64, 64 modes, "
                    "BF16 matrix cores, FP32 accumulation) trained on the batched solver's exit "
                    f"spectra ({runs} runs x {k} lengths); top 1 % + 1 % random re-run by the "
                    "solver",
        "screen_points": args.screen_points, "train_runs": runs, "lengths_per_run": k,
        "parallelism": "single GPU",
    }
    extra = {
        "metric": "scan runs/hour vs plain FP64 at stated field error",
        "unit": "scan points/hour",
        "aggregate": "scan points resolved per hour (generation, training, screening and "
                     "verification included)",
        "plain_runs_per_hour": plain_rate,
        "speedup_vs_plain": value / plain_rate,
        "plain_sample": f"{args.plain_sample} screened points marched alone "
                        f"({plain_done} to their end, at most {args.max_attempts} attempts)",
        "generation": {"seconds": gen_s, "rows": int(len(spectra)),
                       "runs_dropped": int((~solved.ok[:, -1]).sum()),
                       "attempts_median": float(np.median(solved.attempts))},
        "training": {"seconds": train_s, "device_ms": train_ms, "steps": int(len(losses)),
                     "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-20:].mean()),
                     "held_out_error_decades_median": float(np.median(err_held)),
                     "nearest_neighbour_error_decades_median": float(np.median(err_nn))},
        "screening": {"seconds": screening.seconds,
                      "points_per_second": args.screen_points / screening.seconds},
        "verification": {"seconds": verify_s, "runs": int(len(picks)), "reached": int(ok.sum()),
                         "error_decades_median": float(np.median(err)),
                         "error_decades_p90": float(np.quantile(err, 0.9)),
                         "within_0.1_decade": float(np.mean(err <= 0.1)),
                         "density_error_median": float(np.median(sg.density_error(
                             pred[ok], verified.exits[ok, 0]))),
                         "top_above_random_p99": float(np.mean(true_uv[:n_top][top_ok] >= threshold)),
                         "true_uv_top_mean": float(true_uv[:n_top][top_ok].mean()),
                         "true_uv_random_mean": float(true_uv[n_top:][rand_ok].mean())},
        "field_error_verified": "0 (verified points are the solver's own exit fields)",
    }
    return value, 1e3 * total_s / max(args.screen_points, 1), args.screen_points, config, None, \
        extra, None


def run_fno(args):
    """The surrogate network's training step and inference alone (the full shape: 4
    layers, width 64, 64 modes, 2048 bins, batch 64) on synthetic rows: a step is one
    Adam step over a batch; beside it the forward rate in samples per second. The FLOP
    count is the network's dense products (the BF16 matrix cores) per sample."""
    from oracle import fno as fno_oracle  # initial parameters only (a seeded draw)
    from solprop_amd import surrogate as sg

    shape = sg.Shape()
    batch = 64
    rng = np.random.default_rng(args.seed)
    rows = 4096
    spectra = rng.uniform(0, 1, (rows, shape.n)).astype(np.float32)
    params = rng.uniform(0, 1, (rows, shape.n_params)).astype(np.float32)
    targets = rng.uniform(-0.1, 0.1, (rows, shape.n)).astype(np.float32)
    net = sg.Network(shape, max_batch=batch)
    o = fno_oracle.Shape(in_ch=shape.in_ch, width=shape.width, modes=shape.modes,
                         layers=shape.layers, hidden=shape.hidden, n=shape.n)
    net.set_parameters(fno_oracle.init(o, seed=args.seed))
    net.load_dataset(spectra, params, targets)
    per_epoch = rows // batch
    warm = max(1, args.warmup)
    net.train(epochs=max(1, (warm + per_epoch - 1) // per_epoch), batch=batch, seed=1)
    epochs = max(1, (args.steps + per_epoch - 1) // per_epoch)
    losses, ms = net.train(epochs=epochs, batch=batch, seed=2)
    steps = len(losses)
    t0 = time.perf_counter()
    net.forward(spectra, params)
    infer_s = time.perf_counter() - t0
    net.close()
    W, M, n, H, L = shape.width, shape.modes, shape.n, shape.hidden, shape.layers
    # Dense products a sample, forward: per layer the truncated transform and its inverse
    # (W x n x 2M each) and the 1x1 convolution (W x W x n); the lift's head (H x W x n).
    forward = 2.0 * (L * (2 * W * n * 2 * M + W * W * n) + H * W * n)
    # Training: the forward, then per product its two gradients.
    train = 3.0 * forward
    value = steps * batch / (ms * 1e-3)
    achieved = train * batch * steps / (ms * 1e-3) / 1e12
    config = {"workload": "surrogate network training: 1-D Fourier neural operator, 4 layers, "
                          "width 64, 64 modes, 2048 bins, batch 64, Adam, BF16 matrix-core "
                          "products with FP32 accumulation, synthetic rows",
              "batch": batch, "parallelism": "single GPU"}
    roofline = {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                "frac": achieved / 2500.0, "traffic": None, "kernel": "training step",
                "flops_per_sample": train,
                "timing": "device time of the training steps (events around the loop)"}
    extra = {"metric": "surrogate training samples/s", "unit": "samples/s",
             "inference_samples_per_s": rows / infer_s,
             "inference_tflops": forward * rows / infer_s / 1e12, "train_steps": steps,
             "dtype_note": "bf16 products, fp32 accumulation and optimizer"}
    return value, ms / steps, steps, config, roofline, extra, None


# Algorithmic bytes a launch moves, per point of the volume: a transform pass reads and
# writes each complex point once (32); the t rows' forward pass also reads the prefactor
# (48); the register passes read and write the low-storage registers (pass 1: y, source,
# operator in, y, g, B, C out; 2: source, g, B, y in, B, C, g out; 3: source, g, B,
# operator, y in, B, C out; 4: operator, B, source in, y out).
CARTESIAN_PHASE_BYTES = {
    "record upload": 0.0, "t inverse": 32.0, "x pass": 32.0, "y response": 32.0,
    "t forward": 48.0, "ls pass 1": 112.0, "ls pass 2": 112.0, "ls pass 3": 112.0,
    "ls pass 4": 64.0, "norms": 0.0,
}
CARTESIAN_LAUNCHES = {"t inverse": 4, "x pass": 8, "y response": 4, "t forward": 4,
                      "ls pass 1": 1, "ls pass 2": 1, "ls pass 3": 1, "ls pass 4": 1}


def run_cartesian(args):
    """The free-space Kerr beam on a Cartesian (ny, nx, nt) grid, fixed RK4IP steps of
    0.5 um (the reference's ``kerr_beam_*`` cases, benchmarks/registry.py:348-379, on
    power-of-two axes): a step is one low-storage step of the whole volume, four source
    evaluations of two rank-3 transforms each. The rate is point-steps per second (the
    volume's samples times the steps)."""
    import torch

    from solprop_amd.native import Workspace
    from solprop_amd.problem import cartesian_problem

    nx, ny, nt = (int(v) for v in args.cartesian_grid.split(","))
    total = 1 + args.warmup + args.steps + args.profile_steps
    t_build = time.perf_counter()
    problem = cartesian_problem(nx=nx, ny=ny, n_t=nt, spacing=1.05e-6, steps=total,
                                length=0.5e-6 * total)
    build_s = time.perf_counter() - t_build
    volume = nx * ny * nt
    ws = Workspace(problem)
    march = ws.march(0.0, problem.field0)
    edges = problem.schedule

    def take(count, start):
        batch = march.advance(float(edges[start + count]), work_limit=count)
        assert len(batch.attempts) == count
        return start + count

    at = take(1 + args.warmup, 0)  # the first step captures the step's graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    at = take(args.steps, at)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ws.profile(True)
    if args.profile_steps > 0:
        at = take(args.profile_steps, at)
    torch.cuda.synchronize()
    trial_us, trials = ws.timing()
    phases = ws.phases()
    device_bytes = ws.counters()["device_bytes"]
    ws.close()
    value = volume * args.steps / elapsed
    kernels = {}
    for name, us in phases.items():
        per = CARTESIAN_PHASE_BYTES.get(name, 0.0) * volume
        kernels[name] = {"avg_us": us, "bytes": per,
                         "GB/s": per / (us * 1e-6) / 1e9 if us > 0 and per > 0 else None}
    # The dominant kernel by its share of the step (launches a step times duration).
    per_step = CARTESIAN_LAUNCHES
    shares = {k: phases[k] * n for k, n in per_step.items() if k in phases}
    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None, "kernel": None}
    if shares:
        dominant = max(shares, key=shares.get)
        achieved = kernels[dominant]["GB/s"]
        roofline = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("cartesian", dominant),
            "kernel": dominant, "bytes_per_launch": kernels[dominant]["bytes"],
            "avg_launch_us": phases[dominant], "trial_us": trial_us, "kernels": kernels,
            "profiled_trials": trials,
            "step_bytes": sum(CARTESIAN_PHASE_BYTES[k] * n for k, n in per_step.items())
            * volume,
            "timing": "HIP events between consecutive launches on the workspace stream, "
                      "over the profiled steps that follow the timed region",
        }
    config = {"workload": f"free-space Kerr beam in fused silica on a Cartesian "
                          f"{ny} x {nx} x {nt} envelope grid (y, x, t), 260 nm 30 fs 0.1 uJ "
                          "Gaussian of 40 um waist, 1.05 um transverse spacing, 250 fs window, "
                          "fixed low-storage RK4IP steps of 0.5 um",
              "grid": [ny, nx, nt], "points": volume, "parallelism": "single GPU"}
    extra = {"metric": "Cartesian point-steps/s", "unit": "point-steps/s",
             "device_bytes": device_bytes, "host_build_s": build_s,
             "reference_published": {
                 "ms_per_step": 208.0, "grid": [768, 768, 600], "device": "AMD Instinct MI325X",
                 "point_steps_per_s": 768 * 768 * 600 / 0.208,
                 "source": "docs/dev/benchmarks.md:533-535 (a low-storage step at 768 x 768 "
                           "x 600, after the 2026-11-05 changes)"}}
    return value, elapsed * 1e3 / args.steps, args.steps, config, roofline, extra, problem


def cartesian_cpu_baseline(seconds: float = 12.0):
    """The oracle's NumPy replay of the same route on a 128 x 128 x 256 sample of the
    workload (one thread), steps until about `seconds` have passed."""
    from oracle import cartesian as cart
    from oracle import low_storage as ls
    from solprop_amd.problem import cartesian_problem

    problem = cartesian_problem(nx=128, ny=128, n_t=256, spacing=1.05e-6 * 4, steps=64,
                                length=0.5e-6 * 64)
    state = problem.field0.copy()
    norm_mask = problem.norm_mask.astype(bool)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds and steps < 64:
        z = problem.schedule[steps]
        dz = problem.schedule[steps + 1] - z
        state, _, _ = ls.step(state, z, dz, lambda u, _z: cart.source(problem, u),
                              lambda a, b: cart.flows(problem, a, b), norm_mask)
        steps += 1
    elapsed = time.perf_counter() - t0
    points = 128 * 128 * 256
    return {"value": points * steps / elapsed, "unit": "point-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"the oracle's NumPy replay (numpy.fft, one thread) of {steps} steps on "
                      f"a 128 x 128 x 256 grid of the same beam, {elapsed:.1f} s"}


def run_gpu(args):
    import torch

    from solprop_amd.native import Workspace

    if args.workload == "envelope":
        problem, config = workload_envelope()
    else:
        if args.workload == "capillary":
            problem, config = workload_capillary(args.raman_share)
        else:
            collar = None
            if args.collar:
                width_mm, alpha0 = (float(v) for v in args.collar.split(","))
                collar = (width_mm * 1e-3, alpha0)
            problem, config = workload_radial(args.raman_share, collar)
            if collar:
                config["workload"] += (f"; transverse collar over the outer {width_mm:g} mm, "
                                       f"peak {alpha0:g} / m")
        if args.raman_share > 0.0:
            config["workload"] += (f"; molecular Raman response taking {args.raman_share:g} "
                                   "of the third-order response")
        if args.low_storage_steps > 0:
            # The low-storage RK4IP replay of a uniform schedule instead of the adaptive
            # march: four source evaluations a step, every step accepted.
            import dataclasses

            edges = np.linspace(0.0, problem.z_end, args.low_storage_steps + 1)
            problem = dataclasses.replace(problem, method="low_storage", schedule=edges)
            config["stepper"] = (f"low-storage RK4IP replay, {args.low_storage_steps} uniform "
                                 "steps")
            config["workload"] += (f"; low-storage RK4IP replay of {args.low_storage_steps} "
                                   "uniform steps instead of the adaptive march")
    if args.copy_probe_mib > 0:
        # A plain device copy of known size, three times, for scaling memory counters.
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ws = Workspace(problem)
    # Capture every trial graph the march will use (each stencil width as the order
    # ramps up, with and without a commit at its head) before anything is timed: a march
    # of `--graph-warmup` attempts from the launch field, then the launch field loaded
    # again (loading keeps the captured graphs).
    if args.graph_warmup > 0:
        ws.march(0.0, problem.field0).advance(problem.z_end, work_limit=args.graph_warmup)
    session = ws.march(0.0, problem.field0)

    def take(count):
        accepted = attempts = 0
        done = False
        while accepted < count and not done:
            batch = session.advance(problem.z_end, work_limit=count - accepted)
            accepted += int(batch.attempts[:, 3].sum())
            attempts += len(batch.attempts)
            done = batch.done
        return accepted, attempts

    take(args.warmup)
    torch.cuda.synchronize()
    c0 = ws.counters()
    z0 = session.z
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps, attempts = take(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    c1 = ws.counters()
    z1 = session.z
    # The kernels' own durations: the march continues for up to `--profile-steps`
    # accepted steps with an event between consecutive kernels of every trial (the
    # same kernels as direct launches, since graph-recorded events cannot be timed).
    ws.profile(True)
    take(args.profile_steps)
    torch.cuda.synchronize()
    trial_us, trials = ws.timing()
    phases = ws.phases()
    ws.close()
    config["z_span_m"] = [z0, z1]
    value = steps / elapsed
    if phases:
        algorithmic = trial_bytes(problem, problem.adams_order)
        flops = trial_flops(problem)
        kernels = {}
        for name, us in phases.items():
            row = {"avg_us": us, "bytes": algorithmic[name],
                   "GB/s": algorithmic[name] / (us * 1e-6) / 1e9 if us > 0 else None}
            if name in flops:
                row["flops"] = flops[name]
                row["TFLOP/s"] = flops[name] / (us * 1e-6) / 1e12 if us > 0 else None
            kernels[name] = row
        dominant = max(phases, key=phases.get)
        if dominant in flops:
            achieved = kernels[dominant]["TFLOP/s"]
            bound, peak, unit = "mfma", FP64_MFMA_PEAK_TFLOPS, "TFLOP/s"
            per_launch = {"flops_per_launch": flops[dominant]}
        else:
            achieved = kernels[dominant]["GB/s"]
            bound, peak, unit = "hbm", HBM_PEAK_GBS, "GB/s"
            per_launch = {"bytes_per_launch": algorithmic[dominant]}
        roofline = {
            "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": None if achieved is None else achieved / peak,
            "traffic": measured_traffic(args.workload, dominant), "kernel": dominant,
            **per_launch,
            "avg_launch_us": phases[dominant],
            "trial_us": trial_us, "kernels": kernels, "profiled_trials": trials,
            "timing": "HIP events between consecutive kernels on the workspace stream, "
                      "over the profiled trials that follow the timed region; the fused "
                      "plasma column's current-and-response pass by its own start and end "
                      "events (hipExtLaunchKernel)",
        }
    else:
        roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": None, "traffic": None, "kernel": "trial graph",
                    "trial_us": trial_us}
    extra = {"attempts": attempts, "evaluations": c1["evaluations"] - c0["evaluations"]}
    return value, elapsed * 1e3 / max(steps, 1), steps, config, roofline, extra, problem


def host_cores() -> int:
    """The host threads the CPU baseline uses: this process's CPU affinity, capped by
    OMP_NUM_THREADS where the host sets it (a GPU host's share of its cores)."""
    cores = len(os.sched_getaffinity(0))
    limit = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(cores, limit) if limit > 0 else cores)


def cpu_baseline(problem, seconds: float = 12.0, single_seconds: float = 6.0):
    """The oracle's march on this host, on the first steps of the same workload.

    Every core this process may use (``host_cores``): scipy.fft runs that many workers
    and the BLAS pool that many threads. The one-thread figure of the same march (one
    FFT worker, BLAS held to one thread) is reported beside it.
    """
    import scipy.fft
    from threadpoolctl import threadpool_limits

    cores = host_cores()
    with scipy.fft.set_workers(cores), threadpool_limits(limits=cores):
        result = _cpu_baseline(problem, seconds, cores)
    with scipy.fft.set_workers(1), threadpool_limits(limits=1):
        single = _cpu_baseline(problem, single_seconds, 1)
    result["single_thread"] = {"value": single["value"], "sample": single["sample"]}
    return result


def _cpu_baseline(problem, seconds, cores):
    y0 = problem.field0.reshape(problem.shape).copy()
    steps = [0]
    t0 = time.perf_counter()

    def on_attempt(z, dz, norm, accepted):
        if accepted:
            steps[0] += 1
        if time.perf_counter() - t0 > seconds:
            raise StopIteration

    if problem.method == "adams":
        from oracle import adams

        adams.march(problem, y0, on_attempt=on_attempt)
        what = "exponential Adams march"
    else:
        from oracle import rk

        try:
            rk.march(problem, y0, on_attempt=on_attempt)
        except StopIteration:
            pass
        what = "DOPRI5 march (every stage, no FSAL)"
    elapsed = time.perf_counter() - t0
    return {
        "value": steps[0] / elapsed,
        "unit": "accepted steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the oracle's NumPy/SciPy {what} from z=0 for {elapsed:.1f} s: "
                  f"{steps[0]} accepted steps, {cores} thread{'s' if cores > 1 else ''} "
                  "(scipy.fft workers and BLAS threads)",
    }


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=300)
    parser.add_argument("--warmup", type=int, default=30)
    parser.add_argument("--workload",
                        choices=["capillary", "radial", "envelope", "scan", "surrogate", "fno",
                                 "cartesian"],
                        default="capillary")
    parser.add_argument("--train-runs", type=int, default=2048,
                        help="surrogate workload: solver runs generating training spectra")
    parser.add_argument("--lengths-per-run", type=int, default=8)
    parser.add_argument("--screen-points", type=int, default=100000)
    parser.add_argument("--epochs", type=int, default=30)
    parser.add_argument("--lr", type=float, default=1e-3)
    parser.add_argument("--chunk", type=int, default=2048,
                        help="surrogate workload: instances in one batched march")
    parser.add_argument("--max-attempts", type=int, default=40000)
    parser.add_argument("--seed", type=int, default=0)
    parser.add_argument("--instances", type=int, default=512,
                        help="scan workload: instances in the batched march")
    parser.add_argument("--plain-sample", type=int, default=16,
                        help="scan workload: instances also marched alone, for the plain rate")
    parser.add_argument("--to-completion", action="store_true",
                        help="scan workload: then march every instance to its end "
                             "(runs per hour)")
    parser.add_argument("--cartesian-grid", default="512,512,1024",
                        help="cartesian workload: nx,ny,nt (powers of two, at most 4096)")
    parser.add_argument("--profile-steps", type=int, default=100)
    parser.add_argument("--graph-warmup", type=int, default=40,
                        help="attempts of an untimed march from the launch field that "
                             "capture every trial graph before the warm-up steps")
    parser.add_argument("--low-storage-steps", type=int, default=0,
                        help="field workloads: replay this many uniform steps with the "
                             "low-storage RK4IP march instead of the adaptive one")
    parser.add_argument("--copy-probe-mib", type=int, default=0,
                        help="run a plain device copy of this many MiB first (counter scaling)")
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--collar", default="",
                        help="radial workload: add a transverse collar, 'WIDTH_MM,ALPHA0' "
                             "(its matrix product runs as a third GEMM)")
    parser.add_argument("--raman-share", type=float, default=0.0,
                        help="add a damped-oscillator Raman response taking this share of "
                             "the third-order response (field workloads; off by default)")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py runs one GPU: one propagation problem runs on one device, and this "
              "build measures the single-GPU path only (--gpus 1).", file=sys.stderr)
        sys.exit(2)
    runner = {"scan": run_scan, "surrogate": run_surrogate, "fno": run_fno,
              "cartesian": run_cartesian}.get(args.workload, run_gpu)
    value, ms, steps, config, roofline, extra, problem = runner(args)
    if args.no_cpu or args.workload in ("scan", "surrogate", "fno"):
        baseline = None
    elif args.workload == "cartesian":
        baseline = cartesian_cpu_baseline()
    else:
        baseline = cpu_baseline(problem)
    line = {
        "metric": extra.pop("metric", "propagation steps/s per GPU"),
        "value": value,
        "unit": extra.pop("unit", "accepted steps/s"),
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (parametric builders: Marcatili modes, ADK helium rate, "
                "Gaussian input; no material database)",
        "config": config,
        **extra,
        "roofline": roofline,
        "cpu_baseline": baseline,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    os.environ.setdefault("PYTHONUNBUFFERED", "1")
    main()
