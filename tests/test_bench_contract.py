"""The measurement script's contract on the host (no GPU): one GPU only, refused with a
message and a nonzero status for more, before anything touches a device; the
algorithmic byte and flop formulas cover every phase the workspace reports."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    run = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert run.returncode == 2
    assert "one GPU" in run.stderr


def test_every_phase_has_algorithmic_bytes():
    sys.path.insert(0, str(ROOT))
    import bench
    from solprop_amd.problem import capillary_problem, radial_problem

    phases = ["adams_predict", "field_scatter", "inverse_fft", "synthesis_gemm", "raman",
              "plasma_rate", "plasma_tile_prefix", "plasma_current",
              "plasma_tile_prefix_current", "node_response", "projection_gemm", "absorber",
              "forward_fft", "field_finish", "adams_correct", "adams_norms"]
    for problem in (capillary_problem(n_t=2048, n_nodes=32, integration="trapezoid"),
                    radial_problem(n_t=1024, n_r=32, integration="trapezoid",
                                   collar=(1e-3, 2e3))):
        table = bench.trial_bytes(problem, 5)
        assert set(phases) <= set(table)
        assert all(v >= 0 for v in table.values())
    flops = bench.trial_flops(problem)
    assert flops["synthesis_gemm"] == 2.0 * 32 * 32 * 1024
    assert flops["absorber"] == 2.0 * 32 * 32 * 1024
