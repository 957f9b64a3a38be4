"""Build recipe of the device library: hipcc for gfx950, in-tree.

``python -m solprop_amd.build`` compiles every HIP source under ``csrc/`` for gfx950
and links ``solprop_amd/libsolprop_mi355x.so``. Host files whose floating-point
decisions must match the reference's to the bit (the controller, the scalar paths)
are compiled with ``g++ -ffp-contract=off``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libsolprop_mi355x.so"
ARCH = "gfx950"

HIP_SOURCES = ["fft.hip", "gemm.hip", "stage.hip", "field.hip", "workspace.hip", "field_ws.hip",
               "stats.hip", "scan.hip", "fft_noalloc.hip",
               "fno.hip", "cartesian.hip"]
HOST_SOURCES = ["capi.cpp", "session.cpp"]
STRICT_SOURCES = ["march.cpp", "trace.cpp", "scalars.cpp", "adams_march.cpp",
                  "adams_trace.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 instructions are kept out of every kernel on this target.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return found


def _run(cmd: list[str]) -> None:
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        lines = [ln for ln in (result.stdout + result.stderr).splitlines()
                 if "packed-fp32-ops" not in ln]
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + "\n".join(lines))


def build(verbose: bool = False) -> Path:
    """Compile and link the library; return its path."""
    work = HERE / "_build"
    work.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objects = []
    missing = [name for name in (*HIP_SOURCES, *HOST_SOURCES, *STRICT_SOURCES)
               if not (CSRC / name).exists()]
    if missing:
        raise FileNotFoundError("listed library sources are missing: " + ", ".join(missing))
    for name in HIP_SOURCES:
        src = CSRC / name
        obj = work / (name + ".o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        objects.append(obj)
    for name in HOST_SOURCES:
        obj = work / (name + ".o")
        jobs.append([hipcc, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off",
                     "-c", str(CSRC / name), "-o", str(obj)])
        objects.append(obj)
    for name in STRICT_SOURCES:
        obj = work / (name + ".o")
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off",
                     "-fno-fast-math", "-c", str(CSRC / name), "-o", str(obj)])
        objects.append(obj)
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        for cmd, _ in zip(jobs, pool.map(_run, jobs), strict=True):
            if verbose:
                print(" ".join(cmd))
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp),
          *[str(o) for o in objects]])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
