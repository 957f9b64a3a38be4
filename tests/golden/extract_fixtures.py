"""Write the golden fixtures from the data files the reference's own tests hold.

Run in a checkout that has the reference at ``/root/reference`` (read-only). Nothing
of the reference is imported or executed: the files below are JSON data, read as
data.

- ``rk_traces.json.gz``: the Runge-Kutta decision corpus
  (``tests/data/rk_traces.json.gz``, 29 scripted scenarios), copied unchanged.
- ``adams_traces.json.gz``: the Adams decision corpus (``tests/data/adams_traces.json.gz``),
  copied unchanged.
- ``luna_plasma_witness.npz``: the input signals, parameters and the Luna plasma stages
  of ``tests/data/luna_plasma.json`` that the plasma column is compared with (the
  SCALAR, ZEROS, PPT and VECTOR cases and their free-current halves).
- ``luna_linear_witness.npz``: every stored quantity of ``tests/data/luna_linear.json``
  (Luna's frequency axes, band masks, beta and alpha of the capillary mode, the linear
  operators, prefactors, gamma mappings and Kerr responses), its input signals and its
  numeric parameters (the capillary's, the GNLSE fibre's and the Kerr case's).
- ``luna_rdw_density.npz``: the density paths of ``tests/data/luna_rdw_media.json`` (the
  full and reduced runs' z axes and helium densities) and its constants, with each
  run's fill (inlet and outlet pressure, length, radius) and the temperature from
  ``tests/data/luna_rdw.json``.
"""

from __future__ import annotations

import json
import shutil
from pathlib import Path

import numpy as np

REFERENCE = Path("/root/reference/tests/data")
HERE = Path(__file__).resolve().parent


def main() -> None:
    for name in ("rk_traces.json.gz", "adams_traces.json.gz"):
        shutil.copyfile(REFERENCE / name, HERE / name)
    data = json.loads((REFERENCE / "luna_plasma.json").read_text())
    luna, signals, parameters = data["luna"], data["signals"], data["parameters"]
    keep = {}
    for key in ("t", "field", "zeroed", "field_x", "field_y"):
        keep["signal_" + key] = np.array(signals[key])
    for case in ("SCALAR", "ZEROS", "PPT", "VECTOR"):
        for suffix in ("", "_FREE"):
            for stage in ("RATE", "FRACTION", "J", "P", "J_X", "J_Y", "P_X", "P_Y"):
                key = f"{case}{suffix}_{stage}"
                if key in luna:
                    keep["luna_" + key] = np.array(luna[key])
    keep["luna_CONSTANTS"] = np.array(luna["CONSTANTS"])
    for key, value in parameters.items():
        if isinstance(value, int | float):
            keep["param_" + key] = np.array(value)
    np.savez_compressed(HERE / "luna_plasma_witness.npz", **keep)

    data = json.loads((REFERENCE / "luna_linear.json").read_text())
    keep = {"luna_" + key: np.array(value) for key, value in data["luna"].items()}
    for key, value in data["signals"].items():
        keep["signal_" + key] = np.array(value)
    for group, values in data["parameters"].items():
        if not isinstance(values, dict):
            continue
        for key, value in values.items():
            if isinstance(value, int | float) or (
                    isinstance(value, list) and all(isinstance(v, int | float) for v in value)):
                keep[f"param_{group}_{key}"] = np.array(value, dtype=float)
    np.savez_compressed(HERE / "luna_linear_witness.npz", **keep)

    media = json.loads((REFERENCE / "luna_rdw_media.json").read_text())
    keep = {key: np.array(media[key]) for key in (
        "CONSTANTS", "HE_DENSITIES", "FULL_Z", "FULL_DENSITY", "REDUCED_Z", "REDUCED_DENSITY")}
    cases = json.loads((REFERENCE / "luna_rdw.json").read_text())
    for case in ("full", "reduced"):
        for key in ("p_in", "p_out", "length", "radius"):
            keep[f"{case.upper()}_{key}"] = np.array(float(cases[case][key]))
    keep["temperature_k"] = np.array(float(cases["temperature_k"]))
    np.savez_compressed(HERE / "luna_rdw_density.npz", **keep)


if __name__ == "__main__":
    main()
