"""Energy and temperature series in the reference's output files.

The reference writes two raw little-endian arrays after a run
(Environment::postRunEvents, Environment.cpp:146-148, via dumpToFile's fwrite,
Environment.cpp:256-266):

* ``energy.bin``: one Float3 per analysed step (three float32: potential,
  kinetic and total energy averaged per particle, Analyzer.cu:137-207). The
  reference's potential is what its kernels accumulate, the bonded terms only
  (LJ contributes nothing, Engine.cu:475-522); ``lj=True`` writes the full
  potential of this build instead.
* ``temperature.bin``: one float32 per thermostat evaluation (Engine.cu:369).

Here the series comes from ``Engine.step_observed`` (lima_step_observed): one
record per observed step, taken in the step itself instead of from a stored
trajectory.  ``logdata.bin`` and ``sim_traindata.bin`` (Environment.cpp:130-141)
come from ``Engine.step_logged`` (lima_step_logged).
"""
from __future__ import annotations

import os

import numpy as np


def energy_series(records: list, system: int, n_particles: int, lj: bool = False) -> np.ndarray:
    """[n_records, 3] float32: potential, kinetic, total energy per particle."""
    out = np.zeros((len(records), 3), np.float32)
    for k, rec in enumerate(records):
        e = rec[system]
        pot = e["potential"] if lj else e["bond"] + e["angle"] + e["dihedral"]
        out[k] = (pot / n_particles, e["kinetic"] / n_particles, (pot + e["kinetic"]) / n_particles)
    return out


def temperature_series(records: list, system: int) -> np.ndarray:
    return np.array([rec[system]["temperature"] for rec in records], np.float32)


def write_reference_dumps(records: list, engine, out_dir: str, system: int = 0, lj: bool = False) -> dict:
    """Writes energy.bin and temperature.bin for one system; returns their paths."""
    s = engine.simulation.systems[system]
    n = s.n_chain + s.n_solvents
    os.makedirs(out_dir, exist_ok=True)
    paths = {"energy": os.path.join(out_dir, "energy.bin"), "temperature": os.path.join(out_dir, "temperature.bin")}
    energy_series(records, system, n, lj).astype("<f4").tofile(paths["energy"])
    temperature_series(records, system).astype("<f4").tofile(paths["temperature"])
    return paths


def write_reference_logs(logdata: np.ndarray, traindata: np.ndarray | None, out_dir: str,
                         critical_error: bool = False) -> dict:
    """Writes logdata.bin (10 float32 per step) and sim_traindata.bin (per step, compound and
    particle slot: 3 Float3 = 9 float32), as Environment::postRunEvents does
    (Environment.cpp:130-141); the training data is skipped after a critical error, as there.
    The arrays are Engine.step_logged's."""
    os.makedirs(out_dir, exist_ok=True)
    paths = {"logdata": os.path.join(out_dir, "logdata.bin")}
    np.ascontiguousarray(logdata, dtype="<f4").tofile(paths["logdata"])
    if traindata is not None and not critical_error:
        paths["traindata"] = os.path.join(out_dir, "sim_traindata.bin")
        np.ascontiguousarray(traindata, dtype="<f4").tofile(paths["traindata"])
    return paths
