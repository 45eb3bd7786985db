"""Helpers that run the reference build (oracle/_ref/ref_driver, compiled
from the reference's own serial sources) on a mesh and load its dumps."""
from __future__ import annotations

import os
import subprocess

import numpy as np

STATE = ("rho", "u", "v", "w", "p")


def run_reference(ref_driver: str, mesh: str, bc: str, out_dir: str, mach: float, alpha: float, cfl: float,
                  iters: int, dumps=(0, 1), ghost_refresh: bool = False, ramp: tuple[float, int] | None = None,
                  geometry: bool = True) -> dict:
    """ramp = (cfl_start, cfl_nstep): CFL ramp from cfl_start to `cfl`."""
    os.makedirs(out_dir, exist_ok=True)
    cmd = [ref_driver, mesh, bc, out_dir, repr(mach), repr(alpha), repr(cfl), str(iters),
           ",".join(str(d) for d in dumps)] + (["ghosts"] if ghost_refresh else []) + \
        ([f"ramp={ramp[0]!r},{ramp[1]}"] if ramp else []) + ([] if geometry else ["nogeo"])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, timeout=900)
    return load_dir(out_dir, iters)


def load_dir(out_dir: str, iters: int) -> dict:
    r = {"dir": out_dir}
    h = os.path.join(out_dir, "history.f64")
    if os.path.exists(h):
        r["history"] = np.fromfile(h).reshape(-1, 7)
    return r


def load_state(out_dir: str, it: int, name: str) -> np.ndarray:
    return np.fromfile(os.path.join(out_dir, f"it{it}_{name}.f64"))


def load_geo(out_dir: str, name: str) -> np.ndarray:
    dt = np.int32 if name.endswith("i32") else np.float64
    return np.fromfile(os.path.join(out_dir, name), dtype=dt)
