"""Autotuner for the layout and launch options of a context (tile size of
the LU-SGS sweeps, waves per sweep workgroup, flat or per-tile sweep
schedule).  Results do not depend on
the options (every per-cell sum has a fixed order), only speed does, so the
tuner times a few iterations of each candidate on the actual mesh and keeps
the fastest.  `python -m minifs_amd.tune` measures a ladder of box+bump
meshes and writes the table `tuned.json` beside this file, which
`tuned_options(n_cells)` reads.
"""
from __future__ import annotations

import argparse
import json
import os
import time

from . import MFS_FP64, MFS_MIXED, Grid, Solver, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(_HERE, "tuned.json")
# (tile_cells, sweep_waves, sweep_by_tiles)
CANDIDATES = ((4096, 4, 0), (4096, 8, 0), (4096, 16, 0), (16384, 8, 0), (2048, 8, 1), (4096, 8, 1), (8192, 16, 1))


def time_options(grid: Grid, params, precision: int, tile_cells: int, sweep_waves: int, sweep_by_tiles: int = 0,
                 steps: int = 10, warmup: int = 2) -> float:
    """Milliseconds per iteration of mfs_step with these options."""
    s = Solver(grid, params, precision=precision, tile_cells=tile_cells, sweep_waves=sweep_waves,
               sweep_by_tiles=sweep_by_tiles)
    try:
        s.step(warmup)
        s.sync()
        t = time.perf_counter()
        s.step(steps)
        s.sync()
        return (time.perf_counter() - t) / steps * 1e3
    finally:
        s.close()


def tune(grid: Grid, params, precision: int = MFS_FP64, candidates=CANDIDATES, steps: int = 10) -> tuple[dict, list]:
    results = []
    for tile, waves, by_tiles in candidates:
        if by_tiles and tile > grid.info["nTCell"]:
            continue  # one tile already holds the mesh
        ms = time_options(grid, params, precision, tile, waves, by_tiles, steps)
        results.append({"tile_cells": tile, "sweep_waves": waves, "sweep_by_tiles": by_tiles,
                        "ms_per_step": round(ms, 4)})
    best = min(results, key=lambda r: r["ms_per_step"])
    return best, results


def tuned_options(n_cells: int, precision: int = MFS_FP64) -> dict:
    """The tuned options of the nearest measured mesh size (defaults when no
    table has been written)."""
    try:
        with open(TABLE) as f:
            table = json.load(f)
    except OSError:
        return {"tile_cells": 0, "sweep_waves": 0, "sweep_by_tiles": 0}
    rows = [r for r in table["rows"] if r["precision"] == precision]
    if not rows:
        return {"tile_cells": 0, "sweep_waves": 0, "sweep_by_tiles": 0}
    import math

    r = min(rows, key=lambda r: abs(math.log(r["cells"]) - math.log(max(n_cells, 1))))
    b = r["best"]
    return {"tile_cells": b["tile_cells"], "sweep_waves": b["sweep_waves"], "sweep_by_tiles": b.get("sweep_by_tiles", 0)}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--sizes", type=int, nargs="+", default=[15, 25, 40, 60, 80, 100], help="hexes per axis")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--write", action="store_true", help=f"write {TABLE}")
    args = ap.parse_args()
    rows = []
    for n in args.sizes:
        p = make_params(mach=0.8, alpha=1.25, cfl=10.0)
        g = Grid.box_bump(n).decouple().preprocess(p)
        for prec in (MFS_FP64, MFS_MIXED):
            best, res = tune(g, p, prec, steps=args.steps)
            row = {"hex_per_axis": n, "cells": g.info["nTCell"], "precision": prec, "best": best, "all": res}
            rows.append(row)
            print(json.dumps(row), flush=True)
        g.close()
    if args.write:
        with open(TABLE, "w") as f:
            json.dump({"workload": "box+bump tetrahedra, Mach 0.8, AoA 1.25, CFL 10", "rows": rows}, f, indent=1)


if __name__ == "__main__":
    main()
