#!/usr/bin/env python3
"""Throughput of the MiniFS solver path on one MI355X.

Workload (default): the box + Gaussian-bump mesh at 100x100x100 hexes x 6
tetrahedra (6.0M cells, 12.06M faces), Mach 0.5, AoA 0, CFL 10, plain FP64,
TDTree-ordered.  One step is one iteration of the reference driver loop
(src/main.cpp:201-291): residual (Roe flux), local time step, LU-SGS
forward/backward sweeps with the flow-field update, Green-Gauss gradients.
The mesh and state are resident in HBM before the timed region.

Prints one JSON line.  `roofline` describes the kernel with the largest
share of the timed region, timed with HIP events on the solver stream;
`cpu_baseline` times the reference's own serial build (oracle/_ref) on a
bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M cell-iters/s FP64 vs FP32/FP64 mixed; served solves/hr + p50/p99 latency"
HBM_PEAK_GBS = 8000.0

# Compulsory bytes per launch of each kernel group (8-byte reals, 4-byte
# indices; every input read once, every output written once): per face,
# per interior cell, per boundary face (ghost), per node.
KERNEL_BYTES = {
    # c1,c2 + normal/area per face; q read + res write per cell; ghost q
    "flux": (40, 80, 40, 0),
    # c1,c2 + normal/area + two distances per face; q, vol read, dt, Diag write; ghost p
    "timestep": (56, 64, 8, 0),
    # flux and time step in one pass (mfs_step): face tables once, q + vol read,
    # res + dt + Diag written, ghost q
    "flux_ts": (56, 104, 40, 0),
    # forward sweep: faces, q, res, Diag read, DQ write, ghost q
    "fwd_sweep": (40, 144, 40, 0),
    # backward sweep + update: as the forward sweep plus q write (K5 + K6)
    "bwd_sweep": (40, 264, 40, 0),
    # node values: 4 weighted entries per tetrahedron (index + weight), q, node write
    "node": (0, 88, 76, 52),
    # Green-Gauss: face nodes + geometry, q, node values, vol, 15 gradient outputs
    "gradient": (52, 128, 80, 40),
}
# whole iteration, plain FP64 (268 nF + 840 nC + 356 nB + 92 nN)
ITER_BYTES = (268, 840, 356, 92)


def alg_bytes(coef, nF, nC, nB, nN):
    return coef[0] * nF + coef[1] * nC + coef[2] * nB + coef[3] * nN


def _cpu_mesh(td: str, grid_n: int, seed: int):
    import minifs_amd as mfs

    g = mfs.Grid.box_bump(grid_n, seed=seed).decouple()
    mesh, bc = os.path.join(td, "m.mfl"), os.path.join(td, "m.bc")
    g.save(mesh, bc)
    nC = g.info["nTCell"]
    g.close()
    return mesh, bc, nC


def cpu_sample(grid_n: int, iters: int, mach: float, alpha: float, cfl: float, seed: int) -> dict | None:
    """Times the reference's own serial build (oracle/_ref/ref_driver, built
    from the reference sources with g++ -O3) on a bounded sample of the same
    workload: `iters` iterations of the n^3 x 6 box+bump mesh, loop time only."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as td:
        mesh, bc, nC = _cpu_mesh(td, grid_n, seed)
        out = os.path.join(td, "ref")
        os.makedirs(out)
        subprocess.run([ref, mesh, bc, out, repr(mach), repr(alpha), repr(cfl), str(iters), ""], check=True,
                       stdout=subprocess.DEVNULL, timeout=1200)
        with open(os.path.join(out, "summary.txt")) as f:
            first = f.readline().split()
        kv = dict(zip(first[0::2], first[1::2]))
        t_loop = float(kv["t_loop"])
        return {
            "value": round(nC * iters / t_loop / 1e6, 4),
            "unit": "M cell-iters/s",
            "cores": 1,
            "kind": "reference",
            "sample": f"reference serial build (oracle/_ref/ref_driver, g++ -O3 -ffp-contract=off), {iters} "
                      f"iterations of the {grid_n}^3x6 = {nC} cell box+bump mesh, same flow condition, "
                      f"loop time only ({t_loop:.1f} s)",
        }


def cpu_port_sample(grid_n: int, iters: int, mach: float, alpha: float, cfl: float, seed: int) -> dict | None:
    """Times the C restatement (oracle/liboracle.so) with OpenMP on all host
    threads: per-cell gathers and level-scheduled sweeps, the arithmetic of
    the reference's serial loop."""
    from oracle import LIB, Oracle

    if not os.path.exists(LIB):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as td:
        mesh, bc, nC = _cpu_mesh(td, grid_n, seed)
        o = Oracle(mesh, bc, mach, alpha, cfl)
        _, t_loop = o.run(iters, threads)
        o.close()
    return {
        "value": round(nC * iters / t_loop / 1e6, 4),
        "unit": "M cell-iters/s",
        "cores": threads,
        "kind": "port",
        "sample": f"C restatement, {threads} OpenMP threads, {iters} iterations of the {grid_n}^3x6 = {nC} cell "
                  f"box+bump mesh, loop time only ({t_loop:.1f} s)",
    }


def copy_calibration(mib: int, reps: int) -> int:
    """A plain device-to-device copy of known size, run under
    `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE): the counters' reading per
    copied byte is the scale factor that turns the solver kernels' readings
    into HBM bytes."""
    import torch

    n = mib * 1024 * 1024 // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                                           "launches": reps}}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100, help="hexes per axis (cells = 6 n^3)")
    ap.add_argument("--mach", type=float, default=0.5)
    ap.add_argument("--alpha", type=float, default=0.0)
    ap.add_argument("--cfl", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--cpu-n", type=int, default=40, help="mesh of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-port-iters", type=int, default=60)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 5)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import minifs_amd as mfs

    t0 = time.time()
    p = mfs.make_params(mach=args.mach, alpha=args.alpha, cfl=args.cfl)
    g = mfs.Grid.box_bump(args.n, seed=args.seed).decouple().preprocess(p)
    info = g.info
    s = mfs.Solver(g, p)
    setup_s = time.time() - t0
    nC, nF, nB, nN = info["nTCell"], info["nTFace"], info["nBFace"], info["nTNode"]

    s.step(args.warmup)
    s.sync()
    s.enable_timing(not args.no_kernel_timing)
    s.reset_timing()
    t1 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    ms_per_step = elapsed / args.steps * 1e3
    value = nC * args.steps / elapsed / 1e6

    kern = {}
    for k in KERNEL_BYTES:
        ms, cnt = s.timing(k)
        if cnt:
            kern[k] = (ms, cnt)
    hist = s.history()
    roof = None
    if kern:
        dom = max(kern, key=lambda k: kern[k][0])
        ms, cnt = kern[dom]
        avg_s = ms / cnt / 1e3
        b = alg_bytes(KERNEL_BYTES[dom], nF, nC, nB, nN)
        achieved = b / avg_s / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_{dom}.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("bytes_per_launch")
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "alg_bytes_per_launch": int(b), "avg_launch_ms": round(ms / cnt, 4)}
    iter_b = alg_bytes(ITER_BYTES, nF, nC, nB, nN)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M cell-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 LU-SGS iteration, box+bump {args.n}^3x6 tets", "cells": nC,
                   "faces": nF, "bfaces": nB, "nodes": nN, "mach": args.mach, "alpha_deg": args.alpha,
                   "cfl": args.cfl, "seed": args.seed, "ordering": "TDTree", "sweep_levels": s.info()["sweep_levels"]},
        "roofline": roof,
        "iteration_roofline": {"alg_bytes_per_iter": int(iter_b),
                               "achieved_GBs": round(iter_b / (ms_per_step / 1e3) / 1e9, 2),
                               "frac": round(iter_b / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)},
        "kernel_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in kern.items()},
        "setup_s": round(setup_s, 2),
        "last_residual": [float(x) for x in hist[-1][:5]] if len(hist) else None,
        "cpu_baseline": None,
        "cpu_port_threaded": None,
    }
    s.close()
    g.close()
    if rank == 0 and args.cpu_n > 0:
        try:
            line["cpu_baseline"] = cpu_sample(args.cpu_n, args.cpu_iters, args.mach, args.alpha, args.cfl, args.seed)
        except Exception as e:  # reported, not fatal: the GPU number stands on its own
            line["cpu_baseline"] = {"error": str(e)}
        try:
            line["cpu_port_threaded"] = cpu_port_sample(args.cpu_n, args.cpu_port_iters, args.mach, args.alpha,
                                                        args.cfl, args.seed)
        except Exception as e:
            line["cpu_port_threaded"] = {"error": str(e)}
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
